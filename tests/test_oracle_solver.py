"""Invariants of the oracle's SolverPPC restatement on one periodic patch, for the parts no
reference code or golden file pins here (SAMRAI same-level fills, the seeded loader, the stage
sequence): exact charge conservation, moments equal on periodic images, deposit partition of
unity over a period, uniform-plasma quiet start, loader statistics.
"""
import numpy as np
import pytest

import oracle_lib as O
from phare_amd import workloads as W


def _sim(wl):
    s = O.Sim(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te)
    wl.setup(s, O.F_B)
    return s


def test_loader_invariants():
    wl = W.uniform1d(ncells=50, ppc=40)
    s = _sim(wl)
    p = s.get_particles()
    assert s.count() == 50 * 40
    assert p["icell"].min() == 0 and p["icell"].max() == 49
    assert np.all((p["delta"] >= 0) & (p["delta"] < 1))
    # cells visited in row-major order, ppc particles each
    assert np.all(np.diff(p["icell"][:, 0]) >= 0)
    assert abs(p["v"].mean()) < 0.05 and abs(p["v"].std() - 0.3) < 0.02
    np.testing.assert_allclose(p["weight"], 1.0 / 40)


def test_loader_density_cutoff():
    wl = W.uniform1d(ncells=20, ppc=5)
    s = O.Sim(1, 1, [20], [0.2])
    n = np.ones(20)
    n[3] = 1e-20
    s.load_maxwellian(0, n, [np.zeros(20)] * 3, [np.ones(20)] * 3, 1.0, 5, 7)
    assert s.count() == 19 * 5
    assert 3 not in set(s.get_particles()["icell"][:, 0])


@pytest.mark.parametrize("dim,order", [(1, 1), (1, 2), (1, 3), (2, 1), (2, 2), (2, 3), (3, 1), (3, 2)])
def test_moments_partition_of_unity_and_periodic_images(dim, order):
    """after moveIons the density summed over one period of primal nodes equals sum(w), and
    nodes that are periodic images of each other hold the same value"""
    if dim == 1:
        wl = W.uniform1d(ncells=40, ppc=20, order=order)
    elif dim == 2:
        wl = W.harris2d(ncells=(24, 20), ppc=8, order=order)
    else:
        wl = W.uniform3d(ncells=(10, 9, 10), ppc=4, order=order)
    s = _sim(wl)
    s.average()
    s.move_ions(wl.dt, 2)
    g = W.GHOST[order]
    n = s.get(O.f_pop(0, 0))
    per = tuple(slice(g, g + wl.ncells[d]) for d in range(dim))
    w = s.get_particles()["weight"]
    np.testing.assert_allclose(n[per].sum(), w.sum(), rtol=1e-12)
    for d in range(dim):
        lo = [slice(None)] * dim
        hi = [slice(None)] * dim
        lo[d] = slice(0, 2 * g + 1)
        hi[d] = slice(wl.ncells[d], wl.ncells[d] + 2 * g + 1)
        np.testing.assert_allclose(n[tuple(lo)], n[tuple(hi)], rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("dim", [1, 2])
def test_advance_conserves_charge_and_count(dim):
    wl = W.uniform1d(ncells=60, ppc=30) if dim == 1 else W.harris2d(ncells=(24, 20), ppc=6)
    s = _sim(wl)
    d0 = s.diagnostics()
    s.advance(wl.dt, 5)
    d1 = s.diagnostics()
    assert d1["count"] == d0["count"]
    assert d1["charge"] == pytest.approx(d0["charge"], rel=1e-14)
    assert d1["kinetic"] == pytest.approx(d0["kinetic"], rel=1e-2)


def test_uniform_plasma_quiet_start():
    """uniform density, uniform B, no drift: E stays at the particle-noise level"""
    wl = W.uniform1d(ncells=100, ppc=200)
    s = _sim(wl)
    s.advance(wl.dt, 10)
    for c in range(3):
        assert np.abs(s.get(O.F_E + c)).max() < 0.1
    np.testing.assert_allclose(s.get(O.F_B + 0)[2:-2], 1.0, atol=1e-12)


def test_two_cell_move_raises():
    """boris.hpp:164-165: |delta| > 2 in a half step is an error (test_exceptions.py uses
    v=1000, dt=0.001, dl=0.3)"""
    s = O.Sim(1, 1, [20], [0.3])
    p = dict(icell=np.array([[5]], np.int32), delta=np.array([[0.5]]), v=np.array([[1000.0, 0, 0]]),
             weight=np.ones(1), charge=np.ones(1))
    s.set_particles(0, 0, p)
    with pytest.raises(RuntimeError):
        s.update_populations(0.001 * 2, 2)
