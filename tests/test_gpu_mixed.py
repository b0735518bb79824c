"""The mixed-precision family (FP32 delta, velocity and weight, int32 cells; FP64 fields,
moments and reductions) against the FP64 oracle on identical seeded inputs.

Tolerances are max-norm relative errors over each array (parity_util.max_rel):
 - moments (n, rho, F) of the initial deposit and after one PPC step: 1e-5;
 - after 20 PPC steps: moments 1e-4, E and B 1e-4 (FP32 trajectories drift from the FP64 ones by
   a few 1e-7 per step; the moments and fields follow them);
 - total energy drift over a long run: within 1e-4 of the FP64 path's (which matches the oracle at
   1e-12), and the total charge exact to FP32 weight rounding (1e-6).
Every test prints the errors it measured.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
import parity_util as U
from phare_amd import phx
from phare_amd import workloads as W

pytestmark = pytest.mark.gpu

MOMENT_TOL_1 = 1e-5
STATE_TOL_20 = 1e-4
DRIFT_TOL = 1e-4

WORKLOADS = {
    "uniform1d": lambda: W.uniform1d(),
    "harris2d-o1": lambda: W.harris2d(ncells=(32, 24), ppc=12, order=1),
    "harris2d-o2": lambda: W.harris2d(ncells=(32, 24), ppc=12, order=2),
    "harris2d-o3": lambda: W.harris2d(ncells=(32, 24), ppc=12, order=3),
    # 3D: the staged path's brick-ordered linear FP32 storage, bricks aligned with the push's
    # chunks (64 ppc on 8^3 cells) and not (8 and 16 ppc on uneven boxes)
    "uniform3d-o1": lambda: W.uniform3d(ncells=(8, 8, 8), ppc=64, order=1),
    "uniform3d-o2": lambda: W.uniform3d(ncells=(12, 11, 10), ppc=8, order=2),
    "uniform3d-o3": lambda: W.uniform3d(ncells=(12, 12, 12), ppc=16, order=3),
}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair(wl):
    g = phx.Level(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te,
                  capacity=wl.n_particles, precision=phx.PRECISION_FP32)
    o = O.Sim(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te)
    wl.setup(g, phx.F_B)
    wl.setup(o, O.F_B)
    return g, o


def _moment_errors(g, o):
    e = {f"moment{k}": U.max_rel(g.get(phx.f_pop(0, k)), o.get(O.f_pop(0, k))) for k in range(5)}
    e["rho"] = U.max_rel(g.get(phx.F_RHO), o.get(O.F_RHO))
    return e


def _field_errors(g, o):
    return {f"field{fid}": U.max_rel(g.get(fid), o.get(fid)) for fid in range(6)}


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_mixed_initial_deposit(name):
    _say(f"[gpu] mixed initial deposit {name}")
    wl = WORKLOADS[name]()
    g, o = _pair(wl)
    e = _moment_errors(g, o)
    _say("  errors", {k: f"{v:.2e}" for k, v in e.items()})
    for k, v in e.items():
        assert v <= MOMENT_TOL_1, f"{k}: {v:.3e}"
    # the count and the total charge survive the FP32 conversion
    d32, d64 = g.diagnostics(), o.diagnostics()
    assert d32["count"] == d64["count"]
    assert abs(d32["charge"] - d64["charge"]) <= 1e-6 * abs(d64["charge"])
    g.close()


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_mixed_one_step(name):
    _say(f"[gpu] mixed one PPC step {name}")
    wl = WORKLOADS[name]()
    g, o = _pair(wl)
    g.advance(wl.dt, 1)
    o.advance(wl.dt, 1)
    e = _moment_errors(g, o)
    f = _field_errors(g, o)
    _say("  moments", {k: f"{v:.2e}" for k, v in e.items()})
    _say("  fields", {k: f"{v:.2e}" for k, v in f.items()})
    for k, v in e.items():
        assert v <= MOMENT_TOL_1, f"{k}: {v:.3e}"
    for k, v in f.items():
        assert v <= STATE_TOL_20, f"{k}: {v:.3e}"
    g.close()


@pytest.mark.parametrize("name", ["uniform1d", "harris2d-o2", "uniform3d-o1"])
def test_mixed_twenty_steps(name):
    _say(f"[gpu] mixed 20 PPC steps {name}")
    wl = WORKLOADS[name]()
    g, o = _pair(wl)
    g.set_sort_interval(7)  # 3D: the device brick re-sort of the FP32 storage
    g.advance(wl.dt, 20)
    o.advance(wl.dt, 20)
    e = _moment_errors(g, o)
    f = _field_errors(g, o)
    _say("  moments", {k: f"{v:.2e}" for k, v in e.items()})
    _say("  fields", {k: f"{v:.2e}" for k, v in f.items()})
    for k, v in list(e.items()) + list(f.items()):
        assert v <= STATE_TOL_20, f"{k}: {v:.3e}"
    assert g.diagnostics()["count"] == o.diagnostics()["count"]
    g.close()


def test_mixed_two_populations():
    """two populations (mass 1 and 4, charge 1 and 2): moments of each and the fields after 5
    PPC steps against the FP64 oracle"""
    _say("[gpu] mixed two populations")
    wl = W.harris2d(ncells=(32, 24), ppc=8, order=2)
    wl.masses, wl.charges = (1.0, 4.0), (1.0, 2.0)
    g, o = _pair(wl)
    g.advance(wl.dt, 5)
    o.advance(wl.dt, 5)
    e = {f"pop{p}.moment{k}": U.max_rel(g.get(phx.f_pop(p, k)), o.get(O.f_pop(p, k)))
         for p in (0, 1) for k in range(5)}
    e.update(_field_errors(g, o))
    _say("  errors", {k: f"{v:.2e}" for k, v in e.items()})
    for k, v in e.items():
        assert v <= STATE_TOL_20, f"{k}: {v:.3e}"
    g.close()


def test_mixed_energy_drift_matches_fp64():
    """1000 PPC steps of a 2D Harris sheet (64x32 cells, 50 ppc, order 2) in both families: the
    total-energy drifts agree within 1e-4 and the total charge is conserved"""
    _say("[gpu] mixed vs plain energy drift, 1000 steps")
    wl = W.harris2d(ncells=(64, 32), ppc=50, order=2)
    drifts = {}
    charges = {}
    for prec in (phx.PRECISION_FP64, phx.PRECISION_FP32):
        lv = phx.Level(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te,
                       capacity=wl.n_particles, precision=prec)
        wl.setup(lv, phx.F_B)
        d0 = lv.diagnostics()
        lv.advance(wl.dt, 1000)
        d1 = lv.diagnostics()
        e0 = d0["kinetic"] + d0["magnetic"]
        e1 = d1["kinetic"] + d1["magnetic"]
        drifts[prec] = (e1 - e0) / e0
        charges[prec] = (d1["charge"] - d0["charge"]) / d0["charge"]
        lv.close()
    _say(f"  drift fp64 {drifts[phx.PRECISION_FP64]:.3e} fp32 {drifts[phx.PRECISION_FP32]:.3e}"
         f" charge fp64 {charges[phx.PRECISION_FP64]:.2e} fp32 {charges[phx.PRECISION_FP32]:.2e}")
    assert abs(drifts[phx.PRECISION_FP32] - drifts[phx.PRECISION_FP64]) <= DRIFT_TOL
    assert abs(charges[phx.PRECISION_FP32]) <= 1e-6


def test_mixed_two_cell_move_is_an_error():
    """the FP32 push raises the reference's two-cell error (boris.hpp:164-165) like the FP64 one;
    particles set after the level's sort are sorted again before the push"""
    _say("[gpu] mixed two-cell move error")
    g = phx.Level(1, 1, [20], [0.3], precision=phx.PRECISION_FP32)
    p = dict(icell=np.array([[5]], np.int32), delta=np.array([[0.5]]),
             v=np.array([[1000.0, 0, 0]]), weight=np.ones(1), charge=np.ones(1))
    g.set_particles(0, phx.DOMAIN, p)
    with pytest.raises(phx.PhxError) as e:
        g.move_ions(0.002, phx.MODE_ALL)
    assert "moved 2 cells" in str(e.value)
    g.close()


def test_mixed_rejects_drop_in_calls():
    """the mixed-precision family is the resident level path: updatePopulations on host-bound
    arrays, ghost arrays and non-periodic patches are refused with an error, not run in FP64"""
    _say("[gpu] mixed family refusals")
    wl = WORKLOADS["harris2d-o2"]()
    g, _ = _pair(wl)
    with pytest.raises(phx.PhxError):
        g.update_populations(wl.dt, phx.MODE_ALL)
    p = U.random_particles(np.random.default_rng(3), 2, 10, wl.ncells, [0, 0], 1)
    with pytest.raises(phx.PhxError):
        g.set_particles(0, phx.PATCH_GHOST, p)
    g.close()
    with pytest.raises(phx.PhxError):
        phx.Level(2, 2, (16, 16), (0.4, 0.4), periodic=False, precision=phx.PRECISION_FP32)
