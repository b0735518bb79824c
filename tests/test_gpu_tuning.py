"""The launch shapes of the particle stage (phx_get_tuning / phx_set_tuning, chosen by
phare_amd/autotune.py): every tuning gives the same physics (identical cells, 1e-12 against the
oracle; only the summation order of the deposit differs), on one patch and on a level of uneven
patches, and out-of-range tunings are refused."""
import sys

import numpy as np
import pytest

import oracle_lib as O
import parity_util as U
from phare_amd import phx
from phare_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


TUNINGS = [
    dict(segment_shape=[4, 16, 1], rows_per_item=8, tile_row_pad=1, sort_interval=3),
    dict(segment_shape=[1, 64, 1], rows_per_item=64, tile_row_pad=32, sort_interval=5),
    dict(segment_shape=[8, 8, 1], rows_per_item=1, tile_row_pad=8, sort_interval=2),
    dict(segment_shape=[2, 32, 1], rows_per_item=48, tile_row_pad=16, sort_interval=4,
         deposit_window=1),
]


@pytest.mark.parametrize("dim,order", [(1, 1), (1, 3), (2, 1), (2, 2), (2, 3)])
def test_deposit_window_matches_the_tile_deposit(dim, order):
    """the mixed-precision push with each lane's particles of one cell summed in a register
    window first: moments within 1e-5 of the per-particle tile deposit after a PPC step"""
    _say(f"[gpu] deposit window dim={dim} order={order}")
    wl = W.harris2d(ncells=(40, 24), ppc=12, order=order) if dim == 2 else W.uniform1d(ncells=60, ppc=50, order=order)
    out = []
    for window in (0, 1):
        g = phx.Level(dim, order, wl.ncells, wl.dx, capacity=wl.n_particles,
                      precision=phx.PRECISION_FP32, apply_tuning_table=False)
        g.set_tuning(deposit_window=window)
        wl.setup(g, phx.F_B)
        g.advance(wl.dt, 2)
        out.append([g.get(phx.f_pop(0, m)) for m in range(5)] + [g.get(phx.F_E + k) for k in range(6)])
        g.close()
    for m, (a, b) in enumerate(zip(*out)):
        U.assert_close(b, a, 1e-5, what=f"array {m}")


@pytest.mark.parametrize("k", range(len(TUNINGS)))
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tunings_keep_the_physics(k, precision):
    _say(f"[gpu] tuning {k} {precision}")
    t = TUNINGS[k]
    wl = W.harris2d(ncells=(40, 24), ppc=10, order=2)
    prec = phx.PRECISION_FP32 if precision == "fp32" else phx.PRECISION_FP64
    g = phx.Level(2, 2, wl.ncells, wl.dx, capacity=wl.n_particles, precision=prec,
                  apply_tuning_table=False)
    g.set_tuning(**t)
    got = g.get_tuning()
    assert got["segment_shape"][:2] == t["segment_shape"][:2] and got["rows_per_item"] == t["rows_per_item"]
    o = O.Sim(2, 2, wl.ncells, wl.dx)
    wl.setup(g, phx.F_B)
    wl.setup(o, O.F_B)
    # a change of tuning on a sorted level re-sorts it and drops its captured steps
    g.advance(wl.dt, 4)
    g.set_tuning(rows_per_item=t["rows_per_item"] * 2)
    g.advance(wl.dt, 6)
    o.advance(wl.dt, 10)
    rtol = U.RTOL if precision == "fp64" else 1e-4
    for fid in range(6):
        U.assert_close(g.get(fid), o.get(fid), rtol, what=f"field {fid}")
    for m in range(5):
        U.assert_close(g.get(phx.f_pop(0, m)), o.get(O.f_pop(0, m)), rtol, what=f"moment {m}")
    if precision == "fp64":
        U.assert_same_particles(g.get_particles(), o.get_particles())
    g.close()


def test_tuning_of_a_level_of_uneven_patches():
    _say("[gpu] tuning of a multi-patch level")
    wl = W.harris2d(ncells=(48, 40), ppc=6, order=2)
    cuts = ((8, 24, 40), (10, 20, 28))
    import itertools

    edges = [[0] + list(c) + [n] for n, c in zip(wl.ncells, cuts)]
    patches = []
    for i, j in itertools.product(range(4), range(4)):
        patches.append(((edges[0][i], edges[1][j]), (edges[0][i + 1] - 1, edges[1][j + 1] - 1)))
    g = phx.Level(2, 2, wl.ncells, wl.dx, capacity=wl.n_particles, patches=patches,
                  apply_tuning_table=False)
    o = O.Sim(2, 2, wl.ncells, wl.dx)
    g.set_tuning(segment_shape=[4, 16, 1], rows_per_item=4, sort_interval=2)
    wl.setup(g, phx.F_B)
    wl.setup(o, O.F_B)
    g.advance(wl.dt, 6)
    o.advance(wl.dt, 6)
    for fid in range(6):
        U.assert_close(g.get(fid), o.get(fid), what=f"field {fid}")
    U.assert_same_particles(g.get_particles(), o.get_particles())
    g.close()


def test_segment_tile_beyond_lds_falls_back_to_direct_deposits(capfd):
    """3D order 3 with segments of 1x1x64 cells: the segment tile (8x8x71 nodes, 5 FP64 slabs)
    does not fit in LDS, so every deposit goes straight to the moment arrays; the run says so once
    on stderr and still equals the oracle's at 1e-12 (3 PPC steps, sorted every step)"""
    _say("[gpu] segment tile beyond LDS")
    wl = W.uniform3d(ncells=(10, 10, 64), ppc=4, order=3)
    g = phx.Level(3, 3, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te, capacity=wl.n_particles,
                  apply_tuning_table=False)
    g.set_tuning(segment_shape=[1, 1, 64], sort_interval=1)
    o = O.Sim(3, 3, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te)
    wl.setup(g, phx.F_B)
    wl.setup(o, O.F_B)
    g.advance(wl.dt, 3)
    o.advance(wl.dt, 3)
    g.sync()
    err = capfd.readouterr().err
    assert err.count("does not fit for segments of 1x1x64 cells") == 1, err
    for fid in range(6):
        U.assert_close(g.get(fid), o.get(fid), what=f"field {fid}")
    for m in range(5):
        U.assert_close(g.get(phx.f_pop(0, m)), o.get(O.f_pop(0, m)), what=f"moment {m}")
    U.assert_same_particles(g.get_particles(), o.get_particles())
    g.close()


def test_tuning_defaults_and_refusals():
    _say("[gpu] tuning defaults and refusals")
    g = phx.Level(2, 2, [40, 24], [0.4, 0.4], capacity=1000, apply_tuning_table=False)
    d = g.get_tuning()
    assert d["segment_shape"][:2] == [2, 32] and d["rows_per_item"] == 32
    assert d["tile_row_pad"] == 16 and d["sort_interval"] == 30 and d["tile_copies"] == 8
    for bad in (dict(segment_shape=[8, 16, 1]), dict(rows_per_item=0), dict(tile_row_pad=0),
                dict(sort_interval=-1), dict(tile_copies=3)):
        with pytest.raises(phx.PhxError) as e:
            g.set_tuning(**bad)
        assert e.value.status == 1, bad
    assert g.get_tuning() == d
    g.close()
