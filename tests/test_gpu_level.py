"""A level of many patches of uneven sizes (phx_create_level): the batched form of SolverPPC's
per-patch loops and of the same-level fills between patches. A periodic domain cut into uneven
patches must reproduce the one-patch level: identical cells, fields, moments and particles within
1e-12 of the oracle's single patch; the patch views (a patch's arrays with their ghost nodes)
agree on the nodes neighbouring patches share (tests/simulator/test_advance.py:27-63 compares
overlaps at atol 5.5e-15); particles migrate to the patch that owns their cell."""
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


def _cut(ncells, splits):
    """patch boxes of a tensor cut of the level box at the given split points per direction"""
    import itertools

    edges = [[0] + list(s) + [n] for n, s in zip(ncells, splits)]
    boxes = []
    for idx in itertools.product(*[range(len(e) - 1) for e in edges]):
        lo = tuple(edges[d][i] for d, i in enumerate(idx))
        hi = tuple(edges[d][i + 1] - 1 for d, i in enumerate(idx))
        boxes.append((lo, hi))
    return boxes


HARRIS_CUT = ((8, 24, 40), (10, 20, 28))  # 48 x 40 cells -> 4 x 4 patches of 8..16 x 8..12 cells


def _level_pair(wl, patches, precision=phx.PRECISION_FP64, interval=3):
    g = phx.Level(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te,
                  capacity=wl.n_particles, precision=precision, patches=patches)
    o = O.Sim(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te)
    g.set_sort_interval(interval)
    wl.setup(g, phx.F_B)
    wl.setup(o, O.F_B)
    return g, o


@pytest.mark.parametrize("order", [1, 2, 3])
def test_uneven_patches_reproduce_the_one_patch_level(order):
    """16 patches of 8..16 x 8..12 cells over a periodic 48 x 40 Harris sheet, 12 PPC steps with a
    sort (and so a migration of the particles between patches) every 3 steps, against the oracle's
    one-patch level at 1e-12"""
    _say(f"[gpu] uneven patches order={order}")
    wl = W.harris2d(ncells=(48, 40), ppc=8, order=order)
    patches = _cut(wl.ncells, HARRIS_CUT)
    g, o = _level_pair(wl, patches)
    g.advance(wl.dt, 12)
    o.advance(wl.dt, 12)
    for fid in range(6):
        U.assert_close(g.get(fid), o.get(fid), what=f"field {fid}")
    for k in range(5):
        U.assert_close(g.get(phx.f_pop(0, k)), o.get(O.f_pop(0, k)), what=f"moment {k}")
    U.assert_close(g.get(phx.F_RHO), o.get(O.F_RHO), what="rho")
    U.assert_same_particles(g.get_particles(), o.get_particles())
    # every patch view: the oracle's one-patch arrays over the patch's nodes, ghosts included
    gw = [2, 4, 4][order - 1]
    for p, (lo, hi) in enumerate(patches):
        for fid in (phx.F_B, phx.F_E + 1, phx.f_pop(0, 0), phx.F_RHO):
            a = g.get_patch(p, fid)
            ref = o.get(fid)[tuple(slice(lo[d], lo[d] + a.shape[d]) for d in range(2))]
            U.assert_close(a, ref, what=f"patch {p} field {fid}")
    g.close()


def test_patch_views_agree_on_shared_nodes():
    """neighbouring patches hold the same value on the nodes they share (the primal nodes of a
    common face, and each other's ghost nodes), at the atol of tests/simulator/test_advance.py"""
    _say("[gpu] patch views agree on shared nodes")
    wl = W.harris2d(ncells=(48, 40), ppc=6, order=2)
    patches = _cut(wl.ncells, HARRIS_CUT)
    g, o = _level_pair(wl, patches)
    g.advance(wl.dt, 4)
    gw = 4
    for fid in (phx.F_B, phx.F_B + 1, phx.F_E, phx.f_pop(0, 2), phx.F_VI):
        views = [(g.get_patch(p, fid), lo) for p, (lo, hi) in enumerate(patches)]
        seen = {}
        for a, lo in views:
            for idx in np.ndindex(a.shape):
                key = (idx[0] + lo[0], idx[1] + lo[1])
                if key in seen:
                    assert abs(seen[key] - a[idx]) <= 5.5e-15, (fid, key)
                else:
                    seen[key] = a[idx]
    g.close()


def test_particles_belong_to_exactly_one_patch():
    """after a run with migrations every particle lies in exactly one patch's domain box, and
    the patches' domain arrays together are the level's particles"""
    _say("[gpu] particles partitioned by patch")
    wl = W.harris2d(ncells=(48, 40), ppc=6, order=2)
    patches = _cut(wl.ncells, HARRIS_CUT)
    g, o = _level_pair(wl, patches, interval=2)
    g.advance(wl.dt, 6)
    total = g.get_particles()
    parts = [g.get_patch_particles(p) for p in range(len(patches))]
    assert sum(len(q["weight"]) for q in parts) == len(total["weight"]) == wl.n_particles
    for p, (lo, hi) in enumerate(patches):
        c = parts[p]["icell"]
        assert np.all((c >= np.array(lo)) & (c <= np.array(hi))), p
    merged = {k: np.concatenate([q[k] for q in parts]) for k in total}
    U.assert_same_particles(merged, total, exact=True)
    g.close()


def _nlgb_reference(patches, p, pg, ncells):
    """makeNonLevelGhostBoxFor (amr_utils.hpp:233-251) restated: the domain box, then the particle
    ghost box intersected with every same-level neighbour, periodic images included"""
    import itertools

    lo, hi = np.array(patches[p][0]), np.array(patches[p][1])
    out = [(tuple(lo), tuple(hi))]
    for q, (qlo, qhi) in enumerate(patches):
        for sh in itertools.product((-1, 0, 1), repeat=len(ncells)):
            if q == p and not any(sh):
                continue
            s = np.array(sh) * np.array(ncells)
            a = np.maximum(lo - pg, np.array(qlo) + s)
            b = np.minimum(hi + pg, np.array(qhi) + s)
            if np.all(b >= a):
                out.append((tuple(a), tuple(b)))
    return out


def test_nonlevel_ghost_boxes_follow_neighbours():
    _say("[gpu] non-level-ghost boxes")
    patches = _cut([48, 40], HARRIS_CUT)
    g = phx.Level(2, 2, [48, 40], [0.4, 0.4], patches=patches)
    for p in (0, 5, 15):
        got = g.patch_nonlevel_ghost_boxes(p)
        ref = _nlgb_reference(patches, p, 2, [48, 40])
        assert sorted(got) == sorted(ref), p
        # their union is the patch's particle ghost box (a periodic level covered by patches)
        lo, hi = np.array(patches[p][0]) - 2, np.array(patches[p][1]) + 2
        cover = np.zeros(tuple(hi - lo + 1), bool)
        for a, b in got:
            cover[tuple(slice(a[d] - lo[d], b[d] - lo[d] + 1) for d in range(2))] = True
        assert cover.all(), p
    g.close()


def test_patch_fields_set_and_get():
    """a field set patch by patch (physical nodes) is the level's field; each patch view's
    ghost nodes after the ghost fill hold the neighbour's (or periodic image's) values"""
    _say("[gpu] patch field set/get")
    patches = _cut([48, 40], HARRIS_CUT)
    g = phx.Level(2, 2, [48, 40], [0.4, 0.4], patches=patches)
    rng = np.random.default_rng(8)
    full = rng.standard_normal(g.shape(phx.F_B))
    for p, (lo, hi) in enumerate(patches):
        shp = g.patch_shape(p, phx.F_B)
        g.set_patch(p, phx.F_B, full[lo[0]:lo[0] + shp[0], lo[1]:lo[1] + shp[1]])
    got = g.get(phx.F_B)
    gw = 4
    phys = (slice(gw, gw + 49), slice(gw, gw + 40))  # Bx: primal x, dual y physical nodes
    assert np.array_equal(got[phys], full[phys])
    g.fill_ghosts(phx.F_B)
    lvl = g.get(phx.F_B)
    for p, (lo, hi) in enumerate(patches):
        a = g.get_patch(p, phx.F_B)
        assert np.array_equal(a, lvl[lo[0]:lo[0] + a.shape[0], lo[1]:lo[1] + a.shape[1]])
    g.close()


def test_uneven_patches_mixed_precision():
    """the mixed-precision family on the uneven patch level: moments and fields within 1e-5 of
    the FP64 oracle after 2 PPC steps (as on one patch)"""
    _say("[gpu] uneven patches, mixed precision")
    wl = W.harris2d(ncells=(48, 40), ppc=8, order=2)
    patches = _cut(wl.ncells, HARRIS_CUT)
    g, o = _level_pair(wl, patches, precision=phx.PRECISION_FP32)
    g.advance(wl.dt, 2)
    o.advance(wl.dt, 2)
    for fid in range(6):
        U.assert_close(g.get(fid), o.get(fid), 1e-5, what=f"field {fid}")
    for k in range(5):
        U.assert_close(g.get(phx.f_pop(0, k)), o.get(O.f_pop(0, k)), 1e-5, what=f"moment {k}")
    g.close()


def test_one_dimensional_patches():
    """a 1D periodic level of five patches of 8..70 cells: the full 1D workload length (100 PPC
    steps) against the oracle's one patch"""
    _say("[gpu] 1D patches")
    wl = W.uniform1d(ncells=200, ppc=40)
    patches = _cut(wl.ncells, ((8, 30, 100, 130),))
    g, o = _level_pair(wl, patches, interval=10)
    g.advance(wl.dt, 100)
    o.advance(wl.dt, 100)
    for fid in range(6):
        U.assert_close(g.get(fid), o.get(fid), what=f"field {fid}")
    U.assert_same_particles(g.get_particles(), o.get_particles())
    g.close()
