"""The oracle's coarse-to-fine particle split (oracle/refcpu_amr.hpp) pinned against the
reference's own Splitter (src/amr/data/particles/refine/split_{1,2,3}d.hpp, compiled in
oracle/_ref) bit for bit, and against what the mathematics fixes: toFineGrid doubles the position
exactly, the refined weights sum to the parent's weight times ratio^dim times the pattern's total
weight, the weighted mean position of the refined particles is the parent's (symmetric patterns),
and only refined particles inside the destination boxes are kept."""
import os

import numpy as np
import pytest

import oracle_lib as O

NEED_REF = pytest.mark.skipif(not os.path.exists(O.REF_SO),
                              reason="oracle/_ref not built (no PHARE source tree here)")

SPLITTERS = [(1, 1, 2), (1, 1, 3), (1, 2, 2), (1, 2, 3), (1, 2, 4), (1, 3, 2), (1, 3, 3), (1, 3, 4),
             (1, 3, 5), (2, 1, 4), (2, 2, 4), (2, 3, 4), (3, 1, 6), (3, 2, 6), (3, 3, 6)]


@NEED_REF
@pytest.mark.parametrize("dim,order,nbr", SPLITTERS)
def test_split_patterns_match_the_reference_splitter(dim, order, nbr):
    rng = np.random.default_rng(100 * dim + 10 * order + nbr)
    for _ in range(40):
        ic = 2 * rng.integers(-10, 10, dim)  # even fine cells (see oracle_lib.split_one)
        de = rng.random(dim)
        v = rng.standard_normal(3)
        w, q = rng.random() + 0.1, 1.0 + rng.integers(0, 2)
        a = O.split_one(dim, order, nbr, ic, de, v, w, q)
        b = O.split_one(dim, order, nbr, ic, de, v, w, q, use_ref=True)
        assert len(a["weight"]) == len(b["weight"]) == nbr
        for k in a:
            assert np.array_equal(a[k], b[k]), k


@NEED_REF
def test_split_box_growth_matches_the_reference():
    for order in (1, 2, 3):
        assert O.ref().ref_split_max_cell_distance(order) == int(np.ceil((order + 1) * 0.5))


def _position(p):
    return p["icell"] + p["delta"]


@pytest.mark.parametrize("dim,order,nbr", SPLITTERS)
def test_split_conserves_weight_and_centroid(dim, order, nbr):
    """every refined particle kept (one box over everything): the weights sum to the parent's
    times 2^dim times the pattern's total weight, and their weighted mean position is the parent's
    position on the fine grid (2 x the coarse position)"""
    rng = np.random.default_rng(7 + nbr)
    n = 50
    p = dict(icell=rng.integers(-5, 5, (n, dim)).astype(np.int32), delta=rng.random((n, dim)),
             v=rng.standard_normal((n, 3)), weight=rng.random(n) + 0.5, charge=np.ones(n))
    big = (tuple([-10 ** 6] * dim), tuple([10 ** 6] * dim))
    out = O.split(dim, order, nbr, p, [big])
    assert len(out["weight"]) == n * nbr
    total = {(1, 2): 1.0, (1, 3): 1.0, (1, 4): 1.0, (1, 5): 1.0, (2, 4): 1.0, (3, 6): 6 * np.float32(.166666)}
    pw = np.float64(total.get((dim, nbr), 1.0))
    w = out["weight"].reshape(n, nbr)
    # the pattern weights: 1D order 2 with 3 particles is 0.468137 + 2 x 0.265931 (not exactly 1)
    np.testing.assert_allclose(w.sum(axis=1) / p["weight"], w.sum(axis=1)[0] / p["weight"][0],
                               rtol=1e-12)
    assert abs(w.sum(axis=1)[0] / p["weight"][0] / 2 ** dim - pw) < 1e-5
    x = _position(out).reshape(n, nbr, dim)
    centroid = (x * w[:, :, None]).sum(axis=1) / w.sum(axis=1)[:, None]
    np.testing.assert_allclose(centroid, 2.0 * _position(p), atol=1e-6)
    # velocities and charges are the parent's
    assert np.array_equal(out["v"].reshape(n, nbr, 3), np.repeat(p["v"][:, None, :], nbr, axis=1))


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_split_keeps_only_refined_particles_in_the_boxes(dim):
    rng = np.random.default_rng(33 + dim)
    n, order = 400, 2
    nbr = {1: 2, 2: 4, 3: 6}[dim]
    p = dict(icell=rng.integers(0, 8, (n, dim)).astype(np.int32), delta=rng.random((n, dim)),
             v=rng.standard_normal((n, 3)), weight=np.ones(n), charge=np.ones(n))
    boxes = [(tuple([2] * dim), tuple([5] * dim)), (tuple([9] + [0] * (dim - 1)), tuple([11] + [15] * (dim - 1)))]
    out = O.split(dim, order, nbr, p, boxes)
    allp = O.split(dim, order, nbr, p, [(tuple([-10 ** 6] * dim), tuple([10 ** 6] * dim))])
    inside = np.zeros(len(allp["weight"]), bool)
    for lo, hi in boxes:
        inside |= np.all((allp["icell"] >= lo) & (allp["icell"] <= hi), axis=1)
    assert len(out["weight"]) == inside.sum() > 0
    for lo, hi in boxes:
        sel = np.all((out["icell"] >= lo) & (out["icell"] <= hi), axis=1)
        ref = np.all((allp["icell"] >= lo) & (allp["icell"] <= hi), axis=1)
        assert sel.sum() == ref.sum()
