"""The oracle's coarse-to-fine field refinement (oracle/refcpu_amr.hpp through oracle_lib: the
reference's DefaultFieldRefiner, FieldMomentsRefiner, ElectricFieldRefiner, MagneticFieldRefiner,
MagneticFieldInitRefiner and MagneticRefinePatchStrategy, ratio 2) pinned by what the mathematics
fixes, on AMR index ranges that cross zero (negative fine and coarse indices):
 - linear refiners: a field linear in space on the coarse nodes is reproduced exactly on the fine
   nodes, primal and dual, in 1D/2D/3D (a wrong start node, a swapped weight pair or a wrong
   parity fails); NaN-only vs overwrite semantics;
 - electric refiner: along the component's dual direction both fine edges take the coarse edge
   they halve, along its primal directions the field is linear -> a field piecewise constant along
   the dual direction and linear along the primal ones is reproduced exactly;
 - magnetic refiners: the shared fine faces take the coarse face they lie on (flux conservation),
   the new faces are left alone;
 - new faces: a divergence-free coarse B (the discrete curl of a random vector potential) refines
   to a fine B whose discrete divergence vanishes in every fine cell of the box, for equal and
   unequal mesh sizes; with equal mesh sizes the same holds for shared faces taken from a
   divergence-free fine field (not copies), which fails when the cross terms or the third-order
   terms are dropped or flip sign; a uniform B stays uniform; in 1D the new face is the mean of
   its neighbours.
Values are dyadic rationals where the check is exact, so the comparisons are exact there."""
import itertools

import numpy as np
import pytest

import oracle_lib as O

G = 2  # ghost nodes of the test arrays


def _shape(n, primal):
    return [n[d] + 2 * G + (1 if primal[d] else 0) for d in range(len(n))]


def _coords(n, lower, primal, h):
    """AMR-position of each local node, per direction"""
    return [(np.arange(_shape(n, primal)[d]) - G + lower[d] + (0.0 if primal[d] else 0.5)) * h
            for d in range(len(n))]


def _index(n, lower, primal):
    return [np.arange(_shape(n, primal)[d]) - G + lower[d] for d in range(len(n))]


CENTERINGS = {1: ["p", "d"], 2: ["pp", "dp", "pd", "dd"], 3: ["ppp", "dpp", "pdp", "ppd", "ddp", "pdd"]}


def _levels(dim):
    # coarse cells [-3, 8), fine cells [-4, 10) (ratio 2, both reach negative AMR indices)
    cl = [-3, -2, -3][:dim]
    cn = [11, 9, 8][:dim]
    fl = [-4, -2, -6][:dim]
    fn = [14, 10, 12][:dim]
    return cl, cn, fl, fn


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_linear_refiners_reproduce_linear_fields(dim):
    cl, cn, fl, fn = _levels(dim)
    coef = [0.75, -1.25, 0.5][:dim]
    for cen in CENTERINGS[dim]:
        primal = [c == "p" for c in cen]
        Xc = np.meshgrid(*_coords(cn, cl, primal, 1.0), indexing="ij")
        Xf = np.meshgrid(*_coords(fn, fl, primal, 0.5), indexing="ij")
        coarse = 0.5 + sum(c * x for c, x in zip(coef, Xc))
        want = 0.5 + sum(c * x for c, x in zip(coef, Xf))
        # the fine box: every fine node whose two coarse neighbours are in the coarse array
        lo = [fl[d] - 1 for d in range(dim)]
        hi = [fl[d] + fn[d] - (0 if primal[d] else 1) + 1 for d in range(dim)]
        sl = tuple(slice(lo[d] - fl[d] + G, hi[d] - fl[d] + G + 1) for d in range(dim))
        nan = np.full(_shape(fn, primal), np.nan)
        out = O.refine_field(O.REFINE_LINEAR, primal, nan, [a - G for a in fl], coarse,
                             [a - G for a in cl], lo, hi)
        np.testing.assert_array_equal(out[sl], want[sl], err_msg=cen)
        mask = np.ones(out.shape, bool)
        mask[sl] = False
        assert np.all(np.isnan(out[mask]))
        # NaN-only: set nodes keep their value; the overwrite refiner replaces them
        seven = np.full(_shape(fn, primal), 7.0)
        kept = O.refine_field(O.REFINE_LINEAR, primal, seven, [a - G for a in fl], coarse,
                              [a - G for a in cl], lo, hi)
        assert np.all(kept == 7.0)
        over = O.refine_field(O.REFINE_LINEAR_OVERWRITE, primal, seven, [a - G for a in fl], coarse,
                              [a - G for a in cl], lo, hi)
        np.testing.assert_array_equal(over[sl], want[sl])


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_electric_refiner(dim):
    rng = np.random.default_rng(7 + dim)
    cl, cn, fl, fn = _levels(dim)
    for comp in range(3):
        primal = [d != comp for d in range(dim)]
        Ic = np.meshgrid(*_index(cn, cl, primal), indexing="ij")
        If = np.meshgrid(*_index(fn, fl, primal), indexing="ij")
        # a dyadic value per coarse index along each dual direction, linear along the primal ones
        # (1D: every direction is a copy of the coarse value the fine node lies in or on)
        tables = [rng.integers(-16, 16, size=64) / 8.0 for _ in range(dim)]
        slope = [0.25, -0.75, 1.5]

        def value(idx, h, coarse_index):
            v = 0.125
            for d in range(dim):
                if primal[d] and dim > 1:
                    v = v + slope[d] * idx[d] * h
                else:
                    v = v + tables[d][coarse_index(idx[d]) + 32]
            return v

        coarse = value(Ic, 1.0, lambda i: i)
        want = value(If, 0.5, lambda i: np.floor_divide(i, 2))
        lo = [fl[d] - 1 for d in range(dim)]
        hi = [fl[d] + fn[d] - (0 if primal[d] else 1) + 1 for d in range(dim)]
        sl = tuple(slice(lo[d] - fl[d] + G, hi[d] - fl[d] + G + 1) for d in range(dim))
        out = O.refine_field(O.REFINE_ELECTRIC, primal, np.full(_shape(fn, primal), np.nan),
                             [a - G for a in fl], coarse, [a - G for a in cl], lo, hi)
        np.testing.assert_array_equal(out[sl], want[sl], err_msg=f"E{comp}")
        seven = np.full(_shape(fn, primal), 7.0)
        assert np.all(O.refine_field(O.REFINE_ELECTRIC, primal, seven, [a - G for a in fl], coarse,
                                     [a - G for a in cl], lo, hi) == 7.0)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_magnetic_shared_faces(dim):
    rng = np.random.default_rng(20 + dim)
    cl, cn, fl, fn = _levels(dim)
    for comp in range(3):
        primal = [d == comp for d in range(dim)]
        coarse = rng.standard_normal(_shape(cn, primal))
        If = _index(fn, fl, primal)
        lo = [fl[d] for d in range(dim)]
        hi = [fl[d] + fn[d] - (0 if primal[d] else 1) for d in range(dim)]
        for op in (O.REFINE_MAGNETIC, O.REFINE_MAGNETIC_INIT):
            start = np.full(_shape(fn, primal), np.nan if op == O.REFINE_MAGNETIC else 3.0)
            out = O.refine_field(op, primal, start, [a - G for a in fl], coarse, [a - G for a in cl],
                                 lo, hi)
            for loc in itertools.product(*[range(s) for s in out.shape]):
                f = [If[d][loc[d]] for d in range(dim)]
                inside = all(lo[d] <= f[d] <= hi[d] for d in range(dim))
                new_face = comp < dim and f[comp] % 2 != 0
                if not inside or new_face:
                    assert out[loc] == start[loc] or (np.isnan(out[loc]) and np.isnan(start[loc])), (comp, f)
                    continue
                c = tuple(f[d] // 2 - (cl[d] - G) for d in range(dim))
                assert out[loc] == coarse[c], (comp, f)
        # NaN-only: a set shared face keeps its value
        seven = np.full(_shape(fn, primal), 7.0)
        assert np.all(O.refine_field(O.REFINE_MAGNETIC, primal, seven, [a - G for a in fl], coarse,
                                     [a - G for a in cl], lo, hi) == 7.0)


def _curl_B(dim, n, rng, h):
    """a discretely divergence-free B on arrays of n cells (+G ghosts): the Yee curl of a random
    dyadic vector potential (A on the edges, B on the faces)"""
    shp = [n[d] + 2 * G + 2 for d in range(dim)] + [1] * (3 - dim)
    A = [rng.integers(-8, 8, size=shp) / 4.0 for _ in range(3)]
    Bs = [_shape(n, [d == c for d in range(dim)]) for c in range(3)]

    def D(a, d, s):
        # forward difference along d over the first s nodes
        lo = [slice(0, s[k]) for k in range(3)]
        hi = list(lo)
        hi[d] = slice(1, s[d] + 1)
        return (a[tuple(hi)] - a[tuple(lo)]) / (h[d] if d < dim else 1.0)

    out = []
    for c in range(3):
        s = Bs[c] + [1] * (3 - dim)
        a1, a2 = (c + 1) % 3, (c + 2) % 3
        # B_c = d_{a1} A_{a2} - d_{a2} A_{a1} (directions beyond dim have no derivative)
        v = np.zeros(s)
        if a1 < dim:
            v = v + D(A[a2], a1, s)
        if a2 < dim:
            v = v - D(A[a1], a2, s)
        out.append(v.reshape(Bs[c]))
    return out


def _fine_div(B, lo_loc, hi_loc, h, dim):
    """discrete divergence of B in the fine cells lo_loc..hi_loc (local cell indices)"""
    div = 0.0
    for d in range(dim):
        a = [slice(lo_loc[k], hi_loc[k] + 1) for k in range(dim)]
        b = list(a)
        b[d] = slice(lo_loc[d] + 1, hi_loc[d] + 2)
        div = div + (B[d][tuple(b)] - B[d][tuple(a)]) / h[d]
    return div


def _rebuilt_divergence(dim, h, oracle=None):
    """max |div B| of the fine cells after the new faces of a divergence-free fine field (equal or
    unequal mesh sizes h) were erased and rebuilt, and the field's scale"""
    rng = np.random.default_rng(31 + dim)
    cl, cn, fl, fn = _levels(dim)
    hf = [x / 2 for x in h]
    clo = [fl[d] + 2 for d in range(dim)]
    chi = [fl[d] + fn[d] - 3 for d in range(dim)]
    lo_loc = [clo[d] - fl[d] + G for d in range(dim)]
    hi_loc = [chi[d] - fl[d] + G for d in range(dim)]
    Bfine = _curl_B(dim, fn, rng, hf)
    erased = [b.copy() for b in Bfine]
    for c in range(dim):
        for loc in itertools.product(*[range(s) for s in erased[c].shape]):
            inside = all(lo_loc[d] <= loc[d] <= hi_loc[d] + (1 if d == c else 0) for d in range(dim))
            if inside and (loc[c] - G + fl[c]) % 2:
                erased[c][loc] = np.nan
    rebuilt = O.magnetic_new_faces(erased, [a - G for a in fl], hf, clo, chi, oracle=oracle)
    div = _fine_div(rebuilt, lo_loc, hi_loc, hf, dim)
    return np.max(np.abs(div)), max(np.max(np.abs(b)) for b in Bfine) / min(hf)


@pytest.mark.parametrize("plant", [{"ORA_PLANT_IJK": "1.,-1."}, {"ORA_PLANT_IJK": "0.,0."},
                                   {"ORA_PLANT_CROSS": "0."}, {"ORA_PLANT_CROSS": "-1."}])
def test_new_faces_check_catches_planted_mistakes(plant, tmp_path):
    """the divergence check above fails for an oracle whose new-face reconstruction has the
    third-order terms of the wrong sign or none, or the first-order cross terms dropped or
    flipped (3D; 2D for the cross terms)"""
    L = O.planted_oracle(plant, tmp_path)
    dims = [3] if "ORA_PLANT_IJK" in plant else [2, 3]
    for dim in dims:
        h = (1.0,) * dim
        good, scale = _rebuilt_divergence(dim, h)
        bad, _ = _rebuilt_divergence(dim, h, oracle=L)
        assert good <= 1e-13 * scale
        assert bad > 1e-3 * scale, (plant, dim, bad)


@pytest.mark.parametrize("dim,h", [(2, (1.0, 1.0)), (2, (1.0, 0.5)), (3, (1.0, 1.0, 1.0)),
                                   (3, (1.0, 0.75, 0.5))])
def test_new_faces_keep_B_divergence_free(dim, h):
    rng = np.random.default_rng(31 + dim)
    cl, cn, fl, fn = _levels(dim)
    hc = list(h)
    hf = [x / 2 for x in h]
    Bc = _curl_B(dim, cn, rng, hc)
    # the coarse field is divergence-free
    cdiv = _fine_div(Bc, [0] * dim, [cn[d] + 2 * G - 1 for d in range(dim)], hc, dim)
    assert np.max(np.abs(cdiv)) < 1e-12
    Bf = []
    # the fine cells refined: a box of even extent inside both arrays
    clo = [fl[d] + 2 for d in range(dim)]
    chi = [fl[d] + fn[d] - 3 for d in range(dim)]
    for c in range(3):
        primal = [d == c for d in range(dim)]
        lo = list(clo)
        hi = [chi[d] + (1 if primal[d] else 0) for d in range(dim)]
        Bf.append(O.refine_field(O.REFINE_MAGNETIC_INIT, primal, np.full(_shape(fn, primal), np.nan),
                                 [a - G for a in fl], Bc[c], [a - G for a in cl], lo, hi))
    Bn = O.magnetic_new_faces(Bf, [a - G for a in fl], hf, clo, chi)
    lo_loc = [clo[d] - fl[d] + G for d in range(dim)]
    hi_loc = [chi[d] - fl[d] + G for d in range(dim)]
    div = _fine_div(Bn, lo_loc, hi_loc, hf, dim)
    scale = max(np.max(np.abs(b)) for b in Bc) / min(hf)
    assert np.max(np.abs(div)) <= 1e-13 * scale, np.max(np.abs(div))
    # every face of those cells is set
    for c in range(dim):
        sl = tuple(slice(lo_loc[d], hi_loc[d] + 1 + (1 if d == c else 0)) for d in range(dim))
        assert not np.any(np.isnan(Bn[c][sl]))
    # shared faces that are not copies of one coarse face (a fine divergence-free field: the
    # faces of each coarse cell balance, sub-face by sub-face values differ): with equal mesh
    # sizes the new faces rebuilt from them close the divergence of every fine cell, which the
    # mean of the two faces along the normal alone (the cross terms dropped) does not, nor do
    # third-order terms of the wrong sign or none (the reference's weights of the cross terms do
    # not depend on the mesh sizes, so for unequal ones this holds only for copied faces, above)
    if len(set(h)) > 1:
        return
    Bfine = _curl_B(dim, fn, rng, hf)
    erased = [b.copy() for b in Bfine]
    for c in range(dim):
        for loc in itertools.product(*[range(s) for s in erased[c].shape]):
            inside = all(lo_loc[d] <= loc[d] <= hi_loc[d] + (1 if d == c else 0) for d in range(dim))
            if inside and (loc[c] - G + fl[c]) % 2:
                erased[c][loc] = np.nan
    rebuilt = O.magnetic_new_faces(erased, [a - G for a in fl], hf, clo, chi)
    div = _fine_div(rebuilt, lo_loc, hi_loc, hf, dim)
    scale = max(np.max(np.abs(b)) for b in Bfine) / min(hf)
    assert np.max(np.abs(div)) <= 1e-13 * scale, np.max(np.abs(div))
    naive = [b.copy() for b in erased]
    for c in range(dim):
        for loc in itertools.product(*[range(s) for s in naive[c].shape]):
            if np.isnan(naive[c][loc]):
                m = list(loc)
                p = list(loc)
                m[c] -= 1
                p[c] += 1
                naive[c][loc] = 0.5 * (naive[c][tuple(m)] + naive[c][tuple(p)])
    assert np.max(np.abs(_fine_div(naive, lo_loc, hi_loc, hf, dim))) > 1e-3 * scale


def test_new_faces_uniform_and_1d():
    # a uniform B stays uniform (3D, unequal mesh): every cross term cancels
    dim = 3
    cl, cn, fl, fn = _levels(dim)
    val = [1.5, -0.25, 0.75]
    Bf = [np.full(_shape(fn, [d == c for d in range(dim)]), val[c]) for c in range(3)]
    clo = [fl[d] + 2 for d in range(dim)]
    chi = [fl[d] + fn[d] - 3 for d in range(dim)]
    Bn = O.magnetic_new_faces(Bf, [a - G for a in fl], [0.5, 0.25, 0.4], clo, chi)
    for c in range(3):
        assert np.all(Bn[c] == val[c])
    # 1D: the new Bx faces (odd index) are the mean of their neighbours; By, Bz untouched
    rng = np.random.default_rng(3)
    fl1, fn1 = [-4], [14]
    B1 = [rng.integers(-8, 8, size=_shape(fn1, [c == 0])) / 4.0 for c in range(3)]
    out = O.magnetic_new_faces(B1, [fl1[0] - G], [0.5], [fl1[0] + 1], [fl1[0] + fn1[0] - 2])
    for loc in range(B1[0].shape[0]):
        f = loc - G + fl1[0]
        if fl1[0] + 1 <= f <= fl1[0] + fn1[0] - 1 and f % 2:
            assert out[0][loc] == 0.5 * (B1[0][loc - 1] + B1[0][loc + 1])
        else:
            assert out[0][loc] == B1[0][loc]
    assert np.array_equal(out[1], B1[1]) and np.array_equal(out[2], B1[2])


def test_new_faces_edge_rule():
    """at the edge of the arrays a new face whose cross terms would read beyond them takes the
    mean of its two neighbours along the normal, and one without both neighbours is left (a box
    aligned to even indices never needs the rule; here the arrays start at odd AMR indices): a
    uniform B stays uniform, the first faces (no lower neighbour) keep their values, some faces
    take the neighbour mean and the others the full reconstruction"""
    dim = 2
    fl, fn = [-3, -1], [14, 10]
    org = [a - G for a in fl]
    lo = list(org)
    hi = [fl[d] + fn[d] + G - 1 for d in range(dim)]
    val = [1.5, -0.25, 0.75]
    Bf = [np.full(_shape(fn, [d == c for d in range(dim)]), val[c]) for c in range(3)]
    Bn = O.magnetic_new_faces(Bf, org, [0.5, 0.5], lo, hi)
    for c in range(3):
        assert np.all(Bn[c] == val[c])
    rng = np.random.default_rng(5)
    Br = [rng.standard_normal(_shape(fn, [d == c for d in range(dim)])) for c in range(3)]
    Bn = O.magnetic_new_faces(Br, org, [0.5, 0.5], lo, hi)
    # Bx at local x 0 (AMR -5, odd: a new face without a lower neighbour) is left
    assert np.array_equal(Bn[0][0, :], Br[0][0, :])
    mean = 0.5 * (Br[0][:-2, :] + Br[0][2:, :])
    new = Bn[0][1:-1, :]
    odd_rows = [i for i in range(1, Br[0].shape[0] - 1) if (i + org[0]) % 2 != 0]
    took_mean = [np.any(new[i - 1] == mean[i - 1]) for i in odd_rows]
    full = [np.any(new[i - 1] != mean[i - 1]) for i in odd_rows]
    assert any(took_mean) and any(full)
