"""Coarse-to-fine field refinement on the GPU against the oracle (ratio 2), bit-exact:
phx_refine_field with each operator (the reference's DefaultFieldRefiner, FieldMomentsRefiner,
ElectricFieldRefiner, MagneticFieldRefiner, MagneticFieldInitRefiner) on fine arrays holding a mix
of NaN and set nodes, and phx_refine_magnetic_new_faces (MagneticRefinePatchStrategy) on random
B components, against oracle_lib.refine_field / magnetic_new_faces (pinned in
test_oracle_refine.py). The fine level sits on coarse cells 3.. (fine AMR lower 6) with its ghost
nodes reaching below the coarse level's origin, so negative coarse indices are read."""
import sys

import numpy as np
import pytest

import oracle_lib as O
from phare_amd import phx

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _levels(dim, order):
    nc = [[30], [16, 12], [12, 11, 10]][dim - 1]
    dx = [0.4, 0.3, 0.25][:dim]
    coarse = phx.Level(dim, order, nc, dx, apply_tuning_table=False)
    clo = [3] * dim
    chi = [n - 4 for n in nc]
    flo = [2 * a for a in clo]
    fn = [2 * (b - a + 1) for a, b in zip(clo, chi)]
    fine = phx.Level(dim, order, fn, [h / 2 for h in dx], lower=flo, periodic=False,
                     apply_tuning_table=False)
    return coarse, fine, flo, fn, [h / 2 for h in dx]


def _primal(fid, dim):
    """centering flags of a field id's quantity (Yee hybrid layout)"""
    if fid in (phx.F_E, phx.F_E + 1, phx.F_E + 2, phx.F_J, phx.F_J + 1, phx.F_J + 2):
        comp = fid - (phx.F_E if fid < phx.F_J else phx.F_J)
        return [d != comp for d in range(dim)]
    if phx.F_B <= fid <= phx.F_B + 2 or phx.F_BOLD <= fid <= phx.F_BOLD + 2:
        comp = fid - (phx.F_B if fid < phx.F_BOLD else phx.F_BOLD)
        return [d == comp for d in range(dim)]
    return [True] * dim


def _cases():
    lin = [phx.F_RHO, phx.F_VI + 1, phx.f_pop(0, 0), phx.F_E, phx.F_B + 1]
    out = [(f, phx.REFINE_LINEAR) for f in lin] + [(f, phx.REFINE_LINEAR_OVERWRITE) for f in lin]
    out += [(phx.F_E + k, phx.REFINE_ELECTRIC) for k in range(3)]
    out += [(phx.F_J + k, phx.REFINE_ELECTRIC) for k in range(3)]
    out += [(phx.F_B + k, phx.REFINE_MAGNETIC) for k in range(3)]
    out += [(phx.F_BOLD + k, phx.REFINE_MAGNETIC_INIT) for k in range(3)]
    return out


@pytest.mark.parametrize("dim,order", [(1, 1), (1, 3), (2, 1), (2, 2), (3, 1), (3, 3)])
def test_refine_field_matches_oracle(dim, order):
    _say(f"[gpu] refine_field dim={dim} order={order}")
    rng = np.random.default_rng(100 * dim + order)
    coarse, fine, flo, fn, hf = _levels(dim, order)
    g = [2, 4, 4][order - 1]
    forg = [a - g for a in flo]
    corg = [-g] * dim
    for fid, op in _cases():
        fine_a = rng.standard_normal(fine.shape(fid))
        fine_a[rng.random(fine_a.shape) < 0.4] = np.nan
        coarse_a = rng.standard_normal(coarse.shape(fid))
        fine.set(fid, fine_a)
        coarse.set(fid, coarse_a)
        # every fine node of the array
        lo = list(forg)
        hi = [forg[d] + fine_a.shape[d] - 1 for d in range(dim)]
        fine.refine_field(fid, coarse, fid, op, lo, hi)
        want = O.refine_field(op, _primal(fid, dim), fine_a, forg, coarse_a, corg, lo, hi)
        np.testing.assert_array_equal(fine.get(fid), want, err_msg=f"field {fid} op {op}")
        # an inner box: nodes outside it keep their values
        fine.set(fid, fine_a)
        lo2 = [a + 3 for a in lo]
        hi2 = [b - 2 for b in hi]
        fine.refine_field(fid, coarse, fid, op, lo2, hi2)
        want = O.refine_field(op, _primal(fid, dim), fine_a, forg, coarse_a, corg, lo2, hi2)
        np.testing.assert_array_equal(fine.get(fid), want, err_msg=f"field {fid} op {op} inner")
    # refused: a box beyond the fine array, an operator on the wrong quantity, mixed quantities
    with pytest.raises(phx.PhxError):
        fine.refine_field(phx.F_RHO, coarse, phx.F_RHO, phx.REFINE_LINEAR, [a - 1 for a in forg],
                          [a + 2 for a in forg])
    with pytest.raises(phx.PhxError):
        fine.refine_field(phx.F_RHO, coarse, phx.F_RHO, phx.REFINE_ELECTRIC, forg, forg)
    with pytest.raises(phx.PhxError):
        fine.refine_field(phx.F_E, coarse, phx.F_E, phx.REFINE_MAGNETIC, forg, forg)
    with pytest.raises(phx.PhxError):
        fine.refine_field(phx.F_E, coarse, phx.F_B, phx.REFINE_LINEAR, forg, forg)
    coarse.close()
    fine.close()


@pytest.mark.parametrize("dim,order", [(1, 2), (2, 1), (2, 3), (3, 1), (3, 2)])
def test_magnetic_new_faces_match_oracle(dim, order):
    _say(f"[gpu] magnetic new faces dim={dim} order={order}")
    rng = np.random.default_rng(300 + 10 * dim + order)
    coarse, fine, flo, fn, hf = _levels(dim, order)
    g = [2, 4, 4][order - 1]
    forg = [a - g for a in flo]
    for base in (phx.F_B, phx.F_BOLD):
        B = [rng.standard_normal(fine.shape(base + k)) for k in range(3)]
        for k in range(3):
            fine.set(base + k, B[k])
        # the widest cell box whose reconstruction stencil stays inside the arrays
        lo = [a + 1 for a in forg]
        hi = [flo[d] + fn[d] + g - 3 for d in range(dim)]
        fine.refine_magnetic_new_faces(base, lo, hi)
        want = O.magnetic_new_faces(B, forg, hf, lo, hi)
        for k in range(3):
            np.testing.assert_array_equal(fine.get(base + k), want[k], err_msg=f"B{k} base {base}")
        # an odd-aligned inner box
        for k in range(3):
            fine.set(base + k, B[k])
        lo2 = [a + 3 for a in lo]
        hi2 = [b - 4 for b in hi]
        fine.refine_magnetic_new_faces(base, lo2, hi2)
        want = O.magnetic_new_faces(B, forg, hf, lo2, hi2)
        for k in range(3):
            np.testing.assert_array_equal(fine.get(base + k), want[k], err_msg=f"B{k} inner")
        # every cell of the arrays: the faces at their edge take the edge rule
        for k in range(3):
            fine.set(base + k, B[k])
        lo3 = list(forg)
        hi3 = [flo[d] + fn[d] + g - 1 for d in range(dim)]
        fine.refine_magnetic_new_faces(base, lo3, hi3)
        want = O.magnetic_new_faces(B, forg, hf, lo3, hi3)
        for k in range(3):
            np.testing.assert_array_equal(fine.get(base + k), want[k], err_msg=f"B{k} edge")
    with pytest.raises(phx.PhxError):
        fine.refine_magnetic_new_faces(phx.F_B, [a - 1 for a in forg], [a + 4 for a in forg])
    with pytest.raises(phx.PhxError):
        fine.refine_magnetic_new_faces(phx.F_E, [a + 1 for a in forg], [a + 4 for a in forg])
    coarse.close()
    fine.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_refined_B_is_divergence_free(dim):
    """the whole magnetic refinement of a new fine level on the GPU: the coarse B the discrete
    curl of a random vector potential, the fine shared faces by the init refiner, the new faces by
    the patch strategy; the fine divergence vanishes in the refined cells"""
    _say(f"[gpu] refined B divergence dim={dim}")
    rng = np.random.default_rng(500 + dim)
    coarse, fine, flo, fn, hf = _levels(dim, 1)
    g = 2
    hc = [2 * h for h in hf]
    shp = [coarse.shape(phx.F_B + k) for k in range(3)]
    big = [max(s[d] for s in shp) + 1 for d in range(dim)] + [1] * (3 - dim)
    A = [rng.integers(-8, 8, size=big) / 4.0 for _ in range(3)]

    def diff(a, d, s):
        lo = [slice(0, s[k]) for k in range(3)]
        hi = list(lo)
        hi[d] = slice(1, s[d] + 1)
        return (a[tuple(hi)] - a[tuple(lo)]) / hc[d]

    for c in range(3):
        s = list(shp[c]) + [1] * (3 - dim)
        a1, a2 = (c + 1) % 3, (c + 2) % 3
        v = np.zeros(s)
        if a1 < dim:
            v = v + diff(A[a2], a1, s)
        if a2 < dim:
            v = v - diff(A[a1], a2, s)
        coarse.set(phx.F_B + c, v.reshape(shp[c]))
    forg = [a - g for a in flo]
    clo = [a + 2 for a in flo]
    chi = [flo[d] + fn[d] - 3 for d in range(dim)]
    for c in range(3):
        fine.set(phx.F_B + c, np.full(fine.shape(phx.F_B + c), np.nan))
        hi = [chi[d] + (1 if d == c else 0) for d in range(dim)]
        fine.refine_field(phx.F_B + c, coarse, phx.F_B + c, phx.REFINE_MAGNETIC_INIT, clo, hi)
    fine.refine_magnetic_new_faces(phx.F_B, clo, chi)
    Bf = [fine.get(phx.F_B + c) for c in range(3)]
    lo_loc = [clo[d] - forg[d] for d in range(dim)]
    hi_loc = [chi[d] - forg[d] for d in range(dim)]
    div = 0.0
    for d in range(dim):
        a = [slice(lo_loc[k], hi_loc[k] + 1) for k in range(dim)]
        b = list(a)
        b[d] = slice(lo_loc[d] + 1, hi_loc[d] + 2)
        div = div + (Bf[d][tuple(b)] - Bf[d][tuple(a)]) / hf[d]
    assert not np.any(np.isnan(div))
    scale = max(np.nanmax(np.abs(x)) for x in Bf) / min(hf)
    assert np.max(np.abs(div)) <= 1e-13 * scale
    coarse.close()
    fine.close()


@pytest.mark.parametrize("dim,order", [(2, 1), (3, 2)])
def test_magnetic_new_faces_edge_rule_matches_oracle(dim, order):
    """a patch starting at an odd AMR cell: over every cell of its arrays the faces at the edge
    take the edge rule (neighbour mean, or left), the others the full reconstruction, as the
    oracle does"""
    _say(f"[gpu] magnetic new faces edge rule dim={dim} order={order}")
    rng = np.random.default_rng(700 + dim)
    fn = [[14], [14, 10], [10, 8, 6]][dim - 1]
    flo = [5] * dim
    hf = [0.2, 0.15, 0.125][:dim]
    fine = phx.Level(dim, order, fn, hf, lower=flo, periodic=False, apply_tuning_table=False)
    g = [2, 4, 4][order - 1]
    forg = [a - g for a in flo]
    B = [rng.standard_normal(fine.shape(phx.F_B + k)) for k in range(3)]
    for k in range(3):
        fine.set(phx.F_B + k, B[k])
    lo = list(forg)
    hi = [flo[d] + fn[d] + g - 1 for d in range(dim)]
    fine.refine_magnetic_new_faces(phx.F_B, lo, hi)
    want = O.magnetic_new_faces(B, forg, hf, lo, hi)
    for k in range(3):
        np.testing.assert_array_equal(fine.get(phx.F_B + k), want[k], err_msg=f"B{k}")
    fine.close()
