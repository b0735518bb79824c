"""The oracle's fine-to-coarse coarsening (oracle_lib.coarsen: MomentsCoarsener and
ElectricFieldCoarsener of the reference, ratio 2) pinned by what the mathematics fixes: a field
linear in space on the fine grid coarsens to the same linear field sampled at the coarse nodes
(primal nodes coincide with fine node 2c; a dual node is the midpoint of fine dual nodes 2c and
2c + 1), so an injection from the wrong fine node or an average along the wrong direction fails;
nodes outside the box keep their values."""
import numpy as np
import pytest

import oracle_lib as O


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("centering", ["ppp", "dpp", "pdp", "ppd"])
def test_coarsen_linear_fields_exactly(dim, centering):
    if "d" in centering[dim:]:
        pytest.skip("no dual direction beyond the dimension")
    g = 2
    fine_lower = [6, 4, 2][:dim]
    coarse_lower = [0] * dim
    nf = [20, 16, 12][:dim]
    nc = [16, 12, 10][:dim]
    dual = [centering[d] == "d" for d in range(dim)]
    dual_dir = dual.index(True) if any(dual) else -1
    coef = [0.7, -1.3, 2.1][:dim]

    def field(n, lower, h):
        shape = [n[d] + 1 + 2 * g - (1 if dual[d] else 0) for d in range(dim)]
        axes = [(np.arange(shape[d]) - g + lower[d] + (0.5 if dual[d] else 0.0)) * h for d in range(dim)]
        X = np.meshgrid(*axes, indexing="ij")
        return 0.25 + sum(c * x for c, x in zip(coef, X))

    fine = field(nf, fine_lower, 0.5)
    coarse = field(nc, coarse_lower, 1.0)
    poison = np.full_like(coarse, 7.0)
    # the coarse nodes under the fine patch's physical region
    lo = [fine_lower[d] // 2 for d in range(dim)]
    hi = [(fine_lower[d] + nf[d]) // 2 - (1 if dual[d] else 0) for d in range(dim)]
    out = O.coarsen(fine, [a - g for a in fine_lower], poison, [a - g for a in coarse_lower], lo, hi,
                    dual_dir)
    sl = tuple(slice(lo[d] - coarse_lower[d] + g, hi[d] - coarse_lower[d] + g + 1) for d in range(dim))
    np.testing.assert_allclose(out[sl], coarse[sl], rtol=0, atol=1e-13)
    mask = np.ones(out.shape, bool)
    mask[sl] = False
    assert np.all(out[mask] == 7.0)
