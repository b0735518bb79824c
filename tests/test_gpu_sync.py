"""Level synchronisation on the GPU against the oracle (ratio 2): the coarsening of E (and the
flux sum) and of the moments from a fine level onto the coarse nodes it covers (phx_coarsen,
bit-exact against oracle_lib.coarsen, itself pinned in test_oracle_coarsen.py), prepareStep's
Bold copy, the flux-sum accumulation, and the reflux (B = Faraday(Bold, Eavg, dt) plus the
magnetic ghost fill) against the oracle's Faraday and fills."""
import sys

import numpy as np
import pytest

import oracle_lib as O
import parity_util as U
from phare_amd import phx

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _levels(dim, order):
    nc = [[30], [16, 12], [12, 11, 10]][dim - 1]
    dx = [0.4, 0.3, 0.3][:dim]
    coarse = phx.Level(dim, order, nc, dx, apply_tuning_table=False)
    clo = [3] * dim
    chi = [n - 4 for n in nc]
    flo = [2 * a for a in clo]
    fn = [2 * (b - a + 1) for a, b in zip(clo, chi)]
    fine = phx.Level(dim, order, fn, [h / 2 for h in dx], lower=flo, periodic=False,
                     apply_tuning_table=False)
    return coarse, fine, clo, chi, flo, fn


# (coarse field, fine field, operator, per-direction dual flags of its quantity)
def _cases(dim):
    e = [(phx.F_E + k, phx.F_E + k, phx.COARSEN_ELECTRIC) for k in range(3)]
    f = [(phx.F_EAVG + k, phx.F_FLUXSUM + k, phx.COARSEN_ELECTRIC) for k in range(3)]
    m = [(phx.F_RHO, phx.F_RHO, phx.COARSEN_MOMENTS), (phx.F_VI + 1, phx.F_VI + 1, phx.COARSEN_MOMENTS),
         (phx.f_pop(0, 2), phx.f_pop(0, 2), phx.COARSEN_MOMENTS)]
    return e + f + m


@pytest.mark.parametrize("dim,order", [(1, 1), (2, 2), (3, 3)])
def test_coarsen_matches_oracle(dim, order):
    _say(f"[gpu] coarsen dim={dim} order={order}")
    rng = np.random.default_rng(40 + dim)
    coarse, fine, clo, chi, flo, fn = _levels(dim, order)
    g = [2, 4, 4][order - 1]
    for cf, ff, op in _cases(dim):
        fine_a = rng.standard_normal(fine.shape(ff))
        coarse_a = rng.standard_normal(coarse.shape(cf))
        fine.set(ff, fine_a)
        coarse.set(cf, coarse_a)
        comp = (cf - phx.F_E) % 3 if op == phx.COARSEN_ELECTRIC else -1
        dual_dir = comp if 0 <= comp < dim else -1
        lo = list(clo)
        hi = [b + (0 if d == dual_dir else 1) for d, b in enumerate(chi)]
        coarse.coarsen(cf, fine, ff, op, lo, hi)
        want = O.coarsen(fine_a, [a - g for a in flo], coarse_a, [-g] * dim, lo, hi, dual_dir)
        assert np.array_equal(coarse.get(cf), want), (cf, ff, op)
    # a box leaving the fine field is refused
    with pytest.raises(phx.PhxError):
        coarse.coarsen(phx.F_RHO, fine, phx.F_RHO, phx.COARSEN_MOMENTS, [0] * dim, [n - 1 for n in chi])
    coarse.close()
    fine.close()


@pytest.mark.parametrize("dim,order", [(1, 2), (2, 2), (3, 1)])
def test_prepare_step_flux_sum_and_reflux(dim, order):
    _say(f"[gpu] flux sum and reflux dim={dim} order={order}")
    rng = np.random.default_rng(90 + dim)
    nc = [[30], [16, 12], [12, 11, 10]][dim - 1]
    dx = [0.4, 0.3, 0.3][:dim]
    g = phx.Level(dim, order, nc, dx, apply_tuning_table=False)
    o = O.Sim(dim, order, nc, dx)
    B = [rng.standard_normal(g.shape(phx.F_B + k)) + (1.0 if k == 0 else 0.0) for k in range(3)]
    for k in range(3):
        g.set(phx.F_B + k, B[k])
    g.prepare_step()
    for k in range(3):
        assert np.array_equal(g.get(phx.F_BOLD + k), B[k])
    # fluxSum = 0; += Eavg * 0.25 twice (two substeps' averaged E)
    g.reset_flux_sum()
    E1 = [rng.standard_normal(g.shape(phx.F_EAVG + k)) for k in range(3)]
    E2 = [rng.standard_normal(g.shape(phx.F_EAVG + k)) for k in range(3)]
    for E in (E1, E2):
        for k in range(3):
            g.set(phx.F_EAVG + k, E[k])
        g.accumulate_flux_sum(0.25)
    for k in range(3):
        want = (np.zeros_like(E1[k]) + E1[k] * 0.25) + E2[k] * 0.25
        assert np.array_equal(g.get(phx.F_FLUXSUM + k), want)
    # the state B moves on; reflux recomputes it from Bold and the (refluxed) Eavg
    for k in range(3):
        g.set(phx.F_B + k, rng.standard_normal(g.shape(phx.F_B + k)))
    dt = 0.01
    g.reflux(dt)
    # the oracle: Bold in its predicted set, Eavg in its averaged set, Faraday into the state B,
    # then the periodic magnetic ghost fill
    for k in range(3):
        o.set(O.F_BPRED + k, B[k])
        o.set(O.F_EAVG + k, E2[k])
    o.faraday(1, 2, 0, dt)
    o.fill_ghosts(O.F_B)
    o.fill_ghosts(O.F_B + 1)
    o.fill_ghosts(O.F_B + 2)
    for k in range(3):
        U.assert_close(g.get(phx.F_B + k), o.get(O.F_B + k), rtol=0, what=f"refluxed B{k}")
    g.close()
