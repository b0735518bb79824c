"""Coarse-to-fine coupling of two levels on the GPU against the oracle: the particle split into
a fine level's domain and level ghost arrays (phx_refine_particles; the oracle's split is pinned
bit for bit against the reference's Splitter in test_oracle_split.py), the lastStep swap of the
level ghost arrays, the moments of fillIonPopMomentGhosts (old ghosts with 1 - alpha, new with
alpha) and the linear time interpolation of fields. Particle sets bit-exact, moments at 1e-12."""
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


def _ghost_boxes(lo, hi, pg):
    """the level ghost region of a patch: its box grown by pg less the box, as disjoint slabs
    (lower and upper slab along each direction, the later directions inside the earlier ones)"""
    dim = len(lo)
    boxes = []
    glo = [a - pg for a in lo]
    ghi = [b + pg for b in hi]
    for d in range(dim):
        inner_lo = [lo[k] if k < d else glo[k] for k in range(dim)]
        inner_hi = [hi[k] if k < d else ghi[k] for k in range(dim)]
        low = (tuple(inner_lo[:d] + [glo[d]] + inner_lo[d + 1:]),
               tuple(inner_hi[:d] + [lo[d] - 1] + inner_hi[d + 1:]))
        high = (tuple(inner_lo[:d] + [hi[d] + 1] + inner_lo[d + 1:]),
                tuple(inner_hi[:d] + [ghi[d]] + inner_hi[d + 1:]))
        boxes += [low, high]
    return boxes


CASES = [(1, 1, 2), (1, 3, 4), (1, 3, 5), (2, 1, 4), (2, 2, 4), (2, 3, 4), (3, 1, 6), (3, 2, 6),
         (3, 3, 6)]


def _setup(dim, order, seed):
    rng = np.random.default_rng(seed)
    nc = [[30], [16, 12], [12, 11, 10]][dim - 1]
    dx = [0.4, 0.3, 0.3][:dim]
    coarse = phx.Level(dim, order, nc, dx, capacity=20000, apply_tuning_table=False)
    # the fine patch: coarse cells 3 .. n - 4 (ratio 2)
    clo = [3] * dim
    chi = [n - 4 for n in nc]
    flo = [2 * a for a in clo]
    fhi = [2 * b + 1 for b in chi]
    fn = [b - a + 1 for a, b in zip(flo, fhi)]
    fine = phx.Level(dim, order, fn, [h / 2 for h in dx], lower=flo, periodic=False, capacity=20000,
                     apply_tuning_table=False)
    n = [4000, 6000, 8000][dim - 1]
    p = U.random_particles(rng, dim, n, nc, [0] * dim, 0, vscale=0.5)
    coarse.set_particles(0, phx.DOMAIN, p)
    return rng, coarse, fine, p, flo, fhi


@pytest.mark.parametrize("dim,order,nbr", CASES)
def test_refine_particles_match_oracle(dim, order, nbr):
    _say(f"[gpu] refine particles dim={dim} order={order} nbr={nbr}")
    rng, coarse, fine, p, flo, fhi = _setup(dim, order, 70 + 10 * dim + order)
    pg = [1, 2, 2][order - 1]
    ghosts = _ghost_boxes(flo, fhi, pg)
    interior = [(tuple(flo), tuple(fhi))]
    fine.refine_particles(coarse, ghosts, kind=phx.LEVEL_GHOST_NEW, nbr=nbr)
    fine.refine_particles(coarse, interior, kind=phx.DOMAIN, nbr=nbr)
    got_g = fine.get_particles(0, phx.LEVEL_GHOST_NEW)
    got_d = fine.get_particles(0, phx.DOMAIN)
    want_g = O.split(dim, order, nbr, p, ghosts)
    want_d = O.split(dim, order, nbr, p, interior)
    assert len(want_g["weight"]) > 0 and len(want_d["weight"]) > 0
    U.assert_same_particles(got_g, want_g, exact=True, what="level ghost split")
    U.assert_same_particles(got_d, want_d, exact=True, what="interior split")
    # the kept refined particles lie in their boxes
    assert np.all((got_d["icell"] >= flo) & (got_d["icell"] <= fhi))
    coarse.close()
    fine.close()


@pytest.mark.parametrize("dim,order", [(1, 2), (2, 2), (3, 1)])
def test_level_ghost_steps_and_alpha_moments(dim, order):
    """firstStep fills levelGhostNew, lastStep swaps it into levelGhostOld (levelGhost a copy,
    levelGhostNew empty); fillIonPopMomentGhosts deposits old x (1 - alpha) + new x alpha: the
    moments against the oracle's deposits of the same arrays"""
    _say(f"[gpu] level ghost old/new dim={dim} order={order}")
    rng, coarse, fine, p, flo, fhi = _setup(dim, order, 300 + dim)
    pg = [1, 2, 2][order - 1]
    ghosts = _ghost_boxes(flo, fhi, pg)
    fine.refine_particles(coarse, ghosts, kind=phx.LEVEL_GHOST_NEW)
    first = fine.get_particles(0, phx.LEVEL_GHOST_NEW)
    fine.level_ghost_last_step()
    assert fine.count(0, phx.LEVEL_GHOST_NEW) == 0
    U.assert_same_particles(fine.get_particles(0, phx.LEVEL_GHOST_OLD), first, exact=True, what="old")
    U.assert_same_particles(fine.get_particles(0, phx.LEVEL_GHOST), first, exact=True, what="levelGhost")
    # the coarse level moves on: new ghosts from displaced coarse particles
    p2 = dict(p)
    p2["delta"] = (p["delta"] + 0.37) % 1.0
    coarse.set_particles(0, phx.DOMAIN, p2)
    fine.refine_particles(coarse, ghosts, kind=phx.LEVEL_GHOST_NEW)
    second = fine.get_particles(0, phx.LEVEL_GHOST_NEW)
    alpha = 0.375
    o = O.Sim(dim, order, [b - a + 1 for a, b in zip(flo, fhi)], [0.2, 0.15, 0.15][:dim], lower=flo)
    o.set_particles(0, 3, first)
    o.set_particles(0, 4, second)
    fine.deposit(0, phx.LEVEL_GHOST_OLD, 1.0 - alpha)
    fine.deposit(0, phx.LEVEL_GHOST_NEW, alpha)
    o.deposit(0, 3, 1.0 - alpha)
    o.deposit(0, 4, alpha)
    for m in range(5):
        U.assert_close(fine.get(phx.f_pop(0, m)), o.get(O.f_pop(0, m)), what=f"moment {m}")
    coarse.close()
    fine.close()


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_time_interpolation(dim):
    _say(f"[gpu] time interpolation dim={dim}")
    rng = np.random.default_rng(dim)
    nc = [[20], [12, 10], [8, 7, 6]][dim - 1]
    a = phx.Level(dim, 2, nc, [0.3] * dim, periodic=False, apply_tuning_table=False)
    b = phx.Level(dim, 2, nc, [0.3] * dim, periodic=False, apply_tuning_table=False)
    for fid in (phx.F_B + 1, phx.F_E):
        a.set(fid, rng.standard_normal(a.shape(fid)))
        b.set(fid, rng.standard_normal(b.shape(fid)))
    # into a third set of a: (1 - alpha) B_a + alpha B_b on a sub-box, the rest untouched
    dst = phx.F_BPRED + 1
    before = rng.standard_normal(a.shape(dst))
    a.set(dst, before)
    shape = a.shape(dst)
    lo = [1] * dim
    hi = [s - 3 for s in shape]
    alpha = 0.3
    a.time_interpolate(dst, a, phx.F_B + 1, b, phx.F_B + 1, lo, hi, alpha)
    want = before.copy()
    sl = tuple(slice(l, h + 1) for l, h in zip(lo, hi))
    want[sl] = (1.0 - alpha) * a.get(phx.F_B + 1)[sl] + alpha * b.get(phx.F_B + 1)[sl]
    assert np.array_equal(a.get(dst), want)
    # shapes must agree: Ex (dual along x) into a Bx slot (primal along x) is refused
    with pytest.raises(phx.PhxError):
        a.time_interpolate(phx.F_BPRED, a, phx.F_E, b, phx.F_E, lo, hi, alpha)
    a.close()
    b.close()
