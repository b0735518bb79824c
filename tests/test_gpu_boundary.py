"""The pusher- and interpolator-level entry points of the C ABI (phx_push, phx_gather,
phx_deposit), the level-ghost scenario of the reference's IonUpdater test, the staged (non
captured) level path of phx_advance, and the first-error freeze of a captured step, against the
oracle on identical inputs (FP64: cell indices identical, values within 1e-12; a single push is
bit-exact)."""
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


def _pair(dim, order, ncells, dx, **kw):
    return phx.Level(dim, order, ncells, dx, **kw), O.Sim(dim, order, ncells, dx, **kw)


def _random_em(rng, g, o, eset=2, scale=0.3):
    base = {0: 0, 1: 6, 2: 12}[eset]
    for k in range(6):
        a = scale * rng.standard_normal(g.shape(base + k))
        if k >= 3:
            a += 1.0 if k == 3 else 0.4
        g.set(base + k, a)
        o.set(base + k, a)


def _key_order(p):
    """lexsort index of a particle set (cell, delta, velocity): pairs per-particle outputs"""
    keys = [p["v"][:, 2], p["v"][:, 1], p["v"][:, 0]]
    for d in reversed(range(p["delta"].shape[1])):
        keys.append(p["delta"][:, d])
    for d in reversed(range(p["icell"].shape[1])):
        keys.append(p["icell"][:, d])
    return np.lexsort(keys)


def _slice(p, a, b):
    return {k: v[a:b] for k, v in p.items()}


@pytest.mark.parametrize("dim,order", [(1, 1), (1, 3), (2, 2), (3, 1), (3, 2)])
def test_push_with_selectors_matches_oracle(dim, order):
    """Pusher::move (boris.hpp:93-138) with the level-ghost selectors of the updater: first =
    inGhostBox, second = inDomainBox; the partition ends and the three groups bit for bit"""
    _say(f"[gpu] phx_push dim={dim} order={order}")
    rng = np.random.default_rng(5100 + 10 * dim + order)
    nc = [[23], [14, 11], [9, 10, 11]][dim - 1]
    dx = [0.3, 0.25, 0.2][:dim]
    g, o = _pair(dim, order, nc, dx)
    _random_em(rng, g, o)
    pg = [1, 2, 2][order - 1]
    # particles in the ghost box: after the prePush they stay inside the particle arrays' cell
    # map (the ghost box grown by one cell, particles_data.hpp:76-80), where the reference's
    # partition is defined
    p = U.random_particles(rng, dim, 2500, nc, [0] * dim, pg, vscale=1.5)
    g.set_particles(0, phx.LEVEL_GHOST, p)
    o.set_particles(0, 2, p)
    ghost = [(tuple([-pg] * dim), tuple(n - 1 + pg for n in nc))]
    domain = [(tuple([0] * dim), tuple(n - 1 for n in nc))]
    dt = 0.05
    ends_g = g.push(dt, ghost, domain, kind=phx.LEVEL_GHOST)
    ends_o = o.push(dt, ghost, domain, kind=2)
    assert ends_g == ends_o
    a, b = g.get_particles(0, phx.LEVEL_GHOST), o.get_particles(0, 2)
    e2, e1 = ends_o
    for lo, hi, what in ((0, e2, "both"), (e2, e1, "first only"), (e1, len(b["weight"]), "neither")):
        U.assert_same_particles(_slice(a, lo, hi), _slice(b, lo, hi), exact=True, what=what)
    g.close()


@pytest.mark.parametrize("dim,order", [(1, 2), (2, 3), (3, 2)])
def test_gather_matches_oracle(dim, order):
    """Interpolator::operator()(particle, em, layout) at every particle, bit for bit"""
    _say(f"[gpu] phx_gather dim={dim} order={order}")
    rng = np.random.default_rng(5200 + 10 * dim + order)
    nc = [[23], [14, 11], [9, 10, 11]][dim - 1]
    dx = [0.3, 0.25, 0.2][:dim]
    g, o = _pair(dim, order, nc, dx)
    _random_em(rng, g, o, eset=1)
    p = U.random_particles(rng, dim, 3000, nc, [0] * dim, 1)
    g.set_particles(0, phx.DOMAIN, p)
    o.set_particles(0, 0, p)
    Eg, Bg = g.gather(phx.EM_PRED)
    Eo, Bo = o.gather(1)
    ig = _key_order(g.get_particles())
    io = _key_order(o.get_particles())
    assert np.array_equal(Eg[ig], Eo[io]) and np.array_equal(Bg[ig], Bo[io])
    g.close()


@pytest.mark.parametrize("dim,order", [(1, 3), (2, 2), (3, 1)])
def test_deposit_matches_oracle(dim, order):
    """Interpolator deposit (interpolator.hpp:468-510): domain at coef 1 then the level ghosts at
    coef 0.5 twice (fillIonPopMomentGhosts' old/new pair), added into the moments"""
    _say(f"[gpu] phx_deposit dim={dim} order={order}")
    rng = np.random.default_rng(5300 + 10 * dim + order)
    nc = [[23], [14, 11], [9, 10, 11]][dim - 1]
    dx = [0.3, 0.25, 0.2][:dim]
    g, o = _pair(dim, order, nc, dx)
    p = U.random_particles(rng, dim, 3000, nc, [0] * dim, 0)
    lg = U.random_particles(rng, dim, 500, nc, [0] * dim, 1)
    for sim, kinds in ((g, (phx.DOMAIN, phx.LEVEL_GHOST)), (o, (0, 2))):
        sim.set_particles(0, kinds[0], p)
        sim.set_particles(0, kinds[1], lg)
        sim.deposit(0, kinds[0], 1.0)
        sim.deposit(0, kinds[1], 0.5)
        sim.deposit(0, kinds[1], 0.5)
    for k in range(5):
        U.assert_close(g.get(phx.f_pop(0, k)), o.get(O.f_pop(0, k)), what=f"moment {k}")
    g.close()


def _updater_scenario(sim, order, kinds):
    """the fixture of tests/core/numerics/ion_updater/test_updater.cpp:376-527: 1D, 100 cells of
    0.1, two populations (mass 1, charge 1), density 1, v = 0, vth = 0.1, 1000 ppc, E = B = 0;
    level ghosts: copies of the first (order 1) or first two cells' particles shifted into the
    left ghost cells; no patch ghosts. Domain and level ghosts deposited, then updateIons."""
    n = np.ones(100)
    V = [np.zeros(100)] * 3
    T = [np.full(100, 0.1)] * 3
    for pop in (0, 1):
        sim.load_maxwellian(pop, n, V, T, 1.0, 1000, 1000 + pop)
        d = sim.get_particles(pop, kinds[0])
        shift = 1 if order == 1 else 2
        keep = d["icell"][:, 0] < shift
        lg = {k: v[keep].copy() for k, v in d.items()}
        lg["icell"][:, 0] -= shift
        sim.set_particles(pop, kinds[2], lg)
        sim.deposit(pop, kinds[0], 1.0)
        sim.deposit(pop, kinds[2], 1.0)
    sim.update_ions()


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("mode", [1, 2])
def test_updater_level_ghost_scenario(order, mode):
    """the reference's IonUpdater test (test_updater.cpp:744-889) on both sides: after
    updatePopulations, the old/new level-ghost deposit (alpha = 0.5) and updateIons, every
    physical node's moments and Vi have changed, density is within 0.07 of the prescribed 1 on
    the cells' left nodes, no NaN; in domain_only the particle arrays are untouched. The HIP
    path equals the oracle (moments 1e-12, particle sets bit for bit)."""
    _say(f"[gpu] updater level-ghost scenario order={order} mode={mode}")
    # a patch of a wider level (not periodic): the left border touches the coarser level
    g = phx.Level(1, order, [100], [0.1], masses=(1.0, 1.0), periodic=False, capacity=120000)
    o = O.Sim(1, order, [100], [0.1], masses=(1.0, 1.0))
    fields = {}
    for sim, kinds, tag in ((g, (phx.DOMAIN, phx.PATCH_GHOST, phx.LEVEL_GHOST), "g"), (o, (0, 1, 2), "o")):
        _updater_scenario(sim, order, kinds)
        before = {f: sim.get(f) for f in [phx.f_pop(p, k) for p in (0, 1) for k in range(5)] + [27, 28, 29]}
        parts0 = [(sim.get_particles(p, kinds[2]), sim.get_particles(p, kinds[1])) for p in (0, 1)]
        sim.update_populations(0.01, mode, 0)
        # fillIonPopMomentGhosts: levelGhostNew x alpha + levelGhostOld x (1 - alpha), the old and
        # new arrays holding the unpushed copies
        for p in (0, 1):
            lg_pushed = sim.get_particles(p, kinds[2])
            sim.set_particles(p, kinds[2], parts0[p][0])
            sim.deposit(p, kinds[2], 0.5)
            sim.deposit(p, kinds[2], 0.5)
            if mode == 1:
                U.assert_same_particles(lg_pushed, parts0[p][0], exact=True, what="levelGhost untouched")
                assert len(sim.get_particles(p, kinds[1])["weight"]) == 0
        sim.update_ions()
        after = {f: sim.get(f) for f in before}
        g0 = 4 if order > 1 else 2
        for f in before:
            phys = slice(g0, g0 + 101)
            assert np.all(np.isfinite(after[f][phys])), (tag, f)
            assert np.all(np.abs(after[f][phys] - before[f][phys]) > 0.0), (tag, f)
        for p in (0, 1):
            dens = after[phx.f_pop(p, 0)][g0:g0 + 100]
            assert np.abs(dens - 1.0).max() <= 0.07, (tag, p, np.abs(dens - 1.0).max())
        fields[tag] = after
    for f in fields["g"]:
        a, b = fields["g"][f], fields["o"][f]
        # Vi is 0/0 where no particle reaches a ghost node, on both sides
        assert np.array_equal(np.isnan(a), np.isnan(b)), f
        U.assert_close(np.nan_to_num(a), np.nan_to_num(b), what=f"field {f}")
    for p in (0, 1):
        U.assert_same_particles(g.get_particles(p, phx.DOMAIN), o.get_particles(p, 0), exact=True,
                                what=f"pop {p} domain")
    g.close()


def test_advance_staged_path_with_patch_ghosts():
    """phx_advance on a periodic level holding patch-ghost particles takes the staged sequence
    (updatePopulations, fills, updateIons between the field stages: no captured step) with a sort
    inside the run; 8 PPC steps against the oracle at 1e-12"""
    _say("[gpu] advance staged path")
    wl = W.harris2d(ncells=(24, 16), ppc=8, order=2)
    rng = np.random.default_rng(61)
    g = phx.Level(2, 2, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te, capacity=8000)
    o = O.Sim(2, 2, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te)
    g.set_sort_interval(3)
    pg = U.random_particles(rng, 2, 40, wl.ncells, [0, 0], 2, vscale=0.3)
    outside = np.any((pg["icell"] < 0) | (pg["icell"] >= np.array(wl.ncells)), axis=1)
    pg = {k: v[outside] for k, v in pg.items()}
    for sim, fb, kinds in ((g, phx.F_B, (phx.DOMAIN, phx.PATCH_GHOST)), (o, O.F_B, (0, 1))):
        wl.setup(sim, fb)
        sim.set_particles(0, kinds[1], pg)
    g.advance(wl.dt, 8)
    o.advance(wl.dt, 8)
    for fid in range(6):
        U.assert_close(g.get(fid), o.get(fid), what=f"field {fid}")
    for k in range(5):
        U.assert_close(g.get(phx.f_pop(0, k)), o.get(O.f_pop(0, k)), what=f"moment {k}")
    U.assert_same_particles(g.get_particles(), o.get_particles())
    g.close()


def test_advance_stops_at_the_first_two_cell_move():
    """a two-cell move in the first moveIons of a captured step (boris.hpp:164-165): phx_advance
    reports it and every later kernel of the run returns at entry, so that the state is the one
    the failing stage left (the reference throws out of moveIons_, solver_ppc.hpp:558-563): the
    E/B state and the particles equal their values before the call"""
    _say("[gpu] advance stops at the first error")
    rng = np.random.default_rng(3)
    g = phx.Level(1, 1, [40], [0.2], capacity=2000)
    p = U.random_particles(rng, 1, 1200, [40], [0], 0, vscale=0.2)
    p["charge"][:] = 1.0
    p["v"][17] = (1000.0, 0.0, 0.0)
    g.set_particles(0, phx.DOMAIN, p)
    g.set(phx.F_B, np.ones(g.shape(phx.F_B)))
    g.initialize()
    before = {f: g.get(f) for f in range(6)}
    parts = g.get_particles()
    with pytest.raises(phx.PhxError) as e:
        g.advance(0.01, 6)
    assert "moved 2 cells" in str(e.value) and "PrePush" in str(e.value)
    for f in range(6):
        assert np.array_equal(g.get(f), before[f]), f
    U.assert_same_particles(g.get_particles(), parts, exact=True)
    g.close()
