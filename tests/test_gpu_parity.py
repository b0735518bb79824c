"""Parity of the HIP path (libphx.so through the C ABI) with the CPU oracle on identical seeded
inputs. The plain FP64 path must give identical cell indices, and deltas, velocities, moments,
E and B within 1e-12 relative (max-norm over each array); a single push is bit-exact.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
import parity_util as U
from phare_amd import phx
from phare_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [(1, 1), (1, 2), (1, 3), (2, 1), (2, 2), (2, 3), (3, 1), (3, 2), (3, 3)]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair(dim, order, ncells, dx, masses=(1.0,), **kw):
    g = phx.Level(dim, order, ncells, dx, masses=masses, **kw)
    o = O.Sim(dim, order, ncells, dx, masses=masses, **kw)
    return g, o


def _ncells(dim):
    return [[23], [14, 11], [9, 10, 11]][dim - 1]


def _random_em(rng, g, o, eset=2, scale=0.3):
    base = {0: 0, 1: 6, 2: 12}[eset]
    for k in range(6):
        a = scale * rng.standard_normal(g.shape(base + k))
        if k >= 3:
            a += 1.0 if k == 3 else 0.4
        g.set(base + k, a)
        o.set(base + k, a)


@pytest.mark.parametrize("dim,order", CASES)
@pytest.mark.parametrize("mode", [1, 2])
def test_update_populations_matches_oracle(dim, order, mode):
    _say(f"[gpu] update_populations dim={dim} order={order} mode={mode}")
    rng = np.random.default_rng(1000 + 10 * dim + order + 100 * mode)
    nc = _ncells(dim)
    dx = [0.3, 0.25, 0.2][:dim]
    g, o = _pair(dim, order, nc, dx)
    _random_em(rng, g, o)
    pg = [1, 2, 2][order - 1]
    p = U.random_particles(rng, dim, 3000, nc, [0] * dim, 0, vscale=1.5)
    g.set_particles(0, phx.DOMAIN, p)
    o.set_particles(0, 0, p)
    dt = 0.05
    g.update_populations(dt, mode)
    o.update_populations(dt, mode)
    for k in range(5):
        U.assert_close(g.get(phx.f_pop(0, k)), o.get(O.f_pop(0, k)), what=f"moment {k}")
    if mode == 2:
        # a single push is bit-exact; the arrays hold the same particle sets
        U.assert_same_particles(g.get_particles(0, phx.DOMAIN), o.get_particles(0, 0), exact=True,
                                what="domain")
        U.assert_same_particles(g.get_particles(0, phx.PATCH_GHOST), o.get_particles(0, 1),
                                exact=True, what="patchGhost")
    else:
        # domain_only leaves the particles untouched (ion_updater.hpp:181)
        U.assert_same_particles(g.get_particles(0, phx.DOMAIN), p, exact=True, what="untouched")
    g.close()


@pytest.mark.parametrize("dim,order", [(1, 1), (2, 2), (3, 1)])
def test_level_ghost_particles_match_oracle(dim, order):
    """level-ghost selection (inGhostBox first, inDomainBox / inGhostLayer second) and their
    export into the domain (ion_updater.hpp:195-217, 275-288)"""
    _say(f"[gpu] level ghosts dim={dim} order={order}")
    rng = np.random.default_rng(2000 + 10 * dim + order)
    nc = _ncells(dim)
    dx = [0.3, 0.25, 0.2][:dim]
    for mode in (1, 2):
        g, o = _pair(dim, order, nc, dx)
        _random_em(rng, g, o)
        pg = [1, 2, 2][order - 1]
        dom = U.random_particles(rng, dim, 800, nc, [0] * dim, 0, vscale=2.0)
        # level ghosts in the ghost layer (outside the domain, inside the ghost box)
        lg = U.random_particles(rng, dim, 2000, nc, [0] * dim, pg, vscale=2.0)
        inside = np.all((lg["icell"] >= 0) & (lg["icell"] < np.array(nc)), axis=1)
        lg = {k: v[~inside] for k, v in lg.items()}
        for sim, kinds in ((g, (phx.DOMAIN, phx.LEVEL_GHOST)), (o, (0, 2))):
            sim.set_particles(0, kinds[0], dom)
            sim.set_particles(0, kinds[1], lg)
        # dt small enough that no particle moves two cells in a half step
        g.update_populations(0.03, mode)
        o.update_populations(0.03, mode)
        for k in range(5):
            U.assert_close(g.get(phx.f_pop(0, k)), o.get(O.f_pop(0, k)), what=f"moment {k}")
        if mode == 2:
            U.assert_same_particles(g.get_particles(0, phx.DOMAIN), o.get_particles(0, 0), exact=True,
                                    what="domain")
            U.assert_same_particles(g.get_particles(0, phx.LEVEL_GHOST), o.get_particles(0, 2),
                                    exact=True, what="levelGhost")
        g.close()


@pytest.mark.parametrize("dim,order", CASES)
def test_field_operators_match_oracle(dim, order):
    _say(f"[gpu] field operators dim={dim} order={order}")
    rng = np.random.default_rng(3000 + 10 * dim + order)
    nc = _ncells(dim)
    dx = [0.3, 0.25, 0.2][:dim]
    for hyper in (0, 1):
        g, o = _pair(dim, order, nc, dx, eta=0.01, nu=0.02, hyper_mode=hyper, Te=0.2)
        for eset in (0, 1, 2):
            _random_em(rng, g, o, eset)
        for fid in list(range(18, 24)) + [24, 26, 27, 28, 29]:
            a = rng.standard_normal(g.shape(fid))
            g.set(fid, a)
            o.set(fid, a)
        rho = 1.0 + 0.3 * rng.random(g.shape(phx.F_RHO))
        g.set(phx.F_RHO, rho)
        o.set(O.F_RHO, rho)
        g.faraday(0, 2, 1, 0.07)
        o.faraday(0, 2, 1, 0.07)
        g.ampere(1)
        o.ampere(1)
        g.electrons()
        o.electrons()
        g.ohm(1, 1)
        o.ohm(1, 1)
        for fid in list(range(6, 12)) + list(range(18, 25)):
            U.assert_close(g.get(fid), o.get(fid), rtol=0, what=f"field {fid} hyper {hyper}")
        # in-place Faraday (corrector) and the averaged set with its ghost fill
        g.faraday(0, 2, 0, 0.03)
        o.faraday(0, 2, 0, 0.03)
        g.average()
        o.average()
        for fid in range(0, 18):
            U.assert_close(g.get(fid), o.get(fid), rtol=0, what=f"field {fid} after average")
        g.close()


@pytest.mark.parametrize("dim,order", CASES)
def test_fills_match_oracle(dim, order):
    _say(f"[gpu] periodic fills dim={dim} order={order}")
    rng = np.random.default_rng(4000 + 10 * dim + order)
    nc = _ncells(dim)
    dx = [0.3, 0.25, 0.2][:dim]
    g, o = _pair(dim, order, nc, dx)
    for k in range(5):
        a = rng.random(g.shape(phx.f_pop(0, k)))
        g.set(phx.f_pop(0, k), a)
        o.set(O.f_pop(0, k), a)
    g.fill_moment_borders()
    o.fill_moment_borders()
    for k in range(5):
        U.assert_close(g.get(phx.f_pop(0, k)), o.get(O.f_pop(0, k)), rtol=0, what=f"border sum {k}")
    g.update_ions()
    o.update_ions()
    g.fill_ion_borders()
    o.fill_ion_borders()
    for fid in (25, 26, 27, 28, 29):
        U.assert_close(g.get(fid), o.get(fid), rtol=0, what=f"ion total {fid}")
    for base in (0, 3, 18):
        for k in range(3):
            a = rng.standard_normal(g.shape(base + k))
            g.set(base + k, a)
            o.set(base + k, a)
        g.fill_ghosts(base)
        for k in range(3):
            o.fill_ghosts(base + k)
            U.assert_close(g.get(base + k), o.get(base + k), rtol=0, what=f"ghost fill {base + k}")
    g.close()


def test_maxwellian_loader_bitwise():
    """the seeded loader reproduces the oracle's particles bit for bit (same libstdc++ draws)"""
    _say("[gpu] maxwellian loader")
    wl = W.harris2d(ncells=(20, 16), ppc=7)
    g = phx.Level(2, 2, wl.ncells, wl.dx)
    o = O.Sim(2, 2, wl.ncells, wl.dx)
    n, V, T = wl.profiles()
    g.load_maxwellian(0, n, V, T, 1.0, wl.ppc, 12334)
    o.load_maxwellian(0, n, V, T, 1.0, wl.ppc, 12334)
    # the HIP arrays store particles in their own (chunk-interleaved) order: compare the sets
    U.assert_same_particles(g.get_particles(), o.get_particles(), exact=True, what="loader")
    g.close()


def test_aos_records_round_trip():
    """reference Particle<dim> records in and out (particle.hpp:37-61; 56/64/80 B)"""
    _say("[gpu] AoS round trip")
    for dim, size in ((1, 56), (2, 64), (3, 80)):
        dt = phx.particle_dtype(dim)
        assert dt.itemsize == size
        rng = np.random.default_rng(dim)
        nc = _ncells(dim)
        rec = np.zeros(100, dtype=dt)
        rec["weight"] = rng.random(100)
        rec["charge"] = 1.0
        rec["iCell"] = rng.integers(0, 9, (100, dim))
        rec["delta"] = rng.random((100, dim))
        rec["v"] = rng.standard_normal((100, 3))
        g = phx.Level(dim, 1, nc, [0.2] * dim)
        g.set_particles_aos(0, phx.DOMAIN, rec)
        back = g.get_particles_aos(0, phx.DOMAIN, dt)
        order = lambda r: np.lexsort((r["weight"],))
        back, ref = back[order(back)], rec[order(rec)]
        for k in ("weight", "charge", "iCell", "delta", "v"):
            assert np.array_equal(back[k], ref[k]), (dim, k)
        g.close()


def test_two_cell_move_is_an_error():
    """boris.hpp:164-165 -> DictionaryException "Updater::updatePopulations"
    (tests/simulator/test_exceptions.py: v=1000, dt=0.001, dl=0.3)"""
    _say("[gpu] two-cell move error")
    g = phx.Level(1, 1, [20], [0.3])
    p = dict(icell=np.array([[5]], np.int32), delta=np.array([[0.5]]), v=np.array([[1000.0, 0, 0]]),
             weight=np.ones(1), charge=np.ones(1))
    g.set_particles(0, phx.DOMAIN, p)
    with pytest.raises(phx.PhxError) as e:
        g.update_populations(0.002, phx.MODE_ALL)
    assert "moved 2 cells" in str(e.value)
    g.close()


def _run_pair(wl, steps):
    g = phx.Level(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te,
                  capacity=wl.n_particles)
    o = O.Sim(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te)
    wl.setup(g, phx.F_B)
    wl.setup(o, O.F_B)
    for fid in list(range(0, 6)) + [18, 19, 20, 25]:
        U.assert_close(g.get(fid), o.get(fid), what=f"initial field {fid}")
    g.advance(wl.dt, steps)
    o.advance(wl.dt, steps)
    g.sync()
    return g, o


def _check_state(g, o, wl, rtol=U.RTOL):
    for fid in range(0, 6):
        U.assert_close(g.get(fid), o.get(fid), rtol, what=f"field {fid}")
    for k in range(5):
        U.assert_close(g.get(phx.f_pop(0, k)), o.get(O.f_pop(0, k)), rtol, what=f"moment {k}")
    U.assert_close(g.get(phx.F_RHO), o.get(O.F_RHO), rtol, what="rho")
    U.assert_same_particles(g.get_particles(), o.get_particles(), rtol)


def test_advance_uniform1d_full():
    """the 1D CPU-runnable workload in full: 200 cells, 20k particles, 100 PPC steps"""
    _say("[gpu] advance uniform1d 100 steps")
    wl = W.uniform1d()
    g, o = _run_pair(wl, wl.nsteps)
    _check_state(g, o, wl)
    g.close()


@pytest.mark.parametrize("order", [1, 2, 3])
def test_advance_harris2d_small(order):
    _say(f"[gpu] advance harris2d small order={order}")
    wl = W.harris2d(ncells=(32, 24), ppc=12, order=order)
    g, o = _run_pair(wl, 20)
    _check_state(g, o, wl)
    g.close()


def _two_pop_harris(order=2):
    """the small Harris sheet with a second population (mass 4, charge 2) on the same profiles"""
    wl = W.harris2d(ncells=(32, 24), ppc=8, order=order)
    wl.masses, wl.charges = (1.0, 4.0), (1.0, 2.0)
    return wl


def test_advance_two_populations_2d():
    """two populations of different mass and charge (each of one charge, so the segmented push
    derives each charge density from n with its own charge), 12 PPC steps against the oracle"""
    _say("[gpu] advance two populations 2d")
    wl = _two_pop_harris()
    g, o = _run_pair(wl, 12)
    for fid in range(0, 6):
        U.assert_close(g.get(fid), o.get(fid), what=f"field {fid}")
    for pop in (0, 1):
        for k in range(5):
            U.assert_close(g.get(phx.f_pop(pop, k)), o.get(O.f_pop(pop, k)),
                           what=f"pop {pop} moment {k}")
        U.assert_same_particles(g.get_particles(pop), o.get_particles(pop), U.RTOL,
                                what=f"pop {pop} particles")
    U.assert_close(g.get(phx.F_RHO), o.get(O.F_RHO), what="rho")
    g.close()


def test_advance_mixed_charges_2d():
    """a population whose particles carry different charges: the segmented push deposits the
    charge density particle by particle; 10 PPC steps against the oracle"""
    _say("[gpu] advance mixed charges 2d")
    wl = W.harris2d(ncells=(24, 20), ppc=8, order=2)
    rng = np.random.default_rng(77)
    g = phx.Level(2, 2, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te, capacity=20000)
    o = O.Sim(2, 2, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te)
    p = U.random_particles(rng, 2, 4000, wl.ncells, [0, 0], 0, vscale=0.5)
    p["charge"] = np.where(rng.random(4000) < 0.5, 1.0, 2.0)
    for sim, fb in ((g, phx.F_B), (o, O.F_B)):
        sim.set_particles(0, 0, p)
        for c, b in enumerate(wl.b_arrays(lambda c: sim.shape(fb + c))):
            sim.set(fb + c, b)
        sim.initialize()
    g.advance(wl.dt, 10)
    o.advance(wl.dt, 10)
    g.sync()
    _check_state(g, o, wl)
    g.close()


@pytest.mark.skip(reason="one GPU run faulted the GPU in the 3D order-1 case of this test (first 3D "
                         "advance on the captured fused periodic path); the cause is not shown (CPU "
                         "emulations of the library with poison-filled allocations, deferred streams, "
                         "AddressSanitizer and UndefinedBehaviorSanitizer are clean), so the fused 3D "
                         "path stays off; 3D levels advance on the staged path, tested by "
                         "test_advance_uniform3d_staged")
@pytest.mark.parametrize("order", [1, 2, 3])
def test_advance_uniform3d_small(order):
    _say(f"[gpu] advance uniform3d small order={order}")
    wl = W.uniform3d(ncells=(10, 9, 11), ppc=6, order=order)
    g, o = _run_pair(wl, 8)
    _check_state(g, o, wl)
    g.close()


@pytest.mark.parametrize("order", [1, 2, 3])
def test_advance_uniform3d_staged(order):
    """a 3D periodic level advances on the staged path (the drop-in entry points' kernels: linear
    storage push with the in-place periodic re-entry, Faraday/Ampere/Ohm followed by the periodic
    ghost fill, border sum and max of the moments): 6 PPC steps against the oracle at 1e-12"""
    _say(f"[gpu] advance uniform3d staged order={order}")
    wl = W.uniform3d(ncells=(11, 10, 12), ppc=5, order=order)
    g, o = _run_pair(wl, 6)
    _check_state(g, o, wl)
    g.close()


@pytest.mark.parametrize("order", [1, 2, 3])
def test_advance_uniform3d_bricks(order):
    """64 particles per cell on 8^3 (order 1) or 12^3 cells (orders 2, 3: a periodic level needs
    more cells than twice the ghost width): every chunk of the linear 3D push is one brick of
    4x4x4 cells (the loader's particles put in brick order at upload), so the deposit goes through
    the brick tiles with the lanes of a wavefront in 64 different cells; 3 PPC steps against the
    oracle at 1e-12"""
    _say(f"[gpu] advance uniform3d bricks order={order}")
    n = 8 if order == 1 else 12
    wl = W.uniform3d(ncells=(n, n, n), ppc=64, order=order)
    g, o = _run_pair(wl, 3)
    _check_state(g, o, wl)
    g.close()


def test_advance_uniform3d_sorted_every_step():
    """the staged 3D path on cell-segmented storage (one workgroup per 4x4x4-cell segment), sorted
    again after every PPC step with a time step that moves particles across cells: the run stays
    equal to the oracle's at 1e-12 and the storage stays grouped by segment"""
    _say("[gpu] advance uniform3d sorted every step")
    wl = W.uniform3d(ncells=(8, 8, 8), ppc=64, order=1)
    wl.dt = 0.03
    g = phx.Level(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te,
                  capacity=wl.n_particles, apply_tuning_table=False)
    o = O.Sim(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te)
    g.set_sort_interval(1)
    wl.setup(g, phx.F_B)
    wl.setup(o, O.F_B)
    for _ in range(4):
        g.advance(wl.dt, 1)
        o.advance(wl.dt, 1)
        g.sync()
        q = g.get_particles()
        assert np.all(np.diff(_segment_of(q["icell"], wl.ncells)) >= 0)
    _check_state(g, o, wl)
    g.close()

@pytest.mark.parametrize("charges", ["one", "mixed"])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_advance_3d_segmented_one_and_mixed_charges(order, charges):
    """the staged 3D path's segment push for a population of one charge (the LDS tile leaves the
    charge density out, the flush takes q n) and of mixed charges (deposited particle by
    particle): 4 PPC steps sorted every 2 equal the oracle's at 1e-12, identical cells"""
    _say(f"[gpu] advance 3d segmented order {order} charges {charges}")
    wl = W.uniform3d(ncells=(12, 12, 12), ppc=16, order=order)
    wl.dt = 0.02
    rng = np.random.default_rng(5 + order)
    g = phx.Level(3, order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te, capacity=40000,
                  apply_tuning_table=False)
    g.set_sort_interval(2)
    o = O.Sim(3, order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te)
    n = 16 * 12 ** 3
    p = U.random_particles(rng, 3, n, wl.ncells, [0, 0, 0], 0, vscale=0.5)
    p["charge"] = np.ones(n) if charges == "one" else np.where(rng.random(n) < 0.5, 1.0, 2.0)
    for sim, fb in ((g, phx.F_B), (o, O.F_B)):
        sim.set_particles(0, 0, p)
        for c, b in enumerate(wl.b_arrays(lambda c: sim.shape(fb + c))):
            sim.set(fb + c, b)
        sim.initialize()
    g.advance(wl.dt, 4)
    o.advance(wl.dt, 4)
    g.sync()
    _check_state(g, o, wl)
    g.close()


def test_advance_harris2d_full_size_three_steps():
    """the full 2D Harris workload (256x128 cells, 100 ppc, 3.28 M particles, order 2) over 3 PPC
    steps against the oracle: identical cells, 1e-12 on fields, moments and particles"""
    _say("[gpu] advance harris2d full size, 3 steps")
    wl = W.harris2d()
    g, o = _run_pair(wl, 3)
    _check_state(g, o, wl)
    g.close()


def test_move_ions_general_path_matches_fast_path():
    """updatePopulations + fills (the drop-in sequence of moveIons_) gives the same level state
    as the fused periodic path that wraps leavers in place"""
    _say("[gpu] general vs fused moveIons")
    wl = W.harris2d(ncells=(24, 18), ppc=10)
    a = phx.Level(2, 2, wl.ncells, wl.dx)
    b = phx.Level(2, 2, wl.ncells, wl.dx)
    wl.setup(a, phx.F_B)
    wl.setup(b, phx.F_B)
    for lv in (a, b):
        lv.average()
    a.move_ions(0.3, phx.MODE_ALL)  # fused
    b.update_populations(0.3, phx.MODE_ALL)
    b.fill_moment_borders()
    b.fill_ion_ghost_particles()
    b.update_ions()
    b.fill_ion_borders()
    assert b.count(0, phx.PATCH_GHOST) == 0
    # two levels initialised apart: their deposits sum in different orders (LDS atomics), so
    # the states agree to round-off, not bit for bit
    U.assert_same_particles(a.get_particles(), b.get_particles())
    for fid in (25, 26, 27, 28, 29):
        U.assert_close(a.get(fid), b.get(fid), what=f"total {fid}")
    a.close()
    b.close()


def _segment_of(icell, ncells):
    """segment index of each particle's cell: the domain box is cut into segments of at most 64
    cells (a run of 64 cells in 1D, a 2x32 block in 2D, a 4x4x4 block in 3D, clipped to the box
    with the cells left over given to the slower dimensions), numbered row-major"""
    dim = len(ncells)
    pref = [[64], [1, 32], [4, 4, 4]][dim - 1]
    seg, room = [1] * dim, 64
    for d in range(dim - 1, -1, -1):
        want = room if d == 0 else min(pref[d], room)
        seg[d] = max(1, min(want, ncells[d]))
        room //= seg[d]
    nseg = [(ncells[d] + seg[d] - 1) // seg[d] for d in range(dim)]
    s = np.zeros(len(icell), dtype=np.int64)
    for d in range(dim):
        s = s * nseg[d] + icell[:, d] // seg[d]
    return s


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_sort_groups_particles_by_segment(dim):
    """phx_sort_particles keeps the particle set bit for bit and stores the particles segment
    by segment, in segment order"""
    _say(f"[gpu] sort dim={dim}")
    rng = np.random.default_rng(11 + dim)
    ncells = [[300], [40, 100], [12, 10, 9]][dim - 1]
    g = phx.Level(dim, 2, ncells, [0.3] * dim)
    p = U.random_particles(rng, dim, 7000, ncells, [0] * dim, 0)
    g.set_particles(0, phx.DOMAIN, p)
    g.sort_particles()
    assert g.count() == 7000
    q = g.get_particles()
    U.assert_same_particles(q, p, exact=True)
    assert np.all(np.diff(_segment_of(q["icell"], ncells)) >= 0)
    # a second sort copies the sorted slots back into the same storage
    g.sort_particles()
    q = g.get_particles()
    U.assert_same_particles(q, p, exact=True)
    assert np.all(np.diff(_segment_of(q["icell"], ncells)) >= 0)
    g.close()


def _crowded_particles(rng, ncells):
    """two crowded cells in one segment, a crowded cell in the periodic corner, a sparse band of
    columns, every other segment empty"""
    cells = [((3, 70), 3000), ((3, 71), 700), ((ncells[0] - 1, ncells[1] - 1), 800)]
    ic = [np.tile(np.array(c, np.int32), (n, 1)) for c, n in cells]
    band = np.stack([rng.integers(10, 13, 500), rng.integers(0, ncells[1], 500)], 1)
    ic = np.concatenate(ic + [band.astype(np.int32)])
    n = len(ic)
    return dict(icell=ic, delta=rng.random((n, 2)), v=rng.standard_normal((n, 3)),
                weight=0.5 + rng.random(n), charge=np.ones(n))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sorted_push_crowded_cells_and_empty_segments(precision):
    """cell-segmented storage at its extremes: a crowded cell fills most lanes of its segment's
    rows (same-cell LDS deposits within a row), segments with no particle, a crowded cell on the
    periodic corner; both moveIons modes under uniform fields against the oracle (FP64 1e-12,
    mixed precision 1e-5 on the moments)"""
    _say(f"[gpu] sorted push, crowded cells and empty segments, {precision}")
    ncells, dx, dt = (20, 150), (0.4, 0.4), 0.05
    prec = phx.PRECISION_FP32 if precision == "fp32" else phx.PRECISION_FP64
    rng = np.random.default_rng(31)
    p = _crowded_particles(rng, ncells)
    g = phx.Level(2, 2, ncells, dx, capacity=len(p["weight"]), precision=prec)
    o = O.Sim(2, 2, ncells, dx)
    E, B = (0.01, -0.02, 0.005), (0.5, 0.1, 1.0)
    for sim, fe, fb in ((g, phx.F_EAVG, phx.F_BAVG), (o, O.F_EAVG, O.F_BAVG)):
        sim.set_particles(0, phx.DOMAIN, p)
        for c in range(3):
            sim.set(fe + c, np.full(sim.shape(fe + c), E[c]))
            sim.set(fb + c, np.full(sim.shape(fb + c), B[c]))
    g.sort_particles()
    rtol = U.RTOL if precision == "fp64" else 1e-5
    for mode in (phx.MODE_DOMAIN_ONLY, phx.MODE_ALL):
        g.move_ions(dt, mode)
        o.move_ions(dt, mode)
        for k in range(5):
            U.assert_close(g.get(phx.f_pop(0, k)), o.get(O.f_pop(0, k)), rtol,
                           what=f"mode {mode} moment {k}")
    assert g.count() == o.count() == len(p["weight"])
    if precision == "fp64":
        U.assert_same_particles(g.get_particles(), o.get_particles(), U.RTOL)
    g.close()


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_sort_puts_particles_outside_the_domain_last(dim):
    """particles in the ghost layer (outside the domain box) go to the overflow segment after
    every domain segment; the particle set is kept bit for bit, through a second sort as well
    (which copies the sorted slots back into the same storage)"""
    _say(f"[gpu] sort with particles outside the domain dim={dim}")
    rng = np.random.default_rng(41 + dim)
    ncells = [[300], [40, 100], [12, 10, 9]][dim - 1]
    g = phx.Level(dim, 2, ncells, [0.3] * dim)
    p = U.random_particles(rng, dim, 9000, ncells, [0] * dim, 1)
    outside = np.any((p["icell"] < 0) | (p["icell"] >= np.array(ncells)), axis=1)
    assert outside.any() and not outside.all()
    g.set_particles(0, phx.DOMAIN, p)
    for _ in range(2):
        g.sort_particles()
        q = g.get_particles()
        U.assert_same_particles(q, p, exact=True)
        out_q = np.any((q["icell"] < 0) | (q["icell"] >= np.array(ncells)), axis=1)
        n_in = int((~out_q).sum())
        assert n_in == int((~outside).sum())
        assert not out_q[:n_in].any() and out_q[n_in:].all()
        assert np.all(np.diff(_segment_of(q["icell"][:n_in], ncells)) >= 0)
    g.close()


@pytest.mark.parametrize("interval", [1, 3])
def test_advance_with_frequent_sorts(interval):
    """sorting every `interval` PPC steps (the first sort moves the linear upload into segmented
    storage, later ones copy the sorted slots back into it) leaves the run equal to the oracle's"""
    _say(f"[gpu] advance harris2d sorted every {interval} steps")
    wl = W.harris2d(ncells=(32, 24), ppc=12, order=2)
    g = phx.Level(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te,
                  capacity=wl.n_particles)
    o = O.Sim(wl.dim, wl.order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te)
    g.set_sort_interval(interval)
    wl.setup(g, phx.F_B)
    wl.setup(o, O.F_B)
    g.advance(wl.dt, 12)
    o.advance(wl.dt, 12)
    g.sync()
    _check_state(g, o, wl)
    g.close()
