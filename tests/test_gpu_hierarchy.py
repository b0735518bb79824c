"""The two-level subcycled hierarchy (phare_amd/hierarchy.py: MultiPhysicsIntegrator's advance and
synchronisation over a periodic coarser level and one finer patch, the messenger's fills done by
the refine/coarsen entry points), in 1D, 2D and 3D (the 3D coarser level on the staged path). No
oracle restates the reference's driver (SAMRAI schedules):
these tests pin what the mathematics fixes ("parity unpinned" for the driver's sequence itself):
 - a uniform plasma refined into the finer patch keeps its density everywhere on the patch, up to
   the coarse-fine boundary (without the level ghost particles and their alpha-weighted moments
   the boundary nodes would lose about half their density), through coarse steps with subcycling;
 - nothing on the finer patch's physical nodes becomes NaN (every ghost node the stages read is
   filled by a refine operator);
 - a divergence-free coarse B refines to a divergence-free fine B, and both levels keep div B at
   round-off through the steps (Faraday from a single E on each level; the reflux recomputes the
   coarser B from Bold and the coarsened flux sum);
 - the total charge of the coarser level's particles is conserved."""
import sys

import numpy as np
import pytest

from phare_amd import hierarchy as H
from phare_amd import phx
from phare_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _physical(lvl, fid, primal_all=True):
    """the physical nodes of a moment-like (all-primal) field"""
    g = H.GHOSTS[lvl.order]
    a = lvl.get(fid)
    sl = tuple(slice(g, g + n + 1) for n in lvl.ncells)
    return a[sl]


def _uniform(dim, order, ncells, ppc):
    if dim == 1:
        wl = W.uniform1d(ncells=ncells[0], ppc=ppc, order=order)
    else:
        wl = W.Workload("uniform2d", 2, order, list(ncells), [0.3, 0.3], ppc, 2e-3, 10,
                        lambda x, y: np.ones_like(x), [lambda x, y: np.zeros_like(x)] * 3,
                        [lambda x, y: np.full_like(x, 0.3)] * 3,
                        [lambda x, y: np.ones_like(x), lambda x, y: np.zeros_like(x),
                         lambda x, y: np.zeros_like(x)])
    coarse = phx.Level(dim, order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te,
                       capacity=wl.n_particles, apply_tuning_table=False)
    wl.setup(coarse, phx.F_B)
    return wl, coarse


@pytest.mark.parametrize("dim,order", [(1, 1), (1, 2), (2, 1)])
def test_uniform_plasma_keeps_its_density_on_the_finer_patch(dim, order):
    _say(f"[gpu] two-level uniform plasma dim={dim} order={order}")
    ncells = [[64], [24, 20]][dim - 1]
    ppc = [200, 64][dim - 1]
    wl, coarse = _uniform(dim, order, ncells, ppc)
    clo = [[8], [6, 5]][dim - 1]
    chi = [[40], [17, 14]][dim - 1]
    h = H.TwoLevel(coarse, clo, chi)
    h.initialize()
    q0 = coarse.diagnostics()["charge"]
    fine = h.fine
    tol = [0.2, 0.3][dim - 1]
    for step in range(4):
        rho = _physical(fine, phx.F_RHO)
        assert not np.any(np.isnan(rho)), f"NaN density after {step} steps"
        assert np.abs(rho - 1.0).max() < tol, (step, np.abs(rho - 1.0).max())
        # the boundary nodes: no systematic loss (mean over the two outer node layers)
        edge = np.concatenate([rho[:2].ravel(), rho[-2:].ravel()]) if dim == 1 else \
            np.concatenate([rho[:2, :].ravel(), rho[-2:, :].ravel(), rho[:, :2].ravel(), rho[:, -2:].ravel()])
        assert abs(edge.mean() - 1.0) < 0.06, (step, edge.mean())
        for fid in list(range(0, 6)) + [phx.F_J, phx.F_J + 1, phx.F_J + 2]:
            a = fine.get(fid)
            g = H.GHOSTS[order]
            sl = tuple(slice(g, a.shape[d] - g) for d in range(dim))
            assert not np.any(np.isnan(a[sl])), (step, fid)
        h.advance(wl.dt)
    assert abs(coarse.diagnostics()["charge"] - q0) <= 1e-12 * abs(q0)
    _say("  ok")
    h.close()
    coarse.close()


@pytest.mark.parametrize("order", [1, 2])
def test_uniform_plasma_keeps_its_density_on_a_3d_finer_patch(order):
    """3D: a periodic 16^3 coarser level (the staged path) and a 16^3-cell finer patch over coarse
    cells 4..11: the refined uniform plasma keeps its density on the patch (mean over the patch,
    mean over its two outer node layers, every node within the particle noise), no NaN, and the
    coarser level's charge is conserved, over coarse steps with subcycling"""
    _say(f"[gpu] two-level uniform plasma 3D order={order}")
    wl = W.uniform3d(ncells=(16, 16, 16), ppc=16, order=order)
    coarse = phx.Level(3, order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te,
                       capacity=wl.n_particles, apply_tuning_table=False)
    wl.setup(coarse, phx.F_B)
    h = H.TwoLevel(coarse, [4, 4, 4], [11, 11, 11])
    h.initialize()
    q0 = coarse.diagnostics()["charge"]
    fine = h.fine
    g = H.GHOSTS[order]
    for step in range(3):
        rho = _physical(fine, phx.F_RHO)
        assert not np.any(np.isnan(rho)), f"NaN density after {step} steps"
        assert abs(rho.mean() - 1.0) < 0.03, (step, rho.mean())
        assert np.abs(rho - 1.0).max() < 0.6, (step, np.abs(rho - 1.0).max())
        edge = np.concatenate([rho[:2].ravel(), rho[-2:].ravel(), rho[:, :2].ravel(), rho[:, -2:].ravel(),
                               rho[:, :, :2].ravel(), rho[:, :, -2:].ravel()])
        assert abs(edge.mean() - 1.0) < 0.06, (step, edge.mean())
        for fid in list(range(0, 6)) + [phx.F_J, phx.F_J + 1, phx.F_J + 2]:
            a = fine.get(fid)
            sl = tuple(slice(g, a.shape[d] - g) for d in range(3))
            assert not np.any(np.isnan(a[sl])), (step, fid)
        h.advance(wl.dt)
    assert abs(coarse.diagnostics()["charge"] - q0) <= 1e-12 * abs(q0)
    _say("  ok")
    h.close()
    coarse.close()


def _div(B, dx, dim, g):
    """discrete divergence of a Yee B over the physical cells"""
    n = [B[1].shape[0] - 2 * g] if dim == 1 else [B[1].shape[0] - 2 * g, B[0].shape[1] - 2 * g]
    div = 0.0
    for d in range(dim):
        a = [slice(g, g + n[k]) for k in range(dim)]
        b = list(a)
        b[d] = slice(g + 1, g + n[d] + 1)
        div = div + (B[d][tuple(b)] - B[d][tuple(a)]) / dx[d]
    return div


def test_divergence_free_b_on_both_levels():
    """2D: a divergence-free coarse B (the discrete curl of a smooth A_z) with a uniform plasma:
    the finer B from the init refiner and the new faces is divergence-free, and both levels keep
    div B at round-off over coarse steps with subcycling and reflux"""
    _say("[gpu] two-level div B")
    dim, order = 2, 1
    wl, coarse = _uniform(dim, order, [24, 20], 32)
    g = H.GHOSTS[order]
    dx = coarse.dx
    # A_z at primal-primal nodes; Bx = dAz/dy, By = -dAz/dx on the Yee faces
    sx, sy = coarse.shape(phx.F_B + 1)[0], coarse.shape(phx.F_B)[1]
    X = (np.arange(sx + 1) - g) * dx[0]
    Y = (np.arange(sy + 1) - g) * dx[1]
    Lx, Ly = 24 * dx[0], 20 * dx[1]
    A = 0.05 * np.sin(2 * np.pi * X[:, None] / Lx) * np.cos(2 * np.pi * Y[None, :] / Ly)
    bx = (A[:, 1:] - A[:, :-1]) / dx[1]
    by = -(A[1:, :] - A[:-1, :]) / dx[0]
    B0 = [coarse.get(phx.F_B + k) for k in range(3)]
    B0[0] = B0[0] + bx[:B0[0].shape[0], :B0[0].shape[1]]
    B0[1] = B0[1] + by[:B0[1].shape[0], :B0[1].shape[1]]
    for k in range(2):
        coarse.set(phx.F_B + k, B0[k])
    coarse.initialize()
    h = H.TwoLevel(coarse, [6, 5], [17, 14])
    h.initialize()
    fine = h.fine
    for step in range(3):
        for lvl, name in ((coarse, "coarse"), (fine, "fine")):
            B = [lvl.get(phx.F_B + k) for k in range(3)]
            div = _div(B, lvl.dx, dim, g)
            scale = max(np.abs(B[0]).max(), np.abs(B[1]).max()) / min(lvl.dx)
            assert np.abs(div).max() <= 1e-12 * scale, (name, step, np.abs(div).max())
        h.advance(wl.dt)
    h.close()
    coarse.close()


def _div3(B, dx, g):
    """discrete divergence of a 3D Yee B over the physical cells"""
    n = [B[0].shape[0] - 2 * g - 1, B[1].shape[1] - 2 * g - 1, B[2].shape[2] - 2 * g - 1]
    div = 0.0
    for d in range(3):
        a = [slice(g, g + n[k]) for k in range(3)]
        b = list(a)
        b[d] = slice(g + 1, g + n[d] + 1)
        div = div + (B[d][tuple(b)] - B[d][tuple(a)]) / dx[d]
    return div


def test_divergence_free_b_on_both_levels_3d():
    """3D: a coarse B made divergence-free as the discrete curl of A = (Ax(y, z), 0, Az(x, y)) on
    the Yee edges, with a uniform plasma: the finer B from the init refiner and the 3D new faces is
    divergence-free, and both levels keep div B at round-off over coarse steps with subcycling and
    reflux"""
    _say("[gpu] two-level div B 3D")
    order = 1
    wl = W.uniform3d(ncells=(16, 16, 16), ppc=8, order=order)
    coarse = phx.Level(3, order, wl.ncells, wl.dx, masses=wl.masses, Te=wl.Te,
                       capacity=wl.n_particles, apply_tuning_table=False)
    wl.setup(coarse, phx.F_B)
    g = H.GHOSTS[order]
    dx = coarse.dx
    L = [16 * h for h in dx]

    def prim(d, i):
        return (i - g) * dx[d]

    def Ax(y, z):
        return 0.04 * np.sin(2 * np.pi * y / L[1]) * np.cos(2 * np.pi * z / L[2])

    def Az(x, y):
        return 0.05 * np.cos(2 * np.pi * x / L[0]) * np.sin(2 * np.pi * y / L[1])

    B = [coarse.get(phx.F_B + k) for k in range(3)]
    # Bx (primal, dual, dual) = dAz/dy
    i, j, k = np.meshgrid(*[np.arange(s) for s in B[0].shape], indexing="ij")
    B[0] = B[0] + (Az(prim(0, i), prim(1, j + 1)) - Az(prim(0, i), prim(1, j))) / dx[1]
    # By (dual, primal, dual) = dAx/dz - dAz/dx
    i, j, k = np.meshgrid(*[np.arange(s) for s in B[1].shape], indexing="ij")
    B[1] = B[1] + (Ax(prim(1, j), prim(2, k + 1)) - Ax(prim(1, j), prim(2, k))) / dx[2] \
        - (Az(prim(0, i + 1), prim(1, j)) - Az(prim(0, i), prim(1, j))) / dx[0]
    # Bz (dual, dual, primal) = -dAx/dy
    i, j, k = np.meshgrid(*[np.arange(s) for s in B[2].shape], indexing="ij")
    B[2] = B[2] - (Ax(prim(1, j + 1), prim(2, k)) - Ax(prim(1, j), prim(2, k))) / dx[1]
    for c in range(3):
        coarse.set(phx.F_B + c, B[c])
    coarse.initialize()
    h = H.TwoLevel(coarse, [4, 4, 4], [11, 11, 11])
    h.initialize()
    fine = h.fine
    for step in range(3):
        for lvl, name in ((coarse, "coarse"), (fine, "fine")):
            Bl = [lvl.get(phx.F_B + c) for c in range(3)]
            div = _div3(Bl, lvl.dx, g)
            scale = max(np.abs(b).max() for b in Bl) / min(lvl.dx)
            assert np.abs(div).max() <= 1e-12 * scale, (name, step, np.abs(div).max())
        h.advance(wl.dt)
    h.close()
    coarse.close()
