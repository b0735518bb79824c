"""Synthetic workloads of the hybrid-PIC hot path (seeded, reproducible), as per-cell Maxwellian
profiles evaluated at cell centres and magnetic fields evaluated at Yee node positions: what
PHARE's Python input layer hands to MaxwellianParticleInitializer (cell-centred coordinates,
maxwellian_particle_initializer.hpp:149-165) and to Electromag::initialize.

Workloads:
- uniform1d:  1D uniform plasma, 200 cells dx=0.2, n=1, vth=0.3, B=(1,0,0), order 1, 100 ppc,
              dt=1e-3 (the CPU-runnable case);
- harris2d:   2D Harris current sheets, 256x128 cells dx=dy=0.4, order 2, 100 ppc, T=0.5,
              background density 0.2, sheet half-width 1, localized 0.1 perturbation, dt=5e-3.
              The sheet pair and perturbation are the periodic form of
              tests/functional/harris/harris_2d.py:47-105 (one tanh sheet is not periodic);
- uniform3d:  3D uniform Maxwellian, 128^3 cells dx=0.3, vth=0.5, drift (0.2,0,0), B=(1,0.5,0),
              64 ppc, dt=2e-3, orders 1-3.
Sizes can be scaled down (same profiles) for parity tests.
"""
import numpy as np

GHOST = {1: 2, 2: 4, 3: 4}
CENT = {"Bx": "pdd", "By": "dpd", "Bz": "ddp"}


def cell_centres(ncells, dx, lower=None):
    """coordinates of every cell centre, flattened in row-major AMR order (last dim fastest)"""
    lower = lower or [0] * len(ncells)
    axes = [(lower[d] + np.arange(ncells[d]) + 0.5) * dx[d] for d in range(len(ncells))]
    grids = np.meshgrid(*axes, indexing="ij")
    return [g.reshape(-1) for g in grids]


def node_coords(shape, centering, order, dx, lower=None):
    """coordinates of every node of a field array of the given shape and centering string"""
    dim = len(shape)
    lower = lower or [0] * dim
    g = GHOST[order]
    axes = []
    for d in range(dim):
        off = 0.5 if centering[d] == "d" else 0.0
        axes.append((lower[d] + np.arange(shape[d]) - g + off) * dx[d])
    return np.meshgrid(*axes, indexing="ij")


class Workload:
    def __init__(self, name, dim, order, ncells, dx, ppc, dt, nsteps, density, vbulk, vth, bfield,
                 Te=0.1, seed=12334, masses=(1.0,), charges=(1.0,), extra_pops=()):
        self.name, self.dim, self.order = name, dim, order
        self.ncells, self.dx = list(ncells), list(dx)
        self.ppc, self.dt, self.nsteps = ppc, dt, nsteps
        self.density, self.vbulk, self.vth, self.bfield = density, vbulk, vth, bfield
        self.Te, self.seed = Te, seed
        self.masses, self.charges = tuple(masses), tuple(charges)

    @property
    def n_cells(self):
        return int(np.prod(self.ncells))

    @property
    def n_particles(self):
        return self.n_cells * self.ppc * len(self.masses)

    def profiles(self, pop=0):
        X = cell_centres(self.ncells, self.dx)
        n = np.broadcast_to(self.density(*X), X[0].shape).astype(np.float64)
        V = [np.broadcast_to(f(*X), X[0].shape).astype(np.float64) for f in self.vbulk]
        T = [np.broadcast_to(f(*X), X[0].shape).astype(np.float64) for f in self.vth]
        return n, V, T

    def b_arrays(self, shape_of):
        """B components on their Yee nodes; shape_of(component index) -> array shape"""
        out = []
        for c, name in enumerate(("Bx", "By", "Bz")):
            shp = shape_of(c)
            X = node_coords(shp, CENT[name], self.order, self.dx)
            out.append(np.broadcast_to(self.bfield[c](*X), shp).astype(np.float64).copy())
        return out

    def setup(self, sim, F_B):
        """load particles and B into a simulation object (oracle or GPU level), then
        run the root-level initialisation (moments, J, E)"""
        for p in range(len(self.masses)):
            n, V, T = self.profiles(p)
            sim.load_maxwellian(p, n, V, T, self.charges[p], self.ppc, self.seed + p)
        for c, b in enumerate(self.b_arrays(lambda c: sim.shape(F_B + c))):
            sim.set(F_B + c, b)
        sim.initialize()


def _const(v):
    return lambda *X: np.full_like(X[0], v, dtype=np.float64)


def uniform1d(ncells=200, ppc=100, order=1):
    return Workload("uniform1d", 1, order, [ncells], [0.2], ppc, 1e-3, 100, _const(1.0),
                    [_const(0.0)] * 3, [_const(0.3)] * 3, [_const(1.0), _const(0.0), _const(0.0)])


def harris2d(ncells=(256, 128), ppc=100, order=2, dx=(0.4, 0.4)):
    Lx, Ly = ncells[0] * dx[0], ncells[1] * dx[1]
    L, dB, sigma, nb = 1.0, 0.1, 1.0, 0.2

    def density(x, y):
        return nb + 1.0 / np.cosh((y - 0.3 * Ly) / L) ** 2 + 1.0 / np.cosh((y - 0.7 * Ly) / L) ** 2

    def S(y, y0):
        return 0.5 * (1.0 + np.tanh((y - y0) / L))

    def bx(x, y):
        x0, y1, y2 = x - 0.5 * Lx, y - 0.3 * Ly, y - 0.7 * Ly
        d1 = -2 * dB * y1 * np.exp(-(x0 ** 2 + y1 ** 2) / sigma ** 2)
        d2 = 2 * dB * y2 * np.exp(-(x0 ** 2 + y2 ** 2) / sigma ** 2)
        return -1.0 + 2.0 * (S(y, 0.3 * Ly) - S(y, 0.7 * Ly)) + d1 + d2

    def by(x, y):
        x0, y1, y2 = x - 0.5 * Lx, y - 0.3 * Ly, y - 0.7 * Ly
        return (2 * dB * x0 * np.exp(-(x0 ** 2 + y1 ** 2) / sigma ** 2)
                - 2 * dB * x0 * np.exp(-(x0 ** 2 + y2 ** 2) / sigma ** 2))

    return Workload("harris2d", 2, order, list(ncells), list(dx), ppc, 5e-3, 1000, density,
                    [_const(0.0)] * 3, [_const(np.sqrt(0.5))] * 3, [bx, by, _const(0.0)])


def uniform3d(ncells=(128, 128, 128), ppc=64, order=2, dx=0.3):
    return Workload("uniform3d", 3, order, list(ncells), [dx] * 3, ppc, 2e-3, 200, _const(1.0),
                    [_const(0.2), _const(0.0), _const(0.0)], [_const(0.5)] * 3,
                    [_const(1.0), _const(0.5), _const(0.0)])
