"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was built, to the
reference's own header-only numerics (oracle/_ref/libphare_ref.so).

Test infrastructure only: the product (phare_amd/) never imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libphare_ref.so")

# field ids shared with include/phx.h
F_E, F_B, F_EPRED, F_BPRED, F_EAVG, F_BAVG, F_J, F_VE, F_PE, F_RHO, F_MASSD, F_VI = (
    0, 3, 6, 9, 12, 15, 18, 21, 24, 25, 26, 27)
F_POP0 = 32


def f_pop(pop, k):
    """field id of population `pop` moment k (0 n, 1 rho, 2..4 flux x,y,z)"""
    return F_POP0 + 8 * pop + k


# quantity ids of the stand-alone operators: 0..2 B, 3..5 E, 6..8 J, 9 moments
Q_BX, Q_BY, Q_BZ, Q_EX, Q_EY, Q_EZ, Q_JX, Q_JY, Q_JZ, Q_MOM = range(10)

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def _ptrs(arrs):
    return (C.POINTER(C.c_double) * len(arrs))(*[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])


def _i3(v):
    v = list(v) + [0] * (3 - len(v))
    return np.array(v, dtype=np.int32)


def _d3(v):
    v = list(v) + [1.0] * (3 - len(v))
    return np.array(v, dtype=np.float64)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run `make -C oracle`)")
        L = C.CDLL(ORACLE_SO)
        L.ora_create.restype = C.c_void_p
        L.ora_create.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, C.c_int, C.c_int, _dp,
                                 C.c_double, C.c_double, C.c_double, C.c_int]
        L.ora_destroy.argtypes = [C.c_void_p]
        for name in ("ora_field_shape",):
            getattr(L, name).argtypes = [C.c_void_p, C.c_int, _ip]
        L.ora_set_field.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ora_get_field.argtypes = [C.c_void_p, C.c_int, _dp]
        L.ora_num_particles.restype = C.c_long
        L.ora_num_particles.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.ora_set_particles.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_long, _ip, _dp, _dp, _dp, _dp]
        L.ora_get_particles.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _dp, _dp, _dp, _dp]
        L.ora_load_maxwellian.argtypes = [C.c_void_p, C.c_int] + [_dp] * 7 + [
            C.c_double, C.c_uint, C.c_ulonglong, C.c_double]
        for name in ("ora_initialize", "ora_update_ions", "ora_fill_moment_borders",
                     "ora_fill_ion_borders", "ora_electrons", "ora_average"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.ora_update_populations.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int]
        L.ora_move_ions.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.ora_advance.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.ora_faraday.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double]
        L.ora_ampere.argtypes = [C.c_void_p, C.c_int]
        L.ora_ohm.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.ora_fill_ghosts.argtypes = [C.c_void_p, C.c_int]
        L.ora_gather.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.ora_push_error.argtypes = [C.c_void_p, _dp]
        L.ora_push.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _ip, _ip,
                               C.c_int, _ip, _ip, _lp]
        L.ora_deposit.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double]
        L.ora_diagnostics.argtypes = [C.c_void_p, _dp]
        L.ora_weights.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_int), _dp]
        L.ora_layout_shape.argtypes = [C.c_int, C.c_int, _ip, C.c_int, _ip]
        L.ora_s_deposit.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, C.c_long, _ip, _dp, _dp, _dp,
                                    _dp, C.c_double, _dp, _dp, C.c_void_p]
        L.ora_s_laplacian.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, C.c_int, _dp, _dp]
        L.ora_s_project.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.ora_s_faraday.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_double]
        L.ora_s_ampere.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, C.c_void_p, C.c_void_p]
        L.ora_partition.restype = C.c_long
        L.ora_partition.argtypes = [C.c_int, _ip, _ip, C.c_long, _ip, _ip, _ip, _ip, _lp]
        L.ora_boris_trajectory.argtypes = [C.c_int, C.c_long, C.c_double, _dp, C.c_double, C.c_double,
                                           _dp, _dp, _ip, _dp, _dp, _dp]
        L.ora_refine_field.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, _ip, _dp, _ip, _ip, _ip, _ip]
        L.ora_magnetic_new_faces.argtypes = [C.c_int, _dp, _ip, _dp, _ip, _dp, _ip, _ip, _dp, _ip,
                                             _ip]
        L.ora_split.restype = C.c_long
        L.ora_split.argtypes = [C.c_int, C.c_int, C.c_int, C.c_long, _ip, _dp, _dp, _dp, _dp, C.c_int,
                                _ip, _ip, C.c_long, _ip, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def ref():
    """the reference's own numerics, or None where oracle/_ref was not built"""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_SO):
            return None
        R = C.CDLL(REF_SO)
        R.ref_layout_info.argtypes = [C.c_int, C.c_int, _ip, C.c_int, _ip]
        R.ref_weights.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_int), _dp]
        R.ref_gather.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, C.c_void_p, C.c_void_p, C.c_long,
                                 _ip, _dp, _dp, _dp]
        R.ref_deposit.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, C.c_long, _ip, _dp, _dp, _dp, _dp,
                                  C.c_double, _dp, _dp, C.c_void_p]
        R.ref_faraday.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_double]
        R.ref_ampere.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, C.c_void_p, C.c_void_p]
        R.ref_laplacian.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, C.c_int, _dp, _dp]
        R.ref_project.argtypes = [C.c_int, C.c_int, _ip, _dp, _ip, C.c_int, C.c_int, C.c_int, _dp, _dp]
        R.ref_load_maxwellian.restype = C.c_long
        R.ref_load_maxwellian.argtypes = [C.c_int, _ip, _ip, _dp, C.c_void_p, C.c_void_p, C.c_double,
                                          C.c_uint, C.c_ulonglong, C.c_double, C.c_long, _ip, _dp,
                                          _dp, _dp, _dp]
        R.ref_partition.restype = C.c_long
        R.ref_partition.argtypes = [C.c_int, _ip, _ip, C.c_long, _ip, _ip, _ip, _ip, _lp]
        R.ref_split_one.restype = C.c_long
        R.ref_split_one.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _dp, _dp, C.c_double, C.c_double,
                                    _ip, _dp, _dp, _dp, _dp]
        R.ref_split_max_cell_distance.argtypes = [C.c_int]
        _ref = R
    return _ref


def weights(order, dual, icell_local, delta, use_ref=False):
    L = ref() if use_ref else lib()
    start = C.c_int(0)
    w = np.zeros(4)
    (L.ref_weights if use_ref else L.ora_weights)(order, int(dual), int(icell_local), float(delta),
                                                  C.byref(start), w)
    return start.value, w[: order + 1].copy()


def layout_shape(dim, order, ncells, q):
    s = np.zeros(3, dtype=np.int32)
    lib().ora_layout_shape(dim, order, _i3(ncells), q, s)
    return tuple(int(x) for x in s[:dim])


class Sim:
    """One periodic patch driven by the oracle's SolverPPC restatement."""

    def __init__(self, dim, order, ncells, dx, lower=None, masses=(1.0,), Te=0.1, eta=0.0, nu=1e-4,
                 hyper_mode=0, level=0):
        self.dim, self.order = dim, order
        self.ncells = list(ncells)
        self.dx = list(dx)
        lower = lower or [0] * dim
        self.npop = len(masses)
        self.h = lib().ora_create(dim, order, _i3(ncells), _d3(dx), _i3(lower), level, len(masses),
                                  np.array(masses, dtype=np.float64), Te, eta, nu, hyper_mode)
        if not self.h:
            raise RuntimeError("ora_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_destroy(self.h)
            self.h = None

    def shape(self, fid):
        s = np.zeros(3, dtype=np.int32)
        if lib().ora_field_shape(self.h, fid, s):
            raise KeyError(fid)
        return tuple(int(x) for x in s[: self.dim])

    def get(self, fid):
        a = np.zeros(self.shape(fid))
        lib().ora_get_field(self.h, fid, a.reshape(-1))
        return a

    def set(self, fid, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.shape(fid), (a.shape, self.shape(fid))
        lib().ora_set_field(self.h, fid, a.reshape(-1))

    def set_particles(self, pop, kind, p):
        n = len(p["weight"])
        lib().ora_set_particles(self.h, pop, kind, n,
                                np.ascontiguousarray(p["icell"], dtype=np.int32).reshape(-1),
                                np.ascontiguousarray(p["delta"], dtype=np.float64).reshape(-1),
                                np.ascontiguousarray(p["v"], dtype=np.float64).reshape(-1),
                                np.ascontiguousarray(p["weight"], dtype=np.float64),
                                np.ascontiguousarray(p["charge"], dtype=np.float64))

    def count(self, pop=0, kind=0):
        return int(lib().ora_num_particles(self.h, pop, kind))

    def get_particles(self, pop=0, kind=0):
        n = self.count(pop, kind)
        p = dict(icell=np.zeros((n, self.dim), dtype=np.int32), delta=np.zeros((n, self.dim)),
                 v=np.zeros((n, 3)), weight=np.zeros(n), charge=np.zeros(n))
        if n:
            lib().ora_get_particles(self.h, pop, kind, p["icell"].reshape(-1), p["delta"].reshape(-1),
                                    p["v"].reshape(-1), p["weight"], p["charge"])
        return p

    def load_maxwellian(self, pop, n, V, Vth, charge, ppc, seed, cutoff=1e-16):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in [n, *V, *Vth]]
        rc = lib().ora_load_maxwellian(self.h, pop, *arrs, charge, ppc, seed, cutoff)
        assert rc == 0

    def _rc(self, rc):
        if rc:
            e = np.zeros(4)
            lib().ora_push_error(self.h, e)
            raise RuntimeError(f"oracle error {rc}: {e}")

    def initialize(self):
        self._rc(lib().ora_initialize(self.h))

    def update_populations(self, dt, mode, emset=2):
        self._rc(lib().ora_update_populations(self.h, emset, dt, mode))

    def update_ions(self):
        self._rc(lib().ora_update_ions(self.h))

    def fill_moment_borders(self):
        self._rc(lib().ora_fill_moment_borders(self.h))

    def fill_ion_borders(self):
        self._rc(lib().ora_fill_ion_borders(self.h))

    def move_ions(self, dt, mode):
        self._rc(lib().ora_move_ions(self.h, dt, mode))

    def advance(self, dt, nsteps=1):
        self._rc(lib().ora_advance(self.h, dt, nsteps))

    def faraday(self, bset, eset, bnewset, dt):
        self._rc(lib().ora_faraday(self.h, bset, eset, bnewset, dt))

    def ampere(self, bset):
        self._rc(lib().ora_ampere(self.h, bset))

    def electrons(self):
        self._rc(lib().ora_electrons(self.h))

    def ohm(self, bset, eset):
        self._rc(lib().ora_ohm(self.h, bset, eset))

    def average(self):
        self._rc(lib().ora_average(self.h))

    def fill_ghosts(self, fid):
        self._rc(lib().ora_fill_ghosts(self.h, fid))

    def gather(self, emset, pop=0, kind=0):
        n = self.count(pop, kind)
        E = np.zeros((n, 3))
        B = np.zeros((n, 3))
        lib().ora_gather(self.h, emset, pop, kind, E.reshape(-1), B.reshape(-1))
        return E, B

    def push(self, dt, first=(), second=(), pop=0, kind=0, emset=2):
        """BorisPusher::move in place with box-list selectors (empty: noop); returns
        (end of the second selection, end of the first)"""
        def boxes(bl):
            lo = np.zeros(3 * max(1, len(bl)), np.int32)
            hi = np.zeros(3 * max(1, len(bl)), np.int32)
            for b, (l, h) in enumerate(bl):
                lo[3 * b:3 * b + len(l)] = l
                hi[3 * b:3 * b + len(h)] = h
            return len(bl), lo, hi
        nf, flo, fhi = boxes(first)
        ns, slo, shi = boxes(second)
        out = np.zeros(2, np.int64)
        self._rc(lib().ora_push(self.h, emset, pop, kind, dt, nf, flo, fhi, ns, slo, shi, out))
        return int(out[0]), int(out[1])

    def deposit(self, pop=0, kind=0, coef=1.0):
        self._rc(lib().ora_deposit(self.h, pop, kind, coef))

    def diagnostics(self):
        out = np.zeros(4)
        lib().ora_diagnostics(self.h, out)
        return dict(kinetic=out[0], magnetic=out[1], charge=out[2], count=out[3])


def ref_load_maxwellian(ncells, lower, n, V, Vth, charge, ppc, seed, cutoff=1e-16):
    """the loader's particles from the reference's own draws (oracle/_ref), in load order"""
    R = ref()
    dim = len(ncells)
    nc = np.ascontiguousarray(ncells, np.int32)
    lo = np.ascontiguousarray(lower, np.int32)
    dens = np.ascontiguousarray(n, np.float64).reshape(-1)
    Va = [np.ascontiguousarray(a, np.float64).reshape(-1) for a in V]
    Ta = [np.ascontiguousarray(a, np.float64).reshape(-1) for a in Vth]
    Vp = (C.c_void_p * 3)(*[a.ctypes.data for a in Va])
    Tp = (C.c_void_p * 3)(*[a.ctypes.data for a in Ta])
    cap = int(len(dens) * ppc)
    ic = np.zeros(max(1, cap * dim), np.int32)
    de = np.zeros(max(1, cap * dim))
    v = np.zeros(max(1, cap * 3))
    w = np.zeros(max(1, cap))
    q = np.zeros(max(1, cap))
    m = R.ref_load_maxwellian(dim, nc, lo, dens, Vp, Tp, charge, ppc, seed, cutoff, cap, ic, de, v, w, q)
    assert 0 <= m <= cap
    return dict(icell=ic[:m * dim].reshape(m, dim), delta=de[:m * dim].reshape(m, dim),
                v=v[:m * 3].reshape(m, 3), weight=w[:m], charge=q[:m])


def split(dim, order, nbr, p, boxes):
    """the oracle's coarse-to-fine split of particle set p (SoA dict) into the fine AMR boxes
    [(lower, upper), ...]: the refined particles kept, as a SoA dict"""
    n = len(p["weight"])
    lo = np.zeros(3 * len(boxes), np.int32)
    hi = np.zeros(3 * len(boxes), np.int32)
    for i, (a, b) in enumerate(boxes):
        lo[3 * i:3 * i + dim] = a
        hi[3 * i:3 * i + dim] = b
    cap = max(1, n * max(nbr, 6) * len(boxes))
    out = dict(icell=np.zeros((cap, dim), np.int32), delta=np.zeros((cap, dim)), v=np.zeros((cap, 3)),
               weight=np.zeros(cap), charge=np.zeros(cap))
    k = lib().ora_split(dim, order, nbr, n, np.ascontiguousarray(p["icell"], np.int32).reshape(-1),
                        np.ascontiguousarray(p["delta"], np.float64).reshape(-1),
                        np.ascontiguousarray(p["v"], np.float64).reshape(-1),
                        np.ascontiguousarray(p["weight"], np.float64),
                        np.ascontiguousarray(p["charge"], np.float64), len(boxes), lo, hi, cap,
                        out["icell"].reshape(-1), out["delta"].reshape(-1), out["v"].reshape(-1),
                        out["weight"], out["charge"])
    if k < 0:
        raise ValueError(f"no split pattern for dim {dim} order {order} nbr {nbr}")
    return {key: val[:k] for key, val in out.items()}


def split_one(dim, order, nbr, icell, delta, v, w, q, use_ref=False):
    """one particle already on the fine grid split by the oracle's tables (use_ref: by the
    reference's Splitter compiled in oracle/_ref): the refined particles in dispatch order"""
    ic = np.ascontiguousarray(icell, np.int32)
    de = np.ascontiguousarray(delta, np.float64)
    vv = np.ascontiguousarray(v, np.float64)
    if use_ref:
        cap = 32
        out = dict(icell=np.zeros((cap, dim), np.int32), delta=np.zeros((cap, dim)),
                   v=np.zeros((cap, 3)), weight=np.zeros(cap), charge=np.zeros(cap))
        k = ref().ref_split_one(dim, order, nbr, ic, de, vv, float(w), float(q),
                                out["icell"].reshape(-1), out["delta"].reshape(-1),
                                out["v"].reshape(-1), out["weight"], out["charge"])
        return {key: val[:k] for key, val in out.items()}
    # the oracle splits coarse particles: for an even fine cell the coarse particle
    # (cell / 2, delta / 2) maps back onto exactly this fine particle under toFineGrid (halving
    # and doubling are exact in binary; an odd cell would need 1 + delta, which rounds), and one
    # destination box covering everything keeps every refined particle
    if np.any(ic % 2):
        raise ValueError("split_one: the oracle path takes even fine cells")
    cc = ic // 2
    cd = de / 2.0
    p = dict(icell=cc.reshape(1, dim), delta=cd.reshape(1, dim), v=vv.reshape(1, 3),
             weight=np.array([w], float), charge=np.array([q], float))
    big = (tuple([-10 ** 6] * dim), tuple([10 ** 6] * dim))
    return split(dim, order, nbr, p, [big])


def coarsen(fine, fine_origin, coarse, coarse_origin, lower, upper, dual_dir):
    """the fine-to-coarse coarsening of one field, ratio 2, restated from the reference's
    MomentsCoarsener (src/amr/data/field/coarsening/moments_coarsener.hpp:39-75: the coarse node c
    takes the fine node 2c) and ElectricFieldCoarsener (electric_field_coarsener.hpp:38-140: along
    the component's dual direction the mean 0.5 * (f[2c] + f[2c + 1]), along the primal ones the
    fine node 2c). Arrays in local node indices; *_origin: the AMR index of local index 0 per
    direction (lower - ghost width); lower/upper: coarse AMR nodes (inclusive). Returns a copy of
    `coarse` with the box overwritten."""
    out = np.array(coarse, dtype=np.float64, copy=True)
    dim = out.ndim
    ranges = [range(lower[d], upper[d] + 1) for d in range(dim)]
    import itertools

    for c in itertools.product(*ranges):
        f = tuple(2 * c[d] - fine_origin[d] for d in range(dim))
        cl = tuple(c[d] - coarse_origin[d] for d in range(dim))
        if dual_dir < 0:
            out[cl] = fine[f]
        else:
            f1 = tuple(f[d] + (1 if d == dual_dir else 0) for d in range(dim))
            out[cl] = 0.5 * (fine[f] + fine[f1])
    return out


# refine operators of the oracle (values of include/phx.h's PHX_REFINE_*)
REFINE_LINEAR, REFINE_LINEAR_OVERWRITE, REFINE_ELECTRIC, REFINE_MAGNETIC, REFINE_MAGNETIC_INIT = range(5)


def _shape3(a):
    return _i3(list(a.shape) + [1] * (3 - a.ndim))


def refine_field(op, primal, fine, fine_origin, coarse, coarse_origin, lower, upper):
    """the coarse-to-fine refinement of one field, ratio 2 (oracle/refcpu_amr.hpp: the reference's
    DefaultFieldRefiner, FieldMomentsRefiner, ElectricFieldRefiner, MagneticFieldRefiner and
    MagneticFieldInitRefiner). Arrays in local indices; *_origin: the AMR index of local index 0;
    primal: the quantity's centering flags; lower/upper: fine AMR nodes of the quantity
    (inclusive). Returns a copy of `fine` with the box refined."""
    out = np.array(fine, dtype=np.float64, copy=True, order="C")
    dim = out.ndim
    lib().ora_refine_field(dim, op, _i3([1 if p else 0 for p in primal]), out.reshape(-1),
                           _shape3(out), _i3(fine_origin),
                           np.ascontiguousarray(coarse, np.float64).reshape(-1), _shape3(coarse),
                           _i3(coarse_origin), _i3(lower), _i3(upper))
    return out


def planted_oracle(defines, tmpdir):
    """the oracle compiled with mistakes planted through its ORA_PLANT_* macros (defines: name ->
    value), loaded from tmpdir; for tests that show their checks catch those mistakes"""
    import subprocess

    so = os.path.join(str(tmpdir), "liboracle_planted.so")
    cmd = ["g++", "-std=c++20", "-O1", "-fPIC", "-ffp-contract=off", "-shared", "-o", so,
           os.path.join(ROOT, "oracle", "refcpu_capi.cpp")]
    cmd += [f"-D{k}={v}" for k, v in defines.items()]
    subprocess.run(cmd, check=True)
    L = C.CDLL(so)
    L.ora_magnetic_new_faces.argtypes = [C.c_int, _dp, _ip, _dp, _ip, _dp, _ip, _ip, _dp, _ip, _ip]
    return L


def magnetic_new_faces(B, origin, dx, lower, upper, oracle=None):
    """MagneticRefinePatchStrategy::postprocessRefine (oracle/refcpu_amr.hpp) over the fine cell
    box lower..upper (AMR, inclusive): copies of the three B components with the new faces set
    (oracle: a planted_oracle handle instead of the oracle)"""
    out = [np.array(b, dtype=np.float64, copy=True, order="C") for b in B]
    dim = out[0].ndim
    d = np.zeros(3)
    d[:dim] = dx[:dim]
    (oracle or lib()).ora_magnetic_new_faces(dim, out[0].reshape(-1), _shape3(out[0]), out[1].reshape(-1),
                                 _shape3(out[1]), out[2].reshape(-1), _shape3(out[2]), _i3(origin),
                                 d, _i3(lower), _i3(upper))
    return out
