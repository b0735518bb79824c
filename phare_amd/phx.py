"""ctypes binding of include/phx.h (libphx.so, built by phare_amd/build.py).

Loading fails loudly when the HIP library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphx.so")

PHX_OK = 0
PRECISION_FP64, PRECISION_FP32 = 0, 1
MODE_DOMAIN_ONLY, MODE_ALL = 1, 2
DOMAIN, PATCH_GHOST, LEVEL_GHOST, LEVEL_GHOST_OLD, LEVEL_GHOST_NEW = 0, 1, 2, 3, 4
EM, EM_PRED, EM_AVG = 0, 1, 2
F_E, F_B, F_EPRED, F_BPRED, F_EAVG, F_BAVG = 0, 3, 6, 9, 12, 15
F_J, F_VE, F_PE, F_RHO, F_MASSD, F_VI, F_POP0 = 18, 21, 24, 25, 26, 27, 32
F_BOLD, F_FLUXSUM = 64, 67
COARSEN_MOMENTS, COARSEN_ELECTRIC = 0, 1
REFINE_LINEAR, REFINE_LINEAR_OVERWRITE, REFINE_ELECTRIC, REFINE_MAGNETIC, REFINE_MAGNETIC_INIT = range(5)
MAX_POP = 4


def f_pop(pop, k):
    return F_POP0 + 8 * pop + k


class PhxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"phx status {status}: {msg}")
        self.status = status


class Config(C.Structure):
    _fields_ = [("dim", C.c_int), ("interp_order", C.c_int), ("precision", C.c_int),
                ("n_pop", C.c_int), ("mass", C.c_double * MAX_POP), ("eta", C.c_double),
                ("nu", C.c_double), ("hyper_mode", C.c_int), ("Te", C.c_double)]


class Patch(C.Structure):
    _fields_ = [("lower", C.c_int * 3), ("upper", C.c_int * 3), ("dx", C.c_double * 3),
                ("level", C.c_int), ("periodic", C.c_int), ("n_nonlevel_ghost_boxes", C.c_int),
                ("nonlevel_ghost_lower", (C.c_int * 3) * 27),
                ("nonlevel_ghost_upper", (C.c_int * 3) * 27)]


class Tuning(C.Structure):
    _fields_ = [("segment_shape", C.c_int * 3), ("rows_per_item", C.c_int),
                ("tile_row_pad", C.c_int), ("sort_interval", C.c_int), ("tile_copies", C.c_int),
                ("deposit_window", C.c_int)]


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library missing: {LIB_PATH} (run `python -m phare_amd.build`)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.phx_create.argtypes = [C.POINTER(Config), C.POINTER(Patch), C.c_size_t, C.POINTER(vp)]
        L.phx_destroy.argtypes = [vp]
        L.phx_destroy.restype = None
        L.phx_last_error.argtypes = [vp]
        L.phx_last_error.restype = C.c_char_p
        L.phx_synchronize.argtypes = [vp]
        L.phx_stream.argtypes = [vp]
        L.phx_stream.restype = vp
        L.phx_field_shape.argtypes = [vp, C.c_int, _ip]
        L.phx_set_field.argtypes = [vp, C.c_int, _dp]
        L.phx_get_field.argtypes = [vp, C.c_int, _dp]
        L.phx_set_particles.argtypes = [vp, C.c_int, C.c_int, C.c_size_t, _ip, _dp, _dp, _dp, _dp]
        L.phx_particle_count.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_size_t)]
        L.phx_get_particles.argtypes = [vp, C.c_int, C.c_int, _ip, _dp, _dp, _dp, _dp]
        L.phx_set_particles_aos.argtypes = [vp, C.c_int, C.c_int, C.c_size_t, vp]
        L.phx_get_particles_aos.argtypes = [vp, C.c_int, C.c_int, vp]
        L.phx_load_maxwellian.argtypes = [vp, C.c_int] + [_dp] * 7 + [C.c_double, C.c_uint32,
                                                                      C.c_uint64, C.c_double]
        L.phx_update_populations.argtypes = [vp, C.c_int, C.c_double, C.c_int]
        for n in ("phx_update_ions", "phx_fill_moment_borders", "phx_fill_ion_borders",
                  "phx_fill_ion_ghost_particles", "phx_electrons_update", "phx_average",
                  "phx_initialize", "phx_sort_particles"):
            getattr(L, n).argtypes = [vp]
        L.phx_set_sort_interval.argtypes = [vp, C.c_int]
        L.phx_fill_ghosts.argtypes = [vp, C.c_int]
        L.phx_faraday.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double]
        L.phx_ampere.argtypes = [vp, C.c_int]
        L.phx_ohm.argtypes = [vp, C.c_int, C.c_int]
        L.phx_move_ions.argtypes = [vp, C.c_double, C.c_int]
        L.phx_advance.argtypes = [vp, C.c_double, C.c_int]
        L.phx_diagnostics.argtypes = [vp, _dp]
        L.phx_profile_push.argtypes = [vp, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.phx_set_push_timing.argtypes = [vp, C.c_int]
        L.phx_push_timing.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.phx_create_level.argtypes = [C.POINTER(Config), C.POINTER(Patch), C.c_int, _ip, _ip,
                                       C.c_size_t, C.POINTER(vp)]
        L.phx_patch_box.argtypes = [vp, C.c_int, _ip, _ip]
        L.phx_patch_nonlevel_ghost_boxes.argtypes = [vp, C.c_int, C.POINTER(C.c_int), _ip, _ip]
        L.phx_patch_field_shape.argtypes = [vp, C.c_int, C.c_int, _ip]
        L.phx_set_patch_field.argtypes = [vp, C.c_int, C.c_int, _dp]
        L.phx_get_patch_field.argtypes = [vp, C.c_int, C.c_int, _dp]
        L.phx_get_patch_particles.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_size_t), vp, vp,
                                              vp, vp, vp]
        L.phx_add_particles.argtypes = [vp, C.c_int, C.c_int, C.c_size_t, _ip, _dp, _dp, _dp, _dp]
        L.phx_push.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, _ip, _ip, C.c_int,
                               _ip, _ip, C.POINTER(C.c_size_t * 2)]
        L.phx_gather.argtypes = [vp, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.phx_gather_noalloc.argtypes = [vp, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.phx_deposit.argtypes = [vp, C.c_int, C.c_int, C.c_double]
        L.phx_refine_particles.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, _ip, _ip, C.c_int]
        L.phx_level_ghost_last_step.argtypes = [vp]
        L.phx_time_interpolate.argtypes = [vp, C.c_int, vp, C.c_int, vp, C.c_int, _ip, _ip,
                                           C.c_double]
        L.phx_prepare_step.argtypes = [vp]
        L.phx_reset_flux_sum.argtypes = [vp]
        L.phx_accumulate_flux_sum.argtypes = [vp, C.c_double]
        L.phx_reflux.argtypes = [vp, C.c_double]
        L.phx_coarsen.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, _ip, _ip]
        L.phx_refine_field.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, _ip, _ip]
        L.phx_refine_magnetic_new_faces.argtypes = [vp, C.c_int, _ip, _ip]
        L.phx_set_ghosts_nan.argtypes = [vp, C.c_int]
        L.phx_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
        L.phx_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
        _lib = L
    return _lib


EXPORTED = ["phx_create", "phx_destroy", "phx_last_error", "phx_synchronize", "phx_field_shape",
            "phx_set_field", "phx_get_field", "phx_set_particles", "phx_particle_count",
            "phx_get_particles", "phx_set_particles_aos", "phx_get_particles_aos",
            "phx_load_maxwellian", "phx_update_populations", "phx_update_ions",
            "phx_fill_moment_borders", "phx_fill_ion_borders", "phx_fill_ion_ghost_particles",
            "phx_fill_ghosts", "phx_faraday", "phx_ampere", "phx_electrons_update", "phx_ohm",
            "phx_average", "phx_move_ions", "phx_initialize", "phx_advance", "phx_diagnostics",
            "phx_stream", "phx_profile_push", "phx_sort_particles", "phx_set_sort_interval",
            "phx_set_push_timing", "phx_push_timing", "phx_push", "phx_gather", "phx_deposit",
            "phx_create_level", "phx_patch_box", "phx_patch_nonlevel_ghost_boxes",
            "phx_patch_field_shape", "phx_set_patch_field", "phx_get_patch_field",
            "phx_get_patch_particles", "phx_add_particles", "phx_gather_noalloc",
            "phx_get_tuning", "phx_set_tuning", "phx_refine_particles",
            "phx_level_ghost_last_step", "phx_time_interpolate", "phx_prepare_step",
            "phx_reset_flux_sum", "phx_accumulate_flux_sum", "phx_reflux", "phx_coarsen",
            "phx_refine_field", "phx_refine_magnetic_new_faces", "phx_set_ghosts_nan"]


class Level:
    """One patch level resident on the GPU (a periodic level: one patch covering the domain)."""

    def __init__(self, dim, order, ncells, dx, lower=None, masses=(1.0,), Te=0.1, eta=0.0, nu=1e-4,
                 hyper_mode=0, level=0, periodic=True, capacity=0, precision=PRECISION_FP64,
                 patches=None, apply_tuning_table=True):
        """patches: None (one patch covering the box) or a list of (lower, upper) AMR boxes of
        the level's patches (phx_create_level). apply_tuning_table: before the first
        initialisation, apply the autotuner's entry for this dimension, order, precision family
        and particles per cell (phare_amd/tuning_table.json), if there is one"""
        L = lib()
        self.dim, self.order = dim, order
        lower = list(lower or [0] * dim)
        cfg = Config()
        cfg.dim, cfg.interp_order, cfg.precision, cfg.n_pop = dim, order, precision, len(masses)
        for i, m in enumerate(masses):
            cfg.mass[i] = m
        cfg.eta, cfg.nu, cfg.hyper_mode, cfg.Te = eta, nu, hyper_mode, Te
        p = Patch()
        for d in range(dim):
            p.lower[d] = lower[d]
            p.upper[d] = lower[d] + ncells[d] - 1
            p.dx[d] = dx[d]
        p.level, p.periodic, p.n_nonlevel_ghost_boxes = level, int(periodic), 0
        h = C.c_void_p()
        if patches is None:
            rc = L.phx_create(C.byref(cfg), C.byref(p), int(capacity), C.byref(h))
            self.npatch = 1
        else:
            plo = np.zeros(3 * len(patches), np.int32)
            phi = np.zeros(3 * len(patches), np.int32)
            for i, (lo, hi) in enumerate(patches):
                plo[3 * i:3 * i + dim] = lo
                phi[3 * i:3 * i + dim] = hi
            rc = L.phx_create_level(C.byref(cfg), C.byref(p), len(patches), plo, phi, int(capacity),
                                    C.byref(h))
            self.npatch = len(patches)
        if rc != PHX_OK:
            raise PhxError(rc, "phx_create failed")
        self.h = h
        self.npop = len(masses)
        self.ncells = list(ncells)
        self.dx, self.lower, self.capacity = list(dx), lower, int(capacity)
        self.masses, self.Te, self.eta, self.nu = tuple(masses), Te, eta, nu
        self.precision = precision
        self._tuning_pending = apply_tuning_table

    def close(self):
        if getattr(self, "h", None):
            lib().phx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _rc(self, rc):
        if rc != PHX_OK:
            raise PhxError(rc, lib().phx_last_error(self.h).decode())

    def stream(self):
        return lib().phx_stream(self.h)

    def sync(self):
        self._rc(lib().phx_synchronize(self.h))

    def shape(self, fid):
        s = np.zeros(3, dtype=np.int32)
        self._rc(lib().phx_field_shape(self.h, fid, s))
        return tuple(int(x) for x in s[: self.dim])

    def get(self, fid):
        a = np.zeros(self.shape(fid))
        self._rc(lib().phx_get_field(self.h, fid, a.reshape(-1)))
        return a

    def set(self, fid, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.shape(fid), (a.shape, self.shape(fid))
        self._rc(lib().phx_set_field(self.h, fid, a.reshape(-1)))

    def set_particles(self, pop, kind, p):
        n = len(p["weight"])
        self._rc(lib().phx_set_particles(
            self.h, pop, kind, n, np.ascontiguousarray(p["icell"], dtype=np.int32).reshape(-1),
            np.ascontiguousarray(p["delta"], dtype=np.float64).reshape(-1),
            np.ascontiguousarray(p["v"], dtype=np.float64).reshape(-1),
            np.ascontiguousarray(p["weight"], dtype=np.float64),
            np.ascontiguousarray(p["charge"], dtype=np.float64)))

    def count(self, pop=0, kind=DOMAIN):
        n = C.c_size_t(0)
        self._rc(lib().phx_particle_count(self.h, pop, kind, C.byref(n)))
        return int(n.value)

    def get_particles(self, pop=0, kind=DOMAIN):
        n = self.count(pop, kind)
        p = dict(icell=np.zeros((n, self.dim), dtype=np.int32), delta=np.zeros((n, self.dim)),
                 v=np.zeros((n, 3)), weight=np.zeros(n), charge=np.zeros(n))
        if n:
            self._rc(lib().phx_get_particles(self.h, pop, kind, p["icell"].reshape(-1),
                                             p["delta"].reshape(-1), p["v"].reshape(-1), p["weight"],
                                             p["charge"]))
        return p

    def set_particles_aos(self, pop, kind, records):
        buf = np.ascontiguousarray(records)
        self._rc(lib().phx_set_particles_aos(self.h, pop, kind, len(buf), buf.ctypes.data))

    def get_particles_aos(self, pop, kind, dtype):
        n = self.count(pop, kind)
        out = np.zeros(n, dtype=dtype)
        if n:
            self._rc(lib().phx_get_particles_aos(self.h, pop, kind, out.ctypes.data))
        return out

    def load_maxwellian(self, pop, n, V, Vth, charge, ppc, seed, cutoff=1e-16):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1) for a in [n, *V, *Vth]]
        self._rc(lib().phx_load_maxwellian(self.h, pop, *arrs, charge, ppc, seed, cutoff))

    def initialize(self):
        if self._tuning_pending:
            self._tuning_pending = False
            from phare_amd import autotune

            ppc = self.count(0, DOMAIN) / max(1, int(np.prod(self.ncells)))
            t = autotune.lookup(self.dim, self.order,
                                "fp32" if self.precision == PRECISION_FP32 else "fp64", ppc)
            if t is not None:
                self.set_tuning(**t)
        self._rc(lib().phx_initialize(self.h))

    def update_populations(self, dt, mode, emset=EM_AVG):
        self._rc(lib().phx_update_populations(self.h, emset, dt, mode))

    def update_ions(self):
        self._rc(lib().phx_update_ions(self.h))

    def fill_moment_borders(self):
        self._rc(lib().phx_fill_moment_borders(self.h))

    def fill_ion_borders(self):
        self._rc(lib().phx_fill_ion_borders(self.h))

    def fill_ion_ghost_particles(self):
        self._rc(lib().phx_fill_ion_ghost_particles(self.h))

    def fill_ghosts(self, fid):
        self._rc(lib().phx_fill_ghosts(self.h, fid))

    def faraday(self, bset, eset, bnewset, dt):
        self._rc(lib().phx_faraday(self.h, bset, eset, bnewset, dt))

    def ampere(self, bset):
        self._rc(lib().phx_ampere(self.h, bset))

    def electrons(self):
        self._rc(lib().phx_electrons_update(self.h))

    def ohm(self, bset, eset):
        self._rc(lib().phx_ohm(self.h, bset, eset))

    def average(self):
        self._rc(lib().phx_average(self.h))

    def move_ions(self, dt, mode):
        self._rc(lib().phx_move_ions(self.h, dt, mode))

    def advance(self, dt, nsteps=1):
        self._rc(lib().phx_advance(self.h, dt, nsteps))

    def sort_particles(self):
        self._rc(lib().phx_sort_particles(self.h))

    def set_sort_interval(self, steps):
        self._rc(lib().phx_set_sort_interval(self.h, steps))

    def profile_push(self, dt, mode, reps):
        ms = C.c_double(0.0)
        self._rc(lib().phx_profile_push(self.h, dt, mode, reps, C.byref(ms)))
        return ms.value

    def set_push_timing(self, on):
        self._rc(lib().phx_set_push_timing(self.h, int(on)))

    def push_timing(self, mode, pop=0):
        ms, n = C.c_double(0.0), C.c_int(0)
        self._rc(lib().phx_push_timing(self.h, pop, mode, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def patch_box(self, p):
        lo, hi = np.zeros(3, np.int32), np.zeros(3, np.int32)
        self._rc(lib().phx_patch_box(self.h, p, lo, hi))
        return tuple(lo[:self.dim]), tuple(hi[:self.dim])

    def patch_nonlevel_ghost_boxes(self, p):
        n = C.c_int(0)
        lo, hi = np.zeros(81, np.int32), np.zeros(81, np.int32)
        self._rc(lib().phx_patch_nonlevel_ghost_boxes(self.h, p, C.byref(n), lo, hi))
        return [(tuple(lo[3 * b:3 * b + self.dim]), tuple(hi[3 * b:3 * b + self.dim]))
                for b in range(n.value)]

    def patch_shape(self, p, fid):
        s = np.zeros(3, dtype=np.int32)
        self._rc(lib().phx_patch_field_shape(self.h, p, fid, s))
        return tuple(int(x) for x in s[: self.dim])

    def get_patch(self, p, fid):
        a = np.zeros(self.patch_shape(p, fid))
        self._rc(lib().phx_get_patch_field(self.h, p, fid, a.reshape(-1)))
        return a

    def set_patch(self, p, fid, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.patch_shape(p, fid)
        self._rc(lib().phx_set_patch_field(self.h, p, fid, a.reshape(-1)))

    def get_patch_particles(self, p, pop=0):
        n = C.c_size_t(0)
        self._rc(lib().phx_get_patch_particles(self.h, p, pop, C.byref(n), None, None, None, None, None))
        m = int(n.value)
        out = dict(icell=np.zeros((m, self.dim), dtype=np.int32), delta=np.zeros((m, self.dim)),
                   v=np.zeros((m, 3)), weight=np.zeros(m), charge=np.zeros(m))
        if m:
            self._rc(lib().phx_get_patch_particles(
                self.h, p, pop, C.byref(n), out["icell"].ctypes.data, out["delta"].ctypes.data,
                out["v"].ctypes.data, out["weight"].ctypes.data, out["charge"].ctypes.data))
        return out

    def add_particles(self, pop, kind, p):
        n = len(p["weight"])
        self._rc(lib().phx_add_particles(
            self.h, pop, kind, n, np.ascontiguousarray(p["icell"], dtype=np.int32).reshape(-1),
            np.ascontiguousarray(p["delta"], dtype=np.float64).reshape(-1),
            np.ascontiguousarray(p["v"], dtype=np.float64).reshape(-1),
            np.ascontiguousarray(p["weight"], dtype=np.float64),
            np.ascontiguousarray(p["charge"], dtype=np.float64)))

    def push(self, dt, first=(), second=(), pop=0, kind=DOMAIN, emset=EM_AVG):
        """BorisPusher::move in place (phx_push): selectors are lists of (lower, upper) AMR boxes
        (empty: noop); returns (end of the second selection, end of the first); the array is left
        partitioned [both | first only | neither]"""
        def boxes(bl):
            lo = np.zeros(3 * max(1, len(bl)), np.int32)
            hi = np.zeros(3 * max(1, len(bl)), np.int32)
            for b, (l, h) in enumerate(bl):
                lo[3 * b:3 * b + len(l)] = l
                hi[3 * b:3 * b + len(h)] = h
            return len(bl), lo, hi
        nf, flo, fhi = boxes(first)
        ns, slo, shi = boxes(second)
        out = (C.c_size_t * 2)()
        self._rc(lib().phx_push(self.h, pop, kind, emset, dt, nf, flo, fhi, ns, slo, shi, C.byref(out)))
        return int(out[0]), int(out[1])

    def gather(self, emset=EM_AVG, pop=0, kind=DOMAIN):
        """E and B at every particle (phx_gather), in the order of get_particles"""
        n = self.count(pop, kind)
        E, B = np.zeros((n, 3)), np.zeros((n, 3))
        if n:
            self._rc(lib().phx_gather(self.h, emset, pop, kind, E.reshape(-1), B.reshape(-1)))
        return E, B

    def gather_noalloc(self, emset=EM_AVG, pop=0, kind=DOMAIN):
        """phx_gather through phx_gather_noalloc (no device allocation: the library's fixed
        scratch array stages the values)"""
        n = self.count(pop, kind)
        E, B = np.zeros((n, 3)), np.zeros((n, 3))
        self._rc(lib().phx_gather_noalloc(self.h, emset, pop, kind, E.reshape(-1), B.reshape(-1)))
        return E, B

    def get_tuning(self):
        t = Tuning()
        self._rc(lib().phx_get_tuning(self.h, C.byref(t)))
        return dict(segment_shape=list(t.segment_shape), rows_per_item=t.rows_per_item,
                    tile_row_pad=t.tile_row_pad, sort_interval=t.sort_interval,
                    tile_copies=t.tile_copies, deposit_window=t.deposit_window)

    def set_tuning(self, **kw):
        """phx_set_tuning with the given fields changed from the current tuning"""
        cur = self.get_tuning()
        cur.update(kw)
        t = Tuning((C.c_int * 3)(*cur["segment_shape"]), cur["rows_per_item"], cur["tile_row_pad"],
                   cur["sort_interval"], cur["tile_copies"], cur["deposit_window"])
        self._rc(lib().phx_set_tuning(self.h, C.byref(t)))

    def refine_particles(self, coarse, boxes, pop=0, kind=LEVEL_GHOST, nbr=0):
        """phx_refine_particles: split the domain particles of `coarse` (a Level) into array
        `kind` of this (fine) level, keeping those in the fine AMR boxes [(lower, upper), ...]"""
        lo = np.zeros(3 * len(boxes), np.int32)
        hi = np.zeros(3 * len(boxes), np.int32)
        for i, (a, b) in enumerate(boxes):
            lo[3 * i:3 * i + self.dim] = a
            hi[3 * i:3 * i + self.dim] = b
        self._rc(lib().phx_refine_particles(self.h, coarse.h, pop, kind, len(boxes), lo, hi, nbr))

    def level_ghost_last_step(self):
        self._rc(lib().phx_level_ghost_last_step(self.h))

    def time_interpolate(self, fid, old, old_fid, new, new_fid, lower, upper, alpha):
        lo = np.zeros(3, np.int32)
        hi = np.zeros(3, np.int32)
        lo[:self.dim] = lower
        hi[:self.dim] = upper
        self._rc(lib().phx_time_interpolate(self.h, fid, old.h, old_fid, new.h, new_fid, lo, hi,
                                            alpha))

    def prepare_step(self):
        self._rc(lib().phx_prepare_step(self.h))

    def reset_flux_sum(self):
        self._rc(lib().phx_reset_flux_sum(self.h))

    def accumulate_flux_sum(self, coef):
        self._rc(lib().phx_accumulate_flux_sum(self.h, coef))

    def reflux(self, dt):
        self._rc(lib().phx_reflux(self.h, dt))

    def coarsen(self, fid, fine, fine_fid, op, lower, upper):
        """phx_coarsen: this (coarse) level's field fid over the coarse AMR nodes lower..upper
        from the finer level's field fine_fid"""
        lo = np.zeros(3, np.int32)
        hi = np.zeros(3, np.int32)
        lo[:self.dim] = lower
        hi[:self.dim] = upper
        self._rc(lib().phx_coarsen(self.h, fid, fine.h, fine_fid, op, lo, hi))

    def refine_field(self, fid, coarse, coarse_fid, op, lower, upper):
        """phx_refine_field: this (fine) level's field fid over its fine AMR nodes lower..upper
        from the coarser level's field coarse_fid"""
        lo = np.zeros(3, np.int32)
        hi = np.zeros(3, np.int32)
        lo[:self.dim] = lower
        hi[:self.dim] = upper
        self._rc(lib().phx_refine_field(self.h, fid, coarse.h, coarse_fid, op, lo, hi))

    def refine_magnetic_new_faces(self, b_fid, lower, upper):
        """phx_refine_magnetic_new_faces over the fine AMR cell box lower..upper"""
        lo = np.zeros(3, np.int32)
        hi = np.zeros(3, np.int32)
        lo[:self.dim] = lower
        hi[:self.dim] = upper
        self._rc(lib().phx_refine_magnetic_new_faces(self.h, b_fid, lo, hi))

    def set_ghosts_nan(self, fid):
        """phx_set_ghosts_nan: NaN on every node of field fid outside its physical nodes"""
        self._rc(lib().phx_set_ghosts_nan(self.h, fid))

    def deposit(self, pop=0, kind=DOMAIN, coef=1.0):
        self._rc(lib().phx_deposit(self.h, pop, kind, coef))

    def diagnostics(self):
        out = np.zeros(4)
        self._rc(lib().phx_diagnostics(self.h, out))
        return dict(kinetic=out[0], magnetic=out[1], charge=out[2], count=out[3])


def particle_dtype(dim):
    """numpy dtype of the reference's Particle<dim> record (particle.hpp:37-61)"""
    fields = [("weight", "<f8"), ("charge", "<f8"), ("iCell", "<i4", (dim,))]
    if dim in (1, 3):
        fields.append(("_pad", "<i4"))
    fields += [("delta", "<f8", (dim,)), ("v", "<f8", (3,))]
    return np.dtype(fields)
