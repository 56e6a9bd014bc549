"""ctypes binding of libselm_hip.so (include/selm_hip.h).

Python mirror of the reference's Selm / Eulerian_Period / Lagrangian plugin
surface (cuda/SELM.cuh:6-28, cuda/SELM_Eulerian.cuh:6-29,
cuda/SELM_Lagrangian.cuh:10-49): a ``System`` is one Selm (one periodic grid
plus one bead set) and a ``Pool`` holds many of them in one device arena.
Every call goes through the C ABI; there is no Python or CPU fallback: if the
HIP library is missing, ``load()`` raises.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libselm_hip.so")

SELM_OK = 0
SELM_ERR_ARG = 1
SELM_ERR_HIP = 2
SELM_ERR_STATE = 3
SELM_ERR_NOMEM = 4
SELM_ERR_UNSUPPORTED = 5
SELM_ERR_POOL_FULL = 6
SELM_ERR_DEVICE = 7
SELM_ERR_SCRATCH = 8

# device status flags (selm_status / selm_pool_status)
STATUS_TILE_OVERFLOW = 1
STATUS_NONFINITE = 2
STATUS_DISPLACEMENT = 4

PREC_FP64 = 0
PREC_FP32_BULK = 1

STAGES = ["thermal_noise", "bead_update", "spread", "ustar_fft_x", "fft_y", "fft_z_projection", "ifft_y",
          "ifft_x", "bond_forces"]

EXPORTS = [
    "selm_params_default", "selm_create", "selm_destroy", "selm_last_error", "selm_set_beads",
    "selm_set_bonds", "selm_set_external_force", "selm_set_conservative_force", "selm_init",
    "selm_step", "selm_sync", "selm_get_beads", "selm_get_field", "selm_set_field",
    "selm_timestep", "selm_footprint", "selm_stage_times", "selm_mean_mode", "selm_stream",
    "selm_pool_create", "selm_pool_destroy", "selm_pool_last_error", "selm_pool_admit",
    "selm_pool_step", "selm_pool_sync", "selm_pool_get_beads", "selm_pool_get_field",
    "selm_pool_timestep", "selm_pool_retire", "selm_pool_stats", "selm_pool_stream",
    "selm_msd", "selm_set_angles", "selm_set_pair_lj", "selm_fft_passes", "selm_autotune",
    "selm_pool_autotune", "selm_status", "selm_pool_status", "selm_check_beads", "selm_msd_noalloc",
    "selm_set_timestep", "selm_pool_set_field", "selm_pool_set_timestep",
    "selm_spread_shape", "selm_pool_spread_shape",
]


class SelmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("selm error %d: %s" % (code, msg))
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("dims", C.c_int * 3),
        ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("rho", C.c_double), ("mu", C.c_double), ("kB", C.c_double),
        ("T", C.c_double), ("dt", C.c_double),
        ("seed", C.c_uint64),
        ("stochastic", C.c_int), ("precision", C.c_int), ("f4_quirk", C.c_int),
        ("reserved", C.c_int),
    ]


class PoolParams(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_replicas", C.c_int),
                ("precision", C.c_int), ("use_graph", C.c_int)]


_lib = None
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
vp = C.c_void_p


def load(path=None):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError("libselm_hip.so is not built (%s); run __graft_entry__.build()" % path)
    L = C.CDLL(path)
    L.selm_params_default.argtypes = [C.POINTER(Params)]
    L.selm_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.selm_destroy.argtypes = [vp]
    L.selm_last_error.argtypes = [vp]
    L.selm_last_error.restype = C.c_char_p
    L.selm_set_beads.argtypes = [vp, C.c_int, dp, dp, dp]
    L.selm_set_bonds.argtypes = [vp, C.c_int, ip, dp, dp]
    L.selm_set_external_force.argtypes = [vp, C.c_int, ip, dp]
    L.selm_set_conservative_force.argtypes = [vp, dp]
    L.selm_init.argtypes = [vp]
    L.selm_step.argtypes = [vp, C.c_int, vp]
    L.selm_sync.argtypes = [vp]
    L.selm_get_beads.argtypes = [vp, dp, dp, dp, dp]
    L.selm_get_field.argtypes = [vp, C.c_int, dp]
    L.selm_set_field.argtypes = [vp, C.c_int, dp]
    L.selm_timestep.argtypes = [vp]
    L.selm_timestep.restype = C.c_longlong
    L.selm_footprint.argtypes = [C.POINTER(Params), C.c_int, C.c_int, C.POINTER(C.c_size_t)]
    L.selm_stage_times.argtypes = [vp, C.c_int, dp, C.c_int]
    L.selm_mean_mode.argtypes = [vp, dp]
    L.selm_fft_passes.argtypes = [vp]
    L.selm_autotune.argtypes = [vp, C.c_int, ip, dp]
    L.selm_pool_autotune.argtypes = [vp, C.c_int, ip, dp]
    L.selm_stream.argtypes = [vp]
    L.selm_stream.restype = vp
    L.selm_pool_create.argtypes = [C.POINTER(PoolParams), C.POINTER(vp)]
    L.selm_pool_destroy.argtypes = [vp]
    L.selm_pool_last_error.argtypes = [vp]
    L.selm_pool_last_error.restype = C.c_char_p
    L.selm_pool_admit.argtypes = [vp, C.POINTER(Params), C.c_int, dp, dp, dp, C.c_int, ip, dp, dp,
                                  C.c_int, ip, dp, ip]
    L.selm_pool_step.argtypes = [vp, C.c_int]
    L.selm_pool_sync.argtypes = [vp]
    L.selm_status.argtypes = [vp, ip, C.c_int]
    L.selm_check_beads.argtypes = [C.c_int, dp, dp, dp, ip]
    L.selm_pool_status.argtypes = [vp, C.c_int, ip, C.c_int]
    L.selm_pool_get_beads.argtypes = [vp, C.c_int, dp, dp, dp, dp]
    L.selm_pool_get_field.argtypes = [vp, C.c_int, C.c_int, dp]
    L.selm_pool_timestep.argtypes = [vp, C.c_int]
    L.selm_pool_timestep.restype = C.c_longlong
    L.selm_set_timestep.argtypes = [vp, C.c_longlong]
    L.selm_spread_shape.argtypes = [vp, ip]
    L.selm_pool_spread_shape.argtypes = [vp, ip]
    L.selm_pool_set_field.argtypes = [vp, C.c_int, C.c_int, dp]
    L.selm_pool_set_timestep.argtypes = [vp, C.c_int, C.c_longlong]
    L.selm_pool_retire.argtypes = [vp, C.c_int]
    L.selm_pool_stats.argtypes = [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), ip]
    L.selm_pool_stream.argtypes = [vp]
    L.selm_pool_stream.restype = vp
    L.selm_msd.argtypes = [C.POINTER(C.c_float), C.c_int, C.c_int, dp, C.c_int, C.c_int, dp]
    L.selm_msd_noalloc.argtypes = [vp, C.c_int, C.c_int, dp, C.c_int, C.c_int, vp, vp]
    L.selm_set_angles.argtypes = [vp, C.c_int, ip, dp, dp]
    L.selm_set_pair_lj.argtypes = [vp, C.c_int, ip, dp, dp, dp, dp]
    _lib = L
    return L


def _d(a):
    return None if a is None else a.ctypes.data_as(dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(ip)


def f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(shape) if shape is not None else a


def make_params(dims, lo, hi, rho=1.0, mu=1.0, kB=1.0, T=1.0, dt=0.01, seed=123456,
                stochastic=False, precision=PREC_FP64, f4_quirk=True):
    p = Params()
    load().selm_params_default(C.byref(p))
    p.dims = (C.c_int * 3)(*[int(d) for d in dims])
    p.lo = (C.c_double * 3)(*[float(v) for v in lo])
    p.hi = (C.c_double * 3)(*[float(v) for v in hi])
    p.rho, p.mu, p.kB, p.T, p.dt = rho, mu, kB, T, dt
    p.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    p.stochastic = int(bool(stochastic))
    p.precision = int(precision)
    p.f4_quirk = int(bool(f4_quirk))
    return p


def check_beads(x, v, m):
    """selm_check_beads: -1 when the bead inputs are valid, else the first
    offending bead (runs on the host, no device needed)."""
    x = f64(x, (-1, 3))
    v = None if v is None else f64(v, (-1, 3))
    m = f64(m, (-1,))
    bad = C.c_int(-1)
    load().selm_check_beads(x.shape[0], _d(x), _d(v), _d(m), C.byref(bad))
    return bad.value


def footprint(params, M, nb):
    out = C.c_size_t()
    rc = load().selm_footprint(C.byref(params), int(M), int(nb), C.byref(out))
    if rc:
        raise SelmError(rc, "selm_footprint")
    return out.value


class System:
    """One SELM system on the GPU (selm_ctx)."""

    def __init__(self, params):
        self.L = load()
        self.params = params
        self.h = vp()
        rc = self.L.selm_create(C.byref(params), C.byref(self.h))
        if rc:
            raise SelmError(rc, "selm_create")
        d = params.dims
        self.shape = (d[2], d[1], d[0])
        self.N = d[0] * d[1] * d[2]
        self.M = 0

    def _chk(self, rc, what):
        if rc:
            raise SelmError(rc, "%s: %s" % (what, self.L.selm_last_error(self.h).decode()))

    def close(self):
        if self.h:
            self.L.selm_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_beads(self, x, v, m):
        x = f64(x, (-1, 3))
        v = None if v is None else f64(v, (-1, 3))
        m = f64(m, (-1,))
        self.M = x.shape[0]
        self._chk(self.L.selm_set_beads(self.h, self.M, _d(x), _d(v), _d(m)), "set_beads")

    def set_bonds(self, ij, K, r0):
        ij = np.ascontiguousarray(ij, dtype=np.int32).reshape(-1, 2)
        K = f64(K, (-1,))
        r0 = f64(r0, (-1,))
        self._chk(self.L.selm_set_bonds(self.h, ij.shape[0], _i(ij), _d(K), _d(r0)), "set_bonds")

    def set_angles(self, ijk, K, theta0):
        ijk = np.ascontiguousarray(ijk, dtype=np.int32).reshape(-1, 3)
        K = f64(K, (-1,))
        t0 = f64(theta0, (-1,))
        self._chk(self.L.selm_set_angles(self.h, ijk.shape[0], _i(ijk), _d(K), _d(t0)), "set_angles")

    def set_pair_lj(self, type_, eps, sigma, rc, special_lj=(0.0, 0.0, 0.0)):
        t = np.ascontiguousarray(type_, dtype=np.int32).reshape(-1)
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        nt = eps.shape[0]
        sigma = f64(sigma, (nt, nt))
        rc = f64(rc, (nt, nt))
        sp = f64(special_lj, (3,))
        self._chk(self.L.selm_set_pair_lj(self.h, nt, _i(t), _d(f64(eps, (nt, nt))), _d(sigma), _d(rc), _d(sp)),
                  "set_pair_lj")

    def set_external_force(self, idx, F):
        idx = np.ascontiguousarray(idx, dtype=np.int32).reshape(-1)
        F = f64(F, (-1, 3))
        self._chk(self.L.selm_set_external_force(self.h, idx.shape[0], _i(idx), _d(F)), "set_external_force")

    def set_conservative_force(self, f):
        f = f64(f, (-1, 3))
        self._chk(self.L.selm_set_conservative_force(self.h, _d(f)), "set_conservative_force")

    def init(self):
        self._chk(self.L.selm_init(self.h), "init")

    def step(self, n=1, stream=None):
        self._chk(self.L.selm_step(self.h, int(n), stream), "step")

    def sync(self):
        """Wait for the system's work; raises SelmError (code SELM_ERR_DEVICE)
        when a device status flag is raised."""
        self._chk(self.L.selm_sync(self.h), "sync")

    def status(self, clear=False):
        """Device status flags (STATUS_*), after waiting for the work."""
        f = C.c_int(0)
        self._chk(self.L.selm_status(self.h, C.byref(f), int(bool(clear))), "status")
        return f.value

    def beads(self):
        M = self.M
        x = np.zeros((M, 3)); v = np.zeros((M, 3)); fc = np.zeros((M, 3)); ft = np.zeros((M, 3))
        self._chk(self.L.selm_get_beads(self.h, _d(x), _d(v), _d(fc), _d(ft)), "get_beads")
        return x, v, fc, ft

    def field(self):
        out = np.zeros((3,) + self.shape)
        for c in range(3):
            buf = np.zeros(self.shape)
            self._chk(self.L.selm_get_field(self.h, c, _d(buf)), "get_field")
            out[c] = buf
        return out

    def set_field(self, u):
        u = f64(u, (3,) + self.shape)
        for c in range(3):
            buf = np.ascontiguousarray(u[c])
            self._chk(self.L.selm_set_field(self.h, c, _d(buf)), "set_field")

    @property
    def timestep(self):
        return self.L.selm_timestep(self.h)

    def set_timestep(self, n):
        """selm_set_timestep: resume at n completed steps (after init)."""
        self._chk(self.L.selm_set_timestep(self.h, int(n)), "set_timestep")

    def stage_times(self, nsteps):
        ms = np.zeros(len(STAGES))
        self._chk(self.L.selm_stage_times(self.h, int(nsteps), _d(ms), len(STAGES)), "stage_times")
        return dict(zip(STAGES, ms.tolist()))

    def mean_mode(self):
        out = np.zeros(3)
        self._chk(self.L.selm_mean_mode(self.h, _d(out)), "mean_mode")
        return out

    def autotune(self, steps=20):
        """Time the FFT pass shapes, then the spread launch's shapes, on this
        system and keep the fastest (selm_autotune).  Returns {"passes",
        "zcols", "split_at", "forcing_workgroups", "ms_per_step"}."""
        shape = np.zeros(2, np.int32)
        ms = np.zeros(1)
        self._chk(self.L.selm_autotune(self.h, int(steps), _i(shape), _d(ms)), "autotune")
        sp = self.spread_shape()
        return {"passes": int(shape[0]), "zcols": int(shape[1]), "split_at": sp[0],
                "forcing_workgroups": sp[1], "ms_per_step": float(ms[0])}

    def spread_shape(self):
        """(split threshold, forcing workgroups of the spread launch) of the
        next step (selm_spread_shape)."""
        out = np.zeros(2, np.int32)
        self._chk(self.L.selm_spread_shape(self.h, _i(out)), "spread_shape")
        return int(out[0]), int(out[1])

    def fft_passes(self):
        """3 (plane path), 5 (transposing path) or 0 (runtime-length kernels)."""
        n = self.L.selm_fft_passes(self.h)
        self._chk(-n if n < 0 else 0, "fft_passes")
        return n

    @property
    def stream(self):
        return self.L.selm_stream(self.h)


class Pool:
    """Many SELM systems in one device arena, stepped together (selm_pool)."""

    def __init__(self, hbm_budget, max_replicas, precision=PREC_FP32_BULK, use_graph=True):
        self.L = load()
        pp = PoolParams(int(hbm_budget), int(max_replicas), int(precision), int(bool(use_graph)))
        self.h = vp()
        rc = self.L.selm_pool_create(C.byref(pp), C.byref(self.h))
        if rc:
            raise SelmError(rc, "selm_pool_create")
        self.meta = {}

    def _chk(self, rc, what):
        if rc:
            raise SelmError(rc, "%s: %s" % (what, self.L.selm_pool_last_error(self.h).decode()))

    def close(self):
        if self.h:
            self.L.selm_pool_destroy(self.h)
            self.h = vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, params, x, v, m, bonds=None, K=None, r0=None, ext_idx=None, ext_F=None):
        """Admit one system; returns its slot or None when the pool is full."""
        x = f64(x, (-1, 3)); m = f64(m, (-1,))
        v = None if v is None else f64(v, (-1, 3))
        M = x.shape[0]
        if bonds is None:
            ij = np.zeros((0, 2), np.int32); K = np.zeros(0); r0 = np.zeros(0)
        else:
            ij = np.ascontiguousarray(bonds, dtype=np.int32).reshape(-1, 2)
            K = f64(K, (-1,)); r0 = f64(r0, (-1,))
        if ext_idx is None:
            ei = np.zeros(0, np.int32); eF = np.zeros((0, 3))
        else:
            ei = np.ascontiguousarray(ext_idx, dtype=np.int32).reshape(-1); eF = f64(ext_F, (-1, 3))
        slot = C.c_int(-1)
        rc = self.L.selm_pool_admit(self.h, C.byref(params), M, _d(x), _d(v), _d(m), ij.shape[0], _i(ij),
                                    _d(K), _d(r0), ei.shape[0], _i(ei), _d(eF), C.byref(slot))
        if rc == SELM_ERR_POOL_FULL:
            return None
        self._chk(rc, "pool_admit")
        d = params.dims
        self.meta[slot.value] = (M, (d[2], d[1], d[0]))
        return slot.value

    def step(self, n=1):
        self._chk(self.L.selm_pool_step(self.h, int(n)), "pool_step")

    def sync(self):
        """Wait for the pool's work; raises SelmError (SELM_ERR_DEVICE) when an
        admitted system has a status flag raised."""
        self._chk(self.L.selm_pool_sync(self.h), "pool_sync")

    def status(self, slot, clear=False):
        """Device status flags (STATUS_*) of one admitted system."""
        f = C.c_int(0)
        self._chk(self.L.selm_pool_status(self.h, int(slot), C.byref(f), int(bool(clear))), "pool_status")
        return f.value

    def beads(self, slot):
        M, _ = self.meta[slot]
        x = np.zeros((M, 3)); v = np.zeros((M, 3)); fc = np.zeros((M, 3)); ft = np.zeros((M, 3))
        self._chk(self.L.selm_pool_get_beads(self.h, slot, _d(x), _d(v), _d(fc), _d(ft)), "pool_get_beads")
        return x, v, fc, ft

    def field(self, slot):
        _, shape = self.meta[slot]
        out = np.zeros((3,) + shape)
        for c in range(3):
            buf = np.zeros(shape)
            self._chk(self.L.selm_pool_get_field(self.h, slot, c, _d(buf)), "pool_get_field")
            out[c] = buf
        return out

    def timestep(self, slot):
        return self.L.selm_pool_timestep(self.h, slot)

    def set_field(self, slot, u):
        _, shape = self.meta[slot]
        u = f64(u, (3,) + shape)
        for c in range(3):
            buf = np.ascontiguousarray(u[c])
            self._chk(self.L.selm_pool_set_field(self.h, slot, c, _d(buf)), "pool_set_field")

    def set_timestep(self, slot, n):
        self._chk(self.L.selm_pool_set_timestep(self.h, slot, int(n)), "pool_set_timestep")

    def retire(self, slot):
        self._chk(self.L.selm_pool_retire(self.h, slot), "pool_retire")
        self.meta.pop(slot, None)

    def autotune(self, steps=10):
        """selm_pool_autotune: time the FFT pass shapes on the pool's current
        systems, restore their state, keep the fastest."""
        shape = np.zeros(2, np.int32)
        ms = np.zeros(1)
        self._chk(self.L.selm_pool_autotune(self.h, int(steps), _i(shape), _d(ms)), "pool_autotune")
        sp = self.spread_shape()
        return {"passes": int(shape[0]), "zcols": int(shape[1]), "split_at": sp[0],
                "forcing_workgroups": sp[1], "ms_per_step": float(ms[0])}

    def spread_shape(self):
        out = np.zeros(2, np.int32)
        self._chk(self.L.selm_pool_spread_shape(self.h, _i(out)), "pool_spread_shape")
        return int(out[0]), int(out[1])

    def stats(self):
        used = C.c_size_t(); budget = C.c_size_t(); active = C.c_int()
        self._chk(self.L.selm_pool_stats(self.h, C.byref(used), C.byref(budget), C.byref(active)), "pool_stats")
        return used.value, budget.value, active.value

    @property
    def stream(self):
        return self.L.selm_pool_stream(self.h)
