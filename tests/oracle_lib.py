"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

The oracle is test infrastructure: this module is imported only by tests,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# SELM_ORACLE_LIB selects another build of the oracle (the sanitizer build of
# tests/test_sanitizers.py)
LIB_PATH = os.environ.get("SELM_ORACLE_LIB") or os.path.join(ORACLE_DIR, "liboracle.so")


class OracleParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("lo", C.c_double * 3), ("hi", C.c_double * 3),
        ("rho", C.c_double), ("mu", C.c_double), ("kB", C.c_double),
        ("T", C.c_double), ("dt", C.c_double),
        ("seed", C.c_uint64), ("noise", C.c_int), ("f4_quirk", C.c_int),
    ]


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"])


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    vp = C.c_void_p
    L.oracle_create.restype = vp
    L.oracle_create.argtypes = [C.POINTER(OracleParams)]
    L.oracle_destroy.argtypes = [vp]
    L.oracle_set_beads.argtypes = [vp, C.c_int, dp, dp, dp]
    L.oracle_set_bonds.argtypes = [vp, C.c_int, ip, dp, dp]
    L.oracle_set_external_force.argtypes = [vp, C.c_int, ip, dp]
    L.oracle_set_conservative_force.argtypes = [vp, dp]
    L.oracle_init.argtypes = [vp]
    L.oracle_step.argtypes = [vp, C.c_int]
    L.oracle_timestep.argtypes = [vp]
    L.oracle_timestep.restype = C.c_long
    L.oracle_get_beads.argtypes = [vp, dp, dp, dp, dp]
    L.oracle_get_field.argtypes = [vp, C.c_int, dp]
    L.oracle_set_field.argtypes = [vp, C.c_int, dp]
    L.oracle_set_threads.argtypes = [C.c_int]
    L.oracle_peskin_phi.restype = C.c_double
    L.oracle_peskin_phi.argtypes = [C.c_double, C.c_int]
    L.oracle_gamma.restype = C.c_double
    L.oracle_gamma.argtypes = [C.c_double] * 3 + [C.c_int] * 3 + [dp, C.c_int]
    L.oracle_lambda.argtypes = [C.c_double] * 3 + [C.c_int] * 3 + [C.c_double, dp, C.c_int]
    L.oracle_fft_r2c_3d.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
    L.oracle_fft_c2r_3d.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
    L.oracle_fft_c2c_1d.argtypes = [C.c_int, C.c_int, dp, dp]
    u32p = C.POINTER(C.c_uint32)
    L.oracle_philox4x32_10.argtypes = [u32p, u32p, u32p]
    L.oracle_normal_pair.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, dp]
    L.oracle_mode_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, dp]
    L.oracle_bond_forces.argtypes = [vp, dp, dp]
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def phi(r, f4=1):
    return lib().oracle_peskin_phi(float(r), int(f4))


def gamma(rx, ry, rz, u, f4=1):
    u = np.ascontiguousarray(u, dtype=np.float64)
    nz, ny, nx = u.shape
    return lib().oracle_gamma(rx, ry, rz, nx, ny, nz, _dp(u), f4)


def lam(rx, ry, rz, f, u_f, f4=1):
    assert u_f.dtype == np.float64 and u_f.flags.c_contiguous
    nz, ny, nx = u_f.shape
    lib().oracle_lambda(rx, ry, rz, nx, ny, nz, f, _dp(u_f), f4)


def r2c3(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    nz, ny, nx = a.shape
    out = np.zeros((nz, ny, nx // 2 + 1), dtype=np.complex128)
    lib().oracle_fft_r2c_3d(nx, ny, nz, _dp(a), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def c2r3(h, nx):
    h = np.ascontiguousarray(h, dtype=np.complex128)
    nz, ny, xc = h.shape
    assert xc == nx // 2 + 1
    out = np.zeros((nz, ny, nx), dtype=np.float64)
    lib().oracle_fft_c2r_3d(nx, ny, nz, h.ctypes.data_as(C.POINTER(C.c_double)), _dp(out))
    return out


def c2c1(a, sign):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    out = np.zeros_like(a)
    lib().oracle_fft_c2c_1d(len(a), sign, a.ctypes.data_as(C.POINTER(C.c_double)),
                            out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().oracle_philox4x32_10(c, k, o)
    return list(o)


def normal_pair(seed, stream, step, index, slot):
    out = np.zeros(2)
    lib().oracle_normal_pair(seed, stream, step, index, slot, _dp(out))
    return out


def mode_normals(seed, stream, step, index):
    out = np.zeros(6)
    lib().oracle_mode_normals(seed, stream, step, index, _dp(out))
    return out


def set_threads(n):
    lib().oracle_set_threads(int(n))


def msd(frames, box_x, nlags, mult=1):
    """oracle_msd: frames [F][3][natoms] float32 -> MSD[0..nlags-1]."""
    frames = np.ascontiguousarray(frames, dtype=np.float32)
    F, _, n = frames.shape
    out = np.zeros(nlags)
    box = np.array([box_x, box_x, box_x], dtype=np.float64)
    L = lib()
    L.oracle_msd.argtypes = [C.POINTER(C.c_float), C.c_int, C.c_int, C.POINTER(C.c_double), C.c_int, C.c_int,
                             C.POINTER(C.c_double)]
    L.oracle_msd(frames.ctypes.data_as(C.POINTER(C.c_float)), F, n, _dp(box), int(nlags), int(mult), _dp(out))
    return out


class Oracle:
    """One SELM system (Eulerian_Period + bead-spring beads) on the CPU, FP64."""

    def __init__(self, nx, ny, nz, lo, hi, rho=1.0, mu=1.0, kB=1.0, T=1.0, dt=0.01,
                 seed=1, noise=False, f4_quirk=True):
        p = OracleParams()
        p.nx, p.ny, p.nz = nx, ny, nz
        p.lo = (C.c_double * 3)(*lo)
        p.hi = (C.c_double * 3)(*hi)
        p.rho, p.mu, p.kB, p.T, p.dt = rho, mu, kB, T, dt
        p.seed = seed
        p.noise = int(bool(noise))
        p.f4_quirk = int(bool(f4_quirk))
        self.p = p
        self.shape = (nz, ny, nx)
        self.N = nx * ny * nz
        self.h = lib().oracle_create(C.byref(p))
        if not self.h:
            raise ValueError("oracle_create failed")
        self.M = 0

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_beads(self, x, v, m):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1, 3)
        m = np.ascontiguousarray(m, dtype=np.float64).reshape(-1)
        self.M = x.shape[0]
        self._keep = (x, v, m)
        assert lib().oracle_set_beads(self.h, self.M, _dp(x), _dp(v), _dp(m)) == 0

    def set_bonds(self, ij, K, r0):
        ij = np.ascontiguousarray(ij, dtype=np.int32).reshape(-1, 2)
        K = np.ascontiguousarray(K, dtype=np.float64).reshape(-1)
        r0 = np.ascontiguousarray(r0, dtype=np.float64).reshape(-1)
        assert lib().oracle_set_bonds(self.h, ij.shape[0], _ip(ij), _dp(K), _dp(r0)) == 0

    def set_external_force(self, idx, F):
        idx = np.ascontiguousarray(idx, dtype=np.int32).reshape(-1)
        F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1, 3)
        assert lib().oracle_set_external_force(self.h, idx.shape[0], _ip(idx), _dp(F)) == 0

    def set_conservative_force(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1, 3)
        lib().oracle_set_conservative_force(self.h, _dp(f))

    def init(self):
        lib().oracle_init(self.h)

    def step(self, n=1):
        lib().oracle_step(self.h, int(n))

    @property
    def timestep(self):
        return lib().oracle_timestep(self.h)

    def beads(self):
        x = np.zeros((self.M, 3)); v = np.zeros((self.M, 3))
        fc = np.zeros((self.M, 3)); ft = np.zeros((self.M, 3))
        lib().oracle_get_beads(self.h, _dp(x), _dp(v), _dp(fc), _dp(ft))
        return x, v, fc, ft

    def field(self, which=0):
        out = np.zeros((3,) + self.shape)
        lib().oracle_get_field(self.h, which, _dp(out))
        return out

    def set_field(self, which, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64).reshape((3,) + self.shape)
        lib().oracle_set_field(self.h, which, _dp(arr))

    def bond_forces(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        f = np.zeros_like(x)
        lib().oracle_bond_forces(self.h, _dp(x), _dp(f))
        return f
