"""ctypes binding of libselm_host.so (include/selm_host.h) for the tests."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# SELM_HOST_LIB selects another build of the library (the sanitizer build of
# tests/test_sanitizers.py)
PATH = os.environ.get("SELM_HOST_LIB") or os.path.join(ROOT, "monomers2026_amd", "libselm_host.so")
DRIVER = os.path.join(ROOT, "monomers2026_amd", "selm_hip_driver")
DECK_SMALL = os.path.join(ROOT, "tests", "golden", "deck_small")

_lib = None
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


class Model(C.Structure):
    _fields_ = [("rho", C.c_double), ("T", C.c_double), ("mu", C.c_double), ("KB", C.c_double),
                ("nsteps", C.c_int), ("flagStochasticDriving", C.c_int), ("SELM_Seed", C.c_int),
                ("numMeshPtsPerDir", C.c_int * 3), ("saveSkipSimulationData", C.c_int),
                ("writeParticalV", C.c_int), ("writeParticalF", C.c_int), ("writeParticalX", C.c_int),
                ("LagrangianType", C.c_char * 64), ("EulerianType", C.c_char * 64)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(PATH):
            from monomers2026_amd import build
            build.build_host()
        L = C.CDLL(PATH)
        L.selmh_last_error.restype = C.c_char_p
        L.selmh_model_load.argtypes = [C.c_char_p, C.POINTER(Model)]
        L.selmh_deck_load.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        L.selmh_deck_free.argtypes = [C.c_void_p]
        L.selmh_deck_sizes.argtypes = [C.c_void_p, ip, ip, ip, ip]
        L.selmh_deck_box.argtypes = [C.c_void_p, dp, dp, dp]
        L.selmh_deck_atoms.argtypes = [C.c_void_p, dp, dp, dp, ip, ip]
        L.selmh_deck_bonds.argtypes = [C.c_void_p, ip, dp, dp]
        L.selmh_deck_external_force.argtypes = [C.c_void_p, dp]
        L.selmh_deck_forces.argtypes = [C.c_void_p, dp, dp, dp]
        L.selmh_deck_angles.argtypes = [C.c_void_p, ip, dp, dp]
        L.selmh_deck_pair_lj.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.selmh_dcd_open.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_void_p)]
        L.selmh_dcd_frame.argtypes = [C.c_void_p, C.c_longlong, dp, dp]
        L.selmh_dcd_close.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(dp)


def _i(a):
    return a.ctypes.data_as(ip)


class HostError(RuntimeError):
    pass


def model(path):
    m = Model()
    if lib().selmh_model_load(path.encode(), C.byref(m)):
        raise HostError(lib().selmh_last_error().decode())
    return m


class Deck:
    def __init__(self, script):
        self.h = C.c_void_p()
        if lib().selmh_deck_load(script.encode(), C.byref(self.h)):
            raise HostError(lib().selmh_last_error().decode())
        n, nb, na, nt = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        lib().selmh_deck_sizes(self.h, C.byref(n), C.byref(nb), C.byref(na), C.byref(nt))
        self.n, self.nb, self.na, self.nt = n.value, nb.value, na.value, nt.value
        lo, hi, dt = np.zeros(3), np.zeros(3), C.c_double()
        lib().selmh_deck_box(self.h, _d(lo), _d(hi), C.byref(dt))
        self.lo, self.hi, self.dt = lo, hi, dt.value
        self.x = np.zeros((self.n, 3)); self.v = np.zeros((self.n, 3)); self.m = np.zeros(self.n)
        self.type = np.zeros(self.n, np.int32); self.id = np.zeros(self.n, np.int32)
        lib().selmh_deck_atoms(self.h, _d(self.x), _d(self.v), _d(self.m), _i(self.type), _i(self.id))
        self.ij = np.zeros((self.nb, 2), np.int32); self.K = np.zeros(self.nb); self.r0 = np.zeros(self.nb)
        lib().selmh_deck_bonds(self.h, _i(self.ij), _d(self.K), _d(self.r0))
        self.fext = np.zeros((self.n, 3))
        lib().selmh_deck_external_force(self.h, _d(self.fext))
        self.ijk = np.zeros((self.na, 3), np.int32); self.aK = np.zeros(self.na); self.at0 = np.zeros(self.na)
        lib().selmh_deck_angles(self.h, _i(self.ijk), _d(self.aK), _d(self.at0))
        nt = self.nt
        self.lj = None
        eps, sig, rc, sp = np.zeros((nt, nt)), np.zeros((nt, nt)), np.zeros((nt, nt)), np.zeros(3)
        if lib().selmh_deck_pair_lj(self.h, _d(eps), _d(sig), _d(rc), _d(sp)) == 0:
            self.lj = (eps, sig, rc, sp)

    def forces(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        f = np.zeros_like(x)
        e = C.c_double()
        if lib().selmh_deck_forces(self.h, _d(x), _d(f), C.byref(e)):
            raise HostError(lib().selmh_last_error().decode())
        return f, e.value

    def close(self):
        if self.h:
            lib().selmh_deck_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_dcd(path, frames, every, dt, box):
    """frames: list of (step, x [n][3])."""
    h = C.c_void_p()
    n = frames[0][1].shape[0]
    assert lib().selmh_dcd_open(path.encode(), n, every, dt, C.byref(h)) == 0
    b = np.asarray(box, dtype=np.float64)
    for step, x in frames:
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert lib().selmh_dcd_frame(h, step, _d(x), _d(b)) == 0
    lib().selmh_dcd_close(h)


def read_dcd(path):
    """Minimal reader: (icntrl, titles, natoms, frames[(cell, xyz)])."""
    import struct
    b = open(path, "rb").read()
    assert struct.unpack_from("<i", b, 0)[0] == 84 and b[4:8] == b"CORD"
    icntrl = list(struct.unpack_from("<20i", b, 8))
    off = 92
    n2 = struct.unpack_from("<i", b, off)[0]
    ntitle = struct.unpack_from("<i", b, off + 4)[0]
    titles = [b[off + 8 + 80 * t: off + 8 + 80 * (t + 1)] for t in range(ntitle)]
    off += 4 + n2 + 4
    natoms = struct.unpack_from("<i", b, off + 4)[0]
    off += 12
    frames = []
    while off < len(b):
        assert struct.unpack_from("<i", b, off)[0] == 48
        cell = struct.unpack_from("<6d", b, off + 4)
        off += 56
        xyz = []
        for _ in range(3):
            ln = struct.unpack_from("<i", b, off)[0]
            assert ln == 4 * natoms
            xyz.append(np.frombuffer(b, dtype="<f4", count=natoms, offset=off + 4))
            off += 8 + ln
        frames.append((cell, np.stack(xyz, axis=1)))
    return icntrl, titles, natoms, frames
