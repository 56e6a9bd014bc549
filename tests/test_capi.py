"""The C-ABI library loads and exports every entry point include/selm_hip.h
declares; host-only entry points work without a GPU."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "selm_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(selm_[a-z_0-9]+)\s*\(", txt)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("selm_create", "selm_step", "selm_destroy", "selm_pool_admit", "selm_pool_step"):
        assert must in names


def test_library_exports_all_declared():
    from monomers2026_amd import selm
    L = selm.load()
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(_declared()) == sorted(selm.EXPORTS)


def test_host_only_entry_points():
    from monomers2026_amd import selm
    p = selm.make_params((32, 32, 32), (0, 0, 0), (32, 32, 32))
    assert p.seed == 123456 and p.f4_quirk == 1 and p.stochastic == 0
    fp = selm.footprint(p, 64, 63)
    assert fp > 3 * 32 ** 3 * 8 * 3
    p32 = selm.make_params((32, 32, 32), (0, 0, 0), (32, 32, 32), precision=selm.PREC_FP32_BULK)
    assert selm.footprint(p32, 64, 63) < fp


@pytest.mark.parametrize("edge", [3, 7, 11, 14, 1024])
def test_unsupported_edges_rejected(edge):
    from monomers2026_amd import selm
    p = selm.make_params((edge, 32, 32), (0, 0, 0), (edge, 32, 32))
    with pytest.raises(selm.SelmError) as e:
        selm.footprint(p, 1, 0)
    assert e.value.code == selm.SELM_ERR_UNSUPPORTED


def test_check_beads_rejects_non_finite_inputs():
    """selm_check_beads (the check selm_set_beads and selm_pool_admit apply)
    refuses NaN/Inf positions or velocities and masses that are not finite
    and > 0, naming the first offending bead; it runs without a device."""
    import numpy as np
    from monomers2026_amd import selm
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 8, (10, 3))
    v = rng.standard_normal((10, 3))
    m = np.ones(10)
    assert selm.check_beads(x, v, m) == -1
    assert selm.check_beads(x, None, m) == -1
    for bad, (arr, val) in enumerate([("x", np.nan), ("v", np.inf), ("m", 0.0), ("x", -np.inf), ("m", np.nan),
                                      ("m", -1.0), ("v", np.nan)]):
        xx, vv, mm = x.copy(), v.copy(), m.copy()
        i = 1 + bad
        if arr == "x":
            xx[i, bad % 3] = val
        elif arr == "v":
            vv[i, bad % 3] = val
        else:
            mm[i] = val
        assert selm.check_beads(xx, vv, mm) == i, (arr, val)


def test_msd_noalloc_rejects_requests_beyond_its_scratch():
    """selm_msd_noalloc holds its partial sums in a fixed array of 65536
    doubles: nlags = 65537 is the largest that fits; beyond it the call
    returns SELM_ERR_SCRATCH before touching a device (the pointers here are
    never dereferenced).  Invalid arguments are SELM_ERR_ARG as in selm_msd."""
    import numpy as np
    from monomers2026_amd import selm
    L = selm.load()
    box = np.array([6.0] * 3).ctypes.data_as(C.POINTER(C.c_double))
    fake = C.c_void_p(0x1000)
    assert L.selm_msd_noalloc(fake, 100000, 4, box, 65538, 1, fake, None) == selm.SELM_ERR_SCRATCH
    assert L.selm_msd_noalloc(fake, 100000, 4, box, 1 << 20, 1, fake, None) == selm.SELM_ERR_SCRATCH
    assert L.selm_msd_noalloc(None, 100, 4, box, 10, 1, fake, None) == selm.SELM_ERR_ARG
    assert L.selm_msd_noalloc(fake, 100, 4, box, 10, 0, fake, None) == selm.SELM_ERR_ARG


def test_resume_entry_points_reject_null_handles():
    """selm_set_timestep / selm_pool_set_field / selm_pool_set_timestep return
    SELM_ERR_ARG for a null context or pool without touching a device."""
    from monomers2026_amd import selm
    L = selm.load()
    assert L.selm_set_timestep(None, 0) == selm.SELM_ERR_ARG
    assert L.selm_pool_set_timestep(None, 0, 0) == selm.SELM_ERR_ARG
    assert L.selm_pool_set_field(None, 0, 0, None) == selm.SELM_ERR_ARG


def test_footprint_follows_bead_density():
    """The spread buckets hold four times a generous mean of entries per
    tile (at least 256 slots): a dense system's footprint grows with its bead
    count beyond what the per-bead arrays alone add (host-only)."""
    from monomers2026_amd import selm
    p = selm.make_params((24, 24, 24), (0, 0, 0), (24.0,) * 3, precision=selm.PREC_FP32_BULK)
    base = selm.footprint(p, 100, 0)          # 27 tiles: 256 slots each
    sparse = selm.footprint(p, 200, 0)        # still 256 slots
    dense = selm.footprint(p, 8000, 0)        # 9536 slots per tile
    per_bead = (sparse - base) / 100.0
    assert per_bead < 27 * 8 + 400           # bead arrays and the overflow list only
    buckets = dense - base - per_bead * 7900
    # 8000 beads over 27 tiles: 4 * 8 * 8000 / 27 -> 9536 slots of 48 bytes
    # (the per-bead estimate carries the 256-byte alignment of each array)
    assert abs(buckets - 27 * (9536 - 256) * 48) < 0.01 * buckets, buckets
