"""selm_msd_noalloc (device buffers, partial sums in the library's fixed
device array, no allocation) against selm_msd (host buffers, allocating) and
the oracle, through the C ABI: two calls in a row on the same stream, the
second smaller and with lags that have no origin left (NaN), so stale
partial sums of the first call would show."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _walk(F, n, L, seed):
    rng = np.random.default_rng(seed)
    x = np.mod(np.cumsum(rng.normal(0.0, 0.3, (F, n, 3)), axis=0), L)
    return np.ascontiguousarray(x.transpose(0, 2, 1)).astype(np.float32)


def test_msd_noalloc_matches_msd(gpu_lib):
    import torch
    from monomers2026_amd import analysis, selm
    L = gpu_lib
    box_x = 6.0
    box = np.array([box_x] * 3)
    st = torch.cuda.current_stream()
    for call, (F, n, nlags, mult) in enumerate([(500, 300, 41, 3), (60, 20, 15, 4)]):
        print("selm_msd_noalloc call %d: F=%d natoms=%d nlags=%d mult=%d" % (call, F, n, nlags, mult), flush=True)
        fr = _walk(F, n, box_x, 11 + call)
        want = analysis.msd_frames(fr, box_x, nlags, mult)
        d_fr = torch.from_numpy(fr).cuda()
        d_out = torch.full((nlags,), -1.0, dtype=torch.float64, device="cuda")
        rc = L.selm_msd_noalloc(C.c_void_p(d_fr.data_ptr()), F, n, box.ctypes.data_as(C.POINTER(C.c_double)),
                                nlags, mult, C.c_void_p(d_out.data_ptr()), C.c_void_p(st.cuda_stream))
        assert rc == selm.SELM_OK, rc
        torch.cuda.synchronize()
        got = d_out.cpu().numpy()
        assert np.array_equal(got, want, equal_nan=True), (call, got, want)
        o = O.msd(fr, box_x, nlags, mult)
        assert np.array_equal(np.isnan(got), np.isnan(o))
        ok = ~np.isnan(o)
        err = np.abs(got[ok] - o[ok]).max() / np.abs(o[ok]).max()
        print("  bitwise equal to selm_msd; vs oracle %.2e; NaN lags %d" % (err, int(np.isnan(got).sum())),
              flush=True)
        assert err < 1e-12
    assert call == 1 and np.isnan(got).sum() > 0


def test_msd_noalloc_single_lag(gpu_lib):
    """nlags = 1: only msd[0] = 0 is written."""
    import torch
    from monomers2026_amd import selm
    fr = torch.from_numpy(_walk(10, 4, 6.0, 3)).cuda()
    out = torch.full((1,), 7.0, dtype=torch.float64, device="cuda")
    box = np.array([6.0] * 3)
    rc = gpu_lib.selm_msd_noalloc(C.c_void_p(fr.data_ptr()), 10, 4, box.ctypes.data_as(C.POINTER(C.c_double)), 1, 1,
                                  C.c_void_p(out.data_ptr()), None)
    assert rc == selm.SELM_OK
    torch.cuda.synchronize()
    assert out.item() == 0.0
