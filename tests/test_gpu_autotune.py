"""The FFT-shape autotuner (selm_autotune / selm_pool_autotune) and the
shapes it chooses between.

- Tuning times steps on the live systems and restores their device state:
  positions, velocities, forces, the fluid and the step counter are bitwise
  what they were before.
- The chosen shape is one this build has for the grid, and a second call is
  answered from the process cache.
- The plane path's two Z-pass shapes (17 and 9 columns per workgroup) give
  the same steps to FP32 rounding (rel 2e-6)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(s):
    x, v, fc, ft = s.beads()
    return [x, v, fc, ft, s.field(), np.array([s.timestep])]


def test_autotune_restores_state_and_caches(gpu_lib):
    from monomers2026_amd import selm, workloads
    w = workloads.chains1000_64()
    s = w.build_system(selm.PREC_FP32_BULK, True, seed=21)
    s.step(3)
    before = _state(s)
    t = s.autotune(4)
    print("autotune chains1000_64 fp32:", t, flush=True)
    assert t["passes"] in (3, 5) and t["zcols"] in ((9, 17) if t["passes"] == 3 else (0,))
    assert t["ms_per_step"] > 0
    after = _state(s)
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    assert s.fft_passes() == t["passes"]
    t2 = s.autotune(4)
    assert t2["ms_per_step"] == 0.0 and t2["passes"] == t["passes"] and t2["zcols"] == t["zcols"]
    s.step(2)
    assert s.timestep == 5
    s.close()


def test_pool_autotune_restores_state(gpu_lib):
    from monomers2026_amd import selm, workloads
    ws = workloads.ensemble(n=6)
    fps = [selm.footprint(w.params(selm.PREC_FP32_BULK, True, seed=50 + i), w.M, w.ij.shape[0])
           for i, w in enumerate(ws)]
    pool = selm.Pool(int(sum(fps) * 1.05) + (16 << 20), len(ws), precision=selm.PREC_FP32_BULK)
    slots = [pool.admit(w.params(selm.PREC_FP32_BULK, True, seed=50 + i), w.x, w.v, w.m, w.ij, w.K, w.r0)
             for i, w in enumerate(ws)]
    pool.step(2)
    pool.sync()
    before = [(pool.beads(sl), pool.field(sl)) for sl in slots]
    t = pool.autotune(3)
    print("pool autotune:", t, flush=True)
    after = [(pool.beads(sl), pool.field(sl)) for sl in slots]
    for (b0, f0), (b1, f1) in zip(before, after):
        for a, b in zip(b0, b1):
            assert np.array_equal(a, b)
        assert np.array_equal(f0, f1)
    pool.close()


def test_pool_store_shapes_give_identical_bits(gpu_lib, monkeypatch):
    """A pool of six systems of the ensemble mix (edges 16-64) on the plane
    path, stepped with the bulk outputs stored at device scope and stored
    plainly: every system's beads and field agree bit for bit."""
    from monomers2026_amd import selm, workloads
    ws = workloads.ensemble(n=6)
    out = {}
    for wt in ("1", "0"):
        monkeypatch.setenv("SELM_WT_STORES", wt)
        fps = [selm.footprint(w.params(selm.PREC_FP32_BULK, True, seed=70 + i), w.M, w.ij.shape[0])
               for i, w in enumerate(ws)]
        pool = selm.Pool(int(sum(fps) * 1.05) + (16 << 20), len(ws), precision=selm.PREC_FP32_BULK)
        slots = [pool.admit(w.params(selm.PREC_FP32_BULK, True, seed=70 + i), w.x, w.v, w.m, w.ij, w.K, w.r0)
                 for i, w in enumerate(ws)]
        pool.step(6)
        pool.sync()
        out[wt] = [(pool.beads(sl), pool.field(sl)) for sl in slots]
        pool.close()
        print("pool store shape %s: stepped" % wt, flush=True)
    monkeypatch.delenv("SELM_WT_STORES", raising=False)
    for (b0, f0), (b1, f1) in zip(out["1"], out["0"]):
        for a, b in zip(b0, b1):
            assert np.array_equal(a, b)
        assert np.array_equal(f0, f1)


def test_plane_z_shapes_agree(gpu_lib, monkeypatch):
    from monomers2026_amd import selm, workloads
    w = workloads.chains1000_64()
    out = {}
    for cols in ("17", "9"):
        monkeypatch.setenv("SELM_ZQ_COLS", cols)
        s = w.build_system(selm.PREC_FP32_BULK, True, seed=8)
        assert s.fft_passes() == 3
        s.step(10)
        out[cols] = (s.beads()[1], s.field())
        s.close()
    monkeypatch.delenv("SELM_ZQ_COLS", raising=False)
    for a, b in zip(out["17"], out["9"]):
        e = np.abs(a - b).max() / max(np.abs(b).max(), 1.0)
        print("plane z shapes 17 vs 9: rel %.2e" % e, flush=True)
        assert e < 2e-6


def test_spread_shapes_give_identical_bits(gpu_lib, monkeypatch):
    """The spread shapes the autotuner chooses between -- split threshold
    64, 32 or none, and the default, 128 or 1024 thermal-forcing workgroups
    in the spread launch, and the plane path's two store shapes (device-scope
    or plain stores of the bulk outputs) -- give bitwise identical steps: the
    spread sums are order-independent fixed-point sums, each forcing mode is
    drawn by one thread whatever the grid, and a store's scope changes only
    where the value is cached.  4000 beads on 64^3 (split list active)."""
    from monomers2026_amd import selm, workloads
    rng = np.random.default_rng(31)
    x, ij = workloads.chains(rng, 40, 100, 64)
    w = workloads.Workload("dense4000_64", 64, x, ij)
    ref = None
    for sa, wg, wt in (("64", "", "1"), ("32", "", "1"), ("0", "", "1"), ("64", "128", "1"),
                       ("128", "1024", "1"), ("64", "", "0"), ("0", "1024", "0")):
        monkeypatch.setenv("SELM_SPLIT_AT", sa)
        monkeypatch.setenv("SELM_FORCING_WG", wg)
        monkeypatch.setenv("SELM_WT_STORES", wt)
        s = w.build_system(selm.PREC_FP32_BULK, True, seed=9)
        shape = s.spread_shape()
        assert shape[0] == int(sa)
        assert shape[1] == (int(wg) if wg else shape[1]) and shape[1] > 0
        s.step(10)
        got = _state(s)
        assert s.status() == 0
        s.close()
        print("spread shape", shape, "device-scope stores", wt, flush=True)
        if ref is None:
            ref = got
        else:
            for a, b in zip(ref, got):
                assert np.array_equal(a, b), shape
    monkeypatch.delenv("SELM_SPLIT_AT", raising=False)
    monkeypatch.delenv("SELM_FORCING_WG", raising=False)
    monkeypatch.delenv("SELM_WT_STORES", raising=False)


def test_autotune_reports_spread_shape(gpu_lib):
    from monomers2026_amd import selm, workloads
    rng = np.random.default_rng(32)
    x, ij = workloads.chains(rng, 30, 100, 64)
    w = workloads.Workload("dense3000_64", 64, x, ij)
    s = w.build_system(selm.PREC_FP32_BULK, True, seed=10)
    s.step(2)
    before = _state(s)
    t = s.autotune(4)
    print("autotune dense3000_64:", t, flush=True)
    assert t["split_at"] in (0, 32, 64, 128)
    assert (t["forcing_workgroups"] > 0) == (t["passes"] == 3)
    assert s.spread_shape() == (t["split_at"], t["forcing_workgroups"])
    after = _state(s)
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    s.close()
