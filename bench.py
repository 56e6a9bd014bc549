"""Benchmark: SELM step throughput at 128^3 with 20k beads on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload dilute20k_128: 128^3 periodic grid (dx = 1, rho = mu = kB = T = 1,
dt = 0.01), 200 seeded random-walk chains x 100 beads (harmonic K = 100,
r0 = 1, m = 1), 1000 beads with a constant external force (0.1, 0, 0),
thermal forcing on.  One step = one full SELM time step of all 20k beads and
the 2.1M-cell grid.  Inputs are resident in HBM before the timed region; the
timed region replays the captured per-step hipGraph K times between two HIP
events on the system's stream.

Prints one JSON line: value = bead-steps/s of the FP32-bulk path (FP32 grid,
FFT, spread and interpolation; FP64 bead state), with the FP64 path, the
per-stage times, the roofline of the dominant kernel and the host-CPU oracle
baseline beside it.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure


class HipEvents:
    """Minimal hipEvent API over ctypes (events recorded on a given stream)."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.h.hipEventSynchronize.argtypes = [C.c_void_p]
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.h.hipEventDestroy.argtypes = [C.c_void_p]

    def new(self):
        e = C.c_void_p()
        assert self.h.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.h.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b):
        assert self.h.hipEventSynchronize(b) == 0
        t = C.c_float()
        assert self.h.hipEventElapsedTime(C.byref(t), a, b) == 0
        return t.value


def stage_bytes(w, s):
    """Algorithmic HBM bytes per launch of each stage (s = bulk float size).

    Per-unit figures: a grid value is s bytes, a complex mode 2s; each pass
    reads its input once and writes its output once.  Beads: the 4x4x4
    stencil of 3 components (192 values) plus FP64 state."""
    N = w.edge ** 3
    Nc = (w.edge // 2 + 1) * w.edge * w.edge
    M = w.M
    rec = 208 if s == 4 else 368  # spread record written per bead
    return {
        "thermal_noise": 3 * Nc * 2 * s,
        "bead_update": M * (192 * s + 24 + 152 + rec),
        "spread": M * (24 + 3 * s) + 192 * M * s,
        "ustar_fft_x": 3 * N * s + 3 * Nc * 2 * s,
        "fft_y": 2 * 3 * Nc * 2 * s,
        "fft_z_projection": 3 * 3 * Nc * 2 * s,
        "ifft_y": 2 * 3 * Nc * 2 * s,
        "ifft_x": 3 * Nc * 2 * s + 3 * N * s,
        "bond_forces": M * (48 + 24 * 2 + 24),
    }


def time_steps(sysg, ev, steps, warmup):
    sysg.step(warmup)
    sysg.sync()
    a, b = ev.new(), ev.new()
    st = sysg.stream
    t0 = time.perf_counter()
    ev.record(a, st)
    sysg.step(steps)
    ev.record(b, st)
    ms = ev.elapsed_ms(a, b)
    wall = time.perf_counter() - t0
    return ms, wall


def cpu_baseline(w, budget_s=12.0):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    threads = min(16, os.cpu_count() or 1)
    O.set_threads(threads)
    o = O.Oracle(*w.dims, w.lo, w.hi, rho=w.rho, mu=w.mu, kB=w.kB, T=w.T, dt=w.dt, seed=1, noise=True)
    o.set_beads(w.x, w.v, w.m)
    o.set_bonds(w.ij, w.K, w.r0)
    if w.ext_idx is not None:
        o.set_external_force(w.ext_idx, w.ext_F)
    o.init()
    t = time.perf_counter()
    o.step(1)
    per = time.perf_counter() - t
    n = max(2, min(200, int(budget_s / max(per, 1e-3))))
    t = time.perf_counter()
    o.step(n)
    el = time.perf_counter() - t
    o.close()
    return {"value": w.M * n / el, "unit": "bead-steps/s", "cores": threads, "kind": "port",
            "sample": "%d steps of dilute20k_128 (FP64 oracle, OpenMP %d threads) after 1 warm-up step, %.1f s"
                      % (n, threads, el),
            "ms_per_step": 1e3 * el / n}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stage-steps", type=int, default=50)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the SELM step of one system runs on one device and "
              "ensembles shard by system without exchange; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from monomers2026_amd import selm, workloads
    w = workloads.dilute20k_128()
    ev = HipEvents()
    N = w.edge ** 3

    # FP32-bulk headline
    s32 = w.build_system(selm.PREC_FP32_BULK, stochastic=True, seed=2026)
    ms32, wall32 = time_steps(s32, ev, args.steps, args.warmup)
    stages32 = s32.stage_times(args.stage_steps)
    s32.close()
    # FP64 path on the same workload
    s64 = w.build_system(selm.PREC_FP64, stochastic=True, seed=2026)
    steps64 = max(1, args.steps // 2)
    ms64, wall64 = time_steps(s64, ev, steps64, max(1, args.warmup // 2))
    stages64 = s64.stage_times(args.stage_steps)
    s64.close()

    per32 = ms32 / args.steps
    per64 = ms64 / steps64
    bead_steps32 = w.M * args.steps / (ms32 * 1e-3)
    bead_steps64 = w.M * steps64 / (ms64 * 1e-3)

    nst = args.stage_steps
    avg32 = {k: v / nst for k, v in stages32.items()}
    avg64 = {k: v / nst for k, v in stages64.items()}
    dom = max(avg32, key=avg32.get)
    b32 = stage_bytes(w, 4)
    achieved = b32[dom] / (avg32[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "algorithmic_bytes": b32[dom], "kernel_ms": round(avg32[dom], 5)}
    step_bytes = sum(b32.values())
    out = {
        "metric": "bead-steps/s at 128^3 FP32-bulk (GPU), SELM step, 20k beads",
        "value": round(bead_steps32, 1),
        "unit": "bead-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(per32, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-bulk (f32 grid/FFT/spread/interp, f64 beads)",
        "data": "synthetic: seeded random-walk chains, random-init fluid at rest",
        "config": {"workload": "dilute20k_128", "grid": [128, 128, 128], "beads": w.M, "chains": 200,
                   "beads_per_chain": 100, "external_force_beads": 1000, "noise": True, "dt": w.dt},
        "grid_point_steps_per_s": round(N * args.steps / (ms32 * 1e-3), 1),
        "fp64": {"value": round(bead_steps64, 1), "ms_per_step": round(per64, 5), "steps": steps64},
        "fp32_over_fp64": round(bead_steps32 / bead_steps64, 3),
        "wall_s": {"fp32": round(wall32, 4), "fp64": round(wall64, 4)},
        "stage_ms_fp32": {k: round(v, 5) for k, v in avg32.items()},
        "stage_ms_fp64": {k: round(v, 5) for k, v in avg64.items()},
        "step_algorithmic_GBs_fp32": round(step_bytes / (per32 * 1e-3) / 1e9, 1),
        "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
