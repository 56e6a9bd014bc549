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

import numpy as np

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


def stage_bytes(w, s, passes=5):
    """Algorithmic HBM bytes per launch of each stage (s = bulk float size).

    Per-unit figures: a grid value is s bytes, a complex mode 2s; each pass
    reads its input once and writes its output once.  Beads: the 4x4x4
    stencil of 3 components (192 values) plus FP64 state.  With the plane
    FFT path (passes == 3) the X and Y transforms of a plane are one pass, so
    the Y stages move nothing of their own; the spread launch also draws the
    thermal forcing (its 3 Nc complex values written), and the XYi launch
    evaluates the bond forces (no angle or lj/cut terms here), so those two
    stages have no kernel of their own."""
    N = w.edge ** 3
    Nc = (w.edge // 2 + 1) * w.edge * w.edge
    M = w.M
    plane = passes == 3
    ypass = 0 if plane else 2 * 3 * Nc * 2 * s
    noise = 3 * Nc * 2 * s
    spread = M * (24 + 3 * s) + 192 * M * s
    bonds = M * (48 + 24 * 2 + 24)
    return {
        "thermal_noise": 0 if plane else noise,
        # stencil values, x read, bead state (x, v, f, m, t_f, s, f_total),
        # one 48-byte bucket entry per (tile, bead) pair
        "bead_update": M * (192 * s + 24 + 152) + 48 * tile_pairs(w),
        "spread": spread + (noise if plane else 0),
        "ustar_fft_x": 3 * N * s + 3 * Nc * 2 * s,
        "fft_y": ypass,
        "fft_z_projection": 3 * 3 * Nc * 2 * s,
        "ifft_y": ypass,
        "ifft_x": 3 * Nc * 2 * s + 3 * N * s + (bonds if plane else 0),
        "bond_forces": 0 if plane else bonds,
    }


def reference_stage_bytes(w, s):
    """Bytes per step of the reference's own stage decomposition, each stage
    reading its inputs once and writing its outputs once (cuda/SELM.cu:73-90
    with cuda/SELM_Eulerian_Period.cu:96-153): u* 9Ns, div 4Ns, r2c Ns+2sNc,
    the Poisson multiplier 4sNc, c2r Ns+2sNc, grad 7Ns, mean 9Ns, thermal
    field 6sNc + 3(2sNc + Ns) + 6Ns, interpolation M(24 + 192s), spread
    M(24 + 3s) + 192Ms, bead state 152M.  A fused implementation moves fewer
    bytes; this is the yardstick the reference's structure sets."""
    N = w.edge ** 3
    Nc = (w.edge // 2 + 1) * w.edge * w.edge
    M = w.M
    grid = (9 + 4 + 1 + 1 + 7 + 9 + 3 + 6) * N * s + (2 + 4 + 2 + 6 + 6) * s * Nc
    beads = M * (24 + 192 * s) + M * (24 + 3 * s) + 192 * M * s + 152 * M
    return grid + beads


# HBM bytes per launch of each stage's kernel at dilute20k_128 in FP32-bulk,
# from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE passes over this script
# (--fp32-only --copy-calibration 256), each scaled by what the counter reads
# for a copy kernel of known size (FETCH_SIZE x2.0, WRITE_SIZE x1.0); the
# per-kernel table is profiles/pmc_traffic_dilute20k_128.md.  The step runs
# the plane FFT path (k_xyf, k_zq, k_xyi): the Y stages have no kernel, the
# spread launch draws the thermal forcing and the XYi launch evaluates the
# bond forces.
PMC_TRAFFIC_FP32 = {
    "thermal_noise": 0,
    "bead_update": 31809740,
    "spread": 47084664,
    "ustar_fft_x": 68842014,
    "fft_y": 0,
    "fft_z_projection": 83422274,
    "ifft_y": 0,
    "ifft_x": 53554104,
    "bond_forces": 0,
}
COPY_CEILING_GBS = 5970.0  # measured: torch elementwise copy of 256 MiB in the same counter runs

# roofline classes: which stages are the spread, the FFT and the interpolation
CLASSES = {
    "spread": ["spread"],
    "fft": ["ustar_fft_x", "fft_y", "fft_z_projection", "ifft_y", "ifft_x"],
    "interp": ["bead_update"],
}


def tile_pairs(w):
    """(tile, bead) pairs of the spread's buckets at the starting positions:
    per axis the 8-cell tiles met by cells floor(s - 1/2) - 1 .. floor(s) + 2
    (the union of the three staggered stencils)."""
    L = w.edge
    te = 8
    xs = np.asarray(w.x, np.float64)
    lo = (np.floor(xs - 0.5) - 1).astype(np.int64)
    hi = np.floor(xs).astype(np.int64) + 2
    n = np.ones(xs.shape[0], np.int64)
    for a in range(3):
        cells = lo[:, a, None] + np.arange(6)[None, :]
        ok = cells <= hi[:, a, None]
        tiles = np.where(ok, (cells % L) // te, -1)
        n *= np.array([np.unique(r[r >= 0]).size for r in tiles])
    return int(n.sum())


def unique_stencil_bytes(w, s):
    """Compulsory stencil bytes: the distinct (component, cell) values that the
    4x4x4 stencils of all beads touch at the starting positions (box [0, L),
    dx = 1), times s bytes.  Chains overlap their stencils, so this is the
    least the gather can read and the spread can write per step."""
    L = w.edge
    xs = np.asarray(w.x, np.float64)
    l = np.arange(4)
    n = 0
    for d in range(3):
        r = xs - np.array([0.5 if a != d else 0.0 for a in range(3)])
        p = np.floor(r).astype(np.int64) - 1
        cx = (p[:, 0, None, None, None] + l[None, :, None, None]) % L
        cy = (p[:, 1, None, None, None] + l[None, None, :, None]) % L
        cz = (p[:, 2, None, None, None] + l[None, None, None, :]) % L
        n += np.unique(((cz * L + cy) * L + cx).ravel()).size
    return n * s


GRAPH_WARMUP = 5  # steps that capture and instantiate both step graphs (1 step and 4 steps)


def time_steps(sysg, ev, steps, warmup, runs=1):
    """Median over `runs` (odd) timed regions of `steps` graph replays each.
    The warm-up is at least GRAPH_WARMUP steps, so both graphs exist before
    the first timed region."""
    assert runs % 2 == 1
    sysg.step(max(warmup, GRAPH_WARMUP))
    sysg.sync()
    st = sysg.stream
    res = []
    for _ in range(runs):
        a, b = ev.new(), ev.new()
        t0 = time.perf_counter()
        ev.record(a, st)
        sysg.step(steps)
        ev.record(b, st)
        ms = ev.elapsed_ms(a, b)
        res.append((ms, time.perf_counter() - t0))
    res.sort()
    return res[len(res) // 2]


def displacement_error(w, steps=1000, seed=2026):
    """FP32-bulk against the FP64 GPU path (equal to the oracle to 1e-12) on
    the headline workload, outside the timed regions: max-norm bead
    displacement error from x^0 (minimum image), relative to the largest
    FP64 displacement, after 100 and `steps` steps."""
    from monomers2026_amd import selm
    s32 = w.build_system(selm.PREC_FP32_BULK, True, seed=seed)
    s64 = w.build_system(selm.PREC_FP64, True, seed=seed)
    x0 = s64.beads()[0]
    L = float(w.edge)
    out = {"steps": [], "rel_err": [], "max_displacement": []}
    done = 0
    for target in (100, steps):
        s32.step(target - done)
        s64.step(target - done)
        done = target
        d32 = s32.beads()[0] - x0
        d64 = s64.beads()[0] - x0
        d32 -= L * np.round(d32 / L)
        d64 -= L * np.round(d64 / L)
        out["steps"].append(target)
        out["rel_err"].append(float(np.abs(d32 - d64).max() / np.abs(d64).max()))
        out["max_displacement"].append(float(np.abs(d64).max()))
    out["status"] = {"fp32": s32.status(), "fp64": s64.status()}
    s32.close()
    s64.close()
    return out


def served_leg(duration=15.0, rate=20.0):
    """A short fixed-seed run of the served ensemble (Poisson arrivals, edges
    32-128, 100-5000 beads in chains of 50, 200-2000 steps each, FP32 bulk,
    continuous batching over a 16 GiB pool of 64 slots); the serving
    configuration asks for 10 minutes (python -m monomers2026_amd.serve
    --duration 600)."""
    from monomers2026_amd import serve
    rep = serve.run(duration=duration, rate=rate, budget_gb=16.0, slots=64, seed=5, progress_s=0)
    rep["note"] = "%.0f s of arrivals, then drained; the 10-minute run is the serving configuration" % duration
    return rep


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, budget_s=10.0):
    """The FP64 oracle (same stage decomposition as the reference's CPU path)
    timed on this host: all cores (up to 16) for the reported value, and
    1 thread beside it.  Bounded samples of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    ncpu = os.cpu_count() or 1
    out = {}
    for threads in (min(16, ncpu), 1):
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
        out[threads] = (n, el)
    threads = min(16, ncpu)
    n, el = out[threads]
    n1, el1 = out[1]
    return {"value": round(w.M * n / el, 1), "unit": "bead-steps/s", "cores": threads, "kind": "port",
            "sample": "%d steps of dilute20k_128 (FP64 oracle, OpenMP %d threads) after 1 warm-up step, %.1f s"
                      % (n, threads, el),
            "ms_per_step": round(1e3 * el / n, 3),
            "one_thread": {"value": round(w.M * n1 / el1, 1), "ms_per_step": round(1e3 * el1 / n1, 3),
                           "sample": "%d steps, %.1f s" % (n1, el1)},
            "host": {"nproc": ncpu, "model": cpu_model()}}


def copy_calibration(mib):
    """5 launches of an elementwise device copy (torch's neg kernel) over mib
    MiB of float32: each reads and writes exactly mib MiB, the known-size
    reference for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    del a, b


def ensemble_leg(ev, steps=500, warmup=10):
    """256 independent systems (edge in {16,24,32,48,64}, one chain of 32-512
    beads, mu ~ U[0.5, 2]) in one pool, FP32-bulk, noise on: per-step latency
    of the batched step from HIP events around each step, and throughput."""
    from monomers2026_amd import selm, workloads
    ws = workloads.ensemble()
    fps = []
    for i, w in enumerate(ws):
        fps.append(selm.footprint(w.params(selm.PREC_FP32_BULK, True, seed=1000 + i), w.M, w.ij.shape[0]))
    budget = int(sum(fps) * 1.02) + (64 << 20)
    pool = selm.Pool(budget, len(ws), precision=selm.PREC_FP32_BULK)
    for i, w in enumerate(ws):
        sl = pool.admit(w.params(selm.PREC_FP32_BULK, True, seed=1000 + i), w.x, w.v, w.m, w.ij, w.K, w.r0)
        if sl is None:
            raise RuntimeError("ensemble system %d does not fit the pool" % i)
    tune = pool.autotune(10)
    pool.step(warmup)
    pool.sync()
    st = pool.stream
    evs = [ev.new() for _ in range(steps + 1)]
    ev.record(evs[0], st)
    for k in range(steps):
        pool.step(1)
        ev.record(evs[k + 1], st)
    lat = sorted(ev.elapsed_ms(evs[k], evs[k + 1]) for k in range(steps))
    total_ms = ev.elapsed_ms(evs[0], evs[-1])
    used, _, active = pool.stats()
    pool.close()
    M = sum(w.M for w in ws)
    G = sum(w.edge ** 3 for w in ws)

    def pct(q):
        return lat[min(len(lat) - 1, int(round(q * (len(lat) - 1))))]

    return {"systems": active, "beads": M, "grid_points": G, "steps": steps, "hbm_bytes": used,
            "p50_ms": round(pct(0.5), 5), "p99_ms": round(pct(0.99), 5), "max_ms": round(lat[-1], 5),
            "autotune": tune,
            "bead_steps_per_s": round(M * steps / (total_ms * 1e-3), 1),
            "grid_point_steps_per_s": round(G * steps / (total_ms * 1e-3), 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--runs", type=int, default=5, help="timed regions (odd); the median is reported")
    ap.add_argument("--no-serve", action="store_true", help="skip the short served-ensemble leg")
    ap.add_argument("--serve-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--stage-steps", type=int, default=50)
    ap.add_argument("--fp32-only", action="store_true",
                    help="time only the FP32-bulk steps (no FP64 leg, stage split, ensemble or CPU baseline); "
                         "for counter runs under rocprofv3 --pmc")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run 5 launches of a plain device copy of MIB MiB (read MIB, write MIB) "
                         "so that a counter run can scale FETCH_SIZE/WRITE_SIZE to bytes")
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
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)

    # FP32-bulk headline; the FFT pass shapes are picked by the autotuner
    # (timed on the system itself, whose state it restores)
    s32 = w.build_system(selm.PREC_FP32_BULK, stochastic=True, seed=2026)
    tune32 = s32.autotune(20)
    runs = max(1, args.runs | 1)
    ms32, wall32 = time_steps(s32, ev, args.steps, args.warmup, 1 if args.fp32_only else runs)
    if args.fp32_only:
        s32.close()
        print(json.dumps({"fp32_only": True, "steps": args.steps, "ms_per_step": round(ms32 / args.steps, 5)}),
              flush=True)
        return
    print("fp32-bulk: %.4f ms/step" % (ms32 / args.steps), file=sys.stderr, flush=True)
    passes32 = s32.fft_passes()
    stages32 = s32.stage_times(args.stage_steps)
    s32.close()
    # FP64 path on the same workload
    s64 = w.build_system(selm.PREC_FP64, stochastic=True, seed=2026)
    steps64 = max(1, args.steps // 2)
    ms64, wall64 = time_steps(s64, ev, steps64, args.warmup, max(3, (runs // 2) | 1))
    passes64 = s64.fft_passes()
    stages64 = s64.stage_times(args.stage_steps)
    s64.close()
    print("fp64: %.4f ms/step" % (ms64 / steps64), file=sys.stderr, flush=True)

    per32 = ms32 / args.steps
    per64 = ms64 / steps64
    bead_steps32 = w.M * args.steps / (ms32 * 1e-3)
    bead_steps64 = w.M * steps64 / (ms64 * 1e-3)

    nst = args.stage_steps
    avg32 = {k: v / nst for k, v in stages32.items()}
    avg64 = {k: v / nst for k, v in stages64.items()}
    dom = max(avg32, key=avg32.get)
    b32 = stage_bytes(w, 4, passes32)
    achieved = b32[dom] / (avg32[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": PMC_TRAFFIC_FP32[dom],
            "algorithmic_bytes": b32[dom], "kernel_ms": round(avg32[dom], 5),
            "copy_ceiling_GBs": COPY_CEILING_GBS}
    classes = {}
    for c, names in CLASSES.items():
        by = sum(b32[k] for k in names)
        ms = sum(avg32[k] for k in names)
        gbs = by / (ms * 1e-3) / 1e9
        classes[c] = {"algorithmic_bytes": by, "ms": round(ms, 5), "GBs": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4)}
    step_bytes = sum(b32.values())
    ref_bytes = reference_stage_bytes(w, 4)
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
        "timing": "median of %d timed regions of %d steps (FP64: %d of %d), after >= %d warm-up steps"
                  % (runs, args.steps, max(3, (runs // 2) | 1), steps64, GRAPH_WARMUP),
        "grid_point_steps_per_s": round(N * args.steps / (ms32 * 1e-3), 1),
        "fp64": {"value": round(bead_steps64, 1), "ms_per_step": round(per64, 5), "steps": steps64},
        "fp32_over_fp64": round(bead_steps32 / bead_steps64, 3),
        "fft_passes": {"fp32": passes32, "fp64": passes64},
        "autotune_fp32": tune32,
        "wall_s": {"fp32": round(wall32, 4), "fp64": round(wall64, 4)},
        "stage_ms_fp32": {k: round(v, 5) for k, v in avg32.items()},
        "stage_ms_fp64": {k: round(v, 5) for k, v in avg64.items()},
        "class_roofline_fp32": classes,
        "step_algorithmic_GBs_fp32": round(step_bytes / (per32 * 1e-3) / 1e9, 1),
        "step_reference_stages_fp32": {"bytes": ref_bytes,
                                       "GBs": round(ref_bytes / (per32 * 1e-3) / 1e9, 1),
                                       "frac": round(ref_bytes / (per32 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": roof,
        "unique_stencil_bytes_fp32": unique_stencil_bytes(w, 4),
    }
    out["fp32_bulk_displacement_error"] = displacement_error(w)
    print("fp32-bulk displacement error vs fp64: %s" % out["fp32_bulk_displacement_error"]["rel_err"],
          file=sys.stderr, flush=True)
    if not args.no_ensemble:
        out["ensemble256"] = ensemble_leg(ev)
        print("ensemble: p50 %.4f ms p99 %.4f ms" % (out["ensemble256"]["p50_ms"], out["ensemble256"]["p99_ms"]),
              file=sys.stderr, flush=True)
    if not args.no_serve:
        out["served"] = served_leg(args.serve_seconds)
        lat = out["served"]["step_latency"]
        print("served: p50 %.4f ms p99 %.4f ms, %.3g bead-steps/s" % (lat["p50_ms"], lat["p99_ms"],
              out["served"]["bead_steps_per_s"]), file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
