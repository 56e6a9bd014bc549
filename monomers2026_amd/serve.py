"""Served ensembles: a request loop with continuous batching over one
device-resident pool of SELM systems.

Each request is one SELM run: a grid, a bead-spring system, a seed and a
number of steps.  The loop works on requests as follows:
- it keeps a FIFO queue of requests that have arrived;
- at every tick it admits each queued request that fits, by its HBM
  footprint, into the pool's arena and a free slot (a request that does not
  fit waits; later, smaller ones may pass it);
- it advances every admitted system by one step with one batched launch per
  stage;
- a system that has run its steps has its final bead state read back and is
  retired, freeing its slot and memory for the next arrival;
- every `status_every` ticks, and before a result is read, the pool's device
  status flags are checked, and a system whose flags are raised (a bead
  position or velocity gone non-finite, a displacement beyond half the box,
  a spread tile out of range) is evicted: retired with its flags recorded,
  while the other systems carry on unchanged.

Requests that arrive while others are on the GPU therefore join the running
batch at the next tick.  Each system keeps its own step counter, noise
counters and arithmetic, so its result is bit-identical to running it alone
(tests/test_gpu_serve.py).

Per-step latency is the GPU time of one batched step.  It is taken from HIP
events recorded on the pool's stream around every step.

    python -m monomers2026_amd.serve [--duration S] [--rate R] [--budget-gb G]
"""
import argparse
import ctypes as C
import json
import sys
import time
from collections import deque

import numpy as np

from . import selm, workloads


class Request:
    """One run: the system, its seed and steps; arrival in seconds (or ticks)."""

    def __init__(self, rid, workload, steps, seed, arrival=0.0):
        self.rid = rid
        self.w = workload
        self.steps = int(steps)
        self.seed = int(seed)
        self.arrival = arrival
        self.slot = None
        self.admitted_at = None
        self.done_at = None
        self.result = None  # (x, v) after `steps` steps
        self.failed = 0     # device status flags that evicted the run (0: none)


class _Events:
    """hipEvent timing on a stream (ctypes over libamdhip64)."""

    def __init__(self):
        h = C.CDLL("libamdhip64.so")
        h.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        h.hipEventSynchronize.argtypes = [C.c_void_p]
        h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        h.hipEventDestroy.argtypes = [C.c_void_p]
        self.h = h
        self.free = []

    def new(self):
        if self.free:
            return self.free.pop()
        e = C.c_void_p()
        if self.h.hipEventCreate(C.byref(e)) != 0:
            raise RuntimeError("hipEventCreate failed")
        return e

    def record(self, e, stream):
        if self.h.hipEventRecord(e, stream) != 0:
            raise RuntimeError("hipEventRecord failed")

    def elapsed(self, a, b):
        if self.h.hipEventSynchronize(b) != 0:
            raise RuntimeError("hipEventSynchronize failed")
        t = C.c_float()
        if self.h.hipEventElapsedTime(C.byref(t), a, b) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return t.value

    def release(self, e):
        self.free.append(e)


class Server:
    """Continuous batching of SELM runs over one selm_pool."""

    def __init__(self, hbm_budget, max_slots, precision=selm.PREC_FP32_BULK, time_steps=True, status_every=16):
        self.pool = selm.Pool(hbm_budget, max_slots, precision=precision)
        self.precision = precision
        self.queue = deque()
        self.running = {}  # slot -> (request, steps done)
        self.finished = []
        self.evicted = []
        self.status_every = status_every
        self.step_ms = []
        self.bead_steps = 0
        self.ticks = 0
        self.ev = _Events() if time_steps else None
        self._pending = []  # (event before, event after) not yet read

    def submit(self, req):
        self.queue.append(req)

    def _admit(self, now):
        waiting = deque()
        while self.queue:
            r = self.queue.popleft()
            w = r.w
            p = w.params(self.precision, True, seed=r.seed)
            sl = self.pool.admit(p, w.x, w.v, w.m, w.ij, w.K, w.r0, w.ext_idx, w.ext_F)
            if sl is None:
                waiting.append(r)
                continue
            r.slot = sl
            r.admitted_at = now
            self.running[sl] = [r, 0]
        self.queue = waiting

    def _drain_events(self, keep=0):
        while len(self._pending) > keep:
            a, b = self._pending.pop(0)
            self.step_ms.append(self.ev.elapsed(a, b))
            self.ev.release(a)
            self.ev.release(b)

    def _evict(self, sl, flags, now):
        r = self.running.pop(sl)[0]
        r.failed = flags
        r.done_at = now
        self.pool.retire(sl)
        self.evicted.append(r)

    def check_status(self, now=0.0):
        """Wait for the pool and evict every running system whose device
        status flags are raised; returns the evicted requests."""
        try:
            self.pool.sync()
            return []
        except selm.SelmError as e:
            if e.code != selm.SELM_ERR_DEVICE:
                raise
        out = []
        for sl in list(self.running):
            f = self.pool.status(sl)
            if f:
                out.append(self.running[sl][0])
                self._evict(sl, f, now)
        return out

    def tick(self, now=0.0):
        """Admit what fits, advance every running system by one step, and
        retire the systems that have run their steps.  Returns the number of
        systems stepped."""
        self._admit(now)
        if not self.running:
            return 0
        st = self.pool.stream
        if self.ev is not None:
            a, b = self.ev.new(), self.ev.new()
            self.ev.record(a, st)
        self.pool.step(1)
        if self.ev is not None:
            self.ev.record(b, st)
            self._pending.append((a, b))
            self._drain_events(keep=64)
        self.ticks += 1
        n = len(self.running)
        for sl in list(self.running):
            ent = self.running[sl]
            ent[1] += 1
            self.bead_steps += ent[0].w.M
        if self.status_every and self.ticks % self.status_every == 0:
            self.check_status(now)
        for sl in list(self.running):
            ent = self.running[sl]
            if ent[1] == ent[0].steps:
                r = ent[0]
                f = self.pool.status(sl)
                if f:
                    self._evict(sl, f, now)
                    continue
                x, v, _, _ = self.pool.beads(sl)
                r.result = (x, v)
                r.done_at = now
                self.pool.retire(sl)
                del self.running[sl]
                self.finished.append(r)
        return n

    def idle(self):
        return not self.queue and not self.running

    def close(self):
        if self.ev is not None:
            self.pool.sync()
            self._drain_events(0)
        self.pool.close()

    def latency(self):
        lat = sorted(self.step_ms)
        if not lat:
            return None

        def pct(q):
            return lat[min(len(lat) - 1, int(round(q * (len(lat) - 1))))]

        return {"steps": len(lat), "p50_ms": round(pct(0.5), 5), "p99_ms": round(pct(0.99), 5),
                "max_ms": round(lat[-1], 5)}


def random_request(rng, rid, edges=(32, 48, 64, 96, 128), mmin=100, mmax=5000, per_chain=50,
                   smin=200, smax=2000, arrival=0.0):
    """A request of the served-ensemble mix: edge from `edges`, 100-5000 beads
    in chains of 50, 200-2000 steps, mu = 1."""
    L = int(rng.choice(edges))
    nchains = max(1, int(rng.integers(mmin, mmax + 1)) // per_chain)
    x, ij = workloads.chains(rng, nchains, per_chain, L)
    w = workloads.Workload("req%d" % rid, L, x, ij)
    steps = int(rng.integers(smin, smax + 1))
    return Request(rid, w, steps, seed=10007 * rid + 1, arrival=arrival)


def make_requests(rng, duration, rate):
    """Poisson arrivals at `rate` per second over `duration` seconds, each a
    random_request; generated before the clock starts."""
    reqs = []
    t = rng.exponential(1.0 / rate)
    while t < duration:
        reqs.append(random_request(rng, len(reqs), arrival=t))
        t += rng.exponential(1.0 / rate)
        if len(reqs) % 1000 == 0:
            print("generated %d requests" % len(reqs), file=sys.stderr, flush=True)
    return reqs


def run(duration=30.0, rate=20.0, budget_gb=16.0, slots=64, seed=5, precision=selm.PREC_FP32_BULK,
        progress_s=30.0):
    """Replay Poisson arrivals at `rate` requests/s for `duration` seconds of
    wall clock, then drain.  Every `progress_s` seconds (0: never) a line of
    counts goes to stderr.  Returns the report dict."""
    rng = np.random.default_rng(seed)
    reqs = make_requests(rng, duration, rate)
    srv = Server(int(budget_gb * (1 << 30)), slots, precision=precision)
    t0 = time.perf_counter()
    i = 0
    peak = 0
    next_report = progress_s
    while True:
        now = time.perf_counter() - t0
        while i < len(reqs) and reqs[i].arrival <= now:
            srv.submit(reqs[i])
            i += 1
        if i == len(reqs) and srv.idle():
            break
        n = srv.tick(now)
        peak = max(peak, n)
        if n == 0:
            time.sleep(0.0005)
        if progress_s and now >= next_report:
            print("t=%.0f s: %d arrived, %d finished, %d running, %d queued, %d batched steps"
                  % (now, i, len(srv.finished), len(srv.running), len(srv.queue), srv.ticks),
                  file=sys.stderr, flush=True)
            next_report += progress_s
    srv.pool.sync()
    wall = time.perf_counter() - t0
    srv.close()
    lat = srv.latency()
    waits = [r.admitted_at - r.arrival for r in srv.finished]
    return {"requests": len(reqs), "finished": len(srv.finished), "evicted": len(srv.evicted),
            "wall_s": round(wall, 3),
            "arrival_rate_per_s": rate, "hbm_budget_gb": budget_gb, "slots": slots,
            "peak_concurrent": peak, "bead_steps_per_s": round(srv.bead_steps / wall, 1),
            "batched_steps": srv.ticks, "step_latency": lat,
            "admission_wait_s": {"p50": round(float(np.median(waits)), 4) if waits else None,
                                 "p99": round(float(np.percentile(waits, 99)), 4) if waits else None}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--duration", type=float, default=30.0)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--budget-gb", type=float, default=16.0)
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--seed", type=int, default=5)
    args = ap.parse_args()
    rep = run(args.duration, args.rate, args.budget_gb, args.slots, args.seed)
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
