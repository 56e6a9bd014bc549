"""Ensembles in one device arena (selm_pool): every system's result is the
same alone and in the mix, across admissions and retirements.

Bar: bit-identical x, v and u (each system's arithmetic, noise counters and
reduction order are its own; nothing is reduced across systems)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _systems(seed=11, n=4):
    from monomers2026_amd import workloads
    rng = np.random.default_rng(seed)
    edges = [16, 24, 32, 48, 30, 64]
    out = []
    for r in range(n):
        L = edges[r % len(edges)]
        M = int(rng.integers(20, 200))
        x, ij = workloads.chains(rng, 1, M, L)
        w = workloads.Workload("sys%d" % r, L, x, ij, mu=float(rng.uniform(0.5, 2.0)))
        out.append(w)
    return out


def _alone(w, prec, nsteps, seed):
    from monomers2026_amd import selm
    s = w.build_system(prec, True, seed=seed)
    s.step(nsteps)
    x, v, _, _ = s.beads()
    u = s.field()
    s.close()
    return x, v, u


def _admit(pool, w, prec, seed):
    return pool.admit(w.params(prec, True, seed=seed), w.x, w.v, w.m, w.ij, w.K, w.r0)


@pytest.mark.parametrize("prec", [0, 1])
def test_pool_matches_alone(gpu_lib, prec):
    from monomers2026_amd import selm
    ws = _systems()
    pool = selm.Pool(2 << 30, 16, precision=prec)
    slots = [_admit(pool, w, prec, 100 + i) for i, w in enumerate(ws)]
    assert None not in slots
    pool.step(25)
    for i, (w, sl) in enumerate(zip(ws, slots)):
        xa, va, ua = _alone(w, prec, 25, 100 + i)
        x, v, _, _ = pool.beads(sl)
        u = pool.field(sl)
        assert pool.timestep(sl) == 25
        assert np.array_equal(x, xa) and np.array_equal(v, va) and np.array_equal(u, ua), (i, w.edge)
    pool.close()


def test_pool_admit_retire_continuous(gpu_lib):
    from monomers2026_amd import selm
    prec = selm.PREC_FP32_BULK
    ws = _systems(seed=5, n=5)
    pool = selm.Pool(2 << 30, 8, precision=prec)
    s = [_admit(pool, ws[i], prec, 7 + i) for i in range(3)]
    pool.step(10)
    pool.retire(s[1])
    s4 = _admit(pool, ws[3], prec, 10)
    s5 = _admit(pool, ws[4], prec, 11)
    used, budget, active = pool.stats()
    assert active == 4 and 0 < used <= budget
    pool.step(10)
    for idx, sl, n in ((0, s[0], 20), (2, s[2], 20), (3, s4, 10), (4, s5, 10)):
        xa, va, ua = _alone(ws[idx], prec, n, 7 + idx)
        x, v, _, _ = pool.beads(sl)
        assert pool.timestep(sl) == n
        assert np.array_equal(x, xa) and np.array_equal(v, va), idx
        assert np.array_equal(pool.field(sl), ua), idx
    pool.close()


def test_pool_admission_by_footprint(gpu_lib):
    from monomers2026_amd import selm
    ws = _systems(seed=3, n=2)
    w = ws[0]
    p = w.params(selm.PREC_FP32_BULK, True)
    fp = selm.footprint(p, w.M, w.ij.shape[0])
    pool = selm.Pool(int(fp * 1.5), 4, precision=selm.PREC_FP32_BULK)
    a = _admit(pool, w, selm.PREC_FP32_BULK, 1)
    assert a is not None
    b = _admit(pool, w, selm.PREC_FP32_BULK, 2)
    assert b is None  # does not fit next to the first
    pool.retire(a)
    b = _admit(pool, w, selm.PREC_FP32_BULK, 2)
    assert b is not None
    pool.step(3)
    assert pool.timestep(b) == 3
    pool.close()
