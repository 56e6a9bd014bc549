"""Exact resume from a snapshot (x, v, u, step count), through the C ABI.

The thermal draws of step n are keyed by n (counter-based Philox), so a
system rebuilt from a snapshot taken after n steps -- selm_set_beads,
selm_init, selm_set_field, selm_set_timestep(n) -- must continue bit for bit
as the run the snapshot came from.  Checked with noise on in FP64 and
FP32-bulk (plane FFT path at 64^3), for a fresh system, for a system rewound
in place (its spread tiles carry stamps of later steps than the snapshot),
and for a system resumed inside a pool.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_SNAP, N_MORE = 20, 15


def _state(s):
    x, v, fc, ft = s.beads()
    return x, v, fc, ft, s.field()


def _restore(s, w, snap):
    x, v, u, n = snap
    s.set_beads(x, v, w.m)
    s.set_bonds(w.ij, w.K, w.r0)
    if w.ext_idx is not None:
        s.set_external_force(w.ext_idx, w.ext_F)
    s.init()
    s.set_field(u)
    s.set_timestep(n)


def _assert_same(a, b, what):
    for name, p, q in zip(("x", "v", "f_cons", "f_total", "u"), a, b):
        assert np.array_equal(p, q), "%s: %s differs (max %.3e)" % (what, name, np.abs(p - q).max())


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_resume_fresh_and_rewound(gpu_lib, prec):
    from monomers2026_amd import selm, workloads
    P = selm.PREC_FP64 if prec == "fp64" else selm.PREC_FP32_BULK
    w = workloads.chains1000_64()
    a = w.build_system(P, True, seed=77)
    a.step(N_SNAP)
    x, v, _, _ = a.beads()
    snap = (x, v, a.field(), a.timestep)
    assert snap[3] == N_SNAP
    a.step(N_MORE)
    ref = _state(a)
    assert a.timestep == N_SNAP + N_MORE

    # a fresh system
    b = w.build_system(P, True, seed=77)
    _restore(b, w, snap)
    assert b.timestep == N_SNAP
    b.step(N_MORE)
    _assert_same(_state(b), ref, "%s fresh resume" % prec)
    assert b.timestep == N_SNAP + N_MORE
    b.close()

    # the same system rewound to the snapshot and run again
    _restore(a, w, snap)
    a.step(N_MORE)
    _assert_same(_state(a), ref, "%s rewound resume" % prec)
    # the fluid and step count alone rewound (beads left at step 35): the
    # stale tile stamps of steps 21..35 must not be read as current
    a.set_field(snap[2])
    a.set_timestep(N_SNAP)
    a.step(1)
    assert a.timestep == N_SNAP + 1
    assert a.status() == 0
    print("%s resume: bitwise equal after %d + %d steps" % (prec, N_SNAP, N_MORE), flush=True)
    a.close()


def test_set_timestep_rejects(gpu_lib):
    from monomers2026_amd import selm, workloads
    w = workloads.chain64_32()
    s = selm.System(w.params(selm.PREC_FP64, True))
    s.set_beads(w.x, w.v, w.m)
    # before init
    assert s.L.selm_set_timestep(s.h, 3) == selm.SELM_ERR_STATE
    s.set_bonds(w.ij, w.K, w.r0)
    s.init()
    assert s.L.selm_set_timestep(s.h, -1) == selm.SELM_ERR_ARG
    assert s.L.selm_set_timestep(s.h, 1 << 31) == selm.SELM_ERR_ARG
    s.set_timestep((1 << 31) - 2)
    assert s.timestep == (1 << 31) - 2
    s.close()


def test_resume_in_pool(gpu_lib):
    from monomers2026_amd import selm, workloads
    w = workloads.chain64_32()
    other = workloads.chains1000_64()
    P = selm.PREC_FP32_BULK
    pa = w.params(P, True, seed=5)
    po = other.params(P, True, seed=6)
    pool = selm.Pool(2 << 30, 4, precision=P)
    sa = pool.admit(pa, w.x, w.v, w.m, w.ij, w.K, w.r0)
    so = pool.admit(po, other.x, other.v, other.m, other.ij, other.K, other.r0)
    pool.step(N_SNAP)
    x, v, _, _ = pool.beads(sa)
    u = pool.field(sa)
    assert pool.timestep(sa) == N_SNAP
    pool.step(N_MORE)
    ref = pool.beads(sa) + (pool.field(sa),)
    # a second pool: the snapshot admitted beside a different system
    pool2 = selm.Pool(2 << 30, 4, precision=P)
    s2o = pool2.admit(po, other.x, other.v, other.m, other.ij, other.K, other.r0)
    s2 = pool2.admit(pa, x, v, w.m, w.ij, w.K, w.r0)
    pool2.set_field(s2, u)
    pool2.set_timestep(s2, N_SNAP)
    pool2.step(N_MORE)
    got = pool2.beads(s2) + (pool2.field(s2),)
    _assert_same(got, ref, "pool resume")
    assert pool2.timestep(s2) == N_SNAP + N_MORE and pool2.timestep(s2o) == N_MORE
    assert pool2.status(s2) == 0 and pool.status(sa) == 0
    pool.close()
    pool2.close()
