"""GPU parity at the largest grids the library takes, against the FP64 oracle.

Every other parity case has edges of at most 128.  Here:
- (256, 192, 160) runs the compile-time-length FFT kernels of the largest
  size class (lengths up to 256), five passes since nx != ny;
- (512, 20, 18) has the largest edge the library accepts (512), which only
  the runtime-radix kernels handle.
Both start from a random (not divergence-free) u^n with 60 beads in a chain
and thermal forcing on, and step twice.  Bars as in test_gpu_parity: FP64
rel <= 1e-12 in x, v, f_cons, f_total and u; FP32-bulk rel <= 2e-6 in u and
v after one step.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _rel(a, b, floor=1.0):
    return np.abs(a - b).max() / max(np.abs(b).max(), floor)


def _case(dims, seed):
    rng = np.random.default_rng(seed)
    M = 60
    x = rng.uniform(0, min(dims), (M, 3))
    v = rng.standard_normal((M, 3))
    m = rng.uniform(0.5, 2.0, M)
    ij = np.array([(i, i + 1) for i in range(M - 1)], np.int32)
    K = np.full(M - 1, 40.0)
    r0 = np.full(M - 1, 1.2)
    u0 = rng.standard_normal((3, dims[2], dims[1], dims[0]))
    return x, v, m, ij, K, r0, u0


KW = dict(rho=1.1, mu=0.9, kB=1.0, T=1.5, dt=0.02, seed=77)


@pytest.mark.parametrize("dims", [(256, 192, 160), (512, 20, 18)])
def test_largest_grids_vs_oracle(gpu_lib, dims):
    import oracle_lib as O
    from monomers2026_amd import selm
    x, v, m, ij, K, r0, u0 = _case(dims, 3)
    lo, hi = (0.0, 0.0, 0.0), tuple(float(d) for d in dims)
    o = O.Oracle(*dims, lo, hi, noise=True, **KW)
    o.set_beads(x, v, m)
    o.set_bonds(ij, K, r0)
    o.init()
    o.set_field(0, u0)
    o.step(1)
    ref1 = (o.beads(), o.field(0))
    o.step(1)
    ref2 = (o.beads(), o.field(0))
    o.close()
    print("largest grid %s: oracle done" % (dims,), flush=True)
    for prec in (selm.PREC_FP64, selm.PREC_FP32_BULK):
        p = selm.make_params(dims, lo, hi, stochastic=True, precision=prec, **KW)
        s = selm.System(p)
        s.set_beads(x, v, m)
        s.set_bonds(ij, K, r0)
        s.init()
        s.set_field(u0)
        for step, (rb, ru) in ((1, ref1), (2, ref2)):
            s.step(1)
            if step == 1:  # the batch (and its FFT path) is built at the first step
                print("largest grid %s precision %d: fft passes %d" % (dims, prec, s.fft_passes()), flush=True)
            xg, vg, fcg, ftg = s.beads()
            ug = s.field()
            errs = dict(x=_rel(xg, rb[0], float(dims[0])), v=_rel(vg, rb[1]), fcons=_rel(fcg, rb[2]),
                        ftot=_rel(ftg, rb[3]), u=_rel(ug, ru))
            print("largest grid %s precision %d step %d: %s" % (dims, prec, step,
                  " ".join("%s=%.2e" % kv for kv in errs.items())), flush=True)
            if prec == selm.PREC_FP64:
                for k, e in errs.items():
                    assert e <= FP64_TOL, (dims, step, k, e)
            elif step == 1:
                assert errs["u"] < 2e-6 and errs["v"] < 2e-6, (dims, errs)
        assert s.status() == 0
        s.close()
