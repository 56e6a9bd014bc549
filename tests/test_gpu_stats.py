"""GPU step against closed forms of the discrete scheme (noise on unless
stated): bead equipartition in a frozen fluid, the stationary fluid
spectrum, the steady mobility of a forced bead (noise off), the long-time
self-diffusivity of free beads, and the FP32-bulk displacement error over
5000 steps.

Predictions (dx = 1, numpy in physics.py):
  equipartition   <m v^2>/kT = 1 / (1 - b/2), b = Upsilon dt / m
  fluid spectrum  E|U_k,d|^2 / N = 2 kT h_k pi_d(k) / (rho (2 - a_k))
  mobility        v = F (1/Upsilon + K_dd), K_dd = (1/(mu N)) sum |S_d|^2 pi_d / |lambda|
  diffusivity     D_d = kT (1/Upsilon + K_dd[h]): the zero-frequency balance
                  of the discrete step (the bead kick is spread back into the
                  fluid with opposite sign, so only the fluid's own forcing
                  moves the bead through the fluid), with the reference's
                  half-strength x-DC / x-Nyquist forcing planes h_k = 1/2
"""
import numpy as np
import pytest

import physics as P

pytestmark = pytest.mark.gpu


def _system(dims, x, m, rho=1.0, mu=1.0, dt=0.01, noise=True, seed=7, prec=0, ext=None):
    from monomers2026_amd import selm
    L = float(dims[0])
    p = selm.make_params(dims, (0.0, 0.0, 0.0), (L, L, L), rho=rho, mu=mu, dt=dt, seed=seed, stochastic=noise,
                         precision=prec)
    s = selm.System(p)
    s.set_beads(x, None, m)
    if ext is not None:
        s.set_external_force(*ext)
    s.init()
    return s


def test_equipartition_frozen_fluid(gpu_lib):
    dims, M, mu, dt = (16, 16, 16), 4000, 1.0, 0.01
    rng = np.random.default_rng(3)
    s = _system(dims, rng.uniform(0.0, 16.0, (M, 3)), np.ones(M), rho=1e8, mu=mu, dt=dt, seed=31)
    s.step(50)
    acc, n = np.zeros(3), 0
    for _ in range(400):
        s.step(5)
        v = s.beads()[1]
        acc += (v ** 2).mean(axis=0)
        n += 1
    s.close()
    b = P.upsilon(mu) * dt
    pred = 1.0 / (1.0 - b / 2.0)
    meas = acc / n
    print("gpu equipartition <m v^2>/kT", meas, "predicted %.5f" % pred, flush=True)
    # 400 x 4000 x 3 samples 5 steps apart (correlation (1-b)^5 = 0.09): std ~ 0.07%
    assert abs(meas.mean() / pred - 1) < 4e-3
    assert np.all(np.abs(meas / pred - 1) < 6e-3)


@pytest.mark.parametrize("prec", [0, 1])
def test_fluid_spectrum(gpu_lib, prec):
    dims, mu, rho, dt = (16, 16, 16), 10.0, 1.0, 0.01
    s = _system(dims, np.zeros((0, 3)), np.zeros(0), rho=rho, mu=mu, dt=dt, seed=41, prec=prec)
    s.step(500)
    N = 16 ** 3
    acc = np.zeros((3, 16, 16, 16))
    ns = 1500
    for _ in range(ns):
        s.step(3)
        acc += np.abs(np.fft.fftn(s.field(), axes=(1, 2, 3))) ** 2 / N
    s.close()
    meas = acc / ns
    pred = np.array(P.fluid_mode_variance(dims, mu, rho, 1.0, dt))
    h = np.broadcast_to(P.half_strength_planes(dims), pred.shape)
    ok = pred > 0
    r_full = meas[ok & (h == 1.0)].sum() / pred[ok & (h == 1.0)].sum()
    r_half = meas[ok & (h == 0.5)].sum() / pred[ok & (h == 0.5)].sum()
    # per shell of |lambda|: the low-k modes (slow, few) and the high-k ones
    _, lam = P.symbols(dims)
    lo_k = ok & (np.broadcast_to(-lam, pred.shape) < 1.0)
    r_lo = meas[lo_k].sum() / pred[lo_k].sum()
    print("gpu fluid spectrum prec=%d: full %.4f half-planes %.4f low-k %.4f k=0 %.1e"
          % (prec, r_full, r_half, r_lo, meas[:, 0, 0, 0].max()), flush=True)
    assert abs(r_full - 1) < 0.01 and abs(r_half - 1) < 0.015 and abs(r_lo - 1) < 0.1
    assert meas[:, 0, 0, 0].max() < (1e-20 if prec == 0 else 1e-8)


@pytest.mark.parametrize("prec", [0, 1])
def test_single_bead_mobility(gpu_lib, prec):
    dims, mu, m, F = (16, 16, 16), 5.0, 10.0, 0.1
    s = _system(dims, np.array([[5.3, 7.9, 9.4]]), np.array([m]), mu=mu, noise=False, prec=prec,
                ext=(np.array([0], np.int32), np.array([[F, 0.0, 0.0]])))
    s.step(3000)
    x, v, _, _ = s.beads()
    s.close()
    K = P.fluid_mobility(x[0], dims, mu)
    pred = F * (1.0 / P.upsilon(mu) + K[0])
    print("gpu mobility prec=%d: v_x %.6e predicted %.6e (rel %.2e)" % (prec, v[0, 0], pred, v[0, 0] / pred - 1),
          flush=True)
    assert abs(v[0, 0] / pred - 1) < (2e-4 if prec == 0 else 1e-3)
    assert abs(v[0, 1]) < 1e-3 * pred and abs(v[0, 2]) < 1e-3 * pred


def test_self_diffusivity(gpu_lib):
    """1000 free beads in a 16^3 box, mu = 5, m = 10 (b = 0.19): slope of the
    mean-square displacement between lags of 1000 and 4000 steps, in 8
    blocks of 10000 steps (error bar = std of the block estimates / sqrt 8)."""
    dims, M, mu, m, dt = (16, 16, 16), 1000, 5.0, 10.0, 0.01
    L = 16.0
    rng = np.random.default_rng(8)
    s = _system(dims, rng.uniform(0.0, L, (M, 3)), np.full(M, m), mu=mu, dt=dt, seed=77)
    s.step(1000)
    every, nblk, per_blk = 20, 8, 500
    xs = np.empty((nblk * per_blk, M, 3))
    for i in range(nblk * per_blk):
        s.step(every)
        xs[i] = s.beads()[0]
    s.close()
    xs = P.unwrap(xs, L)
    l1, l2 = 50, 200  # lags in samples: 1000 and 4000 steps
    Ds = []
    for b in range(nblk):
        blk = xs[b * per_blk:(b + 1) * per_blk]
        m1 = ((blk[l1:] - blk[:-l1]) ** 2).mean(axis=(0, 1))
        m2 = ((blk[l2:] - blk[:-l2]) ** 2).mean(axis=(0, 1))
        Ds.append((m2 - m1) / (2.0 * (l2 - l1) * every * dt))
    Ds = np.array(Ds)
    D, err = Ds.mean(axis=0), Ds.std(axis=0, ddof=1) / np.sqrt(nblk)
    h = P.half_strength_planes(dims)
    Xs = rng.uniform(0.0, L, (16, 3))
    Kh = np.mean([P.fluid_mobility(X, dims, mu, weights=h) for X in Xs], axis=0)
    Kf = np.mean([P.fluid_mobility(X, dims, mu) for X in Xs], axis=0)
    pred = 1.0 / P.upsilon(mu) + Kh
    fdt = 1.0 / P.upsilon(mu) + Kf
    print("gpu self-diffusivity D", D, "+-", err, "predicted", pred, "(kT x mobility:", fdt, ")", flush=True)
    assert np.all(np.abs(D - pred) < np.maximum(4 * err, 0.02 * pred))
    # the measurement tells the reference's half-strength planes from a
    # forcing that would give kT x mobility along x (12% apart at 16^3)
    assert abs(D[0] - fdt[0]) > 3 * err[0]


def test_fp32_bulk_displacement_long(gpu_lib):
    """64^3 grid, 10 chains x 100 beads, noise on: FP32-bulk bead
    displacements against the FP64 path (itself equal to the FP64 oracle to
    1e-12) at steps 100, 1000 and 5000."""
    from monomers2026_amd import selm, workloads
    w = workloads.chains1000_64()
    s64 = w.build_system(selm.PREC_FP64, True, seed=9)
    s32 = w.build_system(selm.PREC_FP32_BULK, True, seed=9)
    x0 = w.x.copy()
    L = w.edge
    out = {}
    done = 0
    for n in (100, 1000, 5000):
        s64.step(n - done)
        s32.step(n - done)
        done = n
        d64 = P.unwrap(np.stack([x0, s64.beads()[0]]), L)[1] - x0
        d32 = P.unwrap(np.stack([x0, s32.beads()[0]]), L)[1] - x0
        out[n] = np.abs(d32 - d64).max() / np.abs(d64).max()
        print("fp32-bulk chains1000_64 step %d: displacement rel err %.2e (max |d| %.2f)"
              % (n, out[n], np.abs(d64).max()), flush=True)
    s64.close()
    s32.close()
    # Measured (max-norm): 2.1e-7, 7.3e-7, 7.5e-6.  With FP32 butterflies
    # (FP32 arithmetic inside the transforms) they were 1.5e-6, 1.4e-5 and
    # 2.5e-4: the trajectory difference grows with the per-step rounding of
    # the fluid update, which the FP64 arithmetic of the passes cuts from
    # 2.2e-7 to 9e-8 of |u| per step.
    assert out[100] < 1e-6
    assert out[1000] < 1e-5
    assert out[5000] < 5e-5
