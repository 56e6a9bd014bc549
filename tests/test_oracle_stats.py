"""The FP64 oracle against closed forms of the discrete scheme: the
steady-state mobility of a forced bead (noise off), bead equipartition in a
frozen fluid, and the stationary spectrum of the thermally driven fluid.

These pin the oracle's whole step (a bead update, spread, viscous predictor,
projection, interpolation and thermal forcing working together) against
values the mathematics fixes, independently of its own code.  The GPU path
meets the same closed forms in test_gpu_stats.py."""
import numpy as np
import pytest

import oracle_lib as O
import physics as P


def _oracle(dims, M, x, m, rho=1.0, mu=1.0, dt=0.01, noise=False, seed=5, ext=None):
    L = float(dims[0])
    o = O.Oracle(*dims, (0.0, 0.0, 0.0), (L, L, L), rho=rho, mu=mu, kB=1.0, T=1.0, dt=dt, seed=seed, noise=noise)
    o.set_beads(x, np.zeros((M, 3)), m)
    if ext is not None:
        o.set_external_force(*ext)
    o.init()
    return o


def test_single_bead_mobility_oracle():
    """Constant force F on one bead, noise off: the bead's steady velocity is
    F (1/Upsilon + K_xx) with K the periodic Stokes sum of the discrete
    scheme (cuda/SELM_Lagrangian_LAMMPS.cu:138 for Upsilon)."""
    dims, mu, m, F = (16, 16, 16), 5.0, 10.0, 0.1
    X = np.array([[5.3, 7.9, 9.4]])
    o = _oracle(dims, 1, X, np.array([m]), mu=mu, ext=(np.array([0], np.int32), np.array([[F, 0.0, 0.0]])))
    o.step(3000)
    x, v, _, _ = o.beads()
    o.close()
    K = P.fluid_mobility(x[0], dims, mu)
    pred = F * (1.0 / P.upsilon(mu) + K[0])
    print("oracle mobility: v_x %.6e predicted %.6e (rel %.2e); v_y %.1e v_z %.1e"
          % (v[0, 0], pred, v[0, 0] / pred - 1, v[0, 1], v[0, 2]))
    # K is evaluated at the final position; the bead drifts 0.04 cells during
    # the run, slowly enough for the fluid to follow (measured: 2e-6)
    assert abs(v[0, 0] / pred - 1) < 2e-4
    assert abs(v[0, 1]) < 2e-4 * pred and abs(v[0, 2]) < 2e-4 * pred


def test_equipartition_frozen_fluid_oracle():
    """rho = 1e8 freezes the fluid; each bead velocity then follows
    v <- (1 - b) v + coef xi / m with b = Upsilon dt / m, whose stationary
    value is <m v^2>/kT = 1 / (1 - b/2)."""
    dims, M, mu, dt = (8, 8, 8), 300, 1.0, 0.01
    rng = np.random.default_rng(1)
    x = rng.uniform(0.0, 8.0, (M, 3))
    o = _oracle(dims, M, x, np.ones(M), rho=1e8, mu=mu, dt=dt, noise=True, seed=17)
    o.step(50)
    acc, n = np.zeros(3), 0
    for _ in range(600):
        o.step(4)
        v = o.beads()[1]
        acc += (v ** 2).mean(axis=0)
        n += 1
    o.close()
    b = P.upsilon(mu) * dt
    pred = 1.0 / (1.0 - b / 2.0)
    meas = acc / n
    print("oracle equipartition: <m v^2>/kT", meas, "predicted %.5f" % pred)
    # 600 x 300 x 3 nearly independent samples: std of the mean ~ 0.3%
    assert abs(meas.mean() / pred - 1) < 0.012
    assert np.all(np.abs(meas / pred - 1) < 0.02)


def test_fluid_spectrum_oracle():
    """No beads, thermal forcing on: the stationary E|U_k,d|^2 / N matches
    2 kT h_k pi_d(k) / (rho (2 - a_k)), including the half-strength x-DC and
    x-Nyquist planes of the reference's forcing."""
    dims, mu, rho, dt = (8, 8, 8), 10.0, 1.0, 0.01
    o = _oracle(dims, 0, np.zeros((0, 3)), np.zeros(0), rho=rho, mu=mu, dt=dt, noise=True, seed=23)
    o.step(300)
    N = 512
    acc = np.zeros((3,) + dims[::-1])
    ns = 1500
    for _ in range(ns):
        o.step(2)
        U = np.fft.fftn(o.field(0), axes=(1, 2, 3))
        acc += np.abs(U) ** 2 / N
    o.close()
    meas = acc / ns
    pred = np.array(P.fluid_mode_variance(dims, mu, rho, 1.0, dt))
    h = np.broadcast_to(P.half_strength_planes(dims), pred.shape)
    ok = pred > 0
    full = ok & (h == 1.0)
    half = ok & (h == 0.5)
    r_full = meas[full].sum() / pred[full].sum()
    r_half = meas[half].sum() / pred[half].sum()
    print("oracle fluid spectrum: measured/predicted  full modes %.4f  half-strength planes %.4f" % (r_full, r_half))
    assert abs(r_full - 1) < 0.02 and abs(r_half - 1) < 0.03
    assert meas[:, 0, 0, 0].max() < 1e-20  # k = 0 removed
