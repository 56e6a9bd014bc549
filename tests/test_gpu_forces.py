"""Device harmonic angles and lj/cut pairs (selm_set_angles,
selm_set_pair_lj) against the host force terms of the LAMMPS-subset deck,
whose forces are pinned as minus the gradient of its energy
(test_host.py); and a few noise-on steps against the FP64 oracle driven with
those host forces."""
import os

import numpy as np
import pytest

import host_lib as H
import oracle_lib as O

pytestmark = pytest.mark.gpu


def _deck_system(d, prec=0, noise=True, seed=4242):
    from monomers2026_amd import selm
    p = selm.make_params((16, 16, 16), tuple(d.lo), tuple(d.hi), rho=1.0, mu=1.0, kB=1.0, T=0.5, dt=d.dt, seed=seed,
                         stochastic=noise, precision=prec)
    s = selm.System(p)
    s.set_beads(d.x, d.v, d.m)
    s.set_bonds(d.ij, d.K, d.r0)
    return s


def _angles_and_pairs(s, d):
    """The deck's angles and lj/cut tables, handed over as the driver does."""
    s.set_angles(d.ijk, d.aK, d.at0)
    eps, sig, rc, sp = d.lj
    s.set_pair_lj(d.type - 1, eps, sig, rc, sp)


def test_device_forces_match_host_terms(gpu_lib):
    d = H.Deck(os.path.join(H.DECK_SMALL, "Model.LAMMPS_script"))
    s = _deck_system(d, noise=False)
    _angles_and_pairs(s, d)
    idx = np.nonzero(np.abs(d.fext).sum(axis=1))[0]
    s.set_external_force(idx, d.fext[idx])
    s.init()
    x, _, fc, _ = s.beads()
    want, _ = d.forces(x)
    err = np.abs(fc - want).max() / np.abs(want).max()
    print("device angle+lj forces vs host terms: %.2e" % err, flush=True)
    assert err < 1e-13
    # and after moving: a step, then the forces at the new positions
    s.step(3)
    x, _, fc, _ = s.beads()
    want, _ = d.forces(x)
    assert np.abs(fc - want).max() / np.abs(want).max() < 1e-13
    s.close()


def test_device_forces_steps_match_oracle(gpu_lib):
    d = H.Deck(os.path.join(H.DECK_SMALL, "Model.LAMMPS_script"))
    s = _deck_system(d, noise=True)
    _angles_and_pairs(s, d)
    idx = np.nonzero(np.abs(d.fext).sum(axis=1))[0]
    s.set_external_force(idx, d.fext[idx])
    s.init()
    o = O.Oracle(16, 16, 16, tuple(d.lo), tuple(d.hi), rho=1.0, mu=1.0, kB=1.0, T=0.5, dt=d.dt, seed=4242, noise=True)
    o.set_beads(d.x, d.v, d.m)
    o.set_bonds(d.ij, d.K, d.r0)
    o.set_external_force(idx, d.fext[idx])
    o.init()
    o.set_conservative_force(d.forces(o.beads()[0])[0])
    for step in range(20):
        s.step(1)
        o.step(1)
        o.set_conservative_force(d.forces(o.beads()[0])[0])
    xg, vg, _, _ = s.beads()
    xo, vo, _, _ = o.beads()
    ex = np.abs(xg - xo).max() / 8.0
    ev = np.abs(vg - vo).max() / max(np.abs(vo).max(), 1.0)
    print("device forces vs oracle+host forces, 20 steps: x %.2e v %.2e" % (ex, ev), flush=True)
    assert ex < 1e-11 and ev < 1e-11
    s.close()
    o.close()


def _lj_gas(bins, monkeypatch, seed=21):
    """8000 beads on a jittered lattice in a 24^3 box, chains of 20 with
    harmonic bonds, two lj/cut types with rc = 2.5 sigma (9 bins per axis)."""
    from monomers2026_amd import selm
    if bins:
        monkeypatch.delenv("SELM_LJ_BINS", raising=False)
    else:
        monkeypatch.setenv("SELM_LJ_BINS", "0")
    rng = np.random.default_rng(seed)
    L, n = 24.0, 20
    g = (np.arange(n) + 0.5) * (L / n)
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    x = np.mod(x + rng.uniform(-0.1, 0.1, x.shape), L)
    M = x.shape[0]
    ij = np.array([(i, i + 1) for i in range(M - 1) if (i + 1) % 20], np.int32)
    p = selm.make_params((24, 24, 24), (0.0, 0.0, 0.0), (L, L, L), rho=1.0, mu=1.0, kB=1.0, T=0.5, dt=0.002,
                         seed=seed, stochastic=True, precision=selm.PREC_FP64)
    s = selm.System(p)
    s.set_beads(x, np.zeros_like(x), np.ones(M))
    s.set_bonds(ij, np.full(len(ij), 20.0), np.full(len(ij), L / n))
    typ = (np.arange(M) % 3 == 0).astype(np.int32)
    eps = np.array([[1.0, 0.8], [0.8, 0.6]])
    sig = np.array([[1.0, 0.9], [0.9, 0.8]])
    rc = 2.5 * sig
    s.set_pair_lj(typ, eps, sig, rc, (0.0, 0.0, 0.5))
    s.init()
    return s


def test_lj_bins_match_all_pairs(gpu_lib, monkeypatch):
    """lj/cut with neighbour bins against the all-pairs loop (SELM_LJ_BINS=0):
    the same pairs in another summation order, so forces and a few noise-on
    steps agree to rounding; two binned runs agree bit for bit."""
    out = {}
    for name, bins in (("bins", True), ("pairs", False), ("bins2", True)):
        s = _lj_gas(bins, monkeypatch)
        fc0 = s.beads()[2]
        s.step(10)
        out[name] = (fc0, s.beads())
        s.close()
    monkeypatch.delenv("SELM_LJ_BINS", raising=False)
    f_b, f_p = out["bins"][0], out["pairs"][0]
    ef = np.abs(f_b - f_p).max() / np.abs(f_p).max()
    xb, xp = out["bins"][1][0], out["pairs"][1][0]
    ex = np.abs(xb - xp).max() / 24.0
    ev = np.abs(out["bins"][1][1] - out["pairs"][1][1]).max() / np.abs(out["pairs"][1][1]).max()
    print("lj bins vs all pairs: forces %.2e, after 10 steps x %.2e v %.2e" % (ef, ex, ev), flush=True)
    assert ef < 1e-12 and ex < 1e-12 and ev < 1e-11
    for a, b in zip(out["bins"][1], out["bins2"][1]):
        assert np.array_equal(a, b)
