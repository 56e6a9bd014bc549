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
