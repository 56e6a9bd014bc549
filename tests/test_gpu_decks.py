"""The reference's own decks on the GPU (fixtures of selm/test/NucleotideMonomer
and cases/SELM/test, rebuilt as run directories by tests/deck_fixture.py):

- through the C ABI, FP64, noise on, with the device's bond, angle and
  lj/cut forces, 5 steps against the oracle driven by the host forces at
  every step, 1e-12 (normalised as in test_gpu_parity);
- through the deck driver (selm_hip_driver), 5 steps, its particle*.dat
  against the oracle to the digits "%g" writes.
Real units (mu 752.8, rho 493.9, kB T = 2.48), 30^3 and 24^3 meshes over a
6 A box."""
import os
import subprocess

import numpy as np
import pytest

import deck_fixture as DF
import host_lib as H

pytestmark = pytest.mark.gpu

NSTEPS = 5


def _oracle_run(d, m, mesh, nsteps):
    import oracle_lib as O
    o = O.Oracle(*mesh, tuple(d.lo), tuple(d.hi), rho=m.rho, mu=m.mu, kB=m.KB, T=m.T, dt=d.dt,
                 seed=m.SELM_Seed, noise=bool(m.flagStochasticDriving))
    o.set_beads(d.x, d.v, d.m)
    o.set_bonds(d.ij, d.K, d.r0)
    o.init()
    o.set_conservative_force(d.forces(o.beads()[0])[0])
    out = []
    for _ in range(nsteps):
        o.step(1)
        x, v, fc, ft = o.beads()
        o.set_conservative_force(d.forces(x)[0])
        out.append((x, v, o.beads()[2], ft, o.field(0)))
    o.close()
    return out


def _rel(a, b, floor=1.0):
    return np.abs(a - b).max() / max(np.abs(b).max(), floor)


@pytest.mark.parametrize("name", ["deck_nucleotide.json", "deck_1base.json"])
def test_reference_deck_abi_vs_oracle(gpu_lib, tmp_path, name):
    from monomers2026_amd import selm
    fx = DF.load(name)
    run = DF.write_run(fx, str(tmp_path / "run"), nsteps=NSTEPS)
    d = H.Deck(run + "/Model.LAMMPS_script")
    m = H.model(run + "/Model.SELM")
    mesh = tuple(m.numMeshPtsPerDir)
    p = selm.make_params(mesh, tuple(d.lo), tuple(d.hi), rho=m.rho, mu=m.mu, kB=m.KB, T=m.T, dt=d.dt,
                         seed=m.SELM_Seed, stochastic=bool(m.flagStochasticDriving), precision=selm.PREC_FP64)
    s = selm.System(p)
    s.set_beads(d.x, d.v, d.m)
    s.set_bonds(d.ij, d.K, d.r0)
    s.set_angles(d.ijk, d.aK, d.at0)
    eps, sig, rc, sp = d.lj
    s.set_pair_lj(d.type - 1, eps, sig, rc, sp)
    s.init()
    ref = _oracle_run(d, m, mesh, NSTEPS)
    L = float(d.hi[0] - d.lo[0])
    sv = np.sqrt(m.KB * m.T / d.m.min())  # thermal velocity scale
    for k in range(NSTEPS):
        s.step(1)
        x, v, fc, ft = s.beads()
        xo, vo, fco, fto, uo = ref[k]
        fs = max(np.abs(fto).max(), 1.0)
        errs = dict(x=_rel(x, xo, L), v=_rel(v, vo, sv), fcons=_rel(fc, fco, fs), ftot=_rel(ft, fto, fs),
                    u=_rel(s.field(), uo, sv))
        print("%s step %d: %s" % (name, k + 1, " ".join("%s=%.2e" % e for e in errs.items())), flush=True)
        for key, e in errs.items():
            assert e <= 1e-12, (name, k + 1, key, e)
    assert s.status() == 0
    s.close()
    d.close()


def test_nucleotide_deck_through_the_driver(gpu_lib, tmp_path):
    fx = DF.load("deck_nucleotide.json")
    run = DF.write_run(fx, str(tmp_path / "run"), nsteps=NSTEPS, save_every=1, dcd_every=1)
    assert os.path.exists(H.DRIVER), "selm_hip_driver is not built"
    p = subprocess.run([H.DRIVER, run], capture_output=True, text=True, timeout=120)
    print(p.stdout[-300:], p.stderr[-300:], flush=True)
    assert p.returncode == 0, p.stderr
    d = H.Deck(run + "/Model.LAMMPS_script")
    m = H.model(run + "/Model.SELM")
    ref = _oracle_run(d, m, tuple(m.numMeshPtsPerDir), NSTEPS)
    X = np.loadtxt(os.path.join(run, "particleX.dat")).reshape(-1, d.n, 3)
    V = np.loadtxt(os.path.join(run, "particleV.dat")).reshape(-1, d.n, 3)
    assert X.shape[0] == NSTEPS
    for k in range(NSTEPS):
        ex = np.abs(X[k] - ref[k][0]).max() / np.abs(ref[k][0]).max()
        ev = np.abs(V[k] - ref[k][1]).max() / np.abs(ref[k][1]).max()
        print("nucleotide deck driver step %d: x %.2e v %.2e" % (k + 1, ex, ev), flush=True)
        assert ex < 2e-5 and ev < 2e-5
    d.close()
