"""The C++ deck driver (Selm over Eulerian_Period_HIP + Lagrangian_LAMMPS_HIP)
runs a Model.SELM + LAMMPS-subset deck end to end.  The deck has harmonic
bonds, harmonic angles, lj/cut with special-bond weights and fix addforce,
and its thermal forcing is on.  The driver's outputs are compared with the
FP64 oracle driven by the same deck, with the host force terms handed back
every step:
- particleX/V/F.dat, written with "%g", so agreement is to the 6 digits
  printed;
- the DCD frames, float32."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import host_lib as H

pytestmark = pytest.mark.gpu


def _blocks(path, n):
    a = np.loadtxt(path).reshape(-1, n, 3)
    return a


@pytest.mark.parametrize("host_forces", [0, 1])
def test_driver_runs_the_small_deck(gpu_lib, tmp_path, host_forces):
    """host_forces 0: angles and lj/cut on the device (default); 1: on the
    host, handed over every step through selm_set_conservative_force."""
    import oracle_lib as O
    run = tmp_path / "deck"
    shutil.copytree(H.DECK_SMALL, run)
    with open(run / "Model.SELM", "a") as f:
        f.write("SELM_HostForces %d\n" % host_forces)
    assert os.path.exists(H.DRIVER), "selm_hip_driver is not built"
    p = subprocess.run([H.DRIVER, str(run)], capture_output=True, text=True, timeout=120)
    print(p.stdout[-400:], p.stderr[-400:], flush=True)
    assert p.returncode == 0, p.stderr
    assert "Total time average" in p.stdout

    d = H.Deck(os.path.join(H.DECK_SMALL, "Model.LAMMPS_script"))
    m = H.model(os.path.join(H.DECK_SMALL, "Model.SELM"))
    o = O.Oracle(16, 16, 16, tuple(d.lo), tuple(d.hi), rho=m.rho, mu=m.mu, kB=m.KB, T=m.T, dt=d.dt,
                 seed=m.SELM_Seed, noise=True)
    o.set_beads(d.x, d.v, d.m)
    o.set_bonds(d.ij, d.K, d.r0)
    idx = np.nonzero(np.abs(d.fext).sum(axis=1))[0].astype(np.int32)
    o.set_external_force(idx, d.fext[idx])
    o.init()
    x0 = o.beads()[0]
    o.set_conservative_force(d.forces(x0)[0])
    ref = {}
    for step in range(1, m.nsteps + 1):
        o.step(1)
        x, v, _, ft = o.beads()
        if step % m.saveSkipSimulationData == 0:
            ref[step] = (x, v, ft)
        o.set_conservative_force(d.forces(x)[0])
    o.close()

    X = _blocks(run / "particleX.dat", d.n)
    V = _blocks(run / "particleV.dat", d.n)
    F = _blocks(run / "particleF.dat", d.n)
    steps = sorted(ref)
    assert X.shape[0] == V.shape[0] == F.shape[0] == len(steps)
    for k, s in enumerate(steps):
        x, v, ft = ref[s]
        for name, got, want in (("x", X[k], x), ("v", V[k], v), ("f", F[k], ft)):
            err = np.abs(got - want).max() / max(np.abs(want).max(), 1e-12)
            print("driver (host forces %d) vs oracle step %d %s: %.2e" % (host_forces, s, name, err), flush=True)
            assert err < 2e-5, (s, name, err)

    icntrl, titles, nat, frames = H.read_dcd(str(run / "Model.dcd"))
    assert nat == d.n and len(frames) == 1 + m.nsteps // 20
    assert icntrl[0] == len(frames) and icntrl[1] == 0 and icntrl[2] == 20 and icntrl[3] == m.nsteps
    np.testing.assert_allclose(frames[0][1], x0.astype(np.float32))
    np.testing.assert_allclose(frames[-1][1], ref[m.nsteps][0].astype(np.float32), rtol=1e-6, atol=1e-6)
    assert frames[0][0] == (8.0, 0.0, 8.0, 0.0, 0.0, 8.0)
