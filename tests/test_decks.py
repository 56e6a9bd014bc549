"""The reference's two input decks (selm/test/NucleotideMonomer and
cases/SELM/test) as fixtures: a run directory rebuilt from each reads back
to the same atoms, topology and coefficients, and its host forces are minus
the gradient of its energy."""
import numpy as np
import pytest

import deck_fixture as DF
import host_lib as H


@pytest.mark.parametrize("name,n", [("deck_nucleotide.json", 384), ("deck_1base.json", 4)])
def test_rebuilt_deck_reads_back(tmp_path, name, n):
    fx = DF.load(name)
    run = DF.write_run(fx, str(tmp_path / "run"), nsteps=5)
    d = H.Deck(run + "/Model.LAMMPS_script")
    m = H.model(run + "/Model.SELM")
    assert d.n == n and d.dt == fx["dt"]
    np.testing.assert_array_equal(d.x, np.asarray(fx["x"]))
    np.testing.assert_array_equal(d.m, np.asarray(fx["m"]))
    np.testing.assert_array_equal(d.ij, np.asarray(fx["bonds"]).reshape(-1, 2))
    np.testing.assert_array_equal(d.K, np.asarray(fx["bond_K"]))
    np.testing.assert_array_equal(d.r0, np.asarray(fx["bond_r0"]))
    np.testing.assert_array_equal(d.ijk, np.asarray(fx["angles"]).reshape(-1, 3))
    np.testing.assert_array_equal(d.at0, np.asarray(fx["angle_theta0"]))
    eps, sig, rc, sp = d.lj
    np.testing.assert_array_equal(sig, np.asarray(fx["lj_sigma"]))
    np.testing.assert_array_equal(rc, np.asarray(fx["lj_rc"]))
    assert (m.mu, m.rho, m.KB, m.T) == (fx["model"]["mu"], fx["model"]["rho"], fx["model"]["KB"], fx["model"]["T"])
    assert m.SELM_Seed == fx["model"]["SELM_Seed"] and list(m.numMeshPtsPerDir) == fx["model"]["numMeshPtsPerDir"]
    # host forces = -dE/dx by central differences at the deck's positions
    f, e = d.forces(d.x)
    h = 1e-6
    for i in range(min(n, 6)):
        for a in range(3):
            xp, xm = d.x.copy(), d.x.copy()
            xp[i, a] += h
            xm[i, a] -= h
            g = (d.forces(xp)[1] - d.forces(xm)[1]) / (2 * h)
            assert abs(f[i, a] + g) <= 1e-5 * max(1.0, abs(g)), (i, a, f[i, a], g)
    d.close()
