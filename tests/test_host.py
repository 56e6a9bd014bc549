"""Host support library (CPU): the Model.SELM reader's semantics, the
LAMMPS-subset deck reader, the host conservative force terms, and the DCD
writer against the layout of the reference's shipped trajectory."""
import json
import os
import struct

import numpy as np
import pytest

import host_lib as H

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


# --- Model.SELM (cuda/SELM_ParmParse.cuh:19-101, cuda/SELM.cu:17-48) --------
def test_model_selm_semantics(tmp_path):
    p = _write(tmp_path, "Model.SELM", "\n".join([
        "#Required Parameters",
        "mu  752.768",
        "rho 493.936   # trailing comment",
        "   KB  0.008313",
        "T  298.15",
        "EulerianType Eulerian_Period       ",
        "LagrangianType Lagrangian_LAMMPS   ",
        "nsteps 20",
        "nsteps 30            # a later line overrides",
        "",
        "#Optional Parameters",
        "flagStochasticDriving 1",
        "numMeshPtsPerDir 24 24 24 99",
    ]))
    m = H.model(p)
    assert (m.mu, m.rho, m.KB, m.T) == (752.768, 493.936, 0.008313, 298.15)
    assert m.nsteps == 30
    assert m.flagStochasticDriving == 1
    assert m.SELM_Seed == 123456  # default (cuda/SELM.cuh:13)
    assert list(m.numMeshPtsPerDir) == [24, 24, 24]
    assert m.LagrangianType == b"Lagrangian_LAMMPS" and m.EulerianType == b"Eulerian_Period"
    assert m.saveSkipSimulationData == 0 and m.writeParticalX == 0


def test_model_selm_errors(tmp_path):
    base = ["rho 1", "T 1", "nsteps 3", "KB 1", "EulerianType Eulerian_Period",
            "LagrangianType Lagrangian_LAMMPS", "numMeshPtsPerDir 8 8 8"]
    with pytest.raises(H.HostError, match="Key not found: mu"):
        H.model(_write(tmp_path, "a.SELM", "\n".join(base)))
    short = [l if not l.startswith("numMesh") else "numMeshPtsPerDir 8 8" for l in base] + ["mu 1"]
    with pytest.raises(H.HostError, match="Not enough data"):
        H.model(_write(tmp_path, "b.SELM", "\n".join(short)))
    with pytest.raises(H.HostError, match="Unable to open"):
        H.model(str(tmp_path / "missing.SELM"))


def test_model_selm_of_the_small_deck():
    m = H.model(os.path.join(H.DECK_SMALL, "Model.SELM"))
    assert m.nsteps == 60 and m.SELM_Seed == 4242 and list(m.numMeshPtsPerDir) == [16, 16, 16]
    assert (m.saveSkipSimulationData, m.writeParticalX, m.writeParticalV, m.writeParticalF) == (20, 1, 1, 1)


# --- LAMMPS-subset deck ------------------------------------------------------
def test_deck_small_contents():
    d = H.Deck(os.path.join(H.DECK_SMALL, "Model.LAMMPS_script"))
    assert (d.n, d.nb, d.na, d.nt) == (10, 7, 4, 3)
    assert list(d.id) == list(range(1, 11))  # ascending atom ID
    np.testing.assert_array_equal(d.lo, [-4, -4, -4])
    np.testing.assert_array_equal(d.hi, [4, 4, 4])
    assert d.dt == 0.005
    np.testing.assert_array_equal(d.x[0], [-1.0, 0.5, 0.3])  # atom 1 (listed second in the file)
    np.testing.assert_array_equal(d.x[4], [3.6, 3.7, -3.8])  # atom 5
    np.testing.assert_array_equal(d.v[1], [0.0, -0.2, 0.0])
    np.testing.assert_array_equal(d.m, [1.0, 2.0, 1.5, 2.0, 1.0, 2.0, 1.5, 1.0, 1.5, 1.0])
    np.testing.assert_array_equal(d.ij[0], [0, 1])
    np.testing.assert_array_equal(d.ij[5], [7, 4])
    np.testing.assert_array_equal(d.K, [50, 35, 50, 50, 35, 50, 35])
    np.testing.assert_array_equal(d.r0, [0.8, 0.65, 0.8, 0.8, 0.65, 0.8, 0.65])
    # fix addforce on group type 3 (atoms 3, 7, 9)
    ext = np.zeros((10, 3))
    ext[[2, 6, 8]] = [0.05, 0.0, -0.02]
    np.testing.assert_array_equal(d.fext, ext)


@pytest.mark.parametrize("bad, msg", [
    ("fix 2 all langevin 1 1 1 1", "unsupported fix"),
    ("pair_style lj/cut/coul/long 1", "unsupported pair_style"),
    ("boundary p p f", "periodic"),
    ("run 100", "unsupported command"),
    ("bond_coeff 3 1 1", "type out of range"),
])
def test_deck_rejects_what_it_cannot_run(tmp_path, bad, msg):
    src = open(os.path.join(H.DECK_SMALL, "Model.LAMMPS_script")).read()
    data = open(os.path.join(H.DECK_SMALL, "Model.LAMMPS_read_data")).read()
    _write(tmp_path, "Model.LAMMPS_read_data", data)
    p = _write(tmp_path, "Model.LAMMPS_script", src + "\n" + bad + "\n")
    with pytest.raises(H.HostError, match=msg):
        H.Deck(p)


def _energy(d, x):
    return d.forces(x)[1]


def test_forces_are_minus_energy_gradient():
    """Central differences of the deck energy (bonds, angles, lj/cut with the
    special-bond weights) against the returned forces minus fix addforce."""
    d = H.Deck(os.path.join(H.DECK_SMALL, "Model.LAMMPS_script"))
    rng = np.random.default_rng(0)
    x = d.x + rng.uniform(-0.05, 0.05, d.x.shape)
    f, e = d.forces(x)
    fc = f - d.fext
    h = 1e-6
    g = np.zeros_like(x)
    for i in range(d.n):
        for a in range(3):
            xp = x.copy(); xp[i, a] += h
            xm = x.copy(); xm[i, a] -= h
            g[i, a] = (_energy(d, xp) - _energy(d, xm)) / (2 * h)
    np.testing.assert_allclose(fc, -g, rtol=1e-6, atol=1e-6 * np.abs(fc).max())
    # internal forces sum to zero (translation invariance, across the seams too)
    assert np.abs(fc.sum(axis=0)).max() < 1e-10 * np.abs(fc).max()


def test_special_bond_weights():
    """The 1-4 pair (atoms 1 and 10, path 1-2-3-10) carries half its LJ energy
    under special_bonds lj 0 0 0.5 and none under lj 0 0 0."""
    import tempfile
    d = H.Deck(os.path.join(H.DECK_SMALL, "Model.LAMMPS_script"))
    x = d.x.copy()
    r = np.linalg.norm(x[0] - x[9])
    assert r < 0.9  # inside the type 1-1 cutoff
    e_pair = 4 * 0.20 * ((0.60 / r) ** 12 - (0.60 / r) ** 6)
    src = open(os.path.join(H.DECK_SMALL, "Model.LAMMPS_script")).read()
    with tempfile.TemporaryDirectory() as t:
        open(os.path.join(t, "Model.LAMMPS_read_data"), "w").write(
            open(os.path.join(H.DECK_SMALL, "Model.LAMMPS_read_data")).read())
        p = os.path.join(t, "Model.LAMMPS_script")
        open(p, "w").write(src.replace("special_bonds   lj 0.0 0.0 0.5", "special_bonds   lj 0.0 0.0 0.0"))
        e0 = H.Deck(p).forces(x)[1]
    assert abs((d.forces(x)[1] - e0) - 0.5 * e_pair) < 1e-12 * max(1.0, abs(e_pair))


# --- DCD --------------------------------------------------------------------
def test_dcd_layout_matches_the_reference_trajectory(tmp_path):
    """Same header words, record lengths and first-frame bytes as
    cases/Diffusion Coefficient/1base.dcd for the same inputs (4 beads, dump
    every 4000 steps, timestep 0.005, box 6, the deck's initial positions)."""
    g = json.load(open(os.path.join(GOLD, "dcd_1base_head.json")))
    frame0 = bytes.fromhex(g["frame0_hex"])
    # initial positions from the fixture's first frame
    n = g["natoms"]
    xyz = [np.frombuffer(frame0, "<f4", n, 56 + k * (8 + 4 * n) + 4) for k in range(3)]
    x0 = np.stack(xyz, axis=1).astype(np.float64)
    path = str(tmp_path / "t.dcd")
    nfr = 3
    H.write_dcd(path, [(k * 4000, x0 + 0.01 * k) for k in range(nfr)], 4000, 0.005, (6.0, 6.0, 6.0))
    b = open(path, "rb").read()
    assert len(b) == g["header_bytes"] + nfr * g["frame_bytes"]
    assert [struct.unpack_from("<i", b, o)[0] for o in (0, 92, g["header_bytes"] - 12)] == g["record_lengths"]
    icntrl = list(struct.unpack_from("<20i", b, 8))
    want = list(g["icntrl"])
    want[0] = nfr                  # frames written
    want[3] = (nfr - 1) * 4000     # last timestep written
    assert icntrl == want
    assert struct.unpack_from("<i", b, 96)[0] == g["ntitle"]
    assert struct.unpack_from("<i", b, g["header_bytes"] - 8)[0] == n
    assert b[g["header_bytes"]:g["header_bytes"] + g["frame_bytes"]] == frame0
    ic, titles, nat, frames = H.read_dcd(path)
    assert nat == n and len(frames) == nfr
    np.testing.assert_allclose(frames[2][1], (x0 + 0.02).astype(np.float32))
