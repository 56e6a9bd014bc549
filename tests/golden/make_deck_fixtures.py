"""Write deck_nucleotide.json and deck_1base.json: the two input decks the
reference ships, as data.

- selm/test/NucleotideMonomer (Model.SELM, Model.LAMMPS_script,
  Model.LAMMPS_read_data): 384 beads in 96 four-bead molecules, 30^3 mesh,
  real units, seed 123456.
- cases/SELM/test (the same three files): one four-bead molecule, 24^3
  mesh, seed 7324 (the run that wrote cases/Diffusion Coefficient/1base.dcd).

Each deck is read with this repository's own Model.SELM and LAMMPS-subset
readers (libselm_host.so) and written as numbers only: the Model.SELM keys,
box, time step, bead positions, velocities, masses and types, bonds and
angles with their coefficients (expanded per bond and angle), the lj/cut
table and the special-bond weights.  tests/deck_fixture.py rebuilds a run
directory from it.  Run here, where the reference is mounted:

    python tests/golden/make_deck_fixtures.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import host_lib as H  # noqa: E402

REF = "/root/reference/data"
DECKS = {
    "deck_nucleotide.json": REF + "/src/selm-main/test/NucleotideMonomer",
    "deck_1base.json": REF + "/cases/SELM/test",
}


def fixture(src):
    m = H.model(os.path.join(src, "Model.SELM"))
    d = H.Deck(os.path.join(src, "Model.LAMMPS_script"))
    eps, sig, rc, sp = d.lj
    out = {
        "source": os.path.relpath(src, REF),
        "model": {"mu": m.mu, "rho": m.rho, "KB": m.KB, "T": m.T, "nsteps": m.nsteps,
                  "flagStochasticDriving": m.flagStochasticDriving, "SELM_Seed": m.SELM_Seed,
                  "numMeshPtsPerDir": list(m.numMeshPtsPerDir),
                  "EulerianType": m.EulerianType.decode(), "LagrangianType": m.LagrangianType.decode()},
        "lo": d.lo.tolist(), "hi": d.hi.tolist(), "dt": d.dt,
        "x": d.x.tolist(), "v": d.v.tolist(), "m": d.m.tolist(), "type": d.type.tolist(), "id": d.id.tolist(),
        "bonds": d.ij.tolist(), "bond_K": d.K.tolist(), "bond_r0": d.r0.tolist(),
        "angles": d.ijk.tolist(), "angle_K": d.aK.tolist(), "angle_theta0": d.at0.tolist(),
        "lj_eps": eps.tolist(), "lj_sigma": sig.tolist(), "lj_rc": rc.tolist(), "special_lj": sp.tolist(),
        "fext": d.fext.tolist(),
    }
    d.close()
    return out


def main():
    for name, src in DECKS.items():
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(fixture(src), f, separators=(",", ":"))
        print("wrote", name)


if __name__ == "__main__":
    main()
