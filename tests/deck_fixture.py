"""Rebuild a run directory (Model.SELM, Model.LAMMPS_script,
Model.LAMMPS_read_data) from a deck fixture of tests/golden (written by
tests/golden/make_deck_fixtures.py from the reference's decks), in the
command subset this repository's deck reader takes.  Bond and angle types
are the distinct coefficient pairs, masses are per atom type."""
import json
import math
import os

import numpy as np

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _deg_for(rad):
    """A decimal degree value that the reader (deg * pi / 180) maps back to rad exactly."""
    d = rad * 180.0 / math.pi
    for _ in range(64):
        back = d * math.pi / 180.0
        if back == rad:
            return d
        d = math.nextafter(d, math.inf if back < rad else -math.inf)
    raise ValueError("no degree value maps to %r" % rad)


def _types(keys):
    uniq = []
    idx = []
    for k in keys:
        if k not in uniq:
            uniq.append(k)
        idx.append(uniq.index(k) + 1)
    return uniq, idx


def write_run(fx, path, nsteps, save_every=1, dcd_every=1, seed=None, stochastic=None):
    os.makedirs(path, exist_ok=True)
    m = fx["model"]
    with open(os.path.join(path, "Model.SELM"), "w") as f:
        f.write("mu %r\nrho %r\nKB %r\nT %r\n" % (m["mu"], m["rho"], m["KB"], m["T"]))
        f.write("EulerianType %s\nLagrangianType %s\nnsteps %d\n" % (m["EulerianType"], m["LagrangianType"], nsteps))
        f.write("flagStochasticDriving %d\n" % (m["flagStochasticDriving"] if stochastic is None else stochastic))
        f.write("SELM_Seed %d\n" % (m["SELM_Seed"] if seed is None else seed))
        f.write("numMeshPtsPerDir %d %d %d\n" % tuple(m["numMeshPtsPerDir"]))
        f.write("saveSkipSimulationData %d\nwriteParticalX 1\nwriteParticalV 1\nwriteParticalF 1\n" % save_every)
    x, v = np.asarray(fx["x"]), np.asarray(fx["v"])
    mass, typ = np.asarray(fx["m"]), np.asarray(fx["type"])
    n, nt = x.shape[0], int(typ.max())
    btypes, bidx = _types(list(zip(fx["bond_K"], fx["bond_r0"])))
    atypes, aidx = _types(list(zip(fx["angle_K"], fx["angle_theta0"])))
    with open(os.path.join(path, "Model.LAMMPS_read_data"), "w") as f:
        f.write("LAMMPS data file rebuilt from a fixture\n\n")
        f.write("%d atoms\n%d bonds\n%d angles\n\n" % (n, len(fx["bonds"]), len(fx["angles"])))
        f.write("%d atom types\n%d bond types\n%d angle types\n\n" % (nt, max(1, len(btypes)), max(1, len(atypes))))
        for a, nm in enumerate(("x", "y", "z")):
            f.write("%r %r %slo %shi\n" % (fx["lo"][a], fx["hi"][a], nm, nm))
        f.write("\nMasses\n\n")
        for t in range(1, nt + 1):
            f.write("%d %r\n" % (t, float(mass[typ == t][0])))
        f.write("\nAtoms\n\n")
        for i in range(n):
            f.write("%d 1 %d %r %r %r\n" % (fx["id"][i], typ[i], *map(float, x[i])))
        if np.any(v != 0):
            f.write("\nVelocities\n\n")
            for i in range(n):
                f.write("%d %r %r %r\n" % (fx["id"][i], *map(float, v[i])))
        if fx["bonds"]:
            f.write("\nBonds\n\n")
            for q, (a, b) in enumerate(fx["bonds"]):
                f.write("%d %d %d %d\n" % (q + 1, bidx[q], fx["id"][a], fx["id"][b]))
        if fx["angles"]:
            f.write("\nAngles\n\n")
            for q, (a, b, c) in enumerate(fx["angles"]):
                f.write("%d %d %d %d %d\n" % (q + 1, aidx[q], fx["id"][a], fx["id"][b], fx["id"][c]))
    eps, sig, rc = np.asarray(fx["lj_eps"]), np.asarray(fx["lj_sigma"]), np.asarray(fx["lj_rc"])
    with open(os.path.join(path, "Model.LAMMPS_script"), "w") as f:
        f.write("units real\nboundary p p p\natom_style angle\nread_data Model.LAMMPS_read_data\n")
        if btypes:
            f.write("bond_style harmonic\n")
            for t, (K, r0) in enumerate(btypes):
                f.write("bond_coeff %d %r %r\n" % (t + 1, K, r0))
        if atypes:
            f.write("angle_style harmonic\n")
            for t, (K, th) in enumerate(atypes):
                f.write("angle_coeff %d %r %r\n" % (t + 1, K, _deg_for(th)))
        f.write("pair_style lj/cut %r\n" % float(rc.max()))
        for i in range(nt):
            for j in range(i, nt):
                f.write("pair_coeff %d %d %r %r %r\n" % (i + 1, j + 1, float(eps[i, j]), float(sig[i, j]), float(rc[i, j])))
        f.write("special_bonds lj %r %r %r\n" % tuple(fx["special_lj"]))
        f.write("timestep %r\n" % fx["dt"])
        f.write("dump 1 all dcd %d Model.dcd\n" % dcd_every)
    return path
