// lammps_subset.hpp -- the part of a LAMMPS input deck that a SELM run uses.
//
// The reference hands its Model.LAMMPS_script to an embedded LAMMPS
// (cuda/SELM_Lagrangian_LAMMPS.cu:21-27) and from then on uses LAMMPS only
// for three things:
// - the atoms (x, v, per-type mass, box);
// - the timestep;
// - the conservative forces recomputed by setup_minimal after every step
//   (:118-129).
//
// This reader understands the commands the reference decks use
// (cases/SELM/test/Model.LAMMPS_script, selm/test/NucleotideMonomer/):
//
//   variable NAME equal|index|string|universe VALUE   (${NAME} and $X substitution)
//   units, boundary p p p, atom_style, log, thermo*, neighbor, neigh_modify
//       accepted; no effect on the step
//   read_data FILE   header counts, box, Masses, Atoms (angle/bond/molecular/
//       full/atomic styles), Velocities, Bonds, Angles
//   bond_style harmonic     bond_coeff T K r0         E = K (r - r0)^2
//   angle_style harmonic    angle_coeff T K theta0    E = K (theta - theta0)^2, theta0 in degrees
//   pair_style lj/cut RC    pair_coeff I J eps sigma [rc]
//       E = 4 eps ((sigma/r)^12 - (sigma/r)^6) for r < rc; unset pairs mixed
//       geometrically (LAMMPS's default pair_modify mix geometric)
//   special_bonds lj A B C (default 0 0 0: bonded 1-2, 1-3, 1-4 pairs have no LJ)
//   group NAME id|type|molecule LIST (integers and ranges a:b)
//   fix ID GROUP addforce FX FY FZ   constant force on the group
//   timestep DT
//   dump ID GROUP dcd N FILE
//
// Any other command, or a style other than these, is an error.  A deck that
// parses therefore runs with the physics it states.
#pragma once

#include <map>
#include <string>
#include <vector>

namespace selm_host {

struct DumpDcd {
    int every = 0;
    std::string file;
};

struct Deck {
    // box
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    double dt = 0.0;
    bool dt_set = false;
    std::string units = "lj";
    // atoms in ascending atom-ID order
    int natoms = 0, ntypes = 0, nbond_types = 0, nangle_types = 0;
    std::vector<int> id, mol, type;    // 1-based types
    std::vector<double> x, v;          // [natoms][3]
    std::vector<double> mass_of_type;  // [ntypes + 1]
    // topology (zero-based atom indices)
    std::vector<int> bond_type, bond_i, bond_j;
    std::vector<int> angle_type, angle_i, angle_j, angle_k;
    // styles and coefficients
    std::string bond_style, angle_style, pair_style;
    std::vector<double> bond_K, bond_r0;           // [nbond_types + 1]
    std::vector<double> angle_K, angle_theta0;     // radians
    double pair_rc = 0.0;
    std::vector<double> lj_eps, lj_sigma, lj_rc;   // [(ntypes+1)^2], -1 = unset
    std::vector<char> lj_set;
    double special_lj[3] = {0.0, 0.0, 0.0};
    // constant external forces, summed per atom: [natoms][3]
    std::vector<double> fext;
    std::map<std::string, std::vector<int>> groups;  // zero-based indices
    std::vector<DumpDcd> dumps;

    std::vector<double> masses() const;  // per atom
    bool has_host_terms() const;         // angles or pair terms present
};

// Parse the script at `script_path`; read_data paths are taken relative to
// the script's directory.  Throws std::runtime_error with the line on error.
Deck load_deck(const std::string &script_path);

// Host conservative force terms that the device does not evaluate: harmonic
// angles and lj/cut pairs with the special-bond exclusions, at positions x
// (minimum image in the periodic box).  Adds into f [natoms][3].  Returns the
// potential energy.
double host_forces(const Deck &d, const double *x, double *f);

// All conservative terms (bonds, angles, pairs, fix addforce) on the host:
// what LAMMPS's setup_minimal leaves in atom->f.  Overwrites f.
double all_forces(const Deck &d, const double *x, double *f);

}  // namespace selm_host
