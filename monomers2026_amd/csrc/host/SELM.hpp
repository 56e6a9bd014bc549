// SELM.hpp -- the reference's Selm / Eulerian / Lagrangian plugin interface
// (cuda/SELM.cuh:6-28, cuda/SELM_Eulerian.cuh:6-29,
// cuda/SELM_Lagrangian.cuh:10-49), implemented over the C ABI of
// libselm_hip.
//
// Selm reads Model.SELM with the reference's keys and semantics
// (cuda/SELM.cu:17-48):
//   required: rho, T, mu, nsteps, KB, LagrangianType, EulerianType
//   optional: flagStochasticDriving, SELM_Seed, saveSkipSimulationData,
//             writeParticalV/F/X (all default 0; seed 123456)
//   numMeshPtsPerDir (required for Eulerian_Period)
//   SELM_Precision fp64 | fp32_bulk   (this build; default fp64)
//   SELM_F4Quirk 1 | 0                 (this build; default 1, the reference's kernel)
//   SELM_HostForces 0 | 1              (this build; default 0: harmonic angles and
//                                       lj/cut run on the device; 1: on the host,
//                                       handed over every step as LAMMPS does)
//
// The type strings are the reference's:
// - "Lagrangian_LAMMPS" selects Lagrangian_LAMMPS_HIP, which reads
//   Model.LAMMPS_script through the LAMMPS-subset reader.
// - "Eulerian_Period" selects Eulerian_Period_HIP.
//
// Division of work: one libselm_hip step carries the whole time step.  The
// bead update, spread, fluid update, thermal forcing for the next step and
// the device-resident bond and fix-addforce terms all run in one captured
// graph.  So Eulerian::run issues it, and Lagrangian::run and
// Eulerian::compute_f have nothing left to launch.
//
// Lagrangian::compute_f keeps the reference's host duties:
// - the particle{X,V,F}.dat files every saveSkipSimulationData steps
//   (cuda/SELM_Lagrangian_LAMMPS.cu:145-175);
// - `dump dcd` frames;
// - with SELM_HostForces 1, the harmonic angles and lj/cut pairs, added to
//   the device's bond and addforce forces and handed back through
//   selm_set_conservative_force.  By default these terms run on the device
//   with the bonds.
#pragma once

#include <memory>
#include <string>

#include "dcd_writer.hpp"
#include "lammps_subset.hpp"
#include "selm_hip.h"

namespace selm_host {

class Selm;

class Eulerian {
  public:
    int dim[3] = {0, 0, 0};
    double dx = 0.0;
    Selm *selm = nullptr;
    virtual void initial() = 0;
    virtual void compute_f() = 0;
    virtual void run() = 0;
    virtual void final() = 0;
    virtual ~Eulerian() = default;
};

class Lagrangian {
  public:
    int nlocal = 0;
    const double *lo = nullptr, *hi = nullptr;
    int saveSkipSimulationData = 0;
    int writeParticalV = 0, writeParticalF = 0, writeParticalX = 0;
    Selm *selm = nullptr;
    virtual void initial() = 0;
    virtual void compute_f() = 0;
    virtual void run() = 0;
    virtual void final() = 0;
    virtual ~Lagrangian() = default;
};

class Selm {
  public:
    int nsteps = 0;
    long long timestep = 0;
    double total_time = 0.0;
    int SELM_Seed = 123456;
    bool flagStochasticDriving = false;
    double KB = 0, T = 0, deltaT = 0, mu = 0, rho = 0;
    int precision = SELM_PREC_FP64;
    int f4_quirk = 1;
    int host_forces_on = 0;  // SELM_HostForces 1: angles and pairs on the host hook
    Eulerian *eulerian = nullptr;
    Lagrangian *lagrangian = nullptr;
    selm_ctx *ctx = nullptr;  // the system; created by Eulerian::initial

    // Reads `dir`/Model.SELM and, for Lagrangian_LAMMPS, `dir`/Model.LAMMPS_script.
    explicit Selm(const std::string &dir = ".");
    ~Selm();
    void initial();
    void compute_f();
    void run();
    void final();

    // Throws std::runtime_error with selm_last_error's text when rc != 0.
    void check(int rc, const char *what) const;
    std::string dir;
};

class Eulerian_Period_HIP : public Eulerian {
  public:
    void initial() override;
    void compute_f() override {}
    void run() override;
    void final() override;
};

class Lagrangian_LAMMPS_HIP : public Lagrangian {
  public:
    explicit Lagrangian_LAMMPS_HIP(const std::string &script);
    void initial() override;
    void compute_f() override;
    void run() override {}
    void final() override;
    Deck deck;

  private:
    void writedata();
    std::vector<double> x_, v_, fc_, ft_;
    bool host_terms_ = false;  // angles / pairs evaluated on the host each step
    std::vector<std::unique_ptr<DcdWriter>> dcd_;
};

}  // namespace selm_host
