// selm_kernels.hpp -- launch interface between the host engine and kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "selm_desc.hpp"

namespace selm {

enum Stage {
    kStageNoise = 0,  // thermal forcing in k-space (forked stream inside a step)
    kStageBead,
    kStageSpread,
    kStageFftX,  // + viscous predictor u* on the compile-time-length path

    kStageFftY,
    kStageFftZ,
    kStageFftYi,
    kStageFftXi,
    kStageBonds,
    kNumStages
};

// Everything one batched step launch needs (device pointers + totals).
struct StepLaunch {
    const RepDev *reps;   // device array [nrep]
    int nrep;
    bool fp64;
    const long long *st_bead, *st_tile, *st_pts, *st_fx, *st_fy, *st_fz;  // device [nrep+1]
    long long tot_bead, tot_tile, tot_pts, tot_fx, tot_fy, tot_fz;
    size_t lds_fx, lds_fy, lds_fz;   // dynamic LDS bytes per pass
    int twmax;                       // LDS twiddle slots (elements) preceding the data
    bool fast;                       // compile-time-length FFT kernels (selm_fft_kernels.hip)
    bool lj_bins;                    // some replica bins its beads for lj/cut
    int lj_maxbins;                  // the most lj/cut bins of one replica
    bool plane;                      // ... as the three-pass plane path (XYf, Zq, XYi)
    bool wt;                         // plane path: bulk outputs (forcing, spread pool, the three
                                     // passes' results) stored at device scope, written through
                                     // the XCD's L2 (the batch's store shape)
    int okmax;                       // plane path: LDS ints for one plane's tile stamps
    int zq_cols;                     // plane path: (kx, ky) columns per Z-pass workgroup (17 or 9)
    bool aux_bonds;                  // plane path: the XYi launch evaluates the bond forces (no
                                     // replica has angle or lj/cut terms); no bond stage of its own
    const long long *st_split;       // work table of the spread's split-tile workgroups
    long long tot_split;
    int split_at;                    // spread shape: split threshold (0: no split list)
    int noise_blocks;                // spread shape (plane path): thermal-forcing workgroups of
                                     // the spread launch (0: the default for the batch's modes)
    const long long *st_modes;       // thermal-forcing work table (half-spectrum modes)
    long long tot_modes;             // 0 when no replica has thermal forcing (or not fast)
    hipStream_t st2;                 // forked stream for the thermal forcing (five-pass path)
    hipEvent_t ev_fork, ev_join;
};

void launch_step(const StepLaunch &L, hipStream_t st);
// thermal-forcing workgroups of a plane-path spread launch for `modes` modes
// when the batch's spread shape leaves the count open
int spread_noise_default(long long modes);
void launch_stage(const StepLaunch &L, int stage, hipStream_t st);
void launch_bond_forces(const StepLaunch &L, int advance, hipStream_t st);
void launch_fill_f64(double *p, long long n, double v, hipStream_t st);
void launch_fill_f32(float *p, long long n, float v, hipStream_t st);

// selm_fft_kernels.hip
bool fft_fast_supported(const int n[3]);
void fft_fast_items(bool fp64, const int n[3], long long *ix, long long *iy, long long *iz);
void fft_fast_lds(bool fp64, const int n[3], int twmax, size_t *lx, size_t *ly, size_t *lz);
void launch_fft_fast(const StepLaunch &L, int stage, hipStream_t st);
void launch_noise(const StepLaunch &L, hipStream_t st);
bool fft_plane_shape(const int n[3]);
int fft_plane_okf(const int tn[3]);
void fft_plane_lds(bool fp64, const int n[3], int twmax, int okmax, int zcols, size_t *lxy, size_t *lz);
void fft_plane_items(bool fp64, const int n[3], int zcols, long long *ixy, long long *iz);
bool fft_plane_zcols_ok(int zcols);
hipError_t fft_plane_prepare(bool fp64, int twmax, size_t lds);

}  // namespace selm
