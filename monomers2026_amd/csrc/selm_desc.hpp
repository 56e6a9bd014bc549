// selm_desc.hpp -- the per-replica descriptor read by every kernel, and the
// per-stage work tables that let one launch cover many heterogeneous replicas.
#pragma once

#include <stdint.h>

#include "selm_device.hpp"

namespace selm {

// Fluid arrays of one replica (T = double or float):
//   u     [3][nz][ny][nx]   velocity u^n (MAC faces: u_x at (i, j+1/2, k+1/2) dx)
//   fsp   [3][nz][ny][nx]   u* (the FFT input) on the runtime-length FFT path
//   pool  [slot][3][tile]   spread impulse of the tiles that hold beads
//   spec  [3][nz][ny][xc]   half spectrum (xc = nx/2+1), unnormalised DFT
// Bead arrays (always FP64, AoS [M][3] like LAMMPS atom->x):
//   x v fc ftot fext tf sg ; m [M]
struct RepDev {
    // ---- grid
    int n[3];             // nx, ny, nz
    int xc;               // nx/2 + 1
    long long N, Nc;      // nx*ny*nz, xc*ny*nz
    void *u, *fsp, *spec, *spec2;   // spec2: second spectrum layout of the transposing FFT
    void *nbuf;           // thermal forcing of this step in k-space, [3][xc][ny][nz] complex T
    void *pool;           // spread impulse per tile: [tile][3][te0*te1*te2] of T, valid when
    int *tstamp;          // tstamp[tile] == the current step (a bead touched the tile)
    const void *tw[3];    // cplx<T>[n_a]: e^{-2 pi i m / n_a}
    const void *ktab[3];  // per axis, T[4*n_a]: {lam_a, Dre_a, Dim_a, sqrt-part}
    // FP64 copies used by the compile-time-length FFT kernels, whose arithmetic
    // is FP64 in both precisions (the same arrays as tw/ktab in FP64; in
    // FP32-bulk, lam_a = -4 sin^2(theta_a/2))
    const void *twd[3];
    const void *ktabd[3];
    RadixPlan plan[3];
    int bx, by, bz;       // rows / columns per workgroup in the X, Y, Z passes
    // ---- tiles for binning / spread
    int te[3];            // tile edge per axis (cells)
    int tn[3];            // tiles per axis
    int ntiles;
    // ---- beads
    int M;
    double *x, *v, *fc, *ftot, *m, *fext, *tf, *sg;
    int *bcnt;                 // per tile: beads inserted this step (may exceed bcap)
    int *slist;                // [ntiles] split tiles of the step (more than split_at bucket
                               // entries), in the order their count passed split_at
    int split_at;              // > 0: tiles of more than split_at entries are listed in slist
                               // and taken first by the spread (systems of at least
                               // kSplitMinBeads beads; the batch's spread shape sets the
                               // threshold); 0: every tile in its own slot
    int bcap;                  // bucket slots per tile (>= kBucketCap; scaled to the bead density)
    double *bucket;            // [ntiles][bcap][6]: s[3] and t_f[3] of the beads inserted
                               // this step, in insertion order
    int *ovf;                  // [27*M][2] (tile, bead) pairs beyond a full bucket
    int *ctr;                  // [8]: overflow count per step parity [0,1]
    int *status;               // status flags (kStatus*), one word per replica in the batch's array
    int *adj_ptr, *adj_i1, *adj_i2;   // CSR over beads: bonds touching each bead, bond order
    double *adj_K, *adj_r0;
    int *ang_ptr, *ang_a, *ang_b, *ang_c;  // CSR over beads: harmonic angles (end, centre, end) touching each bead
    double *ang_K, *ang_t0;
    int lj_nt;                 // atom types of lj/cut (0: no pair term)
    int *lj_type;              // [M] type of each bead (0-based)
    double *lj_tab;            // [nt][nt][4]: eps, sigma, rc^2, 0
    int *spec_ptr, *spec_j;    // CSR over beads: 1-2, 1-3, 1-4 partners whose lj weight is not 1
    double *spec_w;
    // lj/cut neighbour bins (ljb_n = 0: all pairs): ljb_nb[a] >= 3 bins of
    // edge >= the largest cutoff per axis over [ljb_lo[a], ljb_lo[a] + Lb[a])
    int ljb_nb[3], ljb_n;
    double ljb_lo[3];
    int *ljb_of;               // [M] bin of each bead this step
    int *ljb_cnt;              // [ljb_n] count, then fill cursor; zero between steps
    int *ljb_start;            // [ljb_n + 1] first list entry of each bin
    int *ljb_part;             // [lj_bin_chunks(ljb_n)] bead count, then first list entry, of
                               // each chunk of kLjChunk consecutive bins
    int *ljb_list;             // [M] bead indices by bin, ascending within a bin
    // ---- scalars
    double dx, lo, hi, L;      // lo[0], hi[0] (reference wraps and grids with dim 0 only)
    double Lb[3];              // per-dimension box lengths (bond minimum image)
    double dt, rho, mu;
    double ups;                // 6 pi (2 dx) mu
    double bead_coef;          // sqrt(2 kB T ups dt)
    double noise_amp;          // N * sqrt(2 kB T dt mu / N) / rho
    double inv_dxsq;
    uint64_t seed;
    int noise, f4;
    long long *ts;             // completed steps (device counter)
    double *diag;              // [3] k=0 mode of w (momentum diagnostic), FP64
};

// A bead is binned into every tile its stencils touch: on each axis the union
// of the three staggered stencils spans cells floor(s-1/2)-1 .. floor(s)+2,
// which meets at most 3 tiles (2 unless a ragged tile is narrower than 4).
constexpr int kMaxTilesPerBead = 27;
// lj/cut bins are scanned in chunks of kLjChunk consecutive bins (one
// workgroup each); at most 64 bins per axis gives at most 256 chunks
constexpr int kLjChunk = 1024;
__host__ __device__ inline int lj_bin_chunks(int nbins) { return (nbins + kLjChunk - 1) / kLjChunk; }
constexpr int kBucketCap = 256;  // the fewest bucket slots per tile
constexpr int kSplitAt = 64;      // default split threshold: tiles with more bucket entries are
                                  // taken first by the spread
constexpr int kSplitMinBeads = 2048;

// A replica's ctr[8] (zeroed by init), by step parity: [0], [1] the (tile,
// bead) pairs beyond a full bucket; [4], [5] the split tiles.
constexpr int kCtrSplits = 4;
// Status flags (the same values as SELM_STATUS_* in selm_hip.h).
constexpr int kStatusTileOverflow = 1;   // a tile met more than kMaxTileBeads beads
constexpr int kStatusNonFinite = 2;      // a bead position or velocity became NaN or infinite
constexpr int kStatusDisplacement = 4;   // |v dt| > L/2 for a bead component (cuda/SELM_Kernel.cu:363-365)
// The spread sums each cell as 64-bit fixed point with every term below
// 2^49 in magnitude, so up to 2^14 terms per cell cannot overflow; a tile
// whose bead count exceeds this raises kStatusTileOverflow.
constexpr int kMaxTileBeads = 1 << 14;

// Blocks of a launch are dealt round-robin to the 8 XCDs (each with its own
// L2); remap so that consecutive logical blocks run on the same XCD and share
// cache lines.  G is the (padded) grid size, a multiple of 8.
__device__ __forceinline__ long long xcd_remap(long long b, long long G) {
    return (b & 7) * (G >> 3) + (b >> 3);
}

// Work table for one stage: item g of the launch belongs to replica r with
// start[r] <= g < start[r+1].
__device__ __forceinline__ int find_rep(const long long *start, int nrep, long long g) {
    int lo = 0, hi = nrep - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace selm
