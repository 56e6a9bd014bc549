// selm_kernels.hip -- CDNA4 kernels of one SELM time step.
//
// Every kernel takes the replica descriptor array and a work table, so one
// launch advances any number of heterogeneous replicas (a single system is a
// batch of one).  Per step, in order:
//
//   k_bead_update     16 lanes per bead (4 beads per wave): 4x4x4 Peskin
//                     gather of u^n for the three staggered components, drag +
//                     thermal kick, x/v/f update, the bead's grid coordinate and
//                     impulse, insertion into the buckets of the tiles its
//                     stencils touch
//   k_spread          workgroup per tile: the bucket's coordinates and
//                     impulses staged in LDS, Peskin factors recomputed, every
//                     (bead, component) added by one wave (lanes = stencil
//                     points, distinct cells) as 64-bit fixed point, so the
//                     sums do not depend on the order; the tile goes to the
//                     sparse pool, untouched tiles are marked empty.  On the
//                     plane FFT path the same launch draws the thermal forcing
//   [k_ustar]         u* = u + (mu dt L7 u + f) / rho (runtime-length FFT path
//                     only; the compile-time-length X pass computes u* itself)
//   X, Y, Z+projection, Y, X passes of the 3-D FFT (selm_fft_kernels.hip for
//                     the compile-time lengths, below for the others)
//   [k_lj_bin, k_lj_chunk_sums/scan/apply, k_lj_fill, k_lj_sort]  lj/cut
//                     neighbour bins
//                     (only when a replica has them)
//   k_bond_forces     thread per bead: harmonic bonds and angles, lj/cut pairs
//                     and the external force at x^{n+1}; advances each
//                     replica's step counter (on the plane FFT path without
//                     angle or lj/cut terms the XYi launch does this)
#include <algorithm>

#include "selm_fluid.hpp"
#include "selm_forces.hpp"
#include "selm_kernels.hpp"

namespace selm {

typedef double f64x2 __attribute__((ext_vector_type(2)));

// Tiles met on one axis by the union of a bead's three staggered stencils.
__device__ __forceinline__ int axis_tiles(double s, int n, int te, int *out) {
    const int lo = (int)(floor(s - 0.5) - 1.0);
    const int hi = (int)floor(s) + 2;
    const float rte = 1.0f / (float)te;
    int cnt = 0;
    for (int c = lo; c <= hi && c - lo < 6; c++) {
        const int t = div_small(wrap_near(c, n), te, rte);
        bool dup = false;
        for (int q = 0; q < cnt; q++) dup |= out[q] == t;
        if (!dup) out[cnt++] = t;
    }
    return cnt;
}

// The four 1-D Peskin factors of a stencil at grid coordinate r: base
// p0 = floor(r) - 1 and w[l] = phi(|r - (p0 + l)|), phi as in
// cuda/SELM_Kernel.cu:387-395.  With f = r - floor(r) (exact) the distances
// are 1+f, f, 1-f and 2-f.  Both branches of phi take the same root there:
// 1 + 4r(1-r) at r = f or 1-f, and -7 + 4r(3-r) at r = 1+f or 2-f, all equal
// to 1 + 4f(1-f).  So with s = sqrt(1 + 4f(1-f)):
//   phi(1+f) = (3 - 2f - s)/8    phi(f)   = (3 - 2f + s)/8
//   phi(1-f) = (1 + 2f + s)/8    phi(2-f) = (1 + 2f - s)/8
// (at f = 0, phi(1) = 1/4 by either branch).  At f = 0 the last distance is
// exactly 2, beyond both branches: the reference leaves the factor at 1.0
// there (f4 != 0), the textbook kernel has 0.  One root per axis instead of
// four; the result differs from the branch-by-branch evaluation by rounding.
template <typename T>
__device__ __forceinline__ int peskin4(double r, int f4, T w[4]) {
    const double fl = floor(r);
    const T f = (T)(r - fl);
    const T sq = sqrt((T)1 + (T)4 * f * ((T)1 - f));
    const T a = (T)3 - (T)2 * f, b = (T)1 + (T)2 * f;
    w[0] = (a - sq) * (T)0.125;
    w[1] = (a + sq) * (T)0.125;
    w[2] = (b + sq) * (T)0.125;
    w[3] = f == (T)0 ? (f4 ? (T)1 : (T)0) : (b - sq) * (T)0.125;
    return (int)(fl - 1.0);
}

template <typename T>
__device__ __forceinline__ T pick4(const T w[4], int l) {
    return l == 0 ? w[0] : (l == 1 ? w[1] : (l == 2 ? w[2] : w[3]));
}

// ===========================================================================
// Beads: gather, drag, thermal kick, update.  Semantics of
// cuda/SELM_Kernel.cu:302-373 and mpi/SELM_Lagrangian_LAMMPS.cpp:115-180.
// ===========================================================================
constexpr int kBeadThreads = 256;
constexpr int kBeadsPerBlock = kBeadThreads / 16;

// Bead b of replica R, by the 16 lanes of its group (active: a real bead;
// inactive groups mirror the last bead and store nothing).
//
// The order of the memory operations follows from the vector memory counter
// (vmcnt), which counts loads, stores and atomics alike and lets a wait for
// one of them wait for every older one too:
//  1. the bead's position and state (x, v, f_cons, m);
//  2. the 12 gather loads of the three 4x4x4 stencils;
//  3. the slot claims in the bucket of every tile the stencils touch
//     (returning atomics), whose round trip overlaps the gather's;
//  4. the bead kick's normal, drawn while both are in flight;
//  5. the state stores, then the bucket entries, whose slots have arrived
//     by then.
// No load waits for a store, and the claims' results wait only for the
// loads issued before them.
template <typename T>
__device__ __forceinline__ void bead_update(const RepDev &R, int b, bool active) {
    const int sl = threadIdx.x & 15;  // sub-lane: (li, lj) column of the stencil
    const int nx = R.n[0], ny = R.n[1], nz = R.n[2];
    const long long N = R.N;
    const int f4 = R.f4;
    const SELM_GLOBAL double *gx = gbl((const double *)R.x);

    double xr[3], s[3];
#pragma unroll
    for (int d = 0; d < 3; d++) xr[d] = gx[3 * b + d];
    // the state of component sl (lanes 0..2 of the bead's group use it)
    const int ist = 3 * b + (sl < 3 ? sl : 0);
    const double vi = gbl(R.v)[ist], fci = gbl(R.fc)[ist], mb = gbl(R.m)[b];
#pragma unroll
    for (int d = 0; d < 3; d++) s[d] = (xr[d] - R.lo) / R.dx;
    const uint32_t step = (uint32_t)(R.ts[0] + 1);

    // Gamma: the gather loads of all three components, then their weights
    const int li = sl & 3, lj = sl >> 2;
    const SELM_GLOBAL T *u = gbl((const T *)R.u);
    T uv[3][4], wk[3][4];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        // component d is sampled at its face centre: the other two axes shift by -1/2
        const double rx = s[0] - (d != 0 ? 0.5 : 0.0);
        const double ry = s[1] - (d != 1 ? 0.5 : 0.0);
        const double rz = s[2] - (d != 2 ? 0.5 : 0.0);
        T wx[4], wy[4], fz[4];
        const int p0 = peskin4<T>(rx, f4, wx), p1 = peskin4<T>(ry, f4, wy), p2 = peskin4<T>(rz, f4, fz);
        const T fx = pick4(wx, li), fy = pick4(wy, lj);
        const long long col = (long long)wrap_near(p1 + lj, ny) * nx + wrap_near(p0 + li, nx);
        const SELM_GLOBAL T *ud = u + d * N;
#pragma unroll
        for (int k = 0; k < 4; k++) uv[d][k] = ud[(long long)wrap_near(p2 + k, nz) * ny * nx + col];
#pragma unroll
        for (int k = 0; k < 4; k++) wk[d][k] = (((T)1 * fz[k]) * fy) * fx;
    }

    // Slot claims: the bead goes into the bucket of every tile its stencils
    // touch, the bead's 16 lanes one (tile) insertion each, at most two per
    // lane (27 tiles at most).
    int tx[6], ty[6], tz[6];
    const int cx = axis_tiles(s[0], nx, R.te[0], tx);
    const int cy = axis_tiles(s[1], ny, R.te[1], ty);
    const int cz = axis_tiles(s[2], nz, R.te[2], tz);
    const int ncomb = active ? cx * cy * cz : 0;
    const float rcx = 1.0f / (float)cx, rcy = 1.0f / (float)cy;
    int tq[2], sq[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int q = sl + 16 * h;
        tq[h] = 0;
        sq[h] = 0;
        if (q < ncomb) {
            const int qa = div_small(q, cx, rcx), c = div_small(qa, cy, rcy);
            const int a = q - qa * cx, qb = qa - c * cy;
            tq[h] = (tz[c] * R.tn[1] + ty[qb]) * R.tn[0] + tx[a];
            sq[h] = __hip_atomic_fetch_add(gbl(R.bcnt) + tq[h], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }

    // the bead kick's normal (lanes 0..2: component sl)
    double z0 = 0.0;
    if (active && sl < 3 && R.noise) {
        double z1;
        normal_pair_f64(R.seed, kStreamBead, step, (uint32_t)(3 * b + sl), 0u, z0, z1);
    }

    double gam[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        T acc = (T)0;
#pragma unroll
        for (int k = 0; k < 4; k++) acc += uv[d][k] * wk[d][k];
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) acc += __shfl_xor(acc, m, 64);
        gam[d] = (double)acc;
    }
    if (!active) return;

    double tfl = 0.0;  // lanes 0..2 of the bead's group: t_f of component sl
    if (sl < 3) {
        const int d = sl;
        const double gd = d == 0 ? gam[0] : (d == 1 ? gam[1] : gam[2]);
        const double sd = d == 0 ? s[0] : (d == 1 ? s[1] : s[2]);
        const int i = 3 * b + d;
        const double dt = R.dt;
        double tf = -R.ups * (vi - gd) * dt;
        if (R.noise) tf += z0 * R.bead_coef;
        tfl = tf;
        gbl(R.tf)[i] = tf;
        gbl(R.sg)[i] = sd;
        const double ft = tf + dt * fci;
        gbl(R.ftot)[i] = ft / dt;
        double tmp = vi * dt;
        // a displacement beyond half the box is the reference's sign of a
        // broken run (cuda/SELM_Kernel.cu:363-365, where it is printed)
        int flags = fabs(tmp) > 0.5 * R.L ? kStatusDisplacement : 0;
        tmp += d == 0 ? xr[0] : (d == 1 ? xr[1] : xr[2]);
        if (tmp > R.hi) tmp -= R.L;
        else if (tmp < R.lo) tmp += R.L;
        const double vn = vi + ft / mb;
        flags |= (isfinite(tmp) && isfinite(vn)) ? 0 : kStatusNonFinite;
        if (flags) atomicOr(R.status, flags);
        gbl(R.x)[i] = tmp;
        gbl(R.v)[i] = vn;
    }
    // The bucket entries: what the spread needs of the bead, s at x^n and t_f.
    const int g16 = (threadIdx.x & 63) & ~15;
    const double t0 = __shfl(tfl, g16, 64), t1 = __shfl(tfl, g16 + 1, 64), t2 = __shfl(tfl, g16 + 2, 64);
#pragma unroll
    for (int h = 0; h < 2; h++) {
        if (sl + 16 * h >= ncomb) break;
        const int t = tq[h], slot = sq[h];
        if (slot < R.bcap) {
            SELM_GLOBAL f64x2 *e = gbl((f64x2 *)R.bucket) + 3 * ((long long)t * R.bcap + slot);
            e[0] = f64x2{s[0], s[1]};
            e[1] = f64x2{s[2], t0};
            e[2] = f64x2{t1, t2};
            // a tile of more than split_at entries goes to the split list
            // (once per step: the count passes split_at once; split_at is
            // below kBucketCap <= bcap, so the slot is a bucket slot)
            if (slot == R.split_at && R.split_at > 0)
                gbl(R.slist)[atomicAdd(&R.ctr[kCtrSplits + (step & 1)], 1)] = t;
        } else {
            const int o = atomicAdd(&R.ctr[step & 1], 1);
            gbl(R.ovf)[2 * o] = t;
            gbl(R.ovf)[2 * o + 1] = b;
        }
    }
}

// 5 waves per SIMD: 5 workgroups per CU, so 20k beads (1250 workgroups) are
// resident at once on 256 CUs instead of running in two rounds.  When every
// bead of a wave belongs to one replica (always for a single system), the
// replica index is taken from the first lane, so the descriptor's fields come
// through scalar loads instead of a chain of vector loads; a wave that
// straddles two replicas of a batch takes the per-lane form.
template <typename T>
__global__ __launch_bounds__(kBeadThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_bead_update(
    const RepDev *__restrict__ reps, int nrep, const long long *__restrict__ start, long long total) {
    const long long g0 = (long long)blockIdx.x * kBeadsPerBlock + (threadIdx.x >> 4);
    const bool active = g0 < total;
    const long long g = active ? g0 : total - 1;  // inactive groups mirror the last bead, store nothing
    const int r = find_rep(start, nrep, g);
    const int r0 = __builtin_amdgcn_readfirstlane(r);
    if (__ballot(r != r0) == 0) bead_update<T>(reps[r0], (int)(g - start[r0]), active);
    else bead_update<T>(reps[r], (int)(g - start[r]), active);
}

// ===========================================================================
// Spread (Lambda), workgroup per tile.  Semantics of cuda/SELM_Kernel.cu:424-470:
// each bead adds -t_f * w to the 4x4x4 stencil of each staggered component,
// w = phi_z phi_y phi_x, at the grid coordinate s = (x^n - lo)/dx of the
// bead update.
//
// Each cell is summed in 64-bit fixed point.  Every term -t_f w is rounded
// once to a multiple of 2^-S, with S = 1071 - E and E the largest biased
// FP64 exponent of the impulses t_f of the tile's beads (so |term| 2^S <=
// 2^49); integer sums do not depend on the order of the additions.  The
// tile's beads are therefore taken in whatever order the bucket insertions
// left them, with no sort, and the result is bitwise the same from run to
// run and alone or in a batch.  The tile is converted to T once when it is
// written.
//
// One thread per (bead, component) unit: it reads the bead's entry (s and
// t_f, 48 bytes), computes the 12 Peskin factors in registers with the bead
// update's own arithmetic (peskin4 on the same FP64 coordinate, so the
// factors are bitwise those of the gather) and adds the unit's stencil
// points that lie in the tile with ds_add_u64.  There is no staging and no
// barrier between units: a tile is one exponent pass, the units, and the
// write.
// ===========================================================================
constexpr int kTileCells = 512;   // tile edge <= 8
constexpr int kSpreadThreads = 256;
// Thermal-forcing workgroups of a spread launch: about 8 modes per thread,
// in whole multiples of the 256 CUs (blocks are dealt round-robin, and a CU
// with one forcing workgroup more than the others finishes last), at most
// kSpreadNoiseMax.  Timed at 128^3, 20k beads (1.06e6 modes): 128, 256, 512,
// 768, 1024, 2048 workgroups gave 0.174, 0.149, 0.142, 0.145, 0.145, 0.146 ms
// per step; 520 gave 0.147.  On the 256-system ensemble (1.1e7 modes) a cap
// of 512, 1024, 2048, 8192 gave 1.148, 1.093, 1.105, 1.112 ms per step.
constexpr int kSpreadModesPerThread = 8;
constexpr int kSpreadNoiseMax = 1024;
int spread_noise_default(long long modes) {
    if (modes <= 0) return 0;
    long long b = modes / ((long long)kSpreadThreads * kSpreadModesPerThread);
    b = b / 256 * 256;
    return (int)std::min<long long>(std::max<long long>(b, 256), kSpreadNoiseMax);
}

// round(v 2^S) as a 64-bit integer for |v 2^S| < 2^51 (scale = 2^S): adding
// 1.5 2^52 leaves the rounded integer in the low mantissa bits.
__device__ __forceinline__ unsigned long long fixed_term(double v, double scale) {
    const double t = __builtin_fma(v, scale, 6755399441055744.0);
    return (unsigned long long)(__double_as_longlong(t) - 0x4338000000000000LL);
}

// biased FP64 exponent of |v| (|v| < 2^(e - 1022) for e >= 1; 0 for v = 0)
__device__ __forceinline__ int exp_bits(double v) { return (int)((__double_as_longlong(v) >> 52) & 0x7ff); }

// scale exponent S for a largest biased impulse exponent E
__device__ __forceinline__ int spread_shift(int E) { return max(-1000, min(1000, 1071 - E)); }

struct SpreadTile {
    int nx, ny, nz;  // grid
    int ox, oy, oz;  // tile origin
    int ex, ey, ez;  // tile extent (ragged at the high end)
    int te0, te1;    // tile edge (cell index strides)
};

// One (bead, component) unit: bead coordinate sg (grid units) and impulse
// tf of component d.  The unit's 12 Peskin factors, with the x factors
// times -tf; per axis the 4 stencil cells, wrapped into [0, n) and taken
// relative to the tile origin, with a 4-bit mask of those inside the tile;
// then one fixed-point add per stencil point in the tile, in the order z, y,
// x with the product (w_z w_y) w_x.  Any grid edge >= 4 works: the 4 cells
// of an axis are distinct modulo n, so no point is added twice.
template <typename T>
__device__ __forceinline__ void spread_unit(char *accb, const double sg[3], double tf, int d, int f4,
                                            const SpreadTile &g, double scale) {
    const int n[3] = {g.nx, g.ny, g.nz}, o[3] = {g.ox, g.oy, g.oz}, e[3] = {g.ex, g.ey, g.ez};
    const int stride[3] = {8, 8 * g.te0, 8 * g.te0 * g.te1};  // bytes per cell step
    T w[3][4];
    int off[3][4];
    unsigned msk[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        // component d is sampled at its face centre: the other two axes shift by -1/2
        const double r = sg[a] - (d != a ? 0.5 : 0.0);
        const int p = wrap_near(peskin4<T>(r, f4, w[a]), n[a]);
        unsigned m = 0;
#pragma unroll
        for (int l = 0; l < 4; l++) {
            int c = p + l;
            c -= c >= n[a] ? n[a] : 0;
            c -= o[a];
            m |= ((unsigned)c < (unsigned)e[a] ? 1u : 0u) << l;
            off[a][l] = c * stride[a];
        }
        msk[a] = m;
    }
    const T fd = (T)(-tf);
#pragma unroll
    for (int l = 0; l < 4; l++) w[0][l] = w[0][l] * fd;  // x carries the force
    char *base = accb + 8 * d * kTileCells;
#pragma unroll
    for (int lk = 0; lk < 4; lk++) {
        if (!((msk[2] >> lk) & 1u)) continue;
#pragma unroll
        for (int lj = 0; lj < 4; lj++) {
            if (!((msk[1] >> lj) & 1u)) continue;
            const T wzy = w[2][lk] * w[1][lj];
            char *row = base + off[2][lk] + off[1][lj];
#pragma unroll
            for (int li = 0; li < 4; li++) {
                if (!((msk[0] >> li) & 1u)) continue;
                __hip_atomic_fetch_add((unsigned long long *)(row + off[0][li]), fixed_term((double)(wzy * w[0][li]), scale),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    }
}

constexpr int kSpreadWin = 256;  // overflow pairs a spread workgroup scans per round
struct SpreadShared {
    unsigned long long acc[3 * kTileCells];  // cell sums
    int list[kSpreadWin];                    // an overflow window's beads of this tile
    int nsel, emax;
    int pad[2];
};

// Tile t with cnt bucket insertions this step: the scale from every impulse
// of the tile, then the bucket entries' units and the tile's overflow pairs
// window by window, all adding into one LDS accumulator; the tile is
// rounded to T and written to the pool (WT: at device scope, FP32 only).
// Every thread runs the same sequence of barriers.
template <typename T, bool WT>
__device__ __forceinline__ void spread_tile(const RepDev &R, int t, int cnt, uint32_t step, int par,
                                            SpreadShared &sh) {
    const int tid = threadIdx.x;
    const int nx = R.n[0], ny = R.n[1], nz = R.n[2];
    const int te0 = R.te[0], te1 = R.te[1];
    const int tcells = te0 * te1 * R.te[2];
    const int tx = t % R.tn[0], ty = (t / R.tn[0]) % R.tn[1], tz = t / (R.tn[0] * R.tn[1]);
    const int ox = tx * te0, oy = ty * te1, oz = tz * R.te[2];
    const SpreadTile tl{nx, ny, nz, ox, oy, oz, min(te0, nx - ox), min(te1, ny - oy), min(R.te[2], nz - oz), te0, te1};
    const int nb = min(cnt, R.bcap);
    const int novf = cnt > R.bcap ? min(R.ctr[par], kMaxTilesPerBead * R.M) : 0;
    const SELM_GLOBAL f64x2 *src = gbl((const f64x2 *)R.bucket) + 3 * (long long)t * R.bcap;
    const int f4 = R.f4;
    char *accb = (char *)sh.acc;
    __syncthreads();  // a previous tile's use of LDS is over
    if (tid == 0) sh.emax = 0;
    for (int q = tid; q < 3 * kTileCells; q += kSpreadThreads) sh.acc[q] = 0ull;
    // the scale: the largest exponent of every impulse of the tile (bucket
    // entries' pieces 1 and 2, and the overflow pairs).  Thread tid takes the
    // impulses of units tid, tid + 256, ... (unit u = component u % 3 of
    // entry u / 3), as in the unit pass below, and keeps the entry of its
    // first unit in registers for that pass: one load of the entries for
    // tiles of up to 85.
    const bool has0 = tid < 3 * nb;
    const int k0 = tid / 3, d0 = tid - 3 * k0;
    f64x2 a0 = {0.0, 0.0}, a1 = {0.0, 0.0}, a2 = {0.0, 0.0};
    if (has0) {
        a0 = src[3 * k0];
        a1 = src[3 * k0 + 1];
        a2 = src[3 * k0 + 2];
    }
    const double tf0 = d0 == 0 ? a1.y : (d0 == 1 ? a2.x : a2.y);
    int e = has0 ? exp_bits(tf0) : 0;
    for (int u = tid + kSpreadThreads; u < 3 * nb; u += kSpreadThreads) {
        const int k = u / 3, dd = u - 3 * k;
        const f64x2 e1 = src[3 * k + 1], e2 = src[3 * k + 2];
        e = max(e, exp_bits(dd == 0 ? e1.y : (dd == 1 ? e2.x : e2.y)));
    }
    for (int o = tid; o < novf; o += kSpreadThreads)
        if (R.ovf[2 * o] == t) {
            const int bi = R.ovf[2 * o + 1];
#pragma unroll
            for (int dd = 0; dd < 3; dd++) e = max(e, exp_bits(gbl(R.tf)[3 * bi + dd]));
        }
    // the wave's largest exponent by shuffles, then one LDS atomic per wave
    // (an atomic from every lane becomes a per-lane scalar loop)
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) e = max(e, __shfl_xor(e, m, 64));
    __syncthreads();  // emax = 0 and the zeroed cells are seen by every thread
    if ((tid & 63) == 0) atomicMax(&sh.emax, e);
    __syncthreads();
    const int E = sh.emax;
    const double scale = ldexp(1.0, spread_shift(E));
    if (has0) {
        const double sg[3] = {a0.x, a0.y, a1.x};
        spread_unit<T>(accb, sg, tf0, d0, f4, tl, scale);
    }
    for (int u = tid + kSpreadThreads; u < 3 * nb; u += kSpreadThreads) {
        const int k = u / 3, dd = u - 3 * k;
        const f64x2 e0 = src[3 * k], e1 = src[3 * k + 1], e2 = src[3 * k + 2];
        const double sg[3] = {e0.x, e0.y, e1.x};
        spread_unit<T>(accb, sg, dd == 0 ? e1.y : (dd == 1 ? e2.x : e2.y), dd, f4, tl, scale);
    }
    // the overflow pairs of this tile, window by window
    for (int o0 = 0; o0 < novf; o0 += kSpreadWin) {
        __syncthreads();
        if (tid == 0) sh.nsel = 0;
        __syncthreads();
        if (o0 + tid < novf && R.ovf[2 * (o0 + tid)] == t) sh.list[atomicAdd(&sh.nsel, 1)] = R.ovf[2 * (o0 + tid) + 1];
        __syncthreads();
        const int nq = __builtin_amdgcn_readfirstlane(sh.nsel);
        for (int u = tid; u < 3 * nq; u += kSpreadThreads) {
            const int k = u / 3, dd = u - 3 * k, bi = sh.list[k];
            const SELM_GLOBAL double *sgb = gbl(R.sg) + 3 * bi;
            const double sg[3] = {sgb[0], sgb[1], sgb[2]};
            spread_unit<T>(accb, sg, gbl(R.tf)[3 * bi + dd], dd, f4, tl, scale);
        }
    }
    __syncthreads();
    const double inv = ldexp(1.0, -spread_shift(E));
    SELM_GLOBAL T *pool = gbl((T *)R.pool) + (long long)t * 3 * tcells;
    for (int q = tid; q < 3 * tcells; q += kSpreadThreads) {
        const int dd = q / tcells;
        const T v = (T)((double)(long long)sh.acc[dd * kTileCells + (q - dd * tcells)] * inv);
        if constexpr (WT && sizeof(T) == 4) st_dev(pool + q, v);
        else pool[q] = v;
    }
    if (tid == 0) {
        if (cnt > kMaxTileBeads) atomicOr(R.status, kStatusTileOverflow);
        R.bcnt[t] = 0;  // ready for the next step's insertions
        R.tstamp[t] = (int)step;
    }
}

// One launch, three roles.  The first nnoise workgroups draw the thermal
// forcing of the step into nbuf (plane path; work table st_modes, each
// thread striding over the modes): about two workgroups per CU for most of
// the launch, at the default issue priority, so the spread's VALU-light LDS
// work hides much of the forcing's arithmetic.  The next tot_split workgroups (work
// table st_split) take the split tiles of the step, which the bead update
// listed: workgroup j of a replica takes split tiles j, j + G, ... (G = its
// share); they come early so that the fullest tiles start at once.  The
// remaining workgroups (work table st_tile) take one tile each: every touched
// tile, or with R.split_at > 0 those of at most split_at bucket entries.
// Waves per SIMD: 7 in FP32-bulk (at most 72 VGPRs, a few spilled; 1 µs
// faster per step at 128^3, 20k beads, than 6 with none spilled), 5 in FP64.
// WT: the batch's store shape (FP32-bulk plane path): forcing and pool
// stored at device scope.
template <typename T, bool WT>
__global__ __launch_bounds__(kSpreadThreads) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 7 : 5))) void k_spread(
    const RepDev *__restrict__ reps, int nrep,
                                                           const long long *__restrict__ st_modes,
                                                           long long tot_modes, int nnoise,
                                                           const long long *__restrict__ st_split,
                                                           long long tot_split,
                                                           const long long *__restrict__ st_tile,
                                                           long long tot_tile) {
    __shared__ SpreadShared sh;
    if constexpr (sizeof(T) == 4) {  // the plane path is FP32-bulk; FP64 launches have nnoise = 0
        if ((int)blockIdx.x < nnoise) {
            const long long stride = (long long)nnoise * kSpreadThreads;
            if (nrep == 1) {
                // one system: four modes per call, their table values loaded
                // before the first store
                for (long long m = (long long)blockIdx.x * kSpreadThreads + threadIdx.x; m < tot_modes; m += 4 * stride)
                    noise_modes_rep<T, 4, WT>(reps[0], m, stride, tot_modes, 1);
            } else {
                for (long long m = (long long)blockIdx.x * kSpreadThreads + threadIdx.x; m < tot_modes; m += stride)
                    noise_mode<T, WT>(reps, nrep, st_modes, m, 1);
            }
            return;
        }
    }
    const long long gb = blockIdx.x - nnoise;
    if (gb < tot_split) {
        const int r = find_rep(st_split, nrep, gb);
        const RepDev &R = reps[r];
        const int j = (int)(gb - st_split[r]), G = (int)(st_split[r + 1] - st_split[r]);
        const uint32_t step = (uint32_t)(R.ts[0] + 1);
        const int par = step & 1;
        const int nsplit = R.ctr[kCtrSplits + par];
        if (j >= nsplit) return;
        // issue priority over the thermal forcing
        __builtin_amdgcn_s_setprio(3);
        for (int k = j; k < nsplit; k += G) {
            const int t = gbl(R.slist)[k];
            spread_tile<T, WT>(R, t, R.bcnt[t], step, par, sh);
        }
        return;
    }
    const long long g = gb - tot_split;
    if (g >= tot_tile) return;
    const int r = find_rep(st_tile, nrep, g);
    const RepDev &R = reps[r];
    const int t = (int)(g - st_tile[r]);
    const uint32_t step = (uint32_t)(R.ts[0] + 1);
    const int par = step & 1;
    if (t == 0 && threadIdx.x == 0) {
        // the next step's counters (last used a step ago)
        R.ctr[par ^ 1] = 0;
        R.ctr[kCtrSplits + (par ^ 1)] = 0;
    }
    const int cnt = R.bcnt[t];
    // untouched tiles keep a stale stamp; split tiles have their own workgroups
    if (cnt == 0 || (R.split_at > 0 && cnt > R.split_at)) return;
    // issue priority over the thermal forcing
    __builtin_amdgcn_s_setprio(3);
    spread_tile<T, WT>(R, t, cnt, step, par, sh);
}

// u* over the whole grid into fsp (runtime-length FFT path only)
template <typename T>
__global__ __launch_bounds__(256) void k_ustar(const RepDev *__restrict__ reps, int nrep,
                                               const long long *__restrict__ start, long long total) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const long long n = g - start[r];
    const int nx = R.n[0], ny = R.n[1];
    const long long Nyx = (long long)nx * ny;
    const int z = (int)(n / Nyx), y = (int)((n % Nyx) / nx), x = (int)(n % nx);
    int tyz, lyz;
    spread_row(R, y, z, tyz, lyz);
    const int step = (int)(R.ts[0] + 1);
    T *out = (T *)R.fsp;
#pragma unroll
    for (int d = 0; d < 3; d++) out[d * R.N + n] = ustar_at<T>(R, d, x, y, z, tyz, lyz, step);
}

// ===========================================================================
// FFT passes.  LDS layout is element-major: element e of sequence s at
// buf[e*S + s], so consecutive lanes touch consecutive sequences.
// ===========================================================================
template <typename T>
__device__ __forceinline__ void load_twiddles(cplx<T> *dst, const cplx<T> *src, int n) {
    for (int m = threadIdx.x; m < n; m += blockDim.x) dst[m] = src[m];
}

// ===========================================================================
// Runtime-length FFT passes (any edge that is a product of 2, 3 and 5); the
// compile-time lengths use selm_fft_kernels.hip.
// ===========================================================================
// Forward r2c along x of u* (fsp): rows rho = (d*nz + z)*ny + y of length nx;
// two real rows form one complex sequence z = a + i b, separated after the DFT.
template <typename T>
__global__ __launch_bounds__(256) void k_fft_x_r2c(const RepDev *__restrict__ reps, int nrep,
                                                   const long long *__restrict__ start,
                                                   long long total, int lds_stride_elems) {
    extern __shared__ __align__(16) unsigned char smem[];
    const long long g = blockIdx.x;
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const int nx = R.n[0], xc = R.xc;
    const int BX = R.bx, P = BX / 2, S = P | 1;
    const long long rows = 3LL * R.n[1] * R.n[2];
    const long long row0 = (g - start[r]) * BX;
    cplx<T> *tw = (cplx<T> *)smem;
    cplx<T> *a = tw + lds_stride_elems;
    cplx<T> *b = a + (nx + 2) * S;
    load_twiddles(tw, (const cplx<T> *)R.tw[0], nx);
    const T *in = (const T *)R.fsp;
    for (int idx = threadIdx.x; idx < BX * nx; idx += blockDim.x) {
        int rl = idx / nx, e = idx - rl * nx;
        long long row = row0 + rl;
        T v = row < rows ? in[row * nx + e] : (T)0;
        T *slot = (T *)&a[e * S + (rl >> 1)];
        slot[rl & 1] = v;
    }
    __syncthreads();
    cplx<T> *res = lds_fft<T>(a, b, R.plan[0], S, tw, -1);
    cplx<T> *out = (cplx<T> *)R.spec;
    for (int idx = threadIdx.x; idx < BX * xc; idx += blockDim.x) {
        int rl = idx / xc, k = idx - rl * xc;
        long long row = row0 + rl;
        if (row >= rows) continue;
        int p = rl >> 1;
        cplx<T> zk = res[k * S + p];
        cplx<T> zm = res[((nx - k) % nx) * S + p];
        cplx<T> o;
        if ((rl & 1) == 0) o = mk<T>((zk.x + zm.x) * (T)0.5, (zk.y - zm.y) * (T)0.5);
        else o = mk<T>((zk.y + zm.y) * (T)0.5, (zm.x - zk.x) * (T)0.5);
        out[row * xc + k] = o;
    }
}

// Inverse c2r along x with FFTW semantics: only the real parts of the DC and
// Nyquist coefficients enter.  Writes u^{n+1}.
template <typename T>
__global__ __launch_bounds__(256) void k_fft_x_c2r(const RepDev *__restrict__ reps, int nrep,
                                                   const long long *__restrict__ start,
                                                   long long total, int lds_stride_elems) {
    extern __shared__ __align__(16) unsigned char smem[];
    const long long g = blockIdx.x;
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const int nx = R.n[0], xc = R.xc;
    const int BX = R.bx, P = BX / 2, S = P | 1;
    const long long rows = 3LL * R.n[1] * R.n[2];
    const long long row0 = (g - start[r]) * BX;
    cplx<T> *tw = (cplx<T> *)smem;
    cplx<T> *a = tw + lds_stride_elems;
    cplx<T> *b = a + (nx + 2) * S;
    load_twiddles(tw, (const cplx<T> *)R.tw[0], nx);
    const cplx<T> *in = (const cplx<T> *)R.spec;
    for (int idx = threadIdx.x; idx < BX * xc; idx += blockDim.x) {
        int rl = idx / xc, k = idx - rl * xc;
        long long row = row0 + rl;
        cplx<T> v = row < rows ? in[row * xc + k] : mk<T>((T)0, (T)0);
        if (k == 0 || 2 * k == nx) v.y = (T)0;
        b[((rl & 1) * xc + k) * S + (rl >> 1)] = v;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < nx * P; idx += blockDim.x) {
        int e = idx / P, p = idx - e * P;
        cplx<T> A, B;
        if (e < xc) {
            A = b[e * S + p];
            B = b[(xc + e) * S + p];
        } else {
            A = cconj(b[(nx - e) * S + p]);
            B = cconj(b[(xc + nx - e) * S + p]);
        }
        a[e * S + p] = mk<T>(A.x - B.y, A.y + B.x);
    }
    __syncthreads();
    cplx<T> *res = lds_fft<T>(a, b, R.plan[0], S, tw, +1);
    T *out = (T *)R.u;
    for (int idx = threadIdx.x; idx < BX * nx; idx += blockDim.x) {
        int rl = idx / nx, e = idx - rl * nx;
        long long row = row0 + rl;
        if (row >= rows) continue;
        cplx<T> z = res[e * S + (rl >> 1)];
        out[row * nx + e] = (rl & 1) ? z.y : z.x;
    }
}

// Complex DFT along y, in place on spec; dir = -1 forward, +1 inverse.
template <typename T>
__global__ __launch_bounds__(256) void k_fft_y(const RepDev *__restrict__ reps, int nrep,
                                               const long long *__restrict__ start,
                                               long long total, int dir, int lds_stride_elems) {
    extern __shared__ __align__(16) unsigned char smem[];
    const long long g = blockIdx.x;
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const int ny = R.n[1], nz = R.n[2], xc = R.xc;
    const int BY = R.by;
    const int nbx = (xc + BY - 1) / BY;
    const long long item = g - start[r];
    const int xb = (int)(item % nbx);
    const long long dz = item / nbx;  // d*nz + z
    const int x0 = xb * BY;
    cplx<T> *tw = (cplx<T> *)smem;
    cplx<T> *a = tw + lds_stride_elems;
    cplx<T> *b = a + ny * BY;
    load_twiddles(tw, (const cplx<T> *)R.tw[1], ny);
    cplx<T> *sp = (cplx<T> *)R.spec + dz * ny * xc;
    for (int idx = threadIdx.x; idx < ny * BY; idx += blockDim.x) {
        int y = idx / BY, c = idx - y * BY;
        a[idx] = (x0 + c < xc) ? sp[(long long)y * xc + x0 + c] : mk<T>((T)0, (T)0);
    }
    __syncthreads();
    cplx<T> *res = lds_fft<T>(a, b, R.plan[1], BY, tw, dir);
    for (int idx = threadIdx.x; idx < ny * BY; idx += blockDim.x) {
        int y = idx / BY, c = idx - y * BY;
        if (x0 + c < xc) sp[(long long)y * xc + x0 + c] = res[idx];
    }
}

// Columns along z, all three components of a mode together: forward DFT,
// thermal forcing, projection, k=0 removal, 1/N, inverse DFT.
//
// Thermal forcing (cuda/SELM_Kernel.cu:37-85 + Z2D + gthm_scale :88-98): the
// reference draws per half-spectrum mode (i,j,k) and component d the complex
// value (xi1*re, xi2*im), re = im = coef/sqrt(2) (re = coef, im = 0 for a
// self-conjugate mode), coef = sqrt(sum_a 2(1-cos theta_a))/dx, and maps it
// to real space with an unnormalised c2r scaled by sqrt(2 kB T dt mu / N).
// In the x-frequency planes i = 0 and i = nx/2 the c2r keeps only
// (g(i,j,k) + conj g(i,-j,-k))/2, so that is what is added here; the DFT of
// the real field is N times that, and u* receives f/rho.
//
// Projection (cuda/SELM_Eulerian_Period.cu:109-125): with D_a = e^{i theta_a}-1
// (forward difference), G_a = 1-e^{-i theta_a} (backward difference) and
// lambda = sum_a 2(cos theta_a - 1) = sum_a G_a D_a, the update
// u = u* - G lambda^{-1} D.u* is the DFT image of div -> Poisson -> grad;
// the dx factors cancel.  The k=0 mode is zeroed (p(0)=0, then mean removal).
template <typename T>
__global__ __launch_bounds__(256) void k_fft_z_proj(const RepDev *__restrict__ reps, int nrep,
                                                    const long long *__restrict__ start,
                                                    long long total, int lds_stride_elems) {
    extern __shared__ __align__(16) unsigned char smem[];
    const long long g = blockIdx.x;
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const int nx = R.n[0], ny = R.n[1], nz = R.n[2], xc = R.xc;
    const int BZ = R.bz, S = 3 * BZ;
    const long long ncol = (long long)ny * xc;
    const long long c0 = (g - start[r]) * BZ;
    cplx<T> *tw = (cplx<T> *)smem;
    cplx<T> *a = tw + lds_stride_elems;
    cplx<T> *b = a + nz * S;
    load_twiddles(tw, (const cplx<T> *)R.tw[2], nz);
    cplx<T> *sp = (cplx<T> *)R.spec;
    for (int idx = threadIdx.x; idx < nz * S; idx += blockDim.x) {
        int e = idx / S, rem = idx - e * S;
        int d = rem / BZ, c = rem - d * BZ;
        long long col = c0 + c;
        a[idx] = col < ncol ? sp[((long long)d * nz + e) * ncol + col] : mk<T>((T)0, (T)0);
    }
    __syncthreads();
    cplx<T> *res = lds_fft<T>(a, b, R.plan[2], S, tw, -1);
    cplx<T> *oth = (res == a) ? b : a;

    const T *kx = (const T *)R.ktab[0], *ky = (const T *)R.ktab[1], *kz = (const T *)R.ktab[2];
    const T invN = (T)(1.0 / (double)R.N);
    const uint32_t step = (uint32_t)(R.ts[0] + 1);
    const T amp = (T)R.noise_amp;
    for (int idx = threadIdx.x; idx < nz * BZ; idx += blockDim.x) {
        int e = idx / BZ, c = idx - e * BZ;
        long long col = c0 + c;
        if (col >= ncol) continue;
        int j = (int)(col / xc), i = (int)(col - (long long)j * xc), k = e;
        cplx<T> w[3];
#pragma unroll
        for (int d = 0; d < 3; d++) w[d] = res[e * S + d * BZ + c];
        const T lx = kx[4 * i], ly = ky[4 * j], lz = kz[4 * k];
        const T lam = (lx + ly) + lz;
        if (R.noise) {
            cplx<T> gv[3];
            thermal_mode<T>(R, i, j, k, lam, step, gv);
#pragma unroll
            for (int d = 0; d < 3; d++) w[d] = cadd(w[d], gv[d]);
        }
        if (i == 0 && j == 0 && k == 0) {
            if (R.diag) {
#pragma unroll
                for (int d = 0; d < 3; d++) R.diag[d] = (double)w[d].x;
            }
#pragma unroll
            for (int d = 0; d < 3; d++) res[e * S + d * BZ + c] = mk<T>((T)0, (T)0);
            continue;
        }
        cplx<T> D[3];
        D[0] = mk<T>(kx[4 * i + 1], kx[4 * i + 2]);
        D[1] = mk<T>(ky[4 * j + 1], ky[4 * j + 2]);
        D[2] = mk<T>(kz[4 * k + 1], kz[4 * k + 2]);
        cplx<T> sdiv = cadd(cadd(cmul(D[0], w[0]), cmul(D[1], w[1])), cmul(D[2], w[2]));
        cplx<T> q = mk<T>(sdiv.x / lam, sdiv.y / lam);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            cplx<T> G = mk<T>(-D[d].x, D[d].y);
            cplx<T> ud = csub(w[d], cmul(G, q));
            res[e * S + d * BZ + c] = cscale(ud, invN);
        }
    }
    __syncthreads();
    cplx<T> *res2 = lds_fft<T>(res, oth, R.plan[2], S, tw, +1);
    for (int idx = threadIdx.x; idx < nz * S; idx += blockDim.x) {
        int e = idx / S, rem = idx - e * S;
        int d = rem / BZ, c = rem - d * BZ;
        long long col = c0 + c;
        if (col < ncol) sp[((long long)d * nz + e) * ncol + col] = res2[idx];
    }
}

// ===========================================================================
// Conservative forces at x^{n+1}: LAMMPS bond_style harmonic
// (E = K (r - r0)^2, F = -2K (r - r0) r_hat, minimum image) in bond order,
// then fix addforce.  Block 0 also advances every replica's step counter.
// ===========================================================================
// ---------------------------------------------------------------------------
// lj/cut neighbour bins, rebuilt at every force evaluation (beads move):
// k_lj_bin counts the beads of each bin, k_lj_chunk_* turn the counts into
// list offsets and fill cursors, k_lj_fill files each bead, and k_lj_sort orders
// every bin's list by bead index and clears its count for the next step.
// The lists are thus the same whatever order the atomics ran in.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lj_bin_axis(double x, double lo, double L, int nb) {
    double s = (x - lo) / L;
    s -= floor(s);  // periodic image in [0, 1)
    return min((int)(s * nb), nb - 1);
}

__global__ __launch_bounds__(256) void k_lj_bin(const RepDev *__restrict__ reps, int nrep,
                                                const long long *__restrict__ start, long long total) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    if (R.ljb_n == 0) return;
    const int b = (int)(g - start[r]);
    if (!(isfinite(R.x[3 * b]) && isfinite(R.x[3 * b + 1]) && isfinite(R.x[3 * b + 2]))) {
        // a bead that has left the finite numbers is filed in no bin (and
        // flagged): it would otherwise share one bin with every such bead
        R.ljb_of[b] = -1;
        atomicOr(R.status, kStatusNonFinite);
        return;
    }
    const int ix = lj_bin_axis(R.x[3 * b], R.ljb_lo[0], R.Lb[0], R.ljb_nb[0]);
    const int iy = lj_bin_axis(R.x[3 * b + 1], R.ljb_lo[1], R.Lb[1], R.ljb_nb[1]);
    const int iz = lj_bin_axis(R.x[3 * b + 2], R.ljb_lo[2], R.Lb[2], R.ljb_nb[2]);
    const int bin = (iz * R.ljb_nb[1] + iy) * R.ljb_nb[0] + ix;
    R.ljb_of[b] = bin;
    atomicAdd(&R.ljb_cnt[bin], 1);
}

// Exclusive scan of the bin counts in three launches, every one a grid over
// (chunk of kLjChunk bins, replica): the chunks' totals, a scan of those
// totals (one workgroup per replica, at most 256 chunks), then each chunk's
// own scan from its offset.  Each thread takes 4 consecutive bins.
constexpr int kLjPer = kLjChunk / 256;

// exclusive prefix over the 256 threads of a workgroup; *total = the sum
__device__ __forceinline__ int block_exclusive_scan(int v, int *total) {
    __shared__ int wsum[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int base = 0;
    for (int q = 0; q < w; q++) base += wsum[q];
    *total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return base + x - v;
}

__global__ __launch_bounds__(256) void k_lj_chunk_sums(const RepDev *__restrict__ reps, int nrep) {
    const RepDev &R = reps[blockIdx.y];
    const int n = R.ljb_n, a = (int)blockIdx.x * kLjChunk;
    if (a >= n) return;
    int v = 0;
#pragma unroll
    for (int k = 0; k < kLjPer; k++) {
        const int i = a + (int)threadIdx.x * kLjPer + k;
        v += i < n ? R.ljb_cnt[i] : 0;
    }
    int tot = 0;
    block_exclusive_scan(v, &tot);
    if (threadIdx.x == 0) R.ljb_part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_lj_chunk_scan(const RepDev *__restrict__ reps, int nrep) {
    const RepDev &R = reps[blockIdx.x];
    const int n = R.ljb_n;
    if (n == 0) return;
    const int nc = lj_bin_chunks(n);  // <= 256
    const int t = threadIdx.x;
    const int v = t < nc ? R.ljb_part[t] : 0;
    int tot = 0;
    const int off = block_exclusive_scan(v, &tot);
    if (t < nc) R.ljb_part[t] = off;
    if (t == 0) R.ljb_start[n] = tot;
}

__global__ __launch_bounds__(256) void k_lj_chunk_apply(const RepDev *__restrict__ reps, int nrep) {
    const RepDev &R = reps[blockIdx.y];
    const int n = R.ljb_n, a = (int)blockIdx.x * kLjChunk;
    if (a >= n) return;
    int c[kLjPer];
    int v = 0;
#pragma unroll
    for (int k = 0; k < kLjPer; k++) {
        const int i = a + (int)threadIdx.x * kLjPer + k;
        c[k] = i < n ? R.ljb_cnt[i] : 0;
        v += c[k];
    }
    int tot = 0;
    int off = R.ljb_part[blockIdx.x] + block_exclusive_scan(v, &tot);
#pragma unroll
    for (int k = 0; k < kLjPer; k++) {
        const int i = a + (int)threadIdx.x * kLjPer + k;
        if (i < n) {
            R.ljb_start[i] = off;
            R.ljb_cnt[i] = off;  // fill cursor
        }
        off += c[k];
    }
}

__global__ __launch_bounds__(256) void k_lj_fill(const RepDev *__restrict__ reps, int nrep,
                                                 const long long *__restrict__ start, long long total) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    if (R.ljb_n == 0) return;
    const int b = (int)(g - start[r]);
    const int bin = R.ljb_of[b];
    if (bin >= 0) R.ljb_list[atomicAdd(&R.ljb_cnt[bin], 1)] = b;
}

// a thread per bin (grid over bins x replicas): insertion sort of the bin's
// list by bead index, and the count cleared for the next evaluation
__global__ __launch_bounds__(256) void k_lj_sort(const RepDev *__restrict__ reps, int nrep) {
    const RepDev &R = reps[blockIdx.y];
    const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
    if (i < R.ljb_n) {
        const int a = R.ljb_start[i], e = R.ljb_start[i + 1];
        for (int q = a + 1; q < e; q++) {
            const int v = R.ljb_list[q];
            int p = q - 1;
            while (p >= a && R.ljb_list[p] > v) {
                R.ljb_list[p + 1] = R.ljb_list[p];
                p--;
            }
            R.ljb_list[p + 1] = v;
        }
        R.ljb_cnt[i] = 0;
    }
}

// lj/cut force on bead b from bead j (b != j), added to f
__device__ __forceinline__ void lj_pair(const RepDev &R, int b, int j, int ti, int s0, int s1, double xb, double yb,
                                        double zb, double &f0, double &f1, double &f2) {
    double del[3] = {xb - R.x[3 * j], yb - R.x[3 * j + 1], zb - R.x[3 * j + 2]};
#pragma unroll
    for (int d = 0; d < 3; d++)
        if (fabs(del[d]) > 0.5 * R.Lb[d]) del[d] += (del[d] < 0.0) ? R.Lb[d] : -R.Lb[d];
    const double r2 = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
    const double *tab = R.lj_tab + 4 * (ti * R.lj_nt + R.lj_type[j]);
    if (r2 >= tab[2]) return;
    double wgt = 1.0;
    for (int q = s0; q < s1; q++)
        if (R.spec_j[q] == j) wgt = R.spec_w[q];
    if (wgt == 0.0) return;
    const double s2 = tab[1] * tab[1] / r2;
    const double s6 = s2 * s2 * s2;
    const double fr = wgt * 24.0 * tab[0] * (2.0 * s6 * s6 - s6) / r2;
    f0 += fr * del[0];
    f1 += fr * del[1];
    f2 += fr * del[2];
}

__global__ __launch_bounds__(256) void k_bond_forces(const RepDev *__restrict__ reps, int nrep,
                                                     const long long *__restrict__ start,
                                                     long long total, int advance) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (advance && blockIdx.x == 0) {
        for (int q = threadIdx.x; q < nrep; q += blockDim.x) reps[q].ts[0] += 1;
    }
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const int b = (int)(g - start[r]);
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    bond_terms(R, b, f0, f1, f2);
    // harmonic angles (LAMMPS angle_style harmonic): E = K (theta - theta0)^2
    // at the centre bead; -dE/dx = (2 K dtheta / sin theta) dcos/dx, with sin
    // theta held away from 0 by 0.001 as LAMMPS does.
    for (int q = R.ang_ptr[b]; q < R.ang_ptr[b + 1]; q++) {
        const int ia = R.ang_a[q], ib = R.ang_b[q], ic = R.ang_c[q];
        double u[3], w[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double du = R.x[3 * ia + d] - R.x[3 * ib + d];
            if (fabs(du) > 0.5 * R.Lb[d]) du += (du < 0.0) ? R.Lb[d] : -R.Lb[d];
            double dw = R.x[3 * ic + d] - R.x[3 * ib + d];
            if (fabs(dw) > 0.5 * R.Lb[d]) dw += (dw < 0.0) ? R.Lb[d] : -R.Lb[d];
            u[d] = du;
            w[d] = dw;
        }
        const double ru2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        const double rw2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
        const double ru = sqrt(ru2), rw = sqrt(rw2);
        double cs = (u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) / (ru * rw);
        cs = fmin(1.0, fmax(-1.0, cs));
        double sn = sqrt(1.0 - cs * cs);
        if (sn < 0.001) sn = 0.001;
        const double dth = acos(cs) - R.ang_t0[q];
        const double g = 2.0 * R.ang_K[q] * dth / sn;
        double fa[3], fc[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            fa[d] = g * (w[d] / (ru * rw) - cs * u[d] / ru2);
            fc[d] = g * (u[d] / (ru * rw) - cs * w[d] / rw2);
        }
        if (b == ia) {
            f0 += fa[0];
            f1 += fa[1];
            f2 += fa[2];
        } else if (b == ic) {
            f0 += fc[0];
            f1 += fc[1];
            f2 += fc[2];
        } else {
            f0 -= fa[0] + fc[0];
            f1 -= fa[1] + fc[1];
            f2 -= fa[2] + fc[2];
        }
    }
    // lj/cut (LAMMPS pair_style lj/cut): E = 4 eps ((s/r)^12 - (s/r)^6), r < rc,
    // minimum image; 1-2/1-3/1-4 partners carry the special weight.  With
    // bins, the 27 bins around the bead's own are visited in a fixed order
    // (z, y, x offsets -1, 0, 1) and each bin's beads in ascending index; a
    // bin edge is at least the largest cutoff, so no pair within it is
    // missed.  Without bins every other bead is visited in ascending index.
    if (R.lj_nt > 0) {
        const int ti = R.lj_type[b];
        const int s0 = R.spec_ptr[b], s1 = R.spec_ptr[b + 1];
        const double xb = R.x[3 * b], yb = R.x[3 * b + 1], zb = R.x[3 * b + 2];
        if (R.ljb_n > 0 && R.ljb_of[b] >= 0) {
            const int nbx = R.ljb_nb[0], nby = R.ljb_nb[1], nbz = R.ljb_nb[2];
            const int own = R.ljb_of[b];
            const int ix = own % nbx, iy = (own / nbx) % nby, iz = own / (nbx * nby);
            for (int dz = -1; dz <= 1; dz++) {
                const int cz = (iz + dz + nbz) % nbz;
                for (int dy = -1; dy <= 1; dy++) {
                    const int cy = (iy + dy + nby) % nby;
                    for (int dxb = -1; dxb <= 1; dxb++) {
                        const int cx = (ix + dxb + nbx) % nbx;
                        const int bin = (cz * nby + cy) * nbx + cx;
                        for (int q = R.ljb_start[bin]; q < R.ljb_start[bin + 1]; q++) {
                            const int j = R.ljb_list[q];
                            if (j != b) lj_pair(R, b, j, ti, s0, s1, xb, yb, zb, f0, f1, f2);
                        }
                    }
                }
            }
        } else if (R.ljb_n == 0) {
            for (int j = 0; j < R.M; j++)
                if (j != b) lj_pair(R, b, j, ti, s0, s1, xb, yb, zb, f0, f1, f2);
        }
    }
    R.fc[3 * b + 0] = f0 + R.fext[3 * b + 0];
    R.fc[3 * b + 1] = f1 + R.fext[3 * b + 1];
    R.fc[3 * b + 2] = f2 + R.fext[3 * b + 2];
}


template <typename T>
__global__ void k_fill(T *p, long long n, T v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ===========================================================================
// Host-side launchers.
// ===========================================================================
static inline unsigned nblk(long long n, int per) { return (unsigned)((n + per - 1) / per); }

template <typename T>
static void launch_fft_generic(const StepLaunch &L, int stage, hipStream_t st);

template <typename T>
static void launch_stage_t(const StepLaunch &L, int stage, hipStream_t st) {
    const RepDev *reps = L.reps;
    const int nrep = L.nrep;
    switch (stage) {
        case kStageNoise:
            launch_noise(L, st);
            break;
        case kStageBead:
            if (L.tot_bead > 0)
                hipLaunchKernelGGL(k_bead_update<T>, dim3(nblk(L.tot_bead, kBeadsPerBlock)), dim3(kBeadThreads), 0,
                                   st, reps, nrep, L.st_bead, L.tot_bead);
            break;
        case kStageSpread: {
            // the plane path's thermal forcing (FP32-bulk only)
            const int nnoise =
                L.plane && L.tot_modes > 0 ? (L.noise_blocks > 0 ? L.noise_blocks : spread_noise_default(L.tot_modes)) : 0;
            // device-scope stores only with the plane path (FP32-bulk)
            auto *kern = (sizeof(T) == 4 && L.plane && L.wt) ? k_spread<T, sizeof(T) == 4> : k_spread<T, false>;
            hipLaunchKernelGGL(kern, dim3((unsigned)(nnoise + L.tot_split + L.tot_tile)),
                               dim3(kSpreadThreads), 0, st, reps, nrep, L.st_modes, L.tot_modes, nnoise, L.st_split,
                               L.tot_split, L.st_tile, L.tot_tile);
            if (!L.fast)
                hipLaunchKernelGGL(k_ustar<T>, dim3(nblk(L.tot_pts, 256)), dim3(256), 0, st, reps, nrep, L.st_pts,
                                   L.tot_pts);
            break;
        }
        case kStageFftX:
        case kStageFftY:
        case kStageFftZ:
        case kStageFftYi:
        case kStageFftXi:
            if (L.fast) launch_fft_fast(L, stage, st);
            else launch_fft_generic<T>(L, stage, st);
            break;
        case kStageBonds:
            // on the compile-time-length path the Xi pass advances the step counter
            if (!L.aux_bonds) launch_bond_forces(L, L.fast ? 0 : 1, st);
            break;
        default: break;
    }
}

template <typename T>
static void launch_fft_generic(const StepLaunch &L, int stage, hipStream_t st) {
    const RepDev *reps = L.reps;
    const int nrep = L.nrep;
    switch (stage) {
        case kStageFftX:
            hipLaunchKernelGGL(k_fft_x_r2c<T>, dim3((unsigned)L.tot_fx), dim3(256), L.lds_fx, st, reps, nrep,
                               L.st_fx, L.tot_fx, L.twmax);
            break;
        case kStageFftY:
        case kStageFftYi:
            hipLaunchKernelGGL(k_fft_y<T>, dim3((unsigned)L.tot_fy), dim3(256), L.lds_fy, st, reps, nrep,
                               L.st_fy, L.tot_fy, stage == kStageFftY ? -1 : +1, L.twmax);
            break;
        case kStageFftZ:
            hipLaunchKernelGGL(k_fft_z_proj<T>, dim3((unsigned)L.tot_fz), dim3(256), L.lds_fz, st, reps, nrep,
                               L.st_fz, L.tot_fz, L.twmax);
            break;
        case kStageFftXi:
            hipLaunchKernelGGL(k_fft_x_c2r<T>, dim3((unsigned)L.tot_fx), dim3(256), L.lds_fx, st, reps, nrep,
                               L.st_fx, L.tot_fx, L.twmax);
            break;
        default: break;
    }
}

void launch_stage(const StepLaunch &L, int stage, hipStream_t st) {
    if (L.fp64) launch_stage_t<double>(L, stage, st);
    else launch_stage_t<float>(L, stage, st);
}

// One step.  On the five-pass compile-time-length path the thermal forcing
// and the conservative forces at x^{n+1} run on a forked stream once the bead
// update is done, concurrently with the spread and the first two FFT passes;
// the branch joins before the Z pass, which consumes the forcing.  The bead
// update runs alone: it is latency-bound and slows down beside the forcing.
// The plane path runs every kernel in one chain on one queue: its spread
// launch draws the thermal forcing, and its XYi launch (without angle or
// lj/cut terms) evaluates the bond forces, in workgroups of their own.  On the compile-time-length paths the step counter is
// advanced by the last pass (Xi or XYi), after every kernel that reads it;
// otherwise by the bond kernel at the end of the step.
void launch_step(const StepLaunch &L, hipStream_t st) {
    const bool fork = L.tot_modes > 0 && !L.plane;
    launch_stage(L, kStageBead, st);
    if (!fork) {
        for (int s = kStageSpread; s < kNumStages; s++) launch_stage(L, s, st);
        return;
    }
    // The spread is captured before the forked branch's kernels: the graph
    // keeps the first child of a node on its parent's queue, and a
    // dependency between queues costs several microseconds, so the main
    // chain stays on one queue.
    (void)hipEventRecord(L.ev_fork, st);
    launch_stage(L, kStageSpread, st);
    (void)hipStreamWaitEvent(L.st2, L.ev_fork, 0);
    launch_noise(L, L.st2);
    launch_stage(L, kStageBonds, L.st2);
    (void)hipEventRecord(L.ev_join, L.st2);
    for (int s = kStageFftX; s < kNumStages; s++) {
        if (s == kStageBonds) continue;
        if (s == kStageFftZ) (void)hipStreamWaitEvent(st, L.ev_join, 0);
        launch_stage(L, s, st);
    }
}

void launch_bond_forces(const StepLaunch &L, int advance, hipStream_t st) {
    if (L.lj_bins && L.tot_bead > 0) {
        const unsigned nb = nblk(L.tot_bead, 256);
        hipLaunchKernelGGL(k_lj_bin, dim3(nb), dim3(256), 0, st, L.reps, L.nrep, L.st_bead, L.tot_bead);
        const unsigned nch = (unsigned)lj_bin_chunks(L.lj_maxbins);
        hipLaunchKernelGGL(k_lj_chunk_sums, dim3(nch, L.nrep), dim3(256), 0, st, L.reps, L.nrep);
        hipLaunchKernelGGL(k_lj_chunk_scan, dim3(L.nrep), dim3(256), 0, st, L.reps, L.nrep);
        hipLaunchKernelGGL(k_lj_chunk_apply, dim3(nch, L.nrep), dim3(256), 0, st, L.reps, L.nrep);
        hipLaunchKernelGGL(k_lj_fill, dim3(nb), dim3(256), 0, st, L.reps, L.nrep, L.st_bead, L.tot_bead);
        hipLaunchKernelGGL(k_lj_sort, dim3(nblk(L.lj_maxbins, 256), L.nrep), dim3(256), 0, st, L.reps, L.nrep);
    }
    hipLaunchKernelGGL(k_bond_forces, dim3(L.tot_bead > 0 ? nblk(L.tot_bead, 256) : 1), dim3(256), 0, st,
                       L.reps, L.nrep, L.st_bead, L.tot_bead, advance);
}

void launch_fill_f64(double *p, long long n, double v, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_fill<double>, dim3(nblk(n, 256)), dim3(256), 0, st, p, n, v);
}
void launch_fill_f32(float *p, long long n, float v, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_fill<float>, dim3(nblk(n, 256)), dim3(256), 0, st, p, n, v);
}


}  // namespace selm
