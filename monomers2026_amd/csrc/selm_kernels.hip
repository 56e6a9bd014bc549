// selm_kernels.hip -- CDNA4 kernels of one SELM time step.
//
// Every kernel takes the replica descriptor array and a work table, so one
// launch advances any number of heterogeneous replicas (a single system is a
// batch of one).  Per step, in order:
//
//   k_bead_update     16 lanes per bead (4 beads per wave): 4x4x4 Peskin
//                     gather of u^n for the three staggered components, drag +
//                     thermal kick, x/v/f update, the bead's spread record,
//                     insertion into the buckets of the tiles its stencils touch
//   k_spread          workgroup per tile: bucket sorted by bead index in LDS;
//                     wave d accumulates component d (lanes = stencil points,
//                     distinct cells, LDS adds in bead order) and the tile
//                     goes to the sparse pool; untouched tiles are marked empty
//   [k_ustar]         u* = u + (mu dt L7 u + f) / rho (runtime-length FFT path
//                     only; the compile-time-length X pass computes u* itself)
//   X, Y, Z+projection, Y, X passes of the 3-D FFT (selm_fft_kernels.hip for
//                     the compile-time lengths, below for the others)
//   [k_lj_bin, k_lj_scan, k_lj_fill, k_lj_sort]  lj/cut neighbour bins
//                     (only when a replica has them)
//   k_bond_forces     thread per bead: harmonic bonds and angles, lj/cut pairs
//                     and the external force at x^{n+1}; advances each
//                     replica's step counter
#include "selm_fluid.hpp"
#include "selm_kernels.hpp"

namespace selm {

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

// Peskin factor with the branch decided on the FP64 distance (so the r >= 2
// quirk and the continuity at r -> 2^- survive) and the arithmetic in T.
template <typename T>
__device__ __forceinline__ T phi_t(double r, int f4) {
    if constexpr (sizeof(T) == 8) {
        return peskin_phi(r, f4);
    } else {
        const float q = (float)r;
        if (r < 1.0) return (3.0f - 2.0f * q + sqrtf(1.0f + 4.0f * q * (1.0f - q))) * 0.125f;
        if (r < 2.0) return (5.0f - 2.0f * q - sqrtf(fmaxf(-7.0f + 4.0f * q * (3.0f - q), 0.0f))) * 0.125f;
        return f4 ? 1.0f : 0.0f;
    }
}

// ===========================================================================
// Beads: gather, drag, thermal kick, update.  Semantics of
// cuda/SELM_Kernel.cu:302-373 and mpi/SELM_Lagrangian_LAMMPS.cpp:115-180.
// ===========================================================================
constexpr int kBeadThreads = 256;
constexpr int kBeadsPerBlock = kBeadThreads / 16;

// 5 waves per SIMD: 5 workgroups per CU, so 20k beads (1250 workgroups) are
// resident at once on 256 CUs instead of running in two rounds.
template <typename T>
__global__ __launch_bounds__(kBeadThreads) __attribute__((amdgpu_waves_per_eu(5))) void k_bead_update(
    const RepDev *__restrict__ reps, int nrep,
                                                              const long long *__restrict__ start,
                                                              long long total) {
    const int sl = threadIdx.x & 15;  // sub-lane: (li, lj) column of the stencil
    const long long g0 = (long long)blockIdx.x * kBeadsPerBlock + (threadIdx.x >> 4);
    const bool active = g0 < total;
    const long long g = active ? g0 : total - 1;  // inactive groups mirror the last bead, store nothing
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const int b = (int)(g - start[r]);
    const int nx = R.n[0], ny = R.n[1], nz = R.n[2];
    const long long N = R.N;
    const int f4 = R.f4;
    SELM_GLOBAL SpreadRec<T> *rec = gbl((SpreadRec<T> *)R.brec) + b;
    const SELM_GLOBAL double *gx = gbl((const double *)R.x);

    double s[3];
#pragma unroll
    for (int d = 0; d < 3; d++) s[d] = (gx[3 * b + d] - R.lo) / R.dx;

    const int li = sl & 3, lj = sl >> 2;
    const SELM_GLOBAL T *u = gbl((const T *)R.u);
    double gam[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        // component d is sampled at its face centre: the other two axes shift by -1/2
        const double rx = s[0] - (d != 0 ? 0.5 : 0.0);
        const double ry = s[1] - (d != 1 ? 0.5 : 0.0);
        const double rz = s[2] - (d != 2 ? 0.5 : 0.0);
        const int p0 = (int)(floor(rx) - 1.0), p1 = (int)(floor(ry) - 1.0), p2 = (int)(floor(rz) - 1.0);
        const T fx = phi_t<T>(fabs(rx - (p0 + li)), f4);
        const T fy = phi_t<T>(fabs(ry - (p1 + lj)), f4);
        const long long col = (long long)wrap_near(p1 + lj, ny) * nx + wrap_near(p0 + li, nx);
        const SELM_GLOBAL T *ud = u + d * N;
        T fz[4], uv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            fz[k] = phi_t<T>(fabs(rz - (p2 + k)), f4);
            uv[k] = ud[(long long)wrap_near(p2 + k, nz) * ny * nx + col];
        }
        T acc = (T)0;
#pragma unroll
        for (int k = 0; k < 4; k++) acc += uv[k] * ((((T)1 * fz[k]) * fy) * fx);
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) acc += __shfl_xor(acc, m, 64);
        gam[d] = (double)acc;
        if (active) {
            if (lj == 0) rec->w[d][0][li] = fx;
            if (li == 0) rec->w[d][1][lj] = fy;
            if (sl < 4) rec->w[d][2][sl] = sl == 0 ? fz[0] : (sl == 1 ? fz[1] : (sl == 2 ? fz[2] : fz[3]));
            if (sl == 0) {
                rec->p[d][0] = wrap_near(p0, nx);
                rec->p[d][1] = wrap_near(p1, ny);
                rec->p[d][2] = wrap_near(p2, nz);
            }
        }
    }
    if (!active) return;

    const uint32_t step = (uint32_t)(R.ts[0] + 1);
    {
        // insert into the bucket of every tile the stencils touch; the bead's
        // 16 lanes take one (tile) insertion each
        int tx[6], ty[6], tz[6];
        const int cx = axis_tiles(s[0], nx, R.te[0], tx);
        const int cy = axis_tiles(s[1], ny, R.te[1], ty);
        const int cz = axis_tiles(s[2], nz, R.te[2], tz);
        const int ncomb = cx * cy * cz;
        const float rcx = 1.0f / (float)cx, rcy = 1.0f / (float)cy;
        for (int q = sl; q < ncomb; q += 16) {
            const int qa = div_small(q, cx, rcx), c = div_small(qa, cy, rcy);
            const int a = q - qa * cx, qb = qa - c * cy;
            const int t = (tz[c] * R.tn[1] + ty[qb]) * R.tn[0] + tx[a];
            const int slot = atomicAdd(&R.bcnt[t], 1);
            if (slot < kBucketCap) {
                gbl(R.bucket)[t * kBucketCap + slot] = b;
            } else {
                const int o = atomicAdd(&R.ctr[step & 1], 1);
                gbl(R.ovf)[2 * o] = t;
                gbl(R.ovf)[2 * o + 1] = b;
            }
        }
    }
    if (sl < 3) {
        const int d = sl;
        const double gd = d == 0 ? gam[0] : (d == 1 ? gam[1] : gam[2]);
        const double sd = d == 0 ? s[0] : (d == 1 ? s[1] : s[2]);
        const int i = 3 * b + d;
        const double dt = R.dt;
        const double vi = gbl(R.v)[i];
        double tf = -R.ups * (vi - gd) * dt;
        if (R.noise) {
            double z0, z1;
            normal_pair_f64(R.seed, kStreamBead, step, (uint32_t)i, 0u, z0, z1);
            tf += z0 * R.bead_coef;
        }
        gbl(R.tf)[i] = tf;
        gbl(R.sg)[i] = sd;
        rec->f[d] = (T)(-tf);
        const double ft = tf + dt * gbl(R.fc)[i];
        gbl(R.ftot)[i] = ft / dt;
        double tmp = vi * dt;
        tmp += gx[i];
        if (tmp > R.hi) tmp -= R.L;
        else if (tmp < R.lo) tmp += R.L;
        gbl(R.x)[i] = tmp;
        gbl(R.v)[i] = vi + ft / gbl(R.m)[b];
    }
}

// ===========================================================================
// Spread (Lambda), workgroup per tile.  Semantics of cuda/SELM_Kernel.cu:424-470:
// each bead adds -t_f * w to the 4x4x4 stencil of each staggered component,
// w = ((1*phi_z)*phi_y)*phi_x.  Each cell's sum runs over the tile's beads in
// ascending bead index, so the result does not depend on scheduling or on the
// other replicas of a batch.
// ===========================================================================
// The cell sums are FP64 in both precisions: each product -t_f * w is formed
// in T and added to an FP64 accumulator in LDS, and the tile is rounded to T
// once when it is written.
constexpr int kTileCells = 512;  // tile edge <= 8
constexpr int kSpreadChunk = 32;
constexpr int kSpreadBatch = 8;  // beads whose (cell, value) a lane holds before adding
constexpr int kSpreadThreads = 256;
constexpr int kListCap = 1536;  // with FP32 records the workgroup's LDS stays under 32 KB: 5 per CU

// Records of chunk `base` (nq >= 1 beads of the sorted list) into LDS, by NT
// threads (lt = 0 .. NT-1) in 16-byte pieces.  Every thread issues all of its
// loads (indices clamped into the chunk) before its first store, so a chunk
// costs about one global-load latency.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <typename T, int NT>
__device__ __forceinline__ void spread_stage(SpreadRec<T> *dst, const SpreadRec<T> *__restrict__ brec,
                                             const int *sorted, int base, int nq, int lt) {
    static_assert(sizeof(SpreadRec<T>) % 16 == 0, "records are whole 16-byte pieces");
    constexpr int V = sizeof(SpreadRec<T>) / 16;
    constexpr int PER = (kSpreadChunk * V + NT - 1) / NT;
    const int tot = nq * V;
    u32x4 v[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int idx = min(lt + q * NT, tot - 1);
        const int k = idx / V;
        v[q] = gbl((const u32x4 *)&brec[sorted[base + k]])[idx - k * V];
    }
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int idx = lt + q * NT;
        if (idx < tot) {
            const int k = idx / V;
            ((u32x4 *)&dst[k])[idx - k * V] = v[q];
        }
    }
}

// w[d][0][li] *= f[d] for the nq records of a staged chunk, by NT threads
// (lt = 0 .. NT-1): the adding waves then take the force with the x factor.
template <typename T, int NT>
__device__ __forceinline__ void spread_fold_force(SpreadRec<T> *Q, int nq, int lt) {
    for (int e = lt; e < nq * 12; e += NT) {
        const int k = e / 12, r = e - 12 * k, d = r >> 2, li = r & 3;
        Q[k].w[d][0][li] *= Q[k].f[d];
    }
}

struct SpreadTile {
    int nx, ny, nz;  // grid
    int ox, oy, oz;  // tile origin
    int ex, ey, ez;  // tile extent (ragged at the high end)
    int te0, te1;    // tile edge (cell index strides)
};

// v of lane k (k uniform), as a scalar
template <typename T>
__device__ __forceinline__ T readlane_t(T v, int k) {
    if constexpr (sizeof(T) == 4) {
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
    } else {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), k);
        const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    }
}

// Wave d adds component d of the chunk's nq beads, in list order, lane =
// stencil point (li, lj, lk).  A lane's cell on an axis is (p + l) mod n - o;
// it belongs to the tile if that lies in [0, e).  This is the general form,
// with the wrap taken per lane; spread_add_chunk_fast below covers grids
// whose edges are all >= tile edge + 3.
template <typename T>
__device__ __forceinline__ void spread_add_chunk(double *acc, const SpreadRec<T> *Q, int nq, int d, int lane,
                                                 const SpreadTile &g) {
    const int li = lane & 3, lj = (lane >> 2) & 3, lk = lane >> 4;
    const int lofs = li + g.te0 * (lj + g.te1 * lk);
    // lane m (< 32) holds bead m's stencil base and force for this component;
    // the loop takes them per bead with readlane
    static_assert(kSpreadChunk <= 32, "one bead per lane of the half-wave");
    const SpreadRec<T> &M = Q[min(lane & 31, nq - 1)];
    const int px = M.p[d][0], py = M.p[d][1], pz = M.p[d][2];
    for (int k0 = 0; k0 < nq; k0 += kSpreadBatch) {
        int cell[kSpreadBatch];
        double val[kSpreadBatch];
#pragma unroll
        for (int j = 0; j < kSpreadBatch; j++) {
            const int k = k0 + j;  // < kSpreadChunk; beads past nq are masked below
            const SpreadRec<T> &B = Q[min(k, nq - 1)];
            // the x factor carries the force (spread_fold_force)
            const T w = (B.w[d][2][lk] * B.w[d][1][lj]) * B.w[d][0][li];
            int cx = __builtin_amdgcn_readlane(px, k) + li;
            cx -= cx >= g.nx ? g.nx : 0;
            int cy = __builtin_amdgcn_readlane(py, k) + lj;
            cy -= cy >= g.ny ? g.ny : 0;
            int cz = __builtin_amdgcn_readlane(pz, k) + lk;
            cz -= cz >= g.nz ? g.nz : 0;
            cx -= g.ox;
            cy -= g.oy;
            cz -= g.oz;
            const bool in =
                (unsigned)cx < (unsigned)g.ex && (unsigned)cy < (unsigned)g.ey && (unsigned)cz < (unsigned)g.ez;
            const int c = d * kTileCells + (cz * g.te1 + cy) * g.te0 + cx;
            cell[j] = in && k < nq ? c : -1;
            val[j] = (double)w;
        }
        // in bead order: a cell met by several beads of the batch is added
        // to in the order of the sorted list.  LDS atomic adds (ds_add_f64)
        // from one wave are applied in program order, and the lanes of one
        // add meet distinct cells, so each cell's sum is the sequential one;
        // unlike a read-modify-write, no add waits for the previous one.
#pragma unroll
        for (int j = 0; j < kSpreadBatch; j++)
            if (cell[j] >= 0)
                __hip_atomic_fetch_add(&acc[cell[j]], val[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

// spread_add_chunk for grids whose edges are all >= tile edge + 3.  There
// the offset q = p - o, taken as q - n when q >= e, is the same for the whole
// wave (a scalar), and each lane only adds l.  This gives the cell of the
// general form (so n > 8 >= 5):
// - p in [o, o + e): a wrapped lane (p + l >= n) cannot come back into the
//   tile, since p + l - n < o + e + 3 - n <= o;
// - p >= o + e: the stencil meets the tile only through the wrap, where the
//   cell is p + l - n - o = (q - n) + l;
// - p < o: a wrapped cell (<= 2) would need o <= 2, while wrapping needs
//   p >= n - 3, so o > n - 3 > 2.
// The per-bead work is moved out of the bead loop: lane m first forms, for
// bead m of the chunk, the scalar offset of its first stencil cell in the
// accumulators and the 64-bit mask of its stencil points that lie in the
// tile (bit 16 lk + 4 lj + li).  The loop then takes the offset, the force
// and the mask of bead k with readlanes.
__device__ __forceinline__ unsigned axis_mask4(int q, int e) {
    unsigned m = 0;
#pragma unroll
    for (int l = 0; l < 4; l++) m |= ((unsigned)(q + l) < (unsigned)e ? 1u : 0u) << l;
    return m;
}

// v = bit `lane` of the wave-uniform 64-bit mask ? a : b, as one v_cndmask
// with the mask in an SGPR pair (a plain C++ select of a mask bit compiles
// to a shift, an and, a compare and the select).
__device__ __forceinline__ int lane_select(uint64_t mask, int a, int b) {
    int v;
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(v) : "v"(b), "v"(a), "s"(mask));
    return v;
}

template <typename T>
__device__ __forceinline__ void spread_add_chunk_fast(double *acc, const SpreadRec<T> *Q, int nq, int d, int lane,
                                                      const SpreadTile &g) {
    const int li = lane & 3, lj = (lane >> 2) & 3, lk = lane >> 4;
    const int lofs = 8 * (li + g.te0 * (lj + g.te1 * lk));  // bytes
    static_assert(kSpreadChunk <= 32, "one bead per lane of the half-wave");
    const SpreadRec<T> &M = Q[min(lane & 31, nq - 1)];
    int qx = M.p[d][0] - g.ox, qy = M.p[d][1] - g.oy, qz = M.p[d][2] - g.oz;
    qx -= qx >= g.ex ? g.nx : 0;
    qy -= qy >= g.ey ? g.ny : 0;
    qz -= qz >= g.ez ? g.nz : 0;
    const int cm = 8 * (d * kTileCells + (qz * g.te1 + qy) * g.te0 + qx);  // bytes
    const unsigned mx = axis_mask4(qx, g.ex), my = axis_mask4(qy, g.ey), mz = axis_mask4(qz, g.ez);
    uint32_t mlo = 0, mhi = 0;
#pragma unroll
    for (int z = 0; z < 4; z++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const uint32_t row = ((mz >> z) & (my >> y) & 1u) ? mx : 0u;
            if (z < 2) mlo |= row << (16 * z + 4 * y);
            else mhi |= row << (16 * (z - 2) + 4 * y);
        }
    const int sink = 8 * (3 * kTileCells + lane);  // where a lane outside the tile adds (never read)
    char *accb = (char *)acc;
    // The weights of batch b+1 are read before the adds of batch b are
    // issued: a wave's LDS operations complete in order, so a read issued
    // after the adds would wait for them.  Records past nq are read (they
    // lie in the chunk buffer) but their masks are cleared.  The loop is
    // fully unrolled, so record offsets and readlane lanes are constants.
    T wz[kSpreadBatch], wy[kSpreadBatch], wx[kSpreadBatch];
#pragma unroll
    for (int j = 0; j < kSpreadBatch; j++) {
        wz[j] = Q[j].w[d][2][lk];
        wy[j] = Q[j].w[d][1][lj];
        wx[j] = Q[j].w[d][0][li];  // the x factor carries the force (spread_fold_force)
    }
#pragma unroll
    for (int k0 = 0; k0 < kSpreadChunk; k0 += kSpreadBatch) {
        if (k0 >= nq) break;
        int cell[kSpreadBatch];
        double val[kSpreadBatch];
#pragma unroll
        for (int j = 0; j < kSpreadBatch; j++) {
            const int k = k0 + j;
            const T w = (wz[j] * wy[j]) * wx[j];
            const int c = __builtin_amdgcn_readlane(cm, k);
            const uint64_t m = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)mhi, k) << 32) |
                               (uint32_t)__builtin_amdgcn_readlane((int)mlo, k);
            cell[j] = lane_select(k < nq ? m : 0, c + lofs, sink);
            val[j] = (double)w;
        }
        if (k0 + kSpreadBatch < kSpreadChunk) {
#pragma unroll
            for (int j = 0; j < kSpreadBatch; j++) {
                const SpreadRec<T> &B = Q[k0 + kSpreadBatch + j];
                wz[j] = B.w[d][2][lk];
                wy[j] = B.w[d][1][lj];
                wx[j] = B.w[d][0][li];
            }
        }
        // in bead order: a cell met by several beads of the batch is added
        // to in the order of the sorted list (LDS atomic adds from one wave
        // are applied in program order; the lanes of one add meet distinct cells)
#pragma unroll
        for (int j = 0; j < kSpreadBatch; j++)
            __hip_atomic_fetch_add((double *)(accb + cell[j]), val[j], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

template <typename T>
__global__ __launch_bounds__(kSpreadThreads) void k_spread(const RepDev *__restrict__ reps, int nrep,
                                                           const long long *__restrict__ start,
                                                           long long total) {
    static_assert(3 * kTileCells * sizeof(double) >= kListCap * sizeof(int), "list shares the accumulators");
    __shared__ int sorted[kListCap];
    __shared__ double acc[3 * kTileCells + 64];  // the unsorted list first, then the cell sums (+ per-lane sinks)
    __shared__ int nlist;
    __shared__ __align__(16) SpreadRec<T> rec[2][kSpreadChunk];  // chunk c in rec[c & 1]
    int *list = (int *)acc;

    const long long g = blockIdx.x;
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const int t = (int)(g - start[r]);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const uint32_t step = (uint32_t)(R.ts[0] + 1);
    const int par = step & 1;
    if (t == 0 && tid == 0) R.ctr[par ^ 1] = 0;  // next step's overflow count (last used a step ago)
    const int cnt = R.bcnt[t];
    if (cnt == 0) return;  // no bead touches this tile: its stamp stays stale
    // Issue priority over the thermal forcing that runs beside the spread on
    // the forked stream: the step waits for the spread's fullest tiles, not
    // for the forcing, which has the first FFT pass to finish in.
    __builtin_amdgcn_s_setprio(3);
    const int nx = R.n[0], ny = R.n[1], nz = R.n[2];
    const int tx = t % R.tn[0], ty = (t / R.tn[0]) % R.tn[1], tz = t / (R.tn[0] * R.tn[1]);
    const int ox = tx * R.te[0], oy = ty * R.te[1], oz = tz * R.te[2];
    const int ex = min(R.te[0], nx - ox), ey = min(R.te[1], ny - oy), ez = min(R.te[2], nz - oz);

    const int nb = min(cnt, kBucketCap);
    if (tid == 0) nlist = nb;
    for (int i = tid; i < nb; i += kSpreadThreads) list[i] = gbl(R.bucket)[t * kBucketCap + i];
    __syncthreads();
    if (cnt > kBucketCap) {
        const int novf = R.ctr[par];
        for (int o = tid; o < novf; o += kSpreadThreads)
            if (R.ovf[2 * o] == t) {
                const int p = atomicAdd(&nlist, 1);
                if (p < kListCap) list[p] = R.ovf[2 * o + 1];
            }
        __syncthreads();
    }
    const int n = min(nlist, kListCap);
    if (tid == 0) {
        if (nlist > kListCap) atomicOr(&R.ctr[4], kStatusListOverflow);
        R.bcnt[t] = 0;  // ready for the next step's insertions
    }
    // order the tile's beads by index (indices are unique within a tile)
    for (int e = tid; e < n; e += kSpreadThreads) {
        const int bi = list[e];
        int rank = 0;
        for (int q = 0; q < n; q++) rank += (list[q] < bi) ? 1 : 0;
        sorted[rank] = bi;
    }
    __syncthreads();  // sorted is complete and the list is read: its space becomes the accumulators
    const SpreadRec<T> *brec = (const SpreadRec<T> *)R.brec;
    spread_stage<T, kSpreadThreads>(rec[0], brec, sorted, 0, min(kSpreadChunk, n), tid);
    for (int c = tid; c < 3 * kTileCells; c += kSpreadThreads) acc[c] = 0.0;
    __syncthreads();
    spread_fold_force<T, kSpreadThreads>(rec[0], min(kSpreadChunk, n), tid);
    __syncthreads();

    // Waves 0..2 add chunk c (wave d: component d, lane: stencil point) while
    // wave 3 stages chunk c+1 into the other buffer.
    const int te0 = R.te[0], te1 = R.te[1];
    // no stencil wraps back into the tile it starts in (see spread_add_chunk)
    const bool fast = nx >= te0 + 3 && ny >= te1 + 3 && nz >= R.te[2] + 3;
    for (int c = 0, base = 0; base < n; c++, base += kSpreadChunk) {
        const int nq = min(kSpreadChunk, n - base);
        if (wv == 3) {
            const int nb2 = base + kSpreadChunk;
            if (nb2 < n) {
                const int nq2 = min(kSpreadChunk, n - nb2);
                spread_stage<T, 64>(rec[(c + 1) & 1], brec, sorted, nb2, nq2, lane);
                // the fold reads what other lanes of this wave just stored:
                // LDS operations of one wave complete in order
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                spread_fold_force<T, 64>(rec[(c + 1) & 1], nq2, lane);
            }
        } else {
            const SpreadTile tl{nx, ny, nz, ox, oy, oz, ex, ey, ez, te0, te1};
            if (fast) spread_add_chunk_fast<T>(acc, rec[c & 1], nq, wv, lane, tl);
            else spread_add_chunk<T>(acc, rec[c & 1], nq, wv, lane, tl);
        }
        __syncthreads();  // chunk c consumed, chunk c+1 staged
    }
    const int tcells = te0 * te1 * R.te[2];
    SELM_GLOBAL T *pool = gbl((T *)R.pool) + (long long)t * 3 * tcells;
    for (int c = tid; c < 3 * tcells; c += kSpreadThreads) {
        const int d = c / tcells, cell = d * kTileCells + (c - d * tcells);
        pool[c] = (T)acc[cell];
    }
    if (tid == 0) R.tstamp[t] = (int)step;
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
// k_lj_bin counts the beads of each bin, k_lj_scan turns the counts into list
// offsets and fill cursors, k_lj_fill files each bead, and k_lj_sort orders
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
    const int ix = lj_bin_axis(R.x[3 * b], R.ljb_lo[0], R.Lb[0], R.ljb_nb[0]);
    const int iy = lj_bin_axis(R.x[3 * b + 1], R.ljb_lo[1], R.Lb[1], R.ljb_nb[1]);
    const int iz = lj_bin_axis(R.x[3 * b + 2], R.ljb_lo[2], R.Lb[2], R.ljb_nb[2]);
    const int bin = (iz * R.ljb_nb[1] + iy) * R.ljb_nb[0] + ix;
    R.ljb_of[b] = bin;
    atomicAdd(&R.ljb_cnt[bin], 1);
}

// one workgroup per replica: exclusive scan of the counts (each thread a
// contiguous range of bins, then a scan of the 256 range sums)
__global__ __launch_bounds__(256) void k_lj_scan(const RepDev *__restrict__ reps, int nrep) {
    const RepDev &R = reps[blockIdx.x];
    const int n = R.ljb_n;
    if (n == 0) return;
    __shared__ int part[256];
    const int t = threadIdx.x, per = (n + 255) / 256;
    const int a = min(n, t * per), e = min(n, a + per);
    int sum = 0;
    for (int i = a; i < e; i++) sum += R.ljb_cnt[i];
    part[t] = sum;
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int i = 0; i < 256; i++) {
            const int v = part[i];
            part[i] = acc;
            acc += v;
        }
        R.ljb_start[n] = acc;
    }
    __syncthreads();
    int off = part[t];
    for (int i = a; i < e; i++) {
        const int c = R.ljb_cnt[i];
        R.ljb_start[i] = off;
        R.ljb_cnt[i] = off;  // fill cursor
        off += c;
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
    R.ljb_list[atomicAdd(&R.ljb_cnt[R.ljb_of[b]], 1)] = b;
}

// one workgroup per replica, a thread per bin: insertion sort of the bin's list
__global__ __launch_bounds__(256) void k_lj_sort(const RepDev *__restrict__ reps, int nrep) {
    const RepDev &R = reps[blockIdx.x];
    for (int i = threadIdx.x; i < R.ljb_n; i += blockDim.x) {
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
    for (int q = R.adj_ptr[b]; q < R.adj_ptr[b + 1]; q++) {
        const int i1 = R.adj_i1[q], i2 = R.adj_i2[q];
        double del[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double dd = R.x[3 * i1 + d] - R.x[3 * i2 + d];
            if (fabs(dd) > 0.5 * R.Lb[d]) dd += (dd < 0.0) ? R.Lb[d] : -R.Lb[d];
            del[d] = dd;
        }
        double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
        double rr = sqrt(rsq);
        double dr = rr - R.adj_r0[q];
        double rk = R.adj_K[q] * dr;
        double fbond = (rr > 0.0) ? -2.0 * rk / rr : 0.0;
        if (b == i1) {
            f0 += del[0] * fbond;
            f1 += del[1] * fbond;
            f2 += del[2] * fbond;
        } else {
            f0 -= del[0] * fbond;
            f1 -= del[1] * fbond;
            f2 -= del[2] * fbond;
        }
    }
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
        if (R.ljb_n > 0) {
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
        } else {
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
        case kStageSpread:
            hipLaunchKernelGGL(k_spread<T>, dim3((unsigned)L.tot_tile), dim3(kSpreadThreads), 0, st, reps, nrep,
                               L.st_tile, L.tot_tile);
            if (!L.fast)
                hipLaunchKernelGGL(k_ustar<T>, dim3(nblk(L.tot_pts, 256)), dim3(256), 0, st, reps, nrep, L.st_pts,
                                   L.tot_pts);
            break;
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
            launch_bond_forces(L, L.fast ? 0 : 1, st);
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

// One step.  On the compile-time-length path the thermal forcing and the
// conservative forces at x^{n+1} run on a forked stream once the bead update
// is done, concurrently with the spread and the first two FFT passes; the
// branch joins before the Z pass, which consumes the forcing.  The bead
// update runs alone: it is latency-bound and slows down beside the forcing.
// There the step counter is advanced by the last pass (Xi), after every
// kernel that reads it; otherwise by the bond kernel at the end of the step.
void launch_step(const StepLaunch &L, hipStream_t st) {
    const bool fork = L.tot_modes > 0;
    launch_stage(L, kStageBead, st);
    if (fork) {
        (void)hipEventRecord(L.ev_fork, st);
        (void)hipStreamWaitEvent(L.st2, L.ev_fork, 0);
        launch_noise(L, L.st2);
        launch_stage(L, kStageBonds, L.st2);
        (void)hipEventRecord(L.ev_join, L.st2);
    }
    for (int s = kStageSpread; s < kNumStages; s++) {
        if (fork && s == kStageBonds) continue;
        if (s == kStageFftZ && fork) (void)hipStreamWaitEvent(st, L.ev_join, 0);
        launch_stage(L, s, st);
    }
}

void launch_bond_forces(const StepLaunch &L, int advance, hipStream_t st) {
    if (L.lj_bins && L.tot_bead > 0) {
        const unsigned nb = nblk(L.tot_bead, 256);
        hipLaunchKernelGGL(k_lj_bin, dim3(nb), dim3(256), 0, st, L.reps, L.nrep, L.st_bead, L.tot_bead);
        hipLaunchKernelGGL(k_lj_scan, dim3(L.nrep), dim3(256), 0, st, L.reps, L.nrep);
        hipLaunchKernelGGL(k_lj_fill, dim3(nb), dim3(256), 0, st, L.reps, L.nrep, L.st_bead, L.tot_bead);
        hipLaunchKernelGGL(k_lj_sort, dim3(L.nrep), dim3(256), 0, st, L.reps, L.nrep);
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

size_t spread_rec_bytes(bool fp64) { return fp64 ? sizeof(SpreadRec<double>) : sizeof(SpreadRec<float>); }

}  // namespace selm
