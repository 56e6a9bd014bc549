// selm_kernels.hip -- CDNA4 kernels of one SELM time step.
//
// Every kernel takes the replica descriptor array and a work table, so one
// launch advances any number of heterogeneous replicas (a single system is a
// batch of one).  Per step, in order:
//
//   k_bead_update   wave per bead: 4x4x4 Peskin gather of u^n for the three
//                   staggered components (64 lanes = 64 stencil points, wave
//                   reduction), drag + thermal kick, x/v/f update, tile key
//   k_bin_scan      workgroup per replica: tile counts -> bucket starts
//   k_bin_scatter   thread per bead: bucket fill (arbitrary order)
//   k_bin_rank      thread per bead: bucket order := ascending bead index
//   k_spread        workgroup per tile: owner-computes segmented reduction
//                   over the beads of the 27 neighbour buckets (fixed order,
//                   no atomics) -> fsp
//   k_ustar         u* = u + (mu dt L7 u + fsp) / rho, in place over fsp
//   k_fft_x_r2c     rows along x, two real rows per complex transform
//   k_fft_y         columns along y
//   k_fft_z_proj    columns along z: forward DFT, thermal forcing in k-space,
//                   discrete Helmholtz projection, k=0 removal, 1/N, inverse
//   k_fft_y         columns along y (inverse)
//   k_fft_x_c2r     rows along x -> u^{n+1}
//   k_bond_forces   thread per bead: harmonic bonds + external force at
//                   x^{n+1}; advances each replica's step counter
#include "selm_kernels.hpp"

namespace selm {

// ===========================================================================
// Beads: gather, drag, thermal kick, update.  Semantics of
// cuda/SELM_Kernel.cu:302-373 and mpi/SELM_Lagrangian_LAMMPS.cpp:115-180.
// ===========================================================================
template <typename T>
__global__ __launch_bounds__(256) void k_bead_update(const RepDev *__restrict__ reps, int nrep,
                                                     const long long *__restrict__ start,
                                                     long long total) {
    const int lane = threadIdx.x & 63;
    const long long g = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (g >= total) return;  // uniform per wave
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const int b = (int)(g - start[r]);
    const int nx = R.n[0], ny = R.n[1], nz = R.n[2];
    const long long N = R.N;
    const int f4 = R.f4;

    double s[3];
#pragma unroll
    for (int d = 0; d < 3; d++) s[d] = (R.x[3 * b + d] - R.lo) / R.dx;

    const int li = lane & 3, lj = (lane >> 2) & 3, lk = lane >> 4;
    const T *u = (const T *)R.u;
    double gam[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        // component d is sampled at its face centre: the other two axes shift by -1/2
        double rx = s[0] - (d != 0 ? 0.5 : 0.0);
        double ry = s[1] - (d != 1 ? 0.5 : 0.0);
        double rz = s[2] - (d != 2 ? 0.5 : 0.0);
        int pi = (int)(floor(rx) - 1.0) + li;
        int pj = (int)(floor(ry) - 1.0) + lj;
        int pk = (int)(floor(rz) - 1.0) + lk;
        double wz = 1.0 * peskin_phi(fabs(rz - pk), f4);
        double wy = wz * peskin_phi(fabs(ry - pj), f4);
        double w = wy * peskin_phi(fabs(rx - pi), f4);
        long long idx = ((long long)wrap_idx(pk, nz) * ny + wrap_idx(pj, ny)) * nx + wrap_idx(pi, nx);
        T val = u[d * N + idx] * (T)w;
        gam[d] = (double)wave_sum<T>(val);
    }

    if (lane == 0) {
        // bin by the base cell of the unshifted coordinate; every stencil of
        // this bead lies within [c-2, c+2] on each axis
        int c0 = wrap_idx((int)floor(s[0]), nx);
        int c1 = wrap_idx((int)floor(s[1]), ny);
        int c2 = wrap_idx((int)floor(s[2]), nz);
        int t = ((c2 / R.te[2]) * R.tn[1] + (c1 / R.te[1])) * R.tn[0] + (c0 / R.te[0]);
        R.key[b] = t;
        atomicAdd(&R.tcount[t], 1);
    }
    if (lane < 3) {
        const int d = lane;
        const double gd = d == 0 ? gam[0] : (d == 1 ? gam[1] : gam[2]);
        const double sd = d == 0 ? s[0] : (d == 1 ? s[1] : s[2]);
        const int i = 3 * b + d;
        const double dt = R.dt;
        const double vi = R.v[i];
        double tf = -R.ups * (vi - gd) * dt;
        if (R.noise) {
            const uint32_t step = (uint32_t)(R.ts[0] + 1);
            double z0, z1;
            normal_pair_f64(R.seed, kStreamBead, step, (uint32_t)i, 0u, z0, z1);
            tf += z0 * R.bead_coef;
        }
        R.tf[i] = tf;
        R.sg[i] = sd;
        double ft = tf + dt * R.fc[i];
        R.ftot[i] = ft / dt;
        double tmp = vi * dt;
        tmp += R.x[i];
        if (tmp > R.hi) tmp -= R.L;
        else if (tmp < R.lo) tmp += R.L;
        R.x[i] = tmp;
        R.v[i] = vi + ft / R.m[b];
    }
}

// ===========================================================================
// Binning: counting sort of beads by owning tile, deterministic bucket order.
// ===========================================================================
__global__ __launch_bounds__(1024) void k_bin_scan(const RepDev *__restrict__ reps, int nrep) {
    __shared__ int sh[1024];
    __shared__ int carry;
    const RepDev &R = reps[blockIdx.x];
    const int nt = R.ntiles;
    const int tid = threadIdx.x;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nt; base += 1024) {
        const int t = base + tid;
        const int c = t < nt ? R.tcount[t] : 0;
        sh[tid] = c;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int v = tid >= off ? sh[tid - off] : 0;
            __syncthreads();
            sh[tid] += v;
            __syncthreads();
        }
        if (t < nt) {
            R.tstart[t] = carry + sh[tid] - c;
            R.tcount[t] = 0;
            R.tcursor[t] = 0;
        }
        __syncthreads();
        if (tid == 1023) carry += sh[1023];
        __syncthreads();
    }
    if (tid == 0) R.tstart[nt] = carry;
}

__global__ __launch_bounds__(256) void k_bin_scatter(const RepDev *__restrict__ reps, int nrep,
                                                     const long long *__restrict__ start,
                                                     long long total) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const int b = (int)(g - start[r]);
    const int t = R.key[b];
    const int pos = R.tstart[t] + atomicAdd(&R.tcursor[t], 1);
    R.permu[pos] = b;
}

__global__ __launch_bounds__(256) void k_bin_rank(const RepDev *__restrict__ reps, int nrep,
                                                  const long long *__restrict__ start,
                                                  long long total) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const int b = (int)(g - start[r]);
    const int t = R.key[b];
    const int st = R.tstart[t], en = R.tstart[t + 1];
    int rank = 0;
    for (int q = st; q < en; q++) rank += (R.permu[q] < b) ? 1 : 0;
    R.perm[st + rank] = b;
}

// ===========================================================================
// Spread (Lambda): deterministic owner-computes reduction per tile.
// Semantics of cuda/SELM_Kernel.cu:424-470: each bead adds -t_f * w to the
// 4x4x4 stencil of each staggered component, w = ((1*phi_z)*phi_y)*phi_x.
// ===========================================================================
template <typename T>
struct BeadRec {
    int p[3][3];     // [comp][axis] first stencil index, wrapped
    T w[3][3][4];    // [comp][axis][offset] 1-D factors
    T f[3];          // -t_f per comp
};

template <typename T>
struct KahanAcc {
    T s, c;
    __device__ void init() { s = (T)0; c = (T)0; }
    __device__ void add(T x) {
        if constexpr (sizeof(T) == 4) {
            // Neumaier's compensated summation
            T t = s + x;
            if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
            else c += (x - t) + s;
            s = t;
        } else {
            s += x;
        }
    }
    __device__ T value() const { return s + c; }
};

constexpr int kSpreadChunk = 128;
constexpr int kSpreadThreads = 256;

template <typename T>
__global__ __launch_bounds__(kSpreadThreads) void k_spread(const RepDev *__restrict__ reps,
                                                           int nrep,
                                                           const long long *__restrict__ start,
                                                           long long total) {
    extern __shared__ __align__(16) unsigned char smem[];
    BeadRec<T> *rec = (BeadRec<T> *)smem;
    __shared__ int nb_start[27], nb_cnt[27], nb_pre[28];
    __shared__ int nnb;

    const long long g = blockIdx.x;
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const int t = (int)(g - start[r]);
    const int tid = threadIdx.x;
    const int nx = R.n[0], ny = R.n[1], nz = R.n[2];
    const int tx = t % R.tn[0], ty = (t / R.tn[0]) % R.tn[1], tz = t / (R.tn[0] * R.tn[1]);

    if (tid == 0) {
        // unique neighbour tiles per axis, fixed order
        int lx[3], ly[3], lz[3], cx = 0, cy = 0, cz = 0;
        for (int o = -1; o <= 1; o++) {
            int a = wrap_idx(tx + o, R.tn[0]);
            bool dup = false;
            for (int q = 0; q < cx; q++) dup |= lx[q] == a;
            if (!dup) lx[cx++] = a;
            a = wrap_idx(ty + o, R.tn[1]);
            dup = false;
            for (int q = 0; q < cy; q++) dup |= ly[q] == a;
            if (!dup) ly[cy++] = a;
            a = wrap_idx(tz + o, R.tn[2]);
            dup = false;
            for (int q = 0; q < cz; q++) dup |= lz[q] == a;
            if (!dup) lz[cz++] = a;
        }
        int q = 0, acc = 0;
        for (int c = 0; c < cz; c++)
            for (int bq = 0; bq < cy; bq++)
                for (int a = 0; a < cx; a++) {
                    int nt = (lz[c] * R.tn[1] + ly[bq]) * R.tn[0] + lx[a];
                    nb_start[q] = R.tstart[nt];
                    nb_cnt[q] = R.tstart[nt + 1] - R.tstart[nt];
                    nb_pre[q] = acc;
                    acc += nb_cnt[q];
                    q++;
                }
        nb_pre[q] = acc;
        nnb = q;
    }
    __syncthreads();
    const int ncand = nb_pre[nnb];

    // owned cells: tid and tid + 256 of the te0 x te1 x te2 tile
    const int ex = min(R.te[0], nx - tx * R.te[0]);
    const int ey = min(R.te[1], ny - ty * R.te[1]);
    const int ez = min(R.te[2], nz - tz * R.te[2]);
    const int ncell = ex * ey * ez;
    int cxs[2], cys[2], czs[2];
    bool own[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        int c = tid + q * kSpreadThreads;
        own[q] = c < ncell;
        int lx = c % ex, ly = (c / ex) % ey, lz = c / (ex * ey);
        cxs[q] = tx * R.te[0] + lx;
        cys[q] = ty * R.te[1] + ly;
        czs[q] = tz * R.te[2] + lz;
    }
    KahanAcc<T> acc[2][3];
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
        for (int d = 0; d < 3; d++) acc[q][d].init();

    const int f4 = R.f4;
    for (int base = 0; base < ncand; base += kSpreadChunk) {
        const int nq = min(kSpreadChunk, ncand - base);
        if (tid < nq) {
            const int cq = base + tid;
            int s = 0;
            while (nb_pre[s + 1] <= cq) s++;
            const int b = R.perm[nb_start[s] + (cq - nb_pre[s])];
            BeadRec<T> &B = rec[tid];
            const double sg0 = R.sg[3 * b], sg1 = R.sg[3 * b + 1], sg2 = R.sg[3 * b + 2];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                B.f[d] = (T)(-R.tf[3 * b + d]);
#pragma unroll
                for (int e = 0; e < 3; e++) {
                    double re = (e == 0 ? sg0 : (e == 1 ? sg1 : sg2)) - (d != e ? 0.5 : 0.0);
                    int p = (int)(floor(re) - 1.0);
                    int ne = e == 0 ? nx : (e == 1 ? ny : nz);
                    B.p[d][e] = wrap_idx(p, ne);
#pragma unroll
                    for (int o = 0; o < 4; o++) B.w[d][e][o] = (T)peskin_phi(fabs(re - (p + o)), f4);
                }
            }
        }
        __syncthreads();
        for (int k = 0; k < nq; k++) {
            const BeadRec<T> &B = rec[k];
#pragma unroll
            for (int d = 0; d < 3; d++) {
#pragma unroll
                for (int q = 0; q < 2; q++) {
                    int ox = cxs[q] - B.p[d][0];
                    ox += ox < 0 ? nx : 0;
                    int oy = cys[q] - B.p[d][1];
                    oy += oy < 0 ? ny : 0;
                    int oz = czs[q] - B.p[d][2];
                    oz += oz < 0 ? nz : 0;
                    if (ox < 4 && oy < 4 && oz < 4) {
                        T w = (B.w[d][2][oz] * B.w[d][1][oy]) * B.w[d][0][ox];
                        acc[q][d].add(B.f[d] * w);
                    }
                }
            }
        }
        __syncthreads();
    }
    T *fsp = (T *)R.fsp;
    const long long N = R.N;
#pragma unroll
    for (int q = 0; q < 2; q++) {
        if (!own[q]) continue;
        long long n = ((long long)czs[q] * ny + cys[q]) * nx + cxs[q];
#pragma unroll
        for (int d = 0; d < 3; d++) fsp[d * N + n] = acc[q][d].value();
    }
}

// ===========================================================================
// u* = u + (mu * L7(u) * dt + f) / rho  (cuda/SELM_Kernel.cu:155-195), in
// place over the spread buffer.
// ===========================================================================
template <typename T>
__global__ __launch_bounds__(256) void k_ustar(const RepDev *__restrict__ reps, int nrep,
                                               const long long *__restrict__ start,
                                               long long total) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const long long n = g - start[r];
    const int nx = R.n[0], ny = R.n[1], nz = R.n[2];
    const long long N = R.N, Nyx = (long long)nx * ny;
    const int k = (int)(n / Nyx);
    const int j = (int)((n % Nyx) / nx);
    const int i = (int)(n % nx);
    const long long i1 = (i != nx - 1) ? 1 : -i;
    const long long j1 = (j != ny - 1) ? nx : -(long long)j * nx;
    const long long k1 = (k != nz - 1) ? Nyx : -(long long)k * Nyx;
    const long long i0 = (i != 0) ? -1 : nx - 1;
    const long long j0 = (j != 0) ? -(long long)nx : (long long)(ny - 1) * nx;
    const long long k0 = (k != 0) ? -Nyx : (long long)(nz - 1) * Nyx;
    const T dxsq = (T)(R.dx * R.dx);
    const T mu = (T)R.mu, dt = (T)R.dt, rho = (T)R.rho;
    const T *u = (const T *)R.u;
    T *f = (T *)R.fsp;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const T *ud = u + d * N;
        T uc = ud[n];
        T lap = (ud[n + i1] + ud[n + i0] + ud[n + j1] + ud[n + j0] + ud[n + k1] + ud[n + k0] - (T)6 * uc) / dxsq;
        f[d * N + n] = uc + (mu * lap * dt + f[d * N + n]) / rho;
    }
}

// ===========================================================================
// FFT passes.  LDS layout is element-major: element e of sequence s at
// buf[e*S + s], so consecutive lanes touch consecutive sequences.
// ===========================================================================
template <typename T>
__device__ __forceinline__ void load_twiddles(cplx<T> *dst, const cplx<T> *src, int n) {
    for (int m = threadIdx.x; m < n; m += blockDim.x) dst[m] = src[m];
}

// Forward r2c along x: rows rho = (d*nz + z)*ny + y of length nx; two real
// rows form one complex sequence z = a + i b, separated after the DFT.
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
            const int k1 = (nz - k) % nz, j1 = (ny - j) % ny, i1 = (nx - i) % nx;
            const bool selfconj = (k1 == k && j1 == j && i1 == i);
            T coef = sqrt(-lam * (T)R.inv_dxsq);
            T re = selfconj ? coef : coef / sqrt((T)2);
            T im = selfconj ? (T)0 : re;
            const bool plane = (i == 0) || (2 * i == nx);
            const uint32_t nmode = (uint32_t)((long long)k * ncol + col);
            const uint32_t npart = (uint32_t)(((long long)k1 * ny + j1) * xc + i);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                T z0, z1;
                Normal<T>::pair(R.seed, kStreamFluid, step, nmode, (uint32_t)d, z0, z1);
                cplx<T> gk = mk<T>(z0 * re, z1 * im);
                if (plane) {
                    T y0, y1;
                    Normal<T>::pair(R.seed, kStreamFluid, step, npart, (uint32_t)d, y0, y1);
                    gk = mk<T>((gk.x + y0 * re) * (T)0.5, (gk.y - y1 * im) * (T)0.5);
                }
                w[d] = cadd(w[d], cscale(gk, amp));
            }
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
void launch_step_t(const StepLaunch &L, hipStream_t st) {
    const RepDev *reps = L.reps;
    const int nrep = L.nrep;
    if (L.tot_bead > 0) {
        hipLaunchKernelGGL(k_bead_update<T>, dim3(nblk(L.tot_bead, 4)), dim3(256), 0, st, reps, nrep,
                           L.st_bead, L.tot_bead);
    }
    hipLaunchKernelGGL(k_bin_scan, dim3(nrep), dim3(1024), 0, st, reps, nrep);
    if (L.tot_bead > 0) {
        hipLaunchKernelGGL(k_bin_scatter, dim3(nblk(L.tot_bead, 256)), dim3(256), 0, st, reps, nrep,
                           L.st_bead, L.tot_bead);
        hipLaunchKernelGGL(k_bin_rank, dim3(nblk(L.tot_bead, 256)), dim3(256), 0, st, reps, nrep,
                           L.st_bead, L.tot_bead);
    }
    hipLaunchKernelGGL(k_spread<T>, dim3((unsigned)L.tot_tile), dim3(kSpreadThreads),
                       kSpreadChunk * sizeof(BeadRec<T>), st, reps, nrep, L.st_tile, L.tot_tile);
    hipLaunchKernelGGL(k_ustar<T>, dim3(nblk(L.tot_pts, 256)), dim3(256), 0, st, reps, nrep, L.st_pts,
                       L.tot_pts);
    hipLaunchKernelGGL(k_fft_x_r2c<T>, dim3((unsigned)L.tot_fx), dim3(256), L.lds_fx, st, reps, nrep,
                       L.st_fx, L.tot_fx, L.twmax);
    hipLaunchKernelGGL(k_fft_y<T>, dim3((unsigned)L.tot_fy), dim3(256), L.lds_fy, st, reps, nrep,
                       L.st_fy, L.tot_fy, -1, L.twmax);
    hipLaunchKernelGGL(k_fft_z_proj<T>, dim3((unsigned)L.tot_fz), dim3(256), L.lds_fz, st, reps, nrep,
                       L.st_fz, L.tot_fz, L.twmax);
    hipLaunchKernelGGL(k_fft_y<T>, dim3((unsigned)L.tot_fy), dim3(256), L.lds_fy, st, reps, nrep,
                       L.st_fy, L.tot_fy, +1, L.twmax);
    hipLaunchKernelGGL(k_fft_x_c2r<T>, dim3((unsigned)L.tot_fx), dim3(256), L.lds_fx, st, reps, nrep,
                       L.st_fx, L.tot_fx, L.twmax);
    hipLaunchKernelGGL(k_bond_forces, dim3(L.tot_bead > 0 ? nblk(L.tot_bead, 256) : 1), dim3(256), 0, st,
                       reps, nrep, L.st_bead, L.tot_bead, 1);
}

void launch_step(const StepLaunch &L, hipStream_t st) {
    if (L.fp64) launch_step_t<double>(L, st);
    else launch_step_t<float>(L, st);
}

void launch_stage(const StepLaunch &L, int stage, hipStream_t st) {
    // one stage of the step, for per-stage timing; same kernels as launch_step
    const RepDev *reps = L.reps;
    const int nrep = L.nrep;
    const bool f64 = L.fp64;
#define SELM_T(kern) (f64 ? (const void *)kern<double> : (const void *)kern<float>)
    switch (stage) {
        case kStageBead:
            if (L.tot_bead > 0) {
                if (f64)
                    hipLaunchKernelGGL(k_bead_update<double>, dim3(nblk(L.tot_bead, 4)), dim3(256), 0, st, reps,
                                       nrep, L.st_bead, L.tot_bead);
                else
                    hipLaunchKernelGGL(k_bead_update<float>, dim3(nblk(L.tot_bead, 4)), dim3(256), 0, st, reps,
                                       nrep, L.st_bead, L.tot_bead);
            }
            break;
        case kStageBin:
            hipLaunchKernelGGL(k_bin_scan, dim3(nrep), dim3(1024), 0, st, reps, nrep);
            if (L.tot_bead > 0) {
                hipLaunchKernelGGL(k_bin_scatter, dim3(nblk(L.tot_bead, 256)), dim3(256), 0, st, reps, nrep,
                                   L.st_bead, L.tot_bead);
                hipLaunchKernelGGL(k_bin_rank, dim3(nblk(L.tot_bead, 256)), dim3(256), 0, st, reps, nrep,
                                   L.st_bead, L.tot_bead);
            }
            break;
        case kStageSpread:
            if (f64)
                hipLaunchKernelGGL(k_spread<double>, dim3((unsigned)L.tot_tile), dim3(kSpreadThreads),
                                   kSpreadChunk * sizeof(BeadRec<double>), st, reps, nrep, L.st_tile, L.tot_tile);
            else
                hipLaunchKernelGGL(k_spread<float>, dim3((unsigned)L.tot_tile), dim3(kSpreadThreads),
                                   kSpreadChunk * sizeof(BeadRec<float>), st, reps, nrep, L.st_tile, L.tot_tile);
            break;
        case kStageUstar:
            if (f64)
                hipLaunchKernelGGL(k_ustar<double>, dim3(nblk(L.tot_pts, 256)), dim3(256), 0, st, reps, nrep,
                                   L.st_pts, L.tot_pts);
            else
                hipLaunchKernelGGL(k_ustar<float>, dim3(nblk(L.tot_pts, 256)), dim3(256), 0, st, reps, nrep,
                                   L.st_pts, L.tot_pts);
            break;
        case kStageFftX:
            if (f64)
                hipLaunchKernelGGL(k_fft_x_r2c<double>, dim3((unsigned)L.tot_fx), dim3(256), L.lds_fx, st, reps,
                                   nrep, L.st_fx, L.tot_fx, L.twmax);
            else
                hipLaunchKernelGGL(k_fft_x_r2c<float>, dim3((unsigned)L.tot_fx), dim3(256), L.lds_fx, st, reps,
                                   nrep, L.st_fx, L.tot_fx, L.twmax);
            break;
        case kStageFftY:
        case kStageFftYi: {
            int dir = stage == kStageFftY ? -1 : +1;
            if (f64)
                hipLaunchKernelGGL(k_fft_y<double>, dim3((unsigned)L.tot_fy), dim3(256), L.lds_fy, st, reps,
                                   nrep, L.st_fy, L.tot_fy, dir, L.twmax);
            else
                hipLaunchKernelGGL(k_fft_y<float>, dim3((unsigned)L.tot_fy), dim3(256), L.lds_fy, st, reps,
                                   nrep, L.st_fy, L.tot_fy, dir, L.twmax);
            break;
        }
        case kStageFftZ:
            if (f64)
                hipLaunchKernelGGL(k_fft_z_proj<double>, dim3((unsigned)L.tot_fz), dim3(256), L.lds_fz, st,
                                   reps, nrep, L.st_fz, L.tot_fz, L.twmax);
            else
                hipLaunchKernelGGL(k_fft_z_proj<float>, dim3((unsigned)L.tot_fz), dim3(256), L.lds_fz, st,
                                   reps, nrep, L.st_fz, L.tot_fz, L.twmax);
            break;
        case kStageFftXi:
            if (f64)
                hipLaunchKernelGGL(k_fft_x_c2r<double>, dim3((unsigned)L.tot_fx), dim3(256), L.lds_fx, st, reps,
                                   nrep, L.st_fx, L.tot_fx, L.twmax);
            else
                hipLaunchKernelGGL(k_fft_x_c2r<float>, dim3((unsigned)L.tot_fx), dim3(256), L.lds_fx, st, reps,
                                   nrep, L.st_fx, L.tot_fx, L.twmax);
            break;
        case kStageBonds:
            hipLaunchKernelGGL(k_bond_forces, dim3(L.tot_bead > 0 ? nblk(L.tot_bead, 256) : 1), dim3(256), 0,
                               st, reps, nrep, L.st_bead, L.tot_bead, 1);
            break;
        default: break;
    }
#undef SELM_T
}

void launch_bond_forces(const StepLaunch &L, int advance, hipStream_t st) {
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
