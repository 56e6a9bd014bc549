// selm_fft_kernels.hip -- the 3-D FFT passes of the step on the specialised
// (compile-time length) path.  Each kernel switches on the replica's edge
// length per workgroup, so one launch still covers a heterogeneous batch.
//
//   k_fx_r2c    rows along x of u* (evaluated on load from u^n and the sparse
//               spread): two real rows packed as one complex sequence
//   k_fy        columns along y (forward or inverse), in place on spec
//   k_fz_proj   columns along z: forward DFT, thermal forcing in k-space,
//               Helmholtz projection, k=0 removal, 1/N, inverse DFT
//   k_fx_c2r    rows along x back to real u^{n+1}
//
// Spectrum layout: spec[d][z][y][xc], xc = nx/2+1, unnormalised DFT.
#include "selm_fft.hpp"
#include "selm_fluid.hpp"
#include "selm_kernels.hpp"

namespace selm {

template <typename T>
struct FastCfg;
template <>
struct FastCfg<float> {
    static constexpr int BX = 32, BY = 16, BZ = 8;           // rows / columns per workgroup
    static constexpr int NTX = 256, NTY = 256, NTZ = 256;     // threads per workgroup
};
template <>
struct FastCfg<double> {
    static constexpr int BX = 16, BY = 8, BZ = 8;
    static constexpr int NTX = 256, NTY = 256, NTZ = 256;
};

template <typename T>
__device__ __forceinline__ void tw_to_lds(cplx<T> *dst, const void *src, int n) {
    const cplx<T> *s = (const cplx<T> *)src;
    for (int m = threadIdx.x; m < n; m += blockDim.x) dst[m] = s[m];
}

// ------------------------------------------------------------------ X forward
template <typename T, int N>
__device__ __forceinline__ void fx_r2c_body(const RepDev &R, long long item, cplx<T> *tw, cplx<T> *a) {
    constexpr int BX = FastCfg<T>::BX, P = BX / 2, S = P + 1, NT = FastCfg<T>::NTX;
    constexpr int XC = N / 2 + 1;
    const int ny = R.n[1], nz = R.n[2];
    const long long rows = 3LL * ny * nz;
    const long long row0 = item * BX;
    const int step = (int)(R.ts[0] + 1);
    // rows (d, z, y) of u*, evaluated on load from u^n and the sparse spread
    for (int idx = threadIdx.x; idx < BX * N; idx += NT) {
        const int rl = idx / N, e = idx - rl * N;
        const long long row = row0 + rl;
        T v = (T)0;
        if (row < rows) {
            const int y = (int)(row % ny);
            const long long dzr = row / ny;
            const int z = (int)(dzr % nz), d = (int)(dzr / nz);
            int tyz, lyz;
            spread_row(R, y, z, tyz, lyz);
            v = ustar_at<T>(R, d, e, y, z, tyz, lyz, step);
        }
        ((T *)&a[e * S + (rl >> 1)])[rl & 1] = v;
    }
    __syncthreads();
    auto nold = [](int, int) { return cplx<T>{}; };
    auto nost = [](int, int, cplx<T>) {};
    fft_run<T, N, S, NT, false, false>(a, tw, -1, nold, nost);
    cplx<T> *out = (cplx<T> *)R.spec;
    for (int idx = threadIdx.x; idx < BX * XC; idx += NT) {
        const int rl = idx / XC, k = idx - rl * XC;
        const long long row = row0 + rl;
        if (row >= rows) continue;
        const int p = rl >> 1;
        const cplx<T> zk = a[k * S + p];
        const cplx<T> zm = a[((N - k) % N) * S + p];
        cplx<T> o;
        if ((rl & 1) == 0) o = mk<T>((zk.x + zm.x) * (T)0.5, (zk.y - zm.y) * (T)0.5);
        else o = mk<T>((zk.y + zm.y) * (T)0.5, (zm.x - zk.x) * (T)0.5);
        out[row * XC + k] = o;
    }
}

// ------------------------------------------------------------------ X inverse
template <typename T, int N>
__device__ __forceinline__ void fx_c2r_body(const RepDev &R, long long item, cplx<T> *tw, cplx<T> *a) {
    constexpr int BX = FastCfg<T>::BX, P = BX / 2, S = P + 1, NT = FastCfg<T>::NTX;
    constexpr int XC = N / 2 + 1;
    const long long rows = 3LL * R.n[1] * R.n[2];
    const long long row0 = item * BX;
    const cplx<T> *in = (const cplx<T> *)R.spec;
    // Z_e = A_e + i B_e with the Hermitian extension; only the real parts of
    // the DC and Nyquist coefficients enter (c2r semantics)
    for (int idx = threadIdx.x; idx < P * N; idx += NT) {
        const int p = idx / N, e = idx - p * N;
        const long long ra = row0 + 2 * p, rb = ra + 1;
        const int src = e < XC ? e : N - e;
        cplx<T> A = ra < rows ? in[ra * XC + src] : mk<T>((T)0, (T)0);
        cplx<T> B = rb < rows ? in[rb * XC + src] : mk<T>((T)0, (T)0);
        if (e == 0 || 2 * e == N) {
            A.y = (T)0;
            B.y = (T)0;
        }
        if (e >= XC) {
            A.y = -A.y;
            B.y = -B.y;
        }
        a[e * S + p] = mk<T>(A.x - B.y, A.y + B.x);
    }
    __syncthreads();
    auto nold = [](int, int) { return cplx<T>{}; };
    auto nost = [](int, int, cplx<T>) {};
    fft_run<T, N, S, NT, false, false>(a, tw, +1, nold, nost);
    T *out = (T *)R.u;
    for (int idx = threadIdx.x; idx < BX * N; idx += NT) {
        const int rl = idx / N, e = idx - rl * N;
        const long long row = row0 + rl;
        if (row >= rows) continue;
        const cplx<T> z = a[e * S + (rl >> 1)];
        out[row * N + e] = (rl & 1) ? z.y : z.x;
    }
}

// ------------------------------------------------------------------ Y
template <typename T, int N>
__device__ __forceinline__ void fy_body(const RepDev &R, long long item, int dir, cplx<T> *tw, cplx<T> *a) {
    constexpr int S = FastCfg<T>::BY, NT = FastCfg<T>::NTY;
    const int xc = R.xc;
    const int nbx = (xc + S - 1) / S;
    const int xb = (int)(item % nbx);
    const long long dz = item / nbx;
    const int x0 = xb * S;
    cplx<T> *sp = (cplx<T> *)R.spec + dz * N * xc + x0;
    const int valid = min(S, xc - x0);
    auto ld = [&](int e, int s) { return s < valid ? sp[(long long)e * xc + s] : mk<T>((T)0, (T)0); };
    auto st = [&](int e, int s, cplx<T> v) {
        if (s < valid) sp[(long long)e * xc + s] = v;
    };
    fft_run<T, N, S, NT, true, true>(a, tw, dir, ld, st);
}

// ------------------------------------------------------------------ Z + projection
template <typename T, int N>
__device__ __forceinline__ void fz_body(const RepDev &R, long long item, cplx<T> *tw, cplx<T> *a) {
    constexpr int BZ = FastCfg<T>::BZ, S = 3 * BZ, NT = FastCfg<T>::NTZ;
    const int nx = R.n[0], ny = R.n[1], xc = R.xc;
    const long long ncol = (long long)ny * xc;
    const long long c0 = item * BZ;
    cplx<T> *sp = (cplx<T> *)R.spec;
    auto ld = [&](int e, int s) {
        const int d = s / BZ, c = s - d * BZ;
        const long long col = c0 + c;
        return col < ncol ? sp[((long long)d * N + e) * ncol + col] : mk<T>((T)0, (T)0);
    };
    auto st = [&](int e, int s, cplx<T> v) {
        const int d = s / BZ, c = s - d * BZ;
        const long long col = c0 + c;
        if (col < ncol) sp[((long long)d * N + e) * ncol + col] = v;
    };
    auto nold = [](int, int) { return cplx<T>{}; };
    auto nost = [](int, int, cplx<T>) {};
    fft_run<T, N, S, NT, true, false>(a, tw, -1, ld, nost);

    const T *kx = (const T *)R.ktab[0], *ky = (const T *)R.ktab[1], *kz = (const T *)R.ktab[2];
    const T invN = (T)(1.0 / (double)R.N);
    const uint32_t step = (uint32_t)(R.ts[0] + 1);
    const T amp = (T)R.noise_amp;
    for (int idx = threadIdx.x; idx < N * BZ; idx += NT) {
        const int e = idx / BZ, c = idx - e * BZ;
        const long long col = c0 + c;
        if (col >= ncol) continue;
        const int j = (int)(col / xc), i = (int)(col - (long long)j * xc), k = e;
        cplx<T> w[3];
#pragma unroll
        for (int d = 0; d < 3; d++) w[d] = a[e * S + d * BZ + c];
        const T lam = (kx[4 * i] + ky[4 * j]) + kz[4 * k];
        if (R.noise) {
            const int k1 = (N - k) % N, j1 = (ny - j) % ny, i1 = (nx - i) % nx;
            const bool selfconj = (k1 == k && j1 == j && i1 == i);
            const T coef = sqrt(-lam * (T)R.inv_dxsq);
            const T re = selfconj ? coef : coef / sqrt((T)2);
            const T im = selfconj ? (T)0 : re;
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
            for (int d = 0; d < 3; d++) a[e * S + d * BZ + c] = mk<T>((T)0, (T)0);
            continue;
        }
        const cplx<T> D0 = mk<T>(kx[4 * i + 1], kx[4 * i + 2]);
        const cplx<T> D1 = mk<T>(ky[4 * j + 1], ky[4 * j + 2]);
        const cplx<T> D2 = mk<T>(kz[4 * k + 1], kz[4 * k + 2]);
        const cplx<T> sdiv = cadd(cadd(cmul(D0, w[0]), cmul(D1, w[1])), cmul(D2, w[2]));
        const cplx<T> q = mk<T>(sdiv.x / lam, sdiv.y / lam);
        const cplx<T> G[3] = {mk<T>(-D0.x, D0.y), mk<T>(-D1.x, D1.y), mk<T>(-D2.x, D2.y)};
#pragma unroll
        for (int d = 0; d < 3; d++) a[e * S + d * BZ + c] = cscale(csub(w[d], cmul(G[d], q)), invN);
    }
    __syncthreads();
    fft_run<T, N, S, NT, false, true>(a, tw, +1, nold, st);
}

// ------------------------------------------------------------------ kernels
#define SELM_CASE_X(n) \
    case n: fx_r2c_body<T, n>(R, item, tw, a); break;
template <typename T>
__global__ __launch_bounds__(FastCfg<T>::NTX) void k_fx_r2c(const RepDev *__restrict__ reps, int nrep,
                                                            const long long *__restrict__ start,
                                                            long long total, int twmax) {
    extern __shared__ __align__(16) unsigned char smem[];
    const long long g = xcd_remap(blockIdx.x, gridDim.x);
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const long long item = g - start[r];
    cplx<T> *tw = (cplx<T> *)smem;
    cplx<T> *a = tw + twmax;
    tw_to_lds<T>(tw, R.tw[0], R.n[0]);
    switch (R.n[0]) { SELM_FFT_FOR_EACH_N(SELM_CASE_X) default: break; }
}
#undef SELM_CASE_X

#define SELM_CASE_XI(n) \
    case n: fx_c2r_body<T, n>(R, item, tw, a); break;
template <typename T>
__global__ __launch_bounds__(FastCfg<T>::NTX) void k_fx_c2r(const RepDev *__restrict__ reps, int nrep,
                                                            const long long *__restrict__ start,
                                                            long long total, int twmax) {
    extern __shared__ __align__(16) unsigned char smem[];
    const long long g = xcd_remap(blockIdx.x, gridDim.x);
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const long long item = g - start[r];
    cplx<T> *tw = (cplx<T> *)smem;
    cplx<T> *a = tw + twmax;
    tw_to_lds<T>(tw, R.tw[0], R.n[0]);
    switch (R.n[0]) { SELM_FFT_FOR_EACH_N(SELM_CASE_XI) default: break; }
}
#undef SELM_CASE_XI

#define SELM_CASE_Y(n) \
    case n: fy_body<T, n>(R, item, dir, tw, a); break;
template <typename T>
__global__ __launch_bounds__(FastCfg<T>::NTY) void k_fy(const RepDev *__restrict__ reps, int nrep,
                                                        const long long *__restrict__ start, long long total,
                                                        int dir, int twmax) {
    extern __shared__ __align__(16) unsigned char smem[];
    const long long g = xcd_remap(blockIdx.x, gridDim.x);
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const long long item = g - start[r];
    cplx<T> *tw = (cplx<T> *)smem;
    cplx<T> *a = tw + twmax;
    tw_to_lds<T>(tw, R.tw[1], R.n[1]);
    __syncthreads();
    switch (R.n[1]) { SELM_FFT_FOR_EACH_N(SELM_CASE_Y) default: break; }
}
#undef SELM_CASE_Y

#define SELM_CASE_Z(n) \
    case n: fz_body<T, n>(R, item, tw, a); break;
template <typename T>
__global__ __launch_bounds__(FastCfg<T>::NTZ) void k_fz_proj(const RepDev *__restrict__ reps, int nrep,
                                                             const long long *__restrict__ start,
                                                             long long total, int twmax) {
    extern __shared__ __align__(16) unsigned char smem[];
    const long long g = xcd_remap(blockIdx.x, gridDim.x);
    if (g >= total) return;
    const int r = find_rep(start, nrep, g);
    const RepDev &R = reps[r];
    const long long item = g - start[r];
    cplx<T> *tw = (cplx<T> *)smem;
    cplx<T> *a = tw + twmax;
    tw_to_lds<T>(tw, R.tw[2], R.n[2]);
    __syncthreads();
    switch (R.n[2]) { SELM_FFT_FOR_EACH_N(SELM_CASE_Z) default: break; }
}
#undef SELM_CASE_Z

// ------------------------------------------------------------------ host side
static inline bool fast_len(int n) {
    switch (n) {
#define SELM_CASE_LEN(x) case x:
        SELM_FFT_FOR_EACH_N(SELM_CASE_LEN)
#undef SELM_CASE_LEN
        return true;
        default: return false;
    }
}

bool fft_fast_supported(const int n[3]) { return fast_len(n[0]) && fast_len(n[1]) && fast_len(n[2]); }

void fft_fast_blocks(bool fp64, int *bx, int *by, int *bz) {
    if (fp64) {
        *bx = FastCfg<double>::BX;
        *by = FastCfg<double>::BY;
        *bz = FastCfg<double>::BZ;
    } else {
        *bx = FastCfg<float>::BX;
        *by = FastCfg<float>::BY;
        *bz = FastCfg<float>::BZ;
    }
}

// LDS bytes of each fast pass for edges n (twiddle slots twmax first)
void fft_fast_lds(bool fp64, const int n[3], int twmax, size_t *lx, size_t *ly, size_t *lz) {
    const size_t csz = fp64 ? 16 : 8;
    int BX, BY, BZ;
    fft_fast_blocks(fp64, &BX, &BY, &BZ);
    *lx = ((size_t)twmax + (size_t)(BX / 2 + 1) * n[0]) * csz;
    *ly = ((size_t)twmax + (size_t)BY * n[1]) * csz;
    *lz = ((size_t)twmax + (size_t)3 * BZ * n[2]) * csz;
}

static inline unsigned pad8(long long n) { return (unsigned)((n + 7) & ~7LL); }

template <typename T>
static void launch_fast_t(const StepLaunch &L, int stage, hipStream_t st) {
    switch (stage) {
        case kStageFftX:
            hipLaunchKernelGGL(k_fx_r2c<T>, dim3(pad8(L.tot_fx)), dim3(FastCfg<T>::NTX), L.lds_fx, st, L.reps,
                               L.nrep, L.st_fx, L.tot_fx, L.twmax);
            break;
        case kStageFftY:
        case kStageFftYi:
            hipLaunchKernelGGL(k_fy<T>, dim3(pad8(L.tot_fy)), dim3(FastCfg<T>::NTY), L.lds_fy, st, L.reps,
                               L.nrep, L.st_fy, L.tot_fy, stage == kStageFftY ? -1 : +1, L.twmax);
            break;
        case kStageFftZ:
            hipLaunchKernelGGL(k_fz_proj<T>, dim3(pad8(L.tot_fz)), dim3(FastCfg<T>::NTZ), L.lds_fz, st,
                               L.reps, L.nrep, L.st_fz, L.tot_fz, L.twmax);
            break;
        case kStageFftXi:
            hipLaunchKernelGGL(k_fx_c2r<T>, dim3(pad8(L.tot_fx)), dim3(FastCfg<T>::NTX), L.lds_fx, st, L.reps,
                               L.nrep, L.st_fx, L.tot_fx, L.twmax);
            break;
        default: break;
    }
}

void launch_fft_fast(const StepLaunch &L, int stage, hipStream_t st) {
    if (L.fp64) launch_fast_t<double>(L, stage, st);
    else launch_fast_t<float>(L, stage, st);
}

}  // namespace selm
