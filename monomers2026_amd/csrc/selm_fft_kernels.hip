// selm_fft_kernels.hip -- the 3-D FFT of the step on the compile-time-length
// path, as a transposing pipeline: every pass reads whole contiguous rows of
// its transform axis and writes its output transposed, so that the next
// pass's axis is contiguous in memory.  Every global access is a contiguous
// run of >= 256 bytes (full rows, or row-blocks of RB consecutive elements).
//
//   pass   reads                      transform               writes
//   Xf     u^n rows (x fastest), u*   r2c along x             B[d][z][kx][y]
//   Yf     B rows (y fastest)         DFT along y             C[d][kx][ky][z]
//   Z      C rows (z fastest)         DFT along z, thermal    C in place
//                                     forcing, projection,
//                                     k=0 removal, 1/N, iDFT
//   Yi     C (RB z per row-block)     iDFT along y            B rows
//   Xi     B (RB y per row-block)     c2r along x             u^{n+1} rows
//
// B lives in spec, C in spec2.  Each kernel switches on the replica's edge
// length per workgroup, so one launch still covers a heterogeneous batch.
//
// Precision: arrays in HBM hold T (float in FP32-bulk, double in FP64).
// Every butterfly, twiddle product, u* and projection is evaluated in FP64
// (type C).  The data is rounded to T where it is stored or exchanged
// through LDS (FastCfg<T>::L).  The FP32-bulk path therefore keeps FP32's
// HBM traffic while its transforms lose about one FP32 rounding per
// exchange, rather than one per butterfly level.
#include <algorithm>

#include "selm_fft.hpp"
#include "selm_fluid.hpp"
#include "selm_forces.hpp"
#include "selm_kernels.hpp"

namespace selm {

template <typename T>
struct FastCfg;
using C = double;  // arithmetic type of every pass

template <>
struct FastCfg<float> {
    using L = float;               // LDS element type
    static constexpr int RB = 32;  // rows per block in Xf/Xi, z per block in Yf/Yi (256 B runs)
    static constexpr int Q = 16;   // (kx, ky) columns per Z workgroup, times 3 components
    static constexpr int NT = 256, NTZ = 512, NTZ9 = 256;
};
template <>
struct FastCfg<double> {
    using L = double;
    static constexpr int RB = 16;
    static constexpr int Q = 8;
    static constexpr int NT = 256, NTZ = 512, NTZ9 = 256;
};

// Twiddles of a length-N axis, loaded into registers (TWI = ceil(N / NT) per
// thread) so that a kernel can issue its data loads before it writes them to
// LDS: a write to LDS right after the load would wait for the load.
template <int N, int NT>
struct TwRegs {
    static constexpr int TWI = (N + NT - 1) / NT;
    cplx<C> v[TWI];
    __device__ __forceinline__ void load(const void *src) {
        const SELM_GLOBAL cplx<C> *s = gbl((const cplx<C> *)src);
#pragma unroll
        for (int k = 0; k < TWI; k++) v[k] = ldc(s + min((int)threadIdx.x + k * NT, N - 1));
    }
    __device__ __forceinline__ void store(cplx<C> *dst) const {
#pragma unroll
        for (int k = 0; k < TWI; k++)
            if ((int)threadIdx.x + k * NT < N) dst[threadIdx.x + k * NT] = v[k];
    }
};


template <typename T>
__device__ __forceinline__ cplx<T> czero() {
    return mk<T>((T)0, (T)0);
}

// Element i < COUNT is read by ld(i) and written by st(i, v); thread t takes
// i = t, t + NT, ...  Each thread issues BATCH reads before it writes any,
// so the latencies of its global loads overlap instead of adding up.
// Indices past COUNT re-read element COUNT-1 and are not written.
template <int NT, int COUNT, int BATCH, class LD, class ST>
__device__ __forceinline__ void copy_batched(const LD &ld, const ST &st) {
    constexpr int ITER = (COUNT + NT - 1) / NT;
    using V = decltype(ld(0));
#pragma unroll
    for (int b0 = 0; b0 < ITER; b0 += BATCH) {
        V buf[BATCH];
#pragma unroll
        for (int k = 0; k < BATCH; k++) buf[k] = ld(min((int)threadIdx.x + (b0 + k) * NT, COUNT - 1));
#pragma unroll
        for (int k = 0; k < BATCH; k++) {
            const int i = (int)threadIdx.x + (b0 + k) * NT;
            if (b0 + k < ITER && i < COUNT) st(i, buf[k]);
        }
    }
}

// ------------------------------------------------------------------ Xf
// Workgroup: component d, plane z, rows y0 .. y0+RB-1 of u*; rows are paired
// into complex sequences (even row = real part, odd row = imaginary part).
//
// u* (the arithmetic of ustar_at, cuda/SELM_Kernel.cu:155-195) is formed on
// load.  The rows y0-1 .. y0+RB of plane z are staged in LDS first, so the
// x and y neighbours come from LDS; the z neighbours and the spread pool are
// read per element, a batch of elements at a time, with the tiles' stamps
// read into LDS beforehand so no load waits on another.
template <typename T>
struct XfLoads {
    T zp, zm, f;  // u at z+1 and z-1, the tile's spread value
};

template <typename T, int N>
__device__ __forceinline__ void xf_body(const RepDev &R, long long item, cplx<C> *tw,
                                        cplx<typename FastCfg<T>::L> *a) {
    using L = typename FastCfg<T>::L;
    constexpr int RB = FastCfg<T>::RB, P = RB / 2, S = P + 1, NT = FastCfg<T>::NT, XC = N / 2 + 1;
    // the twiddles are written to LDS after the data loads are issued
    TwRegs<N, NT> twr;
    twr.load(R.twd[0]);
    const int ny = R.n[1], nz = R.n[2];
    const int nyb = (ny + RB - 1) / RB;
    const int yb = (int)(item % nyb);
    const long long dz = item / nyb;
    const int z = (int)(dz % nz), d = (int)(dz / nz);
    const int y0 = yb * RB;
    const int step = (int)(R.ts[0] + 1);
    const long long Nyx = (long long)N * ny;
    const T *ud = (const T *)R.u + d * R.N;
    T *X = (T *)(a + S * N);  // rows y0-1 .. y0+RB of plane z (periodic in y)

    // stamps of the tiles under rows y0 .. yl: tile edges are >= 4 cells
    constexpr int FW = N / 4 + 2, FH = RB / 4 + 2;
    __shared__ int okf[FH * FW];
    const int te0 = R.te[0], te1 = R.te[1], te2 = R.te[2], tn0 = R.tn[0], tn1 = R.tn[1];
    const int tz = z / te2;
    const int yl = min(y0 + RB, ny) - 1;
    const int ty0 = y0 / te1, nty = yl / te1 - ty0 + 1;
    // stamps loaded here, written to LDS after the rows' loads are issued
    constexpr int OKI = (FH * FW + NT - 1) / NT;
    int okv[OKI];
#pragma unroll
    for (int k = 0; k < OKI; k++) {
        const int i = min((int)threadIdx.x + k * NT, nty * tn0 - 1);
        const int q = i / tn0, tx = i - q * tn0;
        okv[k] = R.tstamp[(tz * tn1 + ty0 + q) * tn0 + tx];
    }
    constexpr int SROWS = (RB + 2) * N;
    copy_batched<NT, SROWS, (SROWS + NT - 1) / NT>(
        [&](int idx) {
            const int r = idx / N, e = idx - r * N;
            int y = y0 - 1 + r;
            y += y < 0 ? ny : 0;
            y -= y >= ny ? ny : 0;
            return ud[(long long)z * Nyx + (long long)y * N + e];
        },
        [&](int idx, T v) { X[idx] = v; });
    twr.store(tw);
#pragma unroll
    for (int k = 0; k < OKI; k++) {
        const int i = (int)threadIdx.x + k * NT;
        if (i < nty * tn0) {
            const int q = i / tn0, tx = i - q * tn0;
            okf[q * FW + tx] = okv[k] == step;
        }
    }
    __syncthreads();

    const T *up = ud + (long long)(z + 1 < nz ? z + 1 : 0) * Nyx;
    const T *um = ud + (long long)(z > 0 ? z - 1 : nz - 1) * Nyx;
    const T *pool = (const T *)R.pool;
    const int tcells = te0 * te1 * te2;
    const C dxsq = (C)(R.dx * R.dx), mu = (C)R.mu, dt = (C)R.dt, rho = (C)R.rho;
    copy_batched<NT, RB * N, 8>(
        [&](int idx) {
            const int rl = idx / N, e = idx - rl * N;
            const int y = min(y0 + rl, ny - 1);  // rows past the grid are loaded, then zeroed
            const int ty = y / te1, tx = e / te0;
            const long long t = (long long)(tz * tn1 + ty) * tn0 + tx;
            const int lyz = ((z - tz * te2) * te1 + (y - ty * te1)) * te0;
            const long long row = (long long)y * N + e;
            XfLoads<T> v;
            v.zp = up[row];
            v.zm = um[row];
            // only the tiles a bead touched this step are read
            v.f = okf[(y / te1 - ty0) * FW + tx] ? pool[(t * 3 + d) * tcells + lyz + (e - tx * te0)] : (T)0;
            return v;
        },
        [&](int idx, XfLoads<T> v) {
            const int rl = idx / N, e = idx - rl * N;
            const int y = y0 + rl;
            C us = (C)0;
            if (y < ny) {
                // neighbour order x+, x-, y+, y-, z+, z- (as ustar_at)
                const T *row = X + (rl + 1) * N;
                const C uc = (C)row[e];
                const C lap = ((C)row[e + 1 < N ? e + 1 : 0] + (C)row[e > 0 ? e - 1 : N - 1] +
                               (C)X[(rl + 2) * N + e] + (C)X[rl * N + e] + (C)v.zp + (C)v.zm - (C)6 * uc) /
                              dxsq;
                us = uc + (mu * lap * dt + (C)v.f) / rho;
            }
            ((L *)&a[e * S + (rl >> 1)])[rl & 1] = (L)us;
        });
    __syncthreads();
    auto nold = [](int, int) { return cplx<C>{}; };
    auto nost = [](int, int, cplx<C>) {};
    fft_run<C, L, N, S, P, NT, false, false>(a, tw, -1, nold, nost);
    // B[d][z][kx][y]: RB consecutive y per kx
    cplx<T> *out = (cplx<T> *)R.spec + ((long long)d * nz + z) * XC * ny;
    for (int idx = threadIdx.x; idx < XC * RB; idx += NT) {
        const int k = idx / RB, rl = idx - k * RB;
        const int y = y0 + rl;
        if (y >= ny) continue;
        const int p = rl >> 1;
        const cplx<C> zk = ccast<C>(a[k * S + p]);
        const cplx<C> zm = ccast<C>(a[((N - k) % N) * S + p]);
        cplx<C> o;
        if ((rl & 1) == 0) o = mk<C>((zk.x + zm.x) * 0.5, (zk.y - zm.y) * 0.5);
        else o = mk<C>((zk.y + zm.y) * 0.5, (zm.x - zk.x) * 0.5);
        out[(long long)k * ny + y] = ccast<T>(o);
    }
}

// ------------------------------------------------------------------ Yf / Yi
// Workgroup: component d, x-frequency kx, planes z0 .. z0+RB-1.
// Forward: B rows (y contiguous) -> DFT_y -> C[d][kx][ky][z0..] (RB z contiguous).
// Inverse: C -> iDFT_y -> B rows.
template <typename T, int N, bool FWD>
__device__ __forceinline__ void y_body(const RepDev &R, long long item, cplx<C> *tw,
                                       cplx<typename FastCfg<T>::L> *a) {
    using L = typename FastCfg<T>::L;
    constexpr int RB = FastCfg<T>::RB, S = RB + 1, NT = FastCfg<T>::NT;
    // the twiddles are written to LDS after the data loads are issued
    TwRegs<N, NT> twr;
    twr.load(R.twd[1]);
    const int nz = R.n[2], xc = R.xc;
    const int nzb = (nz + RB - 1) / RB;
    const int zb = (int)(item % nzb);
    const long long dk = item / nzb;
    const int kx = (int)(dk % xc), d = (int)(dk / xc);
    const int z0 = zb * RB;
    const int zv = min(RB, nz - z0);
    cplx<T> *B = (cplx<T> *)R.spec + (long long)d * nz * xc * N;             // [z][kx][y]
    cplx<T> *Cc = (cplx<T> *)R.spec2 + ((long long)d * xc + kx) * N * nz;  // [ky][z]
    if (FWD) {
        copy_batched<NT, RB * N, 16>(
            [&](int idx) {
                const int q = idx / N, y = idx - q * N;
                const cplx<T> v = B[((long long)(z0 + min(q, zv - 1)) * xc + kx) * N + y];
                return q < zv ? v : czero<T>();
            },
            [&](int idx, cplx<T> v) {
                const int q = idx / N, y = idx - q * N;
                a[y * S + q] = ccast<L>(v);
            });
    } else {
        copy_batched<NT, N * RB, 16>(
            [&](int idx) {
                const int ky = idx / RB, q = idx - ky * RB;
                const cplx<T> v = Cc[(long long)ky * nz + z0 + min(q, zv - 1)];
                return q < zv ? v : czero<T>();
            },
            [&](int idx, cplx<T> v) {
                const int ky = idx / RB, q = idx - ky * RB;
                a[ky * S + q] = ccast<L>(v);
            });
    }
    twr.store(tw);
    __syncthreads();
    auto nold = [](int, int) { return cplx<C>{}; };
    auto nost = [](int, int, cplx<C>) {};
    fft_run<C, L, N, S, RB, NT, false, false>(a, tw, FWD ? -1 : +1, nold, nost);
    if (FWD) {
        for (int idx = threadIdx.x; idx < N * RB; idx += NT) {
            const int ky = idx / RB, q = idx - ky * RB;
            if (q < zv) Cc[(long long)ky * nz + z0 + q] = ccast<T>(a[ky * S + q]);
        }
    } else {
        for (int idx = threadIdx.x; idx < RB * N; idx += NT) {
            const int q = idx / N, y = idx - q * N;
            if (q < zv) B[((long long)(z0 + q) * xc + kx) * N + y] = ccast<T>(a[y * S + q]);
        }
    }
}

// ------------------------------------------------------------------ Z + projection
// Workgroup: x-frequency kx, y-frequencies ky0 .. ky0+Q-1, all three
// components; rows of C are contiguous along z.
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
template <typename T, int N>
__device__ __forceinline__ void z_body(const RepDev &R, long long item, cplx<C> *tw,
                                       cplx<typename FastCfg<T>::L> *a) {
    using L = typename FastCfg<T>::L;
    constexpr int Q = FastCfg<T>::Q, S = 3 * Q + 1, NT = FastCfg<T>::NTZ;
    // the twiddles are written to LDS after the data loads are issued
    TwRegs<N, NT> twr;
    twr.load(R.twd[2]);
    const int ny = R.n[1], xc = R.xc;
    const int nyb = (ny + Q - 1) / Q;
    const int jb = (int)(item % nyb);
    const int kx = (int)(item / nyb);
    const int j0 = jb * Q;
    const int jv = min(Q, ny - j0);
    cplx<T> *Cs = (cplx<T> *)R.spec2;
    // sequence s = d*Q + q <-> row C[d][kx][j0+q][*]
    copy_batched<NT, 3 * Q * N, 16>(
        [&](int idx) {
            const int s = idx / N, e = idx - s * N;
            const int d = s / Q, q = s - d * Q;
            const cplx<T> v = Cs[(((long long)d * xc + kx) * ny + j0 + min(q, jv - 1)) * N + e];
            return q < jv ? v : czero<T>();
        },
        [&](int idx, cplx<T> v) {
            const int s = idx / N, e = idx - s * N;
            a[e * S + s] = ccast<L>(v);
        });
    twr.store(tw);
    __syncthreads();
    auto nold = [](int, int) { return cplx<C>{}; };
    auto nost = [](int, int, cplx<C>) {};
    fft_run<C, L, N, S, 3 * Q, NT, false, false>(a, tw, -1, nold, nost);

    const C *kxt = (const C *)R.ktabd[0], *kyt = (const C *)R.ktabd[1], *kzt = (const C *)R.ktabd[2];
    const C invN = 1.0 / (double)R.N;
    const int i = kx;
    const long long Nc = R.Nc;
    const cplx<T> *nb = (const cplx<T> *)R.nbuf + ((long long)kx * ny + j0) * N;
    // consecutive threads take consecutive kz of one row: the noise rows are read contiguously
    for (int idx = threadIdx.x; idx < N * Q; idx += NT) {
        const int q = idx / N, e = idx - q * N;
        if (q >= jv) continue;
        const int j = j0 + q, k = e;
        cplx<C> w[3];
#pragma unroll
        for (int d = 0; d < 3; d++) w[d] = ccast<C>(a[e * S + d * Q + q]);
        const C lam = (kxt[4 * i] + kyt[4 * j]) + kzt[4 * k];
        if (R.noise) {
#pragma unroll
            for (int d = 0; d < 3; d++) w[d] = cadd(w[d], ccast<C>(nb[d * Nc + (long long)q * N + e]));
        }
        if (i == 0 && j == 0 && k == 0) {
            if (R.diag) {
#pragma unroll
                for (int d = 0; d < 3; d++) R.diag[d] = (double)w[d].x;
            }
#pragma unroll
            for (int d = 0; d < 3; d++) a[e * S + d * Q + q] = ccast<L>(czero<C>());
            continue;
        }
        const cplx<C> D0 = mk<C>(kxt[4 * i + 1], kxt[4 * i + 2]);
        const cplx<C> D1 = mk<C>(kyt[4 * j + 1], kyt[4 * j + 2]);
        const cplx<C> D2 = mk<C>(kzt[4 * k + 1], kzt[4 * k + 2]);
        const cplx<C> sdiv = cadd(cadd(cmul(D0, w[0]), cmul(D1, w[1])), cmul(D2, w[2]));
        const cplx<C> qv = mk<C>(sdiv.x / lam, sdiv.y / lam);
        const cplx<C> G[3] = {mk<C>(-D0.x, D0.y), mk<C>(-D1.x, D1.y), mk<C>(-D2.x, D2.y)};
#pragma unroll
        for (int d = 0; d < 3; d++) a[e * S + d * Q + q] = ccast<L>(cscale(csub(w[d], cmul(G[d], qv)), invN));
    }
    __syncthreads();
    fft_run<C, L, N, S, 3 * Q, NT, false, false>(a, tw, +1, nold, nost);
    for (int idx = threadIdx.x; idx < 3 * Q * N; idx += NT) {
        const int s = idx / N, e = idx - s * N;
        const int d = s / Q, q = s - d * Q;
        if (q < jv) Cs[(((long long)d * xc + kx) * ny + j0 + q) * N + e] = ccast<T>(a[e * S + s]);
    }
}

// ------------------------------------------------------------------ Xi
// Workgroup: component d, plane z, rows y0 .. y0+RB-1.  Z_e = A_e + i B_e
// with the Hermitian extension; only the real parts of the DC and Nyquist
// coefficients enter (c2r semantics).  Writes u^{n+1}.
template <typename T, int N>
__device__ __forceinline__ void xi_body(const RepDev &R, long long item, cplx<C> *tw,
                                        cplx<typename FastCfg<T>::L> *a) {
    using L = typename FastCfg<T>::L;
    constexpr int RB = FastCfg<T>::RB, P = RB / 2, S = P + 1, NT = FastCfg<T>::NT, XC = N / 2 + 1;
    // the twiddles are written to LDS after the data loads are issued
    TwRegs<N, NT> twr;
    twr.load(R.twd[0]);
    constexpr int SS = RB + 1;
    const int ny = R.n[1], nz = R.n[2];
    const int nyb = (ny + RB - 1) / RB;
    const int yb = (int)(item % nyb);
    const long long dz = item / nyb;
    const int z = (int)(dz % nz), d = (int)(dz / nz);
    const int y0 = yb * RB;
    cplx<L> *stg = a + N * S;  // staging [kx][rl], stride SS
    const cplx<T> *in = (const cplx<T> *)R.spec + ((long long)d * nz + z) * XC * ny;
    copy_batched<NT, XC * RB, 16>(
        [&](int idx) {
            const int k = idx / RB, rl = idx - k * RB;
            const cplx<T> v = in[(long long)k * ny + min(y0 + rl, ny - 1)];
            return y0 + rl < ny ? v : czero<T>();
        },
        [&](int idx, cplx<T> v) {
            const int k = idx / RB, rl = idx - k * RB;
            stg[k * SS + rl] = ccast<L>(v);
        });
    twr.store(tw);
    __syncthreads();
    for (int idx = threadIdx.x; idx < P * N; idx += NT) {
        const int p = idx / N, e = idx - p * N;
        const int src = e < XC ? e : N - e;
        cplx<L> A = stg[src * SS + 2 * p], B = stg[src * SS + 2 * p + 1];
        if (e == 0 || 2 * e == N) {
            A.y = (L)0;
            B.y = (L)0;
        }
        if (e >= XC) {
            A.y = -A.y;
            B.y = -B.y;
        }
        a[e * S + p] = mk<L>(A.x - B.y, A.y + B.x);
    }
    __syncthreads();
    auto nold = [](int, int) { return cplx<C>{}; };
    auto nost = [](int, int, cplx<C>) {};
    fft_run<C, L, N, S, P, NT, false, false>(a, tw, +1, nold, nost);
    T *out = (T *)R.u + ((long long)d * nz + z) * ny * N;
    // the step is complete once this pass ends: advance the replica's counter
    // (no kernel that runs from here to the end of the step reads it)
    if (item == 0 && threadIdx.x == 0) R.ts[0] += 1;
    for (int idx = threadIdx.x; idx < RB * N; idx += NT) {
        const int rl = idx / N, e = idx - rl * N;
        const int y = y0 + rl;
        if (y >= ny) continue;
        const cplx<L> v = a[e * S + (rl >> 1)];
        out[(long long)y * N + e] = (T)((rl & 1) ? v.y : v.x);
    }
}

// ------------------------------------------------------------------ plane path
// When the x and y edges of every replica are equal (N <= 128) and a whole
// xy-plane fits in LDS, the X and Y transforms of a plane run in one
// workgroup and the step takes three passes instead of five:
//
//   pass   reads                      transform               writes
//   XYf    u^n plane z, u*            r2c along x, DFT y      P[d][z][kx][ky]
//   Zq     P (Q consecutive ky per    DFT along z, thermal    P in place
//          (d, z)), thermal forcing   forcing, projection,
//                                     k=0 removal, 1/N, iDFT
//   XYi    P plane z                  iDFT y, c2r along x     u^{n+1} plane z
//
// P lives in spec; the thermal forcing nbuf is stored as [d][kz][kx][ky] to
// match.  The step runs as one chain of kernels on one queue: the spread
// launch draws the thermal forcing in workgroups of its own beside the tiles,
// and, as a 128^3 grid has 384 plane workgroups (three quarters of the 512
// that two per CU hold), the XYi launch evaluates the bond forces at x^{n+1}
// for the next step's bead update in auxiliary workgroups in the free slots
// (systems without angle or lj/cut terms; the others keep their own kernel).
// Both XY passes keep the plane in one LDS buffer of N x (N/2+1)
// elements (element-major, stride N/2+1): the u plane is staged in it, u* is
// formed into registers and written back as N/2 complex sequences (rows 2p
// and 2p+1 as real and imaginary parts), transformed along x, unpacked into
// the half spectrum of every row, and transformed along y.  The x-frequency
// columns kx = 0 and kx = N/2 of a real plane are real sequences in y, so
// they share one complex sequence (column 0 holds X(0) + i X(N/2)) and the Y
// transform runs on N/2 sequences; the two are separated on output.
template <int NMAX>
struct PlaneNT {
    static constexpr int v = NMAX <= 64 ? 256 : 512;
};
// The XY kernels are built for FP32-bulk only: two 512-thread workgroups per
// CU (a 128-edge plane is 66.5 KB of LDS) need <= 128 VGPRs, i.e. 4 waves per
// SIMD.  An FP64 plane takes 133 KB and its FP64 staging does not fit the
// register budget; FP64 keeps the transposing passes.

// Elements [0, COUNT) of ld, one thread's share (i = tid + k NT) held in
// registers as V, then handed to st after the workgroup barrier: for
// rewriting an LDS buffer in place under a new index map.
template <int NT, int COUNT, class V, class LD, class ST>
__device__ __forceinline__ void lds_remap(const LD &ld, const ST &st) {
    constexpr int IT = (COUNT + NT - 1) / NT;
    V buf[IT];
#pragma unroll
    for (int k = 0; k < IT; k++) {
        const int i = (int)threadIdx.x + k * NT;
        if (i < COUNT) buf[k] = ld(i);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < IT; k++) {
        const int i = (int)threadIdx.x + k * NT;
        if (i < COUNT) st(i, buf[k]);
    }
    __syncthreads();
}

template <typename T, int N, int NT, bool WT>
__device__ __forceinline__ void xyf_body(const RepDev &R, long long item, cplx<C> *twx,
                                         cplx<C> *twy, int *okf, int2 *tab, cplx<typename FastCfg<T>::L> *a) {
    using L = typename FastCfg<T>::L;
    static_assert(sizeof(T) == 4 && sizeof(L) == 4, "the plane path is FP32-bulk");
    constexpr int P = N / 2, S = P + 1, NN = N * N, IT = (NN + NT - 1) / NT;
    const int nz = R.n[2];
    const int z = (int)(item % nz), d = (int)(item / nz);
    const int step = (int)(R.ts[0] + 1);
    const SELM_GLOBAL T *ud = gbl((const T *)R.u) + d * R.N;
    T *U = (T *)a;  // plane z of u_d, [y][x]

    // The x and y twiddles and the stamps of the tile row under plane z are
    // loaded first and written to LDS after the plane's loads are issued, so
    // that no load waits for another.  The pool addressing is split into a
    // row part and a column part: the spread value of cell (x, y) is
    // pool[row(y).x + col(x).x], valid when okf[row(y).y + col(x).y].
    TwRegs<N, NT> twxr, twyr;
    twxr.load(R.twd[0]);
    twyr.load(R.twd[1]);
    const int te0 = R.te[0], te1 = R.te[1], te2 = R.te[2], tn0 = R.tn[0], tn1 = R.tn[1];
    const int tz = z / te2;
    const int tcells = te0 * te1 * te2;
    const int ntr = tn0 * tn1;
    const SELM_GLOBAL int *stamp = gbl((const int *)R.tstamp) + tz * tn0 * tn1;
    // tiles of at least 8 cells per edge: at most OKI stamps per thread
    constexpr int OKI = (((N + 7) / 8) * ((N + 7) / 8) + NT - 1) / NT;
    int okv[OKI];
#pragma unroll
    for (int k = 0; k < OKI; k++) okv[k] = stamp[min((int)threadIdx.x + k * NT, ntr - 1)];
    int2 *rowt = tab, *colt = tab + N;
    for (int i = threadIdx.x; i < N; i += NT) {
        const int ty = i / te1, tx = i / te0;
        rowt[i] = make_int2(((tz * tn1 + ty) * tn0 * 3 + d) * tcells + ((z - tz * te2) * te1 + (i - ty * te1)) * te0,
                            ty * tn0);
        colt[i] = make_int2(tx * 3 * tcells + (i - tx * te0), tx);
    }
    const SELM_GLOBAL T *uz = ud + (long long)z * NN;
    // the plane in 16-byte pieces: N is even, so a plane and its offset in u
    // are whole float4s, and U starts 16-byte aligned in LDS
    static_assert(NN % 4 == 0, "float4 staging");
    constexpr int NQ = NN / 4, IQ = (NQ + NT - 1) / NT;
    {
        typedef float f32x4 __attribute__((ext_vector_type(4)));
        const SELM_GLOBAL f32x4 *uz4 = (const SELM_GLOBAL f32x4 *)uz;
        f32x4 *U4 = (f32x4 *)U;
        copy_batched<NT, NQ, (IQ < 16 ? IQ : 16)>([&](int idx) -> f32x4 { return uz4[idx]; },
                                                   [&](int idx, f32x4 v) { U4[idx] = v; });
    }
    twxr.store(twx);
    twyr.store(twy);
#pragma unroll
    for (int k = 0; k < OKI; k++)
        if ((int)threadIdx.x + k * NT < ntr) okf[threadIdx.x + k * NT] = okv[k] == step;
    for (int i = threadIdx.x + OKI * NT; i < ntr; i += NT) okf[i] = stamp[i] == step;  // smaller tiles
    __syncthreads();

    // u* (the arithmetic of ustar_at, cuda/SELM_Kernel.cu:155-195) of this
    // thread's elements, in registers; the z neighbours and the spread pool
    // are loaded BT elements at a time.  u* = u + c_lap * (sum of the six
    // neighbours - 6u) + c_f * f is evaluated in FP32: the Laplacian's terms
    // are O(u) and enter scaled by mu dt / (rho dx^2) (0.01 at the headline
    // workload), so the FP32 sums add well under one FP32 rounding of u*.
    const SELM_GLOBAL T *up = ud + (long long)(z + 1 < nz ? z + 1 : 0) * NN;
    const SELM_GLOBAL T *um = ud + (long long)(z > 0 ? z - 1 : nz - 1) * NN;
    const SELM_GLOBAL T *pool = gbl((const T *)R.pool);
    const float c_lap = (float)(R.mu * R.dt / (R.rho * R.dx * R.dx)), c_f = (float)(1.0 / R.rho);
    constexpr int BT = IT < 4 ? IT : 4;
    L us[IT];
    if constexpr (NT % N == 0 && NN % NT == 0) {
        // Each thread keeps one column x and takes rows y0 + RS k: the column
        // terms are computed once and every LDS and global address is the
        // thread's base plus a constant.
        constexpr int RS = NT / N;
        const int x = (int)threadIdx.x % N, y0 = (int)threadIdx.x / N;
        const int xp = x + 1 < N ? x + 1 : 0, xm = x > 0 ? x - 1 : N - 1;
        const int2 ct = colt[x];
#pragma unroll
        for (int k0 = 0; k0 < IT; k0 += BT) {
            XfLoads<T> v[BT];
            int ok[BT];
#pragma unroll
            for (int j = 0; j < BT; j++) {
                const int k = k0 + j, y = y0 + RS * k, idx = (int)threadIdx.x + NT * k;
                const int2 rt = rowt[y];
                v[j].zp = up[idx];
                v[j].zm = um[idx];
                // only the tiles a bead touched this step are read: another
                // cell's load reads pool element 0 (one cached line), and its
                // value is replaced by 0 (no branch in the batch)
                ok[j] = okf[rt.y + ct.y];
                v[j].f = pool[ok[j] ? rt.x + ct.x : 0];
            }
#pragma unroll
            for (int j = 0; j < BT; j++) {
                const int k = k0 + j, y = y0 + RS * k;
                const T *row = U + y * N;
                // rows y +- 1 wrap only for the first and last k
                const float uyp = k < IT - 1 ? row[N + x] : U[(y + 1 < N ? y + 1 : 0) * N + x];
                const float uym = k > 0 ? row[x - N] : U[(y > 0 ? y - 1 : N - 1) * N + x];
                const float uc = row[x];
                // neighbour order x+, x-, y+, y-, z+, z- (as ustar_at)
                const float nb = (((((row[xp] + row[xm]) + uyp) + uym) + v[j].zp) + v[j].zm) - 6.0f * uc;
                us[k] = uc + c_lap * nb + c_f * (ok[j] ? v[j].f : 0.0f);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();  // every thread is done with U
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const int y = y0 + RS * k;
            ((L *)&a[x * S + (y >> 1)])[y & 1] = us[k];
        }
    } else {
#pragma unroll
    for (int k0 = 0; k0 < IT; k0 += BT) {
        // opaque to the optimiser, so that the index arithmetic of a batch is
        // done in that batch rather than hoisted (and spilled) for all of them
        int base = (int)threadIdx.x + k0 * NT;
        asm volatile("" : "+v"(base));
        XfLoads<T> v[BT];
        int ok[BT];
#pragma unroll
        for (int j = 0; j < BT; j++) {
            const int idx = min(base + j * NT, NN - 1);
            const int y = idx / N, x = idx - y * N;
            const int2 rt = rowt[y], ct = colt[x];
            v[j].zp = up[idx];
            v[j].zm = um[idx];
            // only the tiles a bead touched this step are read (as above)
            ok[j] = okf[rt.y + ct.y];
            v[j].f = pool[ok[j] ? rt.x + ct.x : 0];
        }
#pragma unroll
        for (int j = 0; j < BT; j++) {
            const int idx = base + j * NT;
            if (k0 + j < IT && idx < NN) {
                const int y = idx / N, x = idx - y * N;
                // neighbour order x+, x-, y+, y-, z+, z- (as ustar_at)
                const T *row = U + y * N;
                const float uc = row[x];
                const float nb = (((((row[x + 1 < N ? x + 1 : 0] + row[x > 0 ? x - 1 : N - 1]) +
                                     U[(y + 1 < N ? y + 1 : 0) * N + x]) +
                                    U[(y > 0 ? y - 1 : N - 1) * N + x]) +
                                   v[j].zp) +
                                  v[j].zm) -
                                 6.0f * uc;
                us[k0 + j] = uc + c_lap * nb + c_f * (ok[j] ? v[j].f : 0.0f);
            }
        }
    }
    __syncthreads();  // every thread is done with U
#pragma unroll
    for (int k = 0; k < IT; k++) {
        const int idx = (int)threadIdx.x + k * NT;
        if (idx < NN) {
            const int y = idx / N, x = idx - y * N;
            ((L *)&a[x * S + (y >> 1)])[y & 1] = us[k];
        }
    }
    }
    __syncthreads();
    auto nold = [](int, int) { return cplx<C>{}; };
    auto nost = [](int, int, cplx<C>) {};
    fft_run<C, L, N, S, P, NT, false, false>(a, twx, -1, nold, nost);

    // unpack: element y of sequence s (1 <= s < P) = X_y(s); sequence 0 =
    // X_y(0) + i X_y(P), both real.  Z_p(k) = X_2p(k) + i X_2p+1(k).
    if constexpr (NT % P == 0 && (N * P) % NT == 0 && (NT / P) % 2 == 0) {
        // s = tid % P and the parity of y are the same for all of a
        // thread's elements: the case split is made once per thread
        constexpr int IT2 = N * P / NT, YS = NT / P;
        const int s = (int)threadIdx.x % P, y0 = (int)threadIdx.x / P;
        const bool s0 = s == 0, odd = (y0 & 1) != 0;
        const int c1 = s, c2 = s0 ? P : N - s;
        cplx<L> buf[IT2];
#pragma unroll
        for (int k = 0; k < IT2; k++) {
            const int p = (y0 + YS * k) >> 1;
            const cplx<L> zk = a[c1 * S + p], zm = a[c2 * S + p];
            // in FP32: a sum of two FP32 values halved is rounded once, as in FP64
            const L hx = odd ? (zk.y + zm.y) * (L)0.5 : (zk.x + zm.x) * (L)0.5;
            const L hy = odd ? (zm.x - zk.x) * (L)0.5 : (zk.y - zm.y) * (L)0.5;
            buf[k] = s0 ? (odd ? mk<L>(zk.y, zm.y) : mk<L>(zk.x, zm.x)) : mk<L>(hx, hy);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < IT2; k++) a[(y0 + YS * k) * S + s] = buf[k];
        __syncthreads();
    } else
    lds_remap<NT, N * P, cplx<L>>(
        [&](int i) {
            const int s = i % P, y = i / P, p = y >> 1;
            if (s == 0) {
                const cplx<L> z0 = a[p], zn = a[P * S + p];
                return (y & 1) ? mk<L>(z0.y, zn.y) : mk<L>(z0.x, zn.x);
            }
            // in FP32: a sum of two FP32 values halved is rounded once, as in FP64
            const cplx<L> zk = a[s * S + p];
            const cplx<L> zm = a[(N - s) * S + p];
            return (y & 1) ? mk<L>((zk.y + zm.y) * (L)0.5, (zm.x - zk.x) * (L)0.5)
                           : mk<L>((zk.x + zm.x) * (L)0.5, (zk.y - zm.y) * (L)0.5);
        },
        [&](int i, cplx<L> v) {
            const int s = i % P, y = i / P;
            a[y * S + s] = v;
        });
    fft_run<C, L, N, S, P, NT, false, false>(a, twy, -1, nold, nost);

    // P[d][z][kx][ky]; the kx = 0 and kx = P columns are separated from
    // sequence 0: V(ky) = Y0(ky) + i YP(ky) with Y0, YP transforms of real
    // sequences, so Y0 = (V(ky) + conj V(-ky))/2, YP = (V(ky) - conj V(-ky))/(2i)
    SELM_GLOBAL cplx<T> *out = gbl((cplx<T> *)R.spec) + ((long long)d * nz + z) * (P + 1) * N;
    for (int idx = threadIdx.x; idx < (P + 1) * N; idx += NT) {
        const int kx = idx / N, ky = idx - kx * N;
        cplx<L> o;
        if (kx == 0 || kx == P) {
            const cplx<L> V = a[ky * S];
            const cplx<L> Vm = a[((N - ky) % N) * S];
            o = kx == 0 ? mk<L>((V.x + Vm.x) * (L)0.5, (V.y - Vm.y) * (L)0.5)
                        : mk<L>((V.y + Vm.y) * (L)0.5, (Vm.x - V.x) * (L)0.5);
        } else {
            o = a[ky * S + kx];
        }
        if constexpr (WT) stc_dev(out + idx, o);
        else stc(out + idx, o);
    }
}

// Z pass on the plane layout.  Workgroup: Q consecutive (kx, ky) columns in
// the flat order c = kx * ny + ky, all three components; for each (d, z)
// the Q values are consecutive in memory.  Two shapes are built, chosen per
// batch (selm_autotune times both): Q = 17 with 512 threads (a 128^3 grid
// has 490 workgroups, within the 512 that two per CU hold at once; 16
// columns would make 520, and the last 8 would run after the others) and
// Q = 9 with 256 threads (925 workgroups, four per CU).  The arithmetic is
// that of z_body (thermal forcing, projection, k=0 removal, 1/N).
template <int Q>
struct ZqShape;
template <>
struct ZqShape<17> {
    static constexpr int NT = 512;
};
template <>
struct ZqShape<9> {
    static constexpr int NT = 256;
};

// Dynamic LDS: the z twiddles (twmax slots), the Q columns of the three
// components (a, (3Q+1) N elements), then the wavenumber tables of the
// workgroup's modes (zt, 5Q + 3N doubles): per column q {lam_x + lam_y,
// Dre_x, Dim_x, Dre_y, Dim_y}, per kz {lam_z, Dre_z, Dim_z}.  The twiddles
// and tables are loaded before the columns and written to LDS after the
// column loads are issued, so that no load waits for another.
template <typename T, int N, int Q, bool WT>
__device__ __forceinline__ void zq_body(const RepDev &R, long long item, cplx<C> *tw,
                                        cplx<typename FastCfg<T>::L> *a) {
    using L = typename FastCfg<T>::L;
    constexpr int S = 3 * Q + 1, NT = ZqShape<Q>::NT;
    constexpr int IP = (N * Q + NT - 1) / NT;  // projection elements per thread
    constexpr int NTAB = 5 * Q + 3 * N, TWI = (N + NT - 1) / NT, TBI = (NTAB + NT - 1) / NT;
    const int ny = R.n[1], xc = R.xc;
    const int pl = xc * ny;  // columns of one (d, z) plane of P
    const int c0 = (int)item * Q;
    const int cv = min(Q, pl - c0);
    SELM_GLOBAL cplx<T> *Ps = gbl((cplx<T> *)R.spec) + c0;
    C *zt = (C *)(a + S * N);
    const float rny = 1.0f / (float)ny;
    cplx<C> twv[TWI];
    {
        const SELM_GLOBAL cplx<C> *tws = gbl((const cplx<C> *)R.twd[2]);
#pragma unroll
        for (int k = 0; k < TWI; k++) twv[k] = ldc(tws + min((int)threadIdx.x + k * NT, N - 1));
    }
    C tbv[TBI];
    {
        const SELM_GLOBAL C *kxt = gbl((const C *)R.ktabd[0]), *kyt = gbl((const C *)R.ktabd[1]),
                           *kzt = gbl((const C *)R.ktabd[2]);
#pragma unroll
        for (int k = 0; k < TBI; k++) {
            const int v = min((int)threadIdx.x + k * NT, NTAB - 1);
            if (v < 5 * Q) {
                const int q = v / 5, f = v - 5 * q;
                const int c = c0 + min(q, cv - 1);
                const int i = div_small(c, ny, rny), j = c - i * ny;
                if (f == 0) tbv[k] = kxt[4 * i] + kyt[4 * j];
                else tbv[k] = f < 3 ? kxt[4 * i + f] : kyt[4 * j + f - 2];
            } else {
                const int w = v - 5 * Q, kz = w / 3;
                tbv[k] = kzt[4 * kz + (w - 3 * kz)];
            }
        }
    }
    // sequence s = d*Q + q <-> column c0+q of component d; element e = z
    copy_batched<NT, 3 * Q * N, 16>(
        [&](int idx) {
            const int q = idx % Q, dz = idx / Q;
            const cplx<T> v = ldc(Ps + (long long)dz * pl + min(q, cv - 1));
            return q < cv ? v : czero<T>();
        },
        [&](int idx, cplx<T> v) {
            const int q = idx % Q, dz = idx / Q;
            const int d = dz / N, e = dz - d * N;
            a[e * S + d * Q + q] = ccast<L>(v);
        });
#pragma unroll
    for (int k = 0; k < TWI; k++)
        if ((int)threadIdx.x + k * NT < N) tw[threadIdx.x + k * NT] = twv[k];
#pragma unroll
    for (int k = 0; k < TBI; k++)
        if ((int)threadIdx.x + k * NT < NTAB) zt[threadIdx.x + k * NT] = tbv[k];
    __syncthreads();
    auto nold = [](int, int) { return cplx<C>{}; };
    auto nost = [](int, int, cplx<C>) {};
    fft_run<C, L, N, S, 3 * Q, NT, false, false>(a, tw, -1, nold, nost);
    // the forcing of this thread's modes, all loads in flight together (held
    // through the forward transform they would not fit the register budget)
    cplx<T> nv[IP][3];
    if (R.noise) {
        const SELM_GLOBAL cplx<T> *nb = gbl((const cplx<T> *)R.nbuf) + c0;
#pragma unroll
        for (int k = 0; k < IP; k++) {
            const int idx = min((int)threadIdx.x + k * NT, N * Q - 1);
            const int q = min(idx % Q, cv - 1), e = idx / Q;
#pragma unroll
            for (int d = 0; d < 3; d++) nv[k][d] = ldc(nb + d * R.Nc + (long long)e * pl + q);
        }
    }

    const C invN = 1.0 / (double)R.N;
#pragma unroll
    for (int kk = 0; kk < IP; kk++) {
        const int idx = (int)threadIdx.x + kk * NT;
        if (idx >= N * Q) continue;
        const int q = idx % Q, e = idx / Q;
        if (q >= cv) continue;
        const int c = c0 + q, k = e;
        cplx<C> w[3];
#pragma unroll
        for (int d = 0; d < 3; d++) w[d] = ccast<C>(a[e * S + d * Q + q]);
        const C *zq = zt + 5 * q, *zk = zt + 5 * Q + 3 * k;
        const C lam = zq[0] + zk[0];
        if (R.noise) {
#pragma unroll
            for (int d = 0; d < 3; d++) w[d] = cadd(w[d], ccast<C>(nv[kk][d]));
        }
        if (c == 0 && k == 0) {
            if (R.diag) {
#pragma unroll
                for (int d = 0; d < 3; d++) R.diag[d] = (double)w[d].x;
            }
#pragma unroll
            for (int d = 0; d < 3; d++) a[e * S + d * Q + q] = ccast<L>(czero<C>());
            continue;
        }
        const cplx<C> D0 = mk<C>(zq[1], zq[2]);
        const cplx<C> D1 = mk<C>(zq[3], zq[4]);
        const cplx<C> D2 = mk<C>(zk[1], zk[2]);
        const cplx<C> sdiv = cadd(cadd(cmul(D0, w[0]), cmul(D1, w[1])), cmul(D2, w[2]));
        const cplx<C> qv = mk<C>(sdiv.x / lam, sdiv.y / lam);
        const cplx<C> G[3] = {mk<C>(-D0.x, D0.y), mk<C>(-D1.x, D1.y), mk<C>(-D2.x, D2.y)};
#pragma unroll
        for (int d = 0; d < 3; d++) a[e * S + d * Q + q] = ccast<L>(cscale(csub(w[d], cmul(G[d], qv)), invN));
    }
    __syncthreads();
    fft_run<C, L, N, S, 3 * Q, NT, false, false>(a, tw, +1, nold, nost);
    for (int idx = threadIdx.x; idx < 3 * Q * N; idx += NT) {
        const int q = idx % Q, dz = idx / Q;
        const int d = dz / N, e = dz - d * N;
        if (q < cv) {
            if constexpr (WT) stc_dev(Ps + (long long)dz * pl + q, ccast<T>(a[e * S + d * Q + q]));
            else stc(Ps + (long long)dz * pl + q, ccast<T>(a[e * S + d * Q + q]));
        }
    }
}

template <typename T, int N, int NT, bool WT>
__device__ __forceinline__ void xyi_body(const RepDev &R, long long item, cplx<C> *twx,
                                         cplx<C> *twy, cplx<typename FastCfg<T>::L> *a) {
    using L = typename FastCfg<T>::L;
    constexpr int P = N / 2, S = P + 1, NN = N * N;
    const int nz = R.n[2];
    const int z = (int)(item % nz), d = (int)(item / nz);
    const SELM_GLOBAL cplx<T> *in = gbl((const cplx<T> *)R.spec) + ((long long)d * nz + z) * (P + 1) * N;  // [kx][ky]
    constexpr int NI = (P - 1) * N, ITI = (NI + NT - 1) / NT, I0 = (N + NT - 1) / NT;
    // the twiddles are loaded first and written to LDS after the spectrum's
    // loads are issued
    TwRegs<N, NT> twxr, twyr;
    twxr.load(R.twd[0]);
    twyr.load(R.twd[1]);
    // sequence 0 = Y0h + i YPh, the Hermitian parts of the kx = 0 and kx = P
    // columns (what a c2r keeps of them), whose inverses are real; its loads
    // are issued first and complete under the batched copy below
    cplx<T> c0[I0][4];
#pragma unroll
    for (int k = 0; k < I0; k++) {
        const int ky = min((int)threadIdx.x + k * NT, N - 1), km = (N - ky) % N;
        c0[k][0] = ldc(in + ky);
        c0[k][1] = ldc(in + km);
        c0[k][2] = ldc(in + P * N + ky);
        c0[k][3] = ldc(in + P * N + km);
    }
    // columns 1 <= kx < P: sequence kx of the inverse Y transform
    copy_batched<NT, NI, (ITI < 16 ? ITI : 16)>([&](int idx) { return ldc(in + N + idx); },
                                              [&](int idx, cplx<T> v) {
                                                  const int kx = 1 + idx / N, ky = idx % N;
                                                  a[ky * S + kx] = ccast<L>(v);
                                              });
#pragma unroll
    for (int k = 0; k < I0; k++) {
        const int ky = (int)threadIdx.x + k * NT;
        if (ky < N) {
            const cplx<T> A = c0[k][0], Am = c0[k][1], B = c0[k][2], Bm = c0[k][3];
            const cplx<L> A0 = mk<L>((A.x + Am.x) * (L)0.5, (A.y - Am.y) * (L)0.5);
            const cplx<L> B0 = mk<L>((B.x + Bm.x) * (L)0.5, (B.y - Bm.y) * (L)0.5);
            a[ky * S] = mk<L>(A0.x - B0.y, A0.y + B0.x);
        }
    }
    twxr.store(twx);
    twyr.store(twy);
    __syncthreads();
    auto nold = [](int, int) { return cplx<C>{}; };
    auto nost = [](int, int, cplx<C>) {};
    fft_run<C, L, N, S, P, NT, false, false>(a, twy, +1, nold, nost);

    // pack rows 2p, 2p+1 as Z_e = A_e + i B_e with the Hermitian extension
    // along x; element y of sequence 0 is X_y(0) + i X_y(P)
    if constexpr (NT % N == 0 && (N * P) % NT == 0) {
        // e = tid % N is the same for all of a thread's elements: column,
        // conjugation and the DC/Nyquist case are chosen once per thread
        constexpr int IT2 = N * P / NT, PS = NT / N;
        const int e = (int)threadIdx.x % N, p0 = (int)threadIdx.x / N;
        const bool edge = e == 0 || e == P, nyq = e == P;
        const int col = edge ? 0 : (e < P ? e : N - e);
        const L sy = edge ? (L)0 : (e > P ? (L)-1 : (L)1);  // imaginary part kept, conjugated or dropped
        cplx<L> buf[IT2];
#pragma unroll
        for (int k = 0; k < IT2; k++) {
            const int p = p0 + PS * k;
            const cplx<L> W0 = a[(2 * p) * S + col], W1 = a[(2 * p + 1) * S + col];
            const L ax = nyq ? W0.y : W0.x, ay = sy * W0.y;
            const L bx = nyq ? W1.y : W1.x, by = sy * W1.y;
            buf[k] = mk<L>(ax - by, ay + bx);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < IT2; k++) a[e * S + p0 + PS * k] = buf[k];
        __syncthreads();
    } else
    lds_remap<NT, N * P, cplx<L>>(
        [&](int i) {
            const int e = i % N, p = i / N;  // consecutive lanes read consecutive elements
            const int y0 = 2 * p, y1 = 2 * p + 1;
            cplx<L> A, B;
            if (e == 0 || e == P) {
                const cplx<L> W0 = a[y0 * S], W1 = a[y1 * S];
                A = mk<L>(e == 0 ? W0.x : W0.y, (L)0);
                B = mk<L>(e == 0 ? W1.x : W1.y, (L)0);
            } else if (e < P) {
                A = a[y0 * S + e];
                B = a[y1 * S + e];
            } else {
                A = cconj(a[y0 * S + N - e]);
                B = cconj(a[y1 * S + N - e]);
            }
            return mk<L>(A.x - B.y, A.y + B.x);
        },
        [&](int i, cplx<L> v) {
            const int e = i % N, p = i / N;
            a[e * S + p] = v;
        });
    fft_run<C, L, N, S, P, NT, false, false>(a, twx, +1, nold, nost);
    SELM_GLOBAL T *out = gbl((T *)R.u) + ((long long)d * nz + z) * NN;
    // the step is complete once this pass ends: advance the replica's counter
    // (no kernel that runs from here to the end of the step reads it)
    if (item == 0 && threadIdx.x == 0) R.ts[0] += 1;
    for (int idx = threadIdx.x; idx < NN; idx += NT) {
        const int y = idx / N, x = idx - y * N;
        const cplx<L> v = a[x * S + (y >> 1)];
        if constexpr (WT) st_dev(gbl(out + idx), (T)((y & 1) ? v.y : v.x));
        else out[idx] = (T)((y & 1) ? v.y : v.x);
    }
}

// ------------------------------------------------------------------ thermal forcing
// One mode of the thermal forcing: its three components into nbuf.
// k_noise: modes in the row order of C ([kx][ky][kz], kz fastest); with plane = 1 in
// the order of the plane path's P ([kz][kx][ky], ky fastest).  Runs on a
// forked stream beside the spread and the first FFT pass.  The grid is
// capped (kNoiseBlocks) and each thread strides over the modes, so the
// forcing holds a few workgroup slots per CU for longer instead of taking
// every slot at once while the spread's workgroups wait for them.
constexpr long long kNoiseBlocks = 1024;  // 4 per CU of the MI355X's 256
template <typename T>
__global__ __launch_bounds__(256) void k_noise(const RepDev *__restrict__ reps, int nrep,
                                               const long long *__restrict__ start, long long total,
                                               int plane) {
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (long long)gridDim.x * blockDim.x)
        noise_mode<T>(reps, nrep, start, g, plane);
}

// ------------------------------------------------------------------ kernels
// Each kernel is instantiated per size class NMAX (64, 128, 256) and holds
// only the lengths up to NMAX.  The register budget is then set by the
// largest length the batch can hold, not by 256.
#define SELM_FFT_KERNEL(kname, axis, BODY, NTH, ...)                                                  \
    template <typename T, int NMAX, bool WT = false>                                                  \
    __global__ __launch_bounds__(FastCfg<T>::NTH) __VA_ARGS__ void kname(const RepDev *__restrict__ reps, int nrep, \
                                                            const long long *__restrict__ start,      \
                                                            long long total, int twmax) {             \
        extern __shared__ __align__(16) unsigned char smem[];                                         \
        const long long g = xcd_remap(blockIdx.x, gridDim.x);                                         \
        if (g >= total) return;                                                                       \
        const int r = find_rep(start, nrep, g);                                                       \
        const RepDev &R = reps[r];                                                                    \
        const long long item = g - start[r];                                                          \
        cplx<C> *tw = (cplx<C> *)smem;                                                                \
        auto *a = (cplx<typename FastCfg<T>::L> *)(tw + twmax);                                       \
        switch (R.n[axis]) {                                                                          \
            SELM_FFT_FOR_EACH_N(BODY)                                                                 \
            default: break;                                                                           \
        }                                                                                             \
    }

#define SELM_CASE_XF(n) \
    case n: if constexpr (n <= NMAX) xf_body<T, n>(R, item, tw, a); break;
#define SELM_CASE_YF(n) \
    case n: if constexpr (n <= NMAX) y_body<T, n, true>(R, item, tw, a); break;
#define SELM_CASE_YI(n) \
    case n: if constexpr (n <= NMAX) y_body<T, n, false>(R, item, tw, a); break;
#define SELM_CASE_Z(n) \
    case n: if constexpr (n <= NMAX) z_body<T, n>(R, item, tw, a); break;
#define SELM_CASE_XI(n) \
    case n: if constexpr (n <= NMAX) xi_body<T, n>(R, item, tw, a); break;

#define SELM_CASE_ZQ(n) \
    case n: if constexpr (n <= NMAX) zq_body<T, n, 17, WT>(R, item, tw, a); break;
#define SELM_CASE_ZQ9(n) \
    case n: if constexpr (n <= NMAX) zq_body<T, n, 9, WT>(R, item, tw, a); break;

SELM_FFT_KERNEL(k_xf, 0, SELM_CASE_XF, NT)
// 4 waves per SIMD: two 512-thread workgroups per CU
// The Z pass of the plane path fills its own twiddles (after its first loads).
#define SELM_ZQ_KERNEL(kname, BODY, NTH)                                                              \
    template <typename T, int NMAX, bool WT = false>                                                  \
    __global__ __launch_bounds__(FastCfg<T>::NTH) __attribute__((amdgpu_waves_per_eu(4)))             \
    void kname(const RepDev *__restrict__ reps, int nrep, const long long *__restrict__ start,        \
               long long total, int twmax) {                                                          \
        extern __shared__ __align__(16) unsigned char smem[];                                         \
        const long long g = xcd_remap(blockIdx.x, gridDim.x);                                         \
        if (g >= total) return;                                                                       \
        const int r = find_rep(start, nrep, g);                                                       \
        const RepDev &R = reps[r];                                                                    \
        const long long item = g - start[r];                                                          \
        cplx<C> *tw = (cplx<C> *)smem;                                                                \
        auto *a = (cplx<typename FastCfg<T>::L> *)(tw + twmax);                                       \
        switch (R.n[2]) {                                                                             \
            SELM_FFT_FOR_EACH_N(BODY)                                                                 \
            default: break;                                                                           \
        }                                                                                             \
    }
SELM_ZQ_KERNEL(k_zq, SELM_CASE_ZQ, NTZ)
SELM_ZQ_KERNEL(k_zq9, SELM_CASE_ZQ9, NTZ9)

// XY passes of the plane path.  Dynamic LDS: the x and y twiddles (twmax
// slots each), the pool tables, the tile stamps of one plane (okmax ints),
// then the plane.  The bodies write the twiddles and stamps after issuing
// the plane's loads.
// Auxiliary workgroups (plane_block_role) stride over aux_total work items
// (AUX(i) for each) with aux * NT threads.
#define SELM_PLANE_KERNEL(kname, BODY, AUX)                                                           \
    template <typename T, int NMAX, bool WT>                                                          \
    __global__ __launch_bounds__(PlaneNT<NMAX>::v) __attribute__((amdgpu_waves_per_eu(4)))           \
    void kname(const RepDev *__restrict__ reps, int nrep, const long long *__restrict__ start,        \
               long long total, int twmax, int okmax, const long long *__restrict__ aux_start,        \
               long long aux_total, int aux) {                                                        \
        extern __shared__ __align__(16) unsigned char smem[];                                         \
        int ablk;                                                                                     \
        const long long g = plane_block_role(total, aux, &ablk);                                      \
        if (g < 0) {                                                                                  \
            constexpr long long NTA = PlaneNT<NMAX>::v;                                               \
            if (ablk >= 0)                                                                            \
                for (long long i = ablk * NTA + threadIdx.x; i < aux_total; i += aux * NTA) AUX(i);   \
            return;                                                                                   \
        }                                                                                             \
        const int r = find_rep(start, nrep, g);                                                       \
        const RepDev &R = reps[r];                                                                    \
        const long long item = g - start[r];                                                          \
        cplx<C> *twx = (cplx<C> *)smem, *twy = twx + twmax;                                           \
        int2 *tab = (int2 *)(twy + twmax);                                                            \
        int *okf = (int *)(tab + 2 * twmax);                                                          \
        auto *a = (cplx<typename FastCfg<T>::L> *)(okf + okmax);                                      \
        (void)okf;                                                                                    \
        (void)tab;                                                                                    \
        switch (R.n[0]) {                                                                             \
            SELM_FFT_FOR_EACH_N(BODY)                                                                 \
            default: break;                                                                           \
        }                                                                                             \
    }
#define SELM_CASE_XYF(n) \
    case n: if constexpr (n <= NMAX) xyf_body<T, n, PlaneNT<NMAX>::v, WT>(R, item, twx, twy, okf, tab, a); break;
#define SELM_CASE_XYI(n) \
    case n: if constexpr (n <= NMAX) xyi_body<T, n, PlaneNT<NMAX>::v, WT>(R, item, twx, twy, a); break;
#define SELM_AUX_NONE(i) (void)(i)
#define SELM_AUX_BONDS(i) bond_only_forces(reps, nrep, aux_start, i)
SELM_PLANE_KERNEL(k_xyf, SELM_CASE_XYF, SELM_AUX_NONE)
SELM_PLANE_KERNEL(k_xyi, SELM_CASE_XYI, SELM_AUX_BONDS)
SELM_FFT_KERNEL(k_yf, 1, SELM_CASE_YF, NT)
SELM_FFT_KERNEL(k_zp, 2, SELM_CASE_Z, NTZ)
SELM_FFT_KERNEL(k_yi, 1, SELM_CASE_YI, NT)
SELM_FFT_KERNEL(k_xi, 0, SELM_CASE_XI, NT)

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

// workgroups of each pass for one replica
void fft_fast_items(bool fp64, const int n[3], long long *ix, long long *iy, long long *iz) {
    const int RB = fp64 ? FastCfg<double>::RB : FastCfg<float>::RB;
    const int Q = fp64 ? FastCfg<double>::Q : FastCfg<float>::Q;
    const long long xc = n[0] / 2 + 1;
    *ix = 3LL * n[2] * ((n[1] + RB - 1) / RB);
    *iy = 3LL * xc * ((n[2] + RB - 1) / RB);
    *iz = xc * ((n[1] + Q - 1) / Q);
}

// LDS bytes of each pass for edges n (twiddle slots twmax first)
void fft_fast_lds(bool fp64, const int n[3], int twmax, size_t *lx, size_t *ly, size_t *lz) {
    const size_t csz = fp64 ? sizeof(cplx<FastCfg<double>::L>) : sizeof(cplx<FastCfg<float>::L>);
    const size_t twsz = sizeof(cplx<C>);
    const int RB = fp64 ? FastCfg<double>::RB : FastCfg<float>::RB;
    const int Q = fp64 ? FastCfg<double>::Q : FastCfg<float>::Q;
    const size_t xc = n[0] / 2 + 1;
    const size_t xf = (size_t)(RB / 2 + 1) * n[0];
    const size_t xi = xf + xc * (RB + 1);
    const size_t stage = (size_t)(RB + 2) * n[0] * (fp64 ? 8 : 4);  // Xf's rows y0-1 .. y0+RB
    *lx = (size_t)twmax * twsz + std::max(xf * csz + stage, xi * csz);
    *ly = (size_t)twmax * twsz + (size_t)(RB + 1) * n[1] * csz;
    *lz = (size_t)twmax * twsz + (size_t)(3 * Q + 1) * n[2] * csz;
}

static inline unsigned pad8(long long n) { return (unsigned)((n + 7) & ~7LL); }

// Auxiliary workgroups of a plane launch for `work` items at `per` items per
// thread, at most `cap`.
static inline int aux_blocks(long long work, int nt, int per, int cap) {
    if (work <= 0) return 0;
    const long long b = (work + (long long)nt * per - 1) / ((long long)nt * per);
    return (int)std::min<long long>(b, cap);
}

// ---- plane path
// A replica can take the plane path when nx == ny <= 128 (both compile-time
// lengths) and nz has a compile-time length; the XY passes then need the LDS
// below, which the caller checks against the device's per-workgroup limit.
bool fft_plane_shape(const int n[3]) {
    return n[0] == n[1] && n[0] <= 128 && fast_len(n[0]) && fast_len(n[2]);
}

int fft_plane_okf(const int tn[3]) { return (tn[0] * tn[1] + 3) & ~3; }

// LDS bytes of the XY passes and of the Z pass of the plane path
void fft_plane_lds(bool fp64, const int n[3], int twmax, int okmax, int zcols, size_t *lxy, size_t *lz) {
    const size_t csz = fp64 ? sizeof(cplx<FastCfg<double>::L>) : sizeof(cplx<FastCfg<float>::L>);
    const size_t tsz = fp64 ? 8 : 4;
    const size_t N = (size_t)n[0], P = N / 2;
    const int Q = zcols;
    // twiddles x and y, the row and column pool tables, the tile stamps, the plane
    *lxy = 2 * (size_t)twmax * sizeof(cplx<C>) + 2 * (size_t)twmax * 8 + (size_t)okmax * 4 +
           std::max(N * (P + 1) * csz, N * N * tsz);
    // z twiddles, the columns, the wavenumber tables of the workgroup's modes
    *lz = (size_t)twmax * sizeof(cplx<C>) + (size_t)(3 * Q + 1) * n[2] * csz + (size_t)(5 * Q + 3 * n[2]) * 8;
}

// workgroups of the XY passes and of the Z pass for one replica
void fft_plane_items(bool fp64, const int n[3], int zcols, long long *ixy, long long *iz) {
    (void)fp64;
    *ixy = 3LL * n[2];
    *iz = ((long long)(n[0] / 2 + 1) * n[1] + zcols - 1) / zcols;
}

bool fft_plane_zcols_ok(int zcols) { return zcols == 17 || zcols == 9; }

// Allow the XY kernels of the size class their dynamic LDS (above 64 KiB),
// both store shapes.
template <typename T, int NMAX, bool WT>
static hipError_t plane_attr_n(size_t lds) {
    hipError_t e = hipFuncSetAttribute((const void *)k_xyf<T, NMAX, WT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_xyi<T, NMAX, WT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    return e;
}
template <typename T>
static hipError_t plane_attr_t(int nmax, size_t lds) {
    hipError_t e;
    if (nmax <= 64) {
        e = plane_attr_n<T, 64, true>(lds);
        if (e == hipSuccess) e = plane_attr_n<T, 64, false>(lds);
    } else {
        e = plane_attr_n<T, 128, true>(lds);
        if (e == hipSuccess) e = plane_attr_n<T, 128, false>(lds);
    }
    return e;
}

hipError_t fft_plane_prepare(bool fp64, int twmax, size_t lds) {
    return fp64 ? hipErrorNotSupported : plane_attr_t<float>(twmax, lds);
}

template <typename T, int NMAX, bool WT>
static void launch_plane_n(const StepLaunch &L, int stage, hipStream_t st) {
    constexpr int NT = PlaneNT<NMAX>::v;
    switch (stage) {
        case kStageFftX:
            hipLaunchKernelGGL((k_xyf<T, NMAX, WT>), dim3(plane_grid(L.tot_fx, 0)), dim3(NT), L.lds_fx, st, L.reps,
                               L.nrep, L.st_fx, L.tot_fx, L.twmax, L.okmax, nullptr, 0LL, 0);
            break;
        case kStageFftZ:
            if (L.zq_cols == 9)
                hipLaunchKernelGGL((k_zq9<T, NMAX, WT>), dim3(pad8(L.tot_fz)), dim3(FastCfg<T>::NTZ9), L.lds_fz, st,
                                   L.reps, L.nrep, L.st_fz, L.tot_fz, L.twmax);
            else
                hipLaunchKernelGGL((k_zq<T, NMAX, WT>), dim3(pad8(L.tot_fz)), dim3(FastCfg<T>::NTZ), L.lds_fz, st,
                                   L.reps, L.nrep, L.st_fz, L.tot_fz, L.twmax);
            break;
        case kStageFftXi: {
            const long long nb = L.aux_bonds ? L.tot_bead : 0;
            const int aux = aux_blocks(nb, NT, 1, 1 << 20);
            hipLaunchKernelGGL((k_xyi<T, NMAX, WT>), dim3(plane_grid(L.tot_fx, aux)), dim3(NT), L.lds_fx, st, L.reps,
                               L.nrep, L.st_fx, L.tot_fx, L.twmax, L.okmax, L.st_bead, nb, aux);
            break;
        }
        default: break;  // the Y passes are folded into the XY passes
    }
}

template <typename T>
static void launch_plane_t(const StepLaunch &L, int stage, hipStream_t st) {
    if (L.twmax <= 64) {
        if (L.wt) launch_plane_n<T, 64, true>(L, stage, st);
        else launch_plane_n<T, 64, false>(L, stage, st);
    } else {
        if (L.wt) launch_plane_n<T, 128, true>(L, stage, st);
        else launch_plane_n<T, 128, false>(L, stage, st);
    }
}

template <typename T, int NMAX>
static void launch_fast_n(const StepLaunch &L, int stage, hipStream_t st) {
    const int NT = FastCfg<T>::NT;
    switch (stage) {
        case kStageFftX:
            hipLaunchKernelGGL((k_xf<T, NMAX>), dim3(pad8(L.tot_fx)), dim3(NT), L.lds_fx, st, L.reps, L.nrep, L.st_fx,
                               L.tot_fx, L.twmax);
            break;
        case kStageFftY:
            hipLaunchKernelGGL((k_yf<T, NMAX>), dim3(pad8(L.tot_fy)), dim3(NT), L.lds_fy, st, L.reps, L.nrep, L.st_fy,
                               L.tot_fy, L.twmax);
            break;
        case kStageFftZ:
            hipLaunchKernelGGL((k_zp<T, NMAX>), dim3(pad8(L.tot_fz)), dim3(FastCfg<T>::NTZ), L.lds_fz, st, L.reps,
                               L.nrep, L.st_fz, L.tot_fz, L.twmax);
            break;
        case kStageFftYi:
            hipLaunchKernelGGL((k_yi<T, NMAX>), dim3(pad8(L.tot_fy)), dim3(NT), L.lds_fy, st, L.reps, L.nrep, L.st_fy,
                               L.tot_fy, L.twmax);
            break;
        case kStageFftXi:
            hipLaunchKernelGGL((k_xi<T, NMAX>), dim3(pad8(L.tot_fx)), dim3(NT), L.lds_fx, st, L.reps, L.nrep, L.st_fx,
                               L.tot_fx, L.twmax);
            break;
        default: break;
    }
}

// size class from the largest edge of the batch (twmax >= every edge)
template <typename T>
static void launch_fast_t(const StepLaunch &L, int stage, hipStream_t st) {
    if (L.twmax <= 64) launch_fast_n<T, 64>(L, stage, st);
    else if (L.twmax <= 128) launch_fast_n<T, 128>(L, stage, st);
    else launch_fast_n<T, 256>(L, stage, st);
}

void launch_noise(const StepLaunch &L, hipStream_t st) {
    if (L.tot_modes <= 0 || L.plane) return;  // the plane path's spread launch draws the forcing
    const unsigned nb = (unsigned)std::min<long long>((L.tot_modes + 255) / 256, kNoiseBlocks);
    const int plane = L.plane ? 1 : 0;
    if (L.fp64)
        hipLaunchKernelGGL(k_noise<double>, dim3(nb), dim3(256), 0, st, L.reps, L.nrep, L.st_modes, L.tot_modes, plane);
    else
        hipLaunchKernelGGL(k_noise<float>, dim3(nb), dim3(256), 0, st, L.reps, L.nrep, L.st_modes, L.tot_modes, plane);
}

void launch_fft_fast(const StepLaunch &L, int stage, hipStream_t st) {
    if (L.plane) {  // FP32-bulk only
        launch_plane_t<float>(L, stage, st);
        return;
    }
    if (L.fp64) launch_fast_t<double>(L, stage, st);
    else launch_fast_t<float>(L, stage, st);
}

}  // namespace selm
