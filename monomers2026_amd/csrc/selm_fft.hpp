// selm_fft.hpp -- compile-time-specialised batched FFT for CDNA4.
//
// A workgroup transforms NSEQ interleaved sequences of length N (element e of
// sequence s at lds[e*S + s]; the stride S >= NSEQ may be padded against LDS
// bank conflicts).  Each Stockham stage keeps its radix-R
// butterflies in registers: every thread reads all of its inputs, the
// workgroup synchronises, and the outputs are written back in place, so one
// LDS buffer of S*N elements suffices.  The first stage may read its inputs
// straight from global memory and the last may write straight to global
// memory through caller-supplied functors, skipping two LDS round trips.
// N, S and the thread count are template constants, so the index arithmetic
// compiles to shifts and multiplies.
//
// The arithmetic type C and the LDS element type L may differ: with C =
// double and L = float, every butterfly and twiddle product is exact to FP64
// and the data is rounded to FP32 only where it is exchanged through LDS.
#pragma once

#include "selm_device.hpp"

namespace selm {

// ---------------------------------------------------------------------------
// More radix kernels (dft2/3/4/5/8 are in selm_device.hpp)
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void dft6(cplx<T> *v, int dir) {
    // 6 = 2 x 3: n = n1 + 2 n2, k = k2 + 3 k1
    cplx<T> a[3] = {v[0], v[2], v[4]};
    cplx<T> b[3] = {v[1], v[3], v[5]};
    dft3(a, dir);
    dft3(b, dir);
    const T h = (T)0.86602540378443864676;
    b[1] = cmul(b[1], mk<T>((T)0.5, (T)dir * h));
    b[2] = cmul(b[2], mk<T>((T)-0.5, (T)dir * h));
#pragma unroll
    for (int k2 = 0; k2 < 3; k2++) {
        v[k2] = cadd(a[k2], b[k2]);
        v[k2 + 3] = csub(a[k2], b[k2]);
    }
}

template <typename T>
__device__ __forceinline__ cplx<T> w16(int m, int dir) {
    // e^{dir * 2 pi i m / 16}, m in [0, 16)
    const T c[5] = {(T)1.0, (T)0.92387953251128675613, (T)0.70710678118654752440,
                    (T)0.38268343236508977173, (T)0.0};
    const int q = m >> 2, r = m & 3;
    T x, y;
    switch (q & 3) {  // rotate by q quarter turns
        case 0: x = c[r]; y = c[4 - r]; break;
        case 1: x = -c[4 - r]; y = c[r]; break;
        case 2: x = -c[r]; y = -c[4 - r]; break;
        default: x = c[4 - r]; y = -c[r]; break;
    }
    return mk<T>(x, (T)dir * y);
}

template <typename T>
__device__ __forceinline__ void dft16(cplx<T> *v, int dir) {
    // 16 = 4 x 4: n = n1 + 4 n2, k = k2 + 4 k1
#pragma unroll
    for (int n1 = 0; n1 < 4; n1++) dft4(v[n1], v[n1 + 4], v[n1 + 8], v[n1 + 12], dir);
#pragma unroll
    for (int n1 = 1; n1 < 4; n1++)
#pragma unroll
        for (int k2 = 1; k2 < 4; k2++) v[n1 + 4 * k2] = cmul(v[n1 + 4 * k2], w16<T>(n1 * k2, dir));
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) dft4(v[4 * k2], v[4 * k2 + 1], v[4 * k2 + 2], v[4 * k2 + 3], dir);
    // X[k2 + 4 k1] sits at v[4 k2 + k1]: transpose the 4x4 register block
    cplx<T> t[16];
#pragma unroll
    for (int i = 0; i < 16; i++) t[i] = v[i];
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++)
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) v[k2 + 4 * k1] = t[4 * k2 + k1];
}

template <typename T, int R>
__device__ __forceinline__ void dft_r(cplx<T> *v, int dir) {
    if constexpr (R == 2) dft2(v);
    else if constexpr (R == 3) dft3(v, dir);
    else if constexpr (R == 4) dft4(v[0], v[1], v[2], v[3], dir);
    else if constexpr (R == 5) dft5(v, dir);
    else if constexpr (R == 6) dft6(v, dir);
    else if constexpr (R == 8) dft8(v, dir);
    else dft16(v, dir);
}

// ---------------------------------------------------------------------------
// Radix plans for the specialised lengths (stage radices in order)
// ---------------------------------------------------------------------------
template <int N>
struct FftPlan;
#define SELM_FFT_PLAN(n, ...)                          \
    template <>                                        \
    struct FftPlan<n> {                                \
        static constexpr int R[] = {__VA_ARGS__};      \
        static constexpr int NST = sizeof(R) / sizeof(int); \
    };
SELM_FFT_PLAN(16, 16)
SELM_FFT_PLAN(20, 5, 4)
SELM_FFT_PLAN(24, 8, 3)
SELM_FFT_PLAN(30, 6, 5)
SELM_FFT_PLAN(32, 8, 4)
SELM_FFT_PLAN(40, 8, 5)
SELM_FFT_PLAN(48, 16, 3)
SELM_FFT_PLAN(60, 6, 5, 2)
SELM_FFT_PLAN(64, 8, 8)
SELM_FFT_PLAN(80, 16, 5)
SELM_FFT_PLAN(96, 16, 6)
SELM_FFT_PLAN(120, 8, 5, 3)
SELM_FFT_PLAN(128, 16, 8)
SELM_FFT_PLAN(160, 16, 5, 2)
SELM_FFT_PLAN(192, 16, 6, 2)
SELM_FFT_PLAN(256, 16, 16)
#undef SELM_FFT_PLAN

// The lengths with a specialised kernel path.
#define SELM_FFT_FOR_EACH_N(X) \
    X(16) X(20) X(24) X(30) X(32) X(40) X(48) X(60) X(64) X(80) X(96) X(120) X(128) X(160) X(192) X(256)

template <typename A, typename B>
__device__ __forceinline__ cplx<A> ccast(cplx<B> v) {
    return mk<A>((A)v.x, (A)v.y);
}

// One Stockham stage.  Reads element e of sequence s from `ld(e, s)` (global)
// or lds, writes through `st(e, s, v)` (global) or lds.  Arithmetic in C,
// LDS elements of type L, twiddles cplx<C>.
template <typename C, typename L, int N, int S, int NSEQ, int NT, int R, int NS, bool FROM_G, bool TO_G, class LD,
          class ST>
__device__ __forceinline__ void fft_stage(cplx<L> *lds, const cplx<C> *tw, int dir, const LD &ld, const ST &st) {
    constexpr int NB = NSEQ * (N / R);
    constexpr int PER = (NB + NT - 1) / NT;
    constexpr bool FULL = (NB % NT) == 0;
    const int tid = threadIdx.x;
    cplx<C> v[PER][R];
#pragma unroll
    for (int b = 0; b < PER; b++) {
        const int t = tid + b * NT;
        if (FULL || t < NB) {
            const int s = t % NSEQ, j = t / NSEQ;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int e = j + r * (N / R);
                if constexpr (FROM_G) v[b][r] = ld(e, s);
                else v[b][r] = ccast<C>(lds[e * S + s]);
            }
        }
    }
    if constexpr (!FROM_G) __syncthreads();
#pragma unroll
    for (int b = 0; b < PER; b++) {
        const int t = tid + b * NT;
        if (FULL || t < NB) {
            const int s = t % NSEQ, j = t / NSEQ;
            const int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    cplx<C> w = tw[(r * k) * (N / (NS * R))];
                    if (dir > 0) w.y = -w.y;
                    v[b][r] = cmul(v[b][r], w);
                }
            }
            dft_r<C, R>(v[b], dir);
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int e = base + r * NS;
                if constexpr (TO_G) st(e, s, v[b][r]);
                else lds[e * S + s] = ccast<L>(v[b][r]);
            }
        }
    }
    if constexpr (!TO_G) __syncthreads();
}

template <typename C, typename L, int N, int S, int NSEQ, int NT, int I, int NS, bool LDG, bool STG, class LD,
          class ST>
__device__ __forceinline__ void fft_stages(cplx<L> *lds, const cplx<C> *tw, int dir, const LD &ld, const ST &st) {
    using P = FftPlan<N>;
    constexpr int R = P::R[I];
    constexpr bool first = I == 0, last = I == P::NST - 1;
    fft_stage<C, L, N, S, NSEQ, NT, R, NS, first && LDG, last && STG>(lds, tw, dir, ld, st);
    if constexpr (!last) fft_stages<C, L, N, S, NSEQ, NT, I + 1, NS * R, LDG, STG>(lds, tw, dir, ld, st);
}

// Full transform.  LDG: first stage reads through ld (else data is in lds);
// STG: last stage writes through st (else the result is left in lds).
// The caller synchronises before a first stage that reads lds.
template <typename C, typename L, int N, int S, int NSEQ, int NT, bool LDG, bool STG, class LD, class ST>
__device__ __forceinline__ void fft_run(cplx<L> *lds, const cplx<C> *tw, int dir, const LD &ld, const ST &st) {
    fft_stages<C, L, N, S, NSEQ, NT, 0, 1, LDG, STG>(lds, tw, dir, ld, st);
}

struct NoLoad {
    template <typename T>
    __device__ cplx<T> operator()(int, int) const { return cplx<T>{}; }
};

}  // namespace selm
