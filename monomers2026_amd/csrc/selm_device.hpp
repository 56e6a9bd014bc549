// selm_device.hpp -- device-side building blocks for the SELM step on CDNA4.
//
// Complex arithmetic, the Philox4x32-10 counter-based generator,
// radix-2/3/4/5/8 butterflies and a workgroup-cooperative Stockham FFT over
// LDS-resident batches of sequences.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace selm {

template <typename T>
struct cplx {
    T x, y;
};

// Device arrays reached through the replica descriptor are generic pointers
// to the compiler, which then emits flat loads and stores: 64-bit address
// arithmetic per access, and every LDS wait also waits for them.  Casting
// such a pointer to the global address space gives global loads with a
// scalar base and 32-bit offsets.
#define SELM_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ SELM_GLOBAL T *gbl(T *p) {
    return (SELM_GLOBAL T *)p;
}
template <typename T>
__device__ __forceinline__ const SELM_GLOBAL T *gbl(const T *p) {
    return (const SELM_GLOBAL T *)p;
}

template <typename T>
__device__ __forceinline__ cplx<T> mk(T a, T b) {
    cplx<T> r;
    r.x = a;
    r.y = b;
    return r;
}
// complex load / store through a global-address-space pointer
template <typename T>
__device__ __forceinline__ cplx<T> ldc(const SELM_GLOBAL cplx<T> *p) {
    return mk<T>(p->x, p->y);
}
template <typename T>
__device__ __forceinline__ void stc(SELM_GLOBAL cplx<T> *p, cplx<T> v) {
    p->x = v.x;
    p->y = v.y;
}
// Stores of the FP32 plane path's bulk outputs, made coherent at device
// scope (sc1 on gfx950): they are written through the XCD's L2 instead of
// staying dirty in it, so the end of the kernel has no L2 write-back to
// wait for before the next kernel starts.  Measured at 128^3, 20k beads:
// the gaps after the spread, XYf and Z launches fell from 3.4-4.2 us to
// 1.0-2.8 us, 5.6 us per step in all.  (The FP64 transposing passes, whose
// outputs are 8-byte pieces of strided runs, ran 20% slower this way, so
// they keep plain stores.)
template <typename T>
__device__ __forceinline__ void st_dev(SELM_GLOBAL T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// an FP32 complex pair as one 8-byte device-scope store
__device__ __forceinline__ void stc_dev(SELM_GLOBAL cplx<float> *p, cplx<float> v) {
    const unsigned long long bits = ((unsigned long long)__float_as_uint(v.y) << 32) | __float_as_uint(v.x);
    st_dev((SELM_GLOBAL unsigned long long *)p, bits);
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
template <typename T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) { return mk<T>(a.x, -a.y); }
// multiply by +i (dir>0) or -i (dir<0)
template <typename T>
__device__ __forceinline__ cplx<T> mul_i(cplx<T> a, int dir) {
    return dir > 0 ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11).  Counter = {index, step, slot, stream}, key = seed.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    // each 32 x 32 -> 64-bit product is one v_mad_u64_u32 (both halves at
    // once) instead of a v_mul_lo_u32 and a v_mul_hi_u32, and each
    // three-way XOR one v_bitop3_b32 (truth table 0x96 = a ^ b ^ c)
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c[0], p1 = (uint64_t)M1 * c[2];
        uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        uint32_t n0 = __builtin_amdgcn_bitop3_b32(hi1, c[1], k0, 0x96);
        uint32_t n2 = __builtin_amdgcn_bitop3_b32(hi0, c[3], k1, 0x96);
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += W0;
        k1 += W1;
    }
}

constexpr uint32_t kStreamFluid = 0u;
constexpr uint32_t kStreamBead = 1u;

// Box-Muller pair in FP64.  u1 in (0,1] from 53 bits of words 0,1; angle
// 2*pi*t with t in [-1/2,1/2) from 53 bits of words 2,3.
__device__ __forceinline__ void normal_pair_f64(uint64_t seed, uint32_t stream, uint32_t step,
                                                uint32_t index, uint32_t slot, double &z0,
                                                double &z1) {
    uint32_t c[4] = {index, step, slot, stream};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t a = (((uint64_t)c[0] << 32) | c[1]) >> 11;
    uint64_t b = (((uint64_t)c[2] << 32) | c[3]) >> 11;
    double u1 = (double)(a + 1) * 0x1p-53;
    int64_t bc = (int64_t)b - ((int64_t)1 << 52);
    double t = (double)bc * 0x1p-53;
    double rad = sqrt(-2.0 * log(u1));
    double th = 2.0 * 3.14159265358979323846 * t;
    double s, co;
    sincos(th, &s, &co);
    z0 = rad * co;
    z1 = rad * s;
}

// The six normals of one half-spectrum mode of the thermal forcing (3
// components x real and imaginary part) from two Philox blocks: block
// {index, step, 0, stream} gives component 0 from words (0, 1) and
// component 1 from words (2, 3), block {index, step, 1, stream} gives
// component 2 from words (0, 1).  Each pair is a Box-Muller transform of
// two 32-bit uniforms: u1 = (w_a + 1) 2^-32 in (0, 1] and the angle 2 pi t
// with t = (w_b - 2^31) 2^-32 in [-1/2, 1/2).  (cuRAND's normals, which the
// reference draws, also come from 32-bit uniforms.)
__device__ __forceinline__ void mode_blocks(uint64_t seed, uint32_t stream, uint32_t step, uint32_t index,
                                            uint32_t wa[3], uint32_t wb[3]) {
    uint32_t c[4] = {index, step, 0u, stream};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint32_t e[4] = {index, step, 1u, stream};
    philox4x32_10(e, (uint32_t)seed, (uint32_t)(seed >> 32));
    wa[0] = c[0];
    wb[0] = c[1];
    wa[1] = c[2];
    wb[1] = c[3];
    wa[2] = e[0];
    wb[2] = e[1];
}

// Box-Muller pair in FP64 from the 32-bit words (wa, wb): exact uniforms.
__device__ __forceinline__ void bm32_f64(uint32_t wa, uint32_t wb, double &z0, double &z1) {
    const double u1 = (double)((uint64_t)wa + 1) * 0x1p-32;
    const double t = (double)(int32_t)(wb ^ 0x80000000u) * 0x1p-32;
    const double rad = sqrt(-2.0 * log(u1));
    double s, co;
    sincos(2.0 * 3.14159265358979323846 * t, &s, &co);
    z0 = rad * co;
    z1 = rad * s;
}

// The same pair in FP32.  u1 = (wa + 1) 2^-32 and 1 - u1 = (~wa) 2^-32 are
// both formed with a relative error of at most one FP32 rounding, and
// -ln(u1) takes one logarithm: of u1 itself when u1 <= 1/2, and of w = RN(1 - x),
// x = 1 - u1, when u1 > 1/2, with the first-order term of the rounding of
// w, ((1 - w) - x)/w, added back (1 - w and that difference are exact by
// Sterbenz's lemma; the term is below 2^-24 in magnitude, so the hardware
// reciprocal's error in it is far below FP32 rounding).  Neither branch of
// the choice loses the low bits of the uniform, and no lane evaluates two
// logarithms.  The hardware instructions do the transcendental work:
// - ln w = v_log_f32(w) (log2) * RN(ln 2), within 2 ulp;
// - the radius is v_sqrt_f32, within about 1 ulp;
// - v_sin_f32 / v_cos_f32 take the angle in revolutions, so sin(2 pi t)
//   and cos(2 pi t) are one instruction each for |t| <= 1/2.
// Over 2^24 words spread across the 32-bit range, measured on the MI355X
// against FP64:
// - sine and cosine are within 1.8e-7 absolute (ocml's sincospif: 1.1e-7);
// - the logarithm is within 1.9 ulp (ocml's logf: 2.1);
// - the normals are within 4.7e-7 absolute either way, rms 6e-8.
__device__ __forceinline__ void bm32_f32(uint32_t wa, uint32_t wb, float &z0, float &z1) {
    const float u1 = fmaf((float)wa, 0x1p-32f, 0x1p-32f);
    const float x = (float)(~wa) * 0x1p-32f;
    const bool small = wa < 0x80000000u;  // u1 <= 1/2
    const float w = small ? u1 : 1.0f - x;
    const float corr = small ? 0.0f : ((1.0f - w) - x) * __builtin_amdgcn_rcpf(w);
    const float mlog = -(__builtin_amdgcn_logf(w) * 0.693147182f + corr);
    const float t = (float)(int32_t)(wb ^ 0x80000000u) * 0x1p-32f;
    const float rad = __builtin_amdgcn_sqrtf(2.0f * mlog);
    const float sn = __builtin_amdgcn_sinf(t), co = __builtin_amdgcn_cosf(t);
    z0 = rad * co;
    z1 = rad * sn;
}

template <typename T>
__device__ __forceinline__ void bm32(uint32_t wa, uint32_t wb, T &z0, T &z1) {
    if constexpr (sizeof(T) == 8) bm32_f64(wa, wb, z0, z1);
    else bm32_f32(wa, wb, z0, z1);
}

__device__ __forceinline__ int wrap_idx(int p, int n) {
    int q = p % n;
    return q < 0 ? q + n : q;
}

// wrap_idx for the indices a stencil or a tile scan meets, p in [-n, 2n),
// without an integer division; anything else takes the general modulo.
__device__ __forceinline__ int wrap_near(int p, int n) {
    int q = p + (p < 0 ? n : 0);
    q -= q >= n ? n : 0;
    if ((unsigned)q >= (unsigned)n) q = wrap_idx(p, n);
    return q;
}

// floor(a / n) for 0 <= a < 2^22, n >= 1, given rn = 1.0f / n.  The FP32
// quotient is within 1/2 of a/n, so its truncation is off by at most one,
// which the remainder test corrects: no integer division.
__device__ __forceinline__ int div_small(int a, int n, float rn) {
    int q = (int)((float)a * rn);
    const int r = a - q * n;
    q += r >= n ? 1 : 0;
    q -= r < 0 ? 1 : 0;
    return q;
}

// floor(a / n) for 0 <= a < 2^31, n >= 1, given rn = 1.0 / n: as div_small
// with an FP64 quotient.
__device__ __forceinline__ int div_pos(int a, int n, double rn) {
    int q = (int)((double)a * rn);
    const int r = a - q * n;
    q += r >= n ? 1 : 0;
    q -= r < 0 ? 1 : 0;
    return q;
}

// ---------------------------------------------------------------------------
// Radix-R DFTs in registers.  dir = -1: forward (e^{-i}), +1: inverse.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void dft2(cplx<T> *v) {
    cplx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename T>
__device__ __forceinline__ void dft4(cplx<T> &v0, cplx<T> &v1, cplx<T> &v2, cplx<T> &v3, int dir) {
    cplx<T> t0 = cadd(v0, v2), t1 = csub(v0, v2);
    cplx<T> t2 = cadd(v1, v3), t3 = mul_i(csub(v1, v3), dir);
    v0 = cadd(t0, t2);
    v2 = csub(t0, t2);
    v1 = cadd(t1, t3);
    v3 = csub(t1, t3);
}

template <typename T>
__device__ __forceinline__ void dft3(cplx<T> *v, int dir) {
    const T h = (T)0.86602540378443864676;  // sqrt(3)/2
    cplx<T> s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    cplx<T> x0 = cadd(v[0], s);
    cplx<T> t = mk<T>(v[0].x - (T)0.5 * s.x, v[0].y - (T)0.5 * s.y);
    cplx<T> u = mul_i(cscale(d, h), dir);
    v[0] = x0;
    v[1] = cadd(t, u);
    v[2] = csub(t, u);
}

template <typename T>
__device__ __forceinline__ void dft5(cplx<T> *v, int dir) {
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    cplx<T> a1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
    cplx<T> a2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
    cplx<T> x0 = cadd(v[0], cadd(a1, a2));
    cplx<T> r1 = mk<T>(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
    cplx<T> r2 = mk<T>(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
    cplx<T> i1 = mul_i(mk<T>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y), dir);
    cplx<T> i2 = mul_i(mk<T>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y), dir);
    v[0] = x0;
    v[1] = cadd(r1, i1);
    v[4] = csub(r1, i1);
    v[2] = cadd(r2, i2);
    v[3] = csub(r2, i2);
}

template <typename T>
__device__ __forceinline__ void dft8(cplx<T> *v, int dir) {
    const T c = (T)0.70710678118654752440;
    cplx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    cplx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3, dir);
    dft4(o0, o1, o2, o3, dir);
    // W8^k, W8 = e^{dir*2*pi*i/8}
    cplx<T> w1 = mk<T>(c, (T)dir * c);
    cplx<T> w3 = mk<T>(-c, (T)dir * c);
    o1 = cmul(o1, w1);
    o2 = mul_i(o2, dir);
    o3 = cmul(o3, w3);
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}

// Radix plan for one length: stage radices in execution order.
struct RadixPlan {
    int n;
    int nst;
    int r[12];
};

// One Stockham stage over S interleaved sequences held element-major in LDS:
// element e of sequence s lives at buf[e*S + s].  Reads src, writes dst.
template <typename T, int R>
__device__ __forceinline__ void stockham_stage(const cplx<T> *__restrict__ src,
                                               cplx<T> *__restrict__ dst, int n, int S, int Ns,
                                               const cplx<T> *__restrict__ tw, int dir, int tid,
                                               int nthr) {
    const int nb = n / R;
    const int total = S * nb;
    const int twstep = n / (Ns * R);
    for (int t = tid; t < total; t += nthr) {
        int s = t % S;
        int j = t / S;
        int k = j % Ns;
        cplx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = src[(j + r * nb) * S + s];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                cplx<T> w = tw[r * k * twstep];
                if (dir > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        if constexpr (R == 2) dft2(v);
        else if constexpr (R == 3) dft3(v, dir);
        else if constexpr (R == 4) dft4(v[0], v[1], v[2], v[3], dir);
        else if constexpr (R == 5) dft5(v, dir);
        else dft8(v, dir);
        int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) dst[(base + r * Ns) * S + s] = v[r];
    }
}

// Full FFT of S sequences of length plan.n in LDS (ping-pong between a and b).
// tw[m] = e^{-2 pi i m / n}.  Returns the buffer holding the result.  The
// caller must __syncthreads() before (data loaded) and uses the result after
// this function's final barrier.
template <typename T>
__device__ cplx<T> *lds_fft(cplx<T> *a, cplx<T> *b, const RadixPlan &plan, int S,
                            const cplx<T> *tw, int dir) {
    const int tid = threadIdx.x, nthr = blockDim.x;
    int Ns = 1;
    cplx<T> *src = a, *dst = b;
    for (int st = 0; st < plan.nst; st++) {
        int R = plan.r[st];
        switch (R) {
            case 2: stockham_stage<T, 2>(src, dst, plan.n, S, Ns, tw, dir, tid, nthr); break;
            case 3: stockham_stage<T, 3>(src, dst, plan.n, S, Ns, tw, dir, tid, nthr); break;
            case 4: stockham_stage<T, 4>(src, dst, plan.n, S, Ns, tw, dir, tid, nthr); break;
            case 5: stockham_stage<T, 5>(src, dst, plan.n, S, Ns, tw, dir, tid, nthr); break;
            default: stockham_stage<T, 8>(src, dst, plan.n, S, Ns, tw, dir, tid, nthr); break;
        }
        __syncthreads();
        Ns *= R;
        cplx<T> *t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// Wave-wide (64 lanes) sum.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

}  // namespace selm
