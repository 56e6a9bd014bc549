// selm_fluid.hpp -- the viscous predictor evaluated at one MAC face, reading
// the sparse spread pool.  Shared by the compile-time-length X-FFT pass (which
// evaluates it while loading its rows) and the runtime-length path.
#pragma once

#include "selm_desc.hpp"

namespace selm {

// Spread impulse of component d at cell (x, y, z): from the cell's tile when
// a bead touched the tile in this step, else 0.
template <typename T>
__device__ __forceinline__ T spread_at(const RepDev &R, int d, int x, int tyz, int lyz, int step) {
    const int tx = x / R.te[0];
    const int t = tyz + tx;
    if (R.tstamp[t] != step) return (T)0;
    const int tcells = R.te[0] * R.te[1] * R.te[2];
    const T *pool = (const T *)R.pool;
    return pool[((long long)t * 3 + d) * tcells + lyz + (x - tx * R.te[0])];
}

// Per-row part of the pool addressing: tile index without its x part, and
// the in-tile offset of (y, z).
__device__ __forceinline__ void spread_row(const RepDev &R, int y, int z, int &tyz, int &lyz) {
    const int ty = y / R.te[1], tz = z / R.te[2];
    tyz = (tz * R.tn[1] + ty) * R.tn[0];
    lyz = ((z - tz * R.te[2]) * R.te[1] + (y - ty * R.te[1])) * R.te[0];
}

// u* = u + (mu * L7(u) * dt + f) / rho with the periodic 7-point Laplacian
// (cuda/SELM_Kernel.cu:155-195; neighbour order x+, x-, y+, y-, z+, z-).
// Values stored as T, arithmetic in C.
template <typename T, typename C = T>
__device__ __forceinline__ C ustar_at(const RepDev &R, int d, int x, int y, int z, int tyz, int lyz, int step) {
    const int nx = R.n[0], ny = R.n[1], nz = R.n[2];
    const long long Nyx = (long long)nx * ny;
    const T *ud = (const T *)R.u + d * R.N;
    const long long n = (long long)z * Nyx + (long long)y * nx + x;
    const long long i1 = (x != nx - 1) ? 1 : -x;
    const long long i0 = (x != 0) ? -1 : nx - 1;
    const long long j1 = (y != ny - 1) ? nx : -(long long)y * nx;
    const long long j0 = (y != 0) ? -(long long)nx : (long long)(ny - 1) * nx;
    const long long k1 = (z != nz - 1) ? Nyx : -(long long)z * Nyx;
    const long long k0 = (z != 0) ? -Nyx : (long long)(nz - 1) * Nyx;
    const C dxsq = (C)(R.dx * R.dx);
    const C uc = (C)ud[n];
    const C lap = ((C)ud[n + i1] + (C)ud[n + i0] + (C)ud[n + j1] + (C)ud[n + j0] + (C)ud[n + k1] + (C)ud[n + k0] -
                   (C)6 * uc) / dxsq;
    const C f = (C)spread_at<T>(R, d, x, tyz, lyz, step);
    return uc + ((C)R.mu * lap * (C)R.dt + f) / (C)R.rho;
}

// Thermal forcing of mode (i, j, k) as it enters u*'s spectrum, for the
// three components.  The reference (cuda/SELM_Kernel.cu:37-85, then Z2D and
// gthm_scale :88-98) draws per half-spectrum mode and component the complex
// value (xi1*re, xi2*im) with re = im = coef/sqrt(2) (re = coef, im = 0 for a
// self-conjugate mode), coef = sqrt(sum_a 2(1-cos theta_a))/dx, and maps it to
// real space with an unnormalised c2r scaled by sqrt(2 kB T dt mu / N).  In the
// x-frequency planes i = 0 and i = nx/2 the c2r keeps only
// (g(i,j,k) + conj g(i,-j,-k))/2, so that is the value returned; the DFT of
// the real field is N times it, and u* receives f/rho (noise_amp).
// lam = sum_a (lambda table)_a as used by the projection.
template <typename T>
__device__ __forceinline__ void thermal_mode(const RepDev &R, int i, int j, int k, T lam, uint32_t step,
                                             cplx<T> g[3]) {
    const int nx = R.n[0], ny = R.n[1], nz = R.n[2], xc = R.xc;
    const int k1 = k == 0 ? 0 : nz - k, j1 = j == 0 ? 0 : ny - j, i1 = i == 0 ? 0 : nx - i;  // -k mod n
    const bool selfconj = (k1 == k && j1 == j && i1 == i);
    const bool plane = (i == 0) || (2 * i == nx);
    const T coef = sqrt(-lam * (T)R.inv_dxsq);
    const T re = selfconj ? coef : coef / sqrt((T)2);
    const T im = selfconj ? (T)0 : re;
    const T amp = (T)R.noise_amp;
    const uint32_t nmode = (uint32_t)(((long long)k * ny + j) * xc + i);
    const uint32_t npart = (uint32_t)(((long long)k1 * ny + j1) * xc + i);
    uint32_t wa[3], wb[3], pa[3], pb[3];
    mode_blocks(R.seed, kStreamFluid, step, nmode, wa, wb);
    if (plane) mode_blocks(R.seed, kStreamFluid, step, npart, pa, pb);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        T z0, z1;
        bm32<T>(wa[d], wb[d], z0, z1);
        cplx<T> gk = mk<T>(z0 * re, z1 * im);
        if (plane) {
            T y0, y1;
            bm32<T>(pa[d], pb[d], y0, y1);
            gk = mk<T>((gk.x + y0 * re) * (T)0.5, (gk.y - y1 * im) * (T)0.5);
        }
        g[d] = cscale(gk, amp);
    }
}

// Thermal forcing of mode g of a batch's work table (modes of replica r at
// start[r] ..) into nbuf: the row order of C ([kx][ky][kz], kz fastest), or
// with plane = 1 that of the plane path's P ([kz][kx][ky], ky fastest).
// WT: FP32 values stored at device scope (the plane path's store shape).
template <typename T, bool WT = false>
__device__ __forceinline__ void noise_mode_of(const RepDev &R, long long m, int plane) {
    if (!R.noise) return;
    const int ny = R.n[1], nz = R.n[2];
    const int mi = (int)m;  // Nc < 2^31
    int i, j, k;
    if (plane) {
        const int ki = div_pos(mi, ny, 1.0 / ny);
        j = mi - ki * ny;
        k = div_pos(ki, R.xc, 1.0 / R.xc);
        i = ki - k * R.xc;
    } else {
        const int ij = div_pos(mi, nz, 1.0 / nz);
        k = mi - ij * nz;
        i = div_pos(ij, ny, 1.0 / ny);
        j = ij - i * ny;
    }
    const T *kxt = (const T *)R.ktab[0], *kyt = (const T *)R.ktab[1], *kzt = (const T *)R.ktab[2];
    const T lam = (kxt[4 * i] + kyt[4 * j]) + kzt[4 * k];
    cplx<T> gv[3];
    thermal_mode<T>(R, i, j, k, lam, (uint32_t)(R.ts[0] + 1), gv);
    SELM_GLOBAL cplx<T> *nb = gbl((cplx<T> *)R.nbuf);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if constexpr (sizeof(T) == 4) {
            if (WT) stc_dev(nb + d * R.Nc + m, gv[d]);
            else stc(nb + d * R.Nc + m, gv[d]);
        } else {
            stc(nb + d * R.Nc + m, gv[d]);
        }
    }
}

// Thermal forcing of the modes m0, m0 + stride, ... (at most K, below
// tot) of replica R into nbuf, plane layout as noise_mode_of.  The
// wavenumber-table values of all K modes are loaded before the first
// store: stores count in the same vector memory counter as loads, so a
// load issued after them would also wait for their completion.
template <typename T, int K, bool WT = false>
__device__ __forceinline__ void noise_modes_rep(const RepDev &R, long long m0, long long stride, long long tot,
                                                int plane) {
    if (!R.noise) return;
    const int ny = R.n[1], nz = R.n[2];
    const uint32_t step = (uint32_t)(R.ts[0] + 1);
    const T *kxt = (const T *)R.ktab[0], *kyt = (const T *)R.ktab[1], *kzt = (const T *)R.ktab[2];
    int I[K], J[K], Z[K];
    T lam[K];
#pragma unroll
    for (int q = 0; q < K; q++) {
        const long long m = m0 + q * stride;
        const int mi = (int)(m < tot ? m : m0);  // Nc < 2^31
        if (plane) {
            const int ki = div_pos(mi, ny, 1.0 / ny);
            J[q] = mi - ki * ny;
            Z[q] = div_pos(ki, R.xc, 1.0 / R.xc);
            I[q] = ki - Z[q] * R.xc;
        } else {
            const int ij = div_pos(mi, nz, 1.0 / nz);
            Z[q] = mi - ij * nz;
            I[q] = div_pos(ij, ny, 1.0 / ny);
            J[q] = ij - I[q] * ny;
        }
        lam[q] = (kxt[4 * I[q]] + kyt[4 * J[q]]) + kzt[4 * Z[q]];
    }
    SELM_GLOBAL cplx<T> *nb = gbl((cplx<T> *)R.nbuf);
#pragma unroll
    for (int q = 0; q < K; q++) {
        const long long m = m0 + q * stride;
        if (m >= tot) break;
        cplx<T> gv[3];
        thermal_mode<T>(R, I[q], J[q], Z[q], lam[q], step, gv);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if constexpr (sizeof(T) == 4) {
                if (WT) stc_dev(nb + d * R.Nc + m, gv[d]);
                else stc(nb + d * R.Nc + m, gv[d]);
            } else {
                stc(nb + d * R.Nc + m, gv[d]);
            }
        }
    }
}

// When the wave's active lanes all draw modes of one replica (always for a
// single system), the replica index is taken from the first active lane so
// that the descriptor's fields come through scalar loads.
template <typename T, bool WT = false>
__device__ __forceinline__ void noise_mode(const RepDev *__restrict__ reps, int nrep,
                                           const long long *__restrict__ start, long long g, int plane) {
    const int r = find_rep(start, nrep, g);
    const int r0 = __builtin_amdgcn_readfirstlane(r);
    if (__ballot(r != r0) == 0) noise_mode_of<T, WT>(reps[r0], g - start[r0], plane);
    else noise_mode_of<T, WT>(reps[r], g - start[r], plane);
}

}  // namespace selm
