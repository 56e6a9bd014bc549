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
template <typename T>
__device__ __forceinline__ T ustar_at(const RepDev &R, int d, int x, int y, int z, int tyz, int lyz, int step) {
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
    const T dxsq = (T)(R.dx * R.dx);
    const T uc = ud[n];
    const T lap = (ud[n + i1] + ud[n + i0] + ud[n + j1] + ud[n + j0] + ud[n + k1] + ud[n + k0] - (T)6 * uc) / dxsq;
    const T f = spread_at<T>(R, d, x, tyz, lyz, step);
    return uc + ((T)R.mu * lap * (T)R.dt + f) / (T)R.rho;
}

}  // namespace selm
