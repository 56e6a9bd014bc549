// selm_analysis.hip -- mean-square displacement of a stored trajectory.
//
// Semantics of py/Diffusion Coefficient/MSD.py:15-69:
// - The first frame used is startframe = floor(nframes / 5).
// - For lag = dot * mult, dot = 1 .. nlags-1, every bead and every time
//   origin n in [0, F - lag) contributes, with F = nframes - startframe.
// - Per dimension the lag displacement D is taken to its minimum image with
//   half-box X: (D - 2X)^2 if D >= X, (D + 2X)^2 if D <= -X, else D^2.
//   Like the reference, X is half the x edge in every dimension.
// - MSD(lag) = sum / ((F - lag) * natoms).  A lag with no origin gives NaN
//   where the reference script divides by zero.
//
// D and its image shift are evaluated in FP32, from the FP32 coordinates of
// the trajectory, as the reference does on its float32 arrays.  Each squared
// term is formed in FP32.  Terms are summed in FP64 in a fixed order: per
// thread in a fixed stride, then a fixed tree over the workgroup.  The
// reference accumulates in FP32 in (bead, origin, dimension) order, so its
// sums differ in the last FP32 digits.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>

#include "selm_hip.h"

namespace selm {

constexpr int kMsdThreads = 256;

// frames [nframes][3][natoms] FP32 (the per-frame x, y, z records of a DCD)
__global__ __launch_bounds__(kMsdThreads) void k_msd(const float *__restrict__ frames, int natoms, int start,
                                                     int nframes, int mult, float X, double *__restrict__ out) {
    __shared__ double red[kMsdThreads];
    const int dot = blockIdx.x + 1;
    const int lag = dot * mult;
    const int F = nframes - start;
    const long long Fd = (long long)F - lag;
    double acc = 0.0;
    if (Fd > 0) {
        const long long npair = Fd * natoms;
        for (long long p = threadIdx.x; p < npair; p += kMsdThreads) {
            const long long n = p / natoms;
            const int a = (int)(p - n * natoms);
            const float *f0 = frames + (size_t)(start + n) * 3 * natoms;
            const float *f1 = frames + (size_t)(start + n + lag) * 3 * natoms;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                float D = f1[d * natoms + a] - f0[d * natoms + a];
                if (D >= X) D = D - 2.0f * X;
                else if (D <= -X) D = D + 2.0f * X;
                const float sq = D * D;
                acc += (double)sq;
            }
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kMsdThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[dot] = Fd > 0 ? red[0] / ((double)Fd * natoms) : NAN;
}

}  // namespace selm

extern "C" int selm_msd(const float *frames, int nframes, int natoms, const double box[3], int nlags, int mult,
                        double *msd) {
    using namespace selm;
    if (!frames || !box || !msd || nframes < 1 || natoms < 1 || nlags < 1 || mult < 1) return SELM_ERR_ARG;
    msd[0] = 0.0;
    if (nlags == 1) return SELM_OK;
    const int start = nframes / 5;
    const size_t fbytes = (size_t)nframes * 3 * natoms * sizeof(float);
    float *d_frames = nullptr;
    double *d_out = nullptr;
    int rc = SELM_OK;
    if (hipMalloc(&d_frames, fbytes) != hipSuccess || hipMalloc(&d_out, (size_t)nlags * sizeof(double)) != hipSuccess) {
        rc = SELM_ERR_NOMEM;
    } else if (hipMemcpy(d_frames, frames, fbytes, hipMemcpyHostToDevice) != hipSuccess) {
        rc = SELM_ERR_HIP;
    } else {
        const float X = (float)box[0] / 2.0f;
        hipLaunchKernelGGL(k_msd, dim3(nlags - 1), dim3(kMsdThreads), 0, 0, d_frames, natoms, start, nframes, mult,
                           X, d_out);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpy(msd + 1, d_out + 1, (size_t)(nlags - 1) * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            rc = SELM_ERR_HIP;
    }
    if (d_frames) (void)hipFree(d_frames);
    if (d_out) (void)hipFree(d_out);
    return rc;
}
