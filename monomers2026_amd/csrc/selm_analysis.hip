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
// term is formed in FP32.  Terms are summed in FP64 in a fixed order that
// depends on the sizes only:
// - the origins of a lag are cut into nchunk contiguous ranges (msd_chunks);
// - workgroup (dot, c) sums range c: per thread over (origin, bead) pairs in
//   a fixed stride, then a fixed tree over the workgroup, into partial[dot-1][c];
// - k_msd_final sums partial[dot-1][0 .. nchunk-1] in order and normalises.
// The reference accumulates in FP32 in (bead, origin, dimension) order, so
// its sums differ in the last FP32 digits.
//
// Two entry points share these kernels: selm_msd (host buffers; the library
// allocates the device copy of the frames, the output and the partials) and
// selm_msd_noalloc (device buffers; the partials live in msd_scratch, a
// fixed device array of kMsdScratch doubles, and nothing is allocated).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <mutex>

#include "selm_hip.h"

namespace selm {

constexpr int kMsdThreads = 256;
constexpr int kMsdTargetGroups = 2048;   // workgroups per launch to aim for (8 per CU)
constexpr int kMsdMaxChunks = 64;
constexpr int kMsdScratch = 1 << 16;     // doubles of msd_scratch (512 KiB per device)
constexpr int kMsdMaxDevices = 64;

// The partial sums of selm_msd_noalloc.  Each GPU holds its own copy; host
// code reaches it only through the address hipGetSymbolAddress returns for
// the current device (msd_scratch_ptr), never through the name.
static __device__ __align__(256) double msd_scratch[kMsdScratch];

// Origin ranges per lag: enough workgroups to fill the chip, at most one
// range per origin of the shortest lag, at most kMsdMaxChunks.
static int msd_chunks(int nframes, int nlags, int mult) {
    const long long F = nframes - nframes / 5;
    const long long origins = F - mult;   // origins of lag dot = 1
    long long c = (kMsdTargetGroups + nlags - 2) / (nlags - 1);
    if (c > kMsdMaxChunks) c = kMsdMaxChunks;
    if (c > origins) c = origins;
    return c < 1 ? 1 : (int)c;
}

// frames [nframes][3][natoms] FP32 (the per-frame x, y, z records of a DCD).
// Grid (nlags - 1, nchunk): workgroup (dot - 1, c) writes partial[(dot-1)*nchunk + c].
__global__ __launch_bounds__(kMsdThreads) void k_msd_partial(const float *__restrict__ frames, int natoms, int start,
                                                             int nframes, int mult, float X,
                                                             double *__restrict__ partial) {
    __shared__ double red[kMsdThreads];
    const int dot = blockIdx.x + 1;
    const int c = blockIdx.y;
    const int nchunk = gridDim.y;
    const int lag = dot * mult;
    const long long Fd = (long long)(nframes - start) - lag;
    double acc = 0.0;
    if (Fd > 0) {
        const long long n0 = Fd * c / nchunk, n1 = Fd * (c + 1) / nchunk;
        // pair p = (origin n, bead a), p = (n - n0) * natoms + a; thread t takes
        // p = t, t + 256, ... up to (n1 - n0) * natoms
        long long n = n0 + threadIdx.x / natoms;
        int a = threadIdx.x % natoms;
        const int qn = kMsdThreads / natoms, qa = kMsdThreads % natoms;
        for (; n < n1;) {
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
            n += qn;
            a += qa;
            if (a >= natoms) {
                a -= natoms;
                n++;
            }
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kMsdThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)(dot - 1) * nchunk + c] = red[0];
}

// out[0] = 0; out[dot] = sum over c of partial[dot-1][c] / ((F - lag) natoms), or NaN
__global__ void k_msd_final(const double *__restrict__ partial, int nchunk, int nlags, int natoms, int start,
                            int nframes, int mult, double *__restrict__ out) {
    const int dot = blockIdx.x * blockDim.x + threadIdx.x;
    if (dot >= nlags) return;
    if (dot == 0) {
        out[0] = 0.0;
        return;
    }
    const long long Fd = (long long)(nframes - start) - (long long)dot * mult;
    double s = 0.0;
    for (int c = 0; c < nchunk; c++) s += partial[(size_t)(dot - 1) * nchunk + c];
    out[dot] = Fd > 0 ? s / ((double)Fd * natoms) : NAN;
}

static void msd_launch(const float *frames, int nframes, int natoms, float X, int nlags, int mult, int nchunk,
                       double *partial, double *out, hipStream_t st) {
    const int start = nframes / 5;
    if (nlags > 1)
        hipLaunchKernelGGL(k_msd_partial, dim3(nlags - 1, nchunk), dim3(kMsdThreads), 0, st, frames, natoms, start,
                           nframes, mult, X, partial);
    hipLaunchKernelGGL(k_msd_final, dim3((nlags + 255) / 256), dim3(256), 0, st, partial, nchunk, nlags, natoms, start,
                       nframes, mult, out);
}

static bool msd_args_ok(const void *frames, int nframes, int natoms, const double *box, int nlags, int mult,
                        const void *msd) {
    return frames && box && msd && nframes >= 1 && natoms >= 1 && nlags >= 1 && mult >= 1;
}

// Address of this device's msd_scratch, resolved once per device on the
// first call made with that device current.
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMsdMaxDevices];

static int msd_scratch_ptr(double **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMsdMaxDevices) return SELM_ERR_HIP;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(msd_scratch)) != hipSuccess || !p) return SELM_ERR_HIP;
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return SELM_OK;
}

}  // namespace selm

extern "C" int selm_msd(const float *frames, int nframes, int natoms, const double box[3], int nlags, int mult,
                        double *msd) {
    using namespace selm;
    if (!msd_args_ok(frames, nframes, natoms, box, nlags, mult, msd)) return SELM_ERR_ARG;
    msd[0] = 0.0;
    if (nlags == 1) return SELM_OK;
    const int nchunk = msd_chunks(nframes, nlags, mult);
    const size_t fbytes = (size_t)nframes * 3 * natoms * sizeof(float);
    const size_t obytes = (size_t)nlags * sizeof(double);
    const size_t pbytes = (size_t)(nlags - 1) * nchunk * sizeof(double);
    float *d_frames = nullptr;
    double *d_out = nullptr, *d_part = nullptr;
    int rc = SELM_OK;
    if (hipMalloc(&d_frames, fbytes) != hipSuccess || hipMalloc(&d_out, obytes) != hipSuccess ||
        hipMalloc(&d_part, pbytes) != hipSuccess) {
        rc = SELM_ERR_NOMEM;
    } else if (hipMemcpy(d_frames, frames, fbytes, hipMemcpyHostToDevice) != hipSuccess) {
        rc = SELM_ERR_HIP;
    } else {
        msd_launch(d_frames, nframes, natoms, (float)box[0] / 2.0f, nlags, mult, nchunk, d_part, d_out, 0);
        if (hipGetLastError() != hipSuccess || hipMemcpy(msd, d_out, obytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = SELM_ERR_HIP;
    }
    if (d_frames) (void)hipFree(d_frames);
    if (d_out) (void)hipFree(d_out);
    if (d_part) (void)hipFree(d_part);
    return rc;
}

extern "C" int selm_msd_noalloc(const float *d_frames, int nframes, int natoms, const double box[3], int nlags,
                                int mult, double *d_msd, void *stream) {
    using namespace selm;
    if (!msd_args_ok(d_frames, nframes, natoms, box, nlags, mult, d_msd)) return SELM_ERR_ARG;
    const int nchunk = nlags > 1 ? msd_chunks(nframes, nlags, mult) : 1;
    if ((long long)(nlags - 1) * nchunk > kMsdScratch) return SELM_ERR_SCRATCH;
    double *part = nullptr;
    int rc = msd_scratch_ptr(&part);
    if (rc != SELM_OK) return rc;
    msd_launch(d_frames, nframes, natoms, (float)box[0] / 2.0f, nlags, mult, nchunk, part, d_msd,
               (hipStream_t)stream);
    return hipGetLastError() == hipSuccess ? SELM_OK : SELM_ERR_HIP;
}
