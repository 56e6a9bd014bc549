/*
 * selm_oracle.cpp -- FP64 CPU restatement of the reference's SELM time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, and the timed CPU
 * baseline.  The product (monomers2026_amd/) never includes, links or calls
 * this file.
 *
 * The restatement is single-process (the reference's MPI z-slab split with
 * one rank: head/tail halos become the periodic neighbour planes) and follows
 * the reference stage by stage, in the reference's order and arithmetic:
 *
 *   Selm::run step ........ mpi/SELM.cpp:84-91 (== cuda/SELM.cu:73-90)
 *   Lagrangian::run ....... mpi/SELM_Lagrangian_LAMMPS.cpp:90-104
 *   compute_gamma ......... mpi/SELM_Lagrangian_LAMMPS.cpp:161-180
 *   lagrangian_update ..... mpi/SELM_Lagrangian_LAMMPS.cpp:115-159
 *   gamma / lambda ........ mpi/SELM_Lagrangian_LAMMPS.cpp:182-278,
 *                           cuda/SELM_Kernel.cu:375-470
 *   Eulerian::run ......... mpi/SELM_Eulerian_Period.cpp:86-102
 *   u2uAsterisk ........... mpi/SELM_Eulerian_Period.cpp:105-153
 *   div ................... mpi/SELM_Eulerian_Period.cpp:216-238
 *   d2z_data_transform .... cuda/SELM_Kernel.cu:3-34 (layout [z][y][x/2+1])
 *   uAsterisk2u ........... cuda/SELM_Kernel.cu:119-152
 *   Sum / subtract_mean ... mpi/SELM_Eulerian_Period.cpp:202-212,240-250
 *   compute_f (fluid) ..... cuda/SELM_Eulerian_Period.cu:128-153
 *   gthm_generate ......... cuda/SELM_Kernel.cu:37-85
 *   gthm_scale ............ cuda/SELM_Kernel.cu:88-98
 *
 * Where the reference calls a third-party library, the restatement uses its
 * own code with the library's published semantics:
 *   - cuFFT/FFTW D2Z, Z2D: unnormalised 3-D DFTs.  The complex-to-real
 *     transform follows FFTW's rank-3 c2r: complex inverse DFTs along z and y,
 *     then the real inverse along x, which reads only the real parts of the
 *     x-frequency 0 and Nx/2 coefficients.
 *   - cuRAND XORWOW / LAMMPS RanMars: replaced by Philox4x32-10 (Salmon et
 *     al., SC'11) with a Box-Muller transform; the normal for a given
 *     (seed, stream, step, index, slot) is a pure function of those numbers.
 *     The thermal forcing draws the six normals of a mode from two Philox
 *     blocks (mode_normals); a bead's kick draws one pair (normal_pair).
 *   - LAMMPS bond_style harmonic (E = K (r - r0)^2) and fix addforce, applied
 *     at the call site cuda/SELM_Lagrangian_LAMMPS.cu:122 (setup_minimal).
 */
#include "selm_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cd = std::complex<double>;
constexpr double kPi = 3.14159265358979323846;  /* PI, cuda/SELM_Kernel.cuh:3 */

/* ---------------------------------------------------------------------- */
/* Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw 2011). */
/* ---------------------------------------------------------------------- */
inline void mulhilo32(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(M0, c[0], hi0, lo0);
        mulhilo32(M1, c[2], hi1, lo1);
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += W0; k1 += W1;
    }
}

/* Two standard normals from one Philox block.  u1 in (0,1] from the high
 * 53 bits of words (0,1); the angle uses words (2,3) as a 53-bit integer
 * centred to [-2^52, 2^52), i.e. theta = 2*pi*t with t in [-1/2, 1/2). */
void normal_pair(uint64_t seed, uint32_t stream, uint32_t step, uint32_t index,
                 uint32_t slot, double &z0, double &z1) {
    uint32_t c[4] = {index, step, slot, stream};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t a = (((uint64_t)c[0] << 32) | c[1]) >> 11;
    uint64_t b = (((uint64_t)c[2] << 32) | c[3]) >> 11;
    double u1 = (double)(a + 1) * 0x1p-53;
    int64_t bc = (int64_t)b - ((int64_t)1 << 52);
    double t = (double)bc * 0x1p-53;
    double rad = std::sqrt(-2.0 * std::log(u1));
    double th = 2.0 * kPi * t;
    z0 = rad * std::cos(th);
    z1 = rad * std::sin(th);
}

constexpr uint32_t kStreamFluid = 0u;
constexpr uint32_t kStreamBead = 1u;

/* The six normals of one half-spectrum mode (3 components x real and
 * imaginary part) from two Philox blocks: block {index, step, 0, stream}
 * gives z[0], z[1] from words (0, 1) and z[2], z[3] from words (2, 3),
 * block {index, step, 1, stream} gives z[4], z[5] from words (0, 1).  Each
 * pair is a Box-Muller transform of two 32-bit uniforms: u1 = (w_a + 1)
 * 2^-32 in (0, 1], the angle 2*pi*t with t = (w_b - 2^31) 2^-32 in
 * [-1/2, 1/2). */
void mode_normals(uint64_t seed, uint32_t stream, uint32_t step, uint32_t index, double z[6]) {
    uint32_t c[4] = {index, step, 0u, stream}, e[4] = {index, step, 1u, stream};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    philox4x32_10(e, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint32_t wa[3] = {c[0], c[2], e[0]}, wb[3] = {c[1], c[3], e[1]};
    for (int q = 0; q < 3; q++) {
        double u1 = (double)((uint64_t)wa[q] + 1) * 0x1p-32;
        double t = (double)(int32_t)(wb[q] ^ 0x80000000u) * 0x1p-32;
        double rad = std::sqrt(-2.0 * std::log(u1));
        double th = 2.0 * kPi * t;
        z[2 * q] = rad * std::cos(th);
        z[2 * q + 1] = rad * std::sin(th);
    }
}

/* ---------------------------------------------------------------------- */
/* Peskin 4-point 1-D factor with the reference's branch structure          */
/* (cuda/SELM_Kernel.cu:387-395): a factor starts at 1.0 and is multiplied  */
/* only for r < 1 or 1 <= r < 2, so a stencil point at r >= 2 keeps 1.0.    */
/* ---------------------------------------------------------------------- */
inline double phi1(double r, int f4) {
    if (r < 1.0) return (3.0 - 2.0 * r + std::sqrt(1.0 + 4.0 * r * (1.0 - r))) / 8.0;
    if (r < 2.0) return (5.0 - 2.0 * r - std::sqrt(-7.0 + 4.0 * r * (3.0 - r))) / 8.0;
    return f4 ? 1.0 : 0.0;
}

/* gamma: cuda/SELM_Kernel.cu:375-422 (single slab of mpi/...:182-230) */
double gamma_op(double rx, double ry, double rz, int Nx, int Ny, int Nz,
                const double *u, int f4) {
    int p0 = (int)(std::floor(rx) - 1);
    int p1 = (int)(std::floor(ry) - 1);
    int p2 = (int)(std::floor(rz) - 1);
    double rv = 0.0;
    for (int k = 0; k < 4; k++) {
        int pk = p2 + k;
        double delta_k = 1.0 * phi1(std::fabs(rz - pk), f4);
        pk = (Nz + pk) % Nz;
        for (int j = 0; j < 4; j++) {
            int pj = p1 + j;
            double delta_j = delta_k * phi1(std::fabs(ry - pj), f4);
            pj = (Ny + pj) % Ny;
            for (int i = 0; i < 4; i++) {
                int pi = p0 + i;
                double delta_i = delta_j * phi1(std::fabs(rx - pi), f4);
                pi = (Nx + pi) % Nx;
                size_t n = ((size_t)pk * Ny + pj) * Nx + pi;
                rv += u[n] * delta_i;
            }
        }
    }
    return rv;
}

/* lambda: cuda/SELM_Kernel.cu:424-470 (sequential: mpi/...:232-278) */
void lambda_op(double rx, double ry, double rz, int Nx, int Ny, int Nz, double f,
               double *u_f, int f4) {
    int p0 = (int)(std::floor(rx) - 1);
    int p1 = (int)(std::floor(ry) - 1);
    int p2 = (int)(std::floor(rz) - 1);
    for (int k = 0; k < 4; k++) {
        int pk = p2 + k;
        double delta_k = 1.0 * phi1(std::fabs(rz - pk), f4);
        pk = (Nz + pk) % Nz;
        for (int j = 0; j < 4; j++) {
            int pj = p1 + j;
            double delta_j = delta_k * phi1(std::fabs(ry - pj), f4);
            pj = (Ny + pj) % Ny;
            for (int i = 0; i < 4; i++) {
                int pi = p0 + i;
                double delta_i = delta_j * phi1(std::fabs(rx - pi), f4);
                pi = (Nx + pi) % Nx;
                size_t n = ((size_t)pk * Ny + pj) * Nx + pi;
                u_f[n] -= f * delta_i;
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Mixed-radix complex FFT (recursive decimation in time), unnormalised.    */
/* ---------------------------------------------------------------------- */
struct FFT1 {
    int n = 0;
    std::vector<int> fac;
    std::vector<cd> tw[2];  /* [0]: e^{-2 pi i m/n}, [1]: e^{+2 pi i m/n} */
    explicit FFT1(int n_) : n(n_) {
        int m = n;
        for (int p : {4, 2, 3, 5}) {
            while (m % p == 0) { fac.push_back(p); m /= p; }
        }
        for (int p = 7; m > 1; p += 2) {
            while (m % p == 0) { fac.push_back(p); m /= p; }
        }
        tw[0].resize(n);
        tw[1].resize(n);
        for (int k = 0; k < n; k++) {
            double a = 2.0 * kPi * (double)k / (double)n;
            tw[0][k] = cd(std::cos(a), -std::sin(a));
            tw[1][k] = cd(std::cos(a), std::sin(a));
        }
    }
    void rec(const cd *in, size_t stride, cd *out, int len, int fi, int s, cd *scratch) const {
        if (len == 1) { out[0] = in[0]; return; }
        int p = fac[fi];
        int m = len / p;
        for (int q = 0; q < p; q++)
            rec(in + q * stride, stride * p, out + q * m, m, fi + 1, s, scratch);
        int step = n / len;
        const std::vector<cd> &w = tw[s];
        for (int k = 0; k < m; k++) {
            for (int q = 0; q < p; q++) scratch[q] = out[q * m + k] * w[(size_t)q * k * step % n];
            for (int r = 0; r < p; r++) {
                cd acc = 0.0;
                for (int q = 0; q < p; q++) acc += scratch[q] * w[((size_t)q * r * m * step) % n];
                scratch[p + r] = acc;
            }
            for (int r = 0; r < p; r++) out[k + r * m] = scratch[p + r];
        }
    }
    /* sign < 0: forward e^{-i..}; sign > 0: inverse e^{+i..} */
    void exec(const cd *in, size_t stride, cd *out, int sign) const {
        std::vector<cd> scratch(2 * (size_t)n + 16);
        rec(in, stride, out, n, 0, sign > 0 ? 1 : 0, scratch.data());
    }
};

/* 3-D r2c: [z][y][x] real -> [z][y][x/2+1] complex, unnormalised forward. */
void r2c3(int Nx, int Ny, int Nz, const double *in, cd *out) {
    const int Xc = Nx / 2 + 1;
    FFT1 fx(Nx), fy(Ny), fz(Nz);
#pragma omp parallel
    {
        std::vector<cd> a(std::max(Nx, std::max(Ny, Nz))), b(a.size());
#pragma omp for collapse(2)
        for (int k = 0; k < Nz; k++)
            for (int j = 0; j < Ny; j++) {
                const double *row = in + ((size_t)k * Ny + j) * Nx;
                for (int i = 0; i < Nx; i++) a[i] = cd(row[i], 0.0);
                fx.exec(a.data(), 1, b.data(), -1);
                cd *o = out + ((size_t)k * Ny + j) * Xc;
                for (int i = 0; i < Xc; i++) o[i] = b[i];
            }
#pragma omp for collapse(2)
        for (int k = 0; k < Nz; k++)
            for (int i = 0; i < Xc; i++) {
                cd *col = out + (size_t)k * Ny * Xc + i;
                fy.exec(col, Xc, b.data(), -1);
                for (int j = 0; j < Ny; j++) col[(size_t)j * Xc] = b[j];
            }
#pragma omp for collapse(2)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Xc; i++) {
                cd *col = out + (size_t)j * Xc + i;
                fz.exec(col, (size_t)Ny * Xc, b.data(), -1);
                for (int k = 0; k < Nz; k++) col[(size_t)k * Ny * Xc] = b[k];
            }
    }
}

/* 3-D c2r with FFTW rank-3 semantics (see file header); `in` is modified. */
void c2r3(int Nx, int Ny, int Nz, cd *in, double *out) {
    const int Xc = Nx / 2 + 1;
    FFT1 fx(Nx), fy(Ny), fz(Nz);
#pragma omp parallel
    {
        std::vector<cd> a(std::max(Nx, std::max(Ny, Nz))), b(a.size());
#pragma omp for collapse(2)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Xc; i++) {
                cd *col = in + (size_t)j * Xc + i;
                fz.exec(col, (size_t)Ny * Xc, b.data(), +1);
                for (int k = 0; k < Nz; k++) col[(size_t)k * Ny * Xc] = b[k];
            }
#pragma omp for collapse(2)
        for (int k = 0; k < Nz; k++)
            for (int i = 0; i < Xc; i++) {
                cd *col = in + (size_t)k * Ny * Xc + i;
                fy.exec(col, Xc, b.data(), +1);
                for (int j = 0; j < Ny; j++) col[(size_t)j * Xc] = b[j];
            }
#pragma omp for collapse(2)
        for (int k = 0; k < Nz; k++)
            for (int j = 0; j < Ny; j++) {
                const cd *h = in + ((size_t)k * Ny + j) * Xc;
                /* Hermitian extension; DC (and Nyquist for even Nx) keep only their real part */
                a[0] = cd(h[0].real(), 0.0);
                for (int i = 1; i < Xc; i++) {
                    if (2 * i == Nx) { a[i] = cd(h[i].real(), 0.0); continue; }
                    a[i] = h[i];
                    a[Nx - i] = std::conj(h[i]);
                }
                fx.exec(a.data(), 1, b.data(), +1);
                double *row = out + ((size_t)k * Ny + j) * Nx;
                for (int i = 0; i < Nx; i++) row[i] = b[i].real();
            }
    }
}

/* ---------------------------------------------------------------------- */
/* The oracle state: one Selm with Eulerian_Period + a bead-spring system.   */
/* ---------------------------------------------------------------------- */
struct Oracle {
    oracle_params P{};
    int Nx, Ny, Nz, Xc;
    size_t N, Nc;
    double dx;
    std::vector<double> u[2], f, p;
    int cur_u = 0;
    std::vector<cd> work;
    int M = 0;
    std::vector<double> x, v, fc, ftot, m, fext, ftmp, xgrid;
    std::vector<int> bi, bj;
    std::vector<double> bK, br0;
    long timestep = 0;
};

/* LAMMPS bond_style harmonic + fix addforce, at positions xs (AoS). */
void bond_forces(const Oracle &o, const double *xs, double *fo) {
    std::memset(fo, 0, sizeof(double) * 3 * o.M);
    double L[3], half[3];
    for (int d = 0; d < 3; d++) { L[d] = o.P.hi[d] - o.P.lo[d]; half[d] = 0.5 * L[d]; }
    for (size_t b = 0; b < o.bi.size(); b++) {
        int i1 = o.bi[b], i2 = o.bj[b];
        double del[3];
        for (int d = 0; d < 3; d++) {
            double dd = xs[3 * i1 + d] - xs[3 * i2 + d];
            if (std::fabs(dd) > half[d]) dd += (dd < 0.0) ? L[d] : -L[d];
            del[d] = dd;
        }
        double rsq = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
        double r = std::sqrt(rsq);
        double dr = r - o.br0[b];
        double rk = o.bK[b] * dr;
        double fbond = (r > 0.0) ? -2.0 * rk / r : 0.0;
        for (int d = 0; d < 3; d++) {
            fo[3 * i1 + d] += del[d] * fbond;
            fo[3 * i2 + d] -= del[d] * fbond;
        }
    }
    for (int i = 0; i < 3 * o.M; i++) fo[i] += o.fext[i];
}

/* Eulerian_Period::compute_f (cuda/SELM_Eulerian_Period.cu:128-153):
 * thermal field for step `s` (gthm_generate -> Z2D x3 -> gthm_scale), or 0. */
void fluid_compute_f(Oracle &o, long s) {
    const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz, Xc = o.Xc;
    if (!o.P.noise) {
        std::fill(o.f.begin(), o.f.end(), 0.0);
        return;
    }
    const double dxsq = o.dx * o.dx;
    std::vector<cd> g(3 * o.Nc);
#pragma omp parallel for collapse(2)
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Xc; i++) {
                size_t n = ((size_t)k * Ny + j) * Xc + i;
                double coef = (2 * (1 - std::cos(2 * kPi * i / Nx)) + 2 * (1 - std::cos(2 * kPi * j / Ny)) +
                               2 * (1 - std::cos(2 * kPi * k / Nz))) / dxsq;
                coef = std::sqrt(coef);
                int k1 = (Nz - k) % Nz, j1 = (Ny - j) % Ny, i1 = (Nx - i) % Nx;
                double re, im;
                if (k1 != k || j1 != j || i1 != i) { re = coef / std::sqrt(2.0); im = re; }
                else { re = coef; im = 0.0; }
                double z[6];
                mode_normals(o.P.seed, kStreamFluid, (uint32_t)s, (uint32_t)n, z);
                for (int d = 0; d < 3; d++) g[d * o.Nc + n] = cd(z[2 * d] * re, z[2 * d + 1] * im);
            }
    const double scale = std::sqrt(2 * o.P.kB * o.P.T * o.P.dt * o.P.mu / (double)o.N);
    for (int d = 0; d < 3; d++) {
        double *fd = o.f.data() + d * o.N;
        c2r3(Nx, Ny, Nz, g.data() + d * o.Nc, fd);
        for (size_t n = 0; n < o.N; n++) fd[n] *= scale;
    }
}

/* Lagrangian_LAMMPS::run + lagrangian_update (mpi/SELM_Lagrangian_LAMMPS.cpp:90-180). */
void lagrangian_run(Oracle &o, long step) {
    const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz;
    const double dx = o.dx, dt = o.P.dt;
    const double upsilon = 6.0 * kPi * 2 * dx * o.P.mu;
    const double coef = std::sqrt(2.0 * o.P.kB * o.P.T * upsilon * dt);
    const double *u = o.u[o.cur_u].data();
    double *u_f = o.f.data();
    const size_t N = o.N;
    const int NL = 3 * o.M;
    const double lo = o.P.lo[0], hi = o.P.hi[0];
    const int f4 = o.P.f4_quirk;
    /* compute_gamma */
    for (int i = 0; i < NL; i++) o.xgrid[i] = (o.x[i] - lo) / dx;
#pragma omp parallel for
    for (int i = 0; i < o.M; i++) {
        const double *g = &o.xgrid[3 * i];
        o.ftmp[3 * i + 0] = gamma_op(g[0], g[1] - 0.5, g[2] - 0.5, Nx, Ny, Nz, u, f4);
        o.ftmp[3 * i + 1] = gamma_op(g[0] - 0.5, g[1], g[2] - 0.5, Nx, Ny, Nz, u + N, f4);
        o.ftmp[3 * i + 2] = gamma_op(g[0] - 0.5, g[1] - 0.5, g[2], Nx, Ny, Nz, u + 2 * N, f4);
    }
    /* f_tmp = -upsilon (v - gamma u) dt (+ thermal kick) */
    for (int i = 0; i < NL; i++) {
        o.ftmp[i] = -upsilon * (o.v[i] - o.ftmp[i]) * dt;
        if (o.P.noise) {
            double z0, z1;
            normal_pair(o.P.seed, kStreamBead, (uint32_t)step, (uint32_t)i, 0u, z0, z1);
            o.ftmp[i] += z0 * coef;
        }
    }
    /* u_f -= lambda(f_tmp) */
    for (int i = 0; i < o.M; i++) {
        const double *g = &o.xgrid[3 * i];
        lambda_op(g[0], g[1] - 0.5, g[2] - 0.5, Nx, Ny, Nz, o.ftmp[3 * i], u_f, f4);
        lambda_op(g[0] - 0.5, g[1], g[2] - 0.5, Nx, Ny, Nz, o.ftmp[3 * i + 1], u_f + N, f4);
        lambda_op(g[0] - 0.5, g[1] - 0.5, g[2], Nx, Ny, Nz, o.ftmp[3 * i + 2], u_f + 2 * N, f4);
    }
    /* f, x, v update */
    const double L = hi - lo;
    for (int i = 0; i < NL; i++) {
        o.ftmp[i] += dt * o.fc[i];
        o.ftot[i] = o.ftmp[i] / dt;
        double tmp_x = o.v[i] * dt;
        tmp_x += o.x[i];
        if (tmp_x > hi) tmp_x -= L;
        else if (tmp_x < lo) tmp_x += L;
        o.x[i] = tmp_x;
        o.v[i] += o.ftmp[i] / o.m[i / 3];
    }
}

/* Eulerian_Period::run (mpi/SELM_Eulerian_Period.cpp:86-102) on one process. */
void eulerian_run(Oracle &o) {
    const int Nx = o.Nx, Ny = o.Ny, Nz = o.Nz, Xc = o.Xc;
    const size_t N = o.N, Nyx = (size_t)Ny * Nx;
    const double dx = o.dx, dxsq = dx * dx;
    const double dt = o.P.dt, mu = o.P.mu, rho = o.P.rho;
    o.cur_u = (o.cur_u + 1) % 2;
    const double *un = o.u[(o.cur_u + 1) % 2].data();
    double *us = o.u[o.cur_u].data();
    const double *f = o.f.data();
    /* u2uAsterisk: u* = u + (mu lap(u) dt + f) / rho */
    for (int d = 0; d < 3; d++) {
        const double *ud = un + d * N;
        double *usd = us + d * N;
        const double *fd = f + d * N;
#pragma omp parallel for
        for (int k = 0; k < Nz; k++)
            for (int j = 0; j < Ny; j++)
                for (int i = 0; i < Nx; i++) {
                    size_t n = (size_t)k * Nyx + (size_t)j * Nx + i;
                    long i1 = (i != Nx - 1) ? 1 : -i;
                    long j1 = (j != Ny - 1) ? Nx : -(long)j * Nx;
                    long i0 = (i != 0) ? -1 : Nx - 1;
                    long j0 = (j != 0) ? -Nx : (long)(Ny - 1) * Nx;
                    double z_next = (k != Nz - 1) ? ud[n + Nyx] : ud[n - (size_t)k * Nyx];
                    double z_prev = (k != 0) ? ud[n - Nyx] : ud[n + (size_t)(Nz - 1) * Nyx];
                    double lap_u = (ud[n + i1] + ud[n + i0] + ud[n + j1] + ud[n + j0] + z_next + z_prev -
                                    6 * ud[n]) / dxsq;
                    usd[n] = ud[n] + (mu * lap_u * dt + fd[n]) / rho;
                }
    }
    /* div (forward differences) */
    {
        const double *a = us, *b = us + N, *c = us + 2 * N;
#pragma omp parallel for
        for (int k = 0; k < Nz; k++)
            for (int j = 0; j < Ny; j++)
                for (int i = 0; i < Nx; i++) {
                    size_t n = (size_t)k * Nyx + (size_t)j * Nx + i;
                    long i1 = (i != Nx - 1) ? 1 : -i;
                    long j1 = (j != Ny - 1) ? Nx : -(long)j * Nx;
                    long k1 = (k != Nz - 1) ? (long)Nyx : -(long)k * (long)Nyx;
                    o.p[n] = (a[n + i1] - a[n] + b[n + j1] - b[n] + c[n + k1] - c[n]) / dx;
                }
    }
    /* D2Z, d2z_data_transform, Z2D */
    cd *w = o.work.data();
    r2c3(Nx, Ny, Nz, o.p.data(), w);
#pragma omp parallel for
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Xc; i++) {
                size_t n = ((size_t)k * Ny + j) * Xc + i;
                double coef = (2 * (std::cos(2 * kPi * i / Nx) - 1) + 2 * (std::cos(2 * kPi * j / Ny) - 1) +
                               2 * (std::cos(2 * kPi * k / Nz) - 1)) / dxsq;
                if (n != 0) w[n] = cd(w[n].real() / coef, w[n].imag() / coef);
                else w[n] = cd(0.0, 0.0);
            }
    c2r3(Nx, Ny, Nz, w, o.p.data());
    /* uAsterisk2u: u -= (p[n] - p[n - e_d]) / (dx N) */
    const double *pp = o.p.data();
#pragma omp parallel for
    for (int k = 0; k < Nz; k++)
        for (int j = 0; j < Ny; j++)
            for (int i = 0; i < Nx; i++) {
                size_t n = (size_t)k * Nyx + (size_t)j * Nx + i;
                long off[3];
                off[0] = (i != 0) ? -1 : Nx - 1;
                off[1] = (j != 0) ? -Nx : (long)(Ny - 1) * Nx;
                off[2] = (k != 0) ? -(long)Nyx : (long)(Nz - 1) * (long)Nyx;
                for (int d = 0; d < 3; d++) {
                    double g = (pp[n] - pp[n + off[d]]) / (dx * (double)N);
                    us[d * N + n] -= g;
                }
            }
    /* Sum + subtract_mean */
    for (int d = 0; d < 3; d++) {
        double s = 0.0;
        double *ud = us + d * N;
        for (size_t n = 0; n < N; n++) s += ud[n];
        double mean = s / (double)N;
        for (size_t n = 0; n < N; n++) ud[n] -= mean;
    }
}

}  // namespace

/* ====================================================================== */
extern "C" {

void *oracle_create(const oracle_params *p) {
    if (!p || p->nx < 1 || p->ny < 1 || p->nz < 1) return nullptr;
    Oracle *o = new Oracle();
    o->P = *p;
    o->Nx = p->nx; o->Ny = p->ny; o->Nz = p->nz;
    o->Xc = p->nx / 2 + 1;
    o->N = (size_t)p->nx * p->ny * p->nz;
    o->Nc = (size_t)o->Xc * p->ny * p->nz;
    /* dx = (hi[0]-lo[0])/dim[0]: cuda/SELM_Eulerian_Period.cu:18 */
    o->dx = (p->hi[0] - p->lo[0]) / p->nx;
    o->u[0].assign(3 * o->N, 0.0);
    o->u[1].assign(3 * o->N, 0.0);
    o->f.assign(3 * o->N, 0.0);
    o->p.assign(o->N, 0.0);
    o->work.assign(3 * o->Nc, cd(0.0, 0.0));
    return o;
}

void oracle_destroy(void *o) { delete (Oracle *)o; }

int oracle_set_beads(void *op, int M, const double *x, const double *v, const double *m) {
    Oracle &o = *(Oracle *)op;
    if (M < 0) return -1;
    o.M = M;
    o.x.assign(x, x + 3 * M);
    o.v.assign(3 * M, 0.0);
    if (v) o.v.assign(v, v + 3 * M);
    o.m.assign(m, m + M);
    o.fc.assign(3 * M, 0.0);
    o.ftot.assign(3 * M, 0.0);
    o.fext.assign(3 * M, 0.0);
    o.ftmp.assign(3 * M, 0.0);
    o.xgrid.assign(3 * M, 0.0);
    return 0;
}

int oracle_set_bonds(void *op, int nb, const int *ij, const double *K, const double *r0) {
    Oracle &o = *(Oracle *)op;
    o.bi.resize(nb); o.bj.resize(nb); o.bK.resize(nb); o.br0.resize(nb);
    for (int b = 0; b < nb; b++) {
        if (ij[2 * b] < 0 || ij[2 * b] >= o.M || ij[2 * b + 1] < 0 || ij[2 * b + 1] >= o.M) return -1;
        o.bi[b] = ij[2 * b]; o.bj[b] = ij[2 * b + 1]; o.bK[b] = K[b]; o.br0[b] = r0[b];
    }
    return 0;
}

int oracle_set_external_force(void *op, int n, const int *idx, const double *F) {
    Oracle &o = *(Oracle *)op;
    for (int a = 0; a < n; a++) {
        if (idx[a] < 0 || idx[a] >= o.M) return -1;
        for (int d = 0; d < 3; d++) o.fext[3 * idx[a] + d] += F[3 * a + d];
    }
    return 0;
}

int oracle_set_conservative_force(void *op, const double *f) {
    Oracle &o = *(Oracle *)op;
    o.fc.assign(f, f + 3 * o.M);
    return 0;
}

/* Selm::initial (cuda/SELM.cu:55-58): LAMMPS setup(1) wraps positions into
 * [lo, hi) per dimension and computes the forces at x^0; the Eulerian side
 * starts from u = 0 and draws the thermal field used by step 1. */
int oracle_init(void *op) {
    Oracle &o = *(Oracle *)op;
    for (int i = 0; i < o.M; i++)
        for (int d = 0; d < 3; d++) {
            double L = o.P.hi[d] - o.P.lo[d];
            double &xi = o.x[3 * i + d];
            if (xi < o.P.lo[d]) xi += L;
            if (xi >= o.P.hi[d]) xi -= L;
            if (!(xi >= o.P.lo[d] && xi < o.P.hi[d])) {  /* more than one box length outside */
                xi = o.P.lo[d] + ((xi - o.P.lo[d]) - L * std::floor((xi - o.P.lo[d]) / L));
                if (xi >= o.P.hi[d]) xi -= L;
            }
            xi = std::max(xi, o.P.lo[d]);
        }
    bond_forces(o, o.x.data(), o.fc.data());
    std::fill(o.u[0].begin(), o.u[0].end(), 0.0);
    std::fill(o.u[1].begin(), o.u[1].end(), 0.0);
    o.cur_u = 0;
    o.timestep = 0;
    fluid_compute_f(o, 1);
    return 0;
}

int oracle_step(void *op, int nsteps) {
    Oracle &o = *(Oracle *)op;
    for (int s = 0; s < nsteps; s++) {
        o.timestep++;
        lagrangian_run(o, o.timestep);
        eulerian_run(o);
        fluid_compute_f(o, o.timestep + 1);
        bond_forces(o, o.x.data(), o.fc.data());
    }
    return 0;
}

long oracle_timestep(void *op) { return ((Oracle *)op)->timestep; }

void oracle_get_beads(void *op, double *x, double *v, double *f_cons, double *f_total) {
    Oracle &o = *(Oracle *)op;
    size_t b = sizeof(double) * 3 * o.M;
    if (x) std::memcpy(x, o.x.data(), b);
    if (v) std::memcpy(v, o.v.data(), b);
    if (f_cons) std::memcpy(f_cons, o.fc.data(), b);
    if (f_total) std::memcpy(f_total, o.ftot.data(), b);
}

void oracle_get_field(void *op, int which, double *out) {
    Oracle &o = *(Oracle *)op;
    const std::vector<double> &src = (which == 0) ? o.u[o.cur_u] : o.f;
    std::memcpy(out, src.data(), sizeof(double) * 3 * o.N);
}

void oracle_set_field(void *op, int which, const double *in) {
    Oracle &o = *(Oracle *)op;
    std::vector<double> &dst = (which == 0) ? o.u[o.cur_u] : o.f;
    std::memcpy(dst.data(), in, sizeof(double) * 3 * o.N);
}

void oracle_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

double oracle_peskin_phi(double r, int f4) { return phi1(r, f4); }

double oracle_gamma(double rx, double ry, double rz, int nx, int ny, int nz, const double *u, int f4) {
    return gamma_op(rx, ry, rz, nx, ny, nz, u, f4);
}

void oracle_lambda(double rx, double ry, double rz, int nx, int ny, int nz, double f, double *u_f,
                   int f4) {
    lambda_op(rx, ry, rz, nx, ny, nz, f, u_f, f4);
}

void oracle_fft_r2c_3d(int nx, int ny, int nz, const double *in, double *out_cplx) {
    r2c3(nx, ny, nz, in, (cd *)out_cplx);
}

void oracle_fft_c2r_3d(int nx, int ny, int nz, const double *in_cplx, double *out) {
    size_t nc = (size_t)(nx / 2 + 1) * ny * nz;
    std::vector<cd> tmp((const cd *)in_cplx, (const cd *)in_cplx + nc);
    c2r3(nx, ny, nz, tmp.data(), out);
}

void oracle_fft_c2c_1d(int n, int sign, const double *in_cplx, double *out_cplx) {
    FFT1 f(n);
    f.exec((const cd *)in_cplx, 1, (cd *)out_cplx, sign);
}

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox4x32_10(c, key[0], key[1]);
    for (int i = 0; i < 4; i++) out[i] = c[i];
}

void oracle_mode_normals(uint64_t seed, uint32_t stream, uint32_t step, uint32_t index, double out[6]) {
    mode_normals(seed, stream, step, index, out);
}
void oracle_normal_pair(uint64_t seed, uint32_t stream, uint32_t step, uint32_t index, uint32_t slot,
                        double out[2]) {
    normal_pair(seed, stream, step, index, slot, out[0], out[1]);
}

void oracle_bond_forces(void *op, const double *x, double *f) { bond_forces(*(Oracle *)op, x, f); }

// Mean-square displacement of a trajectory, py/Diffusion Coefficient/MSD.py:15-69:
// frames [nframes][3][natoms] float; origins from frame floor(nframes/5);
// lag = dot*mult; per dimension the lag displacement D (FP32) goes to its
// minimum image with X = box[0]/2 (the script's x half-box, used for every
// dimension): (D-2X)^2 if D >= X, (D+2X)^2 if D <= -X, else D^2, each squared
// in FP32; summed in FP64 over beads, origins and dimensions (in that loop
// order); MSD = sum / ((F-lag) * natoms), NaN when no origin is left.
void oracle_msd(const float *frames, int nframes, int natoms, const double box[3], int nlags, int mult,
                double *msd) {
    const int start = nframes / 5;
    const int F = nframes - start;
    const float X = (float)box[0] / 2.0f;
    msd[0] = 0.0;
    for (int dot = 1; dot < nlags; dot++) {
        const int lag = dot * mult;
        const long long Fd = (long long)F - lag;
        if (Fd <= 0) {
            msd[dot] = NAN;
            continue;
        }
        double acc = 0.0;
        for (int a = 0; a < natoms; a++)
            for (long long n = 0; n < Fd; n++)
                for (int d = 0; d < 3; d++) {
                    const float x1 = frames[((size_t)(start + n + lag) * 3 + d) * natoms + a];
                    const float x0 = frames[((size_t)(start + n) * 3 + d) * natoms + a];
                    float D = x1 - x0;
                    if (D >= X) D = D - 2.0f * X;
                    else if (D <= -X) D = D + 2.0f * X;
                    const float sq = D * D;
                    acc += (double)sq;
                }
        msd[dot] = acc / ((double)Fd * natoms);
    }
}

}  // extern "C"
