// selm_engine.cpp -- host side of the C ABI (include/selm_hip.h): replica
// geometry and memory layout, the batched step (work tables, hipGraph), and
// the single-system (selm_ctx) and pool (selm_pool) entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "selm_hip.h"
#include "selm_kernels.hpp"

namespace selm {

static_assert(kStatusTileOverflow == SELM_STATUS_TILE_OVERFLOW && kStatusNonFinite == SELM_STATUS_NONFINITE &&
                  kStatusDisplacement == SELM_STATUS_DISPLACEMENT,
              "device status flags match the C ABI");

constexpr size_t kLdsMax = 65536;
constexpr double kPi = 3.14159265358979323846;
// the step number keys the thermal draws as 32 bits and stamps the spread
// tiles as an int
constexpr long long kMaxTimestep = (1LL << 31) - 2;

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            err = std::string(#expr) + " failed: " + hipGetErrorString(e_);           \
            return SELM_ERR_HIP;                                                      \
        }                                                                             \
    } while (0)

static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

// ---------------------------------------------------------------------------
// Geometry of one grid
// ---------------------------------------------------------------------------
struct Geometry {
    int n[3];
    int xc;
    long long N, Nc;
    RadixPlan plan[3];
    int te[3], tn[3], ntiles;
    int gsplit;  // spread workgroups that take the split tiles
    int bx, by, bz;
};

static bool make_plan(int n, RadixPlan &p) {
    p.n = n;
    p.nst = 0;
    int m = n;
    while (m % 8 == 0 && p.nst < 12) { p.r[p.nst++] = 8; m /= 8; }
    while (m % 4 == 0 && p.nst < 12) { p.r[p.nst++] = 4; m /= 4; }
    while (m % 2 == 0 && p.nst < 12) { p.r[p.nst++] = 2; m /= 2; }
    while (m % 3 == 0 && p.nst < 12) { p.r[p.nst++] = 3; m /= 3; }
    while (m % 5 == 0 && p.nst < 12) { p.r[p.nst++] = 5; m /= 5; }
    return m == 1;
}

static size_t lds_x(const Geometry &G, int twmax, size_t csz) {
    int S = (G.bx / 2) | 1;
    return ((size_t)twmax + 2 * (size_t)(G.n[0] + 2) * S) * csz;
}
static size_t lds_y(const Geometry &G, int twmax, size_t csz) {
    return ((size_t)twmax + 2 * (size_t)G.n[1] * G.by) * csz;
}
static size_t lds_z(const Geometry &G, int twmax, size_t csz) {
    return ((size_t)twmax + 2 * (size_t)G.n[2] * 3 * G.bz) * csz;
}

static int make_geometry(const selm_params &p, bool fp64, Geometry &G, std::string &err) {
    for (int a = 0; a < 3; a++) {
        int n = p.dims[a];
        if (n < 4 || n > 512) {
            err = "grid edge " + std::to_string(n) + " outside [4, 512]";
            return SELM_ERR_UNSUPPORTED;
        }
        G.n[a] = n;
        if (!make_plan(n, G.plan[a])) {
            err = "grid edge " + std::to_string(n) + " is not a product of 2, 3 and 5";
            return SELM_ERR_UNSUPPORTED;
        }
        // tile edge: <= 8 cells, ragged last tile of 0 or >= 2 cells so a
        // stencil (reach 2) can never skip over a whole tile
        int te = n;
        if (n > 8) {
            te = 8;
            for (int c : {8, 7, 6, 5, 4})
                if (n % c != 1) { te = c; break; }
        }
        G.te[a] = te;
        G.tn[a] = (n + te - 1) / te;
    }
    G.xc = G.n[0] / 2 + 1;
    G.N = (long long)G.n[0] * G.n[1] * G.n[2];
    G.Nc = (long long)G.xc * G.n[1] * G.n[2];
    G.ntiles = G.tn[0] * G.tn[1] * G.tn[2];
    G.gsplit = std::max(1, G.ntiles / 8);
    const size_t csz = fp64 ? 16 : 8;
    G.bx = fp64 ? 16 : 32;
    G.by = fp64 ? 8 : 16;
    G.bz = fp64 ? 4 : 8;
    const int twmax = 512;
    while (lds_x(G, twmax, csz) > kLdsMax && G.bx > 2) G.bx /= 2;
    while (lds_y(G, twmax, csz) > kLdsMax && G.by > 1) G.by /= 2;
    while (lds_z(G, twmax, csz) > kLdsMax && G.bz > 1) G.bz /= 2;
    if (lds_x(G, twmax, csz) > kLdsMax || lds_y(G, twmax, csz) > kLdsMax || lds_z(G, twmax, csz) > kLdsMax) {
        err = "grid too large for the LDS budget";
        return SELM_ERR_UNSUPPORTED;
    }
    return SELM_OK;
}

static size_t grid_bytes(const Geometry &G, bool fp64) {
    size_t s = fp64 ? 8 : 4;
    size_t b = 0;
    b += 2 * al(3 * (size_t)G.N * s);     // u, fsp
    b += 3 * al(3 * (size_t)G.Nc * 2 * s);    // spec, spec2, nbuf
    for (int a = 0; a < 3; a++) b += al((size_t)G.n[a] * 2 * s) + al(4 * (size_t)G.n[a] * s);
    if (!fp64)  // FP64 twiddle and wavenumber tables of the FFT passes
        for (int a = 0; a < 3; a++) b += al((size_t)G.n[a] * 2 * 8) + al(4 * (size_t)G.n[a] * 8);
    b += al(8) + al(3 * 8);               // ts, diag
    const size_t tcells = (size_t)G.te[0] * G.te[1] * G.te[2];
    b += al((size_t)G.ntiles * 3 * tcells * s);                   // pool
    b += 2 * al((size_t)G.ntiles * 4) + al((size_t)G.ntiles * 4) + al(8 * 4);
    return b;
}

// nadj: bond entries (2 per bond), nang: angle entries (3 per angle),
// nt: lj/cut atom types, nspec: special-partner entries
// Bucket slots per spread tile for M beads over ntiles tiles: four times a
// generous mean (8 tiles per bead; 3.4 is typical for edge-8 tiles), at
// least kBucketCap and at most kMaxTileBeads, in multiples of 64.  Dense
// systems thus keep their tiles' entries in the buckets; the overflow list
// takes only what a tile holds beyond its bucket.
static int bucket_cap(int M, int ntiles) {
    if (ntiles <= 0) return kBucketCap;
    const double want = 4.0 * 8.0 * (double)M / ntiles;
    const long long c = ((long long)std::ceil(want) + 63) / 64 * 64;
    return (int)std::min<long long>(kMaxTileBeads, std::max<long long>(kBucketCap, c));
}

static size_t bead_bytes(int M, int nadj, bool fp64, int ntiles, int nang = 0, int nt = 0, int nspec = 0,
                         int nbins = 0) {
    size_t b = 0;
    b += al((size_t)ntiles * bucket_cap(M, ntiles) * 48);  // spread buckets
    b += 7 * al(3 * (size_t)M * 8);
    b += al((size_t)M * 8);
    b += al((size_t)kMaxTilesPerBead * M * 8);
    b += al(((size_t)M + 1) * 4);
    b += 2 * al((size_t)nadj * 4) + 2 * al((size_t)nadj * 8);
    b += al(((size_t)M + 1) * 4) + 3 * al((size_t)nang * 4) + 2 * al((size_t)nang * 8);
    b += al((size_t)M * 4) + al((size_t)nt * nt * 4 * 8);
    b += al(((size_t)M + 1) * 4) + al((size_t)nspec * 4) + al((size_t)nspec * 8);
    b += 2 * al((size_t)M * 4) + al((size_t)nbins * 4) + al(((size_t)nbins + 1) * 4);
    b += al((size_t)lj_bin_chunks(nbins) * 4);
    return b;
}

// ---------------------------------------------------------------------------
// One replica (one Selm system)
// ---------------------------------------------------------------------------
struct Replica {
    selm_params P{};
    bool fp64 = true;
    Geometry G{};
    RepDev D{};
    int M = 0, nb = 0;
    char *grid_base = nullptr;
    char *bead_base = nullptr;
    size_t grid_sz = 0, bead_sz = 0;
    std::vector<double> hx, hv, hm, hfext;
    std::vector<int> adj_ptr, adj_i1, adj_i2;
    std::vector<double> adj_K, adj_r0;
    std::vector<int> ang_ptr, ang_a, ang_b, ang_c;  // CSR entries per bead
    std::vector<double> ang_K, ang_t0;
    std::vector<int> ang_raw;                       // [na][3] as given
    std::vector<double> ang_rawK, ang_rawT;
    int lj_nt = 0;
    std::vector<int> lj_type;
    std::vector<double> lj_tab;
    double special_lj[3] = {0, 0, 0};
    std::vector<int> spec_ptr, spec_j;
    std::vector<double> spec_w;
    int ljb_nb[3] = {0, 0, 0}, ljb_n = 0;  // lj/cut neighbour bins (0: all pairs)
    size_t bytes_beads() const {
        return bead_bytes(M, (int)adj_i1.size(), fp64, G.ntiles, (int)ang_a.size(), lj_nt, (int)spec_j.size(),
                          ljb_n);
    }
    bool inited = false;
};

static void carve_grid(Replica &R, char *base) {
    const Geometry &G = R.G;
    size_t s = R.fp64 ? 8 : 4;
    char *p = base;
    auto take = [&](size_t bytes) {
        char *q = p;
        p += al(bytes);
        return q;
    };
    R.grid_base = base;
    R.D.u = take(3 * (size_t)G.N * s);
    R.D.fsp = take(3 * (size_t)G.N * s);
    R.D.spec = take(3 * (size_t)G.Nc * 2 * s);
    R.D.spec2 = take(3 * (size_t)G.Nc * 2 * s);
    R.D.nbuf = take(3 * (size_t)G.Nc * 2 * s);
    for (int a = 0; a < 3; a++) {
        R.D.tw[a] = take((size_t)G.n[a] * 2 * s);
        R.D.ktab[a] = take(4 * (size_t)G.n[a] * s);
    }
    for (int a = 0; a < 3; a++) {
        if (R.fp64) {
            R.D.twd[a] = R.D.tw[a];
            R.D.ktabd[a] = R.D.ktab[a];
        } else {
            R.D.twd[a] = take((size_t)G.n[a] * 2 * 8);
            R.D.ktabd[a] = take(4 * (size_t)G.n[a] * 8);
        }
    }
    R.D.ts = (long long *)take(8);
    R.D.diag = (double *)take(3 * 8);
    const size_t tcells = (size_t)G.te[0] * G.te[1] * G.te[2];
    R.D.pool = take((size_t)G.ntiles * 3 * tcells * s);
    R.D.tstamp = (int *)take((size_t)G.ntiles * 4);
    R.D.bcnt = (int *)take((size_t)G.ntiles * 4);
    R.D.slist = (int *)take((size_t)G.ntiles * 4);
    R.D.ctr = (int *)take(8 * 4);
}

static void carve_beads(Replica &R, char *base) {
    const int M = R.M;
    const size_t nadj = R.adj_i1.size();
    char *p = base;
    auto take = [&](size_t bytes) {
        char *q = p;
        p += al(bytes);
        return q;
    };
    R.bead_base = base;
    R.D.bcap = bucket_cap(M, R.G.ntiles);
    R.D.bucket = (double *)take((size_t)R.G.ntiles * R.D.bcap * 48);
    R.D.x = (double *)take(3 * (size_t)M * 8);
    R.D.v = (double *)take(3 * (size_t)M * 8);
    R.D.fc = (double *)take(3 * (size_t)M * 8);
    R.D.ftot = (double *)take(3 * (size_t)M * 8);
    R.D.fext = (double *)take(3 * (size_t)M * 8);
    R.D.tf = (double *)take(3 * (size_t)M * 8);
    R.D.sg = (double *)take(3 * (size_t)M * 8);
    R.D.m = (double *)take((size_t)M * 8);
    R.D.ovf = (int *)take((size_t)kMaxTilesPerBead * M * 8);
    R.D.adj_ptr = (int *)take(((size_t)M + 1) * 4);
    R.D.adj_i1 = (int *)take(nadj * 4);
    R.D.adj_i2 = (int *)take(nadj * 4);
    R.D.adj_K = (double *)take(nadj * 8);
    R.D.adj_r0 = (double *)take(nadj * 8);
    const size_t nang = R.ang_a.size(), nspec = R.spec_j.size(), nt = (size_t)R.lj_nt;
    R.D.ang_ptr = (int *)take(((size_t)M + 1) * 4);
    R.D.ang_a = (int *)take(nang * 4);
    R.D.ang_b = (int *)take(nang * 4);
    R.D.ang_c = (int *)take(nang * 4);
    R.D.ang_K = (double *)take(nang * 8);
    R.D.ang_t0 = (double *)take(nang * 8);
    R.D.lj_type = (int *)take((size_t)M * 4);
    R.D.lj_tab = (double *)take(nt * nt * 4 * 8);
    R.D.spec_ptr = (int *)take(((size_t)M + 1) * 4);
    R.D.spec_j = (int *)take(nspec * 4);
    R.D.spec_w = (double *)take(nspec * 8);
    R.D.lj_nt = R.lj_nt;
    R.D.ljb_of = (int *)take((size_t)M * 4);
    R.D.ljb_list = (int *)take((size_t)M * 4);
    R.D.ljb_cnt = (int *)take((size_t)R.ljb_n * 4);
    R.D.ljb_start = (int *)take(((size_t)R.ljb_n + 1) * 4);
    R.D.ljb_part = (int *)take((size_t)lj_bin_chunks(R.ljb_n) * 4);
    R.D.ljb_n = R.lj_nt > 0 ? R.ljb_n : 0;
    for (int a = 0; a < 3; a++) {
        R.D.ljb_nb[a] = R.ljb_nb[a];
        R.D.ljb_lo[a] = R.P.lo[a];
    }
}

static int replica_setup(Replica &R, const selm_params &p, bool fp64, std::string &err) {
    R.P = p;
    R.fp64 = fp64;
    int rc = make_geometry(p, fp64, R.G, err);
    if (rc) return rc;
    if (!(p.hi[0] > p.lo[0]) || !(p.dt > 0) || !(p.rho > 0) || !(p.mu >= 0)) {
        err = "invalid box, dt, rho or mu";
        return SELM_ERR_ARG;
    }
    const Geometry &G = R.G;
    RepDev &D = R.D;
    std::memset(&D, 0, sizeof(D));
    for (int a = 0; a < 3; a++) {
        D.n[a] = G.n[a];
        D.plan[a] = G.plan[a];
        D.te[a] = G.te[a];
        D.tn[a] = G.tn[a];
        D.Lb[a] = p.hi[a] - p.lo[a];
    }
    D.xc = G.xc;
    D.N = G.N;
    D.Nc = G.Nc;
    D.bx = G.bx;
    D.by = G.by;
    D.bz = G.bz;
    D.ntiles = G.ntiles;
    // dx = (hi[0]-lo[0])/dim[0] (cuda/SELM_Eulerian_Period.cu:18)
    D.dx = (p.hi[0] - p.lo[0]) / p.dims[0];
    D.lo = p.lo[0];
    D.hi = p.hi[0];
    D.L = p.hi[0] - p.lo[0];
    D.dt = p.dt;
    D.rho = p.rho;
    D.mu = p.mu;
    // cuda/SELM_Lagrangian_LAMMPS.cu:138-139
    D.ups = 6.0 * kPi * 2 * D.dx * p.mu;
    D.bead_coef = std::sqrt(2.0 * p.kB * p.T * D.ups * p.dt);
    // cuda/SELM_Eulerian_Period.cu:141, times N (DFT of the c2r output) / rho (u* += f/rho)
    D.noise_amp = (double)G.N * std::sqrt(2 * p.kB * p.T * p.dt * p.mu / (double)G.N) / p.rho;
    D.inv_dxsq = 1.0 / (D.dx * D.dx);
    D.seed = p.seed;
    D.noise = p.stochastic ? 1 : 0;
    D.f4 = p.f4_quirk ? 1 : 0;
    return SELM_OK;
}

// Twiddle and wavenumber tables of one replica, uploaded once.
static int upload_tables(Replica &R, hipStream_t st, std::string &err) {
    const Geometry &G = R.G;
    for (int a = 0; a < 3; a++) {
        const int n = G.n[a];
        std::vector<double> tw(2 * n), kt(4 * n);
        for (int m = 0; m < n; m++) {
            long double th = 2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)n;
            tw[2 * m] = (double)cosl(th);
            tw[2 * m + 1] = (double)(-sinl(th));
            long double sh = sinl(th / 2);
            if (R.fp64) {
                // lambda term exactly as cuda/SELM_Kernel.cu:19
                kt[4 * m] = 2 * (std::cos(2 * kPi * m / n) - 1);
            } else {
                kt[4 * m] = (double)(-4.0L * sh * sh);
            }
            kt[4 * m + 1] = (double)(-2.0L * sh * sh);  // Re(e^{i th} - 1)
            kt[4 * m + 2] = (double)sinl(th);            // Im(e^{i th} - 1)
            kt[4 * m + 3] = 0.0;
        }
        if (R.fp64) {
            HIPCHK(hipMemcpyAsync((void *)R.D.tw[a], tw.data(), tw.size() * 8, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync((void *)R.D.ktab[a], kt.data(), kt.size() * 8, hipMemcpyHostToDevice, st));
            HIPCHK(hipStreamSynchronize(st));
        } else {
            std::vector<float> twf(tw.begin(), tw.end()), ktf(kt.begin(), kt.end());
            HIPCHK(hipMemcpyAsync((void *)R.D.tw[a], twf.data(), twf.size() * 4, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync((void *)R.D.ktab[a], ktf.data(), ktf.size() * 4, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync((void *)R.D.twd[a], tw.data(), tw.size() * 8, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync((void *)R.D.ktabd[a], kt.data(), kt.size() * 8, hipMemcpyHostToDevice, st));
            HIPCHK(hipStreamSynchronize(st));
        }
    }
    return SELM_OK;
}

}  // namespace selm

extern "C" int selm_check_beads(int M, const double *x, const double *v, const double *m, int *bad) {
    if (bad) *bad = -1;
    if (M < 0 || (M > 0 && (!x || !m))) return SELM_ERR_ARG;
    for (int i = 0; i < M; i++) {
        bool ok = m[i] > 0 && std::isfinite(m[i]);
        for (int d = 0; d < 3; d++)
            ok = ok && std::isfinite(x[3 * (size_t)i + d]) && (!v || std::isfinite(v[3 * (size_t)i + d]));
        if (!ok) {
            if (bad) *bad = i;
            return SELM_ERR_ARG;
        }
    }
    return SELM_OK;
}

namespace selm {

static int host_set_beads(Replica &R, int M, const double *x, const double *v, const double *m,
                          std::string &err) {
    if (M < 0 || (M > 0 && (!x || !m))) {
        err = "selm_set_beads: bad arguments";
        return SELM_ERR_ARG;
    }
    int bad = -1;
    if (selm_check_beads(M, x, v, m, &bad) != SELM_OK) {
        err = "selm_set_beads: bead " + std::to_string(bad) +
              " has a non-finite position or velocity, or a mass that is not finite and > 0";
        return SELM_ERR_ARG;
    }
    R.M = M;
    R.hx.assign(x, x + 3 * (size_t)M);
    if (v) R.hv.assign(v, v + 3 * (size_t)M);
    else R.hv.assign(3 * (size_t)M, 0.0);
    R.hm.assign(m, m + M);
    R.hfext.assign(3 * (size_t)M, 0.0);
    R.adj_ptr.assign((size_t)M + 1, 0);
    R.adj_i1.clear();
    R.adj_i2.clear();
    R.adj_K.clear();
    R.adj_r0.clear();
    R.nb = 0;
    R.ang_raw.clear();
    R.ang_rawK.clear();
    R.ang_rawT.clear();
    R.ang_ptr.assign((size_t)M + 1, 0);
    R.ang_a.clear();
    R.ang_b.clear();
    R.ang_c.clear();
    R.ang_K.clear();
    R.ang_t0.clear();
    R.lj_nt = 0;
    R.lj_type.clear();
    R.lj_tab.clear();
    R.spec_ptr.assign((size_t)M + 1, 0);
    R.spec_j.clear();
    R.spec_w.clear();
    return SELM_OK;
}

// lj/cut special-bond partners (LAMMPS special list from the bond graph):
// 1-2 partners are bonded, 1-3 partners of those, 1-4 of those; a pair takes
// the lowest order it has.  Only partners whose weight is not 1 are stored.
static void build_special(Replica &R) {
    R.spec_ptr.assign((size_t)R.M + 1, 0);
    R.spec_j.clear();
    R.spec_w.clear();
    if (R.lj_nt == 0) return;
    std::vector<std::vector<int>> nb(R.M);
    for (int i = 0; i < R.M; i++)
        for (int q = R.adj_ptr[i]; q < R.adj_ptr[i + 1]; q++) {
            const int j = R.adj_i1[q] == i ? R.adj_i2[q] : R.adj_i1[q];
            if (std::find(nb[i].begin(), nb[i].end(), j) == nb[i].end()) nb[i].push_back(j);
        }
    std::vector<int> order(R.M, 0);
    for (int i = 0; i < R.M; i++) {
        R.spec_ptr[i] = (int)R.spec_j.size();
        std::vector<int> found;
        std::vector<int> frontier{i};
        order[i] = -1;
        for (int o = 1; o <= 3; o++) {
            std::vector<int> next;
            for (int a : frontier)
                for (int b : nb[a])
                    if (order[b] == 0) {
                        order[b] = o;
                        found.push_back(b);
                        next.push_back(b);
                    }
            frontier.swap(next);
        }
        std::sort(found.begin(), found.end());
        for (int j : found) {
            const double w = R.special_lj[order[j] - 1];
            if (w != 1.0) {
                R.spec_j.push_back(j);
                R.spec_w.push_back(w);
            }
        }
        for (int j : found) order[j] = 0;
        order[i] = 0;
    }
    R.spec_ptr[R.M] = (int)R.spec_j.size();
}

static int host_set_angles(Replica &R, int na, const int *ijk, const double *K, const double *t0,
                           std::string &err) {
    if (na < 0 || (na > 0 && (!ijk || !K || !t0))) {
        err = "selm_set_angles: bad arguments";
        return SELM_ERR_ARG;
    }
    std::vector<std::vector<int>> per(R.M);
    for (int q = 0; q < na; q++) {
        const int a = ijk[3 * q], b = ijk[3 * q + 1], c = ijk[3 * q + 2];
        if (a < 0 || b < 0 || c < 0 || a >= R.M || b >= R.M || c >= R.M || a == b || b == c || a == c) {
            err = "selm_set_angles: angle " + std::to_string(q) + " has invalid bead indices";
            return SELM_ERR_ARG;
        }
        per[a].push_back(q);
        per[b].push_back(q);
        per[c].push_back(q);
    }
    R.ang_raw.assign(ijk, ijk + 3 * (size_t)na);
    R.ang_rawK.assign(K, K + na);
    R.ang_rawT.assign(t0, t0 + na);
    R.ang_ptr.assign((size_t)R.M + 1, 0);
    R.ang_a.clear();
    R.ang_b.clear();
    R.ang_c.clear();
    R.ang_K.clear();
    R.ang_t0.clear();
    for (int i = 0; i < R.M; i++) {
        R.ang_ptr[i] = (int)R.ang_a.size();
        for (int q : per[i]) {  // ascending angle index
            R.ang_a.push_back(ijk[3 * q]);
            R.ang_b.push_back(ijk[3 * q + 1]);
            R.ang_c.push_back(ijk[3 * q + 2]);
            R.ang_K.push_back(K[q]);
            R.ang_t0.push_back(t0[q]);
        }
    }
    R.ang_ptr[R.M] = (int)R.ang_a.size();
    return SELM_OK;
}

// Bins for lj/cut: per axis as many as fit with an edge of at least the
// largest cutoff, at most kLjMaxBins; used when every axis has at least 3
// (so the 27 bins around a bead are distinct) and `bins` is set.
constexpr int kLjMaxBins = 64;
static_assert((long long)kLjMaxBins * kLjMaxBins * kLjMaxBins <= 256LL * kLjChunk,
              "the chunk scan of the lj/cut bins takes one workgroup of 256 chunks");
static void lj_bin_geometry(Replica &R, double rcmax, bool bins) {
    R.ljb_n = 0;
    for (int a = 0; a < 3; a++) R.ljb_nb[a] = 0;
    if (!bins || !(rcmax > 0)) return;
    int n = 1;
    for (int a = 0; a < 3; a++) {
        const double La = R.P.hi[a] - R.P.lo[a];
        const int nb = (int)std::min<double>(kLjMaxBins, std::floor(La / rcmax));
        if (nb < 3) return;
        R.ljb_nb[a] = nb;
        n *= nb;
    }
    R.ljb_n = n;
}

static int host_set_lj(Replica &R, int nt, const int *type, const double *eps, const double *sig, const double *rc,
                       const double special[3], bool bins, std::string &err) {
    if (nt < 0 || (nt > 0 && (!type || !eps || !sig || !rc || !special))) {
        err = "selm_set_pair_lj: bad arguments";
        return SELM_ERR_ARG;
    }
    if (nt == 0) {
        R.lj_nt = 0;
        R.lj_type.clear();
        R.lj_tab.clear();
        lj_bin_geometry(R, 0.0, false);
        build_special(R);
        return SELM_OK;
    }
    double Lmin = 1e300;
    for (int a = 0; a < 3; a++) Lmin = std::min(Lmin, R.P.hi[a] - R.P.lo[a]);
    for (int i = 0; i < R.M; i++)
        if (type[i] < 0 || type[i] >= nt) {
            err = "selm_set_pair_lj: bead type out of range";
            return SELM_ERR_ARG;
        }
    R.lj_tab.assign((size_t)nt * nt * 4, 0.0);
    double rcmax = 0.0;
    for (int k = 0; k < nt * nt; k++) {
        if (!(rc[k] >= 0) || 2 * rc[k] >= Lmin) {
            err = "selm_set_pair_lj: cutoff must be >= 0 and below half the box edge";
            return SELM_ERR_ARG;
        }
        R.lj_tab[4 * k] = eps[k];
        R.lj_tab[4 * k + 1] = sig[k];
        R.lj_tab[4 * k + 2] = rc[k] * rc[k];
        rcmax = std::max(rcmax, rc[k]);
    }
    R.lj_nt = nt;
    lj_bin_geometry(R, rcmax, bins);
    R.lj_type.assign(type, type + R.M);
    for (int o = 0; o < 3; o++) R.special_lj[o] = special[o];
    build_special(R);
    return SELM_OK;
}

static int host_set_bonds(Replica &R, int nb, const int *ij, const double *K, const double *r0,
                          std::string &err) {
    if (nb < 0 || (nb > 0 && (!ij || !K || !r0))) {
        err = "selm_set_bonds: bad arguments";
        return SELM_ERR_ARG;
    }
    std::vector<std::vector<int>> per(R.M);
    for (int b = 0; b < nb; b++) {
        int i1 = ij[2 * b], i2 = ij[2 * b + 1];
        if (i1 < 0 || i2 < 0 || i1 >= R.M || i2 >= R.M || i1 == i2) {
            err = "selm_set_bonds: bond " + std::to_string(b) + " has invalid bead indices";
            return SELM_ERR_ARG;
        }
        per[i1].push_back(b);
        per[i2].push_back(b);
    }
    R.adj_ptr.assign((size_t)R.M + 1, 0);
    R.adj_i1.clear();
    R.adj_i2.clear();
    R.adj_K.clear();
    R.adj_r0.clear();
    for (int i = 0; i < R.M; i++) {
        R.adj_ptr[i] = (int)R.adj_i1.size();
        for (int b : per[i]) {  // ascending bond index = LAMMPS bond-list order
            R.adj_i1.push_back(ij[2 * b]);
            R.adj_i2.push_back(ij[2 * b + 1]);
            R.adj_K.push_back(K[b]);
            R.adj_r0.push_back(r0[b]);
        }
    }
    R.adj_ptr[R.M] = (int)R.adj_i1.size();
    R.nb = nb;
    build_special(R);  // the special list follows the bond graph
    return SELM_OK;
}

static int host_set_external(Replica &R, int n, const int *idx, const double *F, std::string &err) {
    if (n < 0 || (n > 0 && (!idx || !F))) {
        err = "selm_set_external_force: bad arguments";
        return SELM_ERR_ARG;
    }
    std::fill(R.hfext.begin(), R.hfext.end(), 0.0);
    for (int a = 0; a < n; a++) {
        if (idx[a] < 0 || idx[a] >= R.M) {
            err = "selm_set_external_force: bead index out of range";
            return SELM_ERR_ARG;
        }
        if (!std::isfinite(F[3 * a]) || !std::isfinite(F[3 * a + 1]) || !std::isfinite(F[3 * a + 2])) {
            err = "selm_set_external_force: non-finite force";
            return SELM_ERR_ARG;
        }
        for (int d = 0; d < 3; d++) R.hfext[3 * (size_t)idx[a] + d] += F[3 * a + d];
    }
    return SELM_OK;
}

// Selm::initial for one replica whose memory is carved: LAMMPS-style wrap of
// x into [lo, hi) per dimension (Domain::pbc during setup(1)), uploads, u = 0.
static int replica_upload_init(Replica &R, hipStream_t st, std::string &err) {
    std::vector<double> xw = R.hx;
    for (int i = 0; i < R.M; i++)
        for (int d = 0; d < 3; d++) {
            double lo = R.P.lo[d], hi = R.P.hi[d], L = hi - lo;
            double &xi = xw[3 * (size_t)i + d];
            if (xi < lo) xi += L;
            if (xi >= hi) xi -= L;
            if (!(xi >= lo && xi < hi)) {  // more than one box length outside
                xi = lo + ((xi - lo) - L * std::floor((xi - lo) / L));
                if (xi >= hi) xi -= L;
            }
            xi = std::max(xi, lo);
        }
    const size_t b3 = 3 * (size_t)R.M * 8;
    if (R.M > 0) {
        HIPCHK(hipMemcpyAsync(R.D.x, xw.data(), b3, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(R.D.v, R.hv.data(), b3, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(R.D.m, R.hm.data(), (size_t)R.M * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(R.D.fext, R.hfext.data(), b3, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemsetAsync(R.D.ftot, 0, b3, st));
        HIPCHK(hipMemsetAsync(R.D.fc, 0, b3, st));
    }
    HIPCHK(hipMemcpyAsync(R.D.adj_ptr, R.adj_ptr.data(), R.adj_ptr.size() * 4, hipMemcpyHostToDevice, st));
    if (!R.adj_i1.empty()) {
        size_t na = R.adj_i1.size();
        HIPCHK(hipMemcpyAsync(R.D.adj_i1, R.adj_i1.data(), na * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(R.D.adj_i2, R.adj_i2.data(), na * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(R.D.adj_K, R.adj_K.data(), na * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(R.D.adj_r0, R.adj_r0.data(), na * 8, hipMemcpyHostToDevice, st));
    }
    if (R.ang_ptr.size() == (size_t)R.M + 1)
        HIPCHK(hipMemcpyAsync(R.D.ang_ptr, R.ang_ptr.data(), R.ang_ptr.size() * 4, hipMemcpyHostToDevice, st));
    if (!R.ang_a.empty()) {
        const size_t n = R.ang_a.size();
        HIPCHK(hipMemcpyAsync(R.D.ang_a, R.ang_a.data(), n * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(R.D.ang_b, R.ang_b.data(), n * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(R.D.ang_c, R.ang_c.data(), n * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(R.D.ang_K, R.ang_K.data(), n * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(R.D.ang_t0, R.ang_t0.data(), n * 8, hipMemcpyHostToDevice, st));
    }
    if (R.lj_nt > 0) {
        HIPCHK(hipMemcpyAsync(R.D.lj_type, R.lj_type.data(), (size_t)R.M * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(R.D.lj_tab, R.lj_tab.data(), R.lj_tab.size() * 8, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(R.D.spec_ptr, R.spec_ptr.data(), R.spec_ptr.size() * 4, hipMemcpyHostToDevice, st));
        if (!R.spec_j.empty()) {
            HIPCHK(hipMemcpyAsync(R.D.spec_j, R.spec_j.data(), R.spec_j.size() * 4, hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(R.D.spec_w, R.spec_w.data(), R.spec_w.size() * 8, hipMemcpyHostToDevice, st));
        }
    }
    const size_t s = R.fp64 ? 8 : 4;
    HIPCHK(hipMemsetAsync(R.D.u, 0, 3 * (size_t)R.G.N * s, st));
    HIPCHK(hipMemsetAsync(R.D.fsp, 0, 3 * (size_t)R.G.N * s, st));
    HIPCHK(hipMemsetAsync(R.D.bcnt, 0, (size_t)R.G.ntiles * 4, st));
    if (R.D.ljb_n > 0) HIPCHK(hipMemsetAsync(R.D.ljb_cnt, 0, (size_t)R.D.ljb_n * 4, st));
    HIPCHK(hipMemsetAsync(R.D.tstamp, 0xff, (size_t)R.G.ntiles * 4, st));
    HIPCHK(hipMemsetAsync(R.D.ctr, 0, 8 * 4, st));
    HIPCHK(hipMemsetAsync(R.D.status, 0, 4, st));
    HIPCHK(hipMemsetAsync(R.D.ts, 0, 8, st));
    HIPCHK(hipMemsetAsync(R.D.diag, 0, 24, st));
    HIPCHK(hipStreamSynchronize(st));
    R.inited = true;
    return SELM_OK;
}

static bool env_is_zero(const char *name) {
    const char *v = std::getenv(name);
    return v && v[0] == '0' && v[1] == 0;
}

static int env_zq_cols() {
    const char *v = std::getenv("SELM_ZQ_COLS");
    return (v && std::strcmp(v, "9") == 0) ? 9 : 17;
}

// a non-negative integer from the environment, else `def`
static int env_count(const char *name, int def) {
    const char *v = std::getenv(name);
    if (!v || !*v) return def;
    char *end = nullptr;
    const long x = std::strtol(v, &end, 10);
    return (end && *end == 0 && x >= 0 && x < (1L << 20)) ? (int)x : def;
}

// ---------------------------------------------------------------------------
// A batch of replicas stepped together
// ---------------------------------------------------------------------------
struct Batch {
    bool fp64 = true;
    bool use_graph = true;
    bool force_generic_fft = false;
    bool allow_plane = true;        // three-pass plane FFT path where the shapes allow it
    int zq_cols = 17;               // plane path: Z-pass columns per workgroup (17 or 9)
    int split_at = kSplitAt;        // spread: tiles of more entries are taken first (0: none)
    int noise_blocks = 0;           // spread launch (plane path): forcing workgroups (0: default)
    int wt_stores = 1;              // plane path: bulk outputs stored at device scope (1) or plainly (0)
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // forked thermal-forcing branch of each step
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::vector<Replica *> reps;
    RepDev *d_reps = nullptr;
    int cap_reps = 0;
    long long *d_tabs = nullptr;
    int cap_tabs = 0;
    StepLaunch L{};
    bool dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // kGraphSteps consecutive steps in one graph: the launch of a graph costs
    // more than the gap between two kernels inside one
    static constexpr int kGraphSteps = 4;
    hipGraph_t graph_n = nullptr;
    hipGraphExec_t exec_n = nullptr;
    // one status word per system slot (RepDev::status points into it), so
    // that a batch's flags come back in one copy
    int *d_status = nullptr;
    int n_status = 0;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (exec_n) (void)hipGraphExecDestroy(exec_n);
        if (graph_n) (void)hipGraphDestroy(graph_n);
        exec = nullptr;
        graph = nullptr;
        exec_n = nullptr;
        graph_n = nullptr;
    }

    static constexpr int kTabs = 8;  // work tables: beads, tiles, points, X, Y, Z, modes, split tiles

    int rebuild(std::string &err) {
        drop_graph();
        const int nrep = (int)reps.size();
        if (nrep > cap_reps) {
            if (d_reps) (void)hipFree(d_reps);
            d_reps = nullptr;
            int cap = std::max(nrep, 2 * cap_reps);
            HIPCHK(hipMalloc(&d_reps, sizeof(RepDev) * cap));
            cap_reps = cap;
        }
        if (kTabs * (nrep + 1) > cap_tabs) {
            if (d_tabs) (void)hipFree(d_tabs);
            d_tabs = nullptr;
            int cap = std::max(kTabs * (nrep + 1), 2 * cap_tabs);
            HIPCHK(hipMalloc(&d_tabs, sizeof(long long) * cap));
            cap_tabs = cap;
        }
        std::vector<RepDev> h(nrep);
        std::vector<long long> tabs(kTabs * (nrep + 1), 0);
        long long *sb = tabs.data(), *stl = sb + (nrep + 1), *sen = stl + (nrep + 1);
        long long *sfx = sen + (nrep + 1), *sfy = sfx + (nrep + 1), *sfz = sfy + (nrep + 1);
        long long *smo = sfz + (nrep + 1), *ssp = smo + (nrep + 1);
        bool any_noise = false;
        const size_t csz = fp64 ? 16 : 8;
        // a split threshold must be a bucket slot (the bead update lists a
        // tile when its count passes it)
        const int sa = split_at > 0 && split_at < kBucketCap ? split_at : 0;
        int twmax = 0;
        for (int r = 0; r < nrep; r++)
            for (int a = 0; a < 3; a++) twmax = std::max(twmax, reps[r]->G.n[a]);
        twmax = (twmax + 1) & ~1;
        size_t lfx = 0, lfy = 0, lfz = 0;
        // compile-time-length FFT kernels when every replica's edges have one
        bool fast = nrep > 0 && !force_generic_fft;
        for (int r = 0; r < nrep; r++) {
            size_t a, b, c;
            fft_fast_lds(fp64, reps[r]->G.n, twmax, &a, &b, &c);
            fast = fast && fft_fast_supported(reps[r]->G.n) && a <= kLdsMax && b <= kLdsMax && c <= kLdsMax;
        }
        // the three-pass plane path (FP32-bulk) when every replica has
        // nx == ny <= 128 and a plane fits in the device's LDS per workgroup
        bool plane = fast && allow_plane && !fp64;
        int okmax = 0;
        size_t lxy = 0, lzq = 0;
        for (int r = 0; r < nrep && plane; r++) {
            plane = fft_plane_shape(reps[r]->G.n);
            okmax = std::max(okmax, fft_plane_okf(reps[r]->G.tn));
        }
        if (plane) {
            for (int r = 0; r < nrep; r++) {
                size_t a, b;
                fft_plane_lds(fp64, reps[r]->G.n, twmax, okmax, zq_cols, &a, &b);
                lxy = std::max(lxy, a);
                lzq = std::max(lzq, b);
            }
            int dev = 0, lim = 0;
            HIPCHK(hipGetDevice(&dev));
            HIPCHK(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
            plane = lxy <= (size_t)lim && lzq <= kLdsMax;
            if (plane && lxy > kLdsMax && fft_plane_prepare(fp64, twmax, lxy) != hipSuccess) {
                (void)hipGetLastError();
                plane = false;
            }
        }
        for (int r = 0; r < nrep; r++) {
            const Replica &R = *reps[r];
            const Geometry &G = R.G;
            h[r] = R.D;
            h[r].M = R.M;
            h[r].split_at = R.M >= kSplitMinBeads ? sa : 0;
            sb[r + 1] = sb[r] + R.M;
            stl[r + 1] = stl[r] + G.ntiles;
            sen[r + 1] = sen[r] + G.N;
            smo[r + 1] = smo[r] + G.Nc;
            ssp[r + 1] = ssp[r] + (R.M >= kSplitMinBeads && sa > 0 ? G.gsplit : 0);
            any_noise = any_noise || R.D.noise;
            if (plane) {
                long long ixy, iz;
                fft_plane_items(fp64, G.n, zq_cols, &ixy, &iz);
                sfx[r + 1] = sfx[r] + ixy;
                sfy[r + 1] = sfy[r];
                sfz[r + 1] = sfz[r] + iz;
                lfx = lxy;
                lfz = lzq;
            } else if (fast) {
                long long ix, iy, iz;
                fft_fast_items(fp64, G.n, &ix, &iy, &iz);
                sfx[r + 1] = sfx[r] + ix;
                sfy[r + 1] = sfy[r] + iy;
                sfz[r + 1] = sfz[r] + iz;
                size_t a, b, c;
                fft_fast_lds(fp64, G.n, twmax, &a, &b, &c);
                lfx = std::max(lfx, a);
                lfy = std::max(lfy, b);
                lfz = std::max(lfz, c);
            } else {
                long long rows = 3LL * G.n[1] * G.n[2];
                sfx[r + 1] = sfx[r] + (rows + G.bx - 1) / G.bx;
                sfy[r + 1] = sfy[r] + 3LL * G.n[2] * ((G.xc + G.by - 1) / G.by);
                sfz[r + 1] = sfz[r] + ((long long)G.n[1] * G.xc + G.bz - 1) / G.bz;
                lfx = std::max(lfx, lds_x(G, twmax, csz));
                lfy = std::max(lfy, lds_y(G, twmax, csz));
                lfz = std::max(lfz, lds_z(G, twmax, csz));
            }
        }
        if ((!plane && lfx > kLdsMax) || lfy > kLdsMax || lfz > kLdsMax) {
            err = "batch exceeds the LDS budget";
            return SELM_ERR_UNSUPPORTED;
        }
        if (nrep > 0) {
            HIPCHK(hipMemcpyAsync(d_reps, h.data(), sizeof(RepDev) * nrep, hipMemcpyHostToDevice, stream));
            HIPCHK(hipMemcpyAsync(d_tabs, tabs.data(), sizeof(long long) * tabs.size(), hipMemcpyHostToDevice,
                                  stream));
            HIPCHK(hipStreamSynchronize(stream));
        }
        L.reps = d_reps;
        L.nrep = nrep;
        L.fp64 = fp64;
        L.st_bead = d_tabs;
        L.st_tile = d_tabs + (nrep + 1);
        L.st_pts = d_tabs + 2 * (nrep + 1);
        L.st_fx = d_tabs + 3 * (nrep + 1);
        L.st_fy = d_tabs + 4 * (nrep + 1);
        L.st_fz = d_tabs + 5 * (nrep + 1);
        L.st_modes = d_tabs + 6 * (nrep + 1);
        L.st_split = d_tabs + 7 * (nrep + 1);
        L.tot_split = ssp[nrep];
        L.tot_bead = sb[nrep];
        L.tot_tile = stl[nrep];
        L.tot_pts = sen[nrep];
        L.tot_fx = sfx[nrep];
        L.tot_fy = sfy[nrep];
        L.tot_fz = sfz[nrep];
        L.tot_modes = (fast && any_noise) ? smo[nrep] : 0;
        L.st2 = stream2;
        L.ev_fork = ev_fork;
        L.ev_join = ev_join;
        L.lds_fx = lfx;
        L.lds_fy = lfy;
        L.lds_fz = lfz;
        L.twmax = twmax;
        L.fast = fast;
        L.lj_bins = false;
        L.lj_maxbins = 0;
        for (int r = 0; r < nrep; r++) {
            L.lj_bins = L.lj_bins || reps[r]->D.ljb_n > 0;
            L.lj_maxbins = std::max(L.lj_maxbins, reps[r]->D.ljb_n);
        }
        L.plane = plane;
        L.wt = plane && wt_stores;
        L.aux_bonds = plane;
        for (int r = 0; r < nrep; r++) L.aux_bonds = L.aux_bonds && reps[r]->ang_a.empty() && reps[r]->lj_nt == 0;
        L.okmax = okmax;
        L.zq_cols = zq_cols;
        L.split_at = sa;
        L.noise_blocks = noise_blocks;
        dirty = false;
        return SELM_OK;
    }

    // both streams at the default priority; nslots status words
    int init_streams(int nslots, std::string &err) {
        HIPCHK(hipMalloc(&d_status, sizeof(int) * nslots));
        HIPCHK(hipMemset(d_status, 0, sizeof(int) * nslots));
        n_status = nslots;
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));

        return SELM_OK;
    }

    int ensure(std::string &err) {
        if (dirty) return rebuild(err);
        return SELM_OK;
    }

    int step(int nsteps, hipStream_t st, std::string &err) {
        int rc = ensure(err);
        if (rc) return rc;
        if (reps.empty() || nsteps <= 0) return SELM_OK;
        if (!st) st = stream;
        if (use_graph) {
            if (!exec) {
                HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                launch_step(L, stream);
                hipGraph_t g = nullptr;
                HIPCHK(hipStreamEndCapture(stream, &g));
                graph = g;
                HIPCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            }
            if (nsteps >= kGraphSteps && !exec_n) {
                HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                for (int s = 0; s < kGraphSteps; s++) launch_step(L, stream);
                hipGraph_t g = nullptr;
                HIPCHK(hipStreamEndCapture(stream, &g));
                graph_n = g;
                HIPCHK(hipGraphInstantiate(&exec_n, graph_n, nullptr, nullptr, 0));
            }
            int s = 0;
            for (; s + kGraphSteps <= nsteps; s += kGraphSteps) HIPCHK(hipGraphLaunch(exec_n, st));
            for (; s < nsteps; s++) HIPCHK(hipGraphLaunch(exec, st));
        } else {
            for (int s = 0; s < nsteps; s++) launch_step(L, st);
            HIPCHK(hipGetLastError());
        }
        return SELM_OK;
    }

    void release() {
        drop_graph();
        if (d_reps) (void)hipFree(d_reps);
        if (d_tabs) (void)hipFree(d_tabs);
        d_reps = nullptr;
        d_tabs = nullptr;
        if (d_status) (void)hipFree(d_status);
        d_status = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        if (stream2) (void)hipStreamDestroy(stream2);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        stream = nullptr;
        stream2 = nullptr;
        ev_fork = ev_join = nullptr;
    }
};

// Text of a status word (SELM_STATUS_* flags).
static std::string status_text(int f) {
    std::string t;
    if (f & kStatusTileOverflow) t += " tile-overflow(a tile met more than 16384 beads: spread sums out of range)";
    if (f & kStatusNonFinite) t += " non-finite(a bead position or velocity is NaN or infinite)";
    if (f & kStatusDisplacement) t += " displacement(|v dt| exceeded half the box length)";
    return t.empty() ? std::string(" none") : t;
}

static int download_beads(Replica &R, hipStream_t st, double *x, double *v, double *fc, double *ft,
                          std::string &err) {
    const size_t b3 = 3 * (size_t)R.M * 8;
    HIPCHK(hipStreamSynchronize(st));
    if (R.M == 0) return SELM_OK;
    if (x) HIPCHK(hipMemcpy(x, R.D.x, b3, hipMemcpyDeviceToHost));
    if (v) HIPCHK(hipMemcpy(v, R.D.v, b3, hipMemcpyDeviceToHost));
    if (fc) HIPCHK(hipMemcpy(fc, R.D.fc, b3, hipMemcpyDeviceToHost));
    if (ft) HIPCHK(hipMemcpy(ft, R.D.ftot, b3, hipMemcpyDeviceToHost));
    return SELM_OK;
}

static int download_field(Replica &R, hipStream_t st, int comp, double *out, std::string &err) {
    if (comp < 0 || comp > 2 || !out) {
        err = "get_field: bad arguments";
        return SELM_ERR_ARG;
    }
    HIPCHK(hipStreamSynchronize(st));
    const size_t N = (size_t)R.G.N;
    if (R.fp64) {
        HIPCHK(hipMemcpy(out, (double *)R.D.u + comp * N, N * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(N);
        HIPCHK(hipMemcpy(tmp.data(), (float *)R.D.u + comp * N, N * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; i++) out[i] = tmp[i];
    }
    return SELM_OK;
}

static int upload_field(Replica &R, hipStream_t st, int comp, const double *u, std::string &err) {
    if (comp < 0 || comp > 2 || !u) {
        err = "set_field: bad arguments";
        return SELM_ERR_ARG;
    }
    HIPCHK(hipStreamSynchronize(st));
    const size_t N = (size_t)R.G.N;
    if (R.fp64) {
        HIPCHK(hipMemcpy((double *)R.D.u + comp * N, u, N * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp(u, u + N);
        HIPCHK(hipMemcpy((float *)R.D.u + comp * N, tmp.data(), N * 4, hipMemcpyHostToDevice));
    }
    return SELM_OK;
}

// Set the count of completed steps to n, which keys the thermal draws of the
// next step (n + 1).  The per-step bookkeeping of the spread is reset with
// it: tile stamps left by an earlier run at a step number the system will
// reach again would otherwise read as current, and the overflow counters are
// indexed by the step's parity.
static int replica_set_timestep(Replica &R, hipStream_t st, long long n, std::string &err) {
    if (n < 0 || n > kMaxTimestep) {
        err = "set_timestep: n = " + std::to_string(n) + " outside [0, " + std::to_string(kMaxTimestep) + "]";
        return SELM_ERR_ARG;
    }
    if (!R.inited) {
        err = "set_timestep before init";
        return SELM_ERR_STATE;
    }
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipMemsetAsync(R.D.tstamp, 0xff, (size_t)R.G.ntiles * 4, st));
    HIPCHK(hipMemsetAsync(R.D.bcnt, 0, (size_t)R.G.ntiles * 4, st));
    HIPCHK(hipMemsetAsync(R.D.ctr, 0, 8 * 4, st));
    HIPCHK(hipMemcpyAsync(R.D.ts, &n, 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return SELM_OK;
}

}  // namespace selm

using namespace selm;

// ===========================================================================
// selm_ctx: one system
// ===========================================================================
// ---------------------------------------------------------------------------
// Autotuning of the FFT pass shapes and of the spread launch's shape
// ---------------------------------------------------------------------------
// First the FFT shape: the three-pass plane path with 17 or 9 Z-pass columns
// per workgroup, and the five-pass transposing path.  Then, with that FFT
// shape, the spread's: the split threshold (tiles of more bucket entries are
// listed by the bead update and taken first by their own workgroups; 32, 64,
// 128 or no list) and, on the plane path, the number of thermal-forcing
// workgroups sharing the spread launch (the default for the batch's modes,
// half or twice it).  The spread sums are order-independent fixed-point
// sums and each forcing mode is drawn by one thread whatever the grid, so
// the spread candidates give identical bits; the FFT shapes agree to FP32
// rounding.
//
// Each candidate is timed over `steps` graph replays after kGraphSteps + 1
// warm-up steps (which instantiate both graphs), in each of kTuneRounds
// rounds that interleave the candidates, and keeps its fastest round (the
// shapes differ by a few percent, less than one round's spread).  The
// device memory holding the batch's state (the regions in `mem`) is copied
// aside first and restored after every candidate, so tuning leaves the
// systems exactly as it found them.  The chosen shape is cached per
// (precision, grid shapes, bead counts) for the process; a cached shape is
// applied without timing.
struct TuneRegion {
    void *ptr;
    size_t bytes;
};

struct TuneShape {
    int passes, zcols;  // FFT: 3 (plane) with zcols 17 or 9, or 5
    int split_at;       // spread: split threshold (0: no split list)
    int noise_blocks;   // spread launch (plane path): forcing workgroups (0: default)
    int wt;             // plane path: bulk outputs at device scope (1) or plain (0)
};

static std::mutex g_tune_mu;
static std::map<std::string, TuneShape> g_tune_cache;

static std::string tune_key(const Batch &B) {
    std::string k = B.fp64 ? "fp64" : "fp32";
    for (const Replica *r : B.reps)
        k += ":" + std::to_string(r->G.n[0]) + "x" + std::to_string(r->G.n[1]) + "x" + std::to_string(r->G.n[2]) +
             "/" + std::to_string(r->M);
    return k;
}

static void apply_shape(Batch &B, const TuneShape &t) {
    B.allow_plane = t.passes == 3;
    if (t.zcols) B.zq_cols = t.zcols;
    B.split_at = t.split_at;
    B.noise_blocks = t.noise_blocks;
    B.wt_stores = t.wt;
    B.dirty = true;
}

static void report_shape(const Batch &B, int shape[2]) {
    shape[0] = B.L.fast ? (B.L.plane ? 3 : 5) : 0;
    shape[1] = B.L.plane ? B.L.zq_cols : 0;
}

// Times each available candidate (kTuneRounds interleaved rounds, best round
// each); returns the index of the fastest in *best and its time in *best_ms.
// `avail[c]` is decided in the first round from the shape the batch takes.
static int tune_candidates(Batch &B, const std::vector<TuneRegion> &mem, void *save, int steps,
                           const std::vector<TuneShape> &cand, bool fft_stage, int *best, double *best_ms,
                           std::string &err) {
    constexpr int kTuneRounds = 3;
    hipStream_t st = B.stream;
    const int nc = (int)cand.size();
    std::vector<double> cms(nc, -1.0);
    std::vector<bool> avail(nc, false);
    int rc = SELM_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    for (int round = 0; round < kTuneRounds && rc == SELM_OK; round++) {
        for (int c = 0; c < nc && rc == SELM_OK; c++) {
            apply_shape(B, cand[c]);
            rc = B.ensure(err);
            if (rc) break;
            if (round == 0) {
                if (fft_stage) {
                    // a shape the batch cannot take falls back to another;
                    // the five-pass candidate (last) is the fallback
                    const int got = B.L.fast ? (B.L.plane ? 3 : 5) : 0;
                    bool any = false;
                    for (int q = 0; q < c; q++) any = any || avail[q];
                    avail[c] = got == cand[c].passes || (c == nc - 1 && !any);
                } else {
                    avail[c] = true;
                }
            }
            if (!avail[c]) continue;
            // warm-up long enough to capture and instantiate both graphs
            // (one step and kGraphSteps steps), so the timed region below
            // holds graph launches only
            rc = B.step(Batch::kGraphSteps + 1, st, err);
            if (rc) break;
            if (B.use_graph && (!B.exec || !B.exec_n)) {
                err = "autotune: graphs not instantiated before timing";
                rc = SELM_ERR_STATE;
                break;
            }
            HIPCHK(hipEventRecord(e0, st));
            rc = B.step(steps, st, err);
            if (rc) break;
            HIPCHK(hipEventRecord(e1, st));
            HIPCHK(hipEventSynchronize(e1));
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, e0, e1));
            const double per = (double)t / steps;
            if (cms[c] < 0.0 || per < cms[c]) cms[c] = per;
            size_t o = 0;
            for (const auto &m : mem) {
                HIPCHK(hipMemcpyAsync(m.ptr, (char *)save + o, m.bytes, hipMemcpyDeviceToDevice, st));
                o += m.bytes;
            }
            HIPCHK(hipStreamSynchronize(st));
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *best = -1;
    for (int c = 0; c < nc; c++)
        if (cms[c] >= 0.0 && (*best < 0 || cms[c] < *best_ms)) {
            *best = c;
            *best_ms = cms[c];
        }
    if (*best < 0) {
        err = "autotune: no candidate shape could be timed";
        return SELM_ERR_STATE;
    }
    return SELM_OK;
}

static int autotune_batch(Batch &B, const std::vector<TuneRegion> &mem, int steps, int shape[2], double *ms,
                          std::string &err) {
    if (steps < 1) return SELM_ERR_ARG;
    const std::string key = tune_key(B);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
            apply_shape(B, it->second);
            int rc = B.ensure(err);
            if (rc) return rc;
            report_shape(B, shape);
            *ms = 0.0;
            return SELM_OK;
        }
    }
    hipStream_t st = B.stream;
    size_t total = 0;
    for (const auto &m : mem) total += m.bytes;
    void *save = nullptr;
    HIPCHK(hipMalloc(&save, std::max<size_t>(total, 256)));
    size_t off = 0;
    for (const auto &m : mem) {
        HIPCHK(hipMemcpyAsync((char *)save + off, m.ptr, m.bytes, hipMemcpyDeviceToDevice, st));
        off += m.bytes;
    }
    HIPCHK(hipStreamSynchronize(st));
    const TuneShape base{3, 17, B.split_at, B.noise_blocks, B.wt_stores};
    std::vector<TuneShape> fft = {{3, 17, base.split_at, base.noise_blocks, base.wt},
                                  {3, 9, base.split_at, base.noise_blocks, base.wt},
                                  {5, 0, base.split_at, base.noise_blocks, base.wt}};
    int best = -1;
    double best_ms = 0.0;
    int rc = tune_candidates(B, mem, save, steps, fft, true, &best, &best_ms, err);
    TuneShape chosen = best >= 0 ? fft[best] : base;
    if (rc == SELM_OK) {
        // the spread's shape, with the chosen FFT shape
        apply_shape(B, chosen);
        rc = B.ensure(err);
        if (rc == SELM_OK) {
            const bool plane = B.L.plane && B.L.tot_modes > 0;
            const int nb0 = plane ? spread_noise_default(B.L.tot_modes) : 0;
            std::vector<TuneShape> sp;
            for (int sa : {kSplitAt, 32, 128, 0}) {
                TuneShape t = chosen;
                t.split_at = sa;
                t.noise_blocks = 0;
                sp.push_back(t);
            }
            if (plane) {
                for (int nb : {nb0 / 2, nb0 * 2}) {
                    TuneShape t = chosen;
                    t.split_at = kSplitAt;
                    t.noise_blocks = std::max(64, nb);
                    sp.push_back(t);
                }
            }
            int sbest = -1;
            double sms = 0.0;
            rc = tune_candidates(B, mem, save, steps, sp, false, &sbest, &sms, err);
            if (rc == SELM_OK) {
                chosen = sp[sbest];
                best_ms = sms;
            }
            // then the plane path's store shape: device-scope stores shorten
            // the gaps between launches, but cost write traffic where the
            // Z pass's column runs are short (small grids)
            if (rc == SELM_OK && B.L.plane) {
                std::vector<TuneShape> ws = {chosen, chosen};
                ws[0].wt = 1;
                ws[1].wt = 0;
                int wbest = -1;
                double wms = 0.0;
                rc = tune_candidates(B, mem, save, steps, ws, false, &wbest, &wms, err);
                if (rc == SELM_OK) {
                    chosen = ws[wbest];
                    best_ms = wms;
                }
            }
        }
    }
    (void)hipFree(save);
    if (rc) return rc;
    apply_shape(B, chosen);
    rc = B.ensure(err);
    if (rc) return rc;
    report_shape(B, shape);
    *ms = best_ms;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = chosen;
    return SELM_OK;
}

struct selm_ctx {
    Batch batch;
    Replica rep;
    void *grid_mem = nullptr;
    void *bead_mem = nullptr;
    std::string err;
    hipStream_t last = nullptr;
};

#define CTX_CHECK(c)                  \
    do {                              \
        if (!(c)) return SELM_ERR_ARG; \
    } while (0)

extern "C" {

int selm_params_default(selm_params *p) {
    if (!p) return SELM_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    p->dims[0] = p->dims[1] = p->dims[2] = 32;
    for (int d = 0; d < 3; d++) {
        p->lo[d] = 0.0;
        p->hi[d] = 32.0;
    }
    p->rho = p->mu = p->kB = p->T = 1.0;
    p->dt = 0.01;
    p->seed = 123456;
    p->stochastic = 0;
    p->precision = SELM_PREC_FP64;
    p->f4_quirk = 1;
    return SELM_OK;
}

int selm_create(const selm_params *p, selm_ctx **out) {
    if (!p || !out) return SELM_ERR_ARG;
    *out = nullptr;
    if (p->precision != SELM_PREC_FP64 && p->precision != SELM_PREC_FP32_BULK) return SELM_ERR_ARG;
    selm_ctx *c = new (std::nothrow) selm_ctx();
    if (!c) return SELM_ERR_NOMEM;
    std::string &err = c->err;
    const bool fp64 = p->precision == SELM_PREC_FP64;
    int rc = replica_setup(c->rep, *p, fp64, err);
    if (rc) {
        delete c;
        return rc;
    }
    c->batch.fp64 = fp64;
    // SELM_FFT_GENERIC=1 selects the runtime-length FFT kernels for every size
    c->batch.force_generic_fft = std::getenv("SELM_FFT_GENERIC") != nullptr;
    // SELM_FFT_PLANES=0 keeps the five-pass transposing path for every shape;
    // SELM_ZQ_COLS=9 starts the plane path with 9 Z-pass columns per workgroup
    c->batch.allow_plane = !env_is_zero("SELM_FFT_PLANES");
    c->batch.zq_cols = env_zq_cols();
    c->batch.split_at = env_count("SELM_SPLIT_AT", kSplitAt);
    c->batch.noise_blocks = env_count("SELM_FORCING_WG", 0);
    c->batch.wt_stores = env_count("SELM_WT_STORES", 1) ? 1 : 0;
    if (c->batch.init_streams(1, err) != SELM_OK) {
        c->batch.release();
        delete c;
        return SELM_ERR_HIP;
    }
    size_t gb = grid_bytes(c->rep.G, fp64);
    hipError_t e = hipMalloc(&c->grid_mem, gb);
    if (e != hipSuccess) {
        c->batch.release();
        delete c;
        return SELM_ERR_NOMEM;
    }
    c->rep.grid_sz = gb;
    carve_grid(c->rep, (char *)c->grid_mem);
    c->rep.D.status = c->batch.d_status;
    rc = upload_tables(c->rep, c->batch.stream, err);
    if (rc) {
        (void)hipFree(c->grid_mem);
        c->batch.release();
        delete c;
        return rc;
    }
    // zero-bead system until selm_set_beads
    host_set_beads(c->rep, 0, nullptr, nullptr, nullptr, err);
    c->batch.reps.push_back(&c->rep);
    c->last = c->batch.stream;
    *out = c;
    return SELM_OK;
}

int selm_destroy(selm_ctx *c) {
    if (!c) return SELM_ERR_ARG;
    if (c->batch.stream) (void)hipStreamSynchronize(c->batch.stream);
    if (c->last && c->last != c->batch.stream) (void)hipStreamSynchronize(c->last);
    c->batch.release();
    if (c->grid_mem) (void)hipFree(c->grid_mem);
    if (c->bead_mem) (void)hipFree(c->bead_mem);
    delete c;
    return SELM_OK;
}

const char *selm_last_error(const selm_ctx *c) { return c ? c->err.c_str() : "null ctx"; }

static int ctx_alloc_beads(selm_ctx *c) {
    std::string &err = c->err;
    if (c->bead_mem) {
        HIPCHK(hipStreamSynchronize(c->batch.stream));
        (void)hipFree(c->bead_mem);
        c->bead_mem = nullptr;
    }
    size_t bb = c->rep.bytes_beads();
    c->rep.bead_sz = bb;
    if (hipMalloc(&c->bead_mem, std::max<size_t>(bb, 256)) != hipSuccess) {
        err = "hipMalloc of bead buffers failed";
        return SELM_ERR_NOMEM;
    }
    carve_beads(c->rep, (char *)c->bead_mem);
    c->batch.dirty = true;
    c->rep.inited = false;
    return SELM_OK;
}

int selm_set_beads(selm_ctx *c, int M, const double *x, const double *v, const double *m) {
    CTX_CHECK(c);
    int rc = host_set_beads(c->rep, M, x, v, m, c->err);
    if (rc) return rc;
    return ctx_alloc_beads(c);
}

int selm_set_bonds(selm_ctx *c, int nb, const int *ij, const double *K, const double *r0) {
    CTX_CHECK(c);
    int rc = host_set_bonds(c->rep, nb, ij, K, r0, c->err);
    if (rc) return rc;
    return ctx_alloc_beads(c);
}

int selm_set_angles(selm_ctx *c, int na, const int *ijk, const double *K, const double *theta0) {
    CTX_CHECK(c);
    int rc = host_set_angles(c->rep, na, ijk, K, theta0, c->err);
    if (rc) return rc;
    return ctx_alloc_beads(c);
}

int selm_set_pair_lj(selm_ctx *c, int ntypes, const int *type, const double *eps, const double *sigma,
                     const double *rc, const double special_lj[3]) {
    CTX_CHECK(c);
    // SELM_LJ_BINS=0 keeps the all-pairs loop (a test switch)
    int r = host_set_lj(c->rep, ntypes, type, eps, sigma, rc, special_lj, !env_is_zero("SELM_LJ_BINS"), c->err);
    if (r) return r;
    return ctx_alloc_beads(c);
}

int selm_set_external_force(selm_ctx *c, int n, const int *idx, const double *F) {
    CTX_CHECK(c);
    int rc = host_set_external(c->rep, n, idx, F, c->err);
    if (rc) return rc;
    c->rep.inited = false;
    return SELM_OK;
}

int selm_set_conservative_force(selm_ctx *c, const double *f) {
    CTX_CHECK(c);
    std::string &err = c->err;
    if (!c->rep.inited) {
        err = "selm_set_conservative_force before selm_init";
        return SELM_ERR_STATE;
    }
    if (c->rep.M == 0) return SELM_OK;
    if (!f) return SELM_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->last));
    HIPCHK(hipMemcpy(c->rep.D.fc, f, 3 * (size_t)c->rep.M * 8, hipMemcpyHostToDevice));
    return SELM_OK;
}

int selm_init(selm_ctx *c) {
    CTX_CHECK(c);
    std::string &err = c->err;
    if (!c->bead_mem) {
        int rc = ctx_alloc_beads(c);
        if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(c->last));
    int rc = replica_upload_init(c->rep, c->batch.stream, err);
    if (rc) return rc;
    rc = c->batch.ensure(err);
    if (rc) return rc;
    launch_bond_forces(c->batch.L, 0, c->batch.stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->batch.stream));
    c->last = c->batch.stream;
    return SELM_OK;
}

int selm_step(selm_ctx *c, int nsteps, void *stream) {
    CTX_CHECK(c);
    std::string &err = c->err;
    if (!c->rep.inited) {
        err = "selm_step before selm_init";
        return SELM_ERR_STATE;
    }
    if (nsteps < 0) return SELM_ERR_ARG;
    hipStream_t st = stream ? (hipStream_t)stream : c->batch.stream;
    if (st != c->last) HIPCHK(hipStreamSynchronize(c->last));
    c->last = st;
    return c->batch.step(nsteps, st, err);
}

int selm_sync(selm_ctx *c) {
    CTX_CHECK(c);
    std::string &err = c->err;
    HIPCHK(hipStreamSynchronize(c->last));
    int f = 0;
    HIPCHK(hipMemcpy(&f, c->batch.d_status, sizeof(int), hipMemcpyDeviceToHost));
    if (f) {
        err = "device status flags " + std::to_string(f) + ":" + status_text(f);
        return SELM_ERR_DEVICE;
    }
    return SELM_OK;
}

int selm_status(selm_ctx *c, int *flags, int clear) {
    CTX_CHECK(c);
    if (!flags) return SELM_ERR_ARG;
    std::string &err = c->err;
    HIPCHK(hipStreamSynchronize(c->last));
    HIPCHK(hipMemcpy(flags, c->batch.d_status, sizeof(int), hipMemcpyDeviceToHost));
    if (clear) HIPCHK(hipMemset(c->batch.d_status, 0, sizeof(int)));
    return SELM_OK;
}

int selm_get_beads(selm_ctx *c, double *x, double *v, double *fc, double *ft) {
    CTX_CHECK(c);
    if (!c->bead_mem) return SELM_OK;
    return download_beads(c->rep, c->last, x, v, fc, ft, c->err);
}

int selm_get_field(selm_ctx *c, int comp, double *u) {
    CTX_CHECK(c);
    return download_field(c->rep, c->last, comp, u, c->err);
}

int selm_set_field(selm_ctx *c, int comp, const double *u) {
    CTX_CHECK(c);
    std::string &err = c->err;
    if (comp < 0 || comp > 2 || !u) return SELM_ERR_ARG;
    if (!c->rep.D.u) {
        err = "selm_set_field before the grid is allocated (selm_set_beads)";
        return SELM_ERR_STATE;
    }
    HIPCHK(hipStreamSynchronize(c->last));
    return upload_field(c->rep, c->last, comp, u, err);
}

int selm_set_timestep(selm_ctx *c, long long n) {
    CTX_CHECK(c);
    std::string &err = c->err;
    HIPCHK(hipStreamSynchronize(c->last));
    return replica_set_timestep(c->rep, c->last, n, err);
}

long long selm_timestep(selm_ctx *c) {
    if (!c || !c->rep.D.ts) return -1;
    long long ts = -1;
    if (hipStreamSynchronize(c->last) != hipSuccess) return -1;
    if (hipMemcpy(&ts, c->rep.D.ts, 8, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return ts;
}

int selm_footprint(const selm_params *p, int M, int nb, size_t *bytes) {
    if (!p || !bytes || M < 0 || nb < 0) return SELM_ERR_ARG;
    Geometry G;
    std::string err;
    const bool fp64 = p->precision == SELM_PREC_FP64;
    int rc = make_geometry(*p, fp64, G, err);
    if (rc) return rc;
    *bytes = grid_bytes(G, fp64) + bead_bytes(M, 2 * nb, fp64, G.ntiles);
    return SELM_OK;
}

int selm_autotune(selm_ctx *c, int steps, int shape[2], double *ms) {
    CTX_CHECK(c);
    if (!shape || !ms) return SELM_ERR_ARG;
    std::string &err = c->err;
    if (!c->rep.inited) {
        err = "selm_autotune before selm_init";
        return SELM_ERR_STATE;
    }
    HIPCHK(hipStreamSynchronize(c->last));
    c->last = c->batch.stream;
    std::vector<TuneRegion> mem{{c->grid_mem, c->rep.grid_sz}};
    if (c->bead_mem && c->rep.bead_sz) mem.push_back({c->bead_mem, c->rep.bead_sz});
    return autotune_batch(c->batch, mem, steps, shape, ms, err);
}

// the spread launch's shape the next step runs: split threshold (0: no
// split list) and thermal-forcing workgroups (0 off the plane path)
static int spread_shape_of(Batch &B, int out[2], std::string &err) {
    int rc = B.ensure(err);
    if (rc) return rc;
    out[0] = B.L.tot_split > 0 ? B.L.split_at : 0;
    out[1] = B.L.plane && B.L.tot_modes > 0
                 ? (B.L.noise_blocks > 0 ? B.L.noise_blocks : spread_noise_default(B.L.tot_modes))
                 : 0;
    return SELM_OK;
}

int selm_spread_shape(selm_ctx *c, int out[2]) {
    CTX_CHECK(c);
    if (!out) return SELM_ERR_ARG;
    return spread_shape_of(c->batch, out, c->err);
}

int selm_fft_passes(selm_ctx *c) {
    CTX_CHECK(c);
    int rc = c->batch.ensure(c->err);
    if (rc) return -rc;
    if (!c->batch.L.fast) return 0;
    return c->batch.L.plane ? 3 : 5;
}

int selm_stage_times(selm_ctx *c, int nsteps, double *ms, int nstages) {
    CTX_CHECK(c);
    std::string &err = c->err;
    if (!ms || nstages < kNumStages || nsteps < 0) return SELM_ERR_ARG;
    if (!c->rep.inited) {
        err = "selm_stage_times before selm_init";
        return SELM_ERR_STATE;
    }
    int rc = c->batch.ensure(err);
    if (rc) return rc;
    hipStream_t st = c->batch.stream;
    HIPCHK(hipStreamSynchronize(c->last));
    c->last = st;
    std::vector<hipEvent_t> ev(kNumStages + 1);
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    for (int s = 0; s < kNumStages; s++) ms[s] = 0.0;
    for (int it = 0; it < nsteps; it++) {
        HIPCHK(hipEventRecord(ev[0], st));
        for (int s = 0; s < kNumStages; s++) {
            launch_stage(c->batch.L, s, st);
            HIPCHK(hipEventRecord(ev[s + 1], st));
        }
        HIPCHK(hipEventSynchronize(ev[kNumStages]));
        for (int s = 0; s < kNumStages; s++) {
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, ev[s], ev[s + 1]));
            ms[s] += t;
        }
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    HIPCHK(hipGetLastError());
    return SELM_OK;
}

int selm_mean_mode(selm_ctx *c, double out[3]) {
    CTX_CHECK(c);
    std::string &err = c->err;
    HIPCHK(hipStreamSynchronize(c->last));
    HIPCHK(hipMemcpy(out, c->rep.D.diag, 24, hipMemcpyDeviceToHost));
    return SELM_OK;
}

void *selm_stream(selm_ctx *c) { return c ? (void *)c->batch.stream : nullptr; }

}  // extern "C"

// ===========================================================================
// selm_pool: many systems in one device arena, stepped together
// ===========================================================================
namespace selm {
struct ArenaAlloc {
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> bytes
    void init(size_t s) {
        size = s;
        free_.clear();
        free_[0] = s;
    }
    bool alloc(size_t bytes, size_t &off) {
        bytes = al(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                off = it->first;
                size_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = al(bytes);
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const {
        size_t f = 0;
        for (auto &kv : free_) f += kv.second;
        return size - f;
    }
    // after compaction: everything below `end` is taken, the rest is free
    void reset_to(size_t end) {
        free_.clear();
        if (end < size) free_[end] = size - end;
    }
};
}  // namespace selm

struct selm_pool {
    Batch batch;
    char *arena = nullptr;
    ArenaAlloc alloc;
    std::vector<Replica *> slots;
    std::vector<size_t> slot_off, slot_bytes;
    std::string err;
    int precision = 0;
    long long compactions = 0;  // arena compactions so far
};

extern "C" {

int selm_pool_create(const selm_pool_params *pp, selm_pool **out) {
    if (!pp || !out || pp->max_replicas <= 0 || pp->hbm_budget < 4096) return SELM_ERR_ARG;
    if (pp->precision != SELM_PREC_FP64 && pp->precision != SELM_PREC_FP32_BULK) return SELM_ERR_ARG;
    *out = nullptr;
    selm_pool *P = new (std::nothrow) selm_pool();
    if (!P) return SELM_ERR_NOMEM;
    P->precision = pp->precision;
    P->batch.fp64 = pp->precision == SELM_PREC_FP64;
    P->batch.use_graph = pp->use_graph != 0;
    P->batch.force_generic_fft = std::getenv("SELM_FFT_GENERIC") != nullptr;
    P->batch.allow_plane = !env_is_zero("SELM_FFT_PLANES");
    P->batch.zq_cols = env_zq_cols();
    P->batch.split_at = env_count("SELM_SPLIT_AT", kSplitAt);
    P->batch.noise_blocks = env_count("SELM_FORCING_WG", 0);
    P->batch.wt_stores = env_count("SELM_WT_STORES", 1) ? 1 : 0;
    if (P->batch.init_streams(pp->max_replicas, P->err) != SELM_OK) {
        P->batch.release();
        delete P;
        return SELM_ERR_HIP;
    }
    size_t bytes = al(pp->hbm_budget);
    if (hipMalloc((void **)&P->arena, bytes) != hipSuccess) {
        P->batch.release();
        delete P;
        return SELM_ERR_NOMEM;
    }
    P->alloc.init(bytes);
    P->slots.assign(pp->max_replicas, nullptr);
    P->slot_off.assign(pp->max_replicas, 0);
    P->slot_bytes.assign(pp->max_replicas, 0);
    *out = P;
    return SELM_OK;
}

int selm_pool_destroy(selm_pool *P) {
    if (!P) return SELM_ERR_ARG;
    (void)hipStreamSynchronize(P->batch.stream);
    for (Replica *r : P->slots) delete r;
    P->batch.release();
    if (P->arena) (void)hipFree(P->arena);
    delete P;
    return SELM_OK;
}

const char *selm_pool_last_error(const selm_pool *P) { return P ? P->err.c_str() : "null pool"; }

static void pool_rebuild_list(selm_pool *P) {
    P->batch.reps.clear();
    for (Replica *r : P->slots)
        if (r) P->batch.reps.push_back(r);
    P->batch.dirty = true;
}

// Move every admitted system down to the start of the arena, in offset
// order, so that the free space becomes one block.  A system's device state
// is addressed only through the pointers that carve_grid/carve_beads derive
// from its base, so a move is a device-to-device copy of its block and a new
// carve; the batch's descriptors are uploaded again before the next step.  A
// block that moves by less than its size is copied in pieces no longer than
// the distance moved, from low to high addresses, so no piece overwrites
// bytes that a later piece still has to read.
static int pool_compact(selm_pool *P, std::string &err) {
    hipStream_t st = P->batch.stream;
    HIPCHK(hipStreamSynchronize(st));
    std::vector<int> order;
    for (size_t q = 0; q < P->slots.size(); q++)
        if (P->slots[q]) order.push_back((int)q);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return P->slot_off[a] < P->slot_off[b]; });
    size_t end = 0;
    for (int q : order) {
        const size_t off = P->slot_off[q], bytes = al(P->slot_bytes[q]);
        if (off != end) {
            const size_t dist = off - end;
            for (size_t c = 0; c < bytes; c += dist)
                HIPCHK(hipMemcpyAsync(P->arena + end + c, P->arena + off + c, std::min(dist, bytes - c),
                                      hipMemcpyDeviceToDevice, st));
            Replica &R = *P->slots[q];
            carve_grid(R, P->arena + end);
            carve_beads(R, P->arena + end + al(R.grid_sz));
            P->slot_off[q] = end;
        }
        end += bytes;
    }
    HIPCHK(hipStreamSynchronize(st));
    P->alloc.reset_to(end);
    pool_rebuild_list(P);
    P->compactions++;
    return SELM_OK;
}

int selm_pool_admit(selm_pool *P, const selm_params *p, int M, const double *x, const double *v,
                    const double *m, int nb, const int *ij, const double *K, const double *r0, int next,
                    const int *idx, const double *F, int *slot) {
    if (!P || !p || !slot) return SELM_ERR_ARG;
    std::string &err = P->err;
    int s = -1;
    for (size_t q = 0; q < P->slots.size(); q++)
        if (!P->slots[q]) {
            s = (int)q;
            break;
        }
    if (s < 0) {
        err = "no free slot";
        return SELM_ERR_POOL_FULL;
    }
    Replica *R = new (std::nothrow) Replica();
    if (!R) return SELM_ERR_NOMEM;
    selm_params q = *p;
    q.precision = P->precision;
    const bool fp64 = P->batch.fp64;
    int rc = replica_setup(*R, q, fp64, err);
    if (!rc) rc = host_set_beads(*R, M, x, v, m, err);
    if (!rc) rc = host_set_bonds(*R, nb, ij, K, r0, err);
    if (!rc) rc = host_set_external(*R, next, idx, F, err);
    if (rc) {
        delete R;
        return rc;
    }
    size_t gb = grid_bytes(R->G, fp64), bb = bead_bytes(R->M, (int)R->adj_i1.size(), fp64, R->G.ntiles);
    size_t off = 0;
    bool ok = P->alloc.alloc(gb + bb, off);
    if (!ok && P->alloc.size - P->alloc.used() >= al(gb + bb)) {
        // enough free space in total, but not in one block
        rc = pool_compact(P, err);
        if (rc) {
            delete R;
            return rc;
        }
        ok = P->alloc.alloc(gb + bb, off);
    }
    if (!ok) {
        err = "not enough free arena space for footprint " + std::to_string(gb + bb);
        delete R;
        return SELM_ERR_POOL_FULL;
    }
    carve_grid(*R, P->arena + off);
    carve_beads(*R, P->arena + off + al(gb));
    R->D.status = P->batch.d_status + s;
    R->grid_sz = gb;
    R->bead_sz = bb;
    rc = upload_tables(*R, P->batch.stream, err);
    if (!rc) rc = replica_upload_init(*R, P->batch.stream, err);
    if (rc) {
        P->alloc.release(off, gb + bb);
        delete R;
        return rc;
    }
    P->slots[s] = R;
    P->slot_off[s] = off;
    P->slot_bytes[s] = gb + bb;
    pool_rebuild_list(P);
    rc = P->batch.ensure(err);
    if (rc) return rc;
    // forces at x^0 for all systems (recomputing the others' is idempotent)
    launch_bond_forces(P->batch.L, 0, P->batch.stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(P->batch.stream));
    *slot = s;
    return SELM_OK;
}

int selm_pool_step(selm_pool *P, int nsteps) {
    if (!P) return SELM_ERR_ARG;
    return P->batch.step(nsteps, P->batch.stream, P->err);
}

int selm_pool_sync(selm_pool *P) {
    if (!P) return SELM_ERR_ARG;
    std::string &err = P->err;
    HIPCHK(hipStreamSynchronize(P->batch.stream));
    std::vector<int> f(P->slots.size(), 0);
    HIPCHK(hipMemcpy(f.data(), P->batch.d_status, sizeof(int) * f.size(), hipMemcpyDeviceToHost));
    std::string bad;
    for (size_t q = 0; q < f.size(); q++)
        if (P->slots[q] && f[q]) bad += " slot " + std::to_string(q) + ":" + status_text(f[q]) + ";";
    if (!bad.empty()) {
        err = "device status flags on" + bad;
        return SELM_ERR_DEVICE;
    }
    return SELM_OK;
}

int selm_pool_status(selm_pool *P, int slot, int *flags, int clear) {
    if (!P || !flags || slot < 0 || slot >= (int)P->slots.size() || !P->slots[slot]) return SELM_ERR_ARG;
    std::string &err = P->err;
    HIPCHK(hipStreamSynchronize(P->batch.stream));
    HIPCHK(hipMemcpy(flags, P->batch.d_status + slot, sizeof(int), hipMemcpyDeviceToHost));
    if (clear) HIPCHK(hipMemset(P->batch.d_status + slot, 0, sizeof(int)));
    return SELM_OK;
}

static Replica *pool_slot(selm_pool *P, int slot) {
    if (!P || slot < 0 || slot >= (int)P->slots.size()) return nullptr;
    return P->slots[slot];
}

int selm_pool_get_beads(selm_pool *P, int slot, double *x, double *v, double *fc, double *ft) {
    Replica *R = pool_slot(P, slot);
    if (!R) return SELM_ERR_ARG;
    return download_beads(*R, P->batch.stream, x, v, fc, ft, P->err);
}

int selm_pool_get_field(selm_pool *P, int slot, int comp, double *u) {
    Replica *R = pool_slot(P, slot);
    if (!R) return SELM_ERR_ARG;
    return download_field(*R, P->batch.stream, comp, u, P->err);
}

long long selm_pool_timestep(selm_pool *P, int slot) {
    Replica *R = pool_slot(P, slot);
    if (!R) return -1;
    long long ts = -1;
    if (hipStreamSynchronize(P->batch.stream) != hipSuccess) return -1;
    if (hipMemcpy(&ts, R->D.ts, 8, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return ts;
}

int selm_pool_set_field(selm_pool *P, int slot, int comp, const double *u) {
    Replica *R = pool_slot(P, slot);
    if (!R) return SELM_ERR_ARG;
    return upload_field(*R, P->batch.stream, comp, u, P->err);
}

int selm_pool_set_timestep(selm_pool *P, int slot, long long n) {
    Replica *R = pool_slot(P, slot);
    if (!R) return SELM_ERR_ARG;
    return replica_set_timestep(*R, P->batch.stream, n, P->err);
}

int selm_pool_retire(selm_pool *P, int slot) {
    Replica *R = pool_slot(P, slot);
    if (!R) return SELM_ERR_ARG;
    std::string &err = P->err;
    HIPCHK(hipStreamSynchronize(P->batch.stream));
    P->alloc.release(P->slot_off[slot], P->slot_bytes[slot]);
    delete R;
    P->slots[slot] = nullptr;
    pool_rebuild_list(P);
    return SELM_OK;
}

int selm_pool_autotune(selm_pool *P, int steps, int shape[2], double *ms) {
    if (!P || !shape || !ms) return SELM_ERR_ARG;
    size_t end = 0;
    for (size_t q = 0; q < P->slots.size(); q++)
        if (P->slots[q]) end = std::max(end, P->slot_off[q] + P->slot_bytes[q]);
    if (end == 0) {
        P->err = "selm_pool_autotune on an empty pool";
        return SELM_ERR_STATE;
    }
    std::vector<TuneRegion> mem{{P->arena, end}};
    return autotune_batch(P->batch, mem, steps, shape, ms, P->err);
}

int selm_pool_spread_shape(selm_pool *P, int out[2]) {
    if (!P || !out) return SELM_ERR_ARG;
    return spread_shape_of(P->batch, out, P->err);
}

int selm_pool_stats(selm_pool *P, size_t *used, size_t *budget, int *active) {
    if (!P) return SELM_ERR_ARG;
    if (used) *used = P->alloc.used();
    if (budget) *budget = P->alloc.size;
    if (active) *active = (int)P->batch.reps.size();
    return SELM_OK;
}

void *selm_pool_stream(selm_pool *P) { return P ? (void *)P->batch.stream : nullptr; }

}  // extern "C"
