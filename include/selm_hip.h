/*
 * selm_hip.h -- C ABI of the MI355X-native SELM fluctuating-hydrodynamics step.
 *
 * The reference couples a SELM step to LAMMPS through two C++ plugin classes
 * that share raw device pointers (Eulerian_Period: cuda/SELM_Eulerian_Period.cu,
 * Lagrangian_LAMMPS: cuda/SELM_Lagrangian_LAMMPS.cu, driven by Selm:
 * cuda/SELM.cu; prefix cuda/ = data/src/selm-main/src/cuda/).  This header is
 * the boundary that replaces them: plain pointers and sizes, int status codes,
 * no exceptions, no C++ or torch types.  One selm_ctx is one Selm system
 * (one periodic grid + one bead set); it owns all its device memory.  A
 * selm_ctx is thread-compatible (one host thread at a time); there is no
 * global state.
 *
 * Arrays of beads are AoS [M][3] doubles, exactly LAMMPS's atom->x/v/f
 * layout.  Grid fields are [z][y][x] with x fastest; component d of the
 * velocity lives on the MAC face (i, j+1/2, k+1/2) dx for d = 0, etc.
 */
#ifndef SELM_HIP_H
#define SELM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (0 = OK); text via selm_last_error / selm_pool_last_error */
#define SELM_OK 0
#define SELM_ERR_ARG 1          /* invalid argument (the reference would throw or crash) */
#define SELM_ERR_HIP 2          /* HIP runtime error (reference: CHECK_RUNTIME -> exit(-1), cuda/check.cuh:6-10) */
#define SELM_ERR_STATE 3        /* call out of order (e.g. step before init) */
#define SELM_ERR_NOMEM 4        /* device allocation failed */
#define SELM_ERR_UNSUPPORTED 5  /* grid edge not a product of 2,3,5 in [4, 512] */
#define SELM_ERR_POOL_FULL 6    /* pool: no slot or not enough budget for this footprint */
#define SELM_ERR_DEVICE 7       /* a kernel raised a status flag (SELM_STATUS_*); see selm_status */
#define SELM_ERR_SCRATCH 8      /* request larger than the fixed scratch of selm_msd_noalloc */

/* Device status flags of a system, raised by the step's kernels and kept
 * until cleared with selm_status / selm_pool_status. */
#define SELM_STATUS_TILE_OVERFLOW 1  /* more than 16384 beads met one 8^3 spread tile: the
                                        fixed-point cell sums may have wrapped */
#define SELM_STATUS_NONFINITE 2      /* a bead position or velocity became NaN or infinite */
#define SELM_STATUS_DISPLACEMENT 4   /* |v dt| > L/2 for some bead component; the reference
                                        prints this condition and carries on
                                        (cuda/SELM_Kernel.cu:363-365,
                                        mpi/SELM_Lagrangian_LAMMPS.cpp:149-151) */

/* precision of the bulk (grid, FFT, spread, interpolation) arithmetic; bead
 * positions, velocities and forces are FP64 in both */
#define SELM_PREC_FP64 0
#define SELM_PREC_FP32_BULK 1

/* per-stage indices for selm_stage_times */
#define SELM_NUM_STAGES 9

typedef struct selm_ctx selm_ctx;
typedef struct selm_pool selm_pool;

/* Inputs of one system.  Maps the reference's input surface:
 *   dims ........ Model.SELM numMeshPtsPerDir (cuda/SELM.cu:44)
 *   lo, hi ...... LAMMPS domain->boxlo/boxhi (cuda/SELM_Lagrangian_LAMMPS.cu:62-63);
 *                 like the reference, dx = (hi[0]-lo[0])/dims[0] and the bead
 *                 wrap uses lo[0]/hi[0] for every dimension
 *   rho mu kB T . Model.SELM rho, mu, KB, T (cuda/SELM.cu:20-24)
 *   dt .......... LAMMPS `timestep` (cuda/SELM.cu:37)
 *   seed ........ Model.SELM SELM_Seed (cuda/SELM.cu:29)
 *   stochastic .. Model.SELM flagStochasticDriving (cuda/SELM.cu:28)
 *   precision ... SELM_PREC_* (new; default FP64 keeps the reference's arithmetic)
 *   f4_quirk .... 1 (default): the reference's Peskin factor, which stays 1.0 at
 *                 a stencil distance r >= 2 (cuda/SELM_Kernel.cu:387-395);
 *                 0: the textbook kernel (factor 0 there)                     */
typedef struct selm_params {
    int dims[3];
    double lo[3], hi[3];
    double rho, mu, kB, T, dt;
    uint64_t seed;
    int stochastic;
    int precision;
    int f4_quirk;
    int reserved;
} selm_params;

/* Fill p with the reference's defaults (SELM_Seed 123456, flagStochasticDriving 0:
 * cuda/SELM.cuh:13-15) and FP64, f4_quirk = 1. */
int selm_params_default(selm_params *p);

/* Eulerian_Period::initial's allocations (cuda/SELM_Eulerian_Period.cu:14-76)
 * plus the bead-side state; replaces `new Selm` + `new Eulerian_Period`
 * (cuda/SELM.cu:17-48). */
int selm_create(const selm_params *p, selm_ctx **out);

/* Selm::final + ~Selm (cuda/SELM.cu:50-63): frees everything the ctx owns. */
int selm_destroy(selm_ctx *ctx);

/* Text of the last error on this ctx ("" if none). */
const char *selm_last_error(const selm_ctx *ctx);

/* Bead state as LAMMPS holds it after read_data (Lagrangian_LAMMPS::initial,
 * cuda/SELM_Lagrangian_LAMMPS.cu:53-69,84-92): x, v AoS [M][3] (v may be NULL
 * for zero velocities, the read_data default), m [M] per-bead mass. Copied in.
 * SELM_ERR_ARG for a non-finite x or v or a mass that is not finite and > 0.
 * selm_init takes positions into the box as LAMMPS's remap does (periodic
 * image in [lo, hi) per dimension). */
int selm_set_beads(selm_ctx *ctx, int M, const double *x, const double *v, const double *m);

/* The check selm_set_beads and selm_pool_admit apply to bead inputs, on the
 * host and without a device (a serving front end can refuse a request before
 * queueing it): SELM_ERR_ARG and *bad = the first offending bead when a
 * position or velocity is not finite or a mass is not finite and > 0
 * (v may be NULL); SELM_OK and *bad = -1 otherwise.  No reference
 * counterpart: LAMMPS's read_data would accept such values. */
int selm_check_beads(int M, const double *x, const double *v, const double *m, int *bad);

/* LAMMPS `bond_style harmonic` bonds (Bonds section + bond_coeff), evaluated
 * where the reference calls integrate->setup_minimal
 * (cuda/SELM_Lagrangian_LAMMPS.cu:122): E = K (r - r0)^2 with the minimum image
 * convention.  ij [nb][2] zero-based bead indices, K [nb], r0 [nb]. */
int selm_set_bonds(selm_ctx *ctx, int nb, const int *ij, const double *K, const double *r0);

/* LAMMPS `angle_style harmonic` angles (Angles section + angle_coeff), on the
 * device at every step: E = K (theta - theta0)^2, theta the angle at the
 * centre bead, theta0 in radians (angle_coeff gives degrees).  ijk [na][3]
 * zero-based (end, centre, end); replaces the reference's LAMMPS evaluation
 * in setup_minimal (cuda/SELM_Lagrangian_LAMMPS.cu:122). */
int selm_set_angles(selm_ctx *ctx, int na, const int *ijk, const double *K, const double *theta0);

/* LAMMPS `pair_style lj/cut` (pair_coeff, special_bonds lj): E = 4 eps
 * ((sigma/r)^12 - (sigma/r)^6) for r < rc under the minimum image, evaluated
 * on the device at every step over all bead pairs.  type [M] zero-based,
 * eps/sigma/rc [ntypes][ntypes] (already mixed), special_lj the weights of
 * 1-2, 1-3 and 1-4 partners through the bonds of selm_set_bonds (weight 0
 * drops the pair).  Replaces the reference's LAMMPS evaluation
 * (cuda/SELM_Lagrangian_LAMMPS.cu:122); ntypes = 0 removes the term. */
int selm_set_pair_lj(selm_ctx *ctx, int ntypes, const int *type, const double *eps, const double *sigma,
                     const double *rc, const double special_lj[3]);

/* LAMMPS `fix addforce` with a constant vector on the beads idx[n]: F [n][3]. */
int selm_set_external_force(selm_ctx *ctx, int n, const int *idx, const double *F);

/* Host-force hook: overwrite the conservative force for the next step, the
 * reference's H2D copy of LAMMPS forces (cuda/SELM_Lagrangian_LAMMPS.cu:129).
 * Forces from selm_set_bonds/external are recomputed after every step. */
int selm_set_conservative_force(selm_ctx *ctx, const double *f);

/* Selm::initial (cuda/SELM.cu:55-58): LAMMPS setup(1) wraps x into [lo, hi)
 * and computes the forces at x^0; the fluid starts at u = 0; step counter 0. */
int selm_init(selm_ctx *ctx);

/* nsteps iterations of Selm::run's loop body (cuda/SELM.cu:73-90): Lagrangian
 * update with u^n, fluid update, new thermal field and forces at x^{n+1}.
 * Enqueued on `stream` (a hipStream_t; NULL = the ctx's own stream) as one
 * captured hipGraph per step; returns without waiting. */
int selm_step(selm_ctx *ctx, int nsteps, void *stream);

/* Wait for all work enqueued by this ctx.  Returns SELM_ERR_DEVICE (text in
 * selm_last_error) when a status flag is raised; the flags stay raised
 * until selm_status clears them.  selm_step itself only enqueues, so this is
 * where the host first learns of them. */
int selm_sync(selm_ctx *ctx);

/* Wait for the ctx's work and return its status flags (SELM_STATUS_*) in
 * *flags; clear != 0 resets them.  Replaces the reference's printf of a
 * displacement beyond half the box (cuda/SELM_Kernel.cu:363-365), which the
 * host never saw. */
int selm_status(selm_ctx *ctx, int *flags, int clear);

/* Bead state after the last step (waits for it): x, v as in LAMMPS;
 * f_cons = conservative force at the current x (what atom->f holds after
 * setup_minimal); f_total = (drag + thermal + dt f_cons)/dt of the last update
 * (what d_f holds after lagrangian_update and particleF.dat records,
 * cuda/SELM_Kernel.cu:357-358).  Any pointer may be NULL. */
int selm_get_beads(selm_ctx *ctx, double *x, double *v, double *f_cons, double *f_total);

/* Velocity component comp (0..2) of u^n as N doubles [z][y][x]
 * (the reference's d_u[cur_u], cuda/SELM_Eulerian.cuh:15-17). */
int selm_get_field(selm_ctx *ctx, int comp, double *u);

/* Overwrite component comp of u^n (N doubles); for restarts and tests. */
int selm_set_field(selm_ctx *ctx, int comp, const double *u);

/* Number of completed steps (Selm::timestep, cuda/SELM.cuh:11). */
long long selm_timestep(selm_ctx *ctx);

/* Set the number of completed steps to n (0 <= n <= 2^31 - 2), after
 * selm_init: the next step is n + 1 and its thermal draws are keyed by n + 1.
 * A system rebuilt from a snapshot of x and v (selm_set_beads, selm_init),
 * u (selm_set_field) and n therefore continues bit for bit as the run the
 * snapshot was taken from.  Replaces Selm::timestep (cuda/SELM.cuh:11) as a
 * restart point; the reference itself has no restart (`variable restart` is
 * unused, selm/test/NucleotideMonomer/Model.LAMMPS_script:3). */
int selm_set_timestep(selm_ctx *ctx, long long n);

/* Device bytes one system with these parameters, M beads and nb bonds holds. */
int selm_footprint(const selm_params *p, int M, int nb, size_t *bytes);

/* Run nsteps steps stage by stage with HIP events around each stage on the
 * ctx stream; ms[s] receives the summed milliseconds of stage s
 * (0 thermal forcing, 1 bead update + binning, 2 spread, 3 viscous u* + FFT x,
 *  4 FFT y, 5 FFT z + projection, 6 inverse FFT y, 7 inverse FFT x,
 *  8 bond forces).  In selm_step the thermal forcing and the bond forces run
 *  on a forked stream beside stages 2-4; here every stage runs alone. */
int selm_stage_times(selm_ctx *ctx, int nsteps, double *ms, int nstages);

/* Autotuner of the FFT pass shapes and of the spread launch's shape (no
 * reference counterpart: the reference's cuFFT plans choose their own,
 * cuda/SELM_Eulerian_Period.cu:26-30, and its spread is one thread per bead,
 * cuda/SELM_Kernel.cu:424-470).  Times `steps` steps of the system under each
 * FFT shape this build has for its grid -- FP32-bulk with nx == ny <= 128:
 * the three-pass plane path with 17 or 9 (kx, ky) columns per Z-pass
 * workgroup, and the five-pass transposing path; one shape otherwise -- and
 * keeps the fastest; then, with that shape, each spread shape: split
 * threshold 32, 64, 128 or none (tiles with more bucket entries are taken
 * first by workgroups of their own) and, on the plane path, half, the
 * default or twice the thermal-forcing workgroups of the spread launch; then,
 * on the plane path, the store shape (bulk outputs stored at device scope,
 * written through the L2, or plainly).  The system's device state is restored after each candidate.  The choice is
 * cached per (precision, grid shapes, bead counts) for the process, and a
 * cached choice is applied without timing.  shape[0] = FFT passes (3, 5; 0
 * for the runtime-length kernels), shape[1] = Z-pass columns per workgroup on
 * the plane path (else 0), *ms = the best milliseconds per step (0 when taken
 * from the cache); selm_spread_shape reports the spread's shape. */
int selm_autotune(selm_ctx *ctx, int steps, int shape[2], double *ms);

/* The spread launch's shape the next step runs: out[0] = split threshold (0:
 * no split list, e.g. for systems of fewer than 2048 beads), out[1] =
 * thermal-forcing workgroups in the spread launch (0 off the plane path). */
int selm_spread_shape(selm_ctx *ctx, int out[2]);

/* Passes of the 3-D FFT in the ctx's step: 3 (plane path: XY, Z, XY; stages
 * 4 and 6 are then empty), 5 (transposing X, Y, Z, Y, X) or 0 (runtime-length
 * kernels); negative on error.  No reference counterpart: the reference's
 * cuFFT plans (cuda/SELM_Eulerian_Period.cu:26-30) are opaque. */
int selm_fft_passes(selm_ctx *ctx);

/* Sum over the grid of u* for each component before the k=0 mode is
 * removed in the last step (FP64): rho * dx^3 * this equals the fluid
 * momentum injected by the beads' drag and thermal kicks. */
int selm_mean_mode(selm_ctx *ctx, double out[3]);

/* The ctx's own stream (hipStream_t). */
void *selm_stream(selm_ctx *ctx);

/* Mean-square displacement of a stored trajectory, with the semantics of the
 * reference's analysis script (py/Diffusion Coefficient/MSD.py:15-69).
 * - frames: [nframes][3][natoms] float, the x, y, z records of each DCD frame.
 * - box[0]: the x edge of the first frame's cell.  Half of it is the minimum
 *   image threshold X in every dimension, as in the script.
 * - Origins start at frame floor(nframes / 5).
 * - msd[0] = 0 and msd[dot] = MSD at lag dot * mult, for dot = 1 .. nlags-1
 *   (NaN where no origin is left).
 * Computed on the GPU; terms in FP32 as the script does, sums in FP64. */
int selm_msd(const float *frames, int nframes, int natoms, const double box[3], int nlags, int mult,
             double *msd);

/* selm_msd without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).
 * - d_frames, d_msd: DEVICE pointers, the same layouts as selm_msd's frames
 *   and msd; box is a host array.  Enqueued on stream (NULL: the null
 *   stream) and not waited for; d_msd holds the result when the stream has
 *   reached this point.  Results are bitwise those of selm_msd.
 * - The partial sums live in a fixed device array of 65536 doubles (512 KiB,
 *   held by each GPU the library runs on until the process ends); a request
 *   needs at most 2048 + nlags - 1 of them, so nlags <= 65537 always fits
 *   and a larger nlags returns SELM_ERR_SCRATCH.
 * - The first call on a device resolves that device's array address (make it
 *   outside graph capture).  Calls on one GPU share its array: calls that
 *   overlap in time on one GPU (different streams, or threads) race.
 * Calls no hipMalloc, hipFree or allocating library routine. */
int selm_msd_noalloc(const float *d_frames, int nframes, int natoms, const double box[3], int nlags, int mult,
                     double *d_msd, void *stream);

/* ---------------------------------------------------------------------------
 * Ensembles and serving (no counterpart in the reference, which runs one
 * system per process: cuda/main.cu:5-12).  A pool owns one device arena of
 * hbm_budget bytes; systems are admitted by their footprint (selm_footprint)
 * into free space, stepped together by one batched launch per stage
 * (captured as one hipGraph per pool composition), and retired to free their
 * space.  Each system's result is independent of what else is in the pool.
 * ------------------------------------------------------------------------- */
typedef struct selm_pool_params {
    size_t hbm_budget;   /* bytes of the device arena */
    int max_replicas;    /* slots */
    int precision;       /* SELM_PREC_*; all systems in a pool share it */
    int use_graph;       /* 1: replay a captured hipGraph per step */
} selm_pool_params;

int selm_pool_create(const selm_pool_params *pp, selm_pool **out);
int selm_pool_destroy(selm_pool *pool);
const char *selm_pool_last_error(const selm_pool *pool);

/* Admit one system and run its Selm::initial.  p->precision is ignored (the
 * pool's precision applies).  Returns SELM_ERR_POOL_FULL if no slot or no
 * contiguous free space of its footprint is available. */
int selm_pool_admit(selm_pool *pool, const selm_params *p, int M, const double *x,
                    const double *v, const double *m, int nb, const int *ij, const double *K,
                    const double *r0, int next, const int *idx, const double *F, int *slot);

/* Advance every admitted system by nsteps steps; returns after enqueueing. */
int selm_pool_step(selm_pool *pool, int nsteps);
/* Wait for the pool's work; SELM_ERR_DEVICE when an admitted system has a
 * status flag raised (the text names the slots). */
int selm_pool_sync(selm_pool *pool);
/* Status flags of one admitted system (as selm_status): a serving loop
 * evicts (retires) a flagged system without disturbing the others. */
int selm_pool_status(selm_pool *pool, int slot, int *flags, int clear);
int selm_pool_get_beads(selm_pool *pool, int slot, double *x, double *v, double *f_cons,
                        double *f_total);
int selm_pool_get_field(selm_pool *pool, int slot, int comp, double *u);
long long selm_pool_timestep(selm_pool *pool, int slot);
/* selm_set_field and selm_set_timestep for one admitted system (resume a
 * snapshot inside a pool). */
int selm_pool_set_field(selm_pool *pool, int slot, int comp, const double *u);
int selm_pool_set_timestep(selm_pool *pool, int slot, long long n);
int selm_pool_retire(selm_pool *pool, int slot);
/* used/budget bytes of the arena, number of admitted systems */
/* selm_autotune for a pool: the shapes are timed on the pool's current
 * systems stepped together, and the pool's device state is restored. */
int selm_pool_autotune(selm_pool *pool, int steps, int shape[2], double *ms);
int selm_pool_spread_shape(selm_pool *pool, int out[2]);
int selm_pool_stats(selm_pool *pool, size_t *used, size_t *budget, int *active);
void *selm_pool_stream(selm_pool *pool);

#ifdef __cplusplus
}
#endif
#endif /* SELM_HIP_H */
