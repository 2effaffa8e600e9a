/*
 * pcr_penta.h — C ABI of the MI355X (gfx950) batched pentadiagonal solver.
 *
 * Method: PCR-Penta parallel cyclic reduction (Ghosh & Mishra, SIAM J. Sci.
 * Comput., DOI 10.1137/20M1311053): floor(log2 n) reduction steps at stride
 * 1, 2, 4, ..., each eliminating the +-1/+-2 couplings, then a 2x2 / 1x1 solve.
 *
 * Row i of a system reads
 *     a_i x_{i-2} + b_i x_{i-1} + c_i x_i + d_i x_{i+1} + e_i x_{i+2} = y_i
 * with row-aligned diagonals (a_0 = a_1 = b_0 = d_{n-1} = e_{n-2} = e_{n-1}
 * are out of range and never affect the solution). This is the argument order
 * (a,b,c,d,e,y) of the reference kernels, performanalysis.h:53-59.
 *
 * All array pointers are DEVICE pointers (hipMalloc'd or equivalent) unless
 * noted. Entry points never allocate device memory and return a status code
 * (PCRP_OK = 0) instead of exiting the process (the reference's checkCuda
 * macro, performanalysis.h:12-20, exits). Solves are asynchronous on
 * `stream` (a hipStream_t; NULL means the null stream) with these
 * exceptions: pcrp_solve_f32r_f64 and pcrp_solve_uneven_f32r_f64 on the
 * window path (long systems, uneven batches, or large batches of n in
 * (1024, 2048]) synchronise their stream once, after the first sweep, to
 * read a locality flag (pcrp_solve_f32r_f64 under stream capture takes the
 * split path instead, with no read-back); the uneven entry points copy
 * their plan tables from host memory on every call, so they cannot be
 * captured. The plan
 * API (pcrp_plan_create / pcrp_plan_execute) does that work once at plan
 * time and executes without any synchronisation or host copy (capturable).
 * The autotuners, pcrp_plan_create and pcrp_heston_solve with use_graph = 1
 * synchronise their stream.
 * The diagonals a..e are never modified; y is overwritten by the solution
 * (the reference overwrites a..e as well when n > 1024, kernel.cu:304-306).
 */
#ifndef PCR_PENTA_H
#define PCR_PENTA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCRP_ABI_VERSION 2   /* 2: pcrp_heston_params gained separate_explicit */

/* Status codes. */
enum {
    PCRP_OK = 0,
    PCRP_ERR_INVALID_SIZE = 1,   /* n < 1, m < 1, or sizes beyond 2^31 elements */
    PCRP_ERR_NULL = 2,           /* a required pointer is NULL */
    PCRP_ERR_LAYOUT = 3,         /* unknown layout code */
    PCRP_ERR_WORKSPACE = 4,      /* ws_bytes < pcrp_workspace_size(...) */
    PCRP_ERR_HIP = 5,            /* a HIP runtime call or kernel launch failed */
    PCRP_ERR_ARG = 6,            /* other invalid argument (max_sweeps, tol, kind ...) */
    PCRP_ERR_UNSUPPORTED = 7,    /* valid request this build cannot serve */
    PCRP_ERR_SCRATCH = 8         /* needs more than PCRP_NOALLOC_SCRATCH_BYTES */
};

/* Batch layouts. Element i of system j lives at
 *   CONTIGUOUS:  j*n + i   (the reference's own layout, kernel.cu:43)
 *   INTERLEAVED: i*m + j   (Annexure A p.2; cuSPARSE gpsvInterleavedBatch) */
enum { PCRP_LAYOUT_CONTIGUOUS = 0, PCRP_LAYOUT_INTERLEAVED = 1 };

/* Precision forms. */
enum {
    PCRP_PREC_F64 = 0,           /* plain FP64 PCR-Penta */
    PCRP_PREC_F32_REFINE = 1     /* FP32 PCR + FP64 residual/update refinement */
};

int pcrp_abi_version(void);
const char *pcrp_strerror(int status);

/* Device workspace (bytes) a solve of this shape needs; 0 when none. */
size_t pcrp_workspace_size(int n, int m, int precision);

/*
 * Plain FP64 batched solve.
 * Replaces the solve half of the reference's host dispatch
 * time_shared_cached_shared_tile_splitting (main.cu:14-89): pcr_penta for
 * n <= 1024 (main.cu:46-54, kernel.cu:32-154) and the split pair
 * shared_cached_shared_tile_splitting + pcr_penta_splitting_2nd_stage for
 * n > 1024 (main.cu:57-86, kernel.cu:158-320, 665-791). Any n >= 1 is
 * accepted (the reference needs n <= 1024 or n = 1024*2^s, main.cu:9-10).
 * Results equal a host FP64 run of the same algorithm bit for bit.
 */
int pcrp_solve_f64(const double *a, const double *b, const double *c,
                   const double *d, const double *e, double *y,
                   int n, int m, int layout,
                   void *ws, size_t ws_bytes, void *stream);

/*
 * Plain FP64 solve with no workspace argument, allocating nothing: no
 * counterpart in the reference; for callers that cannot allocate (code under
 * hipGraph capture, serving loops, harnesses that forbid hipMalloc). Same
 * arguments, results and paths as pcrp_solve_f64; the split path's
 * workspace (n > 1024) is a fixed device array of PCRP_NOALLOC_SCRATCH_BYTES
 * held on every GPU the library runs on until the process ends. Requests
 * whose pcrp_workspace_size exceeds it return PCRP_ERR_SCRATCH (about 96
 * bytes per unknown: up to ~700,000 unknowns when n > 1024). All calls on one
 * GPU share that GPU's array: calls that overlap in time on one GPU race and
 * must be serialised by the caller. The first call on a device looks the
 * array's address up (hipGetSymbolAddress, not a stream operation).
 */
#define PCRP_NOALLOC_SCRATCH_BYTES (64u << 20)
int pcrp_solve_f64_noalloc(const double *a, const double *b, const double *c,
                           const double *d, const double *e, double *y,
                           int n, int m, int layout, void *stream);

/*
 * FP32 inner solve + FP64 iterative refinement (no counterpart in the
 * reference, which is FP64-only): x0 = solve32(y); then up to max_sweeps
 * times: r = y - A x (FP64), dx = solve32(r), x += dx (FP64). A system stops
 * when ||dx_k|| <= tol ||x||, or when the error left after the update,
 * estimated from the observed contraction as 10 ||dx_k||^2 / ||dx_{k-1}||
 * (k >= 2), is <= tol ||x|| (and the contraction is <= 0.1). Norms are
 * max-norms. Inputs and output are FP64. sweeps_out: optional DEVICE array of
 * m ints receiving each system's number of refinement sweeps (max_sweeps + 1
 * flags a system that did not stop, or whose correction or iterate is NaN
 * or infinite).
 * solve32 is PCR on the 2x2-block tridiagonal form of the matrix (rows 2k,
 * 2k+1 as one block row, diagonal-block pivots), which keeps FP32 error near
 * 3e-7 where PCR-Penta's own off-diagonal pivots lose up to 1e-2. Its
 * levels stop early once every multiplier of a level is below 2^-16 (the
 * couplings left are then ~2^-32 relative: diagonally dominant systems need
 * 4-5 levels).
 * Sizes n <= 2048 run in one LDS-resident kernel per system (the FP32
 * factorisation is computed once and its multipliers kept in registers;
 * interleaved batches with more systems than CUs may take two neighbouring
 * systems per workgroup, with the same results), except batches of n > 1024
 * with at least 2^23 unknowns given a workspace of pcrp_workspace_size
 * bytes, which take the window path below.
 * Larger n need a workspace (pcrp_workspace_size with PCRP_PREC_F32_REFINE).
 * Window path (diagonally dominant systems): each sweep is one pass over
 * windows of 450 block rows with 31-block-row halos, computing the FP64
 * residual, the truncated FP32 block solve and x += dx; interleaved batches
 * are first copied to a system-contiguous layout in the workspace. If a
 * window's couplings outlast its halo, the call reads that flag back to the
 * host after the first sweep (one stream synchronisation) and solves the
 * batch with the split path instead: every sweep re-runs the FP32 block
 * solve split like pcrp_solve_f64 (HBM passes, then an LDS stage per residue
 * sub-system) on the FP32 residual. Under stream capture the split path runs
 * directly (no read-back), so a captured call stays asynchronous.
 * Per-system stop flags let converged systems drop out of later sweeps.
 * A block pivot with determinant below 2^-100 gets a zero inverse (the
 * refinement corrects the difference, or flags the system).
 */
int pcrp_solve_f32r_f64(const double *a, const double *b, const double *c,
                        const double *d, const double *e, double *y,
                        int n, int m, int layout, int max_sweeps, double tol,
                        int *sweeps_out, void *ws, size_t ws_bytes, void *stream);

/*
 * Uneven batch, plain FP64: m systems of different sizes in one call. System
 * j occupies rows h_offsets[j] .. h_offsets[j] + h_n[j] - 1 of every array
 * (contiguous, non-overlapping). h_offsets / h_n are HOST arrays, used to
 * plan the routing: systems up to the LDS threshold are solved whole, larger
 * ones run their first reduction steps as HBM pass rounds shared by all large
 * systems, and every system or residue sub-system then goes through one
 * persistent work-queue LDS launch. Results equal solving each system alone
 * with pcrp_solve_f64 (bit for bit). No counterpart in the reference, whose
 * driver runs one (m, n) shape per launch (main.cu:245-260).
 * The plan tables are copied to the workspace on `stream` each call, from a
 * host buffer that lives to the end of the call (hipMemcpyAsync stages
 * pageable memory before returning), so the call cannot be captured in a
 * graph: pcrp_plan_create uploads them once instead.
 */
size_t pcrp_uneven_workspace_size(const int64_t *h_offsets, const int *h_n, int m,
                                  int precision);
int pcrp_solve_uneven_f64(const double *a, const double *b, const double *c,
                          const double *d, const double *e, double *y,
                          const int64_t *h_offsets, const int *h_n, int m,
                          void *ws, size_t ws_bytes, void *stream);

/*
 * Uneven batch, FP32 inner solve + FP64 refinement (no counterpart in the
 * reference). Same layout, planning and workspace rules as
 * pcrp_solve_uneven_f64 (pcrp_uneven_workspace_size with
 * PCRP_PREC_F32_REFINE); refinement arguments and stopping rule as
 * pcrp_solve_f32r_f64, sweeps_out a DEVICE array of m ints. The inner solve
 * is PCR on the 2x2-block form. Diagonally dominant batches take the window
 * path of pcrp_solve_f32r_f64 (systems of up to 1024 rows as one window,
 * larger ones cut into windows; one stream synchronisation after the first
 * sweep); otherwise, and under stream capture: systems up to 2048 rows in
 * one workgroup, larger ones after HBM pass rounds shared by all large
 * systems.
 */
int pcrp_solve_uneven_f32r_f64(const double *a, const double *b, const double *c,
                               const double *d, const double *e, double *y,
                               const int64_t *h_offsets, const int *h_n, int m,
                               int max_sweeps, double tol, int *sweeps_out,
                               void *ws, size_t ws_bytes, void *stream);

/* Seeded generator for an uneven batch laid out as above (DEVICE arrays
 * d_offsets / d_n): the same values as pcrp_generate gives system j of a
 * batch, row i at d_offsets[j] + i. */
int pcrp_generate_uneven(int kind, uint64_t seed, const int64_t *d_offsets, const int *d_n,
                         int m, double *a, double *b, double *c, double *d, double *e,
                         double *y, void *stream);

/*
 * Plan / execute (no counterpart in the reference, whose only "solve call" is
 * a launch pair synchronised after every launch, main.cu:46-87). Of the
 * one-call entry points above, the uneven ones upload their routing tables
 * from host memory on every call, and FP32+refine of long systems reads
 * back, after its first sweep, whether the window path applies (one stream
 * synchronisation); neither can be captured in a hipGraph. A plan does that
 * work once:
 *   pcrp_plan_create: binds a..e (their values at this call: the plan keeps
 *     matrix-dependent data), the batch shape and the workspace; uploads the
 *     tables, runs the window path's locality test on the matrix and, for
 *     interleaved batches, keeps system-contiguous copies of a..e in the
 *     workspace. Synchronises `stream` (PCRP_ERR_ARG on a capturing stream).
 *     Uniform batch: h_offsets = h_n = NULL, n, m, layout as for
 *     pcrp_solve_f64. Uneven batch: h_offsets, h_n (HOST arrays) and m as for
 *     pcrp_solve_uneven_f64 (n and layout ignored).
 *   pcrp_plan_execute: solves y in place with the plan's route, launching
 *     kernels only (no allocation, no host copy, no synchronisation: it can
 *     be captured and replayed). Results equal the one-call entry points'
 *     (FP64 bit for bit; FP32+refine the same sweeps and bits). max_sweeps,
 *     tol, sweeps_out as for pcrp_solve_f32r_f64 (ignored for FP64).
 * The workspace (pcrp_plan_workspace_size bytes) belongs to the plan until
 * pcrp_plan_destroy: executes of one plan must not overlap in time.
 */
typedef struct pcrp_plan pcrp_plan;
enum {
    PCRP_ROUTE_LDS = 0,          /* one workgroup per system */
    PCRP_ROUTE_SPLIT = 1,        /* HBM passes + LDS stage per residue sub-system */
    PCRP_ROUTE_WINDOW = 2,       /* FP32+refine: one windowed pass per sweep */
    PCRP_ROUTE_UNEVEN_F64 = 3,   /* uneven FP64: pass rounds + persistent LDS queue */
    PCRP_ROUTE_UNEVEN_STAGE = 4  /* uneven FP32+refine: pass rounds + LDS stage per sweep */
};
size_t pcrp_plan_workspace_size(int precision, int n, int m, const int64_t *h_offsets,
                                const int *h_n);
int pcrp_plan_create(pcrp_plan **plan, int precision, const double *a, const double *b,
                     const double *c, const double *d, const double *e, int n, int m,
                     int layout, const int64_t *h_offsets, const int *h_n, void *ws,
                     size_t ws_bytes, void *stream);
int pcrp_plan_execute(const pcrp_plan *plan, double *y, int max_sweeps, double tol,
                      int *sweeps_out, void *stream);
int pcrp_plan_route(const pcrp_plan *plan);
int pcrp_plan_destroy(pcrp_plan *plan);

/*
 * Factor once, solve many (no counterpart in the reference, which recomputes
 * every multiplier on every solve: kernel.cu:72-75, 98-103, 122, 129).
 * The matrix part of the solve depends on a..e only; these entry points
 * compute it once into a caller-owned DEVICE buffer and then solve right-hand
 * sides against it.
 *   shared = 0: m systems in `layout`, one factor block per system;
 *   shared = 1: a..e hold ONE system of n rows (row i at element i; m and
 *               layout are ignored by the factor call) whose factor serves
 *               every one of the m right-hand sides of a solve.
 * pcrp_factor_size: bytes of the factor buffer (0 if unsupported).
 * FP64: pcrp_factor_f64 keeps, per reduction step and row, the pair-equation
 * multipliers and Reduction-III's gamma1/gamma2 (after the zero-pivot rule),
 * plus the reduced c, b, d of the final 2x2 / 1x1 solve: (4 floor(log2 n) + 3)
 * n doubles per system. pcrp_solve_factored_f64 then costs 4 FMA per row and
 * step; its result equals pcrp_solve_f64 on the same matrices bit for bit
 * (signed zeros aside). n <= 1024.
 * FP32+refine: pcrp_factor_f32r keeps the 2x2-block FP32 factor of
 * pcrp_solve_f32r_f64's one-workgroup kernel (the level multipliers, the last
 * D^-1, the level count); pcrp_solve_factored_f32r_f64 runs the same
 * refinement (same arguments and stopping rule) without refactoring, and
 * still needs a..e for its FP64 residuals (the same arrays, and the same
 * `shared` choice, as the factor call). n <= 2048.
 * The solve arguments (n, m, layout, shared) must be those of the factor call;
 * the factor buffer is read-only during solves, so solves of one factor may
 * run concurrently on several streams.
 */
size_t pcrp_factor_size(int n, int m, int precision, int shared);
int pcrp_factor_f64(const double *a, const double *b, const double *c,
                    const double *d, const double *e, int n, int m, int layout,
                    int shared, void *factor, size_t factor_bytes, void *stream);
int pcrp_solve_factored_f64(const void *factor, double *y, int n, int m, int layout,
                            int shared, void *stream);
int pcrp_factor_f32r(const double *a, const double *b, const double *c,
                     const double *d, const double *e, int n, int m, int layout,
                     int shared, void *factor, size_t factor_bytes, void *stream);
int pcrp_solve_factored_f32r_f64(const void *factor, const double *a, const double *b,
                                 const double *c, const double *d, const double *e,
                                 double *y, int n, int m, int layout, int shared,
                                 int max_sweeps, double tol, int *sweeps_out, void *stream);

/*
 * Heston ADI driver (the application of the PCR-Penta paper; the reference
 * ships no ADI code). European call under Heston, Douglas scheme, uniform
 * grid: Ns unknowns in S (S = i Smax/(Ns+1), i = 1..Ns; u = 0 at S = 0,
 * u = Smax - K e^{-r tau} at S = Smax) and Nv unknowns in v (v = j vmax/Nv,
 * j = 0..Nv-1; u = S at v = vmax). Fourth-order 5-point central differences
 * in S (second order next to the S boundaries), second-order upwinding plus
 * central diffusion in v, explicit mixed term. Each step solves Nv S-lines
 * (contiguous, n = Ns) and Ns v-lines (interleaved, n = Nv) with
 * pcrp_solve_f64 or pcrp_solve_f32r_f64 (the S-line solve stores Y1 + W,
 * the v-line right-hand side, directly; by default it also computes its own
 * right-hand side, the explicit stage, from the previous step's grid). reuse_factor = 1 factors the one
 * matrix all v-lines share (the v-line coefficients depend on v only) once
 * before the first step, in both precisions; the S-lines keep the full solve
 * (their matrices differ per line, and reading a per-line factor measured
 * slower than recomputing it); needs Nv <= 1024 (FP64) or 2048
 * (FP32+refine), else PCRP_ERR_UNSUPPORTED.
 * u: DEVICE array of Ns*Nv doubles, u[j*Ns + i-1], receives the price at
 * tau = steps*dt. use_graph = 1 captures a sequence of up to 20 time steps
 * in a hipGraph and replays it (plus one graph for the remaining steps;
 * needs a non-NULL stream; the call then synchronises `stream`
 * before releasing the graph); use_graph = 0 enqueues every step.
 */
/* Refinement settings of the FP32 line solves inside pcrp_heston_solve:
 * normwise tolerance and maximum corrections per line solve. */
#define PCRP_HESTON_REFINE_TOL 1e-12
#define PCRP_HESTON_MAX_SWEEPS 8

typedef struct {
    int ns, nv;                 /* unknowns in S and in v */
    double smax, vmax;          /* grid extents */
    double kappa, theta_v, sigma, rho, r, strike;
    double dt;                  /* time step */
    int steps;                  /* number of steps */
    double theta_scheme;        /* Douglas theta (0.5) */
    int reuse_factor;           /* 1: factor the shared v-line matrix once
                                   (pcrp_factor_*), solve its right-hand sides
                                   each step; 0: full line solves every step */
    int separate_explicit;      /* 0: the S-line solve kernels compute the
                                   explicit stage while they load (two launches
                                   per step; Ns <= 1024 FP64, <= 2048 FP32+refine,
                                   else as 1); 1: a separate explicit-stage kernel
                                   first. Both give the same bits. */
} pcrp_heston_params;

size_t pcrp_heston_workspace_size(const pcrp_heston_params *p, int precision);
int pcrp_heston_solve(const pcrp_heston_params *p, int precision, int use_graph, double *u,
                      void *ws, size_t ws_bytes, void *stream);

/*
 * Seeded input generator (device). kind 0: off-diagonals U[-1,1], main
 * U[4.5,5.5]; kind 1: off-diagonals U[-0.5,0.5], main U[2.2,2.6]; RHS
 * U[-1,1]; out-of-range diagonal entries zeroed. Values depend only on
 * (seed, diagonal, system, row), not on the layout.
 * Device counterpart of the reference's initialize kernel (kernel.cu:5-26).
 */
int pcrp_generate(int kind, uint64_t seed, int n, int m, int layout,
                  double *a, double *b, double *c, double *d, double *e, double *y,
                  void *stream);

/* The reference's own test system T^2 (T = tridiag(-1,2,-1)), y = 1:
 * kernel.cu:5-26 (initialize), including its out-of-range b_0 = d_{n-1} = 1. */
int pcrp_init_t2(int n, int m, int layout,
                 double *a, double *b, double *c, double *d, double *e, double *y,
                 void *stream);

/* ---- Autotuner --------------------------------------------------------
 * The reference hard-codes its launch shapes (tile_size = 512, one thread per
 * row, the n <= 1024 threshold: main.cu:16-22, 46-49). Here each size bucket
 * ceil(log2 n) maps to one compiled LDS-kernel variant (kind, threads NT, rows
 * per thread RPT); the table starts at a built-in default and is set by
 * timing every variant on the device. kind: 0 shared pair equations, 1
 * redundant (reference form), 2 in-place pair equations, 3 spaced rows (each
 * pair equation serves two of a thread's rows); +4 = persistent batch form
 * that prefetches the next system; +8 = two interleaved systems per
 * workgroup. */
int pcrp_lds_variant_count(int precision);
int pcrp_lds_variant_info(int precision, int id, int *kind, int *nt, int *rpt,
                          int *max_n, int *lds_bytes);
/* Pin the bucket containing n to variant_id (-1: back to the default). */
int pcrp_tune_set(int precision, int n, int variant_id);
/* Variant the next solve of size n would use (-1 if n is not an LDS size). */
int pcrp_tune_get(int precision, int n);
/* Time every variant that fits n on the caller's batch (y is copied into
 * y_work before each timed launch; a..e, y unchanged), keep the fastest for
 * n's bucket. Synchronises `stream`. times_ms: optional host array of
 * pcrp_lds_variant_count() entries (-1 for variants that do not fit). */
int pcrp_autotune(int precision, int n, int m, int layout,
                  const double *a, const double *b, const double *c,
                  const double *d, const double *e, const double *y,
                  double *y_work, int reps, void *stream,
                  int *best_id, float *best_ms, float *times_ms);
/* Split path (n above the LDS threshold): sub-systems of at most lds_max rows
 * (a power of two <= 1024) go to the LDS kernel, HBM passes fuse up to
 * k_fuse (1..6) reduction steps each, pass tile shape `shape` (0 ..
 * n_shapes-1). The reference fixes these at 1024, 1 and 512-row tiles
 * (main.cu:16-22, kernel.cu:182). Changing lds_max changes
 * pcrp_workspace_size(). */
int pcrp_split_tune_set(int precision, int lds_max, int k_fuse, int shape);
int pcrp_split_tune_get(int precision, int *lds_max, int *k_fuse, int *shape,
                        int *n_shapes);
/* Time every (lds_max, k_fuse, shape) of the split path on the caller's
 * batch (n must be above the current LDS threshold; ws as for pcrp_solve_f64;
 * lds_max from 256 to 1024, those whose workspace fits ws_bytes), keep the
 * fastest (pcrp_split_tune_get reads the chosen lds_max). Synchronises
 * `stream`. */
int pcrp_split_autotune(int precision, int n, int m, int layout,
                        const double *a, const double *b, const double *c,
                        const double *d, const double *e, const double *y,
                        double *y_work, void *ws, size_t ws_bytes, int reps, void *stream,
                        int *best_k, int *best_shape, float *best_ms);
/* FP32 refinement tunables (no counterpart in the reference): win_levels
 * (2..7, default 4) = reduction levels a window of a long system runs, with
 * 2^win_levels - 1 halo block rows per side (more levels: fewer batches fall
 * back to the split path, more redundant halo work); window_min_unknowns
 * (default 2^23) = batch size in unknowns from which systems of n in
 * (1024, 2048] take the window path instead of the one-workgroup kernel.
 * Results meet the same 1e-12 bar at every setting. */
int pcrp_f32r_tune_set(int win_levels, long long window_min_unknowns);
int pcrp_f32r_tune_get(int *win_levels, long long *window_min_unknowns);
/* Time pcrp_solve_f32r_f64 on the caller's batch (n > 1024; ws as for that
 * call) for win_levels 3..6 and, for n <= 2048, against the one-workgroup
 * kernel; keep the fastest (window_min_unknowns moves to include or exclude
 * this batch size). Synchronises `stream`; y is copied into y_work before
 * each timed solve. window_path: 1 if the window path won. */
int pcrp_f32r_autotune(int n, int m, int layout, const double *a, const double *b,
                       const double *c, const double *d, const double *e, const double *y,
                       double *y_work, void *ws, size_t ws_bytes, int reps, void *stream,
                       int *best_levels, int *window_path, float *best_ms);
/* Persist / restore every table (LDS variants, split tunables including the
 * LDS/HBM threshold, FP32 refinement tunables) as text. */
int pcrp_tune_save(const char *path);
int pcrp_tune_load(const char *path);

#ifdef __cplusplus
}
#endif
#endif /* PCR_PENTA_H */
