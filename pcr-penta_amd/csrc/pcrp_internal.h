// Internal host-side interfaces between the C-ABI layer and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <atomic>

namespace pcrp {

// Per-device launch constants (CU counts, resident grids), computed on the
// first call made with that device current and kept per device: safe for
// concurrent callers (two first calls compute the same value) and for
// processes that use several devices. compute(dev) returns the value, or
// a value <= 0 on failure (not cached, returned as is).
constexpr int kMaxDevices = 64;
template <typename F>
inline int device_cached(std::atomic<int> (&cache)[kMaxDevices], F&& compute)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -1;
    int v = cache[dev].load(std::memory_order_relaxed);
    if (v <= 0) {
        v = compute(dev);
        if (v > 0) cache[dev].store(v, std::memory_order_relaxed);
    }
    return v;
}

// Compute units of the current device (-1 if the query fails).
inline int device_cu_count()
{
    static std::atomic<int> cache[kMaxDevices];
    return device_cached(cache, [](int dev) {
        int cus = 0;
        return hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess
                   ? cus
                   : -1;
    });
}

// One batched FP64 solve over a strided view: element i of system j at
// a[j*ss + i*rs]; `steps` reduction steps, final delta = 2^steps.
// Split-path view of the final LDS stage: the batch is the m0 * 2^S residue
// sub-systems of m0 original systems of size n0, stored interleaved in a
// workspace (sub-system q = R*m0 + j holds rows R + 2^S*l of system j, element
// l at q + l*m); x goes back to the caller's y in the caller's layout.
struct SplitInfo {
    int S;            // steps already done in HBM passes (0: not a split solve);
                      // S > 0: the sub-systems sit in the grouped layout (split_slot)
    int n0, m0;       // original system size and count
    int out_layout;   // caller's layout (PCRP_LAYOUT_*) for x
    double* x;        // caller's y (receives x); S == 0: x's array (y's indexing)
    const double* add = nullptr;   // S == 0: x + add[] is stored (y's indexing)
};

struct SolveArgsF64 {
    const double *a, *b, *c, *d, *e;
    double* y;
    int n, m;
    long long rs, ss;
    int steps;
    SplitInfo split;
};

// Work queues of the persistent LDS kernel: one per XCD.
constexpr int kQueues = 8;

// One job of the persistent LDS queue: rows base + (i << stride_log2),
// i < n, of array set `src`; x written to xbase + (i << xstride_log2) of the
// output.
struct LdsJob {
    long long base, xbase;
    int stride_log2, xstride_log2, n, steps, src, pad;
};

// One system of an uneven-batch HBM pass round: rows at off + i, i < n, in
// the caller's arrays, its residue sub-systems at woff in the pass buffers
// (res_slot); k steps this round; the round's blocks for it start at block0.
struct PassSys {
    long long off, woff;
    long long block0;
    int n, k, tiles, pad;
};

// Residue sub-systems of a split system in the pass buffers. After s levels
// the system's rows form Q = 2^s residue sub-systems (residue q: rows
// q + l*2^s, l < L_q = ceil((n - q) / Q)). Groups of G = min(kResGroup, Q)
// consecutive residues are stored together, row-interleaved: element e of
// residue q at (q / G) * G * Lmax + e * G + q % G, Lmax = ceil(n / Q). A pass
// tile of G neighbouring residues then reads 8 G contiguous bytes per row
// (a wave's rows are one contiguous run), and an LDS job reads its residue at
// stride G: G of its rows per 8 G-double run, instead of one row per 128-byte
// line when every residue of the system is interleaved (stride Q).
// Slots of one level lie in [0, Q * Lmax) and are distinct; Q * Lmax grows
// with s, so 2^S * ceil(n / 2^S) slots hold every level of an S-level split.
constexpr int kResGroupLog2 = 2;

__host__ __device__ inline int res_group_log2(int s)
{
    return s < kResGroupLog2 ? s : kResGroupLog2;
}

__host__ __device__ inline long long res_slot(int q, long long e, int s, int n)
{
    const int gl = res_group_log2(s);
    const long long lmax = ((long long)n + (1LL << s) - 1) >> s;
    return ((long long)(q >> gl) * lmax << gl) + (e << gl) + (q & ((1 << gl) - 1));
}

// The same grouping for the uniform split path: after s levels, Q = m0 2^s
// sub-systems (q = R m0 + j: residue R of system j) of at most lmax rows;
// groups of G consecutive sub-systems (G = 4, or 2 or 1 when Q is not a
// multiple of 4) row-interleaved: element e of q at (q / G) G lmax + e G + q % G.
// Slots lie in [0, Q lmax), which grows with s.
__host__ __device__ inline long long split_slot(long long q, long long e, long long Q,
                                                long long lmax)
{
    const int gl = (Q & 3) == 0 ? 2 : ((Q & 1) == 0 ? 1 : 0);
    return ((q >> gl) * lmax << gl) + (e << gl) + (q & ((1 << gl) - 1));
}

__host__ __device__ inline int split_group_log2(long long Q)
{
    return (Q & 3) == 0 ? 2 : ((Q & 1) == 0 ? 1 : 0);
}

// Slots of the pass buffers one system of n rows split S levels takes.
__host__ __device__ inline long long res_region(int n, int S)
{
    return (((long long)n + (1LL << S) - 1) >> S) << S;
}

// A compiled LDS-kernel variant: kind (0 shared pair equations, 1 redundant,
// 2 in-place pair equations, 3 spaced rows; +4 for the persistent prefetching
// batch form, +8 for two interleaved systems per workgroup),
// NT threads x RPT rows per thread, capacity max_n = NT*RPT rows.
struct LdsVariantF64 {
    int kind, nt, rpt, max_n;
    int (*launch)(const SolveArgsF64&, hipStream_t);
    // persistent work-queue form over a job list (jobs read array set
    // sets[6*src .. 6*src+5], write x into x at the job's indices)
    int (*launch_queue)(const double* const* sets, double* x, const LdsJob* jobs, int njobs,
                        const int* qstart, unsigned int* counters, hipStream_t);
};

int lds_f64_variant_count();
const LdsVariantF64* lds_f64_variant(int id);
int lds_f64_default_variant(int n);
int lds_f64_default_variant_for_batch(int n, int m);
int lds_f64_queue_default_variant(int n);
int lds_f64_lds_bytes(const LdsVariantF64& v);
int lds_f64_max_n();

// FP32 PCR + FP64 refinement over the same strided view.
struct SolveArgsF32R {
    const double *a, *b, *c, *d, *e;
    double* y;
    int n, m;
    long long rs, ss;
    int steps;
    int max_sweeps;
    double tol;
    int* sweeps_out;
    // one-workgroup kernel only (lds_f32r_solve / lds_f32r_factor):
    bool shared = false;          // a..e hold one system (row i at i*ars) for every y
    long long ars = 0;            // shared: row stride of a..e
    double* out = nullptr;        // x goes here (null: into y)
    const double* add = nullptr;  // optional: x + add[] is stored (y's indexing)
    const void* fct = nullptr;    // cached FP32 factor (pcrp_factor_f32r)
};

// Variant the tuning table assigns to an LDS solve of size n (pcrp_api.hip).
int lds_tuned_variant(int prec, int n);
// The same for a whole batch: without a tuned entry, the batch-size-aware default.
int lds_tuned_variant(int prec, int n, int m);
// The same for the uneven work queue (jobs of every size up to n).
int lds_tuned_queue_variant(int prec, int n);

// Split path tunables: sub-systems up to lds_max rows go to the LDS kernel;
// HBM passes fuse up to k_fuse steps; shape indexes the pass tile shapes.
struct SplitTune {
    int lds_max, k_fuse, shape;
};
SplitTune split_f64_get();
int split_f64_set(const SplitTune& t);
int split_f64_shape_count();
int split_f64_levels(int n, int lds_max);
size_t split_f64_workspace(int n, int m);
int split_f64_solve(const SolveArgsF64& A, int layout, void* ws, size_t ws_bytes, hipStream_t st);
int launch_pass_uneven(int shape, const double* const* sets, int src, int dst, const PassSys* tab,
                       const int* bmap, long long blocks, int Sin, int kmax, hipStream_t st);
int pass_shape_tb(int shape);
int pass_shape_g(int shape);
int pass_shape_rpt(int shape);
int pass_shape_single(int shape);

// Right-hand side of a solve read from y, and x (+ add[]) stored: every
// solve except the Heston S-lines with the explicit stage folded in
// (HestonStage, pcrp_heston_ops.h).
struct PlainRhs {
    static constexpr bool kFused = false;
};
struct HestonStage;
// Heston S-lines with the explicit stage computed while the rows load:
// A.m contiguous systems of A.n rows (a..e at j*n + i), x + W stored to A.out
// (FP32+refine) or x (FP64). -2: n beyond the one-workgroup kernels.
int lds_f32r_solve_stage(const SolveArgsF32R& A, const HestonStage& src, hipStream_t st);
int lds_f64_solve_stage(const double* a, const double* b, const double* c, const double* d,
                        const double* e, double* x, int n, int m, const HestonStage& src,
                        hipStream_t st);

int lds_f32r_max_n();
int lds_f32r_solve(const SolveArgsF32R& A, hipStream_t st);
// Writes the FP32 block factor of A's systems (of the one system when
// A.shared) to F; a later lds_f32r_solve with A.fct = F skips the factor.
int lds_f32r_factor(const SolveArgsF32R& A, void* F, hipStream_t st);

// FP32 split solve + FP64 refinement for n above lds_f32r_max_n().
size_t split_f32r_workspace(int n, int m);
int split_f32r_solve(const SolveArgsF32R& A, int layout, void* ws, size_t ws_bytes,
                     hipStream_t st);
int split_f32r_passes(const SolveArgsF32R& A, void* ws, size_t ws_bytes, hipStream_t st);

// Plan forms of the uneven routes (pcrp_uneven.hip, pcrp_uneven_f32r.hip):
// tables uploaded once at plan time, launches only at execute.
struct UnevenF64Plan;
UnevenF64Plan* uneven_f64_plan(const double* a, const double* b, const double* c, const double* d,
                               const double* e, const int64_t* h_off, const int* h_n, int m,
                               void* ws, size_t ws_bytes, hipStream_t st, int& rc);
void uneven_f64_plan_free(UnevenF64Plan* U);
int uneven_f64_plan_execute(const UnevenF64Plan* U, double* y, hipStream_t st);
size_t uneven_f64_workspace(const int64_t* h_off, const int* h_n, int m);
struct UnevenF32RPlan;
UnevenF32RPlan* uneven_f32r_plan_tables(const int64_t* h_off, const int* h_n, int m, void* ws,
                                        size_t ws_bytes, hipStream_t st, int& rc);
void uneven_f32r_plan_free(UnevenF32RPlan* U);
int uneven_f32r_plan_execute(const UnevenF32RPlan* U, const double* a, const double* b,
                             const double* c, const double* d, const double* e, double* y, int m,
                             int max_sweeps, double tol, int* sweeps_out, void* ws,
                             hipStream_t st);
// FP32 inner solve + FP64 refinement over an uneven batch (pcrp_uneven_f32r.hip).
size_t uneven_f32r_workspace(const int64_t* h_off, const int* h_n, int m);

// Windowed FP32 refinement for diagonally dominant long / uneven batches
// (pcrp_window_f32r.hip). fell_back: the batch was not solved (a system's
// couplings outlast a window's halo, or the stream is capturing) and y is
// untouched; the caller then runs the pass/stage path.
size_t window_f32r_workspace(long long elems, int m, bool transposed);
size_t window_f32r_workspace_uniform(int n, int m, int layout);
// FP32 refinement tunables: levels a window of a longer system runs (halo
// 2^L - 1 block rows per side; 2..7, default 4) and the batch size (unknowns)
// from which n in (1024, 2048] takes the window path instead of the
// one-workgroup kernel (default 2^23; below, the read-back and the
// transposes cost more than the windows save).
struct F32RTune {
    int win_levels;
    long long window_min_unknowns;
};
F32RTune f32r_tune_get();
int f32r_tune_set(const F32RTune& t);
inline bool window_preferred(int n, int m)
{
    return n > 1024 && (long long)n * m >= f32r_tune_get().window_min_unknowns;
}
int window_f32r_uniform(const SolveArgsF32R& A, void* ws, size_t ws_bytes, hipStream_t st,
                        bool& fell_back);
int window_f32r_uneven(const double* a, const double* b, const double* c, const double* d,
                       const double* e, double* y, const int64_t* h_off, const int* h_n, int m,
                       int max_sweeps, double tol, int* sweeps_out, void* ws, size_t ws_bytes,
                       hipStream_t st, bool& fell_back);

// Factor once / solve many (pcrp_factor.hip). A factor holds, per system,
// (4 steps + 3) n doubles; shared: one factor for every right-hand side.
struct FactoredF64 {
    const double* F;
    bool shared;
    double* y;            // right-hand sides
    double* out;          // x (null: in place in y)
    const double* add;    // optional: x + add[] is stored (same indexing as y)
    int n, m;
    long long rs, ss;     // element i of right-hand side j at j*ss + i*rs
};
long long factor_f64_words(int n);
int factor_f64(const double* a, const double* b, const double* c, const double* d,
               const double* e, int n, int msys, long long rs, long long ss, double* F,
               hipStream_t st);
int solve_factored_f64(const FactoredF64& A, hipStream_t st);
size_t factor_f32r_bytes(int n);

// Plan form of the window path (pcrp_window_f32r.hip): the matrix-only
// preparation once (synchronous), then capturable executes.
struct WindowLayoutOffsets {
    size_t state, stop, flag, sys, lev, wmap, x0, x1, y0, six, bytes;
};
struct WindowPlan {
    WindowLayoutOffsets L;
    bool local, transposed, uniform;
    long long rs, elems, nwin;
    int n, m, nwu, wl;
    void* ws;
};
int window_f32r_plan_uniform(const SolveArgsF32R& A, void* ws, size_t ws_bytes, hipStream_t st,
                             WindowPlan& P);
int window_f32r_plan_uneven(const double* a, const double* b, const double* c, const double* d,
                            const double* e, const int64_t* h_off, const int* h_n, int m,
                            void* ws, size_t ws_bytes, hipStream_t st, WindowPlan& P);
int window_f32r_plan_execute(const WindowPlan& P, const double* a, const double* b,
                             const double* c, const double* d, const double* e, double* y,
                             int max_sweeps, double tol, int* sweeps_out, hipStream_t st);

inline int pcr_steps(long long n)
{
    int s = 0;
    while ((2LL << s) <= n) ++s;
    return n >= 2 ? s : 0;
}

}  // namespace pcrp
