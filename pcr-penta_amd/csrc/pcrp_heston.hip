// Heston ADI driver: the pentadiagonal solver's application in the PCR-Penta
// paper (convection-dominated Heston PDE). The reference repository ships no
// ADI code; this caller builds the line systems and steps the Douglas scheme.
//
// The PDE, the grid and the difference operators are in pcrp_heston_ops.h.
// S-lines are contiguous (m = Nv systems of n = Ns), v-lines interleaved
// (m = Ns systems of n = Nv). Douglas (theta = 1/2):
//   Y0 = U + dt (A0 U + A1 U + A2 U + b(tau_n))
//   (I - th dt A1) Y1 = Y0 - th dt (A1 U + b1(tau_n)) + th dt b1(tau_n+1)
//   (I - th dt A2) Y2 = Y1 - th dt (A2 U + b2(tau_n)) + th dt b2(tau_n+1)
//   U_{n+1} = Y2.
// A step is two launches on one stream with fixed pointers: the S-line
// solve, whose kernel computes its right-hand side R1 and the v-line term W
// (the explicit stage) from the previous grid while it loads its rows and
// stores Y1 + W into the other grid buffer, then the v-line solve in place
// there. With separate_explicit (or line sizes the folded kernels do not
// take) an explicit-stage kernel writes R1 and W first and the grid stays in
// u. A one-thread kernel advances the device-side step index once per
// captured chunk, so one captured hipGraph replays every chunk.
// With reuse_factor the v-line matrix (constant over the run, and one matrix
// for every v-line, since A2 depends on v only) is factored once before the
// first step: FP64 keeps its PCR-Penta multipliers, FP32+refine its 2x2-block
// factor; each step then solves right-hand sides only for the v-lines. The
// S-lines keep the full solve: their matrices differ per line, and reading a
// per-line factor (43 doubles per row at n = 1024, 180 MB per S-line solve)
// took longer than recomputing it (FP64: 36.2 vs 34.9 us per S-line solve).
#include "pcr_penta.h"
#include "pcrp_heston_ops.h"
#include "pcrp_internal.h"

#include <math.h>

namespace pcrp {

namespace {

using Grid = HestonGrid;

// (I - th dt A) row coefficients into the five diagonals, row-aligned:
// entries that would couple to a boundary node or fall outside are zero.
__global__ void setup_lines_kernel(Grid g, double* sa, double* sb, double* sc, double* sd,
                                   double* se, double* va, double* vb, double* vc, double* vd,
                                   double* ve, double* vm)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)g.ns * g.nv) return;
    const int j = (int)(idx / g.ns), i = (int)(idx % g.ns) + 1;
    const double f = -g.th * g.dt;
    double w[5];
    a1_coef(g, i, j, w);
    // S-line j, row i-1 (contiguous: element (i-1) of system j at idx)
    sa[idx] = (i - 2 >= 1) ? f * w[0] : 0.0;
    sb[idx] = (i - 1 >= 1) ? f * w[1] : 0.0;
    sc[idx] = 1.0 + f * w[2];
    sd[idx] = (i + 1 <= g.ns) ? f * w[3] : 0.0;
    se[idx] = (i + 2 <= g.ns) ? f * w[4] : 0.0;
    a2_coef(g, j, w);
    // v-line i-1, row j (interleaved: element j of system i-1 at j*ns + i-1 = idx)
    va[idx] = (j - 2 >= 0) ? f * w[0] : 0.0;
    vb[idx] = (j - 1 >= 0) ? f * w[1] : 0.0;
    vc[idx] = 1.0 + f * w[2];
    vd[idx] = (j + 1 <= g.nv - 1) ? f * w[3] : 0.0;
    ve[idx] = (j + 2 <= g.nv - 1) ? f * w[4] : 0.0;
    if (i == 1) {   // v-line 0 again, contiguous (row j at vm[k*nv + j]): every v-line's matrix
        vm[j] = va[idx];
        vm[g.nv + j] = vb[idx];
        vm[2 * g.nv + j] = vc[idx];
        vm[3 * g.nv + j] = vd[idx];
        vm[4 * g.nv + j] = ve[idx];
    }
}

__global__ void payoff_kernel(Grid g, double* U, int* step)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx == 0) *step = 0;
    if (idx >= (long long)g.ns * g.nv) return;
    const int i = (int)(idx % g.ns) + 1;
    const double s = S_at(g, i) - g.K;
    U[idx] = s > 0.0 ? s : 0.0;
}

// the explicit stage at unknown (i, j): R1 = RHS of the S-line system,
// W = v-line correction (bS0, bS1: the far-S boundary at tau0, tau0 + dt)
__device__ inline void explicit_node(const Grid& g, const double* U, double* R1, double* W, int i,
                                     int j, double bS0, double bS1)
{
    const int idx = j * g.ns + (i - 1);
    double w2[5];
    a2_coef(g, j, w2);
    heston_stage_node(g, U, i, j, bS0, bS1, w2, R1[idx], W[idx]);
}

// far-S boundary values of every step, once per run: bnd[2k] at
// tau = k dt, bnd[2k+1] at tau = k dt + dt (the explicit stage's tau0, tau1;
// the same expressions as computing them per node, without two FP64 exp per
// node and step)
__global__ void boundary_kernel(Grid g, double* bnd, int steps)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= steps) return;
    const double tau0 = k * g.dt, tau1 = tau0 + g.dt;
    bnd[2 * k] = far_s(g, tau0);
    bnd[2 * k + 1] = far_s(g, tau1);
}

// explicit stage: R1 = RHS of the S-line systems, W = v-line correction
// (step index: the device counter plus this step's offset within its chunk).
// Grid: x over S (i - 1), y over v (j, strided by gridDim.y), so j is
// uniform per workgroup.
__global__ void explicit_kernel(Grid g, const double* U, double* R1, double* W, const int* step,
                                int off, const double* __restrict__ bnd)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x) + 1;
    if (i > g.ns) return;
    const int k = *step + off;
    const double bS0 = bnd[2 * k], bS1 = bnd[2 * k + 1];
    for (int j = (int)blockIdx.y; j < g.nv; j += (int)gridDim.y)
        explicit_node(g, U, R1, W, i, j, bS0, bS1);
}


// the next chunk's first step index, once its explicit kernels have read
// this one (one thread: a completion counter across the explicit kernel's
// blocks measured 30 us per step in atomics)
__global__ void advance_kernel(int* step, int by) { *step += by; }

// X = Y1 + W, for grids whose S-line solve cannot store it directly
__global__ void combine_kernel(Grid g, const double* Y1, const double* W, double* X)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)g.ns * g.nv) return;
    X[idx] = Y1[idx] + W[idx];
}

Grid make_grid(const pcrp_heston_params* p)
{
    Grid g;
    g.ns = p->ns; g.nv = p->nv;
    g.smax = p->smax; g.vmax = p->vmax;
    g.hS = p->smax / (p->ns + 1);
    g.hv = p->vmax / p->nv;
    g.kappa = p->kappa; g.thv = p->theta_v; g.sigma = p->sigma; g.rho = p->rho; g.r = p->r;
    g.K = p->strike; g.dt = p->dt; g.th = p->theta_scheme;
    g.r12hS2 = 1.0 / (12.0 * g.hS * g.hS);
    g.r12hS = 1.0 / (12.0 * g.hS);
    g.rhS2 = 1.0 / (g.hS * g.hS);
    g.r2hS = 1.0 / (2.0 * g.hS);
    g.rhv2 = 1.0 / (g.hv * g.hv);
    g.r2hv = 1.0 / (2.0 * g.hv);
    g.rhv = 1.0 / g.hv;
    g.r4hShv = 1.0 / (4.0 * g.hS * g.hv);
    return g;
}

struct HestonWs {
    double *sa, *sb, *sc, *sd, *se, *va, *vb, *vc, *vd, *ve, *R1, *W;
    int* step;
    double* bnd;   // far-S boundary values per step (boundary_kernel)
    double* vm;    // the v-line matrix, contiguous: diagonal k at vm + k*nv
    void* solver_ws;
    size_t solver_bytes;
    void* fV;   // reuse_factor: the shared factor of the v-line matrix
};

size_t solver_ws_bytes(const pcrp_heston_params* p, int precision)
{
    size_t a = pcrp_workspace_size(p->ns, p->nv, precision);
    size_t b = pcrp_workspace_size(p->nv, p->ns, precision);
    return a > b ? a : b;
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// factor buffer of the reuse_factor form (0 when not used)
size_t factor_v_bytes(const pcrp_heston_params* p, int precision)
{
    if (!p->reuse_factor) return 0;
    return align256(pcrp_factor_size(p->nv, p->ns, precision, 1));
}

// the reuse_factor form needs line sizes the factored solves take
bool reuse_supported(const pcrp_heston_params* p, int precision)
{
    return !p->reuse_factor || factor_v_bytes(p, precision) != 0;
}

// two boundary values per time step
size_t bnd_bytes(const pcrp_heston_params* p) { return align256(2 * (size_t)p->steps * sizeof(double)); }

// the contiguous copy of the v-line matrix (five diagonals of nv rows)
size_t vm_bytes(const pcrp_heston_params* p) { return align256(5 * (size_t)p->nv * sizeof(double)); }

size_t heston_ws_bytes(const pcrp_heston_params* p, int precision)
{
    const size_t N = (size_t)p->ns * p->nv;
    return align256(12 * N * sizeof(double) + 256) + bnd_bytes(p) + vm_bytes(p) +
           solver_ws_bytes(p, precision) +
           factor_v_bytes(p, precision);
}

HestonWs carve(const pcrp_heston_params* p, int precision, void* ws)
{
    const size_t N = (size_t)p->ns * p->nv;
    double* d = static_cast<double*>(ws);
    HestonWs h;
    double** slots[12] = {&h.sa, &h.sb, &h.sc, &h.sd, &h.se, &h.va,
                          &h.vb, &h.vc, &h.vd, &h.ve, &h.R1, &h.W};
    for (int k = 0; k < 12; ++k) *slots[k] = d + (size_t)k * N;
    h.step = reinterpret_cast<int*>(d + 12 * N);
    char* c = static_cast<char*>(ws) + align256(12 * N * sizeof(double) + 256);
    h.bnd = reinterpret_cast<double*>(c);
    c += bnd_bytes(p);
    h.vm = reinterpret_cast<double*>(c);
    c += vm_bytes(p);
    h.solver_ws = c;
    h.solver_bytes = solver_ws_bytes(p, precision);
    c += h.solver_bytes;
    h.fV = factor_v_bytes(p, precision) ? c : nullptr;
    return h;
}

// The factor of the reuse_factor form, once, before the first step.
int factor_lines(const Grid& g, int precision, const HestonWs& h, hipStream_t st)
{
    if (!h.fV) return PCRP_OK;
    // the contiguous copy of v-line 0's rows is the matrix of every v-line
    const double* const vm = h.vm;
    const int nv = g.nv;
    if (precision == PCRP_PREC_F64)
        return factor_f64(vm, vm + nv, vm + 2 * nv, vm + 3 * nv, vm + 4 * nv, nv, 1, 1, 0,
                          static_cast<double*>(h.fV), st) == 0
                   ? PCRP_OK
                   : PCRP_ERR_HIP;
    SolveArgsF32R A{vm, vm + nv, vm + 2 * nv, vm + 3 * nv, vm + 4 * nv, nullptr, nv, g.ns, g.ns, 1,
                    pcr_steps(nv), 0, 0.0, nullptr};
    A.shared = true;
    A.ars = 1;
    return lds_f32r_factor(A, h.fV, st) == 0 ? PCRP_OK : PCRP_ERR_HIP;
}

// S-lines: R1 (contiguous, m = nv systems of n = ns) solved, Y1 + W stored in
// U. Returns PCRP_ERR_UNSUPPORTED when the solver has no such store for this
// size (the caller then solves in place and adds W separately).
int solve_s_lines(const Grid& g, int precision, double* U, const HestonWs& h, hipStream_t st)
{
    const int n = g.ns, m = g.nv;
    if (precision == PCRP_PREC_F64) {
        if (split_f64_levels(n, split_f64_get().lds_max) != 0) return PCRP_ERR_UNSUPPORTED;
        SolveArgsF64 A{h.sa, h.sb, h.sc, h.sd, h.se, h.R1, n, m, 1, n, pcr_steps(n),
                       SplitInfo{0, n, m, PCRP_LAYOUT_CONTIGUOUS, U}};
        A.split.add = h.W;
        const LdsVariantF64* v = lds_f64_variant(lds_tuned_variant(PCRP_PREC_F64, n, m));
        return v->launch(A, st) == 0 ? PCRP_OK : PCRP_ERR_HIP;
    }
    if (n > lds_f32r_max_n() || window_preferred(n, m)) return PCRP_ERR_UNSUPPORTED;
    SolveArgsF32R A{h.sa, h.sb, h.sc, h.sd, h.se, h.R1, n, m, 1, n, pcr_steps(n),
                    PCRP_HESTON_MAX_SWEEPS, PCRP_HESTON_REFINE_TOL, nullptr};
    A.out = U;
    A.add = h.W;
    return lds_f32r_solve(A, st) == 0 ? PCRP_OK : PCRP_ERR_HIP;
}

// v-lines: U (interleaved, m = ns systems of n = nv) solved in place
int solve_v_lines(const Grid& g, int precision, double* U, const HestonWs& h, hipStream_t st)
{
    const int n = g.nv, m = g.ns;
    if (h.fV && precision == PCRP_PREC_F64) {
        FactoredF64 A{static_cast<const double*>(h.fV), true, U, nullptr, nullptr, n, m, m, 1};
        return solve_factored_f64(A, st) == 0 ? PCRP_OK : PCRP_ERR_HIP;
    }
    if (h.fV) {
        // the shared matrix from its contiguous copy (coalesced rows, not v-line 0's
        // one-double-per-line gather)
        const double* const vm = h.vm;
        SolveArgsF32R A{vm, vm + n, vm + 2 * n, vm + 3 * n, vm + 4 * n, U, n, m, m, 1,
                        pcr_steps(n), PCRP_HESTON_MAX_SWEEPS, PCRP_HESTON_REFINE_TOL, nullptr};
        A.shared = true;
        A.ars = 1;
        A.fct = h.fV;
        return lds_f32r_solve(A, st) == 0 ? PCRP_OK : PCRP_ERR_HIP;
    }
    if (precision == PCRP_PREC_F64)
        return pcrp_solve_f64(h.va, h.vb, h.vc, h.vd, h.ve, U, n, m, PCRP_LAYOUT_INTERLEAVED,
                              h.solver_ws, h.solver_bytes, st);
    return pcrp_solve_f32r_f64(h.va, h.vb, h.vc, h.vd, h.ve, U, n, m, PCRP_LAYOUT_INTERLEAVED,
                               PCRP_HESTON_MAX_SWEEPS, PCRP_HESTON_REFINE_TOL, nullptr,
                               h.solver_ws, h.solver_bytes, st);
}

// The explicit stage folded into the S-line solve (the default): the S-line
// kernels compute R1 and W from the previous step's grid while they load
// their rows, so a step is two launches. They read rows j-2..j+2 of that grid
// while other lines' x is stored, so the grid alternates between u and a
// second buffer (h.R1, idle in this form): step k reads buffer k & 1.
bool stage_folded(const pcrp_heston_params* p, int precision)
{
    if (p->separate_explicit) return false;
    if (precision == PCRP_PREC_F64) return p->ns <= lds_f64_max_n();
    return p->ns <= lds_f32r_max_n() && !window_preferred(p->ns, p->nv);
}

// S-lines with the stage folded in: grid Uin (never written), Y1 + W to Uout
int solve_s_lines_folded(const Grid& g, int precision, const double* Uin, double* Uout,
                         const HestonWs& h, int off, hipStream_t st)
{
    const int n = g.ns, m = g.nv;
    const HestonStage src{g, Uin, h.step, off, h.bnd};
    if (precision == PCRP_PREC_F64)
        return lds_f64_solve_stage(h.sa, h.sb, h.sc, h.sd, h.se, Uout, n, m, src, st) == 0
                   ? PCRP_OK
                   : PCRP_ERR_HIP;
    SolveArgsF32R A{h.sa, h.sb, h.sc, h.sd, h.se, nullptr, n, m, 1, n, pcr_steps(n),
                    PCRP_HESTON_MAX_SWEEPS, PCRP_HESTON_REFINE_TOL, nullptr};
    A.out = Uout;
    return lds_f32r_solve_stage(A, src, st) == 0 ? PCRP_OK : PCRP_ERR_HIP;
}

// One Douglas step (step index: device counter + off) from grid Uin to grid
// Uout. Folded: two launches, Uin != Uout. Separate: the explicit kernel,
// then the S-line solve (three launches; four when the S-line solve cannot
// store Y1 + W itself), in place (Uin == Uout).
int enqueue_step(const Grid& g, int precision, bool folded, const double* Uin, double* Uout,
                 const HestonWs& h, int off, hipStream_t st)
{
    if (folded) {
        const int rc = solve_s_lines_folded(g, precision, Uin, Uout, h, off, st);
        return rc != PCRP_OK ? rc : solve_v_lines(g, precision, Uout, h, st);
    }
    double* const U = Uout;
    const long long N = (long long)g.ns * g.nv;
    const int bs = 256;
    const unsigned blocks = (unsigned)((N + bs - 1) / bs);
    hipLaunchKernelGGL(explicit_kernel, dim3((g.ns + bs - 1) / bs, g.nv < 65535 ? g.nv : 65535),
                       dim3(bs), 0, st, g, U,
                       h.R1, h.W, h.step, off, h.bnd);
    if (hipGetLastError() != hipSuccess) return PCRP_ERR_HIP;
    int rc = solve_s_lines(g, precision, U, h, st);
    if (rc == PCRP_ERR_UNSUPPORTED) {
        rc = precision == PCRP_PREC_F64
                 ? pcrp_solve_f64(h.sa, h.sb, h.sc, h.sd, h.se, h.R1, g.ns, g.nv,
                                  PCRP_LAYOUT_CONTIGUOUS, h.solver_ws, h.solver_bytes, st)
                 : pcrp_solve_f32r_f64(h.sa, h.sb, h.sc, h.sd, h.se, h.R1, g.ns, g.nv,
                                       PCRP_LAYOUT_CONTIGUOUS, PCRP_HESTON_MAX_SWEEPS,
                                       PCRP_HESTON_REFINE_TOL, nullptr, h.solver_ws,
                                       h.solver_bytes, st);
        if (rc != PCRP_OK) return rc;
        hipLaunchKernelGGL(combine_kernel, dim3(blocks), dim3(bs), 0, st, g, h.R1, h.W, U);
        if (hipGetLastError() != hipSuccess) return PCRP_ERR_HIP;
    } else if (rc != PCRP_OK) {
        return rc;
    }
    return solve_v_lines(g, precision, U, h, st);
}

// `count` consecutive steps starting from grid buffer `parity` (folded form:
// buffer 0 is u, buffer 1 is h.R1; separate form: always u), then the counter
// moves past them
int enqueue_steps(const Grid& g, int precision, bool folded, double* u, const HestonWs& h,
                  int count, int parity, hipStream_t st)
{
    double* const buf[2] = {u, folded ? h.R1 : u};
    for (int k = 0; k < count; ++k) {
        const int b = (parity + k) & 1;
        const int rc = enqueue_step(g, precision, folded, buf[b], buf[b ^ 1], h, k, st);
        if (rc != PCRP_OK) return rc;
    }
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(1), 0, st, h.step, count);
    return hipGetLastError() == hipSuccess ? PCRP_OK : PCRP_ERR_HIP;
}

// Steps per captured graph: one graph launch per chunk instead of per step
// (each graph launch left ~8 us between steps on MI355X), with the counter
// advanced once per chunk. Even, so that every replay of a folded-form chunk
// starts and ends in u.
constexpr int kStepsPerGraph = 20;

// Captures `count` steps into an executable graph.
int capture_steps(const Grid& g, int precision, bool folded, double* u, const HestonWs& h,
                  int count, hipStream_t st, hipGraphExec_t* exec)
{
    hipGraph_t graph = nullptr;
    *exec = nullptr;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return PCRP_ERR_HIP;
    const int rc = enqueue_steps(g, precision, folded, u, h, count, 0, st);
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    int out = rc != PCRP_OK ? rc : (ec != hipSuccess ? PCRP_ERR_HIP : PCRP_OK);
    if (out == PCRP_OK && hipGraphInstantiate(exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        *exec = nullptr;
        out = PCRP_ERR_HIP;
    }
    if (graph) (void)hipGraphDestroy(graph);
    return out;
}

}  // namespace
}  // namespace pcrp

extern "C" {

size_t pcrp_heston_workspace_size(const pcrp_heston_params* p, int precision)
{
    if (!p || p->ns < 2 || p->nv < 2) return 0;
    return pcrp::heston_ws_bytes(p, precision);
}

int pcrp_heston_solve(const pcrp_heston_params* p, int precision, int use_graph, double* u,
                      void* ws, size_t ws_bytes, void* stream)
{
    using namespace pcrp;
    if (!p || !u || !ws) return PCRP_ERR_NULL;
    if (p->ns < 2 || p->nv < 2 || p->steps < 0) return PCRP_ERR_INVALID_SIZE;
    if ((long long)p->ns * p->nv > 0x7fffffffLL) return PCRP_ERR_INVALID_SIZE;
    if (precision != PCRP_PREC_F64 && precision != PCRP_PREC_F32_REFINE) return PCRP_ERR_ARG;
    if (!(p->dt > 0.0) || !(p->smax > 0.0) || !(p->vmax > 0.0)) return PCRP_ERR_ARG;
    if (p->reuse_factor != 0 && p->reuse_factor != 1) return PCRP_ERR_ARG;
    if (p->separate_explicit != 0 && p->separate_explicit != 1) return PCRP_ERR_ARG;
    if (!reuse_supported(p, precision)) return PCRP_ERR_UNSUPPORTED;
    if (ws_bytes < heston_ws_bytes(p, precision)) return PCRP_ERR_WORKSPACE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const Grid g = make_grid(p);
    const HestonWs h = carve(p, precision, ws);
    const long long N = (long long)g.ns * g.nv;
    const int bs = 256;
    const unsigned blocks = (unsigned)((N + bs - 1) / bs);
    hipLaunchKernelGGL(setup_lines_kernel, dim3(blocks), dim3(bs), 0, st, g, h.sa, h.sb, h.sc,
                       h.sd, h.se, h.va, h.vb, h.vc, h.vd, h.ve, h.vm);
    hipLaunchKernelGGL(payoff_kernel, dim3(blocks), dim3(bs), 0, st, g, u, h.step);
    if (p->steps > 0)
        hipLaunchKernelGGL(boundary_kernel, dim3((p->steps + bs - 1) / bs), dim3(bs), 0, st, g,
                           h.bnd, p->steps);
    if (hipGetLastError() != hipSuccess) return PCRP_ERR_HIP;
    const int frc = factor_lines(g, precision, h, st);
    if (frc != PCRP_OK) return frc;
    const bool folded = stage_folded(p, precision);
    // folded form, odd step count: the last step's grid is in h.R1
    const auto finish = [&]() {
        if (!folded || (p->steps & 1) == 0) return PCRP_OK;
        return hipMemcpyAsync(u, h.R1, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, st) ==
                       hipSuccess
                   ? PCRP_OK
                   : PCRP_ERR_HIP;
    };
    if (!use_graph) {
        for (int n = 0; n < p->steps; ++n) {
            const int rc = enqueue_steps(g, precision, folded, u, h, 1, n & 1, st);
            if (rc != PCRP_OK) return rc;
        }
        return finish();
    }
    if (p->steps == 0) return PCRP_OK;
    if (st == nullptr) return PCRP_ERR_ARG;   // the null stream cannot be captured
    // one graph of up to kStepsPerGraph steps replayed, and one for the rest
    const int K = p->steps < kStepsPerGraph ? p->steps : kStepsPerGraph;
    const int full = p->steps / K, rest = p->steps % K;
    hipGraphExec_t exec = nullptr, exec_rest = nullptr;
    int out = capture_steps(g, precision, folded, u, h, K, st, &exec);
    if (out == PCRP_OK && rest > 0)
        out = capture_steps(g, precision, folded, u, h, rest, st, &exec_rest);
    for (int c = 0; c < full && out == PCRP_OK; ++c)
        if (hipGraphLaunch(exec, st) != hipSuccess) out = PCRP_ERR_HIP;
    if (out == PCRP_OK && exec_rest && hipGraphLaunch(exec_rest, st) != hipSuccess)
        out = PCRP_ERR_HIP;
    if (out == PCRP_OK) out = finish();
    // replays finish before the executable graphs are released
    if ((exec || exec_rest) && hipStreamSynchronize(st) != hipSuccess) out = PCRP_ERR_HIP;
    if (exec && hipGraphExecDestroy(exec) != hipSuccess) out = PCRP_ERR_HIP;
    if (exec_rest && hipGraphExecDestroy(exec_rest) != hipSuccess) out = PCRP_ERR_HIP;
    return out;
}

}  // extern "C"
