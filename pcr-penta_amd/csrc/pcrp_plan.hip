// Plan / execute: everything a solve decides or uploads from the matrix and
// the batch shape happens once, at plan time; executes only enqueue kernels.
//
// The reference has no host API at all (its "solve call" is the launch pair
// of main.cu:46-87, synchronised after every launch). The one-call entry
// points here are asynchronous, but two of them do host work per call: the
// uneven batches upload their routing tables (pass rounds, LDS jobs) from
// host memory, and the FP32 refinement of long systems reads back, after its
// first sweep, whether every window of the batch was local (else it solves
// with the split path). A plan does both once: pcrp_plan_create uploads the
// tables, runs the window path's level-finding sweep on the matrix (the
// locality test and each window's level count depend on a..e only) and,
// for interleaved batches, keeps the system-contiguous copies of a..e; it
// synchronises its stream. pcrp_plan_execute then launches kernels only: no
// allocation, no host copy, no synchronisation, so it can be captured in a
// hipGraph and replayed.
#include "pcr_penta.h"
#include "pcrp_internal.h"

#include <new>

struct pcrp_plan {
    int precision;
    bool uneven;
    int n, m, layout;
    const double *a, *b, *c, *d, *e;
    long long rs, ss;
    void* ws;
    size_t ws_bytes;
    int route;
    pcrp::WindowPlan win;
    pcrp::UnevenF64Plan* u64;
    pcrp::UnevenF32RPlan* u32;
};

namespace {

bool stream_capturing(hipStream_t st)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
}

void release(pcrp_plan* p)
{
    if (!p) return;
    if (p->u64) pcrp::uneven_f64_plan_free(p->u64);
    if (p->u32) pcrp::uneven_f32r_plan_free(p->u32);
    delete p;
}

}  // namespace

extern "C" {

size_t pcrp_plan_workspace_size(int precision, int n, int m, const int64_t* h_offsets,
                                const int* h_n)
{
    if (h_offsets || h_n) {
        if (!h_offsets || !h_n) return 0;
        return pcrp_uneven_workspace_size(h_offsets, h_n, m, precision);
    }
    return pcrp_workspace_size(n, m, precision);
}

int pcrp_plan_create(pcrp_plan** plan, int precision, const double* a, const double* b,
                     const double* c, const double* d, const double* e, int n, int m, int layout,
                     const int64_t* h_offsets, const int* h_n, void* ws, size_t ws_bytes,
                     void* stream)
{
    using namespace pcrp;
    if (!plan) return PCRP_ERR_NULL;
    *plan = nullptr;
    if (!a || !b || !c || !d || !e) return PCRP_ERR_NULL;
    if (precision != PCRP_PREC_F64 && precision != PCRP_PREC_F32_REFINE) return PCRP_ERR_ARG;
    const bool uneven = h_offsets || h_n;
    if (uneven && (!h_offsets || !h_n)) return PCRP_ERR_NULL;
    if (m < 1) return PCRP_ERR_INVALID_SIZE;
    if (!uneven && (n < 1 || (long long)n * m > 0x7fffffffLL)) return PCRP_ERR_INVALID_SIZE;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (stream_capturing(st)) return PCRP_ERR_ARG;   // plan creation synchronises its stream
    const size_t need = pcrp_plan_workspace_size(precision, n, m, h_offsets, h_n);
    if (uneven && need == 0) return PCRP_ERR_INVALID_SIZE;
    if (need > 0 && (!ws || ws_bytes < need)) return PCRP_ERR_WORKSPACE;
    pcrp_plan* p = new (std::nothrow) pcrp_plan{};
    if (!p) return PCRP_ERR_HIP;
    p->precision = precision;
    p->uneven = uneven;
    p->n = n;
    p->m = m;
    p->layout = layout;
    p->a = a; p->b = b; p->c = c; p->d = d; p->e = e;
    p->ws = ws;
    p->ws_bytes = ws_bytes;
    int rc = PCRP_OK;
    if (uneven) {
        if (precision == PCRP_PREC_F64) {
            p->route = PCRP_ROUTE_UNEVEN_F64;
            p->u64 = uneven_f64_plan(a, b, c, d, e, h_offsets, h_n, m, ws, ws_bytes, st, rc);
        } else {
            rc = window_f32r_plan_uneven(a, b, c, d, e, h_offsets, h_n, m, ws, ws_bytes, st,
                                         p->win);
            if (rc == PCRP_OK && p->win.local) {
                p->route = PCRP_ROUTE_WINDOW;
            } else if (rc == PCRP_OK) {
                p->route = PCRP_ROUTE_UNEVEN_STAGE;
                p->u32 = uneven_f32r_plan_tables(h_offsets, h_n, m, ws, ws_bytes, st, rc);
            }
        }
    } else {
        if (layout == PCRP_LAYOUT_CONTIGUOUS) { p->rs = 1; p->ss = n; }
        else if (layout == PCRP_LAYOUT_INTERLEAVED) { p->rs = m; p->ss = 1; }
        else rc = PCRP_ERR_LAYOUT;
        if (rc == PCRP_OK && precision == PCRP_PREC_F64) {
            p->route = split_f64_levels(n, split_f64_get().lds_max) == 0 ? PCRP_ROUTE_LDS
                                                                          : PCRP_ROUTE_SPLIT;
        } else if (rc == PCRP_OK) {
            SolveArgsF32R A{a, b, c, d, e, nullptr, n, m, p->rs, p->ss, pcr_steps(n), 0, 0.0,
                            nullptr};
            const bool try_window = n > lds_f32r_max_n() || window_preferred(n, m);
            p->route = n > lds_f32r_max_n() ? PCRP_ROUTE_SPLIT : PCRP_ROUTE_LDS;
            if (try_window) {
                rc = window_f32r_plan_uniform(A, ws, ws_bytes, st, p->win);
                if (rc == PCRP_OK && p->win.local) p->route = PCRP_ROUTE_WINDOW;
            }
        }
    }
    if (rc == PCRP_OK && hipStreamSynchronize(st) != hipSuccess) rc = PCRP_ERR_HIP;
    if (rc != PCRP_OK) {
        release(p);
        return rc;
    }
    *plan = p;
    return PCRP_OK;
}

int pcrp_plan_route(const pcrp_plan* plan) { return plan ? plan->route : -1; }

int pcrp_plan_execute(const pcrp_plan* plan, double* y, int max_sweeps, double tol,
                      int* sweeps_out, void* stream)
{
    using namespace pcrp;
    if (!plan || !y) return PCRP_ERR_NULL;
    const pcrp_plan& p = *plan;
    if (p.precision == PCRP_PREC_F32_REFINE && (max_sweeps < 0 || !(tol >= 0.0)))
        return PCRP_ERR_ARG;
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (p.route) {
    case PCRP_ROUTE_UNEVEN_F64:
        return uneven_f64_plan_execute(p.u64, y, st);
    case PCRP_ROUTE_UNEVEN_STAGE:
        return uneven_f32r_plan_execute(p.u32, p.a, p.b, p.c, p.d, p.e, y, p.m, max_sweeps, tol,
                                        sweeps_out, p.ws, st);
    case PCRP_ROUTE_WINDOW:
        return window_f32r_plan_execute(p.win, p.a, p.b, p.c, p.d, p.e, y, max_sweeps, tol,
                                        sweeps_out, st);
    default:
        break;
    }
    if (p.precision == PCRP_PREC_F64)   // uniform: no tables, no host work
        return pcrp_solve_f64(p.a, p.b, p.c, p.d, p.e, y, p.n, p.m, p.layout, p.ws, p.ws_bytes,
                              stream);
    SolveArgsF32R A{p.a, p.b, p.c, p.d, p.e, y, p.n, p.m, p.rs, p.ss, pcr_steps(p.n), max_sweeps,
                    tol, sweeps_out};
    if (p.route == PCRP_ROUTE_LDS) return lds_f32r_solve(A, st) == 0 ? PCRP_OK : PCRP_ERR_HIP;
    const int rc = split_f32r_passes(A, p.ws, p.ws_bytes, st);
    return rc == 0 ? PCRP_OK : (rc == -3 ? PCRP_ERR_WORKSPACE : PCRP_ERR_HIP);
}

int pcrp_plan_destroy(pcrp_plan* plan)
{
    release(plan);
    return PCRP_OK;
}

}  // extern "C"
