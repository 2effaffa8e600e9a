// FP32 PCR-Penta + FP64 iterative refinement for systems too large for one
// workgroup (n above the LDS FP32 kernel's limit).
//
// No counterpart in the reference, which is FP64-only; the inner solve is the
// reference's split scheme (shared_cached_shared_tile_splitting,
// kernel.cu:158-320, then pcr_penta_splitting_2nd_stage, kernel.cu:665-791)
// in FP32: the HBM passes of the FP64 split path instantiated for float, then
// an FP32 LDS stage that solves every residue sub-system and adds its
// correction to x in FP64.
//
// One refinement sweep (every kernel skips systems already stopped):
//   residual : r = y - A x in FP64 from the FP64 inputs, stored as FP32
//              (first sweep: r = y, x = 0);
//   passes   : S = ceil(log2(n / lds_max)) FP32 reduction steps in HBM;
//   stage    : FP32 LDS solve of the 2^S sub-systems, x += dx (FP64), and
//              per-system max |dx|, max |x| (atomic max on the bit patterns,
//              which order like the values for non-negative doubles);
//   decide   : the stopping rule of the one-workgroup FP32 kernel: stop when
//              max|dx| <= tol max|x|, or when the observed contraction
//              rho = |dx_k| / |dx_{k-1}| (from the second correction on)
//              puts the error left, 10 rho |dx_k|, below tol with rho <= 0.1.
//              A system that is not done after max_sweeps corrections, or
//              whose correction is NaN, reports max_sweeps + 1.
// The sweep loop is unrolled on the host (max_sweeps + 1 solves); the
// stop flags make the launches of converged systems return at once, so the
// call stays asynchronous (no host read-back) and graph-capturable.
#include "pcr_penta.h"
#include "pcrp_device.h"
#include "pcrp_internal.h"

namespace pcrp {

namespace {

// Per-system refinement state (device).
struct SysState {
    unsigned long long dmax, xmax;   // bit patterns of max |dx|, max |x| this sweep
    double dprev;                    // max |dx| of the previous sweep
    int sweep, pad;
};

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

struct Layout {
    size_t state, stop, y0, a32, sets, bytes;
    size_t per;   // elements per pass array
    int S;
};

Layout plan(int n, int m)
{
    Layout L{};
    const SplitTune tune = split_f64_get();
    L.S = split_f64_levels(n, tune.lds_max);
    const size_t N = (size_t)n * m;
    L.per = (size_t)m * ((size_t)((n + (1 << L.S) - 1) >> L.S) << L.S);
    size_t o = 0;
    L.state = o; o = align_up(o + (size_t)m * sizeof(SysState));
    L.stop = o;  o = align_up(o + (size_t)m * sizeof(int));
    L.y0 = o;    o = align_up(o + N * sizeof(double));
    L.a32 = o;   o = align_up(o + 6 * N * sizeof(float));    // a..e and r, caller layout
    L.sets = o;  o = align_up(o + 2 * 6 * L.per * sizeof(float));
    L.bytes = o;
    return L;
}

__device__ __forceinline__ unsigned long long abs_bits(double v)
{
    return (unsigned long long)__double_as_longlong(fabs(v));
}

__global__ void init_state_kernel(SysState* __restrict__ st, int* __restrict__ stop, int m)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) {
        st[j] = SysState{0ull, 0ull, 0.0, 0, 0};
        stop[j] = 0;
    }
}

// Row i of system j at element e (both layouts are dense): i, j from e.
__device__ __forceinline__ void row_of(long long e, int n, int m, int inter, int& i, int& j)
{
    if (inter) { i = (int)(e / m); j = (int)(e - (long long)i * m); }
    else       { j = (int)(e / n); i = (int)(e - (long long)j * n); }
}

// FP32 copies of a..e (entries outside the band zeroed, so no out-of-range
// value can overflow FP32) and of y as the first residual; y0 keeps y.
__global__ void prep_kernel(const double* __restrict__ a, const double* __restrict__ b,
                            const double* __restrict__ c, const double* __restrict__ d,
                            const double* __restrict__ e, const double* __restrict__ y,
                            float* __restrict__ A32, double* __restrict__ y0, int n, int m,
                            int inter)
{
    const long long N = (long long)n * m;
    const long long el = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (el >= N) return;
    int i, j;
    row_of(el, n, m, inter, i, j);
    A32[0 * N + el] = i >= 2 ? (float)a[el] : 0.f;
    A32[1 * N + el] = i >= 1 ? (float)b[el] : 0.f;
    A32[2 * N + el] = (float)c[el];
    A32[3 * N + el] = i + 1 < n ? (float)d[el] : 0.f;
    A32[4 * N + el] = i + 2 < n ? (float)e[el] : 0.f;
    const double yv = y[el];
    A32[5 * N + el] = (float)yv;
    y0[el] = yv;
}

// r = y0 - A x in FP64 (same operation order as the one-workgroup kernel).
__global__ void residual_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                const double* __restrict__ c, const double* __restrict__ d,
                                const double* __restrict__ e, const double* __restrict__ y0,
                                const double* __restrict__ x, float* __restrict__ r32,
                                const int* __restrict__ stop, int n, int m, int inter)
{
    const long long N = (long long)n * m;
    const long long el = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (el >= N) return;
    int i, j;
    row_of(el, n, m, inter, i, j);
    if (stop[j]) return;
    const long long rs = inter ? m : 1;
    const double xm2 = i >= 2 ? x[el - 2 * rs] : 0.0, xm1 = i >= 1 ? x[el - rs] : 0.0;
    const double xp1 = i + 1 < n ? x[el + rs] : 0.0, xp2 = i + 2 < n ? x[el + 2 * rs] : 0.0;
    double ax = c[el] * x[el];
    ax = __builtin_fma(a[el], xm2, ax);
    ax = __builtin_fma(b[el], xm1, ax);
    ax = __builtin_fma(d[el], xp1, ax);
    ax = __builtin_fma(e[el], xp2, ax);
    r32[el] = (float)(y0[el] - ax);
}

// FP32 LDS stage: one workgroup per residue sub-system q = R*m0 + j (element
// l at q + l*Q of the pass output), `steps` more reduction steps, the final
// 2x2 / 1x1 solve, then x += dx at the caller's element of row R + 2^S l.
template <int NT, int RPT>
__global__ __launch_bounds__(NT) void stage_f32_kernel(
    const float* __restrict__ ga, const float* __restrict__ gb, const float* __restrict__ gc,
    const float* __restrict__ gd, const float* __restrict__ ge, const float* __restrict__ gy,
    int Q, int steps, int S, int n0, int m0, long long xrs, long long xss,
    double* __restrict__ x, SysState* __restrict__ st, const int* __restrict__ stop, int first)
{
    constexpr int NMAX = NT * RPT;
    __shared__ float sR[6][NMAX];
    __shared__ unsigned long long red[2][NT / 64];
    const int q = xcd_remap(blockIdx.x, Q);
    const int R = q / m0, j = q - R * m0;
    if (stop[j]) return;   // uniform over the workgroup
    const int n = (n0 - R + (1 << S) - 1) >> S;
    const int t = threadIdx.x;
    Row<float> r[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        if (i < n) {
            const long long g = (long long)q + (long long)i * Q;
            r[k].a = ga[g]; r[k].b = gb[g]; r[k].c = gc[g];
            r[k].d = gd[g]; r[k].e = ge[g]; r[k].y = gy[g];
        } else {
            r[k] = Row<float>{0.f, 0.f, 1.f, 0.f, 0.f, 0.f};
        }
        sR[0][t + NT * k] = r[k].a; sR[1][t + NT * k] = r[k].b; sR[2][t + NT * k] = r[k].c;
        sR[3][t + NT * k] = r[k].d; sR[4][t + NT * k] = r[k].e; sR[5][t + NT * k] = r[k].y;
    }
    __syncthreads();
    auto readR = [&](int i) {
        Row<float> w;
        w.a = sR[0][i]; w.b = sR[1][i]; w.c = sR[2][i];
        w.d = sR[3][i]; w.e = sR[4][i]; w.y = sR[5][i];
        return w;
    };
    int delta = 1;
    for (int s = 0; s < steps; ++s, delta <<= 1) {
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int i = t + NT * k;
            if (i < n) {
                const bool hp1 = i - delta >= 0, hq1 = i + delta <= n - 1;
                const bool hp2 = i - 2 * delta >= 0, hq2 = i + 2 * delta <= n - 1;
                const bool hp3 = i - 3 * delta >= 0, hq3 = i + 3 * delta <= n - 1;
                const Row<float> P1 = readR(hp1 ? i - delta : i);
                const Row<float> P2 = readR(hp2 ? i - 2 * delta : i);
                const Row<float> Q1 = readR(hq1 ? i + delta : i);
                const Row<float> Q2 = readR(hq2 ? i + 2 * delta : i);
                const Pair<float> EL = make_pair(P2, P1, r[k], hp2, hp3, true, hq1);
                const Pair<float> ER = make_pair(r[k], Q1, Q2, true, hp1, hq2, hq3);
                r[k] = reduce_row(r[k], EL, ER, hp1, hq1, hp2, hq2,
                                  [](float&, float&, float&, float&) {});
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int i = t + NT * k;
            if (i < n) {
                sR[0][i] = r[k].a; sR[1][i] = r[k].b; sR[2][i] = r[k].c;
                sR[3][i] = r[k].d; sR[4][i] = r[k].e; sR[5][i] = r[k].y;
            }
        }
        __syncthreads();
    }
    // final 2x2 / 1x1 (kernel.cu:145-153's branches), x += dx in FP64
    unsigned long long dmax = 0ull, xmax = 0ull;
    auto update = [&](int i, float dxf) {
        const long long g = (long long)j * xss + (long long)(R + ((long long)i << S)) * xrs;
        const double dx = (double)dxf;
        const double xn = first ? dx : x[g] + dx;
        x[g] = xn;
        const unsigned long long bd = abs_bits(dx), bx = abs_bits(xn);
        dmax = bd > dmax ? bd : dmax;
        xmax = bx > xmax ? bx : xmax;
    };
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        if (i < n) {
            if (i < n - delta) {
                const float cii = sR[2][i + delta], bii = sR[1][i + delta], yii = sR[5][i + delta];
                const float deno = fmaf(r[k].c, cii, -(bii * r[k].d));
                const float rd = fabsf(deno) >= kTinyF32 ? rcp_f32(deno) : 0.f;
                update(i, fmaf(cii, r[k].y, -(r[k].d * yii)) * rd);
                update(i + delta, fmaf(r[k].c, yii, -(bii * r[k].y)) * rd);
            } else if (i < delta) {
                update(i, fabsf(r[k].c) >= kTinyF32 ? r[k].y * rcp_f32(r[k].c) : 0.f);
            }
        }
    }
    // workgroup max, then one atomic per quantity
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long od = __shfl_xor(dmax, off, 64), ox = __shfl_xor(xmax, off, 64);
        dmax = od > dmax ? od : dmax;
        xmax = ox > xmax ? ox : xmax;
    }
    const int lane = t & 63, wid = t >> 6;
    if (lane == 0) { red[0][wid] = dmax; red[1][wid] = xmax; }
    __syncthreads();
    if (t == 0) {
#pragma unroll
        for (int w = 1; w < NT / 64; ++w) {
            dmax = red[0][w] > dmax ? red[0][w] : dmax;
            xmax = red[1][w] > xmax ? red[1][w] : xmax;
        }
        atomicMax(&st[j].dmax, dmax);
        atomicMax(&st[j].xmax, xmax);
    }
}

__global__ void decide_kernel(SysState* __restrict__ st, int* __restrict__ stop, int m,
                              int max_sweeps, double tol, int* __restrict__ sweeps_out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m || stop[j]) return;
    SysState s = st[j];
    const double dmax = __longlong_as_double((long long)s.dmax);
    const double xmax = __longlong_as_double((long long)s.xmax);
    const bool bad = dmax != dmax || xmax != xmax;
    const bool small = !(dmax > tol * xmax);
    const bool contracting = s.sweep > 1 && dmax <= 0.1 * s.dprev &&
                             10.0 * dmax * dmax <= tol * xmax * s.dprev;
    const bool done = !bad && (small || contracting);
    s.dprev = dmax;
    if (done || bad || s.sweep >= max_sweeps) {
        if (bad) s.sweep = max_sweeps + 1;
        else if (!done) ++s.sweep;          // flag: tolerance not met
        stop[j] = 1;
        if (sweeps_out) sweeps_out[j] = s.sweep;
    } else {
        ++s.sweep;
    }
    s.dmax = 0ull;
    s.xmax = 0ull;
    st[j] = s;
}

constexpr int kStageNT = 256, kStageRPT = 4;

}  // namespace

size_t split_f32r_workspace(int n, int m)
{
    if (split_f64_levels(n, split_f64_get().lds_max) == 0) return 0;
    return plan(n, m).bytes;
}

int split_f32r_solve(const SolveArgsF32R& A, int layout, void* ws, size_t ws_bytes,
                     hipStream_t st)
{
    const int n0 = A.n, m0 = A.m;
    const SplitTune tune = split_f64_get();
    const Layout L = plan(n0, m0);
    const int S = L.S;
    if (S == 0 || ((n0 + (1 << S) - 1) >> S) > kStageNT * kStageRPT) return -2;
    if (!ws || ws_bytes < L.bytes) return -3;
    char* w = static_cast<char*>(ws);
    SysState* state = reinterpret_cast<SysState*>(w + L.state);
    int* stop = reinterpret_cast<int*>(w + L.stop);
    double* y0 = reinterpret_cast<double*>(w + L.y0);
    float* A32 = reinterpret_cast<float*>(w + L.a32);
    float* sets = reinterpret_cast<float*>(w + L.sets);
    const long long N = (long long)n0 * m0;
    const int inter = layout == PCRP_LAYOUT_INTERLEAVED;
    const unsigned eblocks = (unsigned)((N + 255) / 256);
    const unsigned sblocks = (unsigned)((m0 + 255) / 256);
    hipLaunchKernelGGL(init_state_kernel, dim3(sblocks), dim3(256), 0, st, state, stop, m0);
    hipLaunchKernelGGL(prep_kernel, dim3(eblocks), dim3(256), 0, st, A.a, A.b, A.c, A.d, A.e,
                       A.y, A32, y0, n0, m0, inter);
    if (hipGetLastError() != hipSuccess) return -1;
    float* set0[6];
    float* set1[6];
    for (int v = 0; v < 6; ++v) {
        set0[v] = sets + (size_t)v * L.per;
        set1[v] = sets + (size_t)(6 + v) * L.per;
    }
    for (int it = 0; it <= A.max_sweeps; ++it) {
        const float* in[6];
        for (int v = 0; v < 6; ++v) in[v] = A32 + (size_t)v * N;
        long long in_rs = A.rs, in_ss = A.ss;
        int Sin = 0, Qin = m0;
        float* const* dst = set0;
        while (Sin < S) {
            const int k = (S - Sin < tune.k_fuse) ? S - Sin : tune.k_fuse;
            const int rc = launch_pass_f32(tune.shape, in, dst, n0, m0, Sin, k, in_rs, in_ss, Qin,
                                           stop, st);
            if (rc != 0) return rc;
            Sin += k;
            Qin <<= k;
            for (int v = 0; v < 6; ++v) in[v] = dst[v];
            in_rs = Qin;
            in_ss = 1;
            dst = (dst == set0) ? set1 : set0;
        }
        hipLaunchKernelGGL((stage_f32_kernel<kStageNT, kStageRPT>), dim3((unsigned)Qin),
                           dim3(kStageNT), 0, st, in[0], in[1], in[2], in[3], in[4], in[5], Qin,
                           A.steps - S, S, n0, m0, A.rs, A.ss, A.y, state, stop, it == 0 ? 1 : 0);
        hipLaunchKernelGGL(decide_kernel, dim3(sblocks), dim3(256), 0, st, state, stop, m0,
                           A.max_sweeps, A.tol, A.sweeps_out);
        if (it < A.max_sweeps)
            hipLaunchKernelGGL(residual_kernel, dim3(eblocks), dim3(256), 0, st, A.a, A.b, A.c,
                               A.d, A.e, y0, A.y, A32 + 5 * N, stop, n0, m0, inter);
        if (hipGetLastError() != hipSuccess) return -1;
    }
    return 0;
}

}  // namespace pcrp
