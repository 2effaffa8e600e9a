// Factor once, solve many: plain FP64 PCR-Penta with the matrix part of
// every reduction step computed once and kept.
//
// The reference recomputes every multiplier on every solve: alfa1, alfa2
// (kernel.cu:72-75), beta1, beta2 (kernel.cu:98-103) and gamma1, gamma2
// (kernel.cu:122, 129). All six depend on the matrix only. A batch that
// solves the same matrices for many right-hand sides (the Heston driver's 200
// time steps with constant line matrices; every v-line of a Heston step even
// shares one matrix) can keep them: per step s and row (or centre) i
//   al_i, ar_i : multipliers of the pair equation centred on i (alfa1/alfa2
//                of row i + d = beta1/beta2 of row i - d, kernel.cu:73,75,99,101)
//   g1_i, g2_i : gamma1/gamma2 of row i after the zero-pivot rule and the
//                existence guards (kernel.cu:120-133)
// and after the last step the reduced c_i, b_i, d_i of the final 2x2 / 1x1
// solve (kernel.cu:145-153). A right-hand side then needs, per step,
//   K_j  = y_j + al_j y_{j-d} + ar_j y_{j+d}   (kernel.cu:79 / 105: k1, k2)
//   y_i' = g1_i K_{i-d} + g2_i K_{i+d} - y_i  (kernel.cu:140)
// with the operands and the FMA placement of the full solve (make_pair's K
// and reduce_row's y line), so the result equals pcrp_solve_f64 bit for bit
// (signed zeros aside): 4 FMA per row and step instead of ~20 FMA and 6 IEEE
// divisions.
//
// Factor layout, one block per system (or a single block shared by every
// right-hand side): for s = 0 .. steps-1 the arrays al, ar, g1, g2 (n doubles
// each), then c, b, d: (4 steps + 3) n doubles, row-contiguous, so a thread's
// loads for its rows are coalesced across the workgroup.
#include "pcr_penta.h"
#include "pcrp_device.h"
#include "pcrp_internal.h"

namespace pcrp {

long long factor_f64_words(int n)
{
    return (4LL * pcr_steps(n) + 3) * n;
}

namespace {

constexpr int kFactorNT = 1024;   // one row per thread: n <= 1024

// One workgroup per system: the shared pair-equation form of the full solve
// (rows and pair equations in LDS), with every multiplier written out.
__global__ __launch_bounds__(kFactorNT) void factor_f64_kernel(
    const double* __restrict__ ga, const double* __restrict__ gb, const double* __restrict__ gc,
    const double* __restrict__ gd, const double* __restrict__ ge, int n, long long rs,
    long long ss, int steps, double* __restrict__ F)
{
    __shared__ double sR[6][kFactorNT];
    __shared__ double sE[8][kFactorNT];
    const int sys = blockIdx.x, i = threadIdx.x;
    const long long base = (long long)sys * ss;
    double* const Fs = F + (long long)sys * (4LL * steps + 3) * n;
    Row<double> r{0.0, 0.0, 1.0, 0.0, 0.0, 0.0};   // the right-hand side plays no part
    if (i < n) {
        const long long g = base + (long long)i * rs;
        r.a = ga[g]; r.b = gb[g]; r.c = gc[g]; r.d = gd[g]; r.e = ge[g];
    }
    auto readR = [&](int k) {
        Row<double> q;
        q.a = sR[0][k]; q.b = sR[1][k]; q.c = sR[2][k];
        q.d = sR[3][k]; q.e = sR[4][k]; q.y = sR[5][k];
        return q;
    };
    int delta = 1;
    for (int s = 0; s < steps; ++s, delta <<= 1) {
        if (i < n) {
            sR[0][i] = r.a; sR[1][i] = r.b; sR[2][i] = r.c;
            sR[3][i] = r.d; sR[4][i] = r.e; sR[5][i] = r.y;
        }
        __syncthreads();
        if (i < n) {
            // pair equation centred on i (kernel.cu:69-117), as the full solve builds it
            const bool hasL = i - delta >= 0, hasR = i + delta <= n - 1;
            const bool aC = i - 2 * delta >= 0, arC = i + 2 * delta <= n - 1;
            const Row<double> L = readR(hasL ? i - delta : i);
            const Row<double> R = readR(hasR ? i + delta : i);
            const Pair<double> E = make_pair(L, r, R, hasL, aC, hasR, arC);
            sE[0][i] = E.A; sE[1][i] = E.B; sE[2][i] = E.C; sE[3][i] = E.D;
            sE[4][i] = E.E; sE[5][i] = E.K; sE[6][i] = E.al; sE[7][i] = E.ar;
            Fs[(4LL * s + 0) * n + i] = E.al;
            Fs[(4LL * s + 1) * n + i] = E.ar;
        }
        __syncthreads();
        if (i < n) {
            // Reduction-III (kernel.cu:120-140)
            const bool hp1 = i - delta >= 0, hq1 = i + delta <= n - 1;
            const bool hp2 = i - 2 * delta >= 0, hq2 = i + 2 * delta <= n - 1;
            const int il = hp1 ? i - delta : i, ir = hq1 ? i + delta : i;
            Pair<double> EL, ER;
            EL.A = sE[0][il]; EL.B = sE[1][il]; EL.C = sE[2][il];
            EL.D = sE[3][il]; EL.E = sE[4][il]; EL.K = sE[5][il];
            ER.A = sE[0][ir]; ER.B = sE[1][ir]; ER.C = sE[2][ir];
            ER.D = sE[3][ir]; ER.E = sE[4][ir]; ER.K = sE[5][ir];
            double gg[2];
            r = reduce_row(r, EL, ER, hp1, hq1, hp2, hq2,
                           [&](double& alL, double& arL, double& alR, double& arR) {
                               alL = sE[6][il]; arL = sE[7][il];
                               alR = sE[6][ir]; arR = sE[7][ir];
                           },
                           gg);
            Fs[(4LL * s + 2) * n + i] = gg[0];
            Fs[(4LL * s + 3) * n + i] = gg[1];
        }
        __syncthreads();   // sR and sE are rewritten by the next step
    }
    if (i < n) {
        Fs[(4LL * steps + 0) * n + i] = r.c;
        Fs[(4LL * steps + 1) * n + i] = r.b;
        Fs[(4LL * steps + 2) * n + i] = r.d;
    }
}

// Right-hand sides only. NT threads x RPT rows (row i = t + NT*k) of G
// right-hand sides per workgroup (systems j0 .. j0+G-1; G > 1 only with a
// shared factor, whose loads then serve G systems). fstride: doubles between
// two systems' factors (0: one factor for all). x goes to out (y when out is
// y) at the element of y, plus add[] there when add is given.
template <int NT, int RPT, int G>
__global__ __launch_bounds__(NT) void solve_factored_f64_kernel(
    const double* __restrict__ F, long long fstride, const double* gy, double* gx,
    const double* __restrict__ add, int n, int m, long long rs, long long ss, int steps)
{
    constexpr int NMAX = NT * RPT;
    __shared__ double sy[G][NMAX];
    __shared__ double sK[G][NMAX];
    const int t = threadIdx.x;
    const int j0 = blockIdx.x * G;
    const double* __restrict__ Fs = F + (long long)j0 * fstride;
    double y[RPT][G];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int j = j0 + g;
            y[k][g] = (i < n && j < m) ? gy[(long long)j * ss + (long long)i * rs] : 0.0;
        }
    }
    int delta = 1;
    for (int s = 0; s < steps; ++s, delta <<= 1) {
        const double* __restrict__ Fal = Fs + (4LL * s) * n;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int i = t + NT * k;
            if (i < n)
#pragma unroll
                for (int g = 0; g < G; ++g) sy[g][i] = y[k][g];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int i = t + NT * k;
            if (i < n) {
                const double al = Fal[i], ar = Fal[n + i];
                const int il = i - delta >= 0 ? i - delta : i;
                const int ir = i + delta <= n - 1 ? i + delta : i;
#pragma unroll
                for (int g = 0; g < G; ++g)
                    sK[g][i] = __builtin_fma(ar, sy[g][ir], __builtin_fma(al, sy[g][il], y[k][g]));
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int i = t + NT * k;
            if (i < n) {
                const double g1 = Fal[2 * n + i], g2 = Fal[3 * n + i];
                const int il = i - delta >= 0 ? i - delta : i;
                const int ir = i + delta <= n - 1 ? i + delta : i;
#pragma unroll
                for (int g = 0; g < G; ++g)
                    y[k][g] = __builtin_fma(g2, sK[g][ir], __builtin_fma(g1, sK[g][il], -y[k][g]));
            }
        }
        // sy is next written after this step's reads (behind the barrier
        // above), sK after the next step's first barrier
    }
    // final 2x2 / 1x1 solve (kernel.cu:145-153) at delta = 2^steps
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        if (i < n)
#pragma unroll
            for (int g = 0; g < G; ++g) sy[g][i] = y[k][g];
    }
    __syncthreads();
    const double* __restrict__ Fc = Fs + (4LL * steps) * n;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        if (i >= n) continue;
        const double c = Fc[i];
        if (i < n - delta) {
            const double dq = Fc[2 * n + i];
            const double cii = Fc[i + delta], bii = Fc[n + i + delta];
            const double deno = __builtin_fma(c, cii, -(bii * dq));
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int j = j0 + g;
                if (j >= m) continue;
                const double yi = y[k][g], yii = sy[g][i + delta];
                const long long e0 = (long long)j * ss + (long long)i * rs;
                const long long e1 = e0 + (long long)delta * rs;
                const double x0 = __builtin_fma(cii, yi, -(dq * yii)) / deno;
                const double x1 = __builtin_fma(c, yii, -(bii * yi)) / deno;
                gx[e0] = add ? x0 + add[e0] : x0;
                gx[e1] = add ? x1 + add[e1] : x1;
            }
        } else if (i < delta) {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const int j = j0 + g;
                if (j >= m) continue;
                const long long e0 = (long long)j * ss + (long long)i * rs;
                const double x0 = y[k][g] / c;
                gx[e0] = add ? x0 + add[e0] : x0;
            }
        }
    }
}

template <int NT, int RPT, int G>
int launch_solve(const FactoredF64& A, hipStream_t st)
{
    const int blocks = (A.m + G - 1) / G;
    hipLaunchKernelGGL((solve_factored_f64_kernel<NT, RPT, G>), dim3(blocks), dim3(NT), 0, st,
                       A.F, A.shared ? 0LL : factor_f64_words(A.n), A.y, A.out ? A.out : A.y,
                       A.add, A.n, A.m, A.rs, A.ss, pcr_steps(A.n));
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NT, int RPT>
int launch_solve_g(const FactoredF64& A, hipStream_t st)
{
    // a shared factor serves up to four right-hand sides per workgroup while
    // the batch still fills the CUs
    if (A.shared && A.m >= 4 * 256) return launch_solve<NT, RPT, 4>(A, st);
    if (A.shared && A.m >= 2 * 256) return launch_solve<NT, RPT, 2>(A, st);
    return launch_solve<NT, RPT, 1>(A, st);
}

}  // namespace

int factor_f64(const double* a, const double* b, const double* c, const double* d,
               const double* e, int n, int msys, long long rs, long long ss, double* F,
               hipStream_t st)
{
    if (n < 1 || n > kFactorNT) return -2;
    hipLaunchKernelGGL(factor_f64_kernel, dim3(msys), dim3(kFactorNT), 0, st, a, b, c, d, e, n, rs,
                       ss, pcr_steps(n), F);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int solve_factored_f64(const FactoredF64& A, hipStream_t st)
{
    const int n = A.n;
    if (n <= 64) return launch_solve_g<64, 1>(A, st);
    if (n <= 128) return launch_solve_g<128, 1>(A, st);
    if (n <= 256) return launch_solve_g<256, 1>(A, st);
    if (n <= 512) return launch_solve_g<256, 2>(A, st);
    if (n <= 1024) return launch_solve_g<256, 4>(A, st);
    return -2;
}

}  // namespace pcrp

namespace {

bool strides_of(int n, int m, int layout, long long* rs, long long* ss)
{
    if (layout == PCRP_LAYOUT_CONTIGUOUS) { *rs = 1; *ss = n; return true; }
    if (layout == PCRP_LAYOUT_INTERLEAVED) { *rs = m; *ss = 1; return true; }
    return false;
}

}  // namespace

extern "C" {

size_t pcrp_factor_size(int n, int m, int precision, int shared)
{
    if (n < 1 || m < 1 || (long long)n * m > 0x7fffffffLL || (shared != 0 && shared != 1))
        return 0;
    const long long msys = shared ? 1 : m;
    if (precision == PCRP_PREC_F64)
        return n <= pcrp::lds_f64_max_n() ? (size_t)(msys * pcrp::factor_f64_words(n)) * sizeof(double)
                                          : 0;
    if (precision == PCRP_PREC_F32_REFINE)
        return n <= pcrp::lds_f32r_max_n() ? (size_t)msys * pcrp::factor_f32r_bytes(n) : 0;
    return 0;
}

int pcrp_factor_f64(const double* a, const double* b, const double* c, const double* d,
                    const double* e, int n, int m, int layout, int shared, void* factor,
                    size_t factor_bytes, void* stream)
{
    if (n < 1 || m < 1 || (long long)n * m > 0x7fffffffLL) return PCRP_ERR_INVALID_SIZE;
    if (!a || !b || !c || !d || !e || !factor) return PCRP_ERR_NULL;
    if (shared != 0 && shared != 1) return PCRP_ERR_ARG;
    if (n > pcrp::lds_f64_max_n()) return PCRP_ERR_UNSUPPORTED;
    long long rs = 1, ss = 0;
    if (!shared && !strides_of(n, m, layout, &rs, &ss)) return PCRP_ERR_LAYOUT;
    if (factor_bytes < pcrp_factor_size(n, m, PCRP_PREC_F64, shared)) return PCRP_ERR_WORKSPACE;
    return pcrp::factor_f64(a, b, c, d, e, n, shared ? 1 : m, rs, ss,
                            static_cast<double*>(factor), static_cast<hipStream_t>(stream)) == 0
               ? PCRP_OK
               : PCRP_ERR_HIP;
}

int pcrp_solve_factored_f64(const void* factor, double* y, int n, int m, int layout, int shared,
                            void* stream)
{
    if (n < 1 || m < 1 || (long long)n * m > 0x7fffffffLL) return PCRP_ERR_INVALID_SIZE;
    if (!factor || !y) return PCRP_ERR_NULL;
    if (shared != 0 && shared != 1) return PCRP_ERR_ARG;
    if (n > pcrp::lds_f64_max_n()) return PCRP_ERR_UNSUPPORTED;
    pcrp::FactoredF64 A{static_cast<const double*>(factor), shared != 0, y, nullptr, nullptr,
                        n, m, 0, 0};
    if (!strides_of(n, m, layout, &A.rs, &A.ss)) return PCRP_ERR_LAYOUT;
    return pcrp::solve_factored_f64(A, static_cast<hipStream_t>(stream)) == 0 ? PCRP_OK
                                                                              : PCRP_ERR_HIP;
}

int pcrp_factor_f32r(const double* a, const double* b, const double* c, const double* d,
                     const double* e, int n, int m, int layout, int shared, void* factor,
                     size_t factor_bytes, void* stream)
{
    if (n < 1 || m < 1 || (long long)n * m > 0x7fffffffLL) return PCRP_ERR_INVALID_SIZE;
    if (!a || !b || !c || !d || !e || !factor) return PCRP_ERR_NULL;
    if (shared != 0 && shared != 1) return PCRP_ERR_ARG;
    if (n > pcrp::lds_f32r_max_n()) return PCRP_ERR_UNSUPPORTED;
    long long rs = 1, ss = 0;
    if (!shared && !strides_of(n, m, layout, &rs, &ss)) return PCRP_ERR_LAYOUT;
    if (factor_bytes < pcrp_factor_size(n, m, PCRP_PREC_F32_REFINE, shared))
        return PCRP_ERR_WORKSPACE;
    pcrp::SolveArgsF32R A{a, b, c, d, e, nullptr, n, m, rs, ss, pcrp::pcr_steps(n), 0, 0.0,
                          nullptr};
    A.shared = shared != 0;
    A.ars = 1;
    return pcrp::lds_f32r_factor(A, factor, static_cast<hipStream_t>(stream)) == 0 ? PCRP_OK
                                                                                   : PCRP_ERR_HIP;
}

int pcrp_solve_factored_f32r_f64(const void* factor, const double* a, const double* b,
                                 const double* c, const double* d, const double* e, double* y,
                                 int n, int m, int layout, int shared, int max_sweeps, double tol,
                                 int* sweeps_out, void* stream)
{
    if (n < 1 || m < 1 || (long long)n * m > 0x7fffffffLL) return PCRP_ERR_INVALID_SIZE;
    if (!factor || !a || !b || !c || !d || !e || !y) return PCRP_ERR_NULL;
    if ((shared != 0 && shared != 1) || max_sweeps < 0 || !(tol >= 0.0)) return PCRP_ERR_ARG;
    if (n > pcrp::lds_f32r_max_n()) return PCRP_ERR_UNSUPPORTED;
    pcrp::SolveArgsF32R A{a, b, c, d, e, y, n, m, 0, 0, pcrp::pcr_steps(n), max_sweeps, tol,
                          sweeps_out};
    if (!strides_of(n, m, layout, &A.rs, &A.ss)) return PCRP_ERR_LAYOUT;
    A.shared = shared != 0;
    A.ars = 1;
    A.fct = factor;
    return pcrp::lds_f32r_solve(A, static_cast<hipStream_t>(stream)) == 0 ? PCRP_OK
                                                                          : PCRP_ERR_HIP;
}

}  // extern "C"
