// LDS-resident FP32 PCR-Penta with FP64 iterative refinement, one workgroup
// per system (n <= 2048). No counterpart in the reference, which is FP64-only
// (kernel.cu:32-154); the reduction steps follow the same recurrences.
//
//   1. factor: PCR-Penta on the FP32-rounded matrix. Every multiplier of
//      every level (alfa, alfa' of the pair equation centred on the row,
//      g1, g2 of the row) stays in VGPRs, so later solves touch only the
//      right-hand side;
//   2. x = 0, r = y; repeat: dx = PCR32(r) with the cached multipliers,
//      x += dx (FP64), r = y - A x (FP64, A and y exact), until the
//      relative error left is estimated below tol (see the stopping rule
//      below) or max_sweeps refinement sweeps have run.
// A zero (or below 2^-100) FP32 pivot gives a zero multiplier (the FP64 path
// keeps the reference's rule, kernel.cu:122,129): the refinement corrects it.
//
// Reciprocals are approximate (v_rcp_f32): the refinement absorbs the error.
#include "pcrp_device.h"
#include "pcrp_internal.h"

namespace pcrp {

namespace {

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Off-diagonal couplings of a dominant system shrink doubly exponentially
// with the level (squaring per step), so deep-level FP32 pivots reach the
// denormal range, where v_rcp_f32 returns infinity. Pivots below 2^-100 are
// treated as zero: the coupling they would eliminate is below FP32
// resolution, and the FP64 refinement corrects what is left.
constexpr float kTiny = 0x1p-100f;

struct RowF { float a, b, c, d, e; };
struct PairF { float A, B, C, D, E, rC; };   // rC = 1/C (approximate)

template <int NT>
__device__ __forceinline__ double block_max(double v, double* red)
{
    // wave64 reduction then across waves through LDS
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double r = red[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) r = fmax(r, red[w]);
    __syncthreads();
    return r;
}

}  // namespace

template <int NT, int RPT, int LMAX>
__global__ __launch_bounds__(NT) void lds_f32r_kernel(
    const double* __restrict__ ga, const double* __restrict__ gb, const double* __restrict__ gc,
    const double* __restrict__ gd, const double* __restrict__ ge, double* __restrict__ gy,
    int n, int m, long long rs, long long ss, int steps, int max_sweeps, double tol,
    int* __restrict__ sweeps_out)
{
    constexpr int NMAX = NT * RPT;
    __shared__ double sA[5][NMAX];     // exact FP64 matrix (residual)
    __shared__ double sX[NMAX];        // FP64 iterate x (residual neighbours)
    __shared__ float sF[6][NMAX];      // factor: rows, then pair equations; sweeps: y, K, dx
    __shared__ double sRed[2][NT / 64];

    const int t = threadIdx.x;
    const int sys = xcd_remap(blockIdx.x, m);
    const long long base = (long long)sys * ss;

    double Y[RPT], X[RPT];
    RowF r[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        double a = 0.0, b = 0.0, c = 1.0, d = 0.0, e = 0.0, y = 0.0;
        if (i < n) {
            const long long g = base + (long long)i * rs;
            a = ga[g]; b = gb[g]; c = gc[g]; d = gd[g]; e = ge[g]; y = gy[g];
            // out-of-range couplings are not part of the matrix
            if (i < 2) a = 0.0;
            if (i < 1) b = 0.0;
            if (i > n - 2) d = 0.0;
            if (i > n - 3) e = 0.0;
            sA[0][i] = a; sA[1][i] = b; sA[2][i] = c; sA[3][i] = d; sA[4][i] = e;
            sF[0][i] = (float)a; sF[1][i] = (float)b; sF[2][i] = (float)c;
            sF[3][i] = (float)d; sF[4][i] = (float)e;
        }
        Y[k] = y;
        X[k] = 0.0;
        r[k] = RowF{(float)a, (float)b, (float)c, (float)d, (float)e};
    }

    // cached multipliers per level: al, ar (pair equation on this row), g1, g2
    float Mal[LMAX][RPT], Mar[LMAX][RPT], Mg1[LMAX][RPT], Mg2[LMAX][RPT];
    // final solve: x_i = F1*y_i + F2*y_{i+D}; partner x_{i+D} = F3*y_{i+D} + F4*y_i
    float F1[RPT], F2[RPT], F3[RPT], F4[RPT];
    __syncthreads();

    // ---------------- factor (one LDS buffer, four barriers per level) ----------------
#pragma unroll
    for (int s = 0; s < LMAX; ++s) {
        if (s < steps) {
            const int delta = 1 << s;
            PairF E[RPT];
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int j = t + NT * k;
                float al = 0.f, ar = 0.f;
                E[k] = PairF{0.f, 0.f, 1.f, 0.f, 0.f, 1.f};
                if (j < n) {
                    const bool hasL = j - delta >= 0, hasR = j + delta <= n - 1;
                    const bool aC = j - 2 * delta >= 0, arC = j + 2 * delta <= n - 1;
                    const int jl = hasL ? j - delta : j, jr = hasR ? j + delta : j;
                    const float Lb = sF[1][jl], Rd = sF[3][jr];
                    al = (aC && fabsf(Lb) >= kTiny) ? -r[k].a * frcp(Lb) : 0.f;
                    ar = (arC && fabsf(Rd) >= kTiny) ? -r[k].e * frcp(Rd) : 0.f;
                    E[k].A = al * sF[0][jl];
                    E[k].B = hasL ? fmaf(ar, sF[0][jr], fmaf(al, sF[2][jl], r[k].b)) : 0.f;
                    E[k].C = fmaf(ar, sF[1][jr], fmaf(al, sF[3][jl], r[k].c));
                    E[k].D = hasR ? fmaf(ar, sF[2][jr], fmaf(al, sF[4][jl], r[k].d)) : 0.f;
                    E[k].E = ar * sF[4][jr];
                    E[k].rC = fabsf(E[k].C) >= kTiny ? frcp(E[k].C) : 0.f;
                }
                Mal[s][k] = al;
                Mar[s][k] = ar;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int j = t + NT * k;
                if (j < n) {
                    sF[0][j] = E[k].A; sF[1][j] = E[k].B; sF[2][j] = E[k].C;
                    sF[3][j] = E[k].D; sF[4][j] = E[k].E; sF[5][j] = E[k].rC;
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int i = t + NT * k;
                float g1 = 0.f, g2 = 0.f;
                if (i < n) {
                    const bool hp1 = i - delta >= 0, hq1 = i + delta <= n - 1;
                    const int il = hp1 ? i - delta : i, ir = hq1 ? i + delta : i;
                    g1 = hp1 ? r[k].b * sF[5][il] : 0.f;
                    g2 = hq1 ? r[k].d * sF[5][ir] : 0.f;
                    RowF N;
                    N.a = g1 * sF[0][il];
                    N.b = fmaf(g2, sF[0][ir], fmaf(g1, sF[1][il], -r[k].a));
                    N.c = fmaf(g2, sF[1][ir], fmaf(g1, sF[3][il], -r[k].c));
                    N.d = fmaf(g2, sF[3][ir], fmaf(g1, sF[4][il], -r[k].e));
                    N.e = g2 * sF[4][ir];
                    r[k] = N;
                }
                Mg1[s][k] = g1;
                Mg2[s][k] = g2;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int i = t + NT * k;
                if (i < n) {
                    sF[0][i] = r[k].a; sF[1][i] = r[k].b; sF[2][i] = r[k].c;
                    sF[3][i] = r[k].d; sF[4][i] = r[k].e;
                }
            }
            __syncthreads();
        }
    }
    const int D = 1 << steps;   // final delta
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        F1[k] = F2[k] = F3[k] = F4[k] = 0.f;
        if (i < n) {
            if (i < n - D) {
                const float cii = sF[2][i + D], bii = sF[1][i + D];
                const float det = fmaf(r[k].c, cii, -(bii * r[k].d));
                const float rd = fabsf(det) >= kTiny ? frcp(det) : 0.f;
                F1[k] = cii * rd; F2[k] = -r[k].d * rd; F3[k] = r[k].c * rd; F4[k] = -bii * rd;
            } else if (i < D) {
                F1[k] = fabsf(r[k].c) >= kTiny ? frcp(r[k].c) : 0.f;
            }
        }
    }
    __syncthreads();

    // ---------------- solve + refinement sweeps ----------------
    float* sy = sF[0];
    float* sk = sF[1];
    float* sdx = sF[2];
    double res[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) res[k] = Y[k];
    int sweep = 0;
    double dprev = 0.0;
    for (;;) {
        float v[RPT];
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            v[k] = (float)res[k];
            const int i = t + NT * k;
            if (i < n) sy[i] = v[k];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < LMAX; ++s) {
            if (s < steps) {
                const int delta = 1 << s;
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const int j = t + NT * k;
                    if (j < n) {
                        const int jl = j - delta >= 0 ? j - delta : j;
                        const int jr = j + delta <= n - 1 ? j + delta : j;
                        sk[j] = fmaf(Mar[s][k], sy[jr], fmaf(Mal[s][k], sy[jl], v[k]));
                    }
                }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < RPT; ++k) {
                    const int i = t + NT * k;
                    if (i < n) {
                        const int il = i - delta >= 0 ? i - delta : i;
                        const int ir = i + delta <= n - 1 ? i + delta : i;
                        v[k] = fmaf(Mg2[s][k], sk[ir], fmaf(Mg1[s][k], sk[il], -v[k]));
                        sy[i] = v[k];
                    }
                }
                __syncthreads();
            }
        }
        // final 2x2 / 1x1 on the reduced right-hand side
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int i = t + NT * k;
            if (i < n) {
                if (i < n - D) {
                    const float yii = sy[i + D];
                    sdx[i] = fmaf(F1[k], v[k], F2[k] * yii);
                    sdx[i + D] = fmaf(F3[k], yii, F4[k] * v[k]);
                } else if (i < D) {
                    sdx[i] = F1[k] * v[k];
                }
            }
        }
        __syncthreads();
        double dmax = 0.0, xmax = 0.0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int i = t + NT * k;
            if (i < n) {
                const double dx = (double)sdx[i];
                X[k] += dx;
                sX[i] = X[k];
                dmax = fmax(dmax, fabs(dx));
                xmax = fmax(xmax, fabs(X[k]));
            }
        }
        dmax = block_max<NT>(dmax, sRed[0]);   // contains a barrier: sX complete
        xmax = block_max<NT>(xmax, sRed[1]);
        // Stop when the correction is below tol, or when the error left after
        // applying it, estimated from the observed contraction
        // rho = |dx_k| / |dx_{k-1}| as rho * |dx_k| with a safety factor of 10,
        // is below tol (and rho <= 0.1). rho needs two refinement corrections:
        // |dx_1| / |x_0| measures the first solve's error, not the contraction.
        const bool small = !(dmax > tol * xmax);   // also stops on NaN
        const bool contracting = sweep > 1 && dmax <= 0.1 * dprev &&
                                 10.0 * dmax * dmax <= tol * xmax * dprev;
        const bool done = small || contracting;
        dprev = dmax;
        if (done || sweep >= max_sweeps) {
            if (!done) ++sweep;                   // flag: tolerance not met
            break;
        }
        ++sweep;
        // r = y - A x, exact FP64
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int i = t + NT * k;
            if (i < n) {
                const double xm2 = i >= 2 ? sX[i - 2] : 0.0, xm1 = i >= 1 ? sX[i - 1] : 0.0;
                const double xp1 = i + 1 < n ? sX[i + 1] : 0.0, xp2 = i + 2 < n ? sX[i + 2] : 0.0;
                double ax = sA[2][i] * X[k];
                ax = __builtin_fma(sA[0][i], xm2, ax);
                ax = __builtin_fma(sA[1][i], xm1, ax);
                ax = __builtin_fma(sA[3][i], xp1, ax);
                ax = __builtin_fma(sA[4][i], xp2, ax);
                res[k] = Y[k] - ax;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        if (i < n) gy[base + (long long)i * rs] = X[k];
    }
    if (sweeps_out && t == 0) sweeps_out[sys] = sweep;
}

template <int NT, int RPT>
static int launch_f32r(const SolveArgsF32R& A, hipStream_t st)
{
    constexpr int LMAX = (NT * RPT >= 2048) ? 11 : (NT * RPT >= 1024) ? 10 : (NT * RPT >= 512) ? 9
                       : (NT * RPT >= 256) ? 8 : (NT * RPT >= 128) ? 7 : 6;
    static_assert((1 << (LMAX + 1)) > NT * RPT, "LMAX covers floor(log2 NMAX)");
    hipLaunchKernelGGL((lds_f32r_kernel<NT, RPT, LMAX>), dim3(A.m), dim3(NT), 0, st, A.a, A.b,
                       A.c, A.d, A.e, A.y, A.n, A.m, A.rs, A.ss, A.steps, A.max_sweeps, A.tol,
                       A.sweeps_out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int lds_f32r_max_n() { return 2048; }

int lds_f32r_solve(const SolveArgsF32R& A, hipStream_t st)
{
    const int n = A.n;
    if (n <= 64) return launch_f32r<64, 1>(A, st);
    if (n <= 128) return launch_f32r<128, 1>(A, st);
    if (n <= 256) return launch_f32r<128, 2>(A, st);
    if (n <= 512) return launch_f32r<256, 2>(A, st);
    if (n <= 1024) return launch_f32r<256, 4>(A, st);
    if (n <= 2048) return launch_f32r<512, 4>(A, st);
    return -2;
}

}  // namespace pcrp
