// LDS-resident FP32 inner solve with FP64 iterative refinement, one
// workgroup per system (n <= 2048). No counterpart in the reference, which is
// FP64-only (kernel.cu:32-154).
//
// Inner solver: parallel cyclic reduction on the 2x2-block tridiagonal form
// of the pentadiagonal matrix. Rows (2k, 2k+1) form block row k:
//     L_k X_{k-1} + D_k X_k + U_k X_{k+1} = Y_k,   X_k = (x_2k, x_2k+1),
//     L_k = [[a_2k, b_2k], [0, a_2k+1]],  D_k = [[c_2k, d_2k], [b_2k+1, c_2k+1]],
//     U_k = [[e_2k, 0], [d_2k+1, e_2k+1]].
// One level at block stride s: alpha = -L_k D_{k-s}^-1, beta = -U_k D_{k+s}^-1,
//     D_k += alpha U_{k-s} + beta L_{k+s},  L_k = alpha L_{k-s},
//     U_k = beta U_{k+s},                   Y_k += alpha Y_{k-s} + beta Y_{k+s},
// for s = 1, 2, 4, ... until every coupling has left the system, then
// X_k = D_k^-1 Y_k. The pivots are the diagonal blocks, which stay dominant
// on a dominant matrix. PCR-Penta's own recurrences divide by off-diagonals
// (-a/b, -e/d, kernel.cu:73,75,99,101); on random dominant matrices some of
// those are 1e-6..1e-8, and in FP32 that leaves 1e-5..1e-2 relative error
// (measured on the oracle's generators up to n = 2^20), so refinement needs
// many sweeps or diverges. The block form stays at ~3e-7 at every size, so
// the refinement converges in two corrections.
//
//   1. factor: every level's alpha, beta (8 floats per block row and level)
//      and the final D^-1 stay in VGPRs, so the sweeps touch only the
//      right-hand side;
//   2. x = 0, r = y; repeat: dx = solve32(r), x += dx (FP64),
//      r = y - A x (FP64, A and y exact), until the error left is estimated
//      below tol (stopping rule below) or max_sweeps corrections have run.
// Reciprocals are approximate (v_rcp_f32): the refinement absorbs the error.
// A block determinant below 2^-100 gives a zero inverse (the refinement then
// flags the system if it cannot converge).
#include "pcrp_device.h"
#include "pcrp_internal.h"

namespace pcrp {

namespace {

// Workgroup max of u and of v (wave64 shuffles, then across waves through
// LDS, one barrier). red is rewritten only after the caller's next barrier.
template <int NT>
__device__ __forceinline__ void block_max2(double& u, double& v, double (&red)[2][NT / 64])
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        u = fmax(u, __shfl_xor(u, off, 64));
        v = fmax(v, __shfl_xor(v, off, 64));
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        red[0][wid] = u;
        red[1][wid] = v;
    }
    __syncthreads();
    u = red[0][0];
    v = red[1][0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) {
        u = fmax(u, red[0][w]);
        v = fmax(v, red[1][w]);
    }
}

}  // namespace

// NT threads, BR block rows per thread (block row k = t + NT*q), LMAX levels
// (>= ceil(log2(NT*BR))).
template <int NT, int BR, int LMAX>
__global__ __launch_bounds__(NT) void lds_f32r_kernel(
    const double* __restrict__ ga, const double* __restrict__ gb, const double* __restrict__ gc,
    const double* __restrict__ gd, const double* __restrict__ ge, double* __restrict__ gy,
    int n, int m, long long rs, long long ss, int max_sweeps, double tol,
    int* __restrict__ sweeps_out)
{
    constexpr int NB = NT * BR;        // block rows
    constexpr int NMAX = 2 * NB;       // rows
    // exact FP64 matrix (residual); rows in parity-split order, row i at
    // lds_pos(i) = (i & 1) * NB + i / 2, so the lanes of a wave (consecutive
    // block rows k, rows 2k + h) touch consecutive doubles: no bank conflicts
    // a..e and y: the residuals read no global memory (re-reading y there
    // cost 22% at n = 1024, waves waiting on the strided loads)
    __shared__ double sA[6][NMAX];
    // factor: D^-1, L, U of every block row (one 16-byte vector each);
    // sweeps: FP64 x in sF[0..1] and the workgroup-max scratch in sF[2]
    // (aliases: n = 1024 takes exactly 80 KB, two workgroups per CU, and
    // n = 2048 exactly the CU's 160 KB)
    __shared__ float4 sF[3][NB];
    __shared__ float2 sY[2][NB];       // right-hand side exchange, ping-pong
    double* const sX = reinterpret_cast<double*>(&sF[0][0]);
    static_assert(NMAX * sizeof(double) <= 2 * NB * sizeof(float4), "x fits in sF[0..1]");
    static_assert(2 * (NT / 64) * sizeof(double) <= NB * sizeof(float4), "scratch fits in sF[2]");
    auto& sRed = *reinterpret_cast<double (*)[2][NT / 64]>(&sF[2][0]);

    const auto lds_pos = [](int i) { return (i & 1) * NB + (i >> 1); };
    const int t = threadIdx.x;
    const int sys = xcd_remap(blockIdx.x, m);
    const long long base = (long long)sys * ss;
    const int nb = (n + 1) >> 1;

    double X[BR][2];
    float v[BR][2];   // FP32 right-hand side of the next inner solve (first: y)
    M2 L[BR], D[BR], U[BR];
#pragma unroll
    for (int q = 0; q < BR; ++q) {
        const int k = t + NT * q;
        double rw[2][5];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = 2 * k + h;
            double a = 0.0, b = 0.0, c = 1.0, d = 0.0, e = 0.0, y = 0.0;
            if (i < n) {
                const long long g = base + (long long)i * rs;
                a = ga[g]; b = gb[g]; c = gc[g]; d = gd[g]; e = ge[g]; y = gy[g];
                // out-of-range couplings are not part of the matrix
                if (i < 2) a = 0.0;
                if (i < 1) b = 0.0;
                if (i > n - 2) d = 0.0;
                if (i > n - 3) e = 0.0;
                const int p = lds_pos(i);
                sA[0][p] = a; sA[1][p] = b; sA[2][p] = c; sA[3][p] = d; sA[4][p] = e;
                sA[5][p] = y;
            }
            rw[h][0] = a; rw[h][1] = b; rw[h][2] = c; rw[h][3] = d; rw[h][4] = e;
            X[q][h] = 0.0;
            v[q][h] = (float)y;
        }
        L[q] = M2{(float)rw[0][0], (float)rw[0][1], 0.f, (float)rw[1][0]};
        D[q] = M2{(float)rw[0][2], (float)rw[0][3], (float)rw[1][1], (float)rw[1][2]};
        U[q] = M2{(float)rw[0][4], 0.f, (float)rw[1][3], (float)rw[1][4]};
    }

    // ---------------- factor: two barriers per level ----------------
    M2 cA[LMAX][BR], cB[LMAX][BR];     // alpha, beta per level
    int levels = 0;
    while ((1 << levels) < nb) ++levels;
#pragma unroll
    for (int s = 0; s < LMAX; ++s) {
        if (s < levels) {
            const int delta = 1 << s;
#pragma unroll
            for (int q = 0; q < BR; ++q) {
                const int k = t + NT * q;
                sF[0][k] = f4(minv(D[q]));
                sF[1][k] = f4(L[q]);
                sF[2][k] = f4(U[q]);
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < BR; ++q) {
                const int k = t + NT * q;
                const bool hl = k - delta >= 0, hr = k + delta < nb;
                const int kl = hl ? k - delta : k, kr = hr ? k + delta : k;
                const M2 DiL = m2(sF[0][kl]), LL = m2(sF[1][kl]), UL = m2(sF[2][kl]);
                const M2 DiR = m2(sF[0][kr]), LR = m2(sF[1][kr]), UR = m2(sF[2][kr]);
                const M2 al = hl ? mmul_neg(L[q], DiL) : kZero2;
                const M2 be = hr ? mmul_neg(U[q], DiR) : kZero2;
                D[q] = madd2(D[q], al, UL, be, LR);
                L[q] = mmul(al, LL);
                U[q] = mmul(be, UR);
                cA[s][q] = al;
                cB[s][q] = be;
            }
            __syncthreads();
        }
    }
    M2 Fin[BR];
#pragma unroll
    for (int q = 0; q < BR; ++q) Fin[q] = minv(D[q]);

    // ---------------- solve + refinement sweeps ----------------
    int sweep = 0;
    double dprev = 0.0;
    for (;;) {
#pragma unroll
        for (int s = 0; s < LMAX; ++s) {
            if (s < levels) {
                const int delta = 1 << s, p = s & 1;
#pragma unroll
                for (int q = 0; q < BR; ++q) {
                    const int k = t + NT * q;
                    sY[p][k] = make_float2(v[q][0], v[q][1]);
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < BR; ++q) {
                    const int k = t + NT * q;
                    const int kl = k - delta >= 0 ? k - delta : k;
                    const int kr = k + delta < nb ? k + delta : k;
                    const float2 yl = sY[p][kl], yr = sY[p][kr];
                    const float l0 = yl.x, l1 = yl.y, r0 = yr.x, r1 = yr.y;
                    const M2& al = cA[s][q];
                    const M2& be = cB[s][q];
                    v[q][0] = fmaf(be.p, r0, fmaf(be.q, r1, fmaf(al.p, l0, fmaf(al.q, l1, v[q][0]))));
                    v[q][1] = fmaf(be.r, r0, fmaf(be.s, r1, fmaf(al.r, l0, fmaf(al.s, l1, v[q][1]))));
                }
            }
        }
        double dmax = 0.0, xmax = 0.0;
#pragma unroll
        for (int q = 0; q < BR; ++q) {
            const int k = t + NT * q;
            const float dx0 = fmaf(Fin[q].p, v[q][0], Fin[q].q * v[q][1]);
            const float dx1 = fmaf(Fin[q].r, v[q][0], Fin[q].s * v[q][1]);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int i = 2 * k + h;
                if (i < n) {
                    const double dx = (double)(h == 0 ? dx0 : dx1);
                    X[q][h] += dx;
                    // a NaN correction or iterate never passes the stopping
                    // rule (fmax would drop it): the system is flagged
                    const bool nan = !(dx == dx) || !(X[q][h] == X[q][h]);
                    dmax = fmax(dmax, nan ? INFINITY : fabs(dx));
                    xmax = fmax(xmax, fabs(X[q][h]));
                }
            }
        }
        // (the next write to sRed follows at least the residual's barrier)
        block_max2<NT>(dmax, xmax, sRed);
        // Stop when the correction is below tol, or when the error left after
        // applying it, estimated from the observed contraction
        // rho = |dx_k| / |dx_{k-1}| as rho * |dx_k| with a safety factor of 10,
        // is below tol (and rho <= 0.1). rho needs two refinement corrections:
        // |dx_1| / |x_0| measures the first solve's error, not the contraction.
        const bool small = dmax <= tol * xmax;
        const bool contracting = sweep > 1 && dmax <= 0.1 * dprev &&
                                 10.0 * dmax * dmax <= tol * xmax * dprev;
        const bool done = small || contracting;
        dprev = dmax;
        if (done || sweep >= max_sweeps) {
            if (!done) ++sweep;                   // flag: tolerance not met
            break;
        }
        ++sweep;
        // r = y - A x, exact FP64 (y is overwritten only at the end)
#pragma unroll
        for (int q = 0; q < BR; ++q) {
            const int k = t + NT * q;
#pragma unroll
            for (int h = 0; h < 2; ++h)
                if (2 * k + h < n) sX[h * NB + k] = X[q][h];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < BR; ++q) {
            const int k = t + NT * q;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int i = 2 * k + h;
                if (i < n) {
                    const double xm2 = i >= 2 ? sX[lds_pos(i - 2)] : 0.0;
                    const double xm1 = i >= 1 ? sX[lds_pos(i - 1)] : 0.0;
                    const double xp1 = i + 1 < n ? sX[lds_pos(i + 1)] : 0.0;
                    const double xp2 = i + 2 < n ? sX[lds_pos(i + 2)] : 0.0;
                    const int p = lds_pos(i);
                    double ax = sA[2][p] * X[q][h];
                    ax = __builtin_fma(sA[0][p], xm2, ax);
                    ax = __builtin_fma(sA[1][p], xm1, ax);
                    ax = __builtin_fma(sA[3][p], xp1, ax);
                    ax = __builtin_fma(sA[4][p], xp2, ax);
                    v[q][h] = (float)(sA[5][p] - ax);
                }
            }
        }
        // no barrier here: sX is next written by the next residual, behind the
        // next sweep's barriers, and the sweep writes only sY
    }
#pragma unroll
    for (int q = 0; q < BR; ++q) {
        const int k = t + NT * q;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = 2 * k + h;
            if (i < n) gy[base + (long long)i * rs] = X[q][h];
        }
    }
    if (sweeps_out && t == 0) sweeps_out[sys] = sweep;
}

template <int NT, int BR>
static int launch_f32r(const SolveArgsF32R& A, hipStream_t st)
{
    constexpr int LMAX = (NT * BR >= 1024) ? 10 : (NT * BR >= 512) ? 9 : (NT * BR >= 256) ? 8
                       : (NT * BR >= 128) ? 7 : 6;
    static_assert((1 << LMAX) >= NT * BR, "LMAX covers ceil(log2 block rows)");
    hipLaunchKernelGGL((lds_f32r_kernel<NT, BR, LMAX>), dim3(A.m), dim3(NT), 0, st, A.a, A.b,
                       A.c, A.d, A.e, A.y, A.n, A.m, A.rs, A.ss, A.max_sweeps, A.tol,
                       A.sweeps_out);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int lds_f32r_max_n() { return 2048; }

int lds_f32r_solve(const SolveArgsF32R& A, hipStream_t st)
{
    const int n = A.n;
    if (n <= 128) return launch_f32r<64, 1>(A, st);
    if (n <= 256) return launch_f32r<128, 1>(A, st);
    // From n = 257 on, two block rows per thread: the cached multipliers of
    // every level stay in registers without spilling (one block row per
    // thread at 512 threads needed a 128-register cap for two workgroups per
    // CU and spilled: 11% slower at n = 1024), and at n = 512 the smaller
    // workgroups are 8% faster on 1024 systems (equal on 16384).
    if (n <= 512) return launch_f32r<128, 2>(A, st);
    if (n <= 1024) return launch_f32r<256, 2>(A, st);
    if (n <= 2048) return launch_f32r<512, 2>(A, st);
    return -2;
}

}  // namespace pcrp
