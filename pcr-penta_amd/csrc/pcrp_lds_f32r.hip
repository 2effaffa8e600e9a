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
#include "pcrp_heston_ops.h"
#include "pcrp_internal.h"

namespace pcrp {

namespace {

// Max of u and of v over the NT threads of one system (wave64 shuffles, then
// across its waves through LDS, one workgroup barrier); t: the thread's index
// within its system. red is rewritten only after the caller's next barrier.
template <int NT>
__device__ __forceinline__ void block_max2(double& u, double& v, double (&red)[2][NT / 64], int t)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        u = fmax(u, __shfl_xor(u, off, 64));
        v = fmax(v, __shfl_xor(v, off, 64));
    }
    const int lane = t & 63, wid = t >> 6;
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

// NT threads per system, BR block rows per thread (block row k = t + NT*q),
// LMAX levels (>= ceil(log2(NT*BR))), G systems per workgroup.
// G > 1 (interleaved batches, m a multiple of G, 16-byte aligned arrays):
// the workgroup solves the neighbouring systems j0 .. j0+G-1, whose values of
// one row sit side by side, so it loads them with 16-byte accesses (one per
// two systems) instead of one 8-byte access per system and row: the
// interleaved layout's one-element-per-line gather is what bounds the load
// phase. The G systems then run the same solve on their own LDS regions;
// every barrier is workgroup-wide and a system that has converged idles
// until the last one stops.
// MODE: kSolve (factor, then the sweeps), kFactor (factor only: the
// multipliers, D^-1 of the last level and the level count go to fct, nothing
// else is written), kCached (the factor is read from fct instead of being
// computed; the sweeps and their FP64 residuals are the same). The matrix
// is read at ga..ge[j*ass + i*ars]: ass = 0 gives every system one shared
// matrix (its factor, one block of fct, serves all of them). x goes to gx at
// y's element, plus add[] there when add is not null.
enum { kSolve = 0, kFactor = 1, kCached = 2 };

// Bytes of one system's cached factor for NB block rows and LMAX levels:
// the level count (16 bytes), alpha and beta of every level, D^-1.
__host__ __device__ constexpr long long f32r_factor_bytes_nb(int NB, int LMAX)
{
    return 16 + 16LL * NB * (2 * LMAX + 1);
}

// Src: PlainRhs (y read from gy, x + add[] stored) or HestonStage (G = 1,
// kSolve: y computed from the previous Douglas step's grid, x + W stored).
template <int NT, int BR, int LMAX, int G, int MODE, int MINW = 1, class Src = PlainRhs>
__global__ __launch_bounds__(NT * G, MINW) void lds_f32r_kernel(
    const double* __restrict__ ga, const double* __restrict__ gb, const double* __restrict__ gc,
    const double* __restrict__ gd, const double* __restrict__ ge, const double* gy,
    double* gx, const double* __restrict__ add, int n, int m, long long rs, long long ss,
    long long ars, long long ass, int max_sweeps, double tol, int* __restrict__ sweeps_out,
    float4* __restrict__ fct, Src src)
{
    static_assert(!Src::kFused || (G == 1 && MODE == kSolve), "folded stage: one system, solve");
    constexpr int NB = NT * BR;        // block rows
    constexpr int NMAX = 2 * NB;       // rows
    static_assert(G == 1 || G % 2 == 0, "systems are loaded in pairs");
    // exact FP64 matrix and y (residual); rows in parity-split order, row i at
    // lds_pos(i) = (i & 1) * NB + i / 2, so the lanes of a wave (consecutive
    // block rows k, rows 2k + h) touch consecutive doubles: no bank conflicts.
    // The residuals read no global memory (re-reading y there cost 22% at
    // n = 1024, waves waiting on the strided loads).
    __shared__ double sA[G][6][NMAX];
    // factor: D^-1, L, U of every block row (one 16-byte vector each);
    // sweeps: FP64 x in sF[.][0..1] and the max-reduction scratch in sF[.][2]
    // (aliases: one system of n = 1024 takes exactly 80 KB, of n = 2048 the
    // CU's 160 KB)
    __shared__ float4 sF[G][3][NB];
    __shared__ float2 sY[G][2][NB];    // right-hand side exchange, ping-pong
    static_assert(NMAX * sizeof(double) <= 2 * NB * sizeof(float4), "x fits in sF[.][0..1]");
    static_assert(2 * (NT / 64) * sizeof(double) + sizeof(int) <= NB * sizeof(float4),
                  "scratch and stop flag fit in sF[.][2]");

    const auto lds_pos = [](int i) { return (i & 1) * NB + (i >> 1); };
    const int gs = G == 1 ? 0 : (int)threadIdx.x / NT;   // system slot
    const int t = (int)threadIdx.x - gs * NT;
    const int grp = xcd_remap(blockIdx.x, m / G);
    const int sys = grp * G + gs;
    const long long base = (long long)sys * ss;
    const long long abase = (long long)sys * ass;
    const int nb = (n + 1) >> 1;
    [[maybe_unused]] const auto rhs = [&] {
        if constexpr (Src::kFused) return src.bind(sys);
        else return 0;
    }();
    // this system's factor block (the shared one when ass == 0)
    float4* const Fc = reinterpret_cast<float4*>(reinterpret_cast<char*>(fct) +
                                                 (ass == 0 ? 0LL : (long long)sys) *
                                                     f32r_factor_bytes_nb(NB, LMAX));
    int* const Fc_levels = reinterpret_cast<int*>(Fc);
    float4* const Fc_al = Fc + 1;
    float4* const Fc_be = Fc_al + (long long)LMAX * NB;
    float4* const Fc_fin = Fc_be + (long long)LMAX * NB;
    double(&A)[6][NMAX] = sA[gs];
    float4(&F)[3][NB] = sF[gs];
    float2(&Yx)[2][NB] = sY[gs];
    double* const sX = reinterpret_cast<double*>(&F[0][0]);
    auto& sRed = *reinterpret_cast<double (*)[2][NT / 64]>(&F[2][0]);

    // out-of-range couplings are not part of the matrix
    const auto band = [n](int f, int i, double v) {
        if (f == 0 && i < 2) return 0.0;
        if (f == 1 && i < 1) return 0.0;
        if (f == 3 && i > n - 2) return 0.0;
        if (f == 4 && i > n - 3) return 0.0;
        return v;
    };
    if constexpr (G > 1) {
        // rows 0..n-1 of the six arrays, G neighbouring values at i*rs + j0;
        // consecutive lanes take consecutive LDS positions p (even rows, then
        // odd rows): walking rows i instead put the even and the odd lanes of
        // a wave on the same banks (2.1e5 conflict cycles per v-line launch)
        const long long j0 = (long long)grp * G;
        const double* const arr[6] = {ga, gb, gc, gd, ge, gy};
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            for (int p = (int)threadIdx.x; p < NMAX; p += NT * G) {
                const int i = p < NB ? 2 * p : 2 * (p - NB) + 1;   // lds_pos(i) == p
                if (i >= n) continue;
                if (f < 5 && ass == 0) {   // one shared matrix row for the G systems
                    const double w = band(f, i, arr[f][(long long)i * ars]);
#pragma unroll
                    for (int h = 0; h < G; ++h) sA[h][f][p] = w;
                    continue;
                }
                const double2* src = reinterpret_cast<const double2*>(arr[f] + (long long)i * rs + j0);
#pragma unroll
                for (int h = 0; h < G / 2; ++h) {
                    const double2 w = src[h];
                    sA[2 * h][f][p] = band(f, i, w.x);
                    sA[2 * h + 1][f][p] = band(f, i, w.y);
                }
            }
        }
        __syncthreads();
    }

    double X[BR][2];
    float v[BR][2];   // FP32 right-hand side of the next inner solve (first: y)
    M2 L[BR], D[BR], U[BR];
    // contiguous rows (unit stride, 16-byte aligned): a block row's two rows
    // are one 16-byte load per array
    const auto al16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    const bool vec2 = G == 1 && rs == 1 && ars == 1 && ((base | abase) & 1) == 0 && al16(ga) &&
                      al16(gb) && al16(gc) && al16(gd) && al16(ge) && al16(gy);
#pragma unroll
    for (int q = 0; q < BR; ++q) {
        const int k = t + NT * q;
        double rw[2][5];
        if constexpr (G == 1) {
            if (vec2 && 2 * k + 1 < n) {
                const long long g = abase + 2LL * k;
                const double2 av = *reinterpret_cast<const double2*>(ga + g);
                const double2 bv = *reinterpret_cast<const double2*>(gb + g);
                const double2 cv = *reinterpret_cast<const double2*>(gc + g);
                const double2 dv = *reinterpret_cast<const double2*>(gd + g);
                const double2 ev = *reinterpret_cast<const double2*>(ge + g);
                double2 yv;
                if constexpr (MODE == kFactor) yv = make_double2(0.0, 0.0);
                else if constexpr (Src::kFused)   // W parked in x's elements until the store
                    yv = make_double2(rhs.r1(2 * k, gx + base + 2LL * k),
                                      rhs.r1(2 * k + 1, gx + base + 2LL * k + 1));
                else yv = *reinterpret_cast<const double2*>(gy + base + 2LL * k);
                const double ra[2] = {av.x, av.y}, rb[2] = {bv.x, bv.y}, rc[2] = {cv.x, cv.y};
                const double rd[2] = {dv.x, dv.y}, re[2] = {ev.x, ev.y}, ry[2] = {yv.x, yv.y};
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int i = 2 * k + h;
                    const double a = band(0, i, ra[h]), b = band(1, i, rb[h]), c = rc[h];
                    const double d = band(3, i, rd[h]), e = band(4, i, re[h]), y = ry[h];
                    const int p = lds_pos(i);
                    A[0][p] = a; A[1][p] = b; A[2][p] = c; A[3][p] = d; A[4][p] = e;
                    A[5][p] = y;
                    rw[h][0] = a; rw[h][1] = b; rw[h][2] = c; rw[h][3] = d; rw[h][4] = e;
                    X[q][h] = 0.0;
                    v[q][h] = (float)y;
                }
                L[q] = M2{(float)rw[0][0], (float)rw[0][1], 0.f, (float)rw[1][0]};
                D[q] = M2{(float)rw[0][2], (float)rw[0][3], (float)rw[1][1], (float)rw[1][2]};
                U[q] = M2{(float)rw[0][4], 0.f, (float)rw[1][3], (float)rw[1][4]};
                continue;
            }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int i = 2 * k + h;
            double a = 0.0, b = 0.0, c = 1.0, d = 0.0, e = 0.0, y = 0.0;
            if (i < n) {
                const int p = lds_pos(i);
                if constexpr (G == 1) {
                    const long long g = abase + (long long)i * ars;
                    a = band(0, i, ga[g]); b = band(1, i, gb[g]); c = gc[g];
                    d = band(3, i, gd[g]); e = band(4, i, ge[g]);
                    if constexpr (MODE == kFactor) y = 0.0;
                    else if constexpr (Src::kFused) y = rhs.r1(i, gx + base + (long long)i * rs);
                    else y = gy[base + (long long)i * rs];
                    A[0][p] = a; A[1][p] = b; A[2][p] = c; A[3][p] = d; A[4][p] = e;
                    A[5][p] = y;
                } else {
                    a = A[0][p]; b = A[1][p]; c = A[2][p]; d = A[3][p]; e = A[4][p]; y = A[5][p];
                }
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
    // Truncated once a level's multipliers are all below kTruncF32 (every
    // system of the workgroup); the sweeps then run the same levels.
    M2 cA[LMAX][BR], cB[LMAX][BR];     // alpha, beta per level
    M2 Fin[BR];
    int levels = 0;
    while ((1 << levels) < nb) ++levels;
    if constexpr (MODE == kCached) {
        // the factor of an earlier kFactor launch on the same matrix
        levels = Fc_levels[0];
#pragma unroll
        for (int s = 0; s < LMAX; ++s)
            if (s < levels)
#pragma unroll
                for (int q = 0; q < BR; ++q) {
                    const int k = t + NT * q;
                    cA[s][q] = m2(Fc_al[(long long)s * NB + k]);
                    cB[s][q] = m2(Fc_be[(long long)s * NB + k]);
                }
#pragma unroll
        for (int q = 0; q < BR; ++q) Fin[q] = m2(Fc_fin[t + NT * q]);
    }
#pragma unroll
    for (int s = 0; s < LMAX; ++s) {
        if (MODE != kCached && s < levels) {
            float am = 0.f;   // largest multiplier entry of this thread's block rows
            const int delta = 1 << s;
#pragma unroll
            for (int q = 0; q < BR; ++q) {
                const int k = t + NT * q;
                F[0][k] = f4(minv(D[q]));
                F[1][k] = f4(L[q]);
                F[2][k] = f4(U[q]);
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < BR; ++q) {
                const int k = t + NT * q;
                const bool hl = k - delta >= 0, hr = k + delta < nb;
                const int kl = hl ? k - delta : k, kr = hr ? k + delta : k;
                const M2 DiL = m2(F[0][kl]), LL = m2(F[1][kl]), UL = m2(F[2][kl]);
                const M2 DiR = m2(F[0][kr]), LR = m2(F[1][kr]), UR = m2(F[2][kr]);
                const M2 al = hl ? mmul_neg(L[q], DiL) : kZero2;
                const M2 be = hr ? mmul_neg(U[q], DiR) : kZero2;
                D[q] = madd2(D[q], al, UL, be, LR);
                L[q] = mmul(al, LL);
                U[q] = mmul(be, UR);
                cA[s][q] = al;
                cB[s][q] = be;
                am = fmaxf(am, fmaxf(maxabs2(al), maxabs2(be)));
            }
            // one flag per wave (a ballot, no shuffles) in sY[.][1], idle
            // until the sweeps
            const int big = __any(am > kTruncF32);
            if ((t & 63) == 0) reinterpret_cast<int*>(&Yx[1][0])[t >> 6] = big;
            __syncthreads();
            int any = 0;
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
                for (int w = 0; w < NT / 64; ++w)
                    any |= reinterpret_cast<const int*>(&sY[g][1][0])[w];
            if (!any) levels = s + 1;   // uniform: the remaining levels are negligible
        }
    }
    if constexpr (MODE != kCached) {
#pragma unroll
        for (int q = 0; q < BR; ++q) Fin[q] = minv(D[q]);
    }
    if constexpr (MODE == kFactor) {
        // one system per workgroup (G == 1): its factor block
#pragma unroll
        for (int s = 0; s < LMAX; ++s)
            if (s < levels)
#pragma unroll
                for (int q = 0; q < BR; ++q) {
                    const int k = t + NT * q;
                    Fc_al[(long long)s * NB + k] = f4(cA[s][q]);
                    Fc_be[(long long)s * NB + k] = f4(cB[s][q]);
                }
#pragma unroll
        for (int q = 0; q < BR; ++q) Fc_fin[t + NT * q] = f4(Fin[q]);
        if (t == 0) Fc_levels[0] = levels;
        return;
    }

    // ---------------- solve + refinement sweeps ----------------
    int sweep = 0;
    double dprev = 0.0;
    bool fin = false;   // this system has stopped (uniform over its threads)
    for (;;) {
#pragma unroll
        for (int s = 0; s < LMAX; ++s) {
            if (s < levels) {
                const int delta = 1 << s, p = s & 1;
                if (!fin) {
#pragma unroll
                    for (int q = 0; q < BR; ++q) {
                        const int k = t + NT * q;
                        Yx[p][k] = make_float2(v[q][0], v[q][1]);
                    }
                }
                __syncthreads();
                if (!fin) {
#pragma unroll
                    for (int q = 0; q < BR; ++q) {
                        const int k = t + NT * q;
                        const int kl = k - delta >= 0 ? k - delta : k;
                        const int kr = k + delta < nb ? k + delta : k;
                        const float2 yl = Yx[p][kl], yr = Yx[p][kr];
                        const float l0 = yl.x, l1 = yl.y, r0 = yr.x, r1 = yr.y;
                        const M2& al = cA[s][q];
                        const M2& be = cB[s][q];
                        v[q][0] = fmaf(be.p, r0, fmaf(be.q, r1, fmaf(al.p, l0, fmaf(al.q, l1, v[q][0]))));
                        v[q][1] = fmaf(be.r, r0, fmaf(be.s, r1, fmaf(al.r, l0, fmaf(al.s, l1, v[q][1]))));
                    }
                }
            }
        }
        double dmax = 0.0, xmax = 0.0;
        if (!fin) {
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
        }
        // (the next write to sRed follows at least the residual's barrier)
        block_max2<NT>(dmax, xmax, sRed, t);
        if (!fin) {
            // Stop when the correction is below tol, or when the error left
            // after applying it, estimated from the observed contraction
            // rho = |dx_k| / |dx_{k-1}| as rho * |dx_k| with a safety factor
            // of 10, is below tol (and rho <= 0.1). rho needs two refinement
            // corrections: |dx_1| / |x_0| measures the first solve's error,
            // not the contraction.
            const bool small = dmax <= tol * xmax;
            const bool contracting = sweep > 1 && dmax <= 0.1 * dprev &&
                                     10.0 * dmax * dmax <= tol * xmax * dprev;
            // dmax = INF marks a NaN (or an overflow); INF <= 0.1 * INF would
            // pass the contraction test
            const bool done = dmax < INFINITY && xmax < INFINITY && (small || contracting);
            dprev = dmax;
            if (done || sweep >= max_sweeps) {
                if (!done) ++sweep;               // flag: tolerance not met
                fin = true;
            } else {
                ++sweep;
            }
        }
        if constexpr (G == 1) {
            if (fin) break;
        } else {
            // every system of the workgroup stopped? Each system's flag sits in
            // its own idle sF[.][2] words, after the max-reduction scratch.
            constexpr int kFlag = 4 * (NT / 64);   // int index past sRed's doubles
            if (t == 0) reinterpret_cast<int*>(&F[2][0])[kFlag] = fin ? 1 : 0;
            __syncthreads();
            bool all = true;
#pragma unroll
            for (int g = 0; g < G; ++g) all = all && reinterpret_cast<const int*>(&sF[g][2][0])[kFlag] != 0;
            if (all) break;
        }
        // r = y - A x, exact FP64 (y is overwritten only at the end)
        if (!fin) {
#pragma unroll
            for (int q = 0; q < BR; ++q) {
                const int k = t + NT * q;
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    if (2 * k + h < n) sX[h * NB + k] = X[q][h];
            }
        }
        __syncthreads();
        if (!fin) {
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
                        double ax = A[2][p] * X[q][h];
                        ax = __builtin_fma(A[0][p], xm2, ax);
                        ax = __builtin_fma(A[1][p], xm1, ax);
                        ax = __builtin_fma(A[3][p], xp1, ax);
                        ax = __builtin_fma(A[4][p], xp2, ax);
                        v[q][h] = (float)(A[5][p] - ax);
                    }
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
            if (i < n) {
                const long long g = base + (long long)i * rs;
                if constexpr (Src::kFused) gx[g] = X[q][h] + gx[g];   // this thread's W
                else gx[g] = add ? X[q][h] + add[g] : X[q][h];
            }
        }
    }
    if (sweeps_out && t == 0) sweeps_out[sys] = sweep;
}

template <int NB>
constexpr int lmax_for()
{
    return (NB >= 1024) ? 10 : (NB >= 512) ? 9 : (NB >= 256) ? 8 : (NB >= 128) ? 7 : 6;
}

// mode kFactor: one workgroup per matrix (A.m systems, or one when shared)
template <int NT, int BR, int G>
static int launch_f32r(const SolveArgsF32R& A, hipStream_t st, int mode = kSolve,
                       void* F = nullptr)
{
    constexpr int LMAX = lmax_for<NT * BR>();
    static_assert((1 << LMAX) >= NT * BR, "LMAX covers ceil(log2 block rows)");
    const long long ars = A.shared ? A.ars : A.rs, ass = A.shared ? 0 : A.ss;
    double* const gx = A.out ? A.out : A.y;
    float4* const f4p = static_cast<float4*>(const_cast<void*>(mode == kFactor ? F : A.fct));
    if (mode == kFactor) {
        if constexpr (G == 1)
            hipLaunchKernelGGL((lds_f32r_kernel<NT, BR, LMAX, 1, kFactor>),
                               dim3(A.shared ? 1 : A.m), dim3(NT), 0, st, A.a, A.b, A.c, A.d,
                               A.e, A.y, gx, A.add, A.n, A.shared ? 1 : A.m, A.rs, A.ss, ars, ass,
                               0, 0.0, nullptr, f4p, PlainRhs{});
        else
            return -2;
    } else if (A.fct) {
        hipLaunchKernelGGL((lds_f32r_kernel<NT, BR, LMAX, G, kCached>), dim3(A.m / G),
                           dim3(NT * G), 0, st, A.a, A.b, A.c, A.d, A.e, A.y, gx, A.add, A.n,
                           A.m, A.rs, A.ss, ars, ass, A.max_sweeps, A.tol, A.sweeps_out, f4p, PlainRhs{});
    } else {
        hipLaunchKernelGGL((lds_f32r_kernel<NT, BR, LMAX, G, kSolve>), dim3(A.m / G),
                           dim3(NT * G), 0, st, A.a, A.b, A.c, A.d, A.e, A.y, gx, A.add, A.n,
                           A.m, A.rs, A.ss, ars, ass, A.max_sweeps, A.tol, A.sweeps_out, f4p, PlainRhs{});
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int lds_f32r_max_n() { return 2048; }

// Block rows (capacity) and levels of the one-workgroup variant for size n:
// the cached factor of a system takes f32r_factor_bytes_nb of them.
static void f32r_shape(int n, int& nb_cap, int& lmax)
{
    nb_cap = n <= 128 ? 64 : n <= 256 ? 128 : n <= 512 ? 256 : n <= 1024 ? 512 : 1024;
    lmax = nb_cap >= 1024 ? 10 : nb_cap >= 512 ? 9 : nb_cap >= 256 ? 8 : nb_cap >= 128 ? 7 : 6;
}

size_t factor_f32r_bytes(int n)
{
    int nb, lm;
    f32r_shape(n, nb, lm);
    return (size_t)f32r_factor_bytes_nb(nb, lm);
}

// Grouped loading needs an interleaved batch (rows m apart, systems adjacent)
// whose size is a multiple of the group and six 16-byte aligned arrays.
static bool groupable(const SolveArgsF32R& A, int G)
{
    const auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    // a cached factor of separate matrices is one block per system (levels may differ)
    if (A.fct && !A.shared) return false;
    return A.ss == 1 && A.rs == A.m && A.m % G == 0 && al(A.y) &&
           (A.shared || (al(A.a) && al(A.b) && al(A.c) && al(A.d) && al(A.e)));
}

int lds_f32r_solve(const SolveArgsF32R& A, hipStream_t st)
{
    const int cus = device_cu_count();
    if (cus <= 0) return -1;
    const int n = A.n;
    if (n <= 128) return launch_f32r<64, 1, 1>(A, st);
    if (n <= 256) return launch_f32r<128, 1, 1>(A, st);
    // From n = 257 on, two block rows per thread: the cached multipliers of
    // every level stay in registers without spilling (one block row per
    // thread at 512 threads needed a 128-register cap for two workgroups per
    // CU and spilled: 11% slower at n = 1024).
    // Pairs of neighbouring systems per workgroup on interleaved batches
    // with more systems than CUs (so that one system per workgroup would put
    // two or more workgroups on some CUs). Measured on MI355X against one
    // system per workgroup: n <= 512 (80 KB per pair, two pairs per CU)
    // 7.5% faster at m = 1024 and 2% at m = 16384, but 9% slower at m = 256
    // (half the CUs idle); n <= 1024 7% faster at m = 300 and 3.5% at
    // m = 512, but only up to m = 2 CUs: a 160 KB pair holds the whole CU,
    // so in a longer batch no second workgroup's levels hide its loads (25%
    // slower at m = 16384). Groups of 4..16 systems of n <= 256 measured
    // slower at m = 16384.
    if (n <= 512)
        return groupable(A, 2) && A.m > cus ? launch_f32r<128, 2, 2>(A, st)
                                            : launch_f32r<128, 2, 1>(A, st);
    if (n <= 1024)
        return groupable(A, 2) && A.m > cus && A.m <= 2 * cus ? launch_f32r<256, 2, 2>(A, st)
                                                              : launch_f32r<256, 2, 1>(A, st);
    if (n <= 2048) return launch_f32r<512, 2, 1>(A, st);
    return -2;
}

template <int NT, int BR>
static int launch_f32r_stage(const SolveArgsF32R& A, const HestonStage& src, hipStream_t st)
{
    constexpr int LMAX = lmax_for<NT * BR>();
    hipLaunchKernelGGL((lds_f32r_kernel<NT, BR, LMAX, 1, kSolve, 1, HestonStage>), dim3(A.m),
                       dim3(NT), 0, st, A.a, A.b, A.c, A.d, A.e, nullptr, A.out, nullptr, A.n,
                       A.m, A.rs, A.ss, A.rs, A.ss, A.max_sweeps, A.tol, A.sweeps_out, nullptr,
                       src);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// The one-system shapes of lds_f32r_solve (contiguous lines never group).
int lds_f32r_solve_stage(const SolveArgsF32R& A, const HestonStage& src, hipStream_t st)
{
    const int n = A.n;
    if (A.rs != 1 || A.ss != n || !A.out) return -2;
    if (n <= 128) return launch_f32r_stage<64, 1>(A, src, st);
    if (n <= 256) return launch_f32r_stage<128, 1>(A, src, st);
    if (n <= 512) return launch_f32r_stage<128, 2>(A, src, st);
    if (n <= 1024) return launch_f32r_stage<256, 2>(A, src, st);
    if (n <= 2048) return launch_f32r_stage<512, 2>(A, src, st);
    return -2;
}

int lds_f32r_factor(const SolveArgsF32R& A, void* F, hipStream_t st)
{
    const int n = A.n;
    if (n <= 128) return launch_f32r<64, 1, 1>(A, st, kFactor, F);
    if (n <= 256) return launch_f32r<128, 1, 1>(A, st, kFactor, F);
    if (n <= 512) return launch_f32r<128, 2, 1>(A, st, kFactor, F);
    if (n <= 1024) return launch_f32r<256, 2, 1>(A, st, kFactor, F);
    if (n <= 2048) return launch_f32r<512, 2, 1>(A, st, kFactor, F);
    return -2;
}

}  // namespace pcrp
