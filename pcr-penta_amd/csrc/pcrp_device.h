// Device-side building blocks shared by the PCR-Penta kernels (gfx950).
//
// Notation (kernel.cu:32-39): row i reads a_i x_{i-2d} + b_i x_{i-d} + c_i x_i
// + d_i x_{i+d} + e_i x_{i+2d} = y_i at stride d (delta).
//
// One reduction step of the reference (kernel.cu:66-142) builds, for every row
// i, eq1 = alfa1*row(i-2d) + row(i-d) + alfa2*row(i) (Reduction-I) and
// eq2 = beta1*row(i) + row(i+d) + beta2*row(i+2d) (Reduction-II), then
// new row(i) = g1*eq1 - row(i) + g2*eq2 (Reduction-III).
// eq1 of row i and eq2 of row i-2d are the same equation centred on j = i-d,
// with the same multipliers and the same operand order, so the kernels build
// it ONCE per centre j ("pair equation" E(j)) and every row reads E(i-d) and
// E(i+d). That is a common-subexpression elimination: it changes no rounding
// (the FP64 results equal the oracle's literal restatement bit for bit) and
// cuts the FP64 divisions per row and step from 6 to 4.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcrp {

// Six coefficients of one row: a..e and the right-hand side y.
template <typename T>
struct Row {
    T a, b, c, d, e, y;
};

// Pair equation centred on j (coefficients on x_{j-2d}, x_{j-d}, x_j, x_{j+d},
// x_{j+2d}: A, B, C, D, E after eliminating x_{j-2d}... see make_pair) plus
// its two multipliers, kept for Reduction-III's zero-pivot rule.
template <typename T>
struct Pair {
    T A, B, C, D, E, K;   // K: right-hand side
    T al, ar;             // alfa (left multiplier), alfa' (right multiplier)
};

// E(j) from rows L = j-d, M = j, R = j+d (kernel.cu:69-117).
//   hasL : j-d >= 0      (row L exists; eq1's two-neighbour branch)
//   aC   : j-2d >= 0     (kernel.cu:72 / 101: the left multiplier is formed)
//   hasR : j+d <= n-1
//   arC  : j+2d <= n-1   (kernel.cu:74 / 103)
// When hasL (hasR) is false, L (R) must be any finite row: its terms are
// multiplied by an exact zero multiplier.
__device__ __forceinline__ Pair<double> make_pair(const Row<double>& L, const Row<double>& M,
                                                  const Row<double>& R, bool hasL, bool aC,
                                                  bool hasR, bool arC)
{
    Pair<double> E;
    const double ql = -M.a / L.b;   // -a[p1]/b[p2]  (kernel.cu:73), = -a[q1]/b[id] (:99)
    const double qr = -M.e / R.d;   // -e[p1]/d[id]  (kernel.cu:75), = -e[q1]/d[q2] (:101)
    E.al = (aC && L.b != 0.0) ? ql : 0.0;
    E.ar = (arC && R.d != 0.0) ? qr : 0.0;
    E.A = E.al * L.a;
    E.B = hasL ? __builtin_fma(E.ar, R.a, __builtin_fma(E.al, L.c, M.b)) : 0.0;  // b1=0 (:86)
    E.C = __builtin_fma(E.ar, R.b, __builtin_fma(E.al, L.d, M.c));
    E.D = hasR ? __builtin_fma(E.ar, R.c, __builtin_fma(E.al, L.e, M.d)) : 0.0;  // d2=0 (:114)
    E.E = E.ar * R.e;
    E.K = __builtin_fma(E.ar, R.y, __builtin_fma(E.al, L.y, M.y));
    return E;
}

// Reduction-III and write-back (kernel.cu:120-140) for row M with
// EL = E(i-d), ER = E(i+d). hp1: i-d >= 0, hq1: i+d <= n-1, hp2: i-2d >= 0,
// hq2: i+2d <= n-1. alL/alR/arL/arR: multipliers of EL/ER, consulted only when
// a pivot EL.C / ER.C is exactly zero (kernel.cu:122,129 keep the multiplier
// assigned last: eq2's if it exists, else eq1's).
// gout (optional): receives the final multipliers g1, g2, which with the pair
// multipliers al, ar are all a right-hand side needs (the factored solve).
template <typename LoadAl>
__device__ __forceinline__ Row<double> reduce_row(const Row<double>& M, const Pair<double>& EL,
                                                  const Pair<double>& ER, bool hp1, bool hq1,
                                                  bool hp2, bool hq2, LoadAl&& load_multipliers,
                                                  double* gout = nullptr)
{
    double g1 = M.b / EL.C;   // kernel.cu:122
    double g2 = M.d / ER.C;   // kernel.cu:129
    if (__builtin_expect((EL.C == 0.0) | (ER.C == 0.0), 0)) {
        double alL, arL, alR, arR;
        load_multipliers(alL, arL, alR, arR);
        if (EL.C == 0.0) g1 = hq1 ? alR : (hp2 ? alL : 0.0);
        if (ER.C == 0.0) g2 = hq2 ? arR : (hp1 ? arL : 0.0);
    }
    g1 = hp1 ? g1 : 0.0;
    g2 = hq1 ? g2 : 0.0;
    if (gout) {
        gout[0] = g1;
        gout[1] = g2;
    }
    Row<double> N;
    N.a = g1 * EL.A;
    N.b = __builtin_fma(g2, ER.A, __builtin_fma(g1, EL.B, -M.a));
    N.c = __builtin_fma(g2, ER.B, __builtin_fma(g1, EL.D, -M.c));
    N.d = __builtin_fma(g2, ER.D, __builtin_fma(g1, EL.E, -M.e));
    N.e = g2 * ER.E;
    N.y = __builtin_fma(g2, ER.K, __builtin_fma(g1, EL.K, -M.y));
    return N;
}

// ---- FP32 helpers (inner solve of the mixed-precision refinement) --------
// Approximate reciprocal (v_rcp_f32, 1 ulp): the FP64 refinement absorbs the
// difference. Pivots below 2^-100 are treated as zero (a zero inverse): deep
// reduction levels of a dominant system push couplings into the FP32
// denormal range, where v_rcp_f32 returns infinity.
constexpr float kTinyF32 = 0x1p-100f;

__device__ __forceinline__ float rcp_f32(float x) { return __builtin_amdgcn_rcpf(x); }

// 2x2 FP32 matrix [[p, q], [r, s]]: the blocks of the block-tridiagonal form
// the FP32 inner solve works on.
struct M2 {
    float p, q, r, s;
};

constexpr M2 kZero2 = {0.f, 0.f, 0.f, 0.f};
constexpr M2 kEye2 = {1.f, 0.f, 0.f, 1.f};

__device__ __forceinline__ M2 mmul(const M2& x, const M2& y)
{
    return M2{fmaf(x.p, y.p, x.q * y.r), fmaf(x.p, y.q, x.q * y.s),
              fmaf(x.r, y.p, x.s * y.r), fmaf(x.r, y.q, x.s * y.s)};
}

// -(x y)
__device__ __forceinline__ M2 mmul_neg(const M2& x, const M2& y)
{
    return M2{-fmaf(x.p, y.p, x.q * y.r), -fmaf(x.p, y.q, x.q * y.s),
              -fmaf(x.r, y.p, x.s * y.r), -fmaf(x.r, y.q, x.s * y.s)};
}

// d + x y + z w
__device__ __forceinline__ M2 madd2(const M2& d, const M2& x, const M2& y, const M2& z,
                                    const M2& w)
{
    return M2{fmaf(z.p, w.p, fmaf(z.q, w.r, fmaf(x.p, y.p, fmaf(x.q, y.r, d.p)))),
              fmaf(z.p, w.q, fmaf(z.q, w.s, fmaf(x.p, y.q, fmaf(x.q, y.s, d.q)))),
              fmaf(z.r, w.p, fmaf(z.s, w.r, fmaf(x.r, y.p, fmaf(x.s, y.r, d.r)))),
              fmaf(z.r, w.q, fmaf(z.s, w.s, fmaf(x.r, y.q, fmaf(x.s, y.s, d.s))))};
}

__device__ __forceinline__ float maxabs2(const M2& m)
{
    return fmaxf(fmaxf(fabsf(m.p), fabsf(m.q)), fmaxf(fabsf(m.r), fabsf(m.s)));
}

// Truncated (incomplete) cyclic reduction for the FP32 inner solve: once
// every multiplier entry of a level (alpha = -L D^-1, beta = -U D^-1) is at
// most this, the couplings left after it are at most that times the ones
// before, and the next level's multipliers about its square (quadratic
// decay of the off-diagonal blocks of block cyclic reduction on diagonally
// dominant matrices, Heller, SIAM J. Numer. Anal. 13 (1976) 484-496): the
// remaining levels would change X by ~2^-32 relative, far below the FP32
// solve's own ~3e-7, so X = D^-1 Y right after that level. Systems whose
// couplings do not decay run every level.
constexpr float kTruncF32 = 0x1p-16f;

__device__ __forceinline__ float4 f4(const M2& m) { return make_float4(m.p, m.q, m.r, m.s); }
__device__ __forceinline__ M2 m2(const float4& v) { return M2{v.x, v.y, v.z, v.w}; }

// Inverse by the adjugate; a determinant below 2^-100 gives the zero matrix.
__device__ __forceinline__ M2 minv(const M2& m)
{
    const float det = fmaf(m.p, m.s, -(m.q * m.r));
    const float rd = fabsf(det) >= kTinyF32 ? rcp_f32(det) : 0.f;
    return M2{m.s * rd, -m.q * rd, -m.r * rd, m.p * rd};
}

// Blocks are dispatched round-robin over the 8 XCDs (each with its own L2).
// Remap so that each XCD works on runs of C consecutive systems at a time:
// in the interleaved layout those share cache lines (C = 16 FP64 or 32 FP32
// values per 128-byte line), so the line is fetched into one L2 instead of
// eight. Bijective on [0, m).
template <int C>
__device__ __forceinline__ int xcd_remap_c(int b, int m)
{
    constexpr int NX = 8;
    const int full = (m / (NX * C)) * (NX * C);
    if (b >= full) return b;
    const int xcd = b % NX, s = b / NX;
    return (s / C) * (NX * C) + xcd * C + (s % C);
}

__device__ __forceinline__ int xcd_remap(int b, int m) { return xcd_remap_c<16>(b, m); }

// ---- seeded generator (same bits as the oracle's oracle_gen_uniform) ------
__device__ __forceinline__ uint64_t splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double gen_uniform(uint64_t key, int diag, uint64_t sys,
                                              uint64_t row, double lo, double hi)
{
    const uint64_t u = splitmix64(key + (((uint64_t)diag << 56) | (sys << 24) | row));
    const double f = (double)(u >> 11) * 0x1.0p-53;
    const double span = hi - lo;
    const double prod = span * f;   // built with -ffp-contract=off: no fusion
    return lo + prod;
}

}  // namespace pcrp
