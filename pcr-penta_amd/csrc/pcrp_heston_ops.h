// Heston ADI stencils: the grid, the difference operators A0, A1, A2 with
// their boundary terms, and the explicit stage of one Douglas step evaluated
// at a single unknown. Shared by the driver (pcrp_heston.hip: matrices, the
// separate explicit-stage kernel) and by the S-line solve kernels that
// compute the explicit stage while they load their rows (HestonStage).
//
// PDE in time-to-maturity tau for a European call:
//   u_tau = 1/2 v S^2 u_SS + rho sigma v S u_Sv + 1/2 sigma^2 v u_vv
//           + r S u_S + kappa (theta_v - v) u_v - r u
// Unknowns: S_i = i hS, i = 1..Ns (hS = Smax/(Ns+1); u = 0 at S = 0 and
// u = Smax - K e^{-r tau} at S = Smax), v_j = j hv, j = 0..Nv-1 (hv = vmax/Nv;
// the PDE holds at v = 0; u = S at v = vmax). Storage U[j*Ns + (i-1)].
//   A1 (S): 1/2 v S^2 u_SS + r S u_S - r u/2, fourth-order 5-point central
//           differences (3-point second order at i = 1 and i = Ns);
//   A2 (v): 1/2 sigma^2 v u_vv (3-point) + kappa(theta_v - v) u_v with
//           second-order upwinding (first order next to v = vmax / at j = 1),
//           - r u/2;
//   A0    : rho sigma v S u_Sv, 4-point central cross difference (explicit).
// Built with -ffp-contract=off: every expression rounds the same way in each
// kernel that evaluates it, so the folded and the separate explicit stage give
// the same bits.
#pragma once

#include <hip/hip_runtime.h>

#include "pcrp_internal.h"

namespace pcrp {

struct HestonGrid {
    int ns, nv;
    double hS, hv, smax, vmax, kappa, thv, sigma, rho, r, K, dt, th;
    // reciprocals of the difference-quotient denominators (host-computed, so
    // the stencil coefficients take multiplications, not FP64 divisions)
    double r12hS2, r12hS, rhS2, r2hS, rhv2, r2hv, rhv, r4hShv;
};

__host__ __device__ inline double S_at(const HestonGrid& g, int i) { return i * g.hS; }
__host__ __device__ inline double v_at(const HestonGrid& g, int j) { return j * g.hv; }

// far-S boundary value Smax - K e^{-r tau} at time-to-maturity tau
__device__ inline double far_s(const HestonGrid& g, double tau)
{
    return g.smax - g.K * exp(-g.r * tau);
}

// value of the grid function at node (i, j), i in 0..ns+1, j in 0..nv
// (boundary nodes from the boundary conditions; bS = far_s(tau))
__device__ inline double node(const HestonGrid& g, const double* U, int i, int j, double bS)
{
    if (i == 0) return 0.0;
    if (i == g.ns + 1) return bS;
    if (j == g.nv) return S_at(g, i);
    return U[j * g.ns + (i - 1)];   // ns * nv < 2^31 (checked at entry)
}

// stencil coefficients of A1 at row i (offsets -2..2)
__host__ __device__ inline void a1_coef(const HestonGrid& g, int i, int j, double w[5])
{
    const double S = S_at(g, i), v = v_at(g, j);
    const double D = 0.5 * v * S * S, C = g.r * S;
    if (i >= 2 && i <= g.ns - 1) {
        const double d2 = D * g.r12hS2, d1 = C * g.r12hS;
        w[0] = -d2 + d1;
        w[1] = 16.0 * d2 - 8.0 * d1;
        w[2] = -30.0 * d2 - 0.5 * g.r;
        w[3] = 16.0 * d2 + 8.0 * d1;
        w[4] = -d2 - d1;
    } else {
        const double d2 = D * g.rhS2, d1 = C * g.r2hS;
        w[0] = 0.0;
        w[1] = d2 - d1;
        w[2] = -2.0 * d2 - 0.5 * g.r;
        w[3] = d2 + d1;
        w[4] = 0.0;
    }
}

// stencil coefficients of A2 at row j (offsets -2..2)
__host__ __device__ inline void a2_coef(const HestonGrid& g, int j, double w[5])
{
    const double v = v_at(g, j);
    const double diff = 0.5 * g.sigma * g.sigma * v * g.rhv2;
    const double mu = g.kappa * (g.thv - v);
    w[0] = w[1] = w[3] = w[4] = 0.0;
    w[2] = -0.5 * g.r;
    if (j >= 1) {   // diffusion (zero at v = 0)
        w[1] += diff;
        w[2] += -2.0 * diff;
        w[3] += diff;
    }
    if (mu >= 0.0) {
        if (j + 2 <= g.nv) {   // second-order forward
            const double f = mu * g.r2hv;
            w[2] += -3.0 * f; w[3] += 4.0 * f; w[4] += -f;
        } else {               // first-order forward next to v = vmax
            const double f = mu * g.rhv;
            w[2] += -f; w[3] += f;
        }
    } else {
        if (j >= 2) {          // second-order backward
            const double f = mu * g.r2hv;
            w[2] += 3.0 * f; w[1] += -4.0 * f; w[0] += f;
        } else {               // first-order backward at j = 1
            const double f = mu * g.rhv;
            w[2] += f; w[1] += -f;
        }
    }
}

// A1 U (with boundary nodes at tau) at unknown (i, j); w: a1_coef(i, j)
__device__ inline double apply_a1(const HestonGrid& g, const double* U, int i, int j, double bS,
                                  const double (&w)[5])
{
    double acc = w[2] * node(g, U, i, j, bS);
    acc += w[1] * node(g, U, i - 1, j, bS);
    acc += w[3] * node(g, U, i + 1, j, bS);
    if (w[0] != 0.0) acc += w[0] * node(g, U, i - 2, j, bS);
    if (w[4] != 0.0) acc += w[4] * node(g, U, i + 2, j, bS);
    return acc;
}

// boundary part b1(tau) of A1 at (i, j): terms on boundary nodes only
__device__ inline double bnd_a1(const HestonGrid& g, int i, int j, double bS,
                                const double (&w)[5])
{
    double acc = 0.0;
    for (int o = -2; o <= 2; ++o) {
        const int ii = i + o;
        if (w[o + 2] != 0.0 && (ii == 0 || ii == g.ns + 1))
            acc += w[o + 2] * node(g, nullptr, ii, j, bS);
    }
    return acc;
}

__device__ inline double apply_a2(const HestonGrid& g, const double* U, int i, int j, double bS,
                                  const double (&w)[5])
{
    double acc = w[2] * node(g, U, i, j, bS);
    for (int o = -2; o <= 2; ++o) {
        if (o == 0 || w[o + 2] == 0.0) continue;
        acc += w[o + 2] * node(g, U, i, j + o, bS);
    }
    return acc;
}

__device__ inline double bnd_a2(const HestonGrid& g, int i, int j, double bS,
                                const double (&w)[5])
{
    double acc = 0.0;
    for (int o = 1; o <= 2; ++o)
        if (w[o + 2] != 0.0 && j + o == g.nv) acc += w[o + 2] * node(g, nullptr, i, j + o, bS);
    return acc;
}

__device__ inline double apply_a0(const HestonGrid& g, const double* U, int i, int j, double bS)
{
    if (j == 0) return 0.0;   // v = 0
    const double c = g.rho * g.sigma * v_at(g, j) * S_at(g, i) * g.r4hShv;
    const double x = node(g, U, i + 1, j + 1, bS) - node(g, U, i + 1, j - 1, bS) -
                     node(g, U, i - 1, j + 1, bS) + node(g, U, i - 1, j - 1, bS);
    return c * x;
}

// Douglas (theta = th), boundary terms b_k(tau) split off each A_k:
//   Y0 = U + dt (A0 U + A1 U + A2 U + b(tau_n))
//   R1 = Y0 - th dt (A1 U + b1(tau_n)) + th dt b1(tau_n+1)   (S-line right-hand side)
//   W  = -th dt (A2 U + b2(tau_n)) + th dt b2(tau_n+1)       (v-line right-hand side - Y1)
// at unknown (i, j); bS0, bS1: the far-S boundary at tau_n, tau_n+1; w2:
// a2_coef(g, j). The b(tau_n) terms enter through node()'s boundary values.
//
// kInterior: unknowns whose whole stencil lies inside the grid (3 <= i <= ns-2,
// 2 <= j <= nv-3) take the same expressions in the same order with node()
// as plain loads at fixed offsets from their own element and the boundary
// sums as the exact zeros they are there (the separate explicit-stage kernel
// keeps the general form, so comparing the two forms pins this one).
template <bool kInterior = false>
__device__ inline void heston_stage_node(const HestonGrid& g, const double* U, int i, int j,
                                         double bS0, double bS1, const double (&w2)[5],
                                         double& r1, double& w)
{
    if constexpr (kInterior) {
        if (i >= 3 && i <= g.ns - 2 && j >= 2 && j <= g.nv - 3) {
            const double* const p = U + (j * g.ns + (i - 1));
            const int ns = g.ns;
            const double u = p[0];
            double w1[5];
            a1_coef(g, i, j, w1);
            const double c = g.rho * g.sigma * v_at(g, j) * S_at(g, i) * g.r4hShv;
            const double x = p[ns + 1] - p[-ns + 1] - p[ns - 1] + p[-ns - 1];
            const double l0 = c * x;
            double l1 = w1[2] * u;
            l1 += w1[1] * p[-1];
            l1 += w1[3] * p[1];
            if (w1[0] != 0.0) l1 += w1[0] * p[-2];
            if (w1[4] != 0.0) l1 += w1[4] * p[2];
            double l2 = w2[2] * u;
            if (w2[0] != 0.0) l2 += w2[0] * p[-2 * ns];
            if (w2[1] != 0.0) l2 += w2[1] * p[-ns];
            if (w2[3] != 0.0) l2 += w2[3] * p[ns];
            if (w2[4] != 0.0) l2 += w2[4] * p[2 * ns];
            const double y0 = u + g.dt * (l0 + l1 + l2);
            const double th = g.th * g.dt;
            r1 = y0 - th * l1 + th * 0.0;
            w = -th * l2 + th * 0.0;
            return;
        }
    }
    const double u = U[j * g.ns + (i - 1)];
    double w1[5];
    a1_coef(g, i, j, w1);
    const double l0 = apply_a0(g, U, i, j, bS0);
    const double l1 = apply_a1(g, U, i, j, bS0, w1);
    const double l2 = apply_a2(g, U, i, j, bS0, w2);
    const double y0 = u + g.dt * (l0 + l1 + l2);
    const double th = g.th * g.dt;
    r1 = y0 - th * l1 + th * bnd_a1(g, i, j, bS1, w1);
    w = -th * l2 + th * bnd_a2(g, i, j, bS1, w2);
}

// The explicit stage of one Douglas step folded into the S-line solve: row i
// (0-based) of S-line j takes R1 at unknown (i + 1, j) computed from U (the
// previous step's grid, never written during the solve). W, computed with it,
// waits in the output element until the solve adds x to it; every kernel
// that does so reads that element back behind at least one workgroup barrier
// (its writer may be another thread of the workgroup). The step index is the
// device counter plus off, as in the separate kernel.
struct HestonStage {
    static constexpr bool kFused = true;
    HestonGrid g;
    const double* U;
    const int* step;
    int off;
    const double* bnd;   // far-S boundary values: bnd[2k] at tau = k dt, bnd[2k+1] one step on

    struct Line {
        HestonGrid g;
        const double* U;
        int j;
        double bS0, bS1;
        double w2[5];   // a2_coef(g, j): the same for every row of the line
        // R1 of row i, W of row i into *wout
        __device__ double r1(int i, double* wout) const
        {
            double r, w;
            heston_stage_node<true>(g, U, i + 1, j, bS0, bS1, w2, r, w);
            *wout = w;
            return r;
        }
    };
    __device__ Line bind(int j) const
    {
        const int k = *step + off;
        Line L{g, U, j, bnd[2 * k], bnd[2 * k + 1], {}};
        a2_coef(g, j, L.w2);
        return L;
    }
};

}  // namespace pcrp
