// Spaced reduction steps through LDS (FP64), shared by the LDS-resident solve
// and the fused steps of the HBM tile passes.
//
// At stride d = 2^S the rows of a tile split into d chains (residue r =
// i mod d); chain positions p (row r + d*p) are cut into blocks of 2*RPT, and
// a thread owns the RPT positions of one parity in one block: p = P0 + 2k with
// P0 = par + 2*RPT*b. Its rows' pair equations are centred on the odd
// positions P0 - 1, P0 + 1, ..., P0 + 2*RPT - 1, and each of those RPT + 1
// centres serves two of its rows (as E(i+d) of one and E(i-d) of the next),
// so a row-step costs (RPT+1)/RPT pair equations instead of 2:
// 2 + 2(RPT+1)/RPT divisions (4.5 at RPT = 4) and 2*RPT + 3 rows read from
// LDS per RPT rows. The same pair-equation and reduction expressions
// (pcrp_device.h) run on the same operands, so the bits are those of the
// redundant form (kernel.cu:66-140 as written).
//
// Layout (per stride): thread t = r + d*(par + 2b) of a tile's NT threads owns
// slot t, and its row k sits at element (k*NT + t)*G + g of each coefficient
// array, g being the tile's sub-system (G sub-systems interleaved, so the
// lanes of a wave run over g fastest). The partner thread (other parity, same
// block) is slot t ^ d; the outer rows are rows of slots t -+ 2d; so every lane
// reads one element per instruction and consecutive lanes read consecutive
// elements (no bank conflicts). A step writes its new rows straight into the
// layout of stride 2d, or, after the last step, in natural order (tile row i
// at nat(i)*G + g, nat a permutation the caller picks for its own reads).
//
// Rows of the tile outside the system (global index l = lofs + i outside
// [0, L)) hold the identity row (0, 0, 1, 0, 0, 0) in every layout, so every
// row read is finite; neighbours past either end of a chain inside the tile
// are replaced by the thread's own first / last row, as the redundant form
// replaces them by the row itself. Both only meet exact zero multipliers, or
// rows whose results are discarded (a pass tile's halo).
#pragma once

#include "pcrp_device.h"

namespace pcrp {

__device__ __forceinline__ Row<double> identity_row() { return Row<double>{0.0, 0.0, 1.0, 0.0, 0.0, 0.0}; }

// Slot-row index (k*NT + slot) of tile row i in the spaced layout of stride 2^S.
template <int NT, int RPT>
__device__ __forceinline__ int spaced_addr(int i, int S)
{
    const int d = 1 << S;
    const int r = i & (d - 1), p = i >> S;
    const int b = p / (2 * RPT), q = p & (2 * RPT - 1);
    return (q >> 1) * NT + r + d * ((q & 1) + 2 * b);
}

// Natural order with no permutation (see spaced_step).
struct NaturalOrder {
    __device__ int operator()(int i) const { return i; }
};

// One step at stride 2^S (2^S <= NT/2) over a tile of NT*RPT rows of each of
// G sub-systems, rows in the spaced layout of that stride. Tile row i is row
// lofs + i of a sub-system of L rows (the guards use these global indices).
// Writes the new rows in the layout of stride 2^(S+1), or, when `natural`, in
// natural order: tile row i at slot-row nat(i), nat a permutation of the tile
// rows (identity by default). All threads of the workgroup call it (two
// barriers).
template <int NT, int RPT, int G, typename Nat = NaturalOrder>
__device__ __forceinline__ void spaced_step(int S, bool natural, int lofs, int L,
                                            double (&sR)[6][NT * RPT * G], Nat nat = Nat())
{
    static_assert(RPT >= 2 && RPT % 2 == 0, "spaced steps need an even number of rows per thread");
    const int tid = threadIdx.x;
    const int t = G == 1 ? tid : tid / G;
    const int g = G == 1 ? 0 : tid - t * G;
    const int d = 1 << S;
    const int r = t & (d - 1), par = (t >> S) & 1, b = t >> (S + 1);
    const int NB = NT >> (S + 1);
    const int P0 = par + 2 * RPT * b;
    const int l0 = lofs + r + d * P0;           // global index of own row 0
    const int lend = l0 + 2 * d * (RPT - 1);    // ... and of own row RPT - 1
    auto rd = [&](int sk) {
        const int a = sk * G + g;
        Row<double> q;
        q.a = sR[0][a]; q.b = sR[1][a]; q.c = sR[2][a];
        q.d = sR[3][a]; q.e = sR[4][a]; q.y = sR[5][a];
        return q;
    };
    auto wr = [&](int sk, const Row<double>& q) {
        const int a = sk * G + g;
        sR[0][a] = q.a; sR[1][a] = q.b; sR[2][a] = q.c;
        sR[3][a] = q.d; sR[4][a] = q.e; sR[5][a] = q.y;
    };
    // odd positions P0 + 2m - 1: m = 1..RPT-1 from the partner at bo + m*NT;
    // m = 0 and m = RPT may lie in the neighbouring block's partner slot
    const int bo = par ? t - d : t + d - NT;
    const int a_o0 = par ? t - d : (b > 0 ? (RPT - 1) * NT + t - d : t);
    const int a_on = par ? (b < NB - 1 ? t + d : (RPT - 1) * NT + t) : (RPT - 1) * NT + t + d;
    const int a_l2 = b > 0 ? (RPT - 1) * NT + t - 2 * d : t;        // position P0 - 2
    const int a_r2 = b < NB - 1 ? t + 2 * d : (RPT - 1) * NT + t;   // position P0 + 2*RPT
    Row<double> N[RPT];
    if (l0 < L && lend >= 0) {   // a thread with no row of the system only writes identity rows
        Pair<double> Ep;
        Row<double> Lw = rd(a_l2);
#pragma unroll
        for (int m = 0; m <= RPT; ++m) {
            // E(m) serves rows m - 1 and m; once row m - 1 is past the end,
            // so is every later row (their slots get identity rows below)
            if (m > 0 && l0 + 2 * d * (m - 1) >= L) break;
            const Row<double> M = rd(m == 0 ? a_o0 : (m == RPT ? a_on : bo + m * NT));
            const Row<double> R = rd(m == RPT ? a_r2 : m * NT + t);
            const int lc = l0 + (2 * m - 1) * d;
            const Pair<double> E = make_pair(Lw, M, R, lc - d >= 0, lc - 2 * d >= 0, lc + d <= L - 1,
                                             lc + 2 * d <= L - 1);
            if (m > 0) {
                const int lo = lc - d;   // own row m - 1, held in Lw
                N[m - 1] = reduce_row(Lw, Ep, E, lo - d >= 0, lo + d <= L - 1, lo - 2 * d >= 0,
                                      lo + 2 * d <= L - 1,
                                      [&](double& alL, double& arL, double& alR, double& arR) {
                                          alL = Ep.al; arL = Ep.ar; alR = E.al; arR = E.ar;
                                      });
            }
            Ep = E;
            Lw = R;
        }
    }
    __syncthreads();   // every read of this step done
    if (natural) {
        const int base = r + d * P0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int l = l0 + 2 * d * k;
            if (l >= 0 && l < L) wr(nat(base + 2 * d * k), N[k]);
        }
    } else {
        // row r + d*(P0 + 2k) at stride 2d: residue r + d*par, block b/2,
        // parity k&1, slot row (RPT/2)*(b&1) + k/2
        const int w0 = (RPT / 2) * (b & 1) * NT + r + d * par + 4 * d * (b >> 1);
        const int w1 = w0 + 2 * d;
        if (__builtin_expect(l0 >= 0 && lend < L, 1)) {   // all RPT rows in the system
#pragma unroll
            for (int k = 0; k < RPT; ++k) wr(((k & 1) ? w1 : w0) + (k >> 1) * NT, N[k]);
        } else {
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int sk = ((k & 1) ? w1 : w0) + (k >> 1) * NT;
                const int l = l0 + 2 * d * k;
                if (l >= 0 && l < L) wr(sk, N[k]);
                else wr(sk, identity_row());
            }
        }
    }
    __syncthreads();
}

}  // namespace pcrp
