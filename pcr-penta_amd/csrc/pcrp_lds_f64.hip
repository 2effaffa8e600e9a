// LDS-resident FP64 PCR-Penta: one workgroup solves one whole system.
//
// Counterpart of pcr_penta (kernel.cu:32-154) and of the second split stage
// pcr_penta_splitting_2nd_stage (kernel.cu:665-791), which is the same solve
// on a strided sub-system. Differences by design:
//   * NT threads each own RPT rows (row i = t + NT*k), so n is not tied to the
//     workgroup size, the rows of a thread stay in VGPRs and the RPT rows give
//     each thread independent division chains to interleave;
//   * steps with stride >= NT couple only rows of the same thread: they run in
//     registers with no LDS traffic and no barrier;
//   * any element addressing (row stride, system stride, base offset), so the
//     reference's contiguous layout, the interleaved layout and the strided
//     sub-systems of the split path all use the same kernel;
//   * a..e are read-only; y is overwritten by x.
// Two ways to run a step through LDS (chosen per size by the autotuner):
//   SHARED  : each centre j builds its pair equation E(j) once (pcrp_device.h),
//             4 FP64 divisions per row-step, LDS = rows + pair equations
//             (14 doubles per row);
//   REDUNDANT: each row builds E(i-d) and E(i+d) itself from four neighbour
//             rows, exactly as kernel.cu:66-117 writes it: 6 divisions per
//             row-step but only the rows in LDS (6 doubles per row), so three
//             times as many workgroups fit per CU.
//   INPLACE : the shared form with each pair equation written over its own
//             row in one 6-double buffer (four barriers per step), so 4
//             divisions per row-step at the redundant form's LDS footprint.
//             A step in which any pair equation has an exactly zero pivot
//             (detected block-wide) runs as a redundant step instead, whose
//             rows are still intact, because the reference's zero-pivot rule
//             needs the neighbours' multipliers.
//   SPACED  : at each LDS stride a thread owns RPT rows two strides apart, so
//             its RPT + 1 pair equations serve all its rows (4.5 divisions per
//             row-step at RPT = 4) and it reads 2*RPT + 3 rows; the rows stay
//             in LDS between steps in a per-stride layout (pcrp_spaced.h).
// All give the same bits (the pair equation is the same expression on the
// same operands either way).
#include "pcrp_device.h"
#include "pcrp_internal.h"
#include "pcrp_spaced.h"
#include "pcrp_heston_ops.h"

namespace pcrp {

enum { KIND_SHARED = 0, KIND_REDUNDANT = 1, KIND_INPLACE = 2, KIND_SPACED = 3 };

template <int KIND>
struct LdsWords {
    static constexpr int value = KIND == KIND_SHARED ? 14 : 6;
};

// Field-wise row copy: keeps loops over row arrays from being turned into a
// memcpy of a private array (which would live in scratch memory).
__device__ __forceinline__ void set_row(Row<double>& d, const Row<double>& s)
{
    d.a = s.a; d.b = s.b; d.c = s.c; d.d = s.d; d.e = s.e; d.y = s.y;
}

// Register-resident step at stride delta = DK*NT: rows k-DK, k, k+DK of the
// same thread are the neighbours. new[k] from E(k-DK), row k, E(k+DK).
template <int NT, int RPT, int DK>
__device__ __forceinline__ void register_step(Row<double> (&r)[RPT], int t, int n)
{
    const int delta = DK * NT;
    Pair<double> E[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int j = t + NT * k;
        const bool hasL = k - DK >= 0;                 // j - delta >= 0
        const bool aC = k - 2 * DK >= 0;               // j - 2 delta >= 0
        const bool hasR = k + DK < RPT && j + delta <= n - 1;
        const bool arC = k + 2 * DK < RPT && j + 2 * delta <= n - 1;
        const Row<double>& L = r[hasL ? k - DK : k];
        const Row<double>& R = r[k + DK < RPT ? k + DK : k];
        E[k] = make_pair(L, r[k], hasR ? R : r[k], hasL, aC, hasR, arC);
    }
    Row<double> N[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        const bool hp1 = k - DK >= 0, hp2 = k - 2 * DK >= 0;
        const bool hq1 = k + DK < RPT && i + delta <= n - 1;
        const bool hq2 = k + 2 * DK < RPT && i + 2 * delta <= n - 1;
        const Pair<double>& EL = E[hp1 ? k - DK : k];
        const Pair<double>& ER = E[k + DK < RPT ? k + DK : k];
        const Pair<double>& ERs = hq1 ? ER : E[k];
        N[k] = reduce_row(r[k], EL, ERs, hp1, hq1, hp2, hq2,
                          [&](double& alL, double& arL, double& alR, double& arR) {
                              alL = EL.al; arL = EL.ar; alR = ERs.al; arR = ERs.ar;
                          });
    }
#pragma unroll
    for (int k = 0; k < RPT; ++k)
        if (t + NT * k < n) set_row(r[k], N[k]);
}

// One INPLACE step through LDS. Builds every own pair equation E(j) from the
// current rows; if any has an exactly zero pivot (C == 0), returns true
// without touching LDS (the caller then runs the redundant step, which keeps
// the reference's zero-pivot rule). Otherwise writes E(j) over row j, forms
// the new rows from E(i-d), E(i+d) and writes them back; returns false.
template <int NT, int RPT, typename ReadR>
__device__ __forceinline__ bool inplace_zero_pivot_step(Row<double> (&r)[RPT], int n, int delta,
                                                        double (&sR)[6][NT * RPT], ReadR&& readR)
{
    const int t = threadIdx.x;
    Pair<double> E[RPT];
    int zero = 0;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int j = t + NT * k;
        if (j < n) {
            const bool hasL = j - delta >= 0, hasR = j + delta <= n - 1;
            const bool aC = j - 2 * delta >= 0, arC = j + 2 * delta <= n - 1;
            const Row<double> L = readR(hasL ? j - delta : j);
            const Row<double> R = readR(hasR ? j + delta : j);
            E[k] = make_pair(L, r[k], R, hasL, aC, hasR, arC);
            zero |= (E[k].C == 0.0);
        }
    }
    if (__syncthreads_or(zero)) return true;   // uniform: every thread takes the same branch
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int j = t + NT * k;
        if (j < n) {
            sR[0][j] = E[k].A; sR[1][j] = E[k].B; sR[2][j] = E[k].C;
            sR[3][j] = E[k].D; sR[4][j] = E[k].E; sR[5][j] = E[k].K;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        if (i < n) {
            const bool hp1 = i - delta >= 0, hq1 = i + delta <= n - 1;
            const bool hp2 = i - 2 * delta >= 0, hq2 = i + 2 * delta <= n - 1;
            const int il = hp1 ? i - delta : i, ir = hq1 ? i + delta : i;
            Pair<double> EL, ER;
            EL.A = sR[0][il]; EL.B = sR[1][il]; EL.C = sR[2][il];
            EL.D = sR[3][il]; EL.E = sR[4][il]; EL.K = sR[5][il];
            ER.A = sR[0][ir]; ER.B = sR[1][ir]; ER.C = sR[2][ir];
            ER.D = sR[3][ir]; ER.E = sR[4][ir]; ER.K = sR[5][ir];
            // no pivot is zero in this step: the multiplier reuse never triggers
            r[k] = reduce_row(r[k], EL, ER, hp1, hq1, hp2, hq2,
                              [&](double& alL, double& arL, double& alR, double& arR) {
                                  alL = arL = alR = arR = 0.0;
                              });
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
    return false;
}

// Solve one system (or sub-system) held at ga..gy[base + i*rs], i < n, with
// `steps` reduction steps; x goes to gx[xbase + i*xrs]. All threads of the
// workgroup call this together (it contains barriers).
// This thread's rows i = t + NT*k of one system into registers.
template <int NT, int RPT>
__device__ __forceinline__ void lds_load_rows(const double* __restrict__ ga,
                                              const double* __restrict__ gb,
                                              const double* __restrict__ gc,
                                              const double* __restrict__ gd,
                                              const double* __restrict__ ge, const double* gy,
                                              long long base, long long rs, int n,
                                              Row<double> (&r)[RPT])
{
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        if (i < n) {
            const long long g = base + (long long)i * rs;
            r[k].a = ga[g]; r[k].b = gb[g]; r[k].c = gc[g];
            r[k].d = gd[g]; r[k].e = ge[g]; r[k].y = gy[g];
        } else {
            r[k] = Row<double>{0.0, 0.0, 1.0, 0.0, 0.0, 0.0};
        }
    }
}

template <int KIND, int NT, int RPT>
__device__ __forceinline__ void lds_solve_regs(
    Row<double> (&r)[RPT], double* gx, long long xbase, long long xrs, int n, int steps,
    double (&sR)[6][NT * RPT],
    double (&sE)[KIND == KIND_SHARED ? 8 : 1][KIND == KIND_SHARED ? NT * RPT : 1],
    const double* __restrict__ add = nullptr);

template <int KIND, int NT, int RPT>
__device__ __forceinline__ void lds_solve_one(
    const double* __restrict__ ga, const double* __restrict__ gb, const double* __restrict__ gc,
    const double* __restrict__ gd, const double* __restrict__ ge, const double* gy,
    double* gx, long long base, long long rs, long long xbase, long long xrs, int n,
    int steps, double (&sR)[6][NT * RPT],
    double (&sE)[KIND == KIND_SHARED ? 8 : 1][KIND == KIND_SHARED ? NT * RPT : 1],
    const double* __restrict__ add = nullptr)
{
    Row<double> r[RPT];
    lds_load_rows<NT, RPT>(ga, gb, gc, gd, ge, gy, base, rs, n, r);
    lds_solve_regs<KIND, NT, RPT>(r, gx, xbase, xrs, n, steps, sR, sE, add);
}

// The solve proper, from rows already in registers. store(i, x) receives the
// solution of row i (called for every row i < n exactly once, by some
// thread: a row's x and its final-stage partner's come from one thread).
template <int KIND, int NT, int RPT, typename Store>
__device__ __forceinline__ void lds_solve_regs_f(
    Row<double> (&r)[RPT], int n, int steps, double (&sR)[6][NT * RPT],
    double (&sE)[KIND == KIND_SHARED ? 8 : 1][KIND == KIND_SHARED ? NT * RPT : 1], Store&& store)
{
    const int t = threadIdx.x;

    auto writeR = [&](int i, const Row<double>& q) {
        sR[0][i] = q.a; sR[1][i] = q.b; sR[2][i] = q.c;
        sR[3][i] = q.d; sR[4][i] = q.e; sR[5][i] = q.y;
    };
    auto readR = [&](int i) {
        Row<double> q;
        q.a = sR[0][i]; q.b = sR[1][i]; q.c = sR[2][i];
        q.d = sR[3][i]; q.e = sR[4][i]; q.y = sR[5][i];
        return q;
    };

    // Steps with stride < NT exchange rows through LDS.
    int s = 0, delta = 1;
    const bool any_lds = steps > 0 && NT > 1;
    if constexpr (KIND == KIND_SPACED) {
        if (any_lds) {
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int i = t + NT * k;
                writeR(spaced_addr<NT, RPT>(i, 0), i < n ? r[k] : identity_row());
            }
            __syncthreads();
            for (; s < steps && delta < NT; ++s, delta <<= 1)
                spaced_step<NT, RPT, 1>(s, s + 1 == steps || 2 * delta >= NT, 0, n, sR);
            // back to one thread per row i = t + NT*k (rows in natural order now)
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int i = t + NT * k;
                r[k] = i < n ? readR(i) : identity_row();
            }
        }
    } else if (any_lds) {
#pragma unroll
        for (int k = 0; k < RPT; ++k)
            if (t + NT * k < n) writeR(t + NT * k, r[k]);
        __syncthreads();
    }
    for (; KIND != KIND_SPACED && s < steps && delta < NT; ++s, delta <<= 1) {
        if constexpr (KIND == KIND_SHARED) {
            // Reduction-I/II once per centre j (kernel.cu:69-117).
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int j = t + NT * k;
                if (j < n) {
                    const bool hasL = j - delta >= 0, hasR = j + delta <= n - 1;
                    const bool aC = j - 2 * delta >= 0, arC = j + 2 * delta <= n - 1;
                    const Row<double> L = readR(hasL ? j - delta : j);
                    const Row<double> R = readR(hasR ? j + delta : j);
                    const Pair<double> E = make_pair(L, r[k], R, hasL, aC, hasR, arC);
                    sE[0][j] = E.A; sE[1][j] = E.B; sE[2][j] = E.C; sE[3][j] = E.D;
                    sE[4][j] = E.E; sE[5][j] = E.K; sE[6][j] = E.al; sE[7][j] = E.ar;
                }
            }
            __syncthreads();
            // Reduction-III (kernel.cu:120-140).
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int i = t + NT * k;
                if (i < n) {
                    const bool hp1 = i - delta >= 0, hq1 = i + delta <= n - 1;
                    const bool hp2 = i - 2 * delta >= 0, hq2 = i + 2 * delta <= n - 1;
                    const int il = hp1 ? i - delta : i, ir = hq1 ? i + delta : i;
                    Pair<double> EL, ER;
                    EL.A = sE[0][il]; EL.B = sE[1][il]; EL.C = sE[2][il];
                    EL.D = sE[3][il]; EL.E = sE[4][il]; EL.K = sE[5][il];
                    ER.A = sE[0][ir]; ER.B = sE[1][ir]; ER.C = sE[2][ir];
                    ER.D = sE[3][ir]; ER.E = sE[4][ir]; ER.K = sE[5][ir];
                    r[k] = reduce_row(r[k], EL, ER, hp1, hq1, hp2, hq2,
                                      [&](double& alL, double& arL, double& alR, double& arR) {
                                          alL = sE[6][il]; arL = sE[7][il];
                                          alR = sE[6][ir]; arR = sE[7][ir];
                                      });
                    writeR(i, r[k]);
                }
            }
            __syncthreads();
        } else if (KIND == KIND_INPLACE && !inplace_zero_pivot_step<NT, RPT>(r, n, delta, sR, readR)) {
            // fast path done inside (pair equations over the rows, 4 barriers)
        } else {
            // kernel.cu:66-140 as written: each row builds eq1 and eq2 itself.
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int i = t + NT * k;
                if (i < n) {
                    const bool hp1 = i - delta >= 0, hq1 = i + delta <= n - 1;
                    const bool hp2 = i - 2 * delta >= 0, hq2 = i + 2 * delta <= n - 1;
                    const bool hp3 = i - 3 * delta >= 0, hq3 = i + 3 * delta <= n - 1;
                    const Row<double> P1 = readR(hp1 ? i - delta : i);
                    const Row<double> P2 = readR(hp2 ? i - 2 * delta : i);
                    const Row<double> Q1 = readR(hq1 ? i + delta : i);
                    const Row<double> Q2 = readR(hq2 ? i + 2 * delta : i);
                    // eq1 = E(i-d) from rows i-2d, i-d, i; eq2 = E(i+d) from rows i, i+d, i+2d
                    const Pair<double> EL = make_pair(P2, P1, r[k], hp2, hp3, true, hq1);
                    const Pair<double> ER = make_pair(r[k], Q1, Q2, true, hp1, hq2, hq3);
                    r[k] = reduce_row(r[k], EL, ER, hp1, hq1, hp2, hq2,
                                      [&](double& alL, double& arL, double& alR, double& arR) {
                                          alL = EL.al; arL = EL.ar; alR = ER.al; arR = ER.ar;
                                      });
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                if (t + NT * k < n) writeR(t + NT * k, r[k]);
            __syncthreads();
        }
    }
    // Steps with stride >= NT: all neighbours are rows of this thread.
#pragma unroll
    for (int dk = 1; dk < RPT; dk <<= 1) {
        if (s < steps) {
            if (dk == 1) register_step<NT, RPT, 1>(r, t, n);
            if constexpr (RPT > 2) if (dk == 2) register_step<NT, RPT, 2>(r, t, n);
            if constexpr (RPT > 4) if (dk == 4) register_step<NT, RPT, 4>(r, t, n);
            if constexpr (RPT > 8) if (dk == 8) register_step<NT, RPT, 8>(r, t, n);
            ++s;
            delta <<= 1;
        }
    }

    // Final 2x2 / 1x1 solve (kernel.cu:145-153).
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        if (i < n) {
            if (i < n - delta) {
                double cii, bii, yii;
                if (delta >= NT) {
                    // partner is row k + delta/NT of this thread (selects, no indexing)
                    const int kk = k + delta / NT;
                    cii = 0.0; bii = 0.0; yii = 0.0;
#pragma unroll
                    for (int q = 0; q < RPT; ++q) {
                        const bool hit = q == kk;
                        cii = hit ? r[q].c : cii;
                        bii = hit ? r[q].b : bii;
                        yii = hit ? r[q].y : yii;
                    }
                } else {
                    cii = sR[2][i + delta]; bii = sR[1][i + delta]; yii = sR[5][i + delta];
                }
                const double deno = __builtin_fma(r[k].c, cii, -(bii * r[k].d));
                const double x0 = __builtin_fma(cii, r[k].y, -(r[k].d * yii)) / deno;
                const double x1 = __builtin_fma(r[k].c, yii, -(bii * r[k].y)) / deno;
                store(i, x0);
                store(i + delta, x1);
            } else if (i < delta) {
                store(i, r[k].y / r[k].c);
            }
        }
    }
}

// x to gx[xbase + i*xrs]; add (optional): x + add[] is stored, at x's own
// element.
template <int KIND, int NT, int RPT>
__device__ __forceinline__ void lds_solve_regs(
    Row<double> (&r)[RPT], double* gx, long long xbase, long long xrs, int n, int steps,
    double (&sR)[6][NT * RPT],
    double (&sE)[KIND == KIND_SHARED ? 8 : 1][KIND == KIND_SHARED ? NT * RPT : 1],
    const double* __restrict__ add)
{
    lds_solve_regs_f<KIND, NT, RPT>(r, n, steps, sR, sE, [&](int i, double x) {
        const long long g = xbase + (long long)i * xrs;
        gx[g] = add ? x + add[g] : x;
    });
}

template <int KIND, int NT, int RPT>
__global__ __launch_bounds__(NT) void lds_f64_kernel(
    const double* __restrict__ ga, const double* __restrict__ gb, const double* __restrict__ gc,
    const double* __restrict__ gd, const double* __restrict__ ge, double* gy,
    int n, int m, long long rs, long long ss, int steps, SplitInfo split)
{
    constexpr int NMAX = NT * RPT;
    __shared__ double sR[6][NMAX];                               // rows a,b,c,d,e,y
    __shared__ double sE[KIND == KIND_SHARED ? 8 : 1][KIND == KIND_SHARED ? NMAX : 1];

    const int sys = xcd_remap(blockIdx.x, m);
    long long base = (long long)sys * ss;
    // where x goes: in place (or to split.x at y's element), or (split path)
    // back to the caller's layout
    double* gx = split.x ? split.x : gy;
    long long xbase = base, xrs = rs;
    if (split.S > 0) {
        // sub-system sys = R*m0 + j: rows R + 2^S*l of original system j, in
        // the grouped layout of the pass buffers (n: the longest sub-system)
        const int R = sys / split.m0, j = sys - R * split.m0;
        base = split_slot(sys, 0, m, n);
        rs = 1LL << split_group_log2(m);
        n = (split.n0 - R + (1 << split.S) - 1) >> split.S;
        gx = split.x;
        if (split.out_layout == 0) {   // contiguous j*n0 + i
            xbase = (long long)j * split.n0 + R;
            xrs = 1LL << split.S;
        } else {                       // interleaved i*m0 + j = sys + l*m
            xbase = sys;
            xrs = m;
        }
    }
    lds_solve_one<KIND, NT, RPT>(ga, gb, gc, gd, ge, gy, gx, base, rs, xbase, xrs, n, steps, sR, sE,
                                 split.S > 0 ? nullptr : split.add);
}

// Heston S-lines with the explicit stage folded in: system sys is S-line j
// (contiguous rows base + i), its right-hand side computed from the previous
// step's grid while the rows load, x + W stored to gx (never the grid read).
template <int KIND, int NT, int RPT>
__global__ __launch_bounds__(NT) void lds_f64_stage_kernel(
    const double* __restrict__ ga, const double* __restrict__ gb, const double* __restrict__ gc,
    const double* __restrict__ gd, const double* __restrict__ ge, double* gx, int n, int m,
    int steps, HestonStage src)
{
    constexpr int NMAX = NT * RPT;
    __shared__ double sR[6][NMAX];
    __shared__ double sE[KIND == KIND_SHARED ? 8 : 1][KIND == KIND_SHARED ? NMAX : 1];
    const int sys = xcd_remap(blockIdx.x, m);
    const long long base = (long long)sys * n;
    const auto rhs = src.bind(sys);
    const int t = threadIdx.x;
    Row<double> r[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        if (i < n) {
            const long long g = base + i;
            r[k].a = ga[g]; r[k].b = gb[g]; r[k].c = gc[g];
            r[k].d = gd[g]; r[k].e = ge[g]; r[k].y = rhs.r1(i, gx + g);   // W parked in x's element
        } else {
            r[k] = Row<double>{0.0, 0.0, 1.0, 0.0, 0.0, 0.0};
        }
    }
    lds_solve_regs_f<KIND, NT, RPT>(r, n, steps, sR, sE, [&](int i, double x) {
        gx[base + i] = x + gx[base + i];   // behind the solve's barriers
    });
}

// Persistent batch form: a resident grid walks the systems, and while it
// solves one system the rows of its next system are already being loaded
// into registers, so the strided HBM loads overlap the LDS steps. Virtual
// block ids b, b + G, b + 2G, ... keep b's XCD (G is a multiple of 8), and
// xcd_remap gives each XCD runs of consecutive systems as in the batch form.
template <int KIND, int NT, int RPT>
__global__ __launch_bounds__(NT) void lds_f64_persist_kernel(
    const double* __restrict__ ga, const double* __restrict__ gb, const double* __restrict__ gc,
    const double* __restrict__ gd, const double* __restrict__ ge, double* gy,
    int n, int m, long long rs, long long ss, int steps)
{
    constexpr int NMAX = NT * RPT;
    __shared__ double sR[6][NMAX];
    __shared__ double sE[KIND == KIND_SHARED ? 8 : 1][KIND == KIND_SHARED ? NMAX : 1];
    const int G = gridDim.x;
    int vb = blockIdx.x;
    if (vb >= m) return;
    Row<double> cur[RPT];
    int sys = xcd_remap(vb, m);
    lds_load_rows<NT, RPT>(ga, gb, gc, gd, ge, gy, (long long)sys * ss, rs, n, cur);
    for (;;) {
        const int nvb = vb + G;
        Row<double> nxt[RPT];
        int nsys = 0;
        if (nvb < m) {
            nsys = xcd_remap(nvb, m);
            lds_load_rows<NT, RPT>(ga, gb, gc, gd, ge, gy, (long long)nsys * ss, rs, n, nxt);
        }
        lds_solve_regs<KIND, NT, RPT>(cur, gy, (long long)sys * ss, rs, n, steps, sR, sE);
        if (nvb >= m) break;
        __syncthreads();   // LDS of this system fully consumed before the next one starts
        for (int k = 0; k < RPT; ++k) set_row(cur[k], nxt[k]);
        vb = nvb;
        sys = nsys;
    }
}

// Persistent work-queue form. Jobs (whole systems or residue sub-systems of
// split systems) are split into one queue per XCD; a workgroup drains the
// queue of the XCD it runs on (hardware register XCC_ID), then helps the
// others, so every wave reaches the exit once all queues are empty. Residue
// sub-systems of one system read row-interleaved groups that share cache
// lines, and the host puts runs of them into the same queue, i.e. behind the
// same L2. src picks the array set a job reads (the caller's arrays or a pass
// buffer); x goes to the job's output indices (the system's own rows).
template <int KIND, int NT, int RPT>
__global__ __launch_bounds__(NT) void lds_f64_queue_kernel(
    const double* const* __restrict__ sets, double* __restrict__ gx, const LdsJob* __restrict__ jobs,
    const int* __restrict__ qstart, unsigned int* __restrict__ counters)
{
    constexpr int NMAX = NT * RPT;
    __shared__ double sR[6][NMAX];
    __shared__ double sE[KIND == KIND_SHARED ? 8 : 1][KIND == KIND_SHARED ? NMAX : 1];
    __shared__ int sJob;
    const int home = (int)(__builtin_amdgcn_s_getreg((3 << 11) | 20) & (kQueues - 1));
    for (int t = 0; t < kQueues; ++t) {
        const int qx = (home + t) & (kQueues - 1);
        const int begin = qstart[qx], len = qstart[qx + 1] - begin;
        for (;;) {
            if (threadIdx.x == 0) sJob = (int)atomicAdd(&counters[qx], 1u);
            __syncthreads();
            const int idx = sJob;
            __syncthreads();
            if (idx >= len) break;
            const LdsJob J = jobs[begin + idx];
            const double* const* S = sets + 6 * J.src;
            // the spaced form works through every row of its tile, so the
            // spaced queue takes it for jobs that fill most of the tile and
            // the redundant form (same LDS) for the rest (uniform per job)
            bool redundant = false;
            if constexpr (KIND == KIND_SPACED) redundant = J.n * 32 <= NT * RPT * 23;
            if (redundant)
                lds_solve_one<KIND == KIND_SPACED ? KIND_REDUNDANT : KIND, NT, RPT>(
                    S[0], S[1], S[2], S[3], S[4], S[5], gx, J.base, 1LL << J.stride_log2, J.xbase,
                    1LL << J.xstride_log2, J.n, J.steps, sR, sE);
            else
                lds_solve_one<KIND, NT, RPT>(S[0], S[1], S[2], S[3], S[4], S[5], gx, J.base,
                                             1LL << J.stride_log2, J.xbase, 1LL << J.xstride_log2,
                                             J.n, J.steps, sR, sE);
            __syncthreads();
        }
    }
}

// Two neighbouring systems j0, j0 + 1 of an interleaved batch per workgroup
// (in place, m even, 16-byte aligned arrays). In the interleaved layout one
// system's rows sit one per 128-byte line, so the load phase runs at the L1
// line rate; here every line fetched serves both systems (16-byte loads and
// stores). The workgroup solves them one after the other: the second system's
// rows wait in registers, and each x goes through LDS column 0 (the final
// stage writes partner rows' x from other threads and reads only columns 1, 2
// and 5) to the thread that stores the pair.
template <int KIND, int NT, int RPT>
__global__ __launch_bounds__(NT) void lds_f64_pair_kernel(
    const double* __restrict__ ga, const double* __restrict__ gb, const double* __restrict__ gc,
    const double* __restrict__ gd, const double* __restrict__ ge, double* gy, int n, int m,
    int steps)
{
    constexpr int NMAX = NT * RPT;
    __shared__ double sR[6][NMAX];
    __shared__ double sE[KIND == KIND_SHARED ? 8 : 1][KIND == KIND_SHARED ? NMAX : 1];
    const int t = threadIdx.x;
    // 8 pairs = 16 systems = one 128-byte line of each row, kept on one XCD
    const long long j0 = 2LL * xcd_remap_c<8>((int)blockIdx.x, m >> 1);
    Row<double> rA[RPT], rB[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        if (i < n) {
            const long long g = (long long)i * m + j0;
            const double2 va = *reinterpret_cast<const double2*>(ga + g);
            const double2 vb = *reinterpret_cast<const double2*>(gb + g);
            const double2 vc = *reinterpret_cast<const double2*>(gc + g);
            const double2 vd = *reinterpret_cast<const double2*>(gd + g);
            const double2 ve = *reinterpret_cast<const double2*>(ge + g);
            const double2 vy = *reinterpret_cast<const double2*>(gy + g);
            rA[k] = Row<double>{va.x, vb.x, vc.x, vd.x, ve.x, vy.x};
            rB[k] = Row<double>{va.y, vb.y, vc.y, vd.y, ve.y, vy.y};
        } else {
            rA[k] = Row<double>{0.0, 0.0, 1.0, 0.0, 0.0, 0.0};
            rB[k] = rA[k];
        }
    }
    const auto to_lds = [&](int i, double x) { sR[0][i] = x; };
    lds_solve_regs_f<KIND, NT, RPT>(rA, n, steps, sR, sE, to_lds);
    __syncthreads();
    double xA[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) xA[k] = t + NT * k < n ? sR[0][t + NT * k] : 0.0;
    __syncthreads();   // every xA read before the second system's rows go to LDS
    lds_solve_regs_f<KIND, NT, RPT>(rB, n, steps, sR, sE, to_lds);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        if (i < n)
            *reinterpret_cast<double2*>(gy + (long long)i * m + j0) = make_double2(xA[k], sR[0][i]);
    }
}

// ---- variant table ---------------------------------------------------------

template <int KIND, int NT, int RPT>
static int launch(const SolveArgsF64& A, hipStream_t st)
{
    hipLaunchKernelGGL((lds_f64_kernel<KIND, NT, RPT>), dim3(A.m), dim3(NT), 0, st,
                       A.a, A.b, A.c, A.d, A.e, A.y, A.n, A.m, A.rs, A.ss, A.steps, A.split);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Persistent launch: one workgroup per resident slot (CUs x blocks per CU).
template <int KIND, int NT, int RPT>
static int launch_queue(const double* const* sets, double* x, const LdsJob* jobs, int njobs,
                        const int* qstart, unsigned int* counters, hipStream_t st)
{
    static std::atomic<int> grids[kMaxDevices];
    const int grid = device_cached(grids, [](int) {
        const int cus = device_cu_count();
        int per = 0;
        if (cus <= 0 ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per, reinterpret_cast<const void*>(&lds_f64_queue_kernel<KIND, NT, RPT>), NT, 0) !=
                hipSuccess)
            return -1;
        return cus * (per > 0 ? per : 1);
    });
    if (grid <= 0) return -1;
    const int g = njobs < grid ? njobs : grid;
    if (g <= 0) return 0;
    hipLaunchKernelGGL((lds_f64_queue_kernel<KIND, NT, RPT>), dim3(g), dim3(NT), 0, st, sets, x,
                       jobs, qstart, counters);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Workgroups of one kernel that fit on a gfx950 CU at once: 4 SIMDs with 512
// VGPRs per lane each (allocated in granules of 8), at most 8 waves per SIMD,
// 160 KB of LDS.
static int resident_blocks_per_cu(int nt, int vgprs, int lds_bytes)
{
    const int waves = (nt + 63) / 64;
    const int v = vgprs < 8 ? 8 : ((vgprs + 7) / 8) * 8;
    int per_simd = 512 / v;
    if (per_simd > 8) per_simd = 8;
    int by_regs = (4 * per_simd) / waves;
    int by_lds = lds_bytes > 0 ? (160 * 1024) / lds_bytes : 32;
    int b = by_regs < by_lds ? by_regs : by_lds;
    return b < 1 ? 1 : b;
}

template <int KIND, int NT, int RPT>
static int launch_persist(const SolveArgsF64& A, hipStream_t st)
{
    static std::atomic<int> grids[kMaxDevices];
    const int grid = device_cached(grids, [](int) {
        const int cus = device_cu_count();
        hipFuncAttributes fa;
        if (cus <= 0 || hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(
                                                      &lds_f64_persist_kernel<KIND, NT, RPT>)) !=
                            hipSuccess)
            return -1;
        const int g = (cus * resident_blocks_per_cu(NT, fa.numRegs, (int)fa.sharedSizeBytes) / 8) * 8;
        return g < 8 ? 8 : g;
    });
    if (grid <= 0) return -1;
    // split views (SplitInfo) use the batch form; this form covers plain batches
    if (A.split.S > 0) return launch<KIND, NT, RPT>(A, st);
    const int g = A.m < grid ? A.m : grid;
    hipLaunchKernelGGL((lds_f64_persist_kernel<KIND, NT, RPT>), dim3(g), dim3(NT), 0, st, A.a, A.b,
                       A.c, A.d, A.e, A.y, A.n, A.m, A.rs, A.ss, A.steps);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int lds_f64_plain_variant_for_batch(int n, int m);

// Paired form: interleaved in-place batches of even m with 16-byte aligned
// arrays; every other batch (and split views) takes the one-system variant
// this batch would get by default.
template <int KIND, int NT, int RPT>
static int launch_pair(const SolveArgsF64& A, hipStream_t st)
{
    const auto al16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
    const bool pairable = A.split.S == 0 && A.split.x == A.y && A.split.add == nullptr &&
                          A.ss == 1 && A.rs == A.m && (A.m & 1) == 0 && al16(A.a) && al16(A.b) &&
                          al16(A.c) && al16(A.d) && al16(A.e) && al16(A.y);
    if (!pairable) {
        const LdsVariantF64* v = lds_f64_variant(lds_f64_plain_variant_for_batch(A.n, A.m));
        return v ? v->launch(A, st) : -1;
    }
    hipLaunchKernelGGL((lds_f64_pair_kernel<KIND, NT, RPT>), dim3(A.m / 2), dim3(NT), 0, st, A.a,
                       A.b, A.c, A.d, A.e, A.y, A.n, A.m, A.steps);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

#define PCRP_2(KIND, NT, RPT) \
    {KIND + 8, NT, RPT, NT * RPT, &launch_pair<KIND, NT, RPT>, &launch_queue<KIND, NT, RPT>}

#define PCRP_P(KIND, NT, RPT) \
    {KIND + 4, NT, RPT, NT * RPT, &launch_persist<KIND, NT, RPT>, &launch_queue<KIND, NT, RPT>}

#define PCRP_V(KIND, NT, RPT) \
    {KIND, NT, RPT, NT * RPT, &launch<KIND, NT, RPT>, &launch_queue<KIND, NT, RPT>}

static const LdsVariantF64 kVariants[] = {
    // shared pair equations, 14 doubles of LDS per row
    PCRP_V(KIND_SHARED, 64, 1),   PCRP_V(KIND_SHARED, 128, 1),  PCRP_V(KIND_SHARED, 256, 1),
    PCRP_V(KIND_SHARED, 512, 1),  PCRP_V(KIND_SHARED, 1024, 1), PCRP_V(KIND_SHARED, 64, 2),
    PCRP_V(KIND_SHARED, 128, 2),  PCRP_V(KIND_SHARED, 256, 2),  PCRP_V(KIND_SHARED, 512, 2),
    PCRP_V(KIND_SHARED, 64, 4),   PCRP_V(KIND_SHARED, 128, 4),  PCRP_V(KIND_SHARED, 256, 4),
    // reference-style redundant pairs, 6 doubles of LDS per row
    PCRP_V(KIND_REDUNDANT, 64, 1),   PCRP_V(KIND_REDUNDANT, 128, 1),
    PCRP_V(KIND_REDUNDANT, 128, 8),
    PCRP_V(KIND_REDUNDANT, 256, 1),  PCRP_V(KIND_REDUNDANT, 512, 1),
    PCRP_V(KIND_REDUNDANT, 1024, 1), PCRP_V(KIND_REDUNDANT, 64, 2),
    PCRP_V(KIND_REDUNDANT, 128, 2),  PCRP_V(KIND_REDUNDANT, 256, 2),
    PCRP_V(KIND_REDUNDANT, 512, 2),  PCRP_V(KIND_REDUNDANT, 64, 4),
    PCRP_V(KIND_REDUNDANT, 128, 4),  PCRP_V(KIND_REDUNDANT, 256, 4),
    // pair equations written over the rows, 6 doubles of LDS per row
    PCRP_V(KIND_INPLACE, 64, 1),    PCRP_V(KIND_INPLACE, 128, 1),  PCRP_V(KIND_INPLACE, 256, 1),
    PCRP_V(KIND_INPLACE, 512, 1),   PCRP_V(KIND_INPLACE, 1024, 1), PCRP_V(KIND_INPLACE, 128, 2),
    PCRP_V(KIND_INPLACE, 256, 2),   PCRP_V(KIND_INPLACE, 512, 2),  PCRP_V(KIND_INPLACE, 64, 4),
    PCRP_V(KIND_INPLACE, 128, 4),   PCRP_V(KIND_INPLACE, 256, 4),  PCRP_V(KIND_INPLACE, 128, 8),
    // persistent batch form with next-system prefetch (kind code + 4)
    PCRP_P(KIND_REDUNDANT, 256, 4), PCRP_P(KIND_REDUNDANT, 1024, 1), PCRP_P(KIND_REDUNDANT, 256, 2),
    PCRP_P(KIND_REDUNDANT, 256, 1), PCRP_P(KIND_REDUNDANT, 128, 4), PCRP_P(KIND_SHARED, 512, 1),
    // two interleaved systems per workgroup (kind code + 8)
    PCRP_2(KIND_REDUNDANT, 256, 4), PCRP_2(KIND_REDUNDANT, 256, 2), PCRP_2(KIND_REDUNDANT, 128, 4),
    // rows two strides apart per thread, rows in LDS between steps
    PCRP_V(KIND_SPACED, 256, 4),    PCRP_V(KIND_SPACED, 128, 8),    PCRP_V(KIND_SPACED, 512, 2),
    PCRP_V(KIND_SPACED, 256, 2),    PCRP_V(KIND_SPACED, 128, 4),    PCRP_V(KIND_SPACED, 64, 4),
    PCRP_2(KIND_SPACED, 256, 4),    PCRP_2(KIND_SPACED, 128, 4),
};

#undef PCRP_V
#undef PCRP_P
#undef PCRP_2

template <int KIND, int NT, int RPT>
static int launch_stage(const double* a, const double* b, const double* c, const double* d,
                        const double* e, double* x, int n, int m, const HestonStage& src,
                        hipStream_t st)
{
    hipLaunchKernelGGL((lds_f64_stage_kernel<KIND, NT, RPT>), dim3(m), dim3(NT), 0, st, a, b, c,
                       d, e, x, n, m, pcr_steps(n), src);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Shapes as the batch defaults for these sizes: the redundant form up to
// n = 512, spaced 256 x 4 above.
int lds_f64_solve_stage(const double* a, const double* b, const double* c, const double* d,
                        const double* e, double* x, int n, int m, const HestonStage& src,
                        hipStream_t st)
{
    if (n <= 256) return launch_stage<KIND_REDUNDANT, 256, 1>(a, b, c, d, e, x, n, m, src, st);
    if (n <= 512) return launch_stage<KIND_REDUNDANT, 256, 2>(a, b, c, d, e, x, n, m, src, st);
    if (n <= 1024) return launch_stage<KIND_SPACED, 256, 4>(a, b, c, d, e, x, n, m, src, st);
    return -2;
}

int lds_f64_variant_count() { return (int)(sizeof(kVariants) / sizeof(kVariants[0])); }

const LdsVariantF64* lds_f64_variant(int id)
{
    return (id >= 0 && id < lds_f64_variant_count()) ? &kVariants[id] : nullptr;
}

int lds_f64_max_n() { return 1024; }

int lds_f64_lds_bytes(const LdsVariantF64& v)
{
    return 8 * v.max_n * ((v.kind & 3) == KIND_SHARED ? LdsWords<KIND_SHARED>::value
                                                : LdsWords<KIND_REDUNDANT>::value);   // INPLACE: 6 too
}

static int find_variant(int kind, int nt, int rpt)
{
    for (int id = 0; id < lds_f64_variant_count(); ++id) {
        const LdsVariantF64& v = kVariants[id];
        if (v.kind == kind && v.nt == nt && v.rpt == rpt) return id;
    }
    return -1;
}

// The redundant form with 256 threads per workgroup and as many rows per
// thread as n needs (tightest capacity first, then the workgroup size closest
// to 256).
static int redundant_default_variant(int n)
{
    int best = -1, best_score = 1 << 30;
    for (int id = 0; id < lds_f64_variant_count(); ++id) {
        const LdsVariantF64& v = kVariants[id];
        if (v.kind != KIND_REDUNDANT || v.max_n < n) continue;
        const int score = v.max_n * 4 + (v.nt > 256 ? v.nt / 256 : 256 / v.nt);
        if (score < best_score) { best = id; best_score = score; }
    }
    return best;
}

// Default for systems of n rows, batch size unknown: the spaced form when n
// fills most of 256 x 4 (past n it skips a thread's trailing rows, but every
// workgroup still walks all strides), else the redundant form above.
int lds_f64_default_variant(int n)
{
    if (n > 736 && n <= 1024) {
        const int id = find_variant(KIND_SPACED, 256, 4);
        if (id >= 0) return id;
    }
    return redundant_default_variant(n);
}

// Default of the uneven work queue, whose jobs have every size up to its
// capacity n: the spaced queue kernel, which runs the spaced form on the jobs
// that fill most of its tile and the redundant form on the others (MI355X,
// uneven batch of 10000 systems: 18.4-18.6 vs 18.7-18.8 ms per solve with
// the redundant form on every job; the spaced form on every job was even
// with it, as the small jobs lost what the large ones gained).
int lds_f64_queue_default_variant(int n)
{
    return lds_f64_default_variant(n);
}

// One-system default for a whole batch of m systems (MI355X, FP64, every
// candidate warmed and timed interleaved):
//   * 736 < n <= 1024: spaced 256 x 4 from m = 2 x CUs on (n = 1024, m = 16384:
//     0.846 vs 0.935 ms interleaved, 0.680 vs 0.814 ms contiguous; m = 512:
//     0.039 vs 0.042 ms; n = 768: 0.60-0.69 vs 0.61-0.71 ms; n = 700 is 2-3%
//     slower); smaller batches are latency-bound and take one row per thread
//     (1024 threads);
//   * 512 < n <= 736, m <= 2 x CUs: one row per thread as well (5-10% faster
//     than 256 x 4 up to m = 512);
//   * 448 < n <= 512: spaced 128 x 4 from m = 16 x CUs (contiguous m = 16384:
//     0.313 vs 0.372 ms), spaced 256 x 2 below (m = 2048: 0.051-0.057 vs
//     0.053-0.061 ms);
//   * 224 < n <= 256, m >= 8 x CUs: spaced 64 x 4 (m = 16384: 0.137 vs 0.172
//     ms contiguous, 0.169 vs 0.185 ms interleaved).
// Sizes in between fill the spaced tiles too little to gain.
int lds_f64_plain_variant_for_batch(int n, int m)
{
    const int cus = device_cu_count();
    int id = -1;
    if (cus > 0) {
        if (n > 736 && n <= 1024)
            id = m >= 2 * cus ? find_variant(KIND_SPACED, 256, 4) : find_variant(KIND_REDUNDANT, 1024, 1);
        else if (n > 512 && n <= 736 && m <= 2 * cus)
            id = find_variant(KIND_REDUNDANT, 1024, 1);
        else if (n > 448 && n <= 512)
            id = m >= 16 * cus ? find_variant(KIND_SPACED, 128, 4) : find_variant(KIND_SPACED, 256, 2);
        else if (n > 224 && n <= 256 && m >= 8 * cus)
            id = find_variant(KIND_SPACED, 64, 4);
    }
    return id >= 0 ? id : lds_f64_default_variant(n);
}

// Default for a whole batch: large interleaved batches of 256 < n <= 832 take
// the redundant paired form and 448 < n <= 512 the spaced paired form (two
// systems per workgroup share every 128-byte line of the load; a batch that
// cannot pair runs the one-system default above). MI355X, m = 16384,
// interleaved: n = 512 0.397 ms paired-spaced 128 x 4 vs 0.415 spaced 128 x 4;
// n = 640-768 0.60-0.68 ms paired-redundant vs 0.62-0.71 one system; from
// n = 896 the spaced one-system form is faster (0.77-0.85 vs 0.80-0.89 ms).
int lds_f64_default_variant_for_batch(int n, int m)
{
    const int cus = device_cu_count();
    if (cus > 0 && m >= 16 * cus && n > 256 && n <= 832) {
        const int id = n > 512 ? find_variant(KIND_REDUNDANT + 8, 256, 4)
                               : (n > 448 ? find_variant(KIND_SPACED + 8, 128, 4)
                                          : find_variant(KIND_REDUNDANT + 8, 128, 4));
        if (id >= 0) return id;
    }
    return lds_f64_plain_variant_for_batch(n, m);
}

}  // namespace pcrp
