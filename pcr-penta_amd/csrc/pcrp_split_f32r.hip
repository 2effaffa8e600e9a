// FP32 inner solve + FP64 iterative refinement for systems too large for one
// workgroup (n above the LDS FP32 kernel's limit).
//
// No counterpart in the reference, which is FP64-only. The inner solve is
// PCR on the 2x2-block tridiagonal form of the matrix (block row k = rows 2k,
// 2k+1; L_k X_{k-1} + D_k X_k + U_k X_{k+1} = Y_k, diagonal-block pivots: the
// form the one-workgroup FP32 kernel uses, for the same accuracy reason),
// split the way the reference splits PCR-Penta (stage 1 in HBM,
// kernel.cu:158-320; strided sub-systems in shared memory, kernel.cu:665-791):
//   passes : S levels in HBM, k_fuse levels per pass over tiles of TB block
//            rows of G neighbouring sub-systems, halo 2^k - 1 block rows per
//            side; output regrouped by residue as in the FP64 passes, so
//            sub-system (R, j) of block rows R + 2^S l lands at q = R*m + j;
//   stage  : one workgroup per sub-system, the remaining levels in LDS, then
//            X = D^-1 Y, and x += dx (FP64) at the caller's elements.
// One refinement sweep (every kernel skips systems already stopped):
//   residual : r = y - A x in FP64 from the FP64 inputs, stored as the FP32
//              block right-hand side (first sweep: r = y, x = 0);
//   passes + stage (factor and solve together: recomputing the FP32 factor
//              moves fewer bytes than storing its multipliers);
//   decide   : the stopping rule of the one-workgroup kernel: stop when
//              max|dx| <= tol max|x|, or when the observed contraction
//              rho = |dx_k| / |dx_{k-1}| (from the second correction on)
//              puts the error left, 10 rho |dx_k|, below tol with rho <= 0.1.
//              A system that is not done after max_sweeps corrections, or
//              whose correction is NaN, reports max_sweeps + 1.
// The sweep loop is unrolled on the host (max_sweeps + 1 solves); the stop
// flags make the launches of converged systems return at once, so the call
// stays asynchronous (no host read-back) and graph-capturable.
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

// Block row of the FP32 inner solve.
struct BRow {
    M2 D, L, U;
    float y0, y1;
};

constexpr int kWords = 14;            // floats per block row
constexpr int kStageNT = 256, kStageBR = 4;
constexpr int kStageMax = kStageNT * kStageBR;   // block rows per LDS sub-system
constexpr int kFuse = 4;              // levels per HBM pass
constexpr int kPassTB = 256, kPassG = 4, kPassRPT = 4;

// LDS record of one block row during a level: D^-1, L, U, Y.
template <int N>
__device__ __forceinline__ void put_rec(float (&s)[kWords][N], int i, const M2& Di, const BRow& r)
{
    s[0][i] = Di.p; s[1][i] = Di.q; s[2][i] = Di.r; s[3][i] = Di.s;
    s[4][i] = r.L.p; s[5][i] = r.L.q; s[6][i] = r.L.r; s[7][i] = r.L.s;
    s[8][i] = r.U.p; s[9][i] = r.U.q; s[10][i] = r.U.r; s[11][i] = r.U.s;
    s[12][i] = r.y0; s[13][i] = r.y1;
}

// One level for block row r from the records of its neighbours il / ir
// (hl / hr: the neighbour exists in the sub-system).
template <int N>
__device__ __forceinline__ BRow level(const BRow& r, const float (&s)[kWords][N], bool hl, int il,
                                      bool hr, int ir)
{
    const M2 DiL{s[0][il], s[1][il], s[2][il], s[3][il]};
    const M2 LL{s[4][il], s[5][il], s[6][il], s[7][il]};
    const M2 UL{s[8][il], s[9][il], s[10][il], s[11][il]};
    const M2 DiR{s[0][ir], s[1][ir], s[2][ir], s[3][ir]};
    const M2 LR{s[4][ir], s[5][ir], s[6][ir], s[7][ir]};
    const M2 UR{s[8][ir], s[9][ir], s[10][ir], s[11][ir]};
    const float yl0 = s[12][il], yl1 = s[13][il], yr0 = s[12][ir], yr1 = s[13][ir];
    const M2 al = hl ? mmul_neg(r.L, DiL) : kZero2;
    const M2 be = hr ? mmul_neg(r.U, DiR) : kZero2;
    BRow o;
    o.D = madd2(r.D, al, UL, be, LR);
    o.L = mmul(al, LL);
    o.U = mmul(be, UR);
    o.y0 = fmaf(be.p, yr0, fmaf(be.q, yr1, fmaf(al.p, yl0, fmaf(al.q, yl1, r.y0))));
    o.y1 = fmaf(be.r, yr0, fmaf(be.s, yr1, fmaf(al.r, yl0, fmaf(al.s, yl1, r.y1))));
    return o;
}

__device__ __forceinline__ BRow load_brow(const float* const* w, long long g)
{
    BRow r;
    r.D = M2{w[0][g], w[1][g], w[2][g], w[3][g]};
    r.L = M2{w[4][g], w[5][g], w[6][g], w[7][g]};
    r.U = M2{w[8][g], w[9][g], w[10][g], w[11][g]};
    r.y0 = w[12][g];
    r.y1 = w[13][g];
    return r;
}

__device__ __forceinline__ void store_brow(float* const* w, long long g, const BRow& r)
{
    w[0][g] = r.D.p; w[1][g] = r.D.q; w[2][g] = r.D.r; w[3][g] = r.D.s;
    w[4][g] = r.L.p; w[5][g] = r.L.q; w[6][g] = r.L.r; w[7][g] = r.L.s;
    w[8][g] = r.U.p; w[9][g] = r.U.q; w[10][g] = r.U.r; w[11][g] = r.U.s;
    w[12][g] = r.y0; w[13][g] = r.y1;
}

struct Arrays14 {
    float* p[kWords];
};

struct Layout {
    size_t state, stop, y0, lvl0, sets, bytes;
    size_t per;        // elements per pass array
    long long nb0;     // block rows per system
    int S;
};

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

Layout plan(int n, int m)
{
    Layout L{};
    L.nb0 = (n + 1) / 2;
    L.S = 0;
    while (((L.nb0 + (1LL << L.S) - 1) >> L.S) > kStageMax) ++L.S;
    const size_t N = (size_t)n * m;
    const size_t NB = (size_t)L.nb0 * m;
    L.per = (size_t)m * ((size_t)((L.nb0 + (1LL << L.S) - 1) >> L.S) << L.S);
    size_t o = 0;
    L.state = o; o = align_up(o + (size_t)m * sizeof(SysState));
    L.stop = o;  o = align_up(o + (size_t)m * sizeof(int));
    L.y0 = o;    o = align_up(o + N * sizeof(double));
    L.lvl0 = o;  o = align_up(o + kWords * NB * sizeof(float));
    L.sets = o;  o = align_up(o + 2 * kWords * L.per * sizeof(float));
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

// Level-0 block rows (block row K of system j at K*m + j) from the FP64
// rows 2K, 2K+1 (entries outside the band zeroed; a missing row 2K+1 of an
// odd n is the identity row); y0 keeps y.
__global__ void prep_kernel(const double* __restrict__ a, const double* __restrict__ b,
                            const double* __restrict__ c, const double* __restrict__ d,
                            const double* __restrict__ e, const double* __restrict__ y,
                            Arrays14 w, double* __restrict__ y0, int n, int m, long long rs,
                            long long ss, long long nb0)
{
    const long long el = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (el >= nb0 * m) return;
    const long long K = el / m;
    const int j = (int)(el - K * m);
    double rw[2][6];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const long long i = 2 * K + h;
        rw[h][0] = 0.0; rw[h][1] = 0.0; rw[h][2] = 1.0; rw[h][3] = 0.0; rw[h][4] = 0.0;
        rw[h][5] = 0.0;
        if (i < n) {
            const long long g = (long long)j * ss + i * rs;
            rw[h][0] = i >= 2 ? a[g] : 0.0;
            rw[h][1] = i >= 1 ? b[g] : 0.0;
            rw[h][2] = c[g];
            rw[h][3] = i + 1 < n ? d[g] : 0.0;
            rw[h][4] = i + 2 < n ? e[g] : 0.0;
            rw[h][5] = y[g];
            y0[g] = rw[h][5];
        }
    }
    BRow r;
    r.L = M2{(float)rw[0][0], (float)rw[0][1], 0.f, (float)rw[1][0]};
    r.D = M2{(float)rw[0][2], (float)rw[0][3], (float)rw[1][1], (float)rw[1][2]};
    r.U = M2{(float)rw[0][4], 0.f, (float)rw[1][3], (float)rw[1][4]};
    r.y0 = (float)rw[0][5];
    r.y1 = (float)rw[1][5];
    store_brow(w.p, el, r);
}

// r = y0 - A x in FP64 (operation order of the one-workgroup kernel), as the
// FP32 right-hand side of level-0 block row K.
__global__ void residual_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                const double* __restrict__ c, const double* __restrict__ d,
                                const double* __restrict__ e, const double* __restrict__ y0,
                                const double* __restrict__ x, float* __restrict__ r0,
                                float* __restrict__ r1, const int* __restrict__ stop, int n, int m,
                                long long rs, long long ss, long long nb0)
{
    const long long el = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (el >= nb0 * m) return;
    const long long K = el / m;
    const int j = (int)(el - K * m);
    if (stop[j]) return;
    float rv[2] = {0.f, 0.f};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const long long i = 2 * K + h;
        if (i < n) {
            const long long g = (long long)j * ss + i * rs;
            const double xm2 = i >= 2 ? x[g - 2 * rs] : 0.0, xm1 = i >= 1 ? x[g - rs] : 0.0;
            const double xp1 = i + 1 < n ? x[g + rs] : 0.0, xp2 = i + 2 < n ? x[g + 2 * rs] : 0.0;
            double ax = c[g] * x[g];
            ax = __builtin_fma(a[g], xm2, ax);
            ax = __builtin_fma(b[g], xm1, ax);
            ax = __builtin_fma(d[g], xp1, ax);
            ax = __builtin_fma(e[g], xp2, ax);
            rv[h] = (float)(y0[g] - ax);
        }
    }
    r0[el] = rv[0];
    r1[el] = rv[1];
}

// Block -> (group, tile), tiles striped over the 8 XCDs with the group index
// fastest (the sub-systems of a group share cache lines).
__device__ __forceinline__ bool tile_of(int b, int groups, int tiles, int& grp, int& tile)
{
    constexpr int NX = 8;
    const int xcd = b % NX, s = b / NX;
    const int tl = s / groups;
    grp = s - tl * groups;
    tile = tl * NX + xcd;
    return tile < tiles;
}

// One HBM pass: k levels over a tile of TB block rows of G sub-systems.
// Current sub-system q = R*m0 + j holds block rows R + 2^Sin l of system j,
// element l at q + l*Qin; output regrouped by residue (q' = (l mod 2^k)*Qin
// + q at l >> k).
template <int TB, int G, int RPT>
__global__ __launch_bounds__(TB* G / RPT) void bpass_kernel(
    Arrays14 in, Arrays14 out, long long nb0, int m0, int Sin, int k, int Qin, int tiles,
    const int* __restrict__ stop)
{
    constexpr int NTH = TB * G / RPT, RSTEP = TB / RPT;
    __shared__ float sR[kWords][TB * G];
    const int H = (1 << k) - 1;
    const int TI = TB - 2 * H;
    int grp, tile;
    if (!tile_of(blockIdx.x, (Qin + G - 1) / G, tiles, grp, tile)) return;
    const int tid = threadIdx.x, lr0 = tid / G, g = tid - lr0 * G;
    const int q = grp * G + g;
    const int R = q / m0;
    const int Lq = (int)((nb0 - R + (1LL << Sin) - 1) >> Sin);
    const bool run = q < Qin && !stop[q - R * m0];
    const long long Qout = (long long)Qin << k;
    const int l0 = tile * TI - H + lr0;
    BRow r[RPT];
    bool valid[RPT];
#pragma unroll
    for (int kk = 0; kk < RPT; ++kk) {
        const int l = l0 + RSTEP * kk;
        valid[kk] = run && l >= 0 && l < Lq;
        if (valid[kk]) {
            r[kk] = load_brow(in.p, (long long)q + (long long)l * Qin);
        } else {
            r[kk].D = kEye2;
            r[kk].L = kZero2;
            r[kk].U = kZero2;
            r[kk].y0 = r[kk].y1 = 0.f;
        }
    }
    for (int s = 0; s < k; ++s) {
        const int delta = 1 << s;
#pragma unroll
        for (int kk = 0; kk < RPT; ++kk) put_rec(sR, tid + NTH * kk, minv(r[kk].D), r[kk]);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < RPT; ++kk) {
            if (valid[kk]) {
                const int l = l0 + RSTEP * kk, lr = lr0 + RSTEP * kk, self = tid + NTH * kk;
                const bool hl = l - delta >= 0, hr = l + delta < Lq;
                // neighbours outside the tile only feed block rows that are not kept
                const int il = (hl && lr - delta >= 0) ? (lr - delta) * G + g : self;
                const int ir = (hr && lr + delta < TB) ? (lr + delta) * G + g : self;
                r[kk] = level(r[kk], sR, hl, il, hr, ir);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int kk = 0; kk < RPT; ++kk) {
        const int l = l0 + RSTEP * kk, lr = lr0 + RSTEP * kk;
        if (valid[kk] && lr >= H && lr < H + TI) {
            const long long go =
                (long long)(l >> k) * Qout + (long long)(l & ((1 << k) - 1)) * Qin + q;
            store_brow(out.p, go, r[kk]);
        }
    }
}

// LDS stage: one workgroup per sub-system q = R*m0 + j (element l at q + l*Q),
// the remaining levels, X = D^-1 Y, x += dx at rows 2K, 2K+1, K = R + 2^S l.
template <int NT, int BR>
__global__ __launch_bounds__(NT) void bstage_kernel(
    Arrays14 in, int Q, int S, long long nb0, int n0, int m0, long long xrs, long long xss,
    double* __restrict__ x, SysState* __restrict__ st, const int* __restrict__ stop, int first)
{
    constexpr int NB = NT * BR;
    __shared__ float sR[kWords][NB];
    __shared__ unsigned long long red[2][NT / 64];
    const int q = xcd_remap(blockIdx.x, Q);
    const int R = q / m0, j = q - R * m0;
    if (stop[j]) return;   // uniform over the workgroup
    const int nb = (int)((nb0 - R + (1LL << S) - 1) >> S);
    const int t = threadIdx.x;
    BRow r[BR];
#pragma unroll
    for (int b = 0; b < BR; ++b) {
        const int l = t + NT * b;
        if (l < nb) {
            r[b] = load_brow(in.p, (long long)q + (long long)l * Q);
        } else {
            r[b].D = kEye2;
            r[b].L = kZero2;
            r[b].U = kZero2;
            r[b].y0 = r[b].y1 = 0.f;
        }
    }
    for (int delta = 1; delta < nb; delta <<= 1) {
#pragma unroll
        for (int b = 0; b < BR; ++b) put_rec(sR, t + NT * b, minv(r[b].D), r[b]);
        __syncthreads();
#pragma unroll
        for (int b = 0; b < BR; ++b) {
            const int l = t + NT * b;
            const bool hl = l - delta >= 0, hr = l + delta < nb;
            r[b] = level(r[b], sR, hl, hl ? l - delta : l, hr, hr ? l + delta : l);
        }
        __syncthreads();
    }
    unsigned long long dmax = 0ull, xmax = 0ull;
#pragma unroll
    for (int b = 0; b < BR; ++b) {
        const int l = t + NT * b;
        if (l < nb) {
            const M2 Di = minv(r[b].D);
            const float d0 = fmaf(Di.p, r[b].y0, Di.q * r[b].y1);
            const float d1 = fmaf(Di.r, r[b].y0, Di.s * r[b].y1);
            const long long K = R + ((long long)l << S);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const long long i = 2 * K + h;
                if (i < n0) {
                    const long long g = (long long)j * xss + i * xrs;
                    const double dx = (double)(h == 0 ? d0 : d1);
                    const double xn = first ? dx : x[g] + dx;
                    x[g] = xn;
                    const unsigned long long bd = abs_bits(dx), bx = abs_bits(xn);
                    dmax = bd > dmax ? bd : dmax;
                    xmax = bx > xmax ? bx : xmax;
                }
            }
        }
    }
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

}  // namespace

size_t split_f32r_workspace(int n, int m)
{
    return plan(n, m).bytes;
}

int split_f32r_solve(const SolveArgsF32R& A, int layout, void* ws, size_t ws_bytes,
                     hipStream_t st)
{
    (void)layout;
    const int n0 = A.n, m0 = A.m;
    const Layout P = plan(n0, m0);
    const int S = P.S;
    if (!ws || ws_bytes < P.bytes) return -3;
    char* w = static_cast<char*>(ws);
    SysState* state = reinterpret_cast<SysState*>(w + P.state);
    int* stop = reinterpret_cast<int*>(w + P.stop);
    double* y0 = reinterpret_cast<double*>(w + P.y0);
    const long long NB = P.nb0 * m0;
    Arrays14 lvl0, set0, set1;
    for (int v = 0; v < kWords; ++v) {
        lvl0.p[v] = reinterpret_cast<float*>(w + P.lvl0) + (size_t)v * NB;
        set0.p[v] = reinterpret_cast<float*>(w + P.sets) + (size_t)v * P.per;
        set1.p[v] = reinterpret_cast<float*>(w + P.sets) + (size_t)(kWords + v) * P.per;
    }
    const unsigned bblocks = (unsigned)((NB + 255) / 256);
    const unsigned sblocks = (unsigned)((m0 + 255) / 256);
    hipLaunchKernelGGL(init_state_kernel, dim3(sblocks), dim3(256), 0, st, state, stop, m0);
    hipLaunchKernelGGL(prep_kernel, dim3(bblocks), dim3(256), 0, st, A.a, A.b, A.c, A.d, A.e,
                       A.y, lvl0, y0, n0, m0, A.rs, A.ss, P.nb0);
    if (hipGetLastError() != hipSuccess) return -1;
    for (int it = 0; it <= A.max_sweeps; ++it) {
        Arrays14 in = lvl0;
        int Sin = 0, Qin = m0;
        bool to0 = true;
        while (Sin < S) {
            const int k = (S - Sin < kFuse) ? S - Sin : kFuse;
            const int H = (1 << k) - 1, TI = kPassTB - 2 * H;
            const long long L = (P.nb0 + (1LL << Sin) - 1) >> Sin;
            const int tiles = (int)((L + TI - 1) / TI);
            const int groups = (Qin + kPassG - 1) / kPassG;
            const long long blocks = (long long)groups * ((tiles + 7) / 8) * 8;
            if (blocks > 0x7fffffffLL) return -2;
            Arrays14 out = to0 ? set0 : set1;
            hipLaunchKernelGGL((bpass_kernel<kPassTB, kPassG, kPassRPT>), dim3((unsigned)blocks),
                               dim3(kPassTB * kPassG / kPassRPT), 0, st, in, out, P.nb0, m0, Sin,
                               k, Qin, tiles, stop);
            Sin += k;
            Qin <<= k;
            in = out;
            to0 = !to0;
        }
        hipLaunchKernelGGL((bstage_kernel<kStageNT, kStageBR>), dim3((unsigned)Qin),
                           dim3(kStageNT), 0, st, in, Qin, S, P.nb0, n0, m0, A.rs, A.ss, A.y,
                           state, stop, it == 0 ? 1 : 0);
        hipLaunchKernelGGL(decide_kernel, dim3(sblocks), dim3(256), 0, st, state, stop, m0,
                           A.max_sweeps, A.tol, A.sweeps_out);
        if (it < A.max_sweeps)
            hipLaunchKernelGGL(residual_kernel, dim3(bblocks), dim3(256), 0, st, A.a, A.b, A.c,
                               A.d, A.e, y0, A.y, lvl0.p[12], lvl0.p[13], stop, n0, m0, A.rs,
                               A.ss, P.nb0);
        if (hipGetLastError() != hipSuccess) return -1;
    }
    return 0;
}

}  // namespace pcrp
