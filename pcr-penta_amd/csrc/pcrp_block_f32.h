// FP32 inner solve of the refinement on the 2x2-block tridiagonal form,
// shared by the uniform split path (pcrp_split_f32r.hip) and the uneven
// batch path (pcrp_uneven_f32r.hip).
//
// Block row K = rows 2K, 2K+1:  L_K X_{K-1} + D_K X_K + U_K X_{K+1} = Y_K with
//     L_K = [[a_2K, b_2K], [0, a_2K+1]],  D_K = [[c_2K, d_2K], [b_2K+1, c_2K+1]],
//     U_K = [[e_2K, 0], [d_2K+1, e_2K+1]]
// (entries outside the band zeroed; the missing row of an odd n is an
// identity row). One level at block stride s:
//     alpha = -L_K D_{K-s}^-1,  beta = -U_K D_{K+s}^-1,
//     D_K += alpha U_{K-s} + beta L_{K+s},  L_K = alpha L_{K-s},
//     U_K = beta U_{K+s},                   Y_K += alpha Y_{K-s} + beta Y_{K+s}.
// The pivots are the diagonal blocks. PCR-Penta's own recurrences divide by
// off-diagonals (kernel.cu:73,75,99,101), which on random dominant matrices
// leaves 1e-5..1e-2 relative error in FP32; this form stays near 3e-7.
#pragma once

#include "pcrp_device.h"
#include "pcrp_internal.h"

namespace pcrp {
namespace {

struct BRow {
    M2 D, L, U;
    float y0, y1;
};

// Block rows in global memory: block row g's matrix part is the 48-byte
// record m[3g .. 3g+2] (D, L, U, each row-major) and its right-hand side
// y[g]. A record is three 16-byte loads, so a workgroup that walks one
// sub-system (records far apart) issues 4 memory requests per block row
// instead of 14 for one array per word; the residual writes y alone,
// coalesced.
struct BRows {
    float4* m;
    float2* y;
};
constexpr size_t kBRowBytes = 3 * sizeof(float4) + sizeof(float2);   // 56

// Per-system refinement state (device).
struct SysState {
    unsigned long long dmax, xmax;   // bit patterns of max |dx|, max |x| this sweep
    double dprev;                    // max |dx| of the previous sweep
    int sweep;                       // index of the sweep being decided
    int last;                        // index of the last sweep run (set on stopping)
};

__device__ __forceinline__ BRow identity_brow()
{
    BRow r;
    r.D = kEye2;
    r.L = kZero2;
    r.U = kZero2;
    r.y0 = r.y1 = 0.f;
    return r;
}

// Block row from the FP64 rows 2K (w[0]) and 2K+1 (w[1]), each {a,b,c,d,e,y}
// with the out-of-band entries already zeroed.
__device__ __forceinline__ BRow brow_from_rows(const double (&w)[2][6])
{
    BRow r;
    r.L = M2{(float)w[0][0], (float)w[0][1], 0.f, (float)w[1][0]};
    r.D = M2{(float)w[0][2], (float)w[0][3], (float)w[1][1], (float)w[1][2]};
    r.U = M2{(float)w[0][4], 0.f, (float)w[1][3], (float)w[1][4]};
    r.y0 = (float)w[0][5];
    r.y1 = (float)w[1][5];
    return r;
}

// LDS records of the block rows during a level: D^-1, L, U (one 16-byte
// vector each) and Y, so a neighbour costs four wide LDS reads.
template <int N>
struct LdsRec {
    float4 Di[N], L[N], U[N];
    float2 Y[N];
};

template <int N>
__device__ __forceinline__ void put_rec(LdsRec<N>& s, int i, const M2& Di, const BRow& r)
{
    s.Di[i] = f4(Di);
    s.L[i] = f4(r.L);
    s.U[i] = f4(r.U);
    s.Y[i] = make_float2(r.y0, r.y1);
}

// One level for block row r from the records of its neighbours il / ir
// (hl / hr: the neighbour exists in the sub-system); am: running maximum of
// the level's multiplier entries.
template <int N>
__device__ __forceinline__ BRow level(const BRow& r, const LdsRec<N>& s, bool hl, int il, bool hr,
                                      int ir, float& am)
{
    const M2 DiL = m2(s.Di[il]), LL = m2(s.L[il]), UL = m2(s.U[il]);
    const M2 DiR = m2(s.Di[ir]), LR = m2(s.L[ir]), UR = m2(s.U[ir]);
    const float2 yl = s.Y[il], yr = s.Y[ir];
    const M2 al = hl ? mmul_neg(r.L, DiL) : kZero2;
    const M2 be = hr ? mmul_neg(r.U, DiR) : kZero2;
    am = fmaxf(am, fmaxf(maxabs2(al), maxabs2(be)));
    BRow o;
    o.D = madd2(r.D, al, UL, be, LR);
    o.L = mmul(al, LL);
    o.U = mmul(be, UR);
    o.y0 = fmaf(be.p, yr.x, fmaf(be.q, yr.y, fmaf(al.p, yl.x, fmaf(al.q, yl.y, r.y0))));
    o.y1 = fmaf(be.r, yr.x, fmaf(be.s, yr.y, fmaf(al.r, yl.x, fmaf(al.s, yl.y, r.y1))));
    return o;
}

__device__ __forceinline__ BRow load_brow(const BRows& w, long long g)
{
    BRow r;
    r.D = m2(w.m[3 * g]);
    r.L = m2(w.m[3 * g + 1]);
    r.U = m2(w.m[3 * g + 2]);
    const float2 y = w.y[g];
    r.y0 = y.x;
    r.y1 = y.y;
    return r;
}

__device__ __forceinline__ void store_brow(const BRows& w, long long g, const BRow& r)
{
    w.m[3 * g] = f4(r.D);
    w.m[3 * g + 1] = f4(r.L);
    w.m[3 * g + 2] = f4(r.U);
    w.y[g] = make_float2(r.y0, r.y1);
}

// Bytes of a set of nb block rows (records, then right-hand sides), rounded
// up to 256 so that sets carved one after the other stay aligned.
__host__ inline size_t brows_bytes(size_t nb) { return (nb * kBRowBytes + 255) / 256 * 256; }

// Carve a set of nb block rows out of workspace memory at p (256-byte
// aligned); returns where the next set starts.
__host__ inline char* carve_brows(char* p, size_t nb, BRows& w)
{
    w.m = reinterpret_cast<float4*>(p);
    w.y = reinterpret_cast<float2*>(p + nb * 3 * sizeof(float4));
    return p + brows_bytes(nb);
}

__device__ __forceinline__ unsigned long long abs_bits(double v)
{
    return (unsigned long long)__double_as_longlong(fabs(v));
}

// Block -> (group, tile), tiles striped over the 8 XCDs with the group index
// fastest (the sub-systems of a group share cache lines).
__device__ __forceinline__ bool btile_of(int b, int groups, int tiles, int& grp, int& tile)
{
    constexpr int NX = 8;
    const int xcd = b % NX, s = b / NX;
    const int tl = s / groups;
    grp = s - tl * groups;
    tile = tl * NX + xcd;
    return tile < tiles;
}

// Workgroups of `kernel` (block_threads threads) resident on the whole GPU at
// once, rounded down to a multiple of 8 so that logical block b and
// b + grid run on the same XCD; 0 if the runtime cannot say. The sweep
// kernels walk their logical blocks in steps of the grid. Sweeps 0 .. 2 (x0
// and the two corrections a dominant system needs) launch one workgroup per
// logical block; later sweeps, which every system has usually left, launch
// only this many, so such a launch costs one round of workgroups that read
// the stop count and exit (the resident grid for every sweep measured 6-8%
// slower on the active sweeps: no dynamic balancing of the tiles).
constexpr int kFullGridSweeps = 3;
template <typename K>
__host__ inline int resident_grid(K kernel, int block_threads)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(kernel),
                                                     block_threads, 0) != hipSuccess)
        return 0;
    const int g = cus * (per > 0 ? per : 1) / 8 * 8;
    return g > 0 ? g : 8;
}

// resident_grid cached per device (one cache per call site).
template <typename K>
__host__ inline int resident_grid(std::atomic<int> (&cache)[kMaxDevices], K kernel,
                                  int block_threads)
{
    return device_cached(cache, [&](int) { return resident_grid(kernel, block_threads); });
}

// Launch grid for `logical` blocks walked by at most `resident` workgroups.
__host__ inline unsigned walk_grid(long long logical, int resident)
{
    return (unsigned)(resident > 0 && logical > resident ? resident : logical);
}

__host__ __forceinline__ long long bblocks_for(int groups, int tiles)
{
    constexpr int NX = 8;
    return (long long)groups * ((tiles + NX - 1) / NX) * NX;
}

// One tile of one HBM pass: block rows l = tile*TI - H + lr (lr < TB) of the
// current sub-system (this thread's slot g of G), k levels in LDS, interior
// rows stored. in_at(l) / out_at(l): element of block row l in / out.
// y_only: the output records already hold this pass's matrix part (the
// matrix is the same in every refinement sweep), so only Y is stored.
// Each thread owns RPT slots (rows lr0 + (TB/RPT)*kk of the same slot g).
template <int TB, int G, int RPT, typename InAt, typename OutAt>
__device__ __forceinline__ void bpass_tile(const BRows& in, const BRows& out, bool run,
                                           bool y_only, int l0, int Lq, int k, int H, int TI,
                                           InAt in_at, OutAt out_at, LdsRec<TB * G>& sR)
{
    constexpr int NTH = TB * G / RPT, RSTEP = TB / RPT;
    const int tid = threadIdx.x, lr0 = tid / G, g = tid - lr0 * G;
    BRow r[RPT];
    bool valid[RPT];
#pragma unroll
    for (int kk = 0; kk < RPT; ++kk) {
        const int l = l0 + RSTEP * kk;
        valid[kk] = run && l >= 0 && l < Lq;
        r[kk] = valid[kk] ? load_brow(in, in_at(l)) : identity_brow();
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
                float unused = 0.f;   // passes run every level the host launched
                r[kk] = level(r[kk], sR, hl, il, hr, ir, unused);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int kk = 0; kk < RPT; ++kk) {
        const int l = l0 + RSTEP * kk, lr = lr0 + RSTEP * kk;
        if (valid[kk] && lr >= H && lr < H + TI) {
            if (y_only) out.y[out_at(l)] = make_float2(r[kk].y0, r[kk].y1);
            else store_brow(out, out_at(l), r[kk]);
        }
    }
}

// LDS solve of one sub-system of nb block rows (block row l at in_at(l)):
// every remaining level, X = D^-1 Y, then x += dx (x = dx when first) at the
// elements x_at(l, h) returns (negative: no such row), and the workgroup's
// max |dx| and max |x| go to *st by atomic max. All threads call it together.
template <int NT, int BR, typename InAt, typename XAt>
__device__ __forceinline__ void bstage_solve(const BRows& in, int nb, InAt in_at, XAt x_at,
                                             double* __restrict__ x, bool first, SysState* st,
                                             LdsRec<NT * BR>& sR,
                                             unsigned long long (&red)[2][NT / 64],
                                             int (&amx)[NT / 64])
{
    const int t = threadIdx.x;
    BRow r[BR];
#pragma unroll
    for (int b = 0; b < BR; ++b) {
        const int l = t + NT * b;
        r[b] = l < nb ? load_brow(in, in_at(l)) : identity_brow();
    }
    // truncated once a level's multipliers are all below kTruncF32
    for (int delta = 1; delta < nb; delta <<= 1) {
#pragma unroll
        for (int b = 0; b < BR; ++b) put_rec(sR, t + NT * b, minv(r[b].D), r[b]);
        __syncthreads();
        float am = 0.f;
#pragma unroll
        for (int b = 0; b < BR; ++b) {
            const int l = t + NT * b;
            const bool hl = l - delta >= 0, hr = l + delta < nb;
            r[b] = level(r[b], sR, hl, hl ? l - delta : l, hr, hr ? l + delta : l, am);
        }
        const int big = __any(am > kTruncF32);   // one flag per wave
        if ((t & 63) == 0) amx[t >> 6] = big;
        __syncthreads();
        int any = 0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) any |= amx[w];
        if (!any) break;   // uniform: the remaining levels are negligible
    }
    unsigned long long dmax = 0ull, xmax = 0ull;
#pragma unroll
    for (int b = 0; b < BR; ++b) {
        const int l = t + NT * b;
        if (l < nb) {
            const M2 Di = minv(r[b].D);
            const float d[2] = {fmaf(Di.p, r[b].y0, Di.q * r[b].y1),
                                fmaf(Di.r, r[b].y0, Di.s * r[b].y1)};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const long long g = x_at(l, h);
                if (g >= 0) {
                    const double dx = (double)d[h];
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
        atomicMax(&st->dmax, dmax);
        atomicMax(&st->xmax, xmax);
    }
    // red is read by thread 0 above; a caller's next sub-system of one block
    // row runs no level, so no barrier would separate its red writes from
    // that read
    __syncthreads();
}

// stop[0 .. m-1]: per-system stop flags; stop[m]: systems still running, so
// every kernel of a sweep after the last system stopped returns at once.
__global__ void init_state_kernel(SysState* __restrict__ st, int* __restrict__ stop, int m)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < m) {
        st[j] = SysState{0ull, 0ull, 0.0, 0, 0};
        stop[j] = 0;
    }
    if (j == 0) stop[m] = m;
}

// Stopping rule of the one-workgroup refinement kernel, per system: stop when
// max|dx| <= tol max|x|, or when the observed contraction rho = |dx_k| /
// |dx_{k-1}| (from the second correction on) puts the error left, 10 rho
// |dx_k|, below tol with rho <= 0.1. Not done after max_sweeps corrections,
// or a NaN or infinite correction: sweeps = max_sweeps + 1.
__global__ void decide_kernel(SysState* __restrict__ st, int* __restrict__ stop, int m,
                              int max_sweeps, double tol, int* __restrict__ sweeps_out)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m || stop[j]) return;
    SysState s = st[j];
    const double dmax = __longlong_as_double((long long)s.dmax);
    const double xmax = __longlong_as_double((long long)s.xmax);
    const bool bad = !(dmax < INFINITY) || !(xmax < INFINITY);   // NaN or overflow
    const bool small = !(dmax > tol * xmax);
    const bool contracting = s.sweep > 1 && dmax <= 0.1 * s.dprev &&
                             10.0 * dmax * dmax <= tol * xmax * s.dprev;
    const bool done = !bad && (small || contracting);
    s.dprev = dmax;
    if (done || bad || s.sweep >= max_sweeps) {
        s.last = s.sweep;
        if (bad) s.sweep = max_sweeps + 1;
        else if (!done) ++s.sweep;          // flag: tolerance not met
        stop[j] = 1;
        atomicSub(&stop[m], 1);
        if (sweeps_out) sweeps_out[j] = s.sweep;
    } else {
        ++s.sweep;
    }
    s.dmax = 0ull;
    s.xmax = 0ull;
    st[j] = s;
}

// r = y0 - A x in FP64 for rows i (< n) of one system at element g of the
// caller's arrays (row stride rs); the operation order of the one-workgroup
// kernel's residual.
__device__ __forceinline__ double residual_row(const double* __restrict__ a,
                                               const double* __restrict__ b,
                                               const double* __restrict__ c,
                                               const double* __restrict__ d,
                                               const double* __restrict__ e,
                                               const double* __restrict__ y0,
                                               const double* __restrict__ x, long long g,
                                               long long i, long long n, long long rs)
{
    const double xm2 = i >= 2 ? x[g - 2 * rs] : 0.0, xm1 = i >= 1 ? x[g - rs] : 0.0;
    const double xp1 = i + 1 < n ? x[g + rs] : 0.0, xp2 = i + 2 < n ? x[g + 2 * rs] : 0.0;
    double ax = c[g] * x[g];
    ax = __builtin_fma(a[g], xm2, ax);
    ax = __builtin_fma(b[g], xm1, ax);
    ax = __builtin_fma(d[g], xp1, ax);
    ax = __builtin_fma(e[g], xp2, ax);
    return y0[g] - ax;
}

// Row i (< n) of one system at element g as {a,b,c,d,e,y}, out-of-band
// entries zeroed.
__device__ __forceinline__ void load_row(const double* __restrict__ a, const double* __restrict__ b,
                                         const double* __restrict__ c, const double* __restrict__ d,
                                         const double* __restrict__ e, const double* __restrict__ y,
                                         long long g, long long i, long long n, double (&w)[6])
{
    w[0] = i >= 2 ? a[g] : 0.0;
    w[1] = i >= 1 ? b[g] : 0.0;
    w[2] = c[g];
    w[3] = i + 1 < n ? d[g] : 0.0;
    w[4] = i + 2 < n ? e[g] : 0.0;
    w[5] = y[g];
}

}  // namespace
}  // namespace pcrp
