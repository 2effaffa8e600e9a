// FP32 inner solve + FP64 refinement by overlapping windows, for long and
// uneven batches of diagonally dominant systems: one kernel per refinement
// sweep, each reading A, y (and x) once and writing x once.
//
// No counterpart in the reference, which is FP64-only (kernel.cu:32-154,
// 158-791). The inner solve is the 2x2-block PCR of pcrp_block_f32.h,
// truncated: on a diagonally dominant matrix the multipliers of block cyclic
// reduction decay quadratically from level to level (Heller, SIAM J. Numer.
// Anal. 13 (1976) 484-496), and once every multiplier of a level is below
// kTruncF32 the couplings left are negligible, so X = D^-1 Y right after it.
// Stopping after level s, block row K depends only on block rows within
// 2^(s+1) - 1 of it, so the solve is local (the truncated-SPIKE observation,
// Polizzi & Sameh, Parallel Computing 32 (2006) 177-194):
//   * a system of nb <= kWinNB block rows is one window, every level it needs
//     (exact, like the one-workgroup kernel);
//   * a longer system is cut into windows of at most T = kWinNB - 2H interior
//     block rows (balanced: ceil(nb / windows)) with H = 2^L - 1 halo block
//     rows on each side (L: the tunable window level count, 5 by default);
//     a window runs at most L levels with neighbours outside the window
//     treated as absent, so the values at its interior rows equal those of
//     the whole-system PCR (the halo rows absorb the edge). A window whose
//     interior multipliers are still above kTruncF32 after L levels sets
//     `nonlocal`, and the batch
//     is solved by the pass/stage path instead (host read-back after the
//     first sweep, which writes only workspace buffers: y is still the
//     right-hand side).
//     A stream under capture cannot read back: the callers then use the
//     pass/stage path directly.
// One sweep, per window: r = y - A x (FP64, rows of the window from global
// memory), FP32 block rows, factor and right-hand side together level by
// level in LDS (the first sweep finds each window's level count, later
// sweeps, same matrix, run that count), dx = D^-1 Y at the interior rows,
// x' = x + dx (FP64) into the other x buffer (windows read x at their halo
// rows, so x is double buffered), max |dx| and max |x'| per system by atomic
// max; then the stopping rule of the other refinement paths (decide_kernel).
// Direct form (uneven and contiguous batches): sweep it writes x0, x1 or y
// for it % 3 = 0, 1, 2; the first sweep reads y and copies it to y0, later
// sweeps read the right-hand side from y0, so the usual two corrections end
// in y and other systems are copied there from the buffer their last sweep
// wrote: per unknown, first sweep 48 B read + 16 B written, later sweeps
// 56 B read + 8 B written. Interleaved uniform batches are first copied to
// system-contiguous arrays (96 B per unknown, LDS-tiled transpose), swept
// there with x alternating between x0 and x1, and x is transposed back
// (16 B).
#include "pcr_penta.h"
#include "pcrp_block_f32.h"
#include "pcrp_internal.h"

#include <algorithm>
#include <mutex>
#include <vector>

namespace pcrp {

namespace {

constexpr int kWinNT = 256, kWinBR = 2;
constexpr int kWinNB = kWinNT * kWinBR;    // block rows per window (LDS records)
// Levels L a window of a longer system may run (tunable, f32r_tune: default
// 4): halo H = 2^L - 1 block rows per side (the reach of L levels), interior
// T = kWinNB - 2H block rows. kWinLMax bounds L (workspace sizes use it).
constexpr int kWinLMax = 7;
constexpr int kWinTMin = kWinNB - 2 * ((1 << kWinLMax) - 1);
inline int win_halo(int L) { return (1 << L) - 1; }
inline int win_interior(int L) { return kWinNB - 2 * win_halo(L); }

// Block row K holds rows 2K - shift and 2K - shift + 1 (rows outside 0..n-1
// are identity rows): a contiguous system at an odd element offset takes
// shift = 1, so that every block row's first element is 16-byte aligned
// (the leading identity row is decoupled: row 0's couplings to rows -2, -1
// are outside the band).
struct WSys {
    long long off;   // element of row 0 (row i at off + i * rs)
    long long w0;    // first window of the system (uneven batches)
    int n, nb, nw, shift;
    int T, pad;      // interior block rows per window (balanced: ceil(nb / nw) <= interior)
};

struct WArgs {
    const double *a, *b, *c, *d, *e;
    const double* rhs;   // right-hand side: y, its copy y0, or its system-contiguous copy
    double* y0;          // first sweep of the direct form: receives y at the interior rows
    const double* xin;   // x of the previous sweep (null in the first)
    double* xout;        // x of this sweep
    long long rs;
    const WSys* sys;
    SysState* st;
    const int* stop;
    int* nonlocal;
    unsigned char* lev;   // levels each window runs (found by the first sweep)
    int m, nwin, uniform;
    int nwu;   // uniform batches: windows per system
    int vec;   // rs == 1 and every array 16-byte aligned
    int wl, wh;   // levels a window of a longer system runs, its halo block rows per side
    const int* wmap = nullptr;   // uneven batches: the system of every window
};

// Interleaved uniform batches (more than one system) are transposed first:
// in place, a window's loads would touch one element per cache line.
inline bool window_transposed(int m, long long rs)
{
    return m > 1 && rs == m;
}

__host__ __device__ inline int windows_for(int nb, int T)
{
    return nb <= kWinNB ? 1 : (nb + T - 1) / T;
}

// Window w -> (system j, window index wi). Uniform batches of systems with
// many windows are window-major (the m systems' windows at one position are
// consecutive) with runs of 16 on one XCD: on the n = 2^20, m = 32 batch 7%
// faster than system-major order; systems of at most 8 windows go
// system-major with runs of 8 on one XCD (neighbouring windows share halo
// rows in that XCD's L2: 2.5% faster at n = 2048, m = 16384). Uneven batches
// are system-major: a table gives each window's system (a binary search over
// the systems' first windows, 14 dependent loads at 10000 systems, cost a
// window its first microseconds; the copy-out kernel spent 0.33 ms on it).
__device__ __forceinline__ void window_of(const WArgs& A, int w, int& j, int& wi)
{
    if (A.uniform) {
        const int nw = A.nwu;
        if (nw <= 8) {
            const int v = xcd_remap_c<8>(w, A.nwin);
            j = v / nw;
            wi = v - j * nw;
        } else {
            const int v = xcd_remap_c<16>(w, A.nwin);
            j = v % A.m;
            wi = v / A.m;
        }
        return;
    }
    j = A.wmap[w];
    wi = (int)(w - A.sys[j].w0);
}

// The window -> system table of an uneven batch: system j's windows are
// w0 .. w0 + nw - 1.
__global__ __launch_bounds__(256) void wmap_kernel(const WSys* __restrict__ sys, int m,
                                                   int* __restrict__ wmap)
{
    for (int j = blockIdx.x; j < m; j += gridDim.x) {
        const long long w0 = sys[j].w0;
        const int nw = sys[j].nw;
        for (int k = threadIdx.x; k < nw; k += blockDim.x) wmap[w0 + k] = j;
    }
}

__device__ __forceinline__ double2 ld2(const double* p)
{
    return *reinterpret_cast<const double2*>(p);
}

// Rows i0 and i0 + 1 of a contiguous system, 2 <= i0 and i0 + 3 < n
// (every coupling in the band, x at i0-2 .. i0+3 in the system), element g0
// 16-byte aligned: one 16-byte load per array for both rows, three for x.
// The residual keeps residual_row's operation order.
__device__ __forceinline__ void load_pair_vec(const WArgs& A, long long g0, double (&rw)[2][6],
                                              double (&xv)[2])
{
    const double2 av = ld2(A.a + g0), bv = ld2(A.b + g0), cv = ld2(A.c + g0);
    const double2 dv = ld2(A.d + g0), ev = ld2(A.e + g0), yv = ld2(A.rhs + g0);
    rw[0][0] = av.x; rw[0][1] = bv.x; rw[0][2] = cv.x; rw[0][3] = dv.x; rw[0][4] = ev.x;
    rw[1][0] = av.y; rw[1][1] = bv.y; rw[1][2] = cv.y; rw[1][3] = dv.y; rw[1][4] = ev.y;
    if (A.xin) {
        const double2 xm = ld2(A.xin + g0 - 2), xc = ld2(A.xin + g0), xp = ld2(A.xin + g0 + 2);
        xv[0] = xc.x;
        xv[1] = xc.y;
        double ax = cv.x * xc.x;
        ax = __builtin_fma(av.x, xm.x, ax);
        ax = __builtin_fma(bv.x, xm.y, ax);
        ax = __builtin_fma(dv.x, xc.y, ax);
        ax = __builtin_fma(ev.x, xp.x, ax);
        rw[0][5] = yv.x - ax;
        ax = cv.y * xc.y;
        ax = __builtin_fma(av.y, xm.y, ax);
        ax = __builtin_fma(bv.y, xc.x, ax);
        ax = __builtin_fma(dv.y, xp.x, ax);
        ax = __builtin_fma(ev.y, xp.y, ax);
        rw[1][5] = yv.y - ax;
    } else {
        xv[0] = yv.x;   // the first sweep copies y to y0
        xv[1] = yv.y;
        rw[0][5] = yv.x;
        rw[1][5] = yv.y;
    }
}

// FIND (first sweep): run levels until the interior multipliers are all
// below kTruncF32 and store the count per window; later sweeps (same matrix,
// same multipliers) run exactly that many levels without the test.
template <int NT, int BR, bool FIND>
__global__ __launch_bounds__(NT) void wsweep_kernel(WArgs A)
{
    constexpr int NB = NT * BR;
    static_assert(NB == kWinNB, "one window per workgroup");
    __shared__ LdsRec<NB> sR;
    __shared__ unsigned long long red[2][NT / 64];
    __shared__ int flg[NT / 64];
    if (A.stop[A.m] == 0) return;   // every system has stopped
    const int t = threadIdx.x;
    // windows w, w + grid, ... (grid a multiple of 8: same XCD)
    for (int w = blockIdx.x; w < A.nwin; w += gridDim.x) {
        int j, wi;
        window_of(A, w, j, wi);
        if (A.stop[j]) continue;   // uniform over the workgroup
        const WSys S = A.sys[j];
        const bool full = S.nw == 1;
        const int kb = full ? 0 : wi * S.T - A.wh;      // block row of window slot 0
        const int k0 = full ? 0 : wi * S.T;             // interior block rows [k0, k1)
        const int k1 = full ? S.nb : min(S.nb, k0 + S.T);
        // block rows the window holds: the interior and its halos, inside the
        // system (the slots past them are idle)
        const int klo = full ? 0 : max(0, k0 - A.wh);
        const int khi = full ? S.nb : min(S.nb, k1 + A.wh);
        int lmax = 0;
        if (full) { while ((1 << lmax) < S.nb) ++lmax; } else { lmax = A.wl; }

        BRow r[BR];
        double xv[BR][2];   // x of the previous sweep at the rows (first sweep: y)
#pragma unroll
        for (int q = 0; q < BR; ++q) {
            const int K = kb + t + NT * q;
            xv[q][0] = xv[q][1] = 0.0;
            if (K < klo || K >= khi) {
                r[q] = identity_brow();
                continue;
            }
            double rw[2][6];
            const long long i0 = 2LL * K - S.shift;
            if (A.vec && ((S.off + i0) & 1) == 0 && i0 >= 2 && i0 + 3 < S.n) {
                load_pair_vec(A, S.off + i0, rw, xv[q]);
                r[q] = brow_from_rows(rw);
                continue;
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const long long i = i0 + h;
                if (i >= 0 && i < S.n) {
                    const long long g = S.off + i * A.rs;
                    load_row(A.a, A.b, A.c, A.d, A.e, A.rhs, g, i, S.n, rw[h]);
                    if (A.xin) {
                        xv[q][h] = A.xin[g];
                        rw[h][5] = residual_row(A.a, A.b, A.c, A.d, A.e, A.rhs, A.xin, g, i, S.n,
                                                A.rs);
                    } else {
                        xv[q][h] = rw[h][5];   // the first sweep copies y to y0
                    }
                } else {   // identity row before row 0 or past row n-1
                    rw[h][0] = rw[h][1] = rw[h][3] = rw[h][4] = rw[h][5] = 0.0;
                    rw[h][2] = 1.0;
                }
            }
            r[q] = brow_from_rows(rw);
        }

        // levels: factor and right-hand side together, until the interior
        // multipliers are all below kTruncF32 (uniform over the workgroup)
        if (!FIND) lmax = A.lev[w];
        int s = 0;
        for (; s < lmax; ++s) {
            const int delta = 1 << s;
            // slots outside [klo, khi) stay identity rows and are nobody's
            // neighbour: skipped (whole waves of a short system's window)
#pragma unroll
            for (int q = 0; q < BR; ++q) {
                const int K = kb + t + NT * q;
                if (K >= klo && K < khi) put_rec(sR, t + NT * q, minv(r[q].D), r[q]);
            }
            __syncthreads();
            float am = 0.f;
#pragma unroll
            for (int q = 0; q < BR; ++q) {
                const int l = t + NT * q, K = kb + l;
                if (K < klo || K >= khi) continue;
                // neighbours outside the system or the window are absent
                // (klo >= kb and khi <= kb + NB: the slots exist)
                const bool hl = K - delta >= klo;
                const bool hr = K + delta < khi;
                float ak = 0.f;
                r[q] = level(r[q], sR, hl, hl ? l - delta : l, hr, hr ? l + delta : l, ak);
                if (FIND && K >= k0 && K < k1) am = fmaxf(am, ak);
            }
            if (FIND) {
                const int big = __any(am > kTruncF32);   // one flag per wave
                if ((t & 63) == 0) flg[t >> 6] = big;
            }
            __syncthreads();
            if (FIND) {
                int any = 0;
#pragma unroll
                for (int v = 0; v < NT / 64; ++v) any |= flg[v];
                if (!any) break;
            }
        }
        if (FIND && t == 0) {
            A.lev[w] = (unsigned char)(s < lmax ? s + 1 : lmax);
            // couplings above the threshold past the halo's reach: not local
            if (s == lmax && !full) *A.nonlocal = 1;
        }

        unsigned long long dmax = 0ull, xmax = 0ull;
#pragma unroll
        for (int q = 0; q < BR; ++q) {
            const int K = kb + t + NT * q;
            if (K >= k0 && K < k1) {
                const M2 Di = minv(r[q].D);
                const float dd[2] = {fmaf(Di.p, r[q].y0, Di.q * r[q].y1),
                                     fmaf(Di.r, r[q].y0, Di.s * r[q].y1)};
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const long long i = 2LL * K - S.shift + h;
                    if (i >= 0 && i < S.n) {
                        const long long g = S.off + i * A.rs;
                        const double dx = (double)dd[h];
                        const double xn = A.xin ? xv[q][h] + dx : dx;
                        if (!A.xin && A.y0) A.y0[g] = xv[q][h];
                        A.xout[g] = xn;
                        const unsigned long long bd = abs_bits(dx), bx = abs_bits(xn);
                        dmax = bd > dmax ? bd : dmax;
                        xmax = bx > xmax ? bx : xmax;
                    }
                }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const unsigned long long od = __shfl_xor(dmax, off, 64);
            const unsigned long long ox = __shfl_xor(xmax, off, 64);
            dmax = od > dmax ? od : dmax;
            xmax = ox > xmax ? ox : xmax;
        }
        if ((t & 63) == 0) {
            red[0][t >> 6] = dmax;
            red[1][t >> 6] = xmax;
        }
        __syncthreads();
        if (t == 0) {
#pragma unroll
            for (int v = 1; v < NT / 64; ++v) {
                dmax = red[0][v] > dmax ? red[0][v] : dmax;
                xmax = red[1][v] > xmax ? red[1][v] : xmax;
            }
            atomicMax(&A.st[j].dmax, dmax);
            atomicMax(&A.st[j].xmax, xmax);
        }
        // Thread 0 reads red above while the other waves may already be at
        // the next window; a window of one block row runs no level (no
        // barrier before its red writes), so red, sR and flg are released
        // here, once per window.
        __syncthreads();
    }
}

// x of every system to y from the buffer its last sweep wrote (sweep it
// writes x0, x1, y for it % 3 = 0, 1, 2), over the window's interior rows.
__global__ __launch_bounds__(256) void wcopy_kernel(WArgs A, const double* __restrict__ x0,
                                                    const double* __restrict__ x1,
                                                    double* __restrict__ y)
{
    for (int w = blockIdx.x; w < A.nwin; w += gridDim.x) {
        int j, wi;
        window_of(A, w, j, wi);
        const int b = A.st[j].last % 3;
        if (b == 2) continue;   // already in y
        const double* __restrict__ x = b == 1 ? x1 : x0;
        const WSys S = A.sys[j];
        const bool full = S.nw == 1;
        const int k0 = full ? 0 : wi * S.T;
        const int k1 = full ? S.nb : min(S.nb, k0 + S.T);
        const long long i0 = max(0LL, 2LL * k0 - S.shift);
        const long long i1 = min((long long)S.n, 2LL * k1 - S.shift);
        for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            const long long g = S.off + i * A.rs;
            y[g] = x[g];
        }
    }
}

// Interleaved -> system-contiguous copies of the six arrays (row i of system
// j from i*m + j to j*n + i): tiles of 64 rows x 32 systems through LDS, so
// both sides are coalesced. blockIdx.z: the array.
struct Six {
    const double* p[6];
};
constexpr int kTrRows = 64, kTrSys = 32;

__global__ __launch_bounds__(256) void wtranspose_in_kernel(Six src, double* __restrict__ dst,
                                                            int n, int m)
{
    __shared__ double tile[kTrRows][kTrSys + 1];
    const int f = blockIdx.z;
    const long long i0 = (long long)blockIdx.x * kTrRows;
    const int j0 = blockIdx.y * kTrSys;
    const double* __restrict__ in = src.p[f];
    const int tj = threadIdx.x % kTrSys, ti = threadIdx.x / kTrSys;   // 32 x 8
    for (int r = ti; r < kTrRows; r += 256 / kTrSys) {
        const long long i = i0 + r;
        if (i < n && j0 + tj < m) tile[r][tj] = in[i * m + j0 + tj];
    }
    __syncthreads();
    const int oi = threadIdx.x % kTrRows, oj = threadIdx.x / kTrRows;   // 64 x 4
    double* __restrict__ out = dst + (long long)f * n * m;
    for (int c = oj; c < kTrSys; c += 256 / kTrRows) {
        const long long i = i0 + oi;
        if (i < n && j0 + c < m) out[(long long)(j0 + c) * n + i] = tile[oi][c];
    }
}

// The result back to the interleaved y: system j's x from the buffer its
// last sweep wrote (odd: x1, even: x0), system-contiguous.
__global__ __launch_bounds__(256) void wtranspose_out_kernel(const double* __restrict__ x0,
                                                             const double* __restrict__ x1,
                                                             const SysState* __restrict__ st,
                                                             double* __restrict__ y, int n, int m)
{
    __shared__ double tile[kTrSys][kTrRows + 1];
    const long long i0 = (long long)blockIdx.x * kTrRows;
    const int j0 = blockIdx.y * kTrSys;
    const int oi = threadIdx.x % kTrRows, oj = threadIdx.x / kTrRows;   // 64 x 4
    for (int c = oj; c < kTrSys; c += 256 / kTrRows) {
        const long long i = i0 + oi;
        const int j = j0 + c;
        if (i < n && j < m) tile[c][oi] = ((st[j].last & 1) ? x1 : x0)[(long long)j * n + i];
    }
    __syncthreads();
    const int tj = threadIdx.x % kTrSys, ti = threadIdx.x / kTrSys;   // 32 x 8
    for (int r = ti; r < kTrRows; r += 256 / kTrSys) {
        const long long i = i0 + r;
        if (i < n && j0 + tj < m) y[i * m + j0 + tj] = tile[tj][r];
    }
}

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Bound on the windows of `elems` rows in m systems at any window level
// count: windows_for(nb, T) <= nb / kWinTMin + 1 and nb <= n / 2 + 1 per system.
size_t lev_capacity(long long elems, int m)
{
    return (size_t)elems / (2 * kWinTMin) + 2 * (size_t)m + 64;
}

// Direct: x buffer and copy of y (16 B per element). Transposed (interleaved
// uniform batches): six system-contiguous arrays and two x buffers (64 B).
using WLayout = WindowLayoutOffsets;

WLayout wlayout(long long elems, int m, bool transposed)
{
    WLayout L{};
    size_t o = 0;
    L.state = o; o = align_up(o + (size_t)m * sizeof(SysState));
    L.stop = o;  o = align_up(o + (size_t)(m + 1) * sizeof(int));
    L.flag = o;  o = align_up(o + sizeof(int));
    L.sys = o;   o = align_up(o + (size_t)m * sizeof(WSys));
    L.lev = o;   o = align_up(o + lev_capacity(elems, m));
    L.wmap = o;  o = align_up(o + lev_capacity(elems, m) * sizeof(int));
    L.x0 = o;    o = align_up(o + (size_t)elems * sizeof(double));
    L.x1 = o;    o = align_up(o + (size_t)elems * sizeof(double));
    L.y0 = o;    if (!transposed) o = align_up(o + (size_t)elems * sizeof(double));
    L.six = o;   if (transposed) o = align_up(o + 6 * (size_t)elems * sizeof(double));
    L.bytes = o;
    return L;
}

bool capturing(hipStream_t st)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
}

// Runs the window path on the systems of `sys` (w0 and nw filled in here).
// transposed: uniform interleaved batch of n rows (rs = m); the windows then
// work on system-contiguous copies. fell_back: a window was not local;
// nothing was written to y.
int window_run(const double* a, const double* b, const double* c, const double* d,
               const double* e, double* y, std::vector<WSys>& sys, bool uniform, long long rs,
               long long elems, bool transposed, int n, int max_sweeps, double tol,
               int* sweeps_out, void* ws, size_t ws_bytes, hipStream_t st, bool& fell_back)
{
    fell_back = true;
    const int m = (int)sys.size();
    if (transposed) {
        rs = 1;
        for (int j = 0; j < m; ++j) sys[j].off = (long long)j * n;
    }
    const int wl = f32r_tune_get().win_levels, wT = win_interior(wl);
    long long nwin = 0;
    for (auto& s : sys) {
        s.shift = rs == 1 ? (int)(s.off & 1) : 0;
        s.nb = (s.n + s.shift + 1) / 2;
        s.nw = windows_for(s.nb, wT);
        s.T = (s.nb + s.nw - 1) / s.nw;
        s.w0 = nwin;
        nwin += s.nw;
    }
    const WLayout L = wlayout(elems, m, transposed);
    // leave it to the pass/stage path
    if (nwin > 0x7fffffffLL || (size_t)nwin > lev_capacity(elems, m)) return PCRP_OK;
    if (!ws || ws_bytes < L.bytes) return PCRP_ERR_WORKSPACE;
    char* w = static_cast<char*>(ws);
    SysState* state = reinterpret_cast<SysState*>(w + L.state);
    int* stop = reinterpret_cast<int*>(w + L.stop);
    int* flag = reinterpret_cast<int*>(w + L.flag);
    WSys* dsys = reinterpret_cast<WSys*>(w + L.sys);
    double* x0 = reinterpret_cast<double*>(w + L.x0);
    double* x1 = reinterpret_cast<double*>(w + L.x1);
    double* y0 = reinterpret_cast<double*>(w + L.y0);
    double* six = reinterpret_cast<double*>(w + L.six);
    unsigned char* lev = reinterpret_cast<unsigned char*>(w + L.lev);
    if (hipMemcpyAsync(dsys, sys.data(), sys.size() * sizeof(WSys), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemsetAsync(flag, 0, sizeof(int), st) != hipSuccess)
        return PCRP_ERR_HIP;
    static std::atomic<int> c_gWin[kMaxDevices];
    const int gWin = resident_grid(c_gWin, wsweep_kernel<kWinNT, kWinBR, false>, kWinNT);
    static std::atomic<int> c_gCopy[kMaxDevices];
    const int gCopy = resident_grid(c_gCopy, wcopy_kernel, 256);
    const unsigned sblocks = (unsigned)((m + 255) / 256);
    hipLaunchKernelGGL(init_state_kernel, dim3(sblocks), dim3(256), 0, st, state, stop, m);
    const dim3 tgrid((unsigned)((n + kTrRows - 1) / kTrRows), (unsigned)((m + kTrSys - 1) / kTrSys));
    if (transposed) {
        if (tgrid.x > 0x7fffffffu || tgrid.y > 65535u) return PCRP_ERR_INVALID_SIZE;
        hipLaunchKernelGGL(wtranspose_in_kernel, dim3(tgrid.x, tgrid.y, 6), dim3(256), 0, st,
                           Six{{a, b, c, d, e, y}}, six, n, m);
    }
    const long long N = elems;
    // transposed: A and y from the system-contiguous copies
    const double* ca = transposed ? six : a;
    WArgs A{ca, transposed ? six + N : b, transposed ? six + 2 * N : c,
            transposed ? six + 3 * N : d, transposed ? six + 4 * N : e, nullptr, nullptr, nullptr,
            nullptr, rs, dsys, state, stop, flag, lev, m, (int)nwin, uniform ? 1 : 0,
            uniform ? sys[0].nw : 0, 0, wl, win_halo(wl)};
    {
        const auto al16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
        A.vec = rs == 1 && al16(A.a) && al16(A.b) && al16(A.c) && al16(A.d) && al16(A.e) &&
                al16(y) && al16(x0) && al16(x1) && al16(y0) && al16(six) ? 1 : 0;
    }
    if (!uniform) {
        int* wmap = reinterpret_cast<int*>(w + L.wmap);
        hipLaunchKernelGGL(wmap_kernel, dim3((unsigned)(m < 4096 ? m : 4096)), dim3(256), 0, st,
                           dsys, m, wmap);
        A.wmap = wmap;
    }
    // Direct form: sweep it writes x0, x1, y for it % 3 = 0, 1, 2; the first
    // sweep reads y and copies it to y0, later sweeps read y0, so the usual
    // two corrections end in y with no final copy. Transposed form: x0, x1
    // alternate and the right-hand side is the transposed copy of y.
    double* const xbuf[3] = {x0, x1, transposed ? x0 : y};
    const int cyc = transposed ? 2 : 3;
    for (int it = 0; it <= max_sweeps; ++it) {
        A.rhs = transposed ? six + 5 * N : (it == 0 ? y : y0);
        A.y0 = transposed || it > 0 ? nullptr : y0;
        A.xin = it == 0 ? nullptr : xbuf[(it - 1) % cyc];
        A.xout = xbuf[it % cyc];
        if (it == 0)
            hipLaunchKernelGGL((wsweep_kernel<kWinNT, kWinBR, true>), dim3((unsigned)nwin),
                               dim3(kWinNT), 0, st, A);
        else
            hipLaunchKernelGGL((wsweep_kernel<kWinNT, kWinBR, false>),
                               dim3(walk_grid(nwin, it < kFullGridSweeps ? 0 : gWin)),
                               dim3(kWinNT), 0, st, A);
        hipLaunchKernelGGL(decide_kernel, dim3(sblocks), dim3(256), 0, st, state, stop, m,
                           max_sweeps, tol, sweeps_out);
        if (hipGetLastError() != hipSuccess) return PCRP_ERR_HIP;
        if (it == 0) {
            int h = 0;
            if (hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return PCRP_ERR_HIP;
            if (h) return PCRP_OK;   // fell_back stays true
        }
    }
    if (transposed) {
        hipLaunchKernelGGL(wtranspose_out_kernel, dim3(tgrid.x, tgrid.y), dim3(256), 0, st, xbuf[0],
                           xbuf[1], state, y, n, m);
    } else {
        hipLaunchKernelGGL(wcopy_kernel, dim3(walk_grid(nwin, gCopy)), dim3(256), 0, st, A,
                           xbuf[0], xbuf[1], y);
    }
    if (hipGetLastError() != hipSuccess) return PCRP_ERR_HIP;
    fell_back = false;
    return PCRP_OK;
}

__global__ void clear_flag_kernel(int* f) { *f = 0; }

// ---- plan form: everything that depends on the matrix only, once ----------
// window_plan prepares the batch (system table, for interleaved uniform
// batches the system-contiguous copies of a..e, each window's level count
// from one level-finding sweep, the locality flag read back: synchronous);
// window_plan_execute then runs only kernels (capturable, no host copy, no
// synchronisation): every sweep reuses the levels, and y is transposed in
// and out when the batch is.
int window_plan(const double* a, const double* b, const double* c, const double* d,
                const double* e, std::vector<WSys>& sys, bool uniform, long long rs,
                long long elems, bool transposed, int n, void* ws, size_t ws_bytes,
                hipStream_t st, WindowPlan& P)
{
    P = WindowPlan{};
    P.local = false;
    const int m = (int)sys.size();
    if (transposed) {
        rs = 1;
        for (int j = 0; j < m; ++j) sys[j].off = (long long)j * n;
    }
    const int wl = f32r_tune_get().win_levels, wT = win_interior(wl);
    long long nwin = 0;
    for (auto& s : sys) {
        s.shift = rs == 1 ? (int)(s.off & 1) : 0;
        s.nb = (s.n + s.shift + 1) / 2;
        s.nw = windows_for(s.nb, wT);
        s.T = (s.nb + s.nw - 1) / s.nw;
        s.w0 = nwin;
        nwin += s.nw;
    }
    const WLayout L = wlayout(elems, m, transposed);
    if (nwin > 0x7fffffffLL || (size_t)nwin > lev_capacity(elems, m)) return PCRP_OK;   // not local
    if (!ws || ws_bytes < L.bytes) return PCRP_ERR_WORKSPACE;
    char* w = static_cast<char*>(ws);
    SysState* state = reinterpret_cast<SysState*>(w + L.state);
    int* stop = reinterpret_cast<int*>(w + L.stop);
    int* flag = reinterpret_cast<int*>(w + L.flag);
    WSys* dsys = reinterpret_cast<WSys*>(w + L.sys);
    double* x0 = reinterpret_cast<double*>(w + L.x0);
    double* six = reinterpret_cast<double*>(w + L.six);
    unsigned char* lev = reinterpret_cast<unsigned char*>(w + L.lev);
    const long long N = elems;
    if (hipMemcpyAsync(dsys, sys.data(), sys.size() * sizeof(WSys), hipMemcpyHostToDevice, st) !=
        hipSuccess)
        return PCRP_ERR_HIP;
    const unsigned sblocks = (unsigned)((m + 255) / 256);
    hipLaunchKernelGGL(clear_flag_kernel, dim3(1), dim3(1), 0, st, flag);
    hipLaunchKernelGGL(init_state_kernel, dim3(sblocks), dim3(256), 0, st, state, stop, m);
    const dim3 tgrid((unsigned)((n + kTrRows - 1) / kTrRows), (unsigned)((m + kTrSys - 1) / kTrSys));
    if (transposed) {
        if (tgrid.x > 0x7fffffffu || tgrid.y > 65535u) return PCRP_ERR_INVALID_SIZE;
        // a..e once: the plan keeps their system-contiguous copies
        hipLaunchKernelGGL(wtranspose_in_kernel, dim3(tgrid.x, tgrid.y, 5), dim3(256), 0, st,
                           Six{{a, b, c, d, e, nullptr}}, six, n, m);
    }
    WArgs A{transposed ? six : a, transposed ? six + N : b, transposed ? six + 2 * N : c,
            transposed ? six + 3 * N : d, transposed ? six + 4 * N : e, nullptr, nullptr, nullptr,
            nullptr, rs, dsys, state, stop, flag, lev, m, (int)nwin, uniform ? 1 : 0,
            uniform ? sys[0].nw : 0, 0, wl, win_halo(wl)};
    if (!uniform) {
        int* wmap = reinterpret_cast<int*>(w + L.wmap);
        hipLaunchKernelGGL(wmap_kernel, dim3((unsigned)(m < 4096 ? m : 4096)), dim3(256), 0, st,
                           dsys, m, wmap);
        A.wmap = wmap;
    }
    // level-finding sweep: its right-hand side is any finite array of the
    // batch's shape (the level counts and the flag depend on the matrix only)
    A.rhs = A.c;
    A.xout = x0;
    hipLaunchKernelGGL((wsweep_kernel<kWinNT, kWinBR, true>), dim3((unsigned)nwin), dim3(kWinNT),
                       0, st, A);
    if (hipGetLastError() != hipSuccess) return PCRP_ERR_HIP;
    int h = 0;
    if (hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return PCRP_ERR_HIP;
    P.local = h == 0;
    P.L = L;
    P.transposed = transposed;
    P.uniform = uniform;
    P.rs = rs;
    P.elems = elems;
    P.nwin = nwin;
    P.n = n;
    P.m = m;
    P.nwu = uniform ? sys[0].nw : 0;
    P.wl = wl;
    P.ws = ws;
    return PCRP_OK;
}

int window_plan_execute(const WindowPlan& P, const double* a, const double* b, const double* c,
                        const double* d, const double* e, double* y, int max_sweeps, double tol,
                        int* sweeps_out, hipStream_t st)
{
    if (!P.local) return PCRP_ERR_ARG;
    char* w = static_cast<char*>(P.ws);
    const WLayout& L = P.L;
    SysState* state = reinterpret_cast<SysState*>(w + L.state);
    int* stop = reinterpret_cast<int*>(w + L.stop);
    int* flag = reinterpret_cast<int*>(w + L.flag);
    WSys* dsys = reinterpret_cast<WSys*>(w + L.sys);
    double* x0 = reinterpret_cast<double*>(w + L.x0);
    double* x1 = reinterpret_cast<double*>(w + L.x1);
    double* y0 = reinterpret_cast<double*>(w + L.y0);
    double* six = reinterpret_cast<double*>(w + L.six);
    unsigned char* lev = reinterpret_cast<unsigned char*>(w + L.lev);
    const bool transposed = P.transposed;
    const long long N = P.elems;
    const int m = P.m, n = P.n;
    static std::atomic<int> c_gWin[kMaxDevices];
    const int gWin = resident_grid(c_gWin, wsweep_kernel<kWinNT, kWinBR, false>, kWinNT);
    static std::atomic<int> c_gCopy[kMaxDevices];
    const int gCopy = resident_grid(c_gCopy, wcopy_kernel, 256);
    const unsigned sblocks = (unsigned)((m + 255) / 256);
    hipLaunchKernelGGL(init_state_kernel, dim3(sblocks), dim3(256), 0, st, state, stop, m);
    const dim3 tgrid((unsigned)((n + kTrRows - 1) / kTrRows), (unsigned)((m + kTrSys - 1) / kTrSys));
    if (transposed)
        hipLaunchKernelGGL(wtranspose_in_kernel, dim3(tgrid.x, tgrid.y, 1), dim3(256), 0, st,
                           Six{{y, nullptr, nullptr, nullptr, nullptr, nullptr}}, six + 5 * N, n,
                           m);
    WArgs A{transposed ? six : a, transposed ? six + N : b, transposed ? six + 2 * N : c,
            transposed ? six + 3 * N : d, transposed ? six + 4 * N : e, nullptr, nullptr, nullptr,
            nullptr, P.rs, dsys, state, stop, flag, lev, m, (int)P.nwin, P.uniform ? 1 : 0,
            P.nwu, 0, P.wl, win_halo(P.wl)};
    if (!P.uniform) A.wmap = reinterpret_cast<const int*>(w + L.wmap);
    {
        const auto al16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
        A.vec = P.rs == 1 && al16(A.a) && al16(A.b) && al16(A.c) && al16(A.d) && al16(A.e) &&
                al16(y) && al16(x0) && al16(x1) && al16(y0) && al16(six) ? 1 : 0;
    }
    double* const xbuf[3] = {x0, x1, transposed ? x0 : y};
    const int cyc = transposed ? 2 : 3;
    for (int it = 0; it <= max_sweeps; ++it) {
        A.rhs = transposed ? six + 5 * N : (it == 0 ? y : y0);
        A.y0 = transposed || it > 0 ? nullptr : y0;
        A.xin = it == 0 ? nullptr : xbuf[(it - 1) % cyc];
        A.xout = xbuf[it % cyc];
        hipLaunchKernelGGL((wsweep_kernel<kWinNT, kWinBR, false>),
                           dim3(walk_grid(P.nwin, it < kFullGridSweeps ? 0 : gWin)), dim3(kWinNT),
                           0, st, A);
        hipLaunchKernelGGL(decide_kernel, dim3(sblocks), dim3(256), 0, st, state, stop, m,
                           max_sweeps, tol, sweeps_out);
        if (hipGetLastError() != hipSuccess) return PCRP_ERR_HIP;
    }
    if (transposed)
        hipLaunchKernelGGL(wtranspose_out_kernel, dim3(tgrid.x, tgrid.y), dim3(256), 0, st, xbuf[0],
                           xbuf[1], state, y, n, m);
    else
        hipLaunchKernelGGL(wcopy_kernel, dim3(walk_grid(P.nwin, gCopy)), dim3(256), 0, st, A,
                           xbuf[0], xbuf[1], y);
    return hipGetLastError() == hipSuccess ? PCRP_OK : PCRP_ERR_HIP;
}

}  // namespace

// (a mutex: solves read them while another thread may tune)
static F32RTune g_f32r = {4, 1LL << 23};
static std::mutex g_f32r_mu;

F32RTune f32r_tune_get()
{
    std::lock_guard<std::mutex> lk(g_f32r_mu);
    return g_f32r;
}

int f32r_tune_set(const F32RTune& t)
{
    if (t.win_levels < 2 || t.win_levels > kWinLMax || t.window_min_unknowns < 0) return -1;
    std::lock_guard<std::mutex> lk(g_f32r_mu);
    g_f32r = t;
    return 0;
}

size_t window_f32r_workspace(long long elems, int m, bool transposed)
{
    return wlayout(elems, m, transposed).bytes;
}

size_t window_f32r_workspace_uniform(int n, int m, int layout)
{
    return wlayout((long long)n * m, m, layout == PCRP_LAYOUT_INTERLEAVED && m > 1).bytes;
}

int window_f32r_uniform(const SolveArgsF32R& A, void* ws, size_t ws_bytes, hipStream_t st,
                        bool& fell_back)
{
    fell_back = true;
    if (capturing(st)) return PCRP_OK;   // the host read-back cannot be captured
    std::vector<WSys> sys(A.m);
    const int nb = (A.n + 1) / 2;
    for (int j = 0; j < A.m; ++j) sys[j] = WSys{(long long)j * A.ss, 0, A.n, nb, 0, 0};
    return window_run(A.a, A.b, A.c, A.d, A.e, A.y, sys, true, A.rs, (long long)A.n * A.m,
                      window_transposed(A.m, A.rs), A.n, A.max_sweeps, A.tol, A.sweeps_out,
                      ws, ws_bytes, st, fell_back);
}

int window_f32r_uneven(const double* a, const double* b, const double* c, const double* d,
                       const double* e, double* y, const int64_t* h_off, const int* h_n, int m,
                       int max_sweeps, double tol, int* sweeps_out, void* ws, size_t ws_bytes,
                       hipStream_t st, bool& fell_back)
{
    fell_back = true;
    if (capturing(st)) return PCRP_OK;
    std::vector<WSys> sys(m);
    long long elems = 0;
    for (int j = 0; j < m; ++j) {
        if (h_n[j] < 1 || h_off[j] < 0) return PCRP_ERR_INVALID_SIZE;
        sys[j] = WSys{h_off[j], 0, h_n[j], (h_n[j] + 1) / 2, 0, 0};
        elems = std::max(elems, (long long)h_off[j] + h_n[j]);
    }
    return window_run(a, b, c, d, e, y, sys, false, 1, elems, false, 0, max_sweeps, tol,
                      sweeps_out, ws, ws_bytes, st, fell_back);
}

int window_f32r_plan_uniform(const SolveArgsF32R& A, void* ws, size_t ws_bytes, hipStream_t st,
                             WindowPlan& P)
{
    std::vector<WSys> sys(A.m);
    const int nb = (A.n + 1) / 2;
    for (int j = 0; j < A.m; ++j) sys[j] = WSys{(long long)j * A.ss, 0, A.n, nb, 0, 0};
    return window_plan(A.a, A.b, A.c, A.d, A.e, sys, true, A.rs, (long long)A.n * A.m,
                       window_transposed(A.m, A.rs), A.n, ws, ws_bytes, st, P);
}

int window_f32r_plan_uneven(const double* a, const double* b, const double* c, const double* d,
                            const double* e, const int64_t* h_off, const int* h_n, int m,
                            void* ws, size_t ws_bytes, hipStream_t st, WindowPlan& P)
{
    std::vector<WSys> sys(m);
    long long elems = 0;
    for (int j = 0; j < m; ++j) {
        if (h_n[j] < 1 || h_off[j] < 0) return PCRP_ERR_INVALID_SIZE;
        sys[j] = WSys{h_off[j], 0, h_n[j], (h_n[j] + 1) / 2, 0, 0};
        elems = std::max(elems, (long long)h_off[j] + h_n[j]);
    }
    return window_plan(a, b, c, d, e, sys, false, 1, elems, false, 0, ws, ws_bytes, st, P);
}

int window_f32r_plan_execute(const WindowPlan& P, const double* a, const double* b,
                             const double* c, const double* d, const double* e, double* y,
                             int max_sweeps, double tol, int* sweeps_out, hipStream_t st)
{
    return window_plan_execute(P, a, b, c, d, e, y, max_sweeps, tol, sweeps_out, st);
}

}  // namespace pcrp
