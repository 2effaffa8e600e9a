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
//            sub-system (R, j) of block rows R + 2^S l is q = R*m + j; the
//            last pass stores each sub-system's block rows consecutively;
//   stage  : one workgroup per sub-system, the remaining levels in LDS, then
//            X = D^-1 Y, and x += dx (FP64) at the caller's elements.
// One refinement sweep (every kernel skips systems already stopped):
//   residual : r = y - A x in FP64 from the FP64 inputs, stored as the FP32
//              block right-hand side (first sweep: r = y, x = 0);
//   passes + stage (factor and solve together: recomputing the FP32 factor
//              moves fewer bytes than storing its multipliers; each pass
//              has its own output set, so from the second sweep on a pass
//              reads the matrix records sweep 0 left there and stores only
//              the right-hand side);
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
#include "pcrp_block_f32.h"
#include "pcrp_internal.h"

#include <algorithm>

namespace pcrp {

namespace {

constexpr int kStageNT = 512, kStageBR = 2;
constexpr int kStageMax = kStageNT * kStageBR;   // block rows per LDS sub-system
constexpr int kFuse = 4;                         // levels per HBM pass
constexpr int kPassTB = 256, kPassG = 4, kPassRPT = 2;

struct Layout {
    size_t state, stop, y0, lvl0, sets, bytes;
    size_t per;        // block rows per pass output set
    long long nb0;     // block rows per system
    int S, rounds;     // HBM levels, passes (one output set each)
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
    L.stop = o;  o = align_up(o + (size_t)(m + 1) * sizeof(int));   // flags + running count
    L.y0 = o;    o = align_up(o + N * sizeof(double));
    L.rounds = (L.S + kFuse - 1) / kFuse;
    L.lvl0 = o;  o = align_up(o + brows_bytes(NB));
    L.sets = o;  o = align_up(o + (size_t)L.rounds * brows_bytes(L.per));
    L.bytes = o;
    return L;
}

// Level-0 block rows (block row K of system j at K*m + j) from the FP64
// rows 2K, 2K+1 of the caller's arrays; y0 keeps y.
__global__ void prep_kernel(const double* __restrict__ a, const double* __restrict__ b,
                            const double* __restrict__ c, const double* __restrict__ d,
                            const double* __restrict__ e, const double* __restrict__ y,
                            BRows w, double* __restrict__ y0, int n, int m, long long rs,
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
        rw[h][0] = rw[h][1] = rw[h][3] = rw[h][4] = rw[h][5] = 0.0;
        rw[h][2] = 1.0;   // identity row past the end of an odd n
        if (i < n) {
            const long long g = (long long)j * ss + i * rs;
            load_row(a, b, c, d, e, y, g, i, n, rw[h]);
            y0[g] = rw[h][5];
        }
    }
    store_brow(w, el, brow_from_rows(rw));
}

// r = y0 - A x in FP64, as the FP32 right-hand side of level-0 block row K.
__global__ void residual_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                const double* __restrict__ c, const double* __restrict__ d,
                                const double* __restrict__ e, const double* __restrict__ y0,
                                const double* __restrict__ x, float2* __restrict__ r,
                                const int* __restrict__ stop, int n, int m,
                                long long rs, long long ss, long long nb0)
{
    if (stop[m] == 0) return;   // every system has stopped
    const long long total = nb0 * m, step = (long long)gridDim.x * blockDim.x;
    for (long long el = (long long)blockIdx.x * blockDim.x + threadIdx.x; el < total; el += step) {
        const long long K = el / m;
        const int j = (int)(el - K * m);
        if (stop[j]) continue;
        float rv[2] = {0.f, 0.f};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const long long i = 2 * K + h;
            if (i < n) {
                const long long g = (long long)j * ss + i * rs;
                rv[h] = (float)residual_row(a, b, c, d, e, y0, x, g, i, n, rs);
            }
        }
        r[el] = make_float2(rv[0], rv[1]);
    }
}

// One HBM pass: k levels over a tile of TB block rows of G sub-systems.
// Current sub-system q = R*m0 + j holds block rows R + 2^Sin l of system j,
// element l at q + l*Qin; output regrouped by residue (q' = (l mod 2^k)*Qin
// + q at l >> k).
// The last pass writes sub-system-major instead (sub-system q' of block rows
// l' at q' * Lst + l'), so the stage, one workgroup per sub-system, reads
// consecutive records.
template <int TB, int G, int RPT>
__global__ __launch_bounds__(TB* G / RPT) void bpass_kernel(
    BRows in, BRows out, long long nb0, int m0, int Sin, int k, int Qin, int tiles,
    long long nblocks, const int* __restrict__ stop, int y_only, int last, long long Lst)
{
    __shared__ LdsRec<TB * G> sR;
    if (stop[m0] == 0) return;   // every system has stopped
    const int H = (1 << k) - 1;
    const int TI = TB - 2 * H;
    // logical blocks b, b + grid, ... (grid a multiple of 8: same XCD)
    for (long long b = blockIdx.x; b < nblocks; b += gridDim.x) {
        int grp, tile;
        if (!btile_of((int)b, (Qin + G - 1) / G, tiles, grp, tile)) continue;
        const int lr0 = threadIdx.x / G, g = threadIdx.x - lr0 * G;
        const int q = grp * G + g;
        const int R = q / m0;
        const int Lq = (int)((nb0 - R + (1LL << Sin) - 1) >> Sin);
        const bool run = q < Qin && !stop[q - R * m0];
        if (!__syncthreads_or(run)) continue;   // every sub-system of this tile has converged
        const long long Qout = (long long)Qin << k;
        bpass_tile<TB, G, RPT>(
            in, out, run, y_only != 0, tile * TI - H + lr0, Lq, k, H, TI,
            [&](int l) { return (long long)q + (long long)l * Qin; },
            [&](int l) {
                const long long qo = (long long)(l & ((1 << k) - 1)) * Qin + q;
                return last ? qo * Lst + (l >> k) : (long long)(l >> k) * Qout + qo;
            },
            sR);
    }
}

// LDS stage: one workgroup per sub-system q = R*m0 + j (element l at
// q * Lst + l), the remaining levels, X = D^-1 Y, x += dx at rows 2K, 2K+1,
// K = R + 2^S l.
template <int NT, int BR>
__global__ __launch_bounds__(NT) void bstage_kernel(
    BRows in, int Q, int S, long long Lst, long long nb0, int n0, int m0, long long xrs,
    long long xss, double* __restrict__ x, SysState* __restrict__ st,
    const int* __restrict__ stop, int first)
{
    __shared__ LdsRec<NT * BR> sR;
    __shared__ unsigned long long red[2][NT / 64];
    __shared__ int amx[NT / 64];
    if (stop[m0] == 0) return;   // every system has stopped
    // logical blocks b, b + grid, ... (grid a multiple of 8: same XCD)
    for (int b = blockIdx.x; b < Q; b += gridDim.x) {
        const int q = xcd_remap_c<32>(b, Q);   // neighbouring sub-systems share x's lines
        const int R = q / m0, j = q - R * m0;
        if (stop[j]) continue;   // uniform over the workgroup
        const int nb = (int)((nb0 - R + (1LL << S) - 1) >> S);
        bstage_solve<NT, BR>(
            in, nb, [&](int l) { return (long long)q * Lst + l; },
            [&](int l, int h) {
                const long long i = 2 * (R + ((long long)l << S)) + h;
                return i < n0 ? (long long)j * xss + i * xrs : -1LL;
            },
            x, first != 0, &st[j], sR, red, amx);
    }
}

}  // namespace

size_t split_f32r_workspace(int n, int m)
{
    return std::max(plan(n, m).bytes,
                    std::max(window_f32r_workspace_uniform(n, m, PCRP_LAYOUT_INTERLEAVED),
                             window_f32r_workspace_uniform(n, m, PCRP_LAYOUT_CONTIGUOUS)));
}

int split_f32r_solve(const SolveArgsF32R& A, int layout, void* ws, size_t ws_bytes,
                     hipStream_t st)
{
    (void)layout;
    // diagonally dominant batches: one windowed pass per sweep
    bool fell_back = true;
    const int wrc = window_f32r_uniform(A, ws, ws_bytes, st, fell_back);
    if (wrc != PCRP_OK) return wrc == PCRP_ERR_WORKSPACE ? -3 : -1;
    if (!fell_back) return 0;
    return split_f32r_passes(A, ws, ws_bytes, st);
}

// The pass/stage form alone (kernel launches only: capturable).
int split_f32r_passes(const SolveArgsF32R& A, void* ws, size_t ws_bytes, hipStream_t st)
{
    const int n0 = A.n, m0 = A.m;
    const Layout P = plan(n0, m0);
    const int S = P.S;
    if (!ws || ws_bytes < P.bytes) return -3;
    char* w = static_cast<char*>(ws);
    SysState* state = reinterpret_cast<SysState*>(w + P.state);
    int* stop = reinterpret_cast<int*>(w + P.stop);
    double* y0 = reinterpret_cast<double*>(w + P.y0);
    const long long NB = P.nb0 * m0;
    BRows lvl0, sets[32];
    if (P.rounds > 32) return -2;
    carve_brows(w + P.lvl0, (size_t)NB, lvl0);
    char* p = w + P.sets;
    for (int r = 0; r < P.rounds; ++r) p = carve_brows(p, P.per, sets[r]);
    if (S < 1) return -2;   // the split path always runs at least one pass
    constexpr int kPassThreads = kPassTB * kPassG / kPassRPT;
    static std::atomic<int> c_gPass[kMaxDevices];
    const int gPass = resident_grid(c_gPass, bpass_kernel<kPassTB, kPassG, kPassRPT>, kPassThreads);
    static std::atomic<int> c_gStage[kMaxDevices];
    const int gStage = resident_grid(c_gStage, bstage_kernel<kStageNT, kStageBR>, kStageNT);
    static std::atomic<int> c_gRes[kMaxDevices];
    const int gRes = resident_grid(c_gRes, residual_kernel, 256);
    const long long Lst = (P.nb0 + (1LL << S) - 1) >> S;   // block rows per stage sub-system
    const unsigned bblocks = (unsigned)((NB + 255) / 256);
    const unsigned sblocks = (unsigned)((m0 + 255) / 256);
    hipLaunchKernelGGL(init_state_kernel, dim3(sblocks), dim3(256), 0, st, state, stop, m0);
    hipLaunchKernelGGL(prep_kernel, dim3(bblocks), dim3(256), 0, st, A.a, A.b, A.c, A.d, A.e,
                       A.y, lvl0, y0, n0, m0, A.rs, A.ss, P.nb0);
    if (hipGetLastError() != hipSuccess) return -1;
    for (int it = 0; it <= A.max_sweeps; ++it) {
        BRows in = lvl0;
        int Sin = 0, Qin = m0;
        for (int r = 0; Sin < S; ++r) {
            const int k = (S - Sin < kFuse) ? S - Sin : kFuse;
            const int H = (1 << k) - 1, TI = kPassTB - 2 * H;
            const long long L = (P.nb0 + (1LL << Sin) - 1) >> Sin;
            const int tiles = (int)((L + TI - 1) / TI);
            const int groups = (Qin + kPassG - 1) / kPassG;
            const long long blocks = (long long)groups * ((tiles + 7) / 8) * 8;
            if (blocks > 0x7fffffffLL) return -2;
            // from the second sweep on, pass r's output set already holds the
            // matrix records of sweep 0: only the right-hand side is stored
            const BRows out = sets[r];
            const unsigned grid = walk_grid(blocks, it < kFullGridSweeps ? 0 : gPass);
            hipLaunchKernelGGL((bpass_kernel<kPassTB, kPassG, kPassRPT>), dim3(grid),
                               dim3(kPassThreads), 0, st, in, out, P.nb0, m0, Sin, k, Qin, tiles,
                               blocks, stop, it > 0 ? 1 : 0, Sin + k == S ? 1 : 0, Lst);
            Sin += k;
            Qin <<= k;
            in = out;
        }
        const unsigned sgrid = walk_grid(Qin, it < kFullGridSweeps ? 0 : gStage);
        hipLaunchKernelGGL((bstage_kernel<kStageNT, kStageBR>), dim3(sgrid), dim3(kStageNT), 0, st,
                           in, Qin, S, Lst, P.nb0, n0, m0, A.rs, A.ss, A.y, state, stop,
                           it == 0 ? 1 : 0);
        hipLaunchKernelGGL(decide_kernel, dim3(sblocks), dim3(256), 0, st, state, stop, m0,
                           A.max_sweeps, A.tol, A.sweeps_out);
        if (it < A.max_sweeps)
            hipLaunchKernelGGL(residual_kernel,
                               dim3(walk_grid(bblocks, it + 1 < kFullGridSweeps ? 0 : gRes)),
                               dim3(256), 0, st,
                               A.a, A.b, A.c,
                               A.d, A.e, y0, A.y, lvl0.y, stop, n0, m0, A.rs,
                               A.ss, P.nb0);
        if (hipGetLastError() != hipSuccess) return -1;
    }
    return 0;
}

}  // namespace pcrp
