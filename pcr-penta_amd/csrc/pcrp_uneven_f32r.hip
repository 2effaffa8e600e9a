// Uneven batches, FP32 inner solve + FP64 iterative refinement: thousands of
// systems of different n in one call.
//
// No counterpart in the reference (FP64-only, one (m, n) at a time,
// main.cu:245-260). Systems are stored contiguously, system j at rows
// off[j] .. off[j]+n[j]-1 (the reference's per-system layout, kernel.cu:43).
// The inner solve is the 2x2-block PCR of pcrp_block_f32.h:
//   * every system becomes nb = ceil(n/2) level-0 block rows, back to back
//     (block row K of system j at boff[j] + K);
//   * a system with nb > 1024 first runs S = ceil(log2(nb / 1024)) levels as
//     HBM pass rounds shared by all such systems (one launch per round, a
//     table maps blocks to systems). For a single system the residue-split
//     position of block row l of sub-system q is q + l*2^Sin, the block row's
//     own index, so passes read and write the same positions;
//   * then every sub-system (or whole small system) is one job of the LDS
//     stage, which solves it and adds dx to x in FP64. A system's last pass
//     round stores each of its stage sub-systems consecutively (sub-system
//     R at pboff[j] + R * lst[j], lst[j] = ceil(nb / 2^S)), so a stage
//     workgroup reads consecutive records; the pass output sets give every
//     system 2^S * lst[j] >= nb positions from pboff[j] on.
// One refinement sweep: residual (FP64, stored as the FP32 block right-hand
// side), pass rounds, stage, per-system stopping decision; the stop flags
// make later launches skip converged systems.
#include "pcr_penta.h"
#include "pcrp_block_f32.h"
#include "pcrp_internal.h"

#include <algorithm>
#include <vector>

namespace pcrp {

namespace {

constexpr int kStageNT = 512, kStageBR = 2;
constexpr int kStageMax = kStageNT * kStageBR;   // block rows per LDS job
constexpr int kFuse = 4;                         // levels per HBM pass round

// Pass shapes: the first round has one sub-system per system (G = 1); later
// rounds put G neighbouring residue sub-systems side by side.
constexpr int kTB0 = 1024, kG0 = 1;
constexpr int kTB1 = 256, kG1 = 4;
constexpr int kRPT = 2;

struct SysInfo {
    long long off, boff;   // first row, first level-0 block row
    long long pboff;       // first position in the pass output sets
    int n, nb, lst, pad;   // lst: block rows per stage sub-system (split systems)
};

struct PassEntry {
    long long ibase, obase, block0;   // input / output position of block row 0
    int nb, k, tiles, sys, last, lst;
};

struct Job {
    int sys, R, S, nb, src, pad;
};

// Level-0 block rows (s[0]) and the output set of each pass round (s[r + 1]).
constexpr int kMaxRounds = 7;   // 28 HBM levels: block rows up to 2^38
struct BRowSets {
    BRows s[kMaxRounds + 1];
};

struct Plan {
    std::vector<SysInfo> sys;
    std::vector<Job> jobs;
    std::vector<std::vector<PassEntry>> rounds;
    std::vector<long long> round_blocks;
    long long rows = 0, nbt = 0, nbt_pad = 0;
    size_t off_state = 0, off_stop = 0, off_sys = 0, off_jobs = 0, off_y0 = 0, off_lvl0 = 0,
           off_sets = 0, bytes = 0;
    std::vector<size_t> round_off;
};

constexpr size_t kAlign = 256;
size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

int levels_for(int nb)
{
    int S = 0;
    while (((nb + (1 << S) - 1) >> S) > kStageMax) ++S;
    return S;
}

int plan_uneven(const int64_t* h_off, const int* h_n, int m, Plan& P)
{
    P.sys.resize(m);
    std::vector<int> S(m);
    for (int j = 0; j < m; ++j) {
        if (h_n[j] < 1 || h_off[j] < 0) return PCRP_ERR_INVALID_SIZE;
        const long long end = h_off[j] + (long long)h_n[j];
        if (end > P.rows) P.rows = end;
        const int nb = (h_n[j] + 1) / 2;
        S[j] = levels_for(nb);
        const int lst = (nb + (1 << S[j]) - 1) >> S[j];
        P.sys[j] = SysInfo{h_off[j], P.nbt, P.nbt_pad, h_n[j], nb, lst, 0};
        P.nbt += nb;
        P.nbt_pad += S[j] > 0 ? ((long long)lst << S[j]) : nb;
    }
    if (P.rows > 0x7fffffffLL) return PCRP_ERR_INVALID_SIZE;
    // stage jobs, largest sub-systems first
    for (int j = 0; j < m; ++j) {
        // the stage reads the output set of the system's last pass round
        // (set r + 1 holds round r's output; set 0 the level-0 block rows)
        const int src = (S[j] + kFuse - 1) / kFuse;
        for (int R = 0; R < (1 << S[j]); ++R) {
            const int nbs = (P.sys[j].nb - R + (1 << S[j]) - 1) >> S[j];
            if (nbs > 0) P.jobs.push_back(Job{j, R, S[j], nbs, src, 0});
        }
    }
    std::stable_sort(P.jobs.begin(), P.jobs.end(),
                     [](const Job& x, const Job& y) { return x.nb > y.nb; });
    // pass rounds: round r covers levels [r*kFuse, min(S, (r+1)*kFuse))
    for (int r = 0;; ++r) {
        const int Sin = r * kFuse;
        const int TB = r == 0 ? kTB0 : kTB1, G = r == 0 ? kG0 : kG1;
        std::vector<PassEntry> tab;
        long long blocks = 0;
        for (int j = 0; j < m; ++j) {
            if (S[j] <= Sin) continue;
            const int k = std::min(kFuse, S[j] - Sin);
            const int H = (1 << k) - 1, TI = TB - 2 * H;
            const int L = (P.sys[j].nb + (1 << Sin) - 1) >> Sin;
            const int tiles = (L + TI - 1) / TI;
            const int groups = ((1 << Sin) + G - 1) / G;
            tab.push_back(PassEntry{r == 0 ? P.sys[j].boff : P.sys[j].pboff, P.sys[j].pboff,
                                    blocks, P.sys[j].nb, k, tiles, j, Sin + k == S[j] ? 1 : 0,
                                    P.sys[j].lst});
            blocks += (long long)groups * ((tiles + 7) / 8) * 8;   // XCD-striped tiles
        }
        if (tab.empty()) break;
        if (r >= kMaxRounds) return PCRP_ERR_INVALID_SIZE;
        if (blocks > 0x7fffffffLL) return PCRP_ERR_INVALID_SIZE;
        P.rounds.push_back(std::move(tab));
        P.round_blocks.push_back(blocks);
    }
    size_t o = 0;
    P.off_state = o; o = align_up(o + (size_t)m * sizeof(SysState));
    P.off_stop = o;  o = align_up(o + (size_t)(m + 1) * sizeof(int));   // flags + running count
    P.off_sys = o;   o = align_up(o + (size_t)m * sizeof(SysInfo));
    P.off_jobs = o;  o = align_up(o + P.jobs.size() * sizeof(Job));
    for (const auto& t : P.rounds) {
        P.round_off.push_back(o);
        o = align_up(o + t.size() * sizeof(PassEntry));
    }
    P.off_y0 = o;    o = align_up(o + (size_t)P.rows * sizeof(double));
    P.off_lvl0 = o;  o = align_up(o + brows_bytes((size_t)P.nbt));
    P.off_sets = o;
    o = align_up(o + P.rounds.size() * brows_bytes((size_t)P.nbt_pad));
    P.bytes = o;
    return PCRP_OK;
}

// System owning level-0 block row e: last j with boff <= e.
__device__ __forceinline__ int sys_of(const SysInfo* __restrict__ sys, int m, long long e)
{
    int lo = 0, hi = m - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sys[mid].boff <= e) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// The same for the block rows e of a workgroup's range starting at e0: one
// binary search per workgroup (on e0, uniform, so scalar loads), then a
// short walk: a workgroup spans few systems (each has at least one block row,
// and the batches this serves have tens to thousands).
__device__ __forceinline__ int sys_of_row(const SysInfo* __restrict__ sys, int m, long long nbt,
                                          long long e0, long long e)
{
    int j = sys_of(sys, m, e0 < nbt ? e0 : nbt - 1);
    while (j + 1 < m && sys[j + 1].boff <= e) ++j;
    return j;
}

__global__ void uprep_kernel(const double* __restrict__ a, const double* __restrict__ b,
                             const double* __restrict__ c, const double* __restrict__ d,
                             const double* __restrict__ e, const double* __restrict__ y,
                             BRows w, double* __restrict__ y0, const SysInfo* __restrict__ sys,
                             int m, long long nbt)
{
    const long long el = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (el >= nbt) return;
    const SysInfo S = sys[sys_of_row(sys, m, nbt, (long long)blockIdx.x * blockDim.x, el)];
    const long long K = el - S.boff;
    double rw[2][6];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const long long i = 2 * K + h;
        rw[h][0] = rw[h][1] = rw[h][3] = rw[h][4] = rw[h][5] = 0.0;
        rw[h][2] = 1.0;   // identity row past the end of an odd n
        if (i < S.n) {
            const long long g = S.off + i;
            load_row(a, b, c, d, e, y, g, i, S.n, rw[h]);
            y0[g] = rw[h][5];
        }
    }
    store_brow(w, el, brow_from_rows(rw));
}

__global__ void uresidual_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                 const double* __restrict__ c, const double* __restrict__ d,
                                 const double* __restrict__ e, const double* __restrict__ y0,
                                 const double* __restrict__ x, float2* __restrict__ r,
                                 const int* __restrict__ stop,
                                 const SysInfo* __restrict__ sys, int m, long long nbt)
{
    if (stop[m] == 0) return;   // every system has stopped
    for (long long e0 = (long long)blockIdx.x * blockDim.x; e0 < nbt;
         e0 += (long long)gridDim.x * blockDim.x) {
        const long long el = e0 + threadIdx.x;
        if (el >= nbt) continue;
        const int j = sys_of_row(sys, m, nbt, e0, el);
        if (stop[j]) continue;
        const SysInfo S = sys[j];
        const long long K = el - S.boff;
        float rv[2] = {0.f, 0.f};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const long long i = 2 * K + h;
            if (i < S.n) rv[h] = (float)residual_row(a, b, c, d, e, y0, x, S.off + i, i, S.n, 1);
        }
        r[el] = make_float2(rv[0], rv[1]);
    }
}

// One pass round: tab lists the systems that still need levels; a block finds
// its system by binary search on block0. Round r reads set r (0: level 0) and
// writes set r + 1.
template <int TB, int G, int RPT>
__global__ __launch_bounds__(TB* G / RPT) void upass_kernel(
    BRows src, BRows dst, const PassEntry* __restrict__ tab, int ntab, int Sin,
    long long nblocks, const int* __restrict__ stop, int m, int y_only)
{
    __shared__ LdsRec<TB * G> sR;
    if (stop[m] == 0) return;   // every system has stopped
    // logical blocks b, b + grid, ... (grid a multiple of 8: same XCD)
    for (long long b = blockIdx.x; b < nblocks; b += gridDim.x) {
        int lo = 0, hi = ntab - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tab[mid].block0 <= b) lo = mid; else hi = mid - 1;
        }
        const PassEntry P = tab[lo];
        const int Qin = 1 << Sin;
        int grp, tile;
        if (!btile_of((int)(b - P.block0), (Qin + G - 1) / G, P.tiles, grp, tile)) continue;
        const int k = P.k;
        const int H = (1 << k) - 1, TI = TB - 2 * H;
        const int lr0 = threadIdx.x / G, g = threadIdx.x - lr0 * G;
        const int q = grp * G + g;
        const int Lq = q < Qin ? (P.nb - q + Qin - 1) >> Sin : 0;
        const bool run = q < Qin && !stop[P.sys];
        if (!__syncthreads_or(run)) continue;   // the system has converged
        // block row l of sub-system q is the system's block row q + l*2^Sin;
        // the last round regroups its output by the stage's sub-system
        auto in_at = [&](int l) { return P.ibase + q + ((long long)l << Sin); };
        auto out_at = [&](int l) {
            if (!P.last) return P.obase + q + ((long long)l << Sin);
            const long long qo = q + ((long long)(l & ((1 << k) - 1)) << Sin);
            return P.obase + qo * P.lst + (l >> k);
        };
        bpass_tile<TB, G, RPT>(src, dst, run, y_only != 0, tile * TI - H + lr0, Lq, k, H, TI,
                               in_at, out_at, sR);
    }
}

template <int NT, int BR>
__global__ __launch_bounds__(NT) void ustage_kernel(BRowSets sets,
                                                    const Job* __restrict__ jobs,
                                                    const SysInfo* __restrict__ sys,
                                                    double* __restrict__ x,
                                                    SysState* __restrict__ st,
                                                    const int* __restrict__ stop, int m,
                                                    int njobs, int first)
{
    __shared__ LdsRec<NT * BR> sR;
    __shared__ unsigned long long red[2][NT / 64];
    __shared__ int amx[NT / 64];
    if (stop[m] == 0) return;   // every system has stopped
    // logical blocks b, b + grid, ... (grid a multiple of 8: same XCD)
    for (int b = blockIdx.x; b < njobs; b += gridDim.x) {
        // consecutive jobs are neighbouring residues of one system (their
        // block rows share cache lines): keep runs of 32 of them on one XCD
        const Job J = jobs[xcd_remap_c<32>(b, njobs)];
        if (stop[J.sys]) continue;   // uniform over the workgroup
        const SysInfo S = sys[J.sys];
        const BRows in = sets.s[J.src];
        bstage_solve<NT, BR>(
            in, J.nb,
            [&](int l) {
                return J.src == 0 ? S.boff + l : S.pboff + (long long)J.R * S.lst + l;
            },
            [&](int l, int h) {
                const long long i = 2 * (J.R + ((long long)l << J.S)) + h;
                return i < S.n ? S.off + i : -1LL;
            },
            x, first != 0, &st[J.sys], sR, red, amx);
    }
}

}  // namespace

size_t uneven_f32r_workspace(const int64_t* h_off, const int* h_n, int m)
{
    Plan P;
    if (plan_uneven(h_off, h_n, m, P) != PCRP_OK) return 0;
    return std::max(P.bytes, window_f32r_workspace(P.rows, m, false));
}

namespace {

// Pass/stage route, tables: the host plan's system, job and pass-round
// tables to the front of the workspace in one copy (from a host buffer that
// must outlive the copy: `host`).
int upload_tables(const Plan& P, void* ws, std::vector<char>& host, hipStream_t st)
{
    char* w = static_cast<char*>(ws);
    host.assign(P.off_y0, 0);
    std::copy(reinterpret_cast<const char*>(P.sys.data()),
              reinterpret_cast<const char*>(P.sys.data() + P.sys.size()), host.begin() + P.off_sys);
    std::copy(reinterpret_cast<const char*>(P.jobs.data()),
              reinterpret_cast<const char*>(P.jobs.data() + P.jobs.size()),
              host.begin() + P.off_jobs);
    for (size_t r = 0; r < P.rounds.size(); ++r)
        std::copy(reinterpret_cast<const char*>(P.rounds[r].data()),
                  reinterpret_cast<const char*>(P.rounds[r].data() + P.rounds[r].size()),
                  host.begin() + P.round_off[r]);
    return hipMemcpyAsync(w + P.off_sys, host.data() + P.off_sys, host.size() - P.off_sys,
                          hipMemcpyHostToDevice, st) == hipSuccess
               ? PCRP_OK
               : PCRP_ERR_HIP;
}

int launch_pass_stage(const Plan& P, const double* a, const double* b, const double* c,
                      const double* d, const double* e, double* y, int m, int max_sweeps,
                      double tol, int* sweeps_out, void* ws, hipStream_t st);

}  // namespace

// Plan form of the pass/stage route: tables uploaded once (synchronous),
// launches only at execute.
struct UnevenF32RPlan {
    Plan P;
};

UnevenF32RPlan* uneven_f32r_plan_tables(const int64_t* h_off, const int* h_n, int m, void* ws,
                                        size_t ws_bytes, hipStream_t st, int& rc)
{
    UnevenF32RPlan* U = new UnevenF32RPlan;
    rc = plan_uneven(h_off, h_n, m, U->P);
    if (rc == PCRP_OK && (!ws || ws_bytes < U->P.bytes)) rc = PCRP_ERR_WORKSPACE;
    std::vector<char> host;
    if (rc == PCRP_OK) rc = upload_tables(U->P, ws, host, st);
    if (rc == PCRP_OK && hipStreamSynchronize(st) != hipSuccess) rc = PCRP_ERR_HIP;
    if (rc != PCRP_OK) {
        delete U;
        return nullptr;
    }
    return U;
}

void uneven_f32r_plan_free(UnevenF32RPlan* U) { delete U; }

int uneven_f32r_plan_execute(const UnevenF32RPlan* U, const double* a, const double* b,
                             const double* c, const double* d, const double* e, double* y, int m,
                             int max_sweeps, double tol, int* sweeps_out, void* ws,
                             hipStream_t st)
{
    return launch_pass_stage(U->P, a, b, c, d, e, y, m, max_sweeps, tol, sweeps_out, ws, st);
}

namespace {

int launch_pass_stage(const Plan& P, const double* a, const double* b, const double* c,
                      const double* d, const double* e, double* y, int m, int max_sweeps,
                      double tol, int* sweeps_out, void* ws, hipStream_t st)
{
    char* w = static_cast<char*>(ws);
    SysState* state = reinterpret_cast<SysState*>(w + P.off_state);
    int* stop = reinterpret_cast<int*>(w + P.off_stop);
    const SysInfo* dsys = reinterpret_cast<const SysInfo*>(w + P.off_sys);
    const Job* djobs = reinterpret_cast<const Job*>(w + P.off_jobs);
    double* y0 = reinterpret_cast<double*>(w + P.off_y0);
    BRowSets sets = {};
    carve_brows(w + P.off_lvl0, (size_t)P.nbt, sets.s[0]);
    char* p = w + P.off_sets;
    for (size_t r = 0; r < P.rounds.size(); ++r)
        p = carve_brows(p, (size_t)P.nbt_pad, sets.s[r + 1]);
    const unsigned bblocks = (unsigned)((P.nbt + 255) / 256);
    const unsigned sblocks = (unsigned)((m + 255) / 256);
    static std::atomic<int> c_gPass0[kMaxDevices];
    const int gPass0 = resident_grid(c_gPass0, upass_kernel<kTB0, kG0, kRPT>, kTB0 * kG0 / kRPT);
    static std::atomic<int> c_gPass1[kMaxDevices];
    const int gPass1 = resident_grid(c_gPass1, upass_kernel<kTB1, kG1, kRPT>, kTB1 * kG1 / kRPT);
    static std::atomic<int> c_gStage[kMaxDevices];
    const int gStage = resident_grid(c_gStage, ustage_kernel<kStageNT, kStageBR>, kStageNT);
    static std::atomic<int> c_gRes[kMaxDevices];
    const int gRes = resident_grid(c_gRes, uresidual_kernel, 256);
    hipLaunchKernelGGL(init_state_kernel, dim3(sblocks), dim3(256), 0, st, state, stop, m);
    hipLaunchKernelGGL(uprep_kernel, dim3(bblocks), dim3(256), 0, st, a, b, c, d, e, y, sets.s[0],
                       y0, dsys, m, P.nbt);
    if (hipGetLastError() != hipSuccess) return PCRP_ERR_HIP;
    for (int it = 0; it <= max_sweeps; ++it) {
        for (size_t r = 0; r < P.rounds.size(); ++r) {
            // from the second sweep on, round r's output set already holds the
            // matrix records of sweep 0: only the right-hand side is stored
            const BRows& src = sets.s[r];
            const BRows& dst = sets.s[r + 1];
            const int y_only = it > 0 ? 1 : 0;
            const PassEntry* tab = reinterpret_cast<const PassEntry*>(w + P.round_off[r]);
            const int ntab = (int)P.rounds[r].size();
            const unsigned nblk = (unsigned)P.round_blocks[r];
            const int full = it < kFullGridSweeps;
            if (r == 0)
                hipLaunchKernelGGL((upass_kernel<kTB0, kG0, kRPT>),
                                   dim3(walk_grid(nblk, full ? 0 : gPass0)),
                                   dim3(kTB0 * kG0 / kRPT), 0, st, src, dst, tab, ntab,
                                   (int)r * kFuse, (long long)nblk, stop, m, y_only);
            else
                hipLaunchKernelGGL((upass_kernel<kTB1, kG1, kRPT>),
                                   dim3(walk_grid(nblk, full ? 0 : gPass1)),
                                   dim3(kTB1 * kG1 / kRPT), 0, st, src, dst, tab, ntab,
                                   (int)r * kFuse, (long long)nblk, stop, m, y_only);
        }
        hipLaunchKernelGGL((ustage_kernel<kStageNT, kStageBR>),
                           dim3(walk_grid((long long)P.jobs.size(),
                                          it < kFullGridSweeps ? 0 : gStage)),
                           dim3(kStageNT), 0,
                           st, sets, djobs, dsys, y, state, stop, m, (int)P.jobs.size(),
                           it == 0 ? 1 : 0);
        hipLaunchKernelGGL(decide_kernel, dim3(sblocks), dim3(256), 0, st, state, stop, m,
                           max_sweeps, tol, sweeps_out);
        if (it < max_sweeps)
            hipLaunchKernelGGL(uresidual_kernel,
                               dim3(walk_grid(bblocks, it + 1 < kFullGridSweeps ? 0 : gRes)),
                               dim3(256), 0, st,
                               a, b, c, d, e, y0, y, sets.s[0].y, stop, dsys, m, P.nbt);
        if (hipGetLastError() != hipSuccess) return PCRP_ERR_HIP;
    }
    return PCRP_OK;
}

}  // namespace

int uneven_f32r_solve(const double* a, const double* b, const double* c, const double* d,
                      const double* e, double* y, const int64_t* h_off, const int* h_n, int m,
                      int max_sweeps, double tol, int* sweeps_out, void* ws, size_t ws_bytes,
                      hipStream_t st)
{
    Plan P;
    int rc = plan_uneven(h_off, h_n, m, P);
    if (rc != PCRP_OK) return rc;
    if (!ws || ws_bytes < std::max(P.bytes, window_f32r_workspace(P.rows, m, false)))
        return PCRP_ERR_WORKSPACE;
    // diagonally dominant batches: one windowed pass per sweep
    bool fell_back = true;
    rc = window_f32r_uneven(a, b, c, d, e, y, h_off, h_n, m, max_sweeps, tol, sweeps_out, ws,
                            ws_bytes, st, fell_back);
    if (rc != PCRP_OK || !fell_back) return rc;
    // the tables are copied from a host buffer that lives to the end of this
    // call (the copy is staged before hipMemcpyAsync returns for pageable memory)
    std::vector<char> host;
    rc = upload_tables(P, ws, host, st);
    if (rc != PCRP_OK) return rc;
    return launch_pass_stage(P, a, b, c, d, e, y, m, max_sweeps, tol, sweeps_out, ws, st);
}

}  // namespace pcrp

extern "C" {

int pcrp_solve_uneven_f32r_f64(const double* a, const double* b, const double* c, const double* d,
                               const double* e, double* y, const int64_t* h_offsets,
                               const int* h_n, int m, int max_sweeps, double tol, int* sweeps_out,
                               void* ws, size_t ws_bytes, void* stream)
{
    if (m < 1) return PCRP_ERR_INVALID_SIZE;
    if (!a || !b || !c || !d || !e || !y || !h_offsets || !h_n) return PCRP_ERR_NULL;
    if (max_sweeps < 0 || !(tol >= 0.0)) return PCRP_ERR_ARG;
    return pcrp::uneven_f32r_solve(a, b, c, d, e, y, h_offsets, h_n, m, max_sweeps, tol,
                                   sweeps_out, ws, ws_bytes, static_cast<hipStream_t>(stream));
}

}  // extern "C"
