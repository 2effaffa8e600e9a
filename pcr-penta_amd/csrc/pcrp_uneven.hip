// Uneven batches: thousands of systems of different n in one call.
//
// No counterpart in the reference (its driver sweeps one (m, n) at a time,
// main.cu:245-260, and requires n = 1024*2^s). Systems are stored
// contiguously, system j at rows off[j] .. off[j]+n[j]-1 of each array (the
// reference's contiguous layout per system, kernel.cu:43).
//
// Routing (planned on the host from the host copies of off/n):
//   * n <= lds_max: the whole system is one job of the LDS work queue;
//   * n > lds_max: S = ceil(log2(n / lds_max)) reduction steps run as HBM
//     passes (k_fuse steps per pass); all large systems share each pass round
//     (one launch per round, a table maps blocks to systems); then every one of
//     its 2^S residue sub-systems is one LDS job. The passes write residue
//     sub-systems to the system's own region of two pass buffers in the
//     grouped residue layout (res_slot: groups of 4 residues row-interleaved),
//     which the next round and the LDS jobs read.
// All LDS jobs go through one persistent kernel with one work queue per XCD:
// a workgroup drains its own XCD's queue, then the others. Runs of 16
// consecutive residue sub-systems of a system (groups whose rows share cache
// lines) stay in one queue, behind one L2; runs go to the least-loaded queue,
// largest first (longest-processing-time order).
// Results are bit-identical to solving each system alone (same arithmetic,
// same guards; the pass rounds are the split path with one system per batch).
#include "pcr_penta.h"
#include "pcrp_internal.h"

#include <algorithm>
#include <vector>

namespace pcrp {

namespace {

// The block -> entry table of a pass round over `tab` (entry i owns blocks
// tab[i].block0 .. tab[i+1].block0 - 1, the last one up to `blocks`).
__global__ void pass_block_map_kernel(const PassSys* __restrict__ tab, int ntab, long long blocks,
                                      int* __restrict__ bmap)
{
    for (int i = blockIdx.x; i < ntab; i += gridDim.x) {
        const long long b0 = tab[i].block0, b1 = i + 1 < ntab ? tab[i + 1].block0 : blocks;
        for (long long b = b0 + threadIdx.x; b < b1; b += blockDim.x) bmap[b] = i;
    }
}

__global__ void reset_counters_kernel(unsigned int* c)
{
    if (threadIdx.x < kQueues) c[threadIdx.x] = 0u;
}

constexpr size_t kAlign = 256;

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

struct UnevenPlan {
    std::vector<LdsJob> jobs;                   // queue 0 jobs, then queue 1, ...
    int qstart[kQueues + 1] = {0};
    std::vector<std::vector<PassSys>> rounds;   // per pass round
    std::vector<long long> round_blocks;
    std::vector<int> round_shape;
    long long total = 0;                        // elements per array (max off + n)
    long long wtotal = 0;                       // pass-buffer slots per array
    std::vector<long long> woff;                // each split system's pass-buffer region
    bool any_split = false;
    // workspace byte offsets
    size_t off_counter = 0, off_sets = 0, off_jobs = 0, off_rounds = 0, off_arrays = 0, bytes = 0;
    std::vector<size_t> round_off, round_map_off;
};

int plan_uneven(const int64_t* h_off, const int* h_n, int m, UnevenPlan& P)
{
    const SplitTune tune = split_f64_get();
    struct Sys { int j, S, nsub; };
    std::vector<Sys> order;
    order.reserve(m);
    for (int j = 0; j < m; ++j) {
        if (h_n[j] < 1 || h_off[j] < 0) return PCRP_ERR_INVALID_SIZE;
        const long long end = h_off[j] + (long long)h_n[j];
        if (end > P.total) P.total = end;
        const int S = split_f64_levels(h_n[j], tune.lds_max);
        order.push_back({j, S, (h_n[j] + (1 << S) - 1) >> S});
        if (S > 0) P.any_split = true;
    }
    if (P.total > 0x7fffffffLL) return PCRP_ERR_INVALID_SIZE;
    // pass-buffer regions of the split systems (grouped residue layout)
    P.woff.assign(m, -1);
    for (const Sys& s : order)
        if (s.S > 0) {
            P.woff[s.j] = P.wtotal;
            P.wtotal += res_region(h_n[s.j], s.S);
        }
    std::stable_sort(order.begin(), order.end(),
                     [](const Sys& x, const Sys& y) { return x.nsub > y.nsub; });
    // LDS jobs, in chunks: a small system alone, or up to 16 consecutive
    // residue sub-systems of a split system (their rows share cache lines).
    // Chunks go to the least-loaded XCD queue, largest first. (Flat chunk
    // records and one pass that writes every job in place: no per-chunk
    // containers, ~4x faster planning at 10000 systems.)
    struct Chunk { long long work; int sys, q0, q1; };   // q0 < 0: the whole system
    std::vector<Chunk> chunks;
    chunks.reserve(order.size() * 2);
    for (const Sys& s : order) {
        const int n = h_n[s.j];
        if (s.S == 0) {
            chunks.push_back({(long long)n, s.j, -1, -1});
            continue;
        }
        const int Q = 1 << s.S;
        for (int q0 = 0; q0 < Q; q0 += 16) {
            const int q1 = std::min(q0 + 16, Q);
            long long work = 0;
            for (int q = q0; q < q1; ++q) work += (n - q + Q - 1) >> s.S;
            chunks.push_back({work, s.j, q0, q1});
        }
    }
    std::vector<unsigned char> qof(chunks.size());
    long long load[kQueues] = {0};
    int count[kQueues] = {0};
    for (size_t c = 0; c < chunks.size(); ++c) {   // already largest first
        int best = 0;
        for (int x = 1; x < kQueues; ++x)
            if (load[x] < load[best]) best = x;
        load[best] += chunks[c].work;
        qof[c] = (unsigned char)best;
        count[best] += chunks[c].q0 < 0 ? 1 : chunks[c].q1 - chunks[c].q0;
    }
    int pos[kQueues];
    for (int x = 0, acc = 0; x < kQueues; ++x) {
        P.qstart[x] = acc;
        pos[x] = acc;
        acc += count[x];
    }
    P.jobs.resize((size_t)P.qstart[kQueues - 1] + count[kQueues - 1]);
    for (size_t c = 0; c < chunks.size(); ++c) {
        const Chunk& ch = chunks[c];
        const int n = h_n[ch.sys];
        const int steps = pcr_steps(n);
        int& at = pos[qof[c]];
        if (ch.q0 < 0) {
            P.jobs[at++] = LdsJob{h_off[ch.sys], h_off[ch.sys], 0, 0, n, steps, 0, 0};
            continue;
        }
        const int S = split_f64_levels(n, tune.lds_max);
        const int rounds = (S + tune.k_fuse - 1) / tune.k_fuse;
        const int src = 1 + ((rounds - 1) & 1);
        // residue q: read from its group in the pass buffer, x to rows q + i 2^S
        for (int q = ch.q0; q < ch.q1; ++q)
            P.jobs[at++] = LdsJob{P.woff[ch.sys] + res_slot(q, 0, S, n), h_off[ch.sys] + q,
                                  res_group_log2(S), S, (n - q + (1 << S) - 1) >> S, steps - S,
                                  src, 0};
    }
    P.qstart[kQueues] = (int)P.jobs.size();
    // HBM pass rounds: round r covers steps [r*k_fuse, min(S, (r+1)*k_fuse))
    for (int r = 0;; ++r) {
        const int Sin = r * tune.k_fuse;
        // first round: one sub-system per system, whole-row tiles; later rounds:
        // several residue sub-systems side by side for coalesced row loads
        const int shape = (r == 0) ? pass_shape_single(tune.shape) : tune.shape;
        const int TB = pass_shape_tb(shape), G = pass_shape_g(shape);
        std::vector<PassSys> tab;
        long long blocks = 0;
        for (int j = 0; j < m; ++j) {
            const int S = split_f64_levels(h_n[j], tune.lds_max);
            if (S <= Sin) continue;
            const int k = std::min(tune.k_fuse, S - Sin);
            const int H = 2 * ((1 << k) - 1), T = TB - 2 * H;
            if (T < 1) return PCRP_ERR_ARG;
            const int L = (h_n[j] + (1 << Sin) - 1) >> Sin;
            const int tiles = (L + T - 1) / T;
            const int groups = ((1 << Sin) + G - 1) / G;
            tab.push_back({h_off[j], P.woff[j], blocks, h_n[j], k, tiles, 0});
            blocks += (long long)groups * ((tiles + 7) / 8) * 8;   // XCD-striped tiles
        }
        if (tab.empty()) break;
        P.rounds.push_back(std::move(tab));
        P.round_blocks.push_back(blocks);
        P.round_shape.push_back(shape);
    }
    // workspace layout
    size_t o = 0;
    P.off_counter = o; o = align_up(o + kQueues * sizeof(unsigned int) + sizeof(P.qstart));
    P.off_sets = o;    o = align_up(o + 18 * sizeof(double*));
    P.off_jobs = o;    o = align_up(o + P.jobs.size() * sizeof(LdsJob));
    for (const auto& t : P.rounds) {
        P.round_off.push_back(o);
        o = align_up(o + t.size() * sizeof(PassSys));
    }
    for (long long blocks : P.round_blocks) {   // written on the device, after the upload
        P.round_map_off.push_back(o);
        o = align_up(o + (size_t)blocks * sizeof(int));
    }
    P.off_arrays = o;
    if (P.any_split) o += 2 * 6 * (size_t)P.wtotal * sizeof(double);
    P.bytes = o;
    return PCRP_OK;
}

}  // namespace

size_t uneven_f64_workspace(const int64_t* h_off, const int* h_n, int m)
{
    UnevenPlan P;
    if (plan_uneven(h_off, h_n, m, P) != PCRP_OK) return 0;
    return P.bytes;
}

namespace {

// y of the next solve into the device pointer table (a plan binds a..e and
// the workspace, not y)
__global__ void set_rhs_kernel(const double** sets, double* y)
{
    if (threadIdx.x == 0) sets[5] = y;
}

// The plan's tables (pointer sets, LDS jobs, pass-round tables) to the front
// of the workspace, from `host` (which must outlive the copy).
int upload_plan(const UnevenPlan& P, const double* a, const double* b, const double* c,
                const double* d, const double* e, double* y, void* ws, std::vector<char>& host,
                hipStream_t st)
{
    char* w = static_cast<char*>(ws);
    // array sets: 0 = caller's arrays, 1 and 2 = pass ping-pong buffers
    const double* sets[18] = {a, b, c, d, e, y};
    for (int v = 0; v < 12; ++v)
        sets[6 + v] = P.any_split
                          ? reinterpret_cast<double*>(w + P.off_arrays) + (size_t)v * P.wtotal
                          : nullptr;
    const size_t up = P.round_map_off.empty() ? P.off_arrays : P.round_map_off[0];
    host.assign(up, 0);
    std::copy(reinterpret_cast<const char*>(P.qstart),
              reinterpret_cast<const char*>(P.qstart) + sizeof(P.qstart),
              host.begin() + P.off_counter + kQueues * sizeof(unsigned int));
    std::copy(reinterpret_cast<const char*>(sets), reinterpret_cast<const char*>(sets) + sizeof sets,
              host.begin() + P.off_sets);
    std::copy(reinterpret_cast<const char*>(P.jobs.data()),
              reinterpret_cast<const char*>(P.jobs.data() + P.jobs.size()), host.begin() + P.off_jobs);
    for (size_t r = 0; r < P.rounds.size(); ++r)
        std::copy(reinterpret_cast<const char*>(P.rounds[r].data()),
                  reinterpret_cast<const char*>(P.rounds[r].data() + P.rounds[r].size()),
                  host.begin() + P.round_off[r]);
    if (hipMemcpyAsync(w, host.data(), host.size(), hipMemcpyHostToDevice, st) != hipSuccess)
        return PCRP_ERR_HIP;
    for (size_t r = 0; r < P.rounds.size(); ++r) {
        const int ntab = (int)P.rounds[r].size();
        hipLaunchKernelGGL(pass_block_map_kernel, dim3((unsigned)(ntab < 4096 ? ntab : 4096)),
                           dim3(256), 0, st, reinterpret_cast<const PassSys*>(w + P.round_off[r]),
                           ntab, P.round_blocks[r], reinterpret_cast<int*>(w + P.round_map_off[r]));
    }
    return hipGetLastError() == hipSuccess ? PCRP_OK : PCRP_ERR_HIP;
}

// The pass rounds and the persistent LDS launch over uploaded tables.
int launch_plan(const UnevenPlan& P, double* y, void* ws, hipStream_t st)
{
    char* w = static_cast<char*>(ws);
    int rc = PCRP_OK;
    const double* const* dsets = reinterpret_cast<const double* const*>(w + P.off_sets);
    // pass rounds: round r reads set (r == 0 ? caller : 1 + (r-1)%2), writes 1 + r%2
    for (size_t r = 0; r < P.rounds.size(); ++r) {
        const int src = r == 0 ? 0 : 1 + (int)((r - 1) & 1);
        const int dst = 1 + (int)(r & 1);
        rc = launch_pass_uneven(P.round_shape[r], dsets, src, dst,
                                reinterpret_cast<const PassSys*>(w + P.round_off[r]),
                                reinterpret_cast<const int*>(w + P.round_map_off[r]),
                                P.round_blocks[r], (int)r * split_f64_get().k_fuse,
                                split_f64_get().k_fuse, st);
        if (rc != 0) return PCRP_ERR_HIP;
    }
    // one persistent LDS launch over every job
    unsigned int* counters = reinterpret_cast<unsigned int*>(w + P.off_counter);
    const int* qstart = reinterpret_cast<const int*>(w + P.off_counter + kQueues * sizeof(unsigned int));
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, st, counters);
    if (hipGetLastError() != hipSuccess) return PCRP_ERR_HIP;
    const LdsVariantF64* v = lds_f64_variant(lds_tuned_queue_variant(PCRP_PREC_F64, split_f64_get().lds_max));
    rc = v->launch_queue(dsets, y, reinterpret_cast<const LdsJob*>(w + P.off_jobs),
                         (int)P.jobs.size(), qstart, counters, st);
    return rc == 0 ? PCRP_OK : PCRP_ERR_HIP;
}

}  // namespace

int uneven_f64_solve(const double* a, const double* b, const double* c, const double* d,
                     const double* e, double* y, const int64_t* h_off, const int* h_n, int m,
                     void* ws, size_t ws_bytes, hipStream_t st)
{
    UnevenPlan P;
    int rc = plan_uneven(h_off, h_n, m, P);
    if (rc != PCRP_OK) return rc;
    if (!ws || ws_bytes < P.bytes) return PCRP_ERR_WORKSPACE;
    // the host buffer lives to the end of this call (hipMemcpyAsync stages
    // pageable memory before it returns)
    std::vector<char> host;
    rc = upload_plan(P, a, b, c, d, e, y, ws, host, st);
    if (rc != PCRP_OK) return rc;
    return launch_plan(P, y, ws, st);
}

// Plan form: tables uploaded once (synchronous), then capturable executes
// that only set the right-hand side pointer and launch.
struct UnevenF64Plan {
    UnevenPlan P;
    void* ws;
};

UnevenF64Plan* uneven_f64_plan(const double* a, const double* b, const double* c, const double* d,
                               const double* e, const int64_t* h_off, const int* h_n, int m,
                               void* ws, size_t ws_bytes, hipStream_t st, int& rc)
{
    UnevenF64Plan* U = new UnevenF64Plan;
    U->ws = ws;
    rc = plan_uneven(h_off, h_n, m, U->P);
    if (rc == PCRP_OK && (!ws || ws_bytes < U->P.bytes)) rc = PCRP_ERR_WORKSPACE;
    std::vector<char> host;
    if (rc == PCRP_OK) rc = upload_plan(U->P, a, b, c, d, e, nullptr, ws, host, st);
    if (rc == PCRP_OK && hipStreamSynchronize(st) != hipSuccess) rc = PCRP_ERR_HIP;
    if (rc != PCRP_OK) {
        delete U;
        return nullptr;
    }
    return U;
}

void uneven_f64_plan_free(UnevenF64Plan* U) { delete U; }

int uneven_f64_plan_execute(const UnevenF64Plan* U, double* y, hipStream_t st)
{
    char* w = static_cast<char*>(U->ws);
    hipLaunchKernelGGL(set_rhs_kernel, dim3(1), dim3(64), 0, st,
                       reinterpret_cast<const double**>(w + U->P.off_sets), y);
    if (hipGetLastError() != hipSuccess) return PCRP_ERR_HIP;
    return launch_plan(U->P, y, U->ws, st);
}

}  // namespace pcrp

extern "C" {

size_t pcrp_uneven_workspace_size(const int64_t* h_offsets, const int* h_n, int m, int precision)
{
    if (!h_offsets || !h_n || m < 1) return 0;
    if (precision == PCRP_PREC_F64) return pcrp::uneven_f64_workspace(h_offsets, h_n, m);
    if (precision == PCRP_PREC_F32_REFINE) return pcrp::uneven_f32r_workspace(h_offsets, h_n, m);
    return 0;
}

int pcrp_solve_uneven_f64(const double* a, const double* b, const double* c, const double* d,
                          const double* e, double* y, const int64_t* h_offsets, const int* h_n,
                          int m, void* ws, size_t ws_bytes, void* stream)
{
    if (m < 1) return PCRP_ERR_INVALID_SIZE;
    if (!a || !b || !c || !d || !e || !y || !h_offsets || !h_n) return PCRP_ERR_NULL;
    return pcrp::uneven_f64_solve(a, b, c, d, e, y, h_offsets, h_n, m, ws, ws_bytes,
                                  static_cast<hipStream_t>(stream));
}

}  // extern "C"
