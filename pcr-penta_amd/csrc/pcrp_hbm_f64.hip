// HBM tile-splitting path for systems larger than the LDS kernel (FP64).
//
// Counterpart of shared_cached_shared_tile_splitting (kernel.cu:158-320) and
// the host split logic of main.cu:57-86. The reference runs s = log2(n/1024)
// full-system reduction steps, one HBM pass per step, in place over a..e,y,
// with one block per system walking its tiles sequentially; then stage 2
// solves the 2^s stride-2^s sub-systems in shared memory.
//
// Here each HBM pass fuses k steps: a workgroup loads a tile of TB rows for G
// neighbouring systems (consecutive in the interleaved layout, so every row
// load is a contiguous G*8-byte run), runs k steps in LDS, and keeps the
// T = TB - 2H interior rows whose dependency cone (H = 2(2^k - 1) rows per
// side) lies in the tile. Tiles run in parallel (no sequential tile walk and
// no shared-memory cache of the previous tile, so no write-after-read race).
// The output goes to a workspace, already split by residue: row l of current
// sub-system q lands in sub-system q' = (l mod 2^k)*Q + q at index l >> k,
// interleaved (element l' of q' at l'*Q' + q'). With the original systems
// numbered j, this makes sub-system (residue R, system j) land at q = R*m + j
// after any sequence of passes, so the next pass and the final LDS stage read
// coalesced, and the final stage writes x back through the same index
// pattern in the caller's interleaved layout. a..e and y are never written
// by the passes.
//
// Per-row arithmetic is make_pair/reduce_row, as in the LDS kernel, with the
// guards evaluated in the current sub-system's coordinates. Those guards map
// one to one onto the whole-system guards (kernel.cu:680-691's stage-2 index
// map), so split and unsplit solves give the same bits.
#include "pcr_penta.h"
#include "pcrp_device.h"
#include "pcrp_internal.h"

namespace pcrp {

template <int TB, int G>
__global__ __launch_bounds__(TB* G) void hbm_pass_f64_kernel(
    const double* __restrict__ ia, const double* __restrict__ ib, const double* __restrict__ ic,
    const double* __restrict__ id, const double* __restrict__ ie, const double* __restrict__ iy,
    double* __restrict__ oa, double* __restrict__ ob, double* __restrict__ oc,
    double* __restrict__ od, double* __restrict__ oe, double* __restrict__ oy,
    int n0, int m0, int Sin, int k, long long in_rs, long long in_ss, int Qin, int tiles)
{
    constexpr int NTH = TB * G;
    __shared__ double sR[6][NTH];

    const int H = 2 * ((1 << k) - 1);
    const int T = TB - 2 * H;
    const int grp = blockIdx.x / tiles, tile = blockIdx.x - grp * tiles;
    const int tid = threadIdx.x, lr = tid / G, g = tid - lr * G;
    const int q = grp * G + g;
    const int l = tile * T - H + lr;
    // current sub-system q = R*m0 + j holds rows R + 2^Sin * l of system j
    const int R = q / m0;
    const int L = (n0 - R + (1 << Sin) - 1) >> Sin;
    const bool valid = q < Qin && l >= 0 && l < L;

    Row<double> r{0.0, 0.0, 1.0, 0.0, 0.0, 0.0};
    if (valid) {
        const long long gi = (long long)q * in_ss + (long long)l * in_rs;
        r.a = ia[gi]; r.b = ib[gi]; r.c = ic[gi]; r.d = id[gi]; r.e = ie[gi]; r.y = iy[gi];
    }
    auto readR = [&](int i) {
        Row<double> w;
        w.a = sR[0][i]; w.b = sR[1][i]; w.c = sR[2][i];
        w.d = sR[3][i]; w.e = sR[4][i]; w.y = sR[5][i];
        return w;
    };
    auto writeR = [&](const Row<double>& w) {
        sR[0][tid] = w.a; sR[1][tid] = w.b; sR[2][tid] = w.c;
        sR[3][tid] = w.d; sR[4][tid] = w.e; sR[5][tid] = w.y;
    };
    writeR(r);
    __syncthreads();
    for (int s = 0; s < k; ++s) {
        const int delta = 1 << s;
        if (valid) {
            const bool hp1 = l - delta >= 0, hq1 = l + delta <= L - 1;
            const bool hp2 = l - 2 * delta >= 0, hq2 = l + 2 * delta <= L - 1;
            const bool hp3 = l - 3 * delta >= 0, hq3 = l + 3 * delta <= L - 1;
            // neighbours outside the tile only feed rows that are not kept
            auto nb = [&](bool exists, int off) {
                const int lj = lr + off;
                return (exists && lj >= 0 && lj < TB) ? lj * G + g : tid;
            };
            const Row<double> P1 = readR(nb(hp1, -delta)), P2 = readR(nb(hp2, -2 * delta));
            const Row<double> Q1 = readR(nb(hq1, delta)), Q2 = readR(nb(hq2, 2 * delta));
            const Pair<double> EL = make_pair(P2, P1, r, hp2, hp3, true, hq1);
            const Pair<double> ER = make_pair(r, Q1, Q2, true, hp1, hq2, hq3);
            r = reduce_row(r, EL, ER, hp1, hq1, hp2, hq2,
                           [&](double& alL, double& arL, double& alR, double& arR) {
                               alL = EL.al; arL = EL.ar; alR = ER.al; arR = ER.ar;
                           });
        }
        __syncthreads();
        writeR(r);
        __syncthreads();
    }
    if (valid && lr >= H && lr < H + T) {
        const int rr = l & ((1 << k) - 1), lp = l >> k;
        const long long Qout = (long long)Qin << k;
        const long long go = (long long)lp * Qout + (long long)rr * Qin + q;
        oa[go] = r.a; ob[go] = r.b; oc[go] = r.c; od[go] = r.d; oe[go] = r.e; oy[go] = r.y;
    }
}

template <int TB, int G>
static int launch_pass(const double* const in[6], double* const out[6], int n0, int m0, int Sin,
                       int k, long long in_rs, long long in_ss, int Qin, hipStream_t st)
{
    const int H = 2 * ((1 << k) - 1);
    const int T = TB - 2 * H;
    if (T < 1) return -2;
    const int L = (n0 + (1 << Sin) - 1) >> Sin;   // longest current sub-system
    const int tiles = (L + T - 1) / T;
    const long long groups = (Qin + G - 1) / G;
    const long long blocks = groups * tiles;
    if (blocks > 0x7fffffffLL) return -2;
    hipLaunchKernelGGL((hbm_pass_f64_kernel<TB, G>), dim3((unsigned)blocks), dim3(TB * G), 0, st,
                       in[0], in[1], in[2], in[3], in[4], in[5], out[0], out[1], out[2], out[3],
                       out[4], out[5], n0, m0, Sin, k, in_rs, in_ss, Qin, tiles);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

using PassFn = int (*)(const double* const*, double* const*, int, int, int, int, long long,
                       long long, int, hipStream_t);

struct PassShape {
    int tb, g;
    PassFn fn;
};

static const PassShape kPassShapes[] = {
    {256, 4, &launch_pass<256, 4>},
    {512, 2, &launch_pass<512, 2>},
    {128, 8, &launch_pass<128, 8>},
    {1024, 1, &launch_pass<1024, 1>},
};

// Tunables of the split path: LDS/HBM threshold (sub-systems at most this
// long go to the LDS kernel), steps fused per HBM pass, pass tile shape.
static SplitTune g_split_f64 = {1024, 4, 0};

SplitTune split_f64_get() { return g_split_f64; }

int split_f64_set(const SplitTune& t)
{
    const int nshapes = (int)(sizeof(kPassShapes) / sizeof(kPassShapes[0]));
    if (t.lds_max < 2 || t.lds_max > lds_f64_max_n() || (t.lds_max & (t.lds_max - 1)) != 0)
        return -1;
    if (t.k_fuse < 1 || t.k_fuse > 6 || t.shape < 0 || t.shape >= nshapes) return -1;
    const int H = 2 * ((1 << t.k_fuse) - 1);
    if (kPassShapes[t.shape].tb - 2 * H < 1) return -1;
    g_split_f64 = t;
    return 0;
}

int split_f64_shape_count() { return (int)(sizeof(kPassShapes) / sizeof(kPassShapes[0])); }

int split_f64_levels(int n, int lds_max)
{
    int S = 0;
    while (((n + (1 << S) - 1) >> S) > lds_max) ++S;
    return S;
}

size_t split_f64_workspace(int n, int m)
{
    const int S = split_f64_levels(n, g_split_f64.lds_max);
    if (S == 0) return 0;
    // two ping-pong sets of six arrays of m * 2^S * ceil(n / 2^S) doubles
    const size_t per = (size_t)m * ((size_t)((n + (1 << S) - 1) >> S) << S);
    return 2 * 6 * per * sizeof(double);
}

int split_f64_solve(const SolveArgsF64& A, int layout, void* ws, size_t ws_bytes, hipStream_t st)
{
    const int n0 = A.n, m0 = A.m;
    const SplitTune tune = g_split_f64;
    const int S = split_f64_levels(n0, tune.lds_max);
    if (S == 0) return -2;
    if (ws_bytes < split_f64_workspace(n0, m0) || ws == nullptr) return -3;
    const size_t per = (size_t)m0 * ((size_t)((n0 + (1 << S) - 1) >> S) << S);
    double* base = static_cast<double*>(ws);
    double* set0[6];
    double* set1[6];
    for (int v = 0; v < 6; ++v) {
        set0[v] = base + (size_t)v * per;
        set1[v] = base + (size_t)(6 + v) * per;
    }
    const PassShape& shape = kPassShapes[tune.shape];
    const double* in[6] = {A.a, A.b, A.c, A.d, A.e, A.y};
    long long in_rs = A.rs, in_ss = A.ss;
    int Sin = 0, Qin = m0;
    double* const* dst = set0;
    while (Sin < S) {
        const int k = (S - Sin < tune.k_fuse) ? S - Sin : tune.k_fuse;
        const int rc = shape.fn(in, dst, n0, m0, Sin, k, in_rs, in_ss, Qin, st);
        if (rc != 0) return rc;
        Sin += k;
        Qin <<= k;
        for (int v = 0; v < 6; ++v) in[v] = dst[v];
        in_rs = Qin;
        in_ss = 1;
        dst = (dst == set0) ? set1 : set0;
    }
    // final stage: every residue sub-system in the LDS kernel, x to the caller's y
    SolveArgsF64 F{in[0], in[1], in[2], in[3], in[4], const_cast<double*>(in[5]),
                   (n0 + (1 << S) - 1) >> S, Qin, in_rs, in_ss, A.steps - S,
                   SplitInfo{S, n0, m0, layout, A.y}};
    const LdsVariantF64* v = lds_f64_variant(lds_tuned_variant(PCRP_PREC_F64, F.n));
    return v->launch(F, st);
}

}  // namespace pcrp
