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
// sub-system q lands in sub-system q' = (l mod 2^k)*Q + q at index l >> k.
// With the original systems numbered j, this makes sub-system (residue R,
// system j) land at q = R*m + j after any sequence of passes. The workspace
// holds the sub-systems in the grouped layout (split_slot): groups of 4
// consecutive sub-systems row-interleaved, so a tile of 4 sub-systems reads
// one contiguous run per wave and the final LDS stage reads each sub-system at
// stride 4 (a layout interleaving all Q sub-systems puts each of a stage
// workgroup's rows in its own 128-byte line). The final stage writes x back
// to the caller's rows. a..e and y are never written by the passes.
//
// Per-row arithmetic is make_pair/reduce_row, as in the LDS kernel, with the
// guards evaluated in the current sub-system's coordinates. Those guards map
// one to one onto the whole-system guards (kernel.cu:680-691's stage-2 index
// map), so split and unsplit solves give the same bits.
#include "pcr_penta.h"
#include "pcrp_device.h"
#include "pcrp_internal.h"
#include "pcrp_spaced.h"

#include <mutex>

namespace pcrp {

// Natural order of a spaced pass's last step: tile row i at i with the bits
// [s, s+3) XORed with the low bits of i >> k (bits below k, so a permutation
// of the tile rows). The grouped store phase reads rows 2^k apart in
// consecutive lanes, which share LDS banks in plain natural order.
struct PassNat {
    int k, s, mask;
    __device__ int operator()(int i) const { return i ^ (((i >> k) & mask) << s); }
};

// Store modes of a spaced pass (SM): 1, the output slot in a group of 4
// residues is g mod 4 (G a multiple of 4); 2, it is j mod 4, j = l mod 2^k (one
// sub-system per tile, at least 2 steps); 0, per-row stores from registers.
template <int G, int SM>
__device__ __forceinline__ PassNat pass_nat(int k)
{
    if (SM == 1 && G == 4 && k >= 3) return PassNat{k, 0, 7};
    if (SM == 1 && G == 8 && k >= 2) return PassNat{k, 0, 3};
    if (SM == 2 && k >= 5) return PassNat{k, 2, 7};
    return PassNat{k, 0, 0};
}

// One tile of one HBM pass: rows l = tile*T - H + lr of current sub-system q
// (valid when 0 <= l < L), k fused steps in LDS, interior rows stored.
// Each thread owns RPT (row, sub-system) slots: slot tid + NTH*kk is row
// lr0 + (TB/RPT)*kk of sub-system g (the same g for all its slots), so its
// RPT rows are independent work for the same division latency.
// in_at(l) / out_at(l) give the element index of row l of this sub-system in
// the input arrays and of its destination in the output arrays.
// SP: the k steps in the spaced form (pcrp_spaced.h; FP64, RPT even, 2^k at
// most the TB/RPT threads of a sub-system), else each row builds both its
// pair equations (kernel.cu:66-140 as written).
// Stores: a row of sub-system q goes to residue q' = q + j*Qstep (j = l mod
// 2^k) at element l >> k, and the output keeps groups of 4 residues
// row-interleaved. With SM != 0 (SP only) the interior rows are stored from
// LDS in the output's own order: each wave writes runs of consecutive
// elements of one group (slot fastest, then the element), instead of pieces
// of 8 groups per instruction when lanes follow the rows (a block whose
// output groups are smaller, SM 2 with one step, stores row by row from LDS).
// out_g(g, l) and len_g(g) give the element index of row l of the tile's
// sub-system g and that sub-system's length (0 when g is past the last one).
template <typename T, int TB, int G, int RPT, bool SP, int SM, typename InAt, typename OutAt,
          typename OutG, typename LenG>
__device__ __forceinline__ void pass_tile(const T* const (&in)[6], T* const (&out)[6],
                                          bool qvalid, int l0, int L, int k, int H, int TI,
                                          InAt in_at, OutAt out_at, T (&sR)[6][TB * G],
                                          OutG out_g, LenG len_g)
{
    static_assert(SM == 0 || SP, "grouped stores need the spaced form");
    constexpr int NTH = TB * G / RPT;
    constexpr int RSTEP = TB / RPT;
    const int tid = threadIdx.x, lr0 = tid / G, g = tid - lr0 * G;
    Row<T> r[RPT];
    bool valid[RPT];
#pragma unroll
    for (int kk = 0; kk < RPT; ++kk) {
        const int l = l0 + RSTEP * kk;
        valid[kk] = qvalid && l >= 0 && l < L;
        r[kk] = Row<T>{T(0), T(0), T(1), T(0), T(0), T(0)};
        if (valid[kk]) {
            const long long gi = in_at(l);
            r[kk].a = in[0][gi]; r[kk].b = in[1][gi]; r[kk].c = in[2][gi];
            r[kk].d = in[3][gi]; r[kk].e = in[4][gi]; r[kk].y = in[5][gi];
        }
    }
    auto readR = [&](int i) {
        Row<T> w;
        w.a = sR[0][i]; w.b = sR[1][i]; w.c = sR[2][i];
        w.d = sR[3][i]; w.e = sR[4][i]; w.y = sR[5][i];
        return w;
    };
    auto writeR = [&](int i, const Row<T>& w) {
        sR[0][i] = w.a; sR[1][i] = w.b; sR[2][i] = w.c;
        sR[3][i] = w.d; sR[4][i] = w.e; sR[5][i] = w.y;
    };
    if constexpr (SP) {
        // rows to the spaced layout of stride 1 (invalid rows are identity rows)
#pragma unroll
        for (int kk = 0; kk < RPT; ++kk)
            writeR(spaced_addr<RSTEP, RPT>(lr0 + RSTEP * kk, 0) * G + g, r[kk]);
        __syncthreads();
        const int lofs = l0 - lr0;   // sub-system row of tile row 0
        const int Lq = qvalid ? L : 0;
        const PassNat nat = pass_nat<G, SM>(k);
        for (int s = 0; s < k; ++s) spaced_step<RSTEP, RPT, G>(s, s + 1 == k, lofs, Lq, sR, nat);
        if constexpr (SM != 0) {
            if (SM == 2 && k < 2) {
#pragma unroll 1
                for (int kk = 0; kk < RPT; ++kk) {
                    const int l = l0 + RSTEP * kk, lr = lr0 + RSTEP * kk;
                    if (valid[kk] && lr >= H && lr < H + TI) {
                        const int a = nat(lr) * G + g;
                        const long long go = out_at(l);
                        out[0][go] = sR[0][a]; out[1][go] = sR[1][a]; out[2][go] = sR[2][a];
                        out[3][go] = sR[3][a]; out[4][go] = sR[4][a]; out[5][go] = sR[5][a];
                    }
                }
                return;
            }
            // items (slot s4, sub-system group gh, element e, residue j), slot
            // fastest; E elements per residue cover the kept rows [kl0, kl0+TI)
            const int K2 = 1 << k;
            const int E = (TI + K2 - 1) >> k;
            const int kl0 = lofs + H;   // = tile*TI >= 0
            constexpr int GH = SM == 1 ? G / 4 : 1;
            const int items = 4 * GH * E * (SM == 1 ? K2 : K2 / 4);
#pragma unroll 1
            for (int u = tid; u < items; u += NTH) {
                const int s4 = u & 3;
                int rest = u >> 2;
                const int gh = rest % GH;
                rest /= GH;
                const int el = rest % E;
                const int jj = rest / E;
                const int gg = SM == 1 ? gh * 4 + s4 : 0;
                const int j = SM == 1 ? jj : jj * 4 + s4;
                const int l = (((kl0 - j + K2 - 1) >> k) + el) * K2 + j;
                if (l < kl0 + TI && l < len_g(gg)) {
                    const int a = nat(l - lofs) * G + gg;
                    const long long go = out_g(gg, l);
                    out[0][go] = sR[0][a]; out[1][go] = sR[1][a]; out[2][go] = sR[2][a];
                    out[3][go] = sR[3][a]; out[4][go] = sR[4][a]; out[5][go] = sR[5][a];
                }
            }
            return;
        } else {
            // (rows outside the sub-system are not stored: any value will do)
#pragma unroll
            for (int kk = 0; kk < RPT; ++kk) r[kk] = readR(tid + NTH * kk);
        }
    }
#pragma unroll
    for (int kk = 0; kk < RPT; ++kk)
        if (!SP) writeR(tid + NTH * kk, r[kk]);
    if (!SP) __syncthreads();
    for (int s = 0; !SP && s < k; ++s) {
        const int delta = 1 << s;
        Row<T> N[RPT];
#pragma unroll
        for (int kk = 0; kk < RPT; ++kk) {
            const int l = l0 + RSTEP * kk, lr = lr0 + RSTEP * kk, self = tid + NTH * kk;
            N[kk] = r[kk];
            if (valid[kk]) {
                const bool hp1 = l - delta >= 0, hq1 = l + delta <= L - 1;
                const bool hp2 = l - 2 * delta >= 0, hq2 = l + 2 * delta <= L - 1;
                const bool hp3 = l - 3 * delta >= 0, hq3 = l + 3 * delta <= L - 1;
                // neighbours outside the tile only feed rows that are not kept
                auto nb = [&](bool exists, int off) {
                    const int lj = lr + off;
                    return (exists && lj >= 0 && lj < TB) ? lj * G + g : self;
                };
                const Row<T> P1 = readR(nb(hp1, -delta)), P2 = readR(nb(hp2, -2 * delta));
                const Row<T> Q1 = readR(nb(hq1, delta)), Q2 = readR(nb(hq2, 2 * delta));
                const Pair<T> EL = make_pair(P2, P1, r[kk], hp2, hp3, true, hq1);
                const Pair<T> ER = make_pair(r[kk], Q1, Q2, true, hp1, hq2, hq3);
                N[kk] = reduce_row(r[kk], EL, ER, hp1, hq1, hp2, hq2,
                                   [&](T& alL, T& arL, T& alR, T& arR) {
                                       alL = EL.al; arL = EL.ar; alR = ER.al; arR = ER.ar;
                                   });
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < RPT; ++kk) {
            r[kk] = N[kk];
            writeR(tid + NTH * kk, r[kk]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int kk = 0; kk < RPT; ++kk) {
        const int l = l0 + RSTEP * kk, lr = lr0 + RSTEP * kk;
        if (valid[kk] && lr >= H && lr < H + TI) {
            const long long go = out_at(l);
            out[0][go] = r[kk].a; out[1][go] = r[kk].b; out[2][go] = r[kk].c;
            out[3][go] = r[kk].d; out[4][go] = r[kk].e; out[5][go] = r[kk].y;
        }
    }
}

// Block -> (group, tile). Groups of G neighbouring sub-systems share cache
// lines (rows are interleaved across sub-systems), so the groups of one tile
// are made consecutive and kept on one XCD: blocks are dealt round-robin to
// the 8 XCDs, so block b runs on XCD b % 8 and the b / 8-th block there takes
// (tile, group) with the group index fastest and tiles striped over XCDs.
// Blocks past the last tile return at once (whole workgroup).
__device__ __forceinline__ bool tile_of_block(int b, int groups, int tiles, int& grp, int& tile)
{
    constexpr int NX = 8;
    const int xcd = b % NX, s = b / NX;
    const int tl = s / groups;
    grp = s - tl * groups;
    tile = tl * NX + xcd;
    return tile < tiles;
}

__host__ __forceinline__ long long blocks_for(int groups, int tiles)
{
    constexpr int NX = 8;
    return (long long)groups * ((tiles + NX - 1) / NX) * NX;
}

// Uniform batch: m0 systems of n0 rows, current sub-systems q = R*m0 + j.
template <typename T, int TB, int G, int RPT, bool SP, int SM>
__global__ __launch_bounds__(TB* G / RPT) void hbm_pass_kernel(
    const T* __restrict__ ia, const T* __restrict__ ib, const T* __restrict__ ic,
    const T* __restrict__ id, const T* __restrict__ ie, const T* __restrict__ iy,
    T* __restrict__ oa, T* __restrict__ ob, T* __restrict__ oc,
    T* __restrict__ od, T* __restrict__ oe, T* __restrict__ oy,
    int n0, int m0, int Sin, int k, long long in_rs, long long in_ss, int Qin, int tiles)
{
    __shared__ T sR[6][TB * G];
    const int H = 2 * ((1 << k) - 1);
    const int TI = TB - 2 * H;   // interior rows kept per tile
    int grp, tile;
    if (!tile_of_block(blockIdx.x, (Qin + G - 1) / G, tiles, grp, tile)) return;
    const int lr0 = threadIdx.x / G, g = threadIdx.x - lr0 * G;
    const int q = grp * G + g;
    const int l0 = tile * TI - H + lr0;
    // current sub-system q = R*m0 + j holds rows R + 2^Sin * l of system j
    const int R = q / m0;
    const int L = (n0 - R + (1 << Sin) - 1) >> Sin;
    const long long Qout = (long long)Qin << k;
    const long long Lin = ((long long)n0 + (1LL << Sin) - 1) >> Sin;
    const long long Lout = ((long long)n0 + (1LL << (Sin + k)) - 1) >> (Sin + k);
    const T* const in[6] = {ia, ib, ic, id, ie, iy};
    T* const out[6] = {oa, ob, oc, od, oe, oy};
    pass_tile<T, TB, G, RPT, SP, SM>(
        in, out, q < Qin, l0, L, k, H, TI,
        [&](int ll) {
            return Sin == 0 ? (long long)q * in_ss + (long long)ll * in_rs
                            : split_slot(q, ll, Qin, Lin);
        },
        [&](int ll) {
            return split_slot((long long)(ll & ((1 << k) - 1)) * Qin + q, ll >> k, Qout, Lout);
        },
        sR,
        [&](int gg, int ll) {
            const int qq = grp * G + gg;
            return split_slot((long long)(ll & ((1 << k) - 1)) * Qin + qq, ll >> k, Qout, Lout);
        },
        [&](int gg) {
            const int qq = grp * G + gg;
            return qq < Qin ? (n0 - qq / m0 + (1 << Sin) - 1) >> Sin : 0;
        });
}

// Uneven batch: systems stored contiguously at their own offsets. The first
// round reads the caller's rows (off + l); every round writes its residue
// sub-systems to the pass buffers in the grouped residue layout at the
// system's own region (woff + res_slot), and later rounds read them there.
// Row l of input residue q (Sin levels) goes to residue q + (l mod 2^k) 2^Sin,
// element l >> k, of the Sin + k levels after the pass.
// tab lists the systems that still need passes in this round; bmap gives
// every block its entry of tab (a binary search over the entries, 13
// dependent loads at 6000 systems, delayed every block's first loads).
template <int TB, int G, int RPT, bool SP, int SM>
__global__ __launch_bounds__(TB* G / RPT) void hbm_pass_uneven_f64_kernel(
    const double* const* __restrict__ sets, int src, int dst, const PassSys* __restrict__ tab,
    const int* __restrict__ bmap, int Sin)
{
    __shared__ double sR[6][TB * G];
    const long long b = blockIdx.x;
    const PassSys P = tab[bmap[b]];
    const int local = (int)(b - P.block0);
    int grp, tile;
    if (!tile_of_block(local, ((1 << Sin) + G - 1) / G, P.tiles, grp, tile)) return;
    const int k = P.k;
    const int H = 2 * ((1 << k) - 1);
    const int T = TB - 2 * H;
    const int lr0 = threadIdx.x / G, g = threadIdx.x - lr0 * G;
    const int Qin = 1 << Sin;
    const int q = grp * G + g;
    const int L = q < Qin ? (P.n - q + Qin - 1) >> Sin : 0;
    const int l0 = tile * T - H + lr0;
    const double* const in[6] = {sets[6 * src + 0], sets[6 * src + 1], sets[6 * src + 2],
                                 sets[6 * src + 3], sets[6 * src + 4], sets[6 * src + 5]};
    double* const out[6] = {const_cast<double*>(sets[6 * dst + 0]), const_cast<double*>(sets[6 * dst + 1]),
                            const_cast<double*>(sets[6 * dst + 2]), const_cast<double*>(sets[6 * dst + 3]),
                            const_cast<double*>(sets[6 * dst + 4]), const_cast<double*>(sets[6 * dst + 5])};
    const int Sout = Sin + k;
    auto in_at = [&](int ll) {
        return Sin == 0 ? P.off + ll : P.woff + res_slot(q, ll, Sin, P.n);
    };
    auto out_at = [&](int ll) {
        return P.woff + res_slot(q + ((ll & ((1 << k) - 1)) << Sin), ll >> k, Sout, P.n);
    };
    pass_tile<double, TB, G, RPT, SP, SM>(
        in, out, q < Qin, l0, L, k, H, T, in_at, out_at, sR,
        [&](int gg, int ll) {
            const int qq = grp * G + gg;
            return P.woff + res_slot(qq + ((ll & ((1 << k) - 1)) << Sin), ll >> k, Sout, P.n);
        },
        [&](int gg) {
            const int qq = grp * G + gg;
            return qq < Qin ? (P.n - qq + Qin - 1) >> Sin : 0;
        });
}

// SP shapes run the spaced form when 2^k fits the TB/RPT threads of a
// sub-system, the redundant form of the same shape otherwise.
template <typename T, int TB, int G, int RPT, bool SP>
static int launch_pass(const T* const in[6], T* const out[6], int n0, int m0, int Sin,
                       int k, long long in_rs, long long in_ss, int Qin, hipStream_t st)
{
    const int H = 2 * ((1 << k) - 1);
    const int TI = TB - 2 * H;
    if (TI < 1) return -2;
    const int L = (n0 + (1 << Sin) - 1) >> Sin;   // longest current sub-system
    const int tiles = (L + TI - 1) / TI;
    const int groups = (Qin + G - 1) / G;
    const long long blocks = blocks_for(groups, tiles);
    if (blocks > 0x7fffffffLL) return -2;
    const auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(TB * G / RPT), 0, st, in[0], in[1],
                           in[2], in[3], in[4], in[5], out[0], out[1], out[2], out[3], out[4],
                           out[5], n0, m0, Sin, k, in_rs, in_ss, Qin, tiles);
    };
    // grouped stores when the output slot is the tile's sub-system mod 4
    constexpr bool SP_OK = SP && G % 4 == 0;
    if (SP && (1 << k) <= TB / RPT) {
        if (SP_OK && Qin % 4 == 0) go(hbm_pass_kernel<T, TB, G, RPT, SP, SP_OK ? 1 : 0>);
        else go(hbm_pass_kernel<T, TB, G, RPT, SP, 0>);
    } else {
        go(hbm_pass_kernel<T, TB, G, RPT, false, 0>);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// kmax: the most steps any system of this round fuses (spaced form only
// when 2^kmax fits the TB/RPT threads of a sub-system).
template <int TB, int G, int RPT, bool SP>
static int launch_pass_uneven_t(const double* const* sets, int src, int dst, const PassSys* tab,
                                const int* bmap, long long blocks, int Sin, int kmax, hipStream_t st)
{
    if (blocks <= 0) return 0;
    if (blocks > 0x7fffffffLL) return -2;
    const auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3((unsigned)blocks), dim3(TB * G / RPT), 0, st, sets, src,
                           dst, tab, bmap, Sin);
    };
    // grouped stores: the output slot is the tile's sub-system mod 4 (groups of
    // 4 or 8 sub-systems, at least 4 residues in) or, in a round of whole
    // systems (one sub-system per tile), the residue j mod 4
    constexpr int SM_G = (SP && G % 4 == 0) ? 1 : (SP && G == 1) ? 2 : 0;
    if (SP && (1 << kmax) <= TB / RPT) {
        if (SM_G == 1 && Sin >= 2) go(hbm_pass_uneven_f64_kernel<TB, G, RPT, SP, SM_G>);
        else if (SM_G == 2 && Sin == 0) go(hbm_pass_uneven_f64_kernel<TB, G, RPT, SP, SM_G>);
        else go(hbm_pass_uneven_f64_kernel<TB, G, RPT, SP, 0>);
    } else {
        go(hbm_pass_uneven_f64_kernel<TB, G, RPT, false, 0>);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

using PassFn = int (*)(const double* const*, double* const*, int, int, int, int, long long,
                       long long, int, hipStream_t);

using PassUnevenFn = int (*)(const double* const*, int, int, const PassSys*, const int*, long long,
                             int, int, hipStream_t);

struct PassShape {
    int tb, g, rpt;
    bool spaced;
    PassFn fn;
    PassUnevenFn uneven;
};

#define PCRP_PASS_SHAPE(TB, G, RPT) \
    {TB, G, RPT, false, &launch_pass<double, TB, G, RPT, false>, &launch_pass_uneven_t<TB, G, RPT, false>}
#define PCRP_PASS_SHAPE_SP(TB, G, RPT) \
    {TB, G, RPT, true, &launch_pass<double, TB, G, RPT, true>, &launch_pass_uneven_t<TB, G, RPT, true>}

static const PassShape kPassShapes[] = {
    // one (row, sub-system) slot per thread
    PCRP_PASS_SHAPE(256, 4, 1),
    PCRP_PASS_SHAPE(512, 2, 1),
    PCRP_PASS_SHAPE(128, 8, 1),
    PCRP_PASS_SHAPE(1024, 1, 1),
    // four slots per thread (256 threads)
    PCRP_PASS_SHAPE(256, 4, 4),
    PCRP_PASS_SHAPE(512, 2, 4),
    PCRP_PASS_SHAPE(128, 8, 4),
    PCRP_PASS_SHAPE(1024, 1, 4),
    // four slots per thread, spaced steps
    PCRP_PASS_SHAPE_SP(256, 4, 4),
    PCRP_PASS_SHAPE_SP(512, 2, 4),
    PCRP_PASS_SHAPE_SP(128, 8, 4),
    PCRP_PASS_SHAPE_SP(1024, 1, 4),
};
#undef PCRP_PASS_SHAPE
#undef PCRP_PASS_SHAPE_SP

int pass_shape_tb(int shape) { return kPassShapes[shape].tb; }
int pass_shape_g(int shape) { return kPassShapes[shape].g; }
int pass_shape_rpt(int shape) { return kPassShapes[shape].rpt; }

// The shape of the same rows-per-thread with whole-row tiles (G = 1), used by
// the first pass round of uneven batches (one sub-system per system there).
int pass_shape_single(int shape)
{
    const int nshapes = (int)(sizeof(kPassShapes) / sizeof(kPassShapes[0]));
    for (int i = 0; i < nshapes; ++i)
        if (kPassShapes[i].g == 1 && kPassShapes[i].rpt == kPassShapes[shape].rpt &&
            kPassShapes[i].spaced == kPassShapes[shape].spaced)
            return i;
    return shape;
}

int launch_pass_uneven(int shape, const double* const* sets, int src, int dst, const PassSys* tab,
                       const int* bmap, long long blocks, int Sin, int kmax, hipStream_t st)
{
    return kPassShapes[shape].uneven(sets, src, dst, tab, bmap, blocks, Sin, kmax, st);
}

// Tunables of the split path: LDS/HBM threshold (sub-systems at most this
// long go to the LDS kernel), steps fused per HBM pass, pass tile shape.
// (a mutex: solves read them while another thread may tune). Default: the
// autotuner's choice on the uneven batch of 10000 systems (MI355X): five
// steps per pass, spaced 256 x 4 tiles of 4 sub-systems (whole-row spaced
// 1024 x 1 tiles in the first uneven round).
static SplitTune g_split_f64 = {1024, 5, 8};
static std::mutex g_split_mu;

SplitTune split_f64_get()
{
    std::lock_guard<std::mutex> lk(g_split_mu);
    return g_split_f64;
}

int split_f64_set(const SplitTune& t)
{
    const int nshapes = (int)(sizeof(kPassShapes) / sizeof(kPassShapes[0]));
    if (t.lds_max < 2 || t.lds_max > lds_f64_max_n() || (t.lds_max & (t.lds_max - 1)) != 0)
        return -1;
    if (t.k_fuse < 1 || t.k_fuse > 6 || t.shape < 0 || t.shape >= nshapes) return -1;
    const int H = 2 * ((1 << t.k_fuse) - 1);
    if (kPassShapes[t.shape].tb - 2 * H < 1) return -1;
    std::lock_guard<std::mutex> lk(g_split_mu);
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
    const int S = split_f64_levels(n, split_f64_get().lds_max);
    if (S == 0) return 0;
    // two ping-pong sets of six arrays of m * 2^S * ceil(n / 2^S) doubles
    const size_t per = (size_t)m * ((size_t)((n + (1 << S) - 1) >> S) << S);
    return 2 * 6 * per * sizeof(double);
}

int split_f64_solve(const SolveArgsF64& A, int layout, void* ws, size_t ws_bytes, hipStream_t st)
{
    const int n0 = A.n, m0 = A.m;
    const SplitTune tune = split_f64_get();
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
    const LdsVariantF64* v = lds_f64_variant(lds_tuned_variant(PCRP_PREC_F64, F.n, F.m));
    return v->launch(F, st);
}

}  // namespace pcrp
