// LDS-resident FP64 PCR-Penta: one workgroup solves one whole system.
//
// Counterpart of pcr_penta (kernel.cu:32-154) and of the second split stage
// pcr_penta_splitting_2nd_stage (kernel.cu:665-791), which is the same solve
// on a strided sub-system. Differences by design:
//   * NT threads each own RPT rows (row i = t + NT*k), so n is not tied to
//     the workgroup size and the rows of one thread stay in VGPRs;
//   * Reduction-I/II are shared per centre j (pair equations, pcrp_device.h):
//     4 FP64 divisions per row and step instead of 6;
//   * rows live in VGPRs; LDS holds the rows (6 doubles) and pair equations
//     (8 doubles) that neighbours read: 2 barriers per step, as kernel.cu:135,142;
//   * any element addressing (row stride, system stride), so both the
//     reference's contiguous layout and the interleaved layout work;
//   * a..e are read-only; y is overwritten by x.
// A sub-system view (used by the split path) is expressed through
// (base offset, row stride, n_sub, steps, final delta) — the reference's
// stage-2 index map g = sys*n + id*G + r (kernel.cu:680) is the special case
// row stride G, offset r.
#include "pcrp_device.h"
#include "pcrp_internal.h"

namespace pcrp {

template <int NT, int RPT>
__global__ __launch_bounds__(NT) void lds_f64_kernel(
    const double* __restrict__ ga, const double* __restrict__ gb, const double* __restrict__ gc,
    const double* __restrict__ gd, const double* __restrict__ ge, double* __restrict__ gy,
    int n, int m, long long rs, long long ss, int steps)
{
    constexpr int NMAX = NT * RPT;
    __shared__ double sR[6][NMAX];   // current rows a,b,c,d,e,y
    __shared__ double sE[8][NMAX];   // pair equations A,B,C,D,E,K,al,ar

    const int t = threadIdx.x;
    const int sys = xcd_remap(blockIdx.x, m);
    const long long base = (long long)sys * ss;

    Row<double> r[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        if (i < n) {
            const long long g = base + (long long)i * rs;
            r[k].a = ga[g]; r[k].b = gb[g]; r[k].c = gc[g];
            r[k].d = gd[g]; r[k].e = ge[g]; r[k].y = gy[g];
            sR[0][i] = r[k].a; sR[1][i] = r[k].b; sR[2][i] = r[k].c;
            sR[3][i] = r[k].d; sR[4][i] = r[k].e; sR[5][i] = r[k].y;
        }
    }
    __syncthreads();

    auto readR = [&](int i) {
        Row<double> q;
        q.a = sR[0][i]; q.b = sR[1][i]; q.c = sR[2][i];
        q.d = sR[3][i]; q.e = sR[4][i]; q.y = sR[5][i];
        return q;
    };
    auto readE = [&](int i) {
        Pair<double> q;
        q.A = sE[0][i]; q.B = sE[1][i]; q.C = sE[2][i];
        q.D = sE[3][i]; q.E = sE[4][i]; q.K = sE[5][i];
        return q;
    };

    int delta = 1;
    for (int s = 0; s < steps; ++s, delta <<= 1) {
        // Reduction-I/II: pair equation centred on each own row j.
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
        // Reduction-III: new row from E(i-d), row i, E(i+d).
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int i = t + NT * k;
            if (i < n) {
                const bool hp1 = i - delta >= 0, hq1 = i + delta <= n - 1;
                const bool hp2 = i - 2 * delta >= 0, hq2 = i + 2 * delta <= n - 1;
                const int il = hp1 ? i - delta : i, ir = hq1 ? i + delta : i;
                const Pair<double> EL = readE(il), ER = readE(ir);
                r[k] = reduce_row(r[k], EL, ER, hp1, hq1, hp2, hq2,
                                  [&](double& alL, double& arL, double& alR, double& arR) {
                                      alL = sE[6][il]; arL = sE[7][il];
                                      alR = sE[6][ir]; arR = sE[7][ir];
                                  });
                sR[0][i] = r[k].a; sR[1][i] = r[k].b; sR[2][i] = r[k].c;
                sR[3][i] = r[k].d; sR[4][i] = r[k].e; sR[5][i] = r[k].y;
            }
        }
        __syncthreads();
    }

    // Final 2x2 / 1x1 solve (kernel.cu:145-153).
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int i = t + NT * k;
        if (i < n) {
            const long long g = base + (long long)i * rs;
            if (i < n - delta) {
                const int ii = i + delta;
                const double cii = sR[2][ii], bii = sR[1][ii], yii = sR[5][ii];
                const double deno = __builtin_fma(r[k].c, cii, -(bii * r[k].d));
                gy[g] = __builtin_fma(cii, r[k].y, -(r[k].d * yii)) / deno;
                gy[g + (long long)delta * rs] = __builtin_fma(r[k].c, yii, -(bii * r[k].y)) / deno;
            } else if (i < delta) {
                gy[g] = r[k].y / r[k].c;
            }
        }
    }
}

template <int NT, int RPT>
static int launch_lds_f64(const SolveArgsF64& A, hipStream_t st)
{
    hipLaunchKernelGGL((lds_f64_kernel<NT, RPT>), dim3(A.m), dim3(NT), 0, st,
                       A.a, A.b, A.c, A.d, A.e, A.y, A.n, A.m, A.rs, A.ss, A.steps);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int lds_f64_max_n() { return 1024; }

// Workgroup shape per size class (NT threads x RPT rows per thread).
int lds_f64_solve(const SolveArgsF64& A, hipStream_t st)
{
    const int n = A.n;
    if (n <= 64) return launch_lds_f64<64, 1>(A, st);
    if (n <= 128) return launch_lds_f64<128, 1>(A, st);
    if (n <= 256) return launch_lds_f64<256, 1>(A, st);
    if (n <= 512) return launch_lds_f64<512, 1>(A, st);
    if (n <= 1024) return launch_lds_f64<1024, 1>(A, st);
    return -2;
}

}  // namespace pcrp
