// C-ABI of the solver (declared in include/pcr_penta.h).
//
// Replaces the reference's host layer: the per-size dispatch of main.cu:46-87
// (n <= 1024 -> pcr_penta, else split stage 1 + stage 2) and the checkCuda
// error convention (performanalysis.h:12-20), which exits the process; here
// every failure is a returned status and nothing synchronises the device.
#include "pcr_penta.h"

#include "pcrp_device.h"
#include "pcrp_internal.h"

using namespace pcrp;

namespace {

bool layout_strides(int n, int m, int layout, long long* rs, long long* ss)
{
    if (layout == PCRP_LAYOUT_CONTIGUOUS) { *rs = 1; *ss = n; return true; }
    if (layout == PCRP_LAYOUT_INTERLEAVED) { *rs = m; *ss = 1; return true; }
    return false;
}

int check_shape(int n, int m)
{
    if (n < 1 || m < 1) return PCRP_ERR_INVALID_SIZE;
    if ((long long)n * (long long)m > 0x7fffffffLL) return PCRP_ERR_INVALID_SIZE;
    return PCRP_OK;
}

__global__ void generate_kernel(int kind, uint64_t key, int n, int m, long long rs, long long ss,
                                double* a, double* b, double* c, double* d, double* e, double* y)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * m) return;
    // consecutive threads walk the fastest-varying dimension of the layout
    int i, j;
    if (ss == 1) { i = (int)(idx / m); j = (int)(idx % m); }
    else { j = (int)(idx / n); i = (int)(idx % n); }
    const long long g = (long long)j * ss + (long long)i * rs;
    const double off = kind == 1 ? 0.5 : 1.0;
    const double clo = kind == 1 ? 2.2 : 4.5, chi = kind == 1 ? 2.6 : 5.5;
    a[g] = i >= 2 ? gen_uniform(key, 0, j, i, -off, off) : 0.0;
    b[g] = i >= 1 ? gen_uniform(key, 1, j, i, -off, off) : 0.0;
    c[g] = gen_uniform(key, 2, j, i, clo, chi);
    d[g] = i <= n - 2 ? gen_uniform(key, 3, j, i, -off, off) : 0.0;
    e[g] = i <= n - 3 ? gen_uniform(key, 4, j, i, -off, off) : 0.0;
    y[g] = gen_uniform(key, 5, j, i, -1.0, 1.0);
}

// kernel.cu:5-26 (initialize): T^2 with y = 1.
__global__ void init_t2_kernel(int n, int m, long long rs, long long ss,
                               double* a, double* b, double* c, double* d, double* e, double* y)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)n * m) return;
    int i, j;
    if (ss == 1) { i = (int)(idx / m); j = (int)(idx % m); }
    else { j = (int)(idx / n); i = (int)(idx % n); }
    const long long g = (long long)j * ss + (long long)i * rs;
    double va = 1.0, vb = -4.0, vc = 6.0, vd = -4.0, ve = 1.0;
    if (i == 0) { va = 0.0; vb = 1.0; vc = 5.0; }
    if (i == 1) va = 0.0;
    if (i == n - 2) ve = 0.0;
    if (i == n - 1) { vc = 5.0; vd = 1.0; ve = 0.0; }
    a[g] = va; b[g] = vb; c[g] = vc; d[g] = vd; e[g] = ve; y[g] = 1.0;
}

uint64_t host_splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int pcrp_abi_version(void) { return PCRP_ABI_VERSION; }

const char* pcrp_strerror(int status)
{
    switch (status) {
        case PCRP_OK: return "ok";
        case PCRP_ERR_INVALID_SIZE: return "invalid size (n < 1, m < 1, or n*m >= 2^31)";
        case PCRP_ERR_NULL: return "required pointer is NULL";
        case PCRP_ERR_LAYOUT: return "unknown layout";
        case PCRP_ERR_WORKSPACE: return "workspace too small";
        case PCRP_ERR_HIP: return "HIP runtime error";
        case PCRP_ERR_ARG: return "invalid argument";
        case PCRP_ERR_UNSUPPORTED: return "unsupported by this build";
        default: return "unknown status";
    }
}

size_t pcrp_workspace_size(int n, int m, int precision)
{
    (void)n; (void)m; (void)precision;
    return 0;
}

int pcrp_solve_f64(const double* a, const double* b, const double* c, const double* d,
                   const double* e, double* y, int n, int m, int layout, void* ws,
                   size_t ws_bytes, void* stream)
{
    (void)ws; (void)ws_bytes;
    int st = check_shape(n, m);
    if (st != PCRP_OK) return st;
    if (!a || !b || !c || !d || !e || !y) return PCRP_ERR_NULL;
    SolveArgsF64 A{a, b, c, d, e, y, n, m, 0, 0, pcr_steps(n)};
    if (!layout_strides(n, m, layout, &A.rs, &A.ss)) return PCRP_ERR_LAYOUT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n <= lds_f64_max_n()) return lds_f64_solve(A, s) == 0 ? PCRP_OK : PCRP_ERR_HIP;
    return PCRP_ERR_UNSUPPORTED;
}

int pcrp_solve_f32r_f64(const double* a, const double* b, const double* c, const double* d,
                        const double* e, double* y, int n, int m, int layout, int max_sweeps,
                        double tol, int* sweeps_out, void* ws, size_t ws_bytes, void* stream)
{
    (void)sweeps_out; (void)ws; (void)ws_bytes; (void)stream;
    int st = check_shape(n, m);
    if (st != PCRP_OK) return st;
    if (!a || !b || !c || !d || !e || !y) return PCRP_ERR_NULL;
    if (max_sweeps < 0 || !(tol >= 0.0)) return PCRP_ERR_ARG;
    long long rs, ss;
    if (!layout_strides(n, m, layout, &rs, &ss)) return PCRP_ERR_LAYOUT;
    return PCRP_ERR_UNSUPPORTED;
}

int pcrp_generate(int kind, uint64_t seed, int n, int m, int layout, double* a, double* b,
                  double* c, double* d, double* e, double* y, void* stream)
{
    int st = check_shape(n, m);
    if (st != PCRP_OK) return st;
    if (!a || !b || !c || !d || !e || !y) return PCRP_ERR_NULL;
    if (kind != 0 && kind != 1) return PCRP_ERR_ARG;
    long long rs, ss;
    if (!layout_strides(n, m, layout, &rs, &ss)) return PCRP_ERR_LAYOUT;
    const long long total = (long long)n * m;
    const int bs = 256;
    hipLaunchKernelGGL(generate_kernel, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0,
                       static_cast<hipStream_t>(stream), kind, host_splitmix64(seed), n, m, rs,
                       ss, a, b, c, d, e, y);
    return hipGetLastError() == hipSuccess ? PCRP_OK : PCRP_ERR_HIP;
}

int pcrp_init_t2(int n, int m, int layout, double* a, double* b, double* c, double* d,
                 double* e, double* y, void* stream)
{
    int st = check_shape(n, m);
    if (st != PCRP_OK) return st;
    if (!a || !b || !c || !d || !e || !y) return PCRP_ERR_NULL;
    long long rs, ss;
    if (!layout_strides(n, m, layout, &rs, &ss)) return PCRP_ERR_LAYOUT;
    const long long total = (long long)n * m;
    const int bs = 256;
    hipLaunchKernelGGL(init_t2_kernel, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0,
                       static_cast<hipStream_t>(stream), n, m, rs, ss, a, b, c, d, e, y);
    return hipGetLastError() == hipSuccess ? PCRP_OK : PCRP_ERR_HIP;
}

}  // extern "C"
