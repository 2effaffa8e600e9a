// Internal host-side interfaces between the C-ABI layer and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace pcrp {

// One batched FP64 solve over a strided view: element i of system j at
// a[j*ss + i*rs]; `steps` reduction steps, final delta = 2^steps.
struct SolveArgsF64 {
    const double *a, *b, *c, *d, *e;
    double* y;
    int n, m;
    long long rs, ss;
    int steps;
};

// A compiled LDS-kernel variant: KIND (0 shared pair equations, 1 redundant),
// NT threads x RPT rows per thread, capacity max_n = NT*RPT rows.
struct LdsVariantF64 {
    int kind, nt, rpt, max_n;
    int (*launch)(const SolveArgsF64&, hipStream_t);
};

int lds_f64_variant_count();
const LdsVariantF64* lds_f64_variant(int id);
int lds_f64_default_variant(int n);
int lds_f64_lds_bytes(const LdsVariantF64& v);
int lds_f64_max_n();

// FP32 PCR + FP64 refinement over the same strided view.
struct SolveArgsF32R {
    const double *a, *b, *c, *d, *e;
    double* y;
    int n, m;
    long long rs, ss;
    int steps;
    int max_sweeps;
    double tol;
    int* sweeps_out;
};

int lds_f32r_max_n();
int lds_f32r_solve(const SolveArgsF32R& A, hipStream_t st);

inline int pcr_steps(long long n)
{
    int s = 0;
    while ((2LL << s) <= n) ++s;
    return n >= 2 ? s : 0;
}

}  // namespace pcrp
