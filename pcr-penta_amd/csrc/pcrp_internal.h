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

int lds_f64_max_n();
int lds_f64_solve(const SolveArgsF64& A, hipStream_t st);

inline int pcr_steps(long long n)
{
    int s = 0;
    while ((2LL << s) <= n) ++s;
    return n >= 2 ? s : 0;
}

}  // namespace pcrp
