/*
 * CPU oracle for the batched PCR-Penta solve — TEST INFRASTRUCTURE ONLY.
 * See pcr_oracle.h for who may use it. Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off -fopenmp; the explicit fma() calls are the only
 * fused operations, so the FP64 results are reproducible bit for bit).
 */
#include "pcr_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* main.cu:38-39 + kernel.cu:63: the reduction loop runs floor(log2 n) times
 * for every n the reference accepts (see pcr_oracle_body.inc). */
int oracle_pcr_steps(int n)
{
    int s = 0;
    while ((2L << s) <= (long)n) ++s;
    return n >= 2 ? s : 0;
}

#define REAL double
#define FMA(x, y, z) fma((x), (y), (z))
#define FN(name) name##_f64
#include "pcr_oracle_body.inc"
#undef REAL
#undef FMA
#undef FN

#define REAL long double
#define FMA(x, y, z) fmal((x), (y), (z))
#define FN(name) name##_f80
#include "pcr_oracle_body.inc"
#undef REAL
#undef FMA
#undef FN

/* ---- seeded, layout-independent generator ------------------------------ */

static inline uint64_t splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* u = splitmix64(splitmix64(seed) + (diag<<56 | sys<<24 | row));
 * value = lo + (hi-lo) * ((u>>11) * 2^-53), evaluated without contraction. */
double oracle_gen_uniform(uint64_t seed, int diag, uint64_t sys, uint64_t row,
                          double lo, double hi)
{
    const uint64_t ctr = ((uint64_t)diag << 56) | (sys << 24) | row;
    const uint64_t u = splitmix64(splitmix64(seed) + ctr);
    const double f = (double)(u >> 11) * 0x1.0p-53;
    const double span = hi - lo;
    const double prod = span * f;
    return lo + prod;
}

void oracle_gen_batch(int kind, uint64_t seed, int n, int m, long rs, long ss,
                      double *a, double *b, double *c, double *d, double *e, double *y)
{
    const double off = kind == 1 ? 0.5 : 1.0;
    const double clo = kind == 1 ? 2.2 : 4.5, chi = kind == 1 ? 2.6 : 5.5;
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < n; ++i) {
            const long g = (long)j * ss + (long)i * rs;
            a[g] = i >= 2 ? oracle_gen_uniform(seed, 0, j, i, -off, off) : 0.0;
            b[g] = i >= 1 ? oracle_gen_uniform(seed, 1, j, i, -off, off) : 0.0;
            c[g] = oracle_gen_uniform(seed, 2, j, i, clo, chi);
            d[g] = i <= n - 2 ? oracle_gen_uniform(seed, 3, j, i, -off, off) : 0.0;
            e[g] = i <= n - 3 ? oracle_gen_uniform(seed, 4, j, i, -off, off) : 0.0;
            y[g] = oracle_gen_uniform(seed, 5, j, i, -1.0, 1.0);
        }
}

/* kernel.cu:19-24: a=1, b=-4, c=6, d=-4, e=1, y=1 with the reference's edge
 * rows (including its out-of-range b[0]=1 and d[n-1]=1). */
void oracle_init_t2(int n, int m, long rs, long ss,
                    double *a, double *b, double *c, double *d, double *e, double *y)
{
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < n; ++i) {
            const long g = (long)j * ss + (long)i * rs;
            a[g] = 1.0; b[g] = -4.0; c[g] = 6.0; d[g] = -4.0; e[g] = 1.0; y[g] = 1.0;
            if (i == 0) { a[g] = 0.0; b[g] = 1.0; c[g] = 5.0; }
            if (i == 1) a[g] = 0.0;
            if (i == n - 2) e[g] = 0.0;
            if (i == n - 1) { c[g] = 5.0; d[g] = 1.0; e[g] = 0.0; }
        }
}

/* ---- batch drivers ------------------------------------------------------ */

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static void gather(const double *src, double *dst, int n, long base, long rs)
{
    for (int i = 0; i < n; ++i) dst[i] = src[base + (long)i * rs];
}

void oracle_pcr_batch_f64(const double *a, const double *b, const double *c,
                          const double *d, const double *e, const double *y,
                          double *x, int n, int m, long rs, long ss, int nthreads)
{
    if (nthreads <= 0) nthreads = oracle_max_threads();
#pragma omp parallel num_threads(nthreads)
    {
        double *buf = (double *)malloc(sizeof(double) * 19 * (size_t)n);
        double *sa = buf, *sb = sa + n, *sc = sb + n, *sd = sc + n, *se = sd + n,
               *sy = se + n, *sx = sy + n, *work = sx + n;
#pragma omp for schedule(static)
        for (int j = 0; j < m; ++j) {
            const long base = (long)j * ss;
            gather(a, sa, n, base, rs); gather(b, sb, n, base, rs);
            gather(c, sc, n, base, rs); gather(d, sd, n, base, rs);
            gather(e, se, n, base, rs); gather(y, sy, n, base, rs);
            oracle_pcr_f64(sa, sb, sc, sd, se, sy, sx, n, work);
            for (int i = 0; i < n; ++i) x[base + (long)i * rs] = sx[i];
        }
        free(buf);
    }
}

int oracle_gbsv_batch_f64(const double *a, const double *b, const double *c,
                          const double *d, const double *e, const double *y,
                          double *x, int n, int m, long rs, long ss, int nthreads)
{
    if (nthreads <= 0) nthreads = oracle_max_threads();
    int bad = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : bad)
    {
        double *buf = (double *)malloc(sizeof(double) * 14 * (size_t)n);
        int *ipiv = (int *)malloc(sizeof(int) * (size_t)n);
        double *sa = buf, *sb = sa + n, *sc = sb + n, *sd = sc + n, *se = sd + n,
               *sy = se + n, *sx = sy + n;
        double *ab = (double *)malloc(sizeof(double) * 7 * (size_t)n);
#pragma omp for schedule(static)
        for (int j = 0; j < m; ++j) {
            const long base = (long)j * ss;
            gather(a, sa, n, base, rs); gather(b, sb, n, base, rs);
            gather(c, sc, n, base, rs); gather(d, sd, n, base, rs);
            gather(e, se, n, base, rs); gather(y, sy, n, base, rs);
            if (oracle_gbsv_f64(sa, sb, sc, sd, se, sy, sx, n, ab, ipiv) != 0) ++bad;
            for (int i = 0; i < n; ++i) x[base + (long)i * rs] = sx[i];
        }
        free(ab);
        free(ipiv);
        free(buf);
    }
    return bad;
}

void oracle_pcr_uneven_f64(const double *a, const double *b, const double *c,
                           const double *d, const double *e, const double *y,
                           double *x, const int64_t *off, const int *n_sys, int m,
                           int nthreads)
{
    if (nthreads <= 0) nthreads = oracle_max_threads();
    int nmax = 1;
    for (int j = 0; j < m; ++j) if (n_sys[j] > nmax) nmax = n_sys[j];
#pragma omp parallel num_threads(nthreads)
    {
        double *work = (double *)malloc(sizeof(double) * 12 * (size_t)nmax);
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < m; ++j) {
            const int64_t o = off[j];
            oracle_pcr_f64(a + o, b + o, c + o, d + o, e + o, y + o, x + o, n_sys[j], work);
        }
        free(work);
    }
}

int oracle_gbsv_uneven_f64(const double *a, const double *b, const double *c,
                           const double *d, const double *e, const double *y,
                           double *x, const int64_t *off, const int *n_sys, int m,
                           int nthreads)
{
    if (nthreads <= 0) nthreads = oracle_max_threads();
    int nmax = 1, bad = 0;
    for (int j = 0; j < m; ++j) if (n_sys[j] > nmax) nmax = n_sys[j];
#pragma omp parallel num_threads(nthreads) reduction(+ : bad)
    {
        double *ab = (double *)malloc(sizeof(double) * 7 * (size_t)nmax);
        int *ipiv = (int *)malloc(sizeof(int) * (size_t)nmax);
#pragma omp for schedule(dynamic, 1)
        for (int j = 0; j < m; ++j) {
            const int64_t o = off[j];
            if (oracle_gbsv_f64(a + o, b + o, c + o, d + o, e + o, y + o, x + o, n_sys[j], ab,
                                ipiv) != 0)
                ++bad;
        }
        free(ipiv);
        free(ab);
    }
    return bad;
}
