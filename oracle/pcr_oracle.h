/*
 * CPU oracle for the batched PCR-Penta solve — TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load this code, and only as the checker or the timed CPU baseline. The
 * shipped solver (libpcrpenta.so) never links or calls it.
 *
 * Contents:
 *   oracle_pcr_steps        step-count rule of main.cu:38-39 / kernel.cu:63
 *   oracle_pcr_f64 / _f80   PCR-Penta restated from kernel.cu:32-154
 *   oracle_gbsv_f64 / _f80  LAPACK dgbsv (kl = ku = 2) banded LU
 *   oracle_gen_*            seeded counter-based input generator
 *   oracle_init_t2          the reference's test matrix, kernel.cu:5-26
 *   oracle_*_batch          batch drivers (OpenMP over systems)
 */
#ifndef PCR_ORACLE_H
#define PCR_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

int oracle_pcr_steps(int n);

void oracle_pcr_f64(const double *a, const double *b, const double *c,
                    const double *d, const double *e, const double *y,
                    double *x, int n, double *work /* 12*n */);
void oracle_pcr_f80(const long double *a, const long double *b, const long double *c,
                    const long double *d, const long double *e, const long double *y,
                    long double *x, int n, long double *work /* 12*n */);

int oracle_gbsv_f64(const double *a, const double *b, const double *c,
                    const double *d, const double *e, const double *y,
                    double *x, int n, double *ab /* 7*n */, int *ipiv /* n */);
int oracle_gbsv_f80(const long double *a, const long double *b, const long double *c,
                    const long double *d, const long double *e, const long double *y,
                    long double *x, int n, long double *ab, int *ipiv);

/* One uniform draw U[lo,hi) keyed by (seed, diag, sys, row): layout-free. */
double oracle_gen_uniform(uint64_t seed, int diag, uint64_t sys, uint64_t row,
                          double lo, double hi);

/* Fill a batch with generator `kind` (0 = strictly dominant: off-diagonals
 * U[-1,1], main U[4.5,5.5]; 1 = weakly dominant: off-diagonals U[-0.5,0.5],
 * main U[2.2,2.6]; RHS U[-1,1]). Out-of-range diagonal entries are zeroed.
 * Element i of system j is at j*ss + i*rs. */
void oracle_gen_batch(int kind, uint64_t seed, int n, int m, long rs, long ss,
                      double *a, double *b, double *c, double *d, double *e, double *y);

/* The reference's T^2 test system (kernel.cu:19-24), y = 1. */
void oracle_init_t2(int n, int m, long rs, long ss,
                    double *a, double *b, double *c, double *d, double *e, double *y);

/* Batch drivers: solve every system of a strided batch with `nthreads`
 * OpenMP threads (<=0: library default). x may alias nothing. */
void oracle_pcr_batch_f64(const double *a, const double *b, const double *c,
                          const double *d, const double *e, const double *y,
                          double *x, int n, int m, long rs, long ss, int nthreads);
int oracle_gbsv_batch_f64(const double *a, const double *b, const double *c,
                          const double *d, const double *e, const double *y,
                          double *x, int n, int m, long rs, long ss, int nthreads);
/* Uneven batch: system j has n_sys[j] rows stored contiguously at off[j]. */
void oracle_pcr_uneven_f64(const double *a, const double *b, const double *c,
                           const double *d, const double *e, const double *y,
                           double *x, const int64_t *off, const int *n_sys, int m,
                           int nthreads);
/* dgbsv (kl = ku = 2) of every system of an uneven batch; returns the number
 * of singular systems. */
int oracle_gbsv_uneven_f64(const double *a, const double *b, const double *c,
                           const double *d, const double *e, const double *y,
                           double *x, const int64_t *off, const int *n_sys, int m,
                           int nthreads);
int oracle_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
