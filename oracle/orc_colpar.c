/*
 * orc_colpar.c - the reference's column-parallel numeric factor and its
 * triangular solve, restated in C with POSIX threads. Test infrastructure
 * and CPU baseline only: nothing in the product calls it.
 *
 * factor_block_parallel / execute_columns (crates/sim-core/src/native_klu.rs:
 * 402-576): T worker threads take the columns k = tid, tid + T, ... in turn.
 * Each scatters A(:, k) into its own work vector (duplicates summed in CSC
 * order), then for every j in U(:, k) except the diagonal, ascending, waits
 * until column j is published (ready[j], acquire) and applies
 * work -= L(:, j) * u_jk when u_jk != 0. It stores U(:, k), clamps the
 * diagonal to +-1e-14 when |d| < 1e-14, divides L(:, k) by it, clears the
 * work entries it touched and publishes column k (release).
 *
 * With the pivots of an earlier factor (row_perm_inv) this is NativeKLU's
 * refactor; with identity pivots it is SparseLuSolver::numeric_factor
 * (sparse_lu.rs:440-537) value for value: the per-column arithmetic, its
 * order and the clamp are the same, and the thread count changes only who
 * computes a column, never what it reads.
 *
 * The solve is solve_internal (sparse_lu.rs:564-618), the same loop as
 * NativeKLU's solve_single_block (native_klu.rs:1442-1489).
 */
#define _POSIX_C_SOURCE 199309L
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>

#include "orc.h"
#include "orc_util.h"

typedef struct {
    int64_t n;
    int nthreads;
    const int64_t *col_perm, *col_perm_inv, *row_perm_inv;
    const int64_t *lcp, *lri, *ucp, *uri;
    const int64_t *ap, *ai;
    const double *ax;
    double *lx, *ux, *ud;
    atomic_char *ready;
} colpar_job;

typedef struct {
    const colpar_job *job;
    int tid;
} colpar_arg;

static inline int64_t pos_of(const colpar_job *J, int64_t orow) {
    int64_t p = J->col_perm[orow];
    return J->row_perm_inv ? J->row_perm_inv[p] : p;
}

/* One column k of execute_columns, work all zero on entry and exit. */
static void column(const colpar_job *J, int64_t k, double *work, int wait) {
    const int64_t n = J->n, oc = J->col_perm_inv[k];
    const int64_t st = J->ap[oc], en = J->ap[oc + 1];
    for (int64_t idx = st; idx < en; idx++) {
        const int64_t orow = J->ai[idx];
        if (orow < n) work[pos_of(J, orow)] += J->ax[idx];
    }
    const int64_t us = J->ucp[k], ue = J->ucp[k + 1] - 1;
    for (int64_t q = us; q < ue; q++) {
        const int64_t j = J->uri[q];
        if (wait)
            while (!atomic_load_explicit(&J->ready[j], memory_order_acquire)) {
            }
        const double u = work[j];
        if (u != 0.0)
            for (int64_t li = J->lcp[j]; li < J->lcp[j + 1]; li++) work[J->lri[li]] -= J->lx[li] * u;
    }
    for (int64_t q = us; q < ue; q++) J->ux[q] = work[J->uri[q]];
    double d = work[k];
    if (fabs(d) < 1e-14) d = d >= 0.0 ? 1e-14 : -1e-14;
    J->ux[ue] = d;
    J->ud[k] = d;
    const int64_t ls = J->lcp[k], le = J->lcp[k + 1];
    for (int64_t li = ls; li < le; li++) J->lx[li] = work[J->lri[li]] / d;
    for (int64_t idx = st; idx < en; idx++) {
        const int64_t orow = J->ai[idx];
        if (orow < n) work[pos_of(J, orow)] = 0.0;
    }
    for (int64_t li = ls; li < le; li++) work[J->lri[li]] = 0.0;
    for (int64_t q = us; q < ue; q++) work[J->uri[q]] = 0.0;
    work[k] = 0.0;
    if (wait) atomic_store_explicit(&J->ready[k], 1, memory_order_release);
}

static void *worker(void *p) {
    const colpar_arg *a = (const colpar_arg *)p;
    const colpar_job *J = a->job;
    double *work = (double *)orc_xcalloc((size_t)(J->n ? J->n : 1), sizeof(double));
    for (int64_t k = a->tid; k < J->n; k += J->nthreads) column(J, k, work, 1);
    free(work);
    return NULL;
}

int orc_colpar_factor(int64_t n, const int64_t *col_perm, const int64_t *col_perm_inv,
                      const int64_t *row_perm_inv, const int64_t *lcp, const int64_t *lri,
                      const int64_t *ucp, const int64_t *uri, const int64_t *ap, const int64_t *ai,
                      const double *ax, double *lx, double *ux, double *ud, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    colpar_job J = {n, nthreads, col_perm, col_perm_inv, row_perm_inv, lcp, lri, ucp, uri,
                    ap, ai, ax, lx, ux, ud, NULL};
    J.ready = (atomic_char *)orc_xcalloc((size_t)(n ? n : 1), sizeof(atomic_char));
    for (int64_t k = 0; k < n; k++) atomic_init(&J.ready[k], 0);
    pthread_t *th = (pthread_t *)orc_xmalloc((size_t)nthreads * sizeof(pthread_t));
    colpar_arg *args = (colpar_arg *)orc_xmalloc((size_t)nthreads * sizeof(colpar_arg));
    int started = 0, rc = 0;
    for (int t = 1; t < nthreads; t++) {
        args[t].job = &J;
        args[t].tid = t;
        if (pthread_create(&th[t], NULL, worker, &args[t]) != 0) { rc = 1; break; }
        started = t;
    }
    if (rc) {
        /* the threads that started wait on columns owned by ones that did
         * not: run every column here instead, then let them finish */
        double *work = (double *)orc_xcalloc((size_t)(n ? n : 1), sizeof(double));
        for (int64_t k = 0; k < n; k++)
            if ((int)(k % nthreads) == 0 || (int)(k % nthreads) > started) column(&J, k, work, 1);
        free(work);
    } else {
        args[0].job = &J;
        args[0].tid = 0;
        worker(&args[0]);
    }
    for (int t = 1; t <= started; t++) pthread_join(th[t], NULL);
    free(args);
    free(th);
    free(J.ready);
    return 0;
}

double orc_colpar_columns(int64_t n, const int64_t *col_perm, const int64_t *col_perm_inv,
                          const int64_t *row_perm_inv, const int64_t *lcp, const int64_t *lri,
                          const int64_t *ucp, const int64_t *uri, const int64_t *ap,
                          const int64_t *ai, const double *ax, double *lx, double *ux, double *ud,
                          const int64_t *cols, int64_t ncols) {
    colpar_job J = {n, 1, col_perm, col_perm_inv, row_perm_inv, lcp, lri, ucp, uri,
                    ap, ai, ax, lx, ux, ud, NULL};
    double *work = (double *)orc_xcalloc((size_t)(n ? n : 1), sizeof(double));
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int64_t i = 0; i < ncols; i++) column(&J, cols[i], work, 0);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(work);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

void orc_lu_solve(int64_t n, const int64_t *col_perm, const int64_t *row_perm_inv,
                  const int64_t *lcp, const int64_t *lri, const double *lx, const int64_t *ucp,
                  const int64_t *uri, const double *ux, double *rhs) {
    double *t = (double *)orc_xcalloc((size_t)(n ? n : 1), sizeof(double));
    for (int64_t i = 0; i < n; i++) {
        const int64_t p = col_perm[i];
        t[row_perm_inv ? row_perm_inv[p] : p] = rhs[i];
    }
    for (int64_t k = 0; k < n; k++) {
        const double z = t[k];
        for (int64_t li = lcp[k]; li < lcp[k + 1]; li++) t[lri[li]] -= lx[li] * z;
    }
    for (int64_t k = n - 1; k >= 0; k--) {
        t[k] /= ux[ucp[k + 1] - 1];
        const double x = t[k];
        for (int64_t q = ucp[k]; q < ucp[k + 1] - 1; q++) t[uri[q]] -= ux[q] * x;
    }
    for (int64_t i = 0; i < n; i++) rhs[i] = t[col_perm[i]];
    free(t);
}
