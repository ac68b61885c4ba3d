/*
 * rsgpu_host.h - host-only entry points of librsgpu.so: the integer analysis
 * and the first (pivoting) factor that the GPU solver runs on the CPU before
 * handing refactors and solves to the device. They allocate no device memory,
 * so they can be exercised on any machine.
 *
 * Reference counterparts:
 *   rs_host_amd_order     - amd::amd_order (crates/sim-core/src/amd.rs:143-161)
 *   rs_host_btf           - btf::btf_decompose (crates/sim-core/src/btf.rs:143-175)
 *   rs_host_factor_create - NativeKluSolver::do_analyze + the first do_factor
 *                           (native_klu.rs:704-761, 1020-1148); kind 1 gives
 *                           SparseLuSolver's no-pivot factor (sparse_lu.rs:302-537)
 */
#ifndef RSGPU_HOST_H
#define RSGPU_HOST_H

#include <stdint.h>

#include "rs_circuit.h"

#ifdef __cplusplus
extern "C" {
#endif

void rs_host_amd_order(int64_t n, const int64_t *ap, const int64_t *ai, int64_t *perm,
                       int64_t *inv_perm);
/* returns num_blocks; block_ptr needs n + 1 entries */
int64_t rs_host_btf(int64_t n, const int64_t *ap, const int64_t *ai, int64_t *row_perm,
                    int64_t *col_perm, int64_t *block_ptr, int64_t *structural_rank);

typedef struct rs_host_factor rs_host_factor;
rs_host_factor *rs_host_factor_create(int32_t kind, int64_t n, const int64_t *ap,
                                      const int64_t *ai, const double *ax, double pivot_tol,
                                      int32_t ordering, int32_t use_btf);
void rs_host_factor_destroy(rs_host_factor *h);
int64_t rs_host_factor_nblocks(const rs_host_factor *h);
void rs_host_factor_block_sizes(const rs_host_factor *h, int64_t b, int64_t *n, int64_t *nl,
                                int64_t *nu);
void rs_host_factor_block_get(const rs_host_factor *h, int64_t b, int64_t *col_perm,
                              int64_t *row_perm, int64_t *lcp, int64_t *lri, int64_t *ucp,
                              int64_t *uri, double *lx, double *ux, double *udiag);
/* out[8] = {factor levels, max level width, L-solve levels, U-solve levels,
 *           scatter runs, off-diagonal entries, nnz(L), nnz(U)} */
void rs_host_factor_plan_stats(const rs_host_factor *h, int64_t *out);

/* bbd_decompose (crates/sim-core/src/bbd.rs:355-485) with
 * GreedyBisectionPartitioner: perm[new] = old (n), block_ptr (blocks + 1
 * entries, at most num_blocks + 1), border_start; returns the block count. */
int64_t rs_host_bbd_decompose(int64_t n, const int64_t *ap, const int64_t *ai, int64_t num_blocks,
                              int64_t *perm, int64_t *block_ptr, int64_t *border_start);
/* Step-program check on the host (no reference counterpart): the first
 * factor (pivots) on ax0, then the refactor on ax1 and the solve of rhs,
 * executed by the host model of the GPU step programs. x equals
 * NativeKluSolver factor(ax0); factor(ax1); solve(rhs) bit for bit.
 * stats[8] = {refactor steps, refactor ops, solve steps, solve ops, refactor
 * slow chunks, solve slow chunks, refactor forwarded ops, solve forwarded ops}.
 * kind | 0x100 builds the programs on the solver's own plan (positions after
 * the first factor's pivots, as the single-matrix solver runs them) instead
 * of the identity-pivot plan of the batched kernel. */
int32_t rs_host_sp_selftest(int32_t kind, int64_t n, const int64_t *ap, const int64_t *ai,
                            const double *ax0, const double *ax1, const double *rhs, double *x,
                            int64_t *stats);
/* Right-looking refactor schedule on the host (no reference counterpart):
 * the first factor (pivots) on ax0, then the refactor on ax1 run phase by
 * phase with each phase's independent work in reverse order. lx (nnz L),
 * ux (nnz U) and ud (n) are in the solver's position order and equal
 * NativeKluSolver factor(ax0); factor(ax1) bit for bit. Returns 6 when the
 * plan has no right-looking schedule. stats[8] = {phases, tasks, updates,
 * long tasks, column levels, nnz(L), nnz(U), value slots}. */
int32_t rs_host_rl_selftest(int32_t kind, int64_t n, const int64_t *ap, const int64_t *ai,
                            const double *ax0, const double *ax1, double *lx, double *ux, double *ud,
                            int64_t *stats);
/* Flat pipelined solve (solve mode 3) on the host (no reference
 * counterpart): the first factor on ax (kind bits 8..15: the ordering of
 * rs_solver_set_config; bit 16: the transposed records, x solving J^T x =
 * rhs, for single-block plans), then the level records of the
 * solver's plan run level by level, products and rows each in reverse order.
 * x equals NativeKluSolver factor(ax); solve(rhs) bit for bit. Returns 6 when
 * the plan has no level records (BTF off-diagonal entries) or a record is out
 * of range. stats[5] = {levels, most entries in a level, entries, row
 * records, levels of the row sweeps}. */
int32_t rs_host_lp_selftest(int32_t kind, int64_t n, const int64_t *ap, const int64_t *ai,
                            const double *ax, const double *rhs, double *x, int64_t *stats);
/* Unmerged CSC pattern of one stamp pass (mode 0 DC, else TRAN) as
 * SparseBuilder::finalize produces it (mna.rs:320-347); ap/ai may be NULL to
 * query n and nnz. Returns 0, or 6 for an unsupported device kind. */
int32_t rs_host_stamp_pattern(const rs_device *dev, int64_t ndev, int64_t node_count,
                              int32_t mode, int64_t *n, int64_t *nnz, int64_t *ap, int64_t *ai);

/* Supernodal schedule on the host (no reference counterpart): the SparseLU
 * symbolic factor of (ap, ai) (sparse_lu.rs:302-437), its supernodal plan
 * and, when ax is given, the factor of ax and the solve of rhs (x) executed
 * in the order of the GPU kernels. lx/ux/ud receive the factor in the
 * symbolic's CSC order. x, lx, ux, ud equal SparseLuSolver factor(ax);
 * solve(rhs) bit for bit. An unsymmetric pattern is planned on its symmetric
 * closure. hot_min > 0: the count of supernodes whose R must hold a position
 * for its diagonal entry to be summed on its own (default 256). flags bit 0:
 * the plan of A^T on A's symbolic factor (x solves A^T x = rhs; lx and ux
 * are not written). stats[13] = {supernodes, levels, contributions, panel
 * doubles, widest supernode, longest R_J, multiply-subtracts of one factor,
 * gather tiles, nnz(L), nnz(U), entries summed on their own, tile records,
 * 1 when the pattern is structurally symmetric}. */
int32_t rs_host_sn_selftest(int64_t n, const int64_t *ap, const int64_t *ai, const double *ax,
                            const double *rhs, double *x, double *lx, double *ux, double *ud,
                            int64_t hot_min, int64_t flags, int64_t *stats);
/* The adaptive transient controller's power x^y (double-double, rounded
 * once; ddpow.hpp) evaluated on the host for n pairs, with near[i] = 1 where
 * the exact power lies within zone ulp of a rounding midpoint (where the
 * device hands the power to the host's libm). No reference counterpart: the
 * checking aid for f64::powf's role in analysis.rs:173, 530-536. */
int32_t rs_host_pow_dd(const double *x, const double *y, int64_t n, double zone, double *out,
                       uint8_t *near);

#ifdef __cplusplus
}
#endif
#endif
