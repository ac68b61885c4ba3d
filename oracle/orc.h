/*
 * orc.h - CPU oracle: a plain-C restatement of RustSpice's DC/TRAN hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or
 * calls this code; it exists so tests, the smoke check and the CPU baseline
 * leg of the benchmark can compare the GPU path against the reference's
 * arithmetic. Every function follows one reference file, cited at its
 * definition; the restatement keeps the reference's loop order, tie-breaks
 * and quirks (unmerged CSC, ground kept in the matrix, 1e-14 pivot clamps,
 * first-factor threshold pivoting inside the static pattern, BE sign, etc.).
 *
 * Parity pinning (what checks each part) is listed in oracle/README.md.
 */
#ifndef ORC_H
#define ORC_H

#include <stdint.h>
#include "../include/rs_circuit.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- AMD (crates/sim-core/src/amd.rs:143-411) -------------------------- */
void orc_amd_order(int64_t n, const int64_t *ap, const int64_t *ai,
                   int64_t *perm, int64_t *inv_perm);

/* ---- BTF (crates/sim-core/src/btf.rs:143-550) -------------------------- */
typedef struct orc_btf {
    int64_t n;
    int64_t *row_perm, *col_perm, *row_perm_inv, *col_perm_inv;
    int64_t *block_ptr; /* num_blocks + 1 */
    int64_t num_blocks, structural_rank, num_singletons, max_block_size;
} orc_btf;
void orc_btf_decompose(int64_t n, const int64_t *ap, const int64_t *ai, orc_btf *out);
void orc_btf_free(orc_btf *b);

/* ---- NativeKLU (crates/sim-core/src/native_klu.rs) --------------------- */
/* Both solver structs start with a uint32_t tag, so the Newton drivers
 * (engine.rs with its Box<dyn LinearSolver>) take either as a void pointer. */
#define ORC_MAGIC_KLU 0x4B4C5530u
#define ORC_MAGIC_BBD 0x42424430u
typedef struct orc_klu orc_klu;
orc_klu *orc_klu_new(int64_t n);
void orc_klu_free(orc_klu *s);
void orc_klu_prepare(orc_klu *s, int64_t n);
int orc_klu_analyze(orc_klu *s, int64_t ap_len, const int64_t *ap, int64_t ai_len,
                    const int64_t *ai);
int orc_klu_factor(orc_klu *s, int64_t ap_len, const int64_t *ap, int64_t ai_len,
                   const int64_t *ai, const double *ax);
int orc_klu_solve(orc_klu *s, int64_t n, double *rhs);
void orc_klu_reset_pattern(orc_klu *s);
void orc_klu_set_config(orc_klu *s, double pivot_tol, int ordering, int use_btf);
void orc_klu_set_sparselu(orc_klu *s); /* sparse_lu.rs semantics */
/* stats: factor_count, refactor_count, nnz_l, nnz_u, nblocks, noffdiag */
void orc_klu_stats(const orc_klu *s, int64_t out[6], double *rcond);
/* symbolic/numeric of block b (for bit-exact comparisons) */
int64_t orc_klu_block_n(const orc_klu *s, int64_t b);
void orc_klu_block_sizes(const orc_klu *s, int64_t b, int64_t *nl, int64_t *nu);
void orc_klu_block_get(const orc_klu *s, int64_t b, int64_t *col_perm, int64_t *row_perm,
                       int64_t *l_col_ptr, int64_t *l_row_idx, int64_t *u_col_ptr,
                       int64_t *u_row_idx, double *l_values, double *u_values,
                       double *u_diag);
/* BTF of the last analyze (nblocks==1 -> no BTF) */
int orc_klu_btf_get(const orc_klu *s, int64_t *row_perm, int64_t *col_perm,
                    int64_t *block_ptr);

/* ---- BBD (bbd.rs), BbdSolver (bbd_solver.rs), DenseSolver (solver.rs) -- */
#define ORC_DENSE_SCHUR_MAX 64 /* DENSE_SCHUR_THRESHOLD, bbd_solver.rs:85 */
typedef struct orc_bbd_decomp {
    int64_t n, num_blocks, border_start, border_size;
    int64_t *perm, *inv_perm, *block_ptr; /* perm[new] = old */
} orc_bbd_decomp;
/* assignment: block id, or -1 for border; returns the block count */
int64_t orc_bbd_partition(int64_t n, const int64_t *ap, const int64_t *ai, int64_t num_blocks,
                          int64_t *assign);
void orc_bbd_decompose(int64_t n, const int64_t *ap, const int64_t *ai, int64_t num_blocks,
                       orc_bbd_decomp *d);
void orc_bbd_decomp_free(orc_bbd_decomp *d);
typedef struct orc_bbd orc_bbd;
orc_bbd *orc_bbd_new(int64_t n, int64_t num_blocks);
void orc_bbd_free(orc_bbd *s);
void orc_bbd_prepare(orc_bbd *s, int64_t n);
void orc_bbd_reset_pattern(orc_bbd *s);
int orc_bbd_analyze(orc_bbd *s, int64_t apl, const int64_t *ap, int64_t ail, const int64_t *ai);
int orc_bbd_factor(orc_bbd *s, int64_t apl, const int64_t *ap, int64_t ail, const int64_t *ai,
                   const double *ax);
int orc_bbd_solve(orc_bbd *s, int64_t n, double *rhs);
int orc_bbd_is_bbd(const orc_bbd *s);
/* out[4] = {blocks, border size, schur kind (1 dense, 0 sparse, -1 none), factor count} */
int orc_bbd_info(const orc_bbd *s, int64_t *out);
/* dense LU with partial pivoting; schur != 0: DenseSchurSolver checks (1e-30) */
int orc_dense_build(int64_t n, int64_t apl, const int64_t *ap, const int64_t *ai, const double *ax,
                    double *lu);
int orc_dense_factor(int64_t n, double *lu, int64_t *piv, int schur);
int orc_dense_solve(int64_t n, const double *lu, const int64_t *piv, double *rhs, int schur);

/* ---- MNA assembly + stamps (mna.rs, stamp.rs, bsim/evaluate.rs) -------- */
enum { ORC_MODE_DC = 0, ORC_MODE_TRAN = 1 };
enum { ORC_BE = 0, ORC_TRAP = 1 };

typedef struct orc_tstate {
    /* keyed by device index (the reference keys by unique instance name) */
    double *cap_v;  /* cap_voltage */
    double *cap_i;  /* cap_current */
    uint8_t *has_v, *has_i;
    int method;
} orc_tstate;

typedef struct orc_mna {
    int64_t n, nnz;
    int64_t *ap, *ai;
    double *ax, *rhs;
} orc_mna;

/* Builds the unmerged CSC + RHS exactly as the engine closure does
 * (engine.rs:262-283 for DC, :485-501 for TRAN). x has x_len entries. */
void orc_build_mna(const rs_device *dev, int64_t ndev, int64_t node_count,
                   const double *x, int64_t x_len, int mode, double t, double dt,
                   const orc_tstate *st, double gmin, double source_scale,
                   orc_mna *out);
void orc_mna_free(orc_mna *m);

/* level-1 MOSFET (bsim/evaluate.rs:285-356 via bsim/mod.rs:441-455):
 * out = {ids, gm, gds, gmbs, ieq} */
/* BSIM model table used by orc_build_mna for MOSFETs of level != 1 */
void orc_set_bsim_models(const rs_bsim_model *m, int64_t n);
/* Planted mistakes (tests only; 0 = reference behaviour, 1 = PMOS current
 * sign flipped in the level-1 model). */
void orc_set_mutation(int m);
const rs_bsim_model *orc_bsim_model(int64_t i);
/* evaluate_bsim_dc / evaluate_bsim4_dc (level 54): out[13] = {ids, gm, gds,
 * gmbs, ieq, region, vth, isub, gsub, igs, igd, gigs, gigd} */
void orc_bsim_eval(const rs_bsim_model *m, int level, double w, double l, double vd, double vg,
                   double vs, double vb, double temp, double out[13]);
void orc_level1(const rs_device *d, double vd, double vg, double vs, double vb,
                double out[5]);
double orc_pulse_eval(const double p[7], double t);

/* ---- Newton + engine (newton.rs, engine.rs, analysis.rs, waveform.rs) -- */
typedef struct orc_newton_result {
    int32_t converged, iterations, reason; /* reason: 0 conv, 1 maxiters, 2 solver */
    double final_norm;
} orc_newton_result;

/* DC operating point (engine.rs:249-311). x_out gets n entries on convergence
 * (the full vector incl. aux, newton.rs:98), otherwise the node_count entries
 * of the x the loop ended with (newton.rs:93-107); *x_len the count written. */
int orc_dc_op(const rs_device *dev, int64_t ndev, int64_t node_count,
              const rs_newton_opts *o, void *solver, double *x_out,
              int64_t x_cap, int64_t *x_len, orc_newton_result *res);

typedef struct orc_tran_result {
    int32_t status;       /* 0 converged, 2 failed (RunStatus) */
    int32_t accepted, rejected, newton_iters_total, factor_calls;
    int32_t dc_iters;
    int64_t npoints;      /* stored points incl. t=tstart */
} orc_tran_result;

/* ---- AC analysis (engine.rs:829-1044, stamp.rs:1525-1747, complex_mna.rs,
 * complex_solver.rs) ------------------------------------------------------ */
int64_t orc_ac_frequencies(int sweep, int64_t points, double fstart, double fstop, double *out,
                           int64_t cap);
int64_t orc_ac(const rs_device *dev, int64_t ndev, int64_t node_count, const double *ac_mag,
               const double *ac_phase, const double *x_dc, int64_t nx, const double *freqs,
               int64_t nfreq, double *xout, double *magph, int64_t cap_n, int64_t *n_out);

/* Initial guess of the next orc_tran (x[0..n) over zeros; n = 0 clears). */
void orc_set_initial_guess(const double *x, int64_t n);

/* TRAN (engine.rs:333-632). times[npoints], sols[npoints*x_stride] */
int orc_tran(const rs_device *dev, int64_t ndev, int64_t node_count,
             const rs_newton_opts *o, double tstep, double tstop, double tstart,
             double tmax, void *solver, int64_t max_points, double *times,
             double *sols, int64_t x_stride, orc_tran_result *res);

/* Fixed-step trapezoidal transient (no reference counterpart): one TRAP
 * Newton solve per step of dt, state updated as on an accepted step. */
int orc_tran_fixed(const rs_device *dev, int64_t ndev, int64_t node_count,
                   const rs_newton_opts *o, double dt, int64_t nsteps,
                   void *solver, double *x_final, int64_t x_cap,
                   int64_t *x_len, int32_t *newton_iters_total);

/* Sweep points of engine.rs:715-741; returns the count (values may be NULL). */
int64_t orc_dc_sweep_points(double start, double stop, double step, double *values,
                            int64_t cap);
typedef struct {
    int32_t status;       /* 0 Converged, 1 MaxIters, 2 Failed, 3 no such source */
    int32_t iterations;   /* sum of the returned pass's iterations per point */
    int64_t npoints, nsolved;
} orc_sweep_result;
/* Engine::run_dc_sweep_result (engine.rs:684-820) on device index `source`:
 * sols[nsolved * x_stride]. */
int orc_dc_sweep(const rs_device *dev, int64_t ndev, int64_t node_count,
                 const rs_newton_opts *o, int64_t source, double start, double stop,
                 double step, void *solver, int64_t max_points, double *sols,
                 int64_t x_stride, orc_sweep_result *res);

/* ---- column-parallel factor and solve (orc_colpar.c) -------------------
 * native_klu.rs:402-576 with nthreads workers; row_perm_inv NULL means
 * identity pivots (SparseLU). orc_colpar_columns recomputes the listed
 * columns on one thread from the L values already in lx and returns the
 * seconds it took. */
int orc_colpar_factor(int64_t n, const int64_t *col_perm, const int64_t *col_perm_inv,
                      const int64_t *row_perm_inv, const int64_t *lcp, const int64_t *lri,
                      const int64_t *ucp, const int64_t *uri, const int64_t *ap, const int64_t *ai,
                      const double *ax, double *lx, double *ux, double *ud, int nthreads);
double orc_colpar_columns(int64_t n, const int64_t *col_perm, const int64_t *col_perm_inv,
                          const int64_t *row_perm_inv, const int64_t *lcp, const int64_t *lri,
                          const int64_t *ucp, const int64_t *uri, const int64_t *ap,
                          const int64_t *ai, const double *ax, double *lx, double *ux, double *ud,
                          const int64_t *cols, int64_t ncols);
void orc_lu_solve(int64_t n, const int64_t *col_perm, const int64_t *row_perm_inv,
                  const int64_t *lcp, const int64_t *lri, const double *lx, const int64_t *ucp,
                  const int64_t *uri, const double *ux, double *rhs);

/* analysis.rs:147-200: out = {error_norm, max_node, accept, factor, raw_lte} */
void orc_lte_milne(const double *xb, const double *xt, int64_t n, double abs_tol,
                   double rel_tol, double out[5]);

#ifdef __cplusplus
}
#endif
#endif
