/*
 * rsgpu.h - C ABI of the MI355X linear-solve and stamping core.
 *
 * Drop-in boundary for RustSpice's `trait LinearSolver`
 * (crates/sim-core/src/solver.rs:141-166). A Rust `impl LinearSolver for
 * GpuSolver` (shown in INTEGRATION.md) forwards each trait method to the
 * function named beside it. Slices become pointer + length; every entry point
 * returns an rs_status that maps one-to-one onto `SolverError`
 * (solver.rs:76-92). Handles are single-owner and not thread-safe, matching
 * the trait's `&mut self` methods; a handle may move between threads (`Send`).
 *
 * All numerics are FP64. Indices are int64 on the API and int32 on the device
 * (checked at analyze time).
 */
#ifndef RSGPU_H
#define RSGPU_H

#include <stdint.h>

#include "rs_circuit.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: SolverError variants (solver.rs:76-92) ---------------- */
typedef enum rs_status {
    RS_OK = 0,
    RS_ERR_ANALYZE_FAILED = 1,   /* SolverError::AnalyzeFailed           */
    RS_ERR_FACTOR_FAILED = 2,    /* SolverError::FactorFailed            */
    RS_ERR_SOLVE_FAILED = 3,     /* SolverError::SolveFailed             */
    RS_ERR_SINGULAR = 4,         /* SolverError::SingularMatrix{pivot}   */
    RS_ERR_ILL_CONDITIONED = 5,  /* SolverError::IllConditioned{rcond}   */
    RS_ERR_INVALID_MATRIX = 6,   /* SolverError::InvalidMatrix{reason}   */
    RS_ERR_DEVICE = 7            /* HIP runtime error (no reference twin) */
} rs_status;

/* ---- solver kinds: SolverType variants this library implements -------------
 * (solver.rs:121-139, factory solver.rs:169-200). Both share the GPU kernels;
 * they differ only in host symbolic choices, exactly as the reference solvers
 * do. */
typedef enum rs_solver_kind {
    RS_SOLVER_NATIVEKLU = 0, /* NativeKluSolver: BTF + per-block AMD, threshold
                                pivoting on the first factor, pivots reused by
                                refactors (native_klu.rs)                        */
    RS_SOLVER_SPARSELU = 1,  /* SparseLuSolver: AMD, no pivoting, 1e-14 pivot
                                perturbation (sparse_lu.rs)                      */
    RS_SOLVER_BBD = 2,       /* BbdSolver with GreedyBisectionPartitioner: SparseLU
                                diagonal blocks, dense (border <= 64) or SparseLU
                                Schur complement, SparseLU fallback for a trivial
                                decomposition (bbd.rs, bbd_solver.rs). Default 4
                                target blocks (SolverType::Bbd, solver.rs:186)    */
    RS_SOLVER_DENSE = 3      /* DenseSolver: row-major LU with partial pivoting
                                (solver.rs:595-724)                              */
} rs_solver_kind;

typedef struct rs_solver rs_solver;

/* NativeKluStats (native_klu.rs:584-594) plus GPU counters. */
typedef struct rs_solver_stats {
    int64_t factor_count;    /* full (pivoting) factors                     */
    int64_t refactor_count;  /* refactors with reused pivots                */
    int64_t nnz_l, nnz_u, nblocks, noffdiag;
    double rcond;            /* min|u_ii| / max|u_ii| (native_klu.rs:1406)  */
    int64_t parallel_factor; /* 1 when the last factor ran on the GPU       */
    int64_t n_factor_levels, max_level_width, n_lsolve_levels, n_usolve_levels;
    double refactor_flops, solve_flops; /* algorithmic work on the pattern  */
    int64_t plan_bytes;      /* device bytes of the index plan               */
} rs_solver_stats;

/* create_solver(kind, n) (solver.rs:169-200) */
rs_solver *rs_solver_create(int32_t kind, int64_t n);
void rs_solver_destroy(rs_solver *h);
/* LinearSolver::prepare (native_klu.rs:1604-1610): a new n resets the pattern */
void rs_prepare(rs_solver *h, int64_t n);
/* LinearSolver::analyze (native_klu.rs:1612-1630): cached while (ap, ai)
 * are unchanged; ap_len must be n + 1 */
int32_t rs_analyze(rs_solver *h, const int64_t *ap, int64_t ap_len, const int64_t *ai,
                   int64_t ai_len);
/* LinearSolver::factor (native_klu.rs:1632-1639): re-analyzes on a pattern
 * change; the first factor of a pattern picks pivots (host), later factors
 * of the same pattern refactor on the GPU with those pivots. ax is the
 * unmerged CSC value array (duplicates summed in order). */
int32_t rs_factor(rs_solver *h, const int64_t *ap, int64_t ap_len, const int64_t *ai,
                  int64_t ai_len, const double *ax, int64_t ax_len);
/* LinearSolver::solve (native_klu.rs:1641-1657): rhs is overwritten with x */
int32_t rs_solve(rs_solver *h, double *rhs, int64_t n);
/* LinearSolver::reset_pattern (native_klu.rs:1661-1672) */
void rs_reset_pattern(rs_solver *h);
/* LinearSolver::name / set_parallel_threads (solver.rs:158-165). The thread
 * count is recorded and ignored: the GPU schedule is fixed by the pattern. */
const char *rs_solver_name(const rs_solver *h);
void rs_set_parallel_threads(rs_solver *h, int64_t threads);
/* NativeKluSolver::set_pivot_tolerance / set_ordering / set_btf
 * (native_klu.rs:664-681) */
void rs_set_config(rs_solver *h, double pivot_tol, int32_t ordering, int32_t use_btf);
int32_t rs_get_stats(const rs_solver *h, rs_solver_stats *out);
/* BbdSolver::new(n, partitioner, num_blocks) target block count (max(2, .));
 * takes effect at the next analyze. */
void rs_set_bbd_blocks(rs_solver *h, int64_t num_blocks);
/* out[4] = {diagonal blocks (1: SparseLU fallback), border size, Schur kind
 * (1 dense, 0 sparse, -1 none), factor count} of an RS_SOLVER_BBD solver. */
int32_t rs_bbd_info(const rs_solver *h, int64_t *out);
/* Dense Schur product of an RS_SOLVER_BBD solver on FP64 MFMA tiles (no
 * reference counterpart: bbd_solver.rs:790-806 accumulates
 * S = B_T - sum_k C_k B_kk^-1 R_k one term at a time). The coupling columns
 * (a block's columns that hold C entries) are packed into K x bs operands and
 * multiplied on v_mfma_f64_16x16x4f64 tiles, split over K and summed in a fixed
 * order: deterministic, but the terms are not added in the reference's order,
 * so results agree with it to rounding (1e-9 relative is the bar), not bit
 * for bit. mode 0 (default): off; 1: when the dense Schur path (border <= 64)
 * has K >= 64 and both packed operands are at least a quarter nonzero;
 * 2: whenever the dense Schur path has K > 0. Takes effect at the next factor.
 * rs_bbd_mfma_info: out[4] = {mode, 1 if the last factor used the tiles,
 * K, nonzero fraction of the packed C operand in parts per thousand}. */
void rs_set_bbd_mfma(rs_solver *h, int32_t mode);
int32_t rs_bbd_mfma_info(const rs_solver *h, int64_t *out);
/* Launch shapes of the refactor and solve kernels (no reference
 * counterpart). By default the library times every shape (1 to 16
 * wavefronts; the column or the right-looking refactor schedule; work
 * vectors or values in LDS when they fit) on the first device refactor and
 * the first solve of each pattern, keeps the fastest and remembers it per
 * pattern for the process; results are bit-identical for every shape.
 * rs_set_launch_shape fixes the shapes instead: threads a power of 2 in
 * 64..1024; ref_mode 0/1 column kernel (1: work vectors in LDS), 2/3
 * right-looking schedule (3: values in LDS), 4 step program; sol_mode 0 level
 * sweeps with the solve vector in global memory, 1 in LDS, 2 step program,
 * 3 pipelined level sweeps (the next levels' row records and first entries
 * loaded while a level computes; solve vector in LDS). A mode is checked
 * against the plan by the next factor when none is built yet, which then
 * fails with RS_ERR_INVALID_MATRIX if it does not fit.
 * rs_launch_shape: out[6] = {refactor threads, refactor mode, solve threads,
 * solve mode, best refactor ns, best solve ns}; returns bit 0 refactor tuned,
 * bit 1 solve tuned, bit 2 shapes fixed by the caller. */
void rs_set_autotune(rs_solver *h, int32_t on);
/* Pivot-growth check of device refactors (no reference counterpart; off by
 * default, as the reference always reuses the first factor's pivots on an
 * unchanged pattern, native_klu.rs:1244-1255). With limit > 0, every device
 * refactor measures max|U| / max|A| on the device; above the limit, or with a
 * non-finite U, the pattern is factored afresh on the host with new pivots
 * (counted as a factor, not a refactor). rs_growth_fallbacks: refactors so
 * replaced; rs_last_growth: the last measured ratio. */
void rs_set_growth_check(rs_solver *h, double limit);
int64_t rs_growth_fallbacks(const rs_solver *h);
double rs_last_growth(const rs_solver *h);
int32_t rs_set_launch_shape(rs_solver *h, int32_t ref_threads, int32_t ref_mode, int32_t sol_threads,
                            int32_t sol_mode);
int32_t rs_launch_shape(const rs_solver *h, int64_t *out);
/* The payload of the last RS_ERR_SINGULAR: SolverError::SingularMatrix's
 * pivot (solver.rs:84, raised by the dense Schur factor, bbd_solver.rs:138-140);
 * -1 if none. RS_ERR_ILL_CONDITIONED's rcond is rs_get_stats' rcond. */
int64_t rs_error_pivot(const rs_solver *h);
/* Text of the last error (SolverError Display, solver.rs:94-117). */
int32_t rs_last_error(const rs_solver *h, char *buf, int64_t len);

/* ---- device-resident hot path (no reference counterpart: the reference has
 * no device boundary). The pattern must be the one last factored; pointers
 * are device memory; `stream` is a hipStream_t (NULL = default stream). ---- */
int32_t rs_refactor_device(rs_solver *h, const double *ax_dev, void *stream);
int32_t rs_solve_device(rs_solver *h, const double *rhs_dev, double *x_dev, void *stream);
/* A^T x = rhs with the stored L/U (no reference counterpart: the adjoint's
 * transposed solve, A^T = Q^T U^T L^T P Q, reusing the factor of the last
 * rs_factor / rs_refactor_device). Host (in place, like rs_solve) and device
 * forms. Needs a single-block level-set factor: BbdSolver, Dense, the
 * supernodal path and BTF-split matrices give RS_ERR_SOLVE_FAILED. Not the
 * reference's arithmetic order (its sums are reduced across a wavefront):
 * an FP64 solve, checked against a dense transposed solve. */
int32_t rs_tsolve(rs_solver *h, double *rhs, int64_t n);
int32_t rs_tsolve_device(rs_solver *h, const double *rhs_dev, double *x_dev, void *stream);
/* rs_solve_device with no device allocation (no reference counterpart): for
 * callers that cannot allocate, such as code under HIP graph capture or a
 * serving loop that forbids hipMalloc. Same kernels and bit-identical
 * results; the solve vector they use besides rhs and x is a static device
 * array of rs_noalloc_max_n() = 2,097,152 doubles (16 MiB per GPU, held
 * until the process ends) instead of the handle's own, and the launch shape
 * is the one the handle holds (no tuning). n above that size, BbdSolver and
 * Dense handles give RS_ERR_SOLVE_FAILED. Every call on one GPU shares that
 * array: calls that overlap on one GPU (any handles, any streams) race. */
int32_t rs_solve_device_noalloc(rs_solver *h, const double *rhs_dev, double *x_dev, void *stream);
int64_t rs_noalloc_max_n(void);
/* Copies the current factor values to host arrays sized by rs_get_stats or
 * rs_get_factor_pattern (lx[nnz_l], ux[nnz_u], udiag[n]); position-space
 * order. A NULL array is skipped. */
int32_t rs_get_factor_values(rs_solver *h, double *lx, double *ux, double *udiag);
/* The pattern behind rs_get_factor_values for a one-block factor (the
 * reference's SymbolicFactor, native_klu.rs:92-107, or SparseLuSolver's
 * perm and L/U pointers, sparse_lu.rs:146-159): sizes[3] = {n, nnz(L),
 * nnz(U)}; col_perm[original column] = position; L
 * strictly lower with rows ascending, U with rows ascending and the diagonal
 * last, both column pointers n + 1 long. Array pointers may be NULL (sizes
 * only). A BbdSolver handle answers for its whole-matrix fallback; several
 * BTF or BBD blocks give RS_ERR_FACTOR_FAILED (with several BTF blocks,
 * sizes still holds the array lengths rs_get_factor_values fills). */
int32_t rs_get_factor_pattern(rs_solver *h, int64_t *sizes, int64_t *col_perm, int64_t *lcp,
                              int64_t *lri, int64_t *ucp, int64_t *uri);

/* ---- many matrices per launch (no reference counterpart: N independent
 * LinearSolver refactor/solve calls, native_klu.rs:1632-1657, as one launch
 * each) ----------------------------------------------------------------------
 * rs_solver_batch_create takes `count` CSC matrices concatenated: matrix m
 * has n[m] columns, n[m] + 1 column pointers starting at 0 in ap_cat, and
 * its nnz = ap[n] row indices and values next in ai_cat / ax_cat. Each is
 * analyzed and factored on the host as rs_factor's first factor would
 * (kind NativeKLU: AMD, BTF, threshold pivots; SparseLU: no pivots).
 * rs_solver_batch_refactor then refactors all of them in one launch from
 * device values in the same concatenated order (each NativeKLU matrix reuses
 * its pivots), rs_solver_batch_solve solves all, one workgroup per matrix,
 * with device vectors concatenated by matrix (rs_solver_batch_layout: sizes
 * = {count, total nnz, total n}, the offsets of each matrix). Results equal
 * a separate solver per matrix bit for bit. Errors are kept for
 * rs_solver_batch_last_error (a failed create still returns a handle). */
typedef struct rs_solver_batch rs_solver_batch;
rs_solver_batch *rs_solver_batch_create(int32_t kind, int64_t count, const int64_t *n,
                                        const int64_t *ap_cat, const int64_t *ai_cat,
                                        const double *ax_cat);
void rs_solver_batch_destroy(rs_solver_batch *b);
int32_t rs_solver_batch_last_error(const rs_solver_batch *b, char *buf, int64_t len);
int32_t rs_solver_batch_layout(const rs_solver_batch *b, int64_t *sizes, int64_t *ax_off,
                               int64_t *vec_off);
int32_t rs_solver_batch_refactor(rs_solver_batch *b, const double *ax_dev, void *stream);
int32_t rs_solver_batch_solve(rs_solver_batch *b, const double *rhs_dev, double *x_dev, void *stream);
/* Factor values of matrix m (lx[nnz_l], ux[nnz_u], udiag[n], position order,
 * as rs_get_factor_values gives them for a single solver). */
int32_t rs_solver_batch_factor_values(rs_solver_batch *b, int64_t m, double *lx, double *ux,
                                      double *udiag);

/* ---- circuit-level boundary: DeviceStamp (stamp.rs:94-121) for a whole
 * instance list, and the Newton drivers of engine.rs / newton.rs ---------- */
typedef struct rs_circuit rs_circuit;

enum rs_stamp_mode { RS_MODE_DC = 0, RS_MODE_TRAN_BE = 1, RS_MODE_TRAN_TRAP = 2 };

/* NewtonResult (newton.rs:50-57); reason: 0 Converged, 1 MaxIters, 2 SolverFailure */
typedef struct rs_newton_result {
    int32_t converged, iterations, reason, pad;
    double final_norm;
} rs_newton_result;

/* TRAN summary (engine.rs:603-631); status: 0 Converged, 2 Failed */
typedef struct rs_tran_result {
    int32_t status, accepted, rejected, newton_iters_total, factor_calls, dc_iters;
    int64_t npoints;
} rs_tran_result;

/* Copies the instance list to the device (node numbers as in rs_circuit.h). */
rs_circuit *rs_circuit_create(const rs_device *dev, int64_t ndev, int64_t node_count);
/* No reference counterpart (the reference rebuilds its circuit): replaces
 * device k's values (value, p[], for a level-1 MOSFET its W, L and model
 * parameters) in place, keeping its kind, nodes and flags, so the stamp
 * pattern, the solvers' analyses and the adjoint's plans stay valid: finite
 * differences and parameter sweeps without a new analysis. BSIM devices are
 * refused (their model index selects a pattern-changing table). */
int32_t rs_circuit_set_device(rs_circuit *c, int64_t k, const rs_device *d);
void rs_circuit_destroy(rs_circuit *c);
int32_t rs_circuit_last_error(const rs_circuit *c, char *buf, int64_t len);
/* Model table of the circuit's BSIM MOSFETs (rs_device.p[RS_MP_LEVEL] != 1,
 * indexed by p[RS_MP_MODEL]); copied to the device. Circuits with BSIM
 * devices stamp on the host-driven Newton path: the reference stamps the
 * gmbs, substrate and gate-current groups only above current thresholds
 * (stamp.rs:441-509), so the pattern can change between iterations. */
int32_t rs_circuit_set_bsim_models(rs_circuit *c, const rs_bsim_model *models, int64_t n);
/* evaluate_bsim_dc / evaluate_bsim4_dc (bsim/evaluate.rs:77-279, 387-620;
 * level 54 is BSIM4, any other level but 1 BSIM3, bsim/mod.rs:430-472) on
 * the GPU for n bias points: bias[7*i..] = {w, l, vd, vg, vs, vb, temp},
 * out[13*i..] = {ids, gm, gds, gmbs, ieq, region (0 cutoff, 1 linear,
 * 2 saturation), vth, isub, gsub, igs, igd, gigs, gigd}. Host arrays. */
int32_t rs_mos_eval(const rs_bsim_model *models, int64_t nmodels, const int32_t *model,
                    const int32_t *level, const double *bias, int64_t n, double *out);
/* rs_stamp with x in device memory and ax / rhs written to device memory on
 * `stream` (no host round trip; same values bit for bit). The pattern is the
 * one rs_circuit_pattern reports afterwards. */
int32_t rs_stamp_device(rs_circuit *c, int32_t mode, const double *x_dev, int64_t x_len, double t,
                        double dt, double gmin, double source_scale, double *ax_dev, int64_t ax_cap,
                        double *rhs_dev, int64_t rhs_cap, int64_t *nnz_out, int64_t *n_out,
                        void *stream);
/* Unmerged CSC pattern the stamp loop produces in `mode` (the engine closure
 * of engine.rs:262-283 / :485-501): call with ap = ai = NULL for the sizes. */
int32_t rs_circuit_pattern(rs_circuit *c, int32_t mode, int64_t *n, int64_t *nnz, int64_t *ap,
                           int64_t *ai);
/* DC sensitivities by the adjoint method (no reference counterpart).
 * x: a converged DC solution (x_len >= node_count, as returned by rs_dc_op);
 * probe: the unknown whose sensitivity is wanted (a node index). Solves
 * J(x)^T lambda = e_probe with J the exact Jacobian of the stamped DC
 * currents (SparseLU factor on the GPU, in AMD order with the unknowns that
 * lack a diagonal entry, the sources' aux rows, ordered last so that no pivot
 * is a structural zero; transposed triangular solves reusing
 * it) and writes dx[probe]/dp for every device into dp[ndev]: resistors
 * d/dR, V and I sources d/dvalue, MOSFETs d/dW, diodes d/dIs, capacitors 0.
 * lambda (n entries) is optional. gmin: the value the DC OP used. */
int32_t rs_dc_adjoint(rs_circuit *c, const double *x, int64_t x_len, int64_t probe, double gmin,
                      double *dp, double *lambda);
/* 1: rs_tran_fixed runs its steps after the first TRAP step as captured
 * hipGraphs (no reference counterpart): one graph per step holds two
 * Newton iterations (device evaluation, assembly, refactor, solve and the
 * Newton loop's convergence test and damping on the device, each kernel a
 * no-op once the step has converged) and the accepted-step capacitor update;
 * the host reads the step's state once per launch. Results are bit-identical
 * to the host-driven loop. Used when the circuit has no BSIM devices, no
 * gmin or source stepping is asked for and the solver is NativeKLU or
 * SparseLU; otherwise the host-driven loop runs. 0 (default): host-driven.
 * rs_circuit_graph_steps: steps run through graphs so far.
 * The same switch runs rs_tran's adaptive loop (engine.rs:463-602) on the
 * device after its first accepted step: graphs of 16 Newton-iteration slots
 * with the step controller on the device (Milne LTE, PI controller and its
 * emergency reductions, breakpoints and settling, BE then TRAP solve per
 * attempt, the accepted points into a ring the host drains once per launch);
 * the controller's powers are evaluated in double-double, and the ~4 % of
 * them within 0.02 ulp of a rounding midpoint are handed to the host's libm
 * (the graph stops until the host writes them back). Results are
 * bit-identical to the host-driven loop. rs_circuit_tran_graph_info: out[3]
 * = {rs_tran runs that used the device controller, graph launches, powers
 * evaluated by the host}. */
void rs_circuit_set_tran_graph(rs_circuit *c, int32_t on);
int64_t rs_circuit_graph_steps(const rs_circuit *c);
int32_t rs_circuit_tran_graph_info(const rs_circuit *c, int64_t *out);
/* Transient sensitivities by the discrete adjoint (no reference
 * counterpart). Runs rs_tran_fixed's transient (DC OP, then nsteps TRAP
 * steps of dt, the accepted-step capacitor history of stamp.rs:615-668
 * between steps) on solver s, keeping a checkpoint (x_k, x_{k-1}, capacitor
 * history) every `every` steps; then, segment by segment from the last
 * checkpoint, recomputes the segment's steps and solves the adjoint steps
 * backward. every <= 0 keeps every state x_0..x_N in host memory instead
 * (nsteps x n doubles) and recomputes nothing; info.checkpoints then counts
 * the stored states. The backward solves: J_k^T lambda_k = e_probe [k = nsteps] - sum over the two later
 * steps of (dF_j/dx_k)^T lambda_j (their capacitor currents read v_k), with
 * J_k the exact Jacobian of step k's stamped currents (SparseLU as above,
 * transposed solves with refinement). dp[ndev] receives dx_N[probe]/dp:
 * resistors d/dR, capacitors d/dC, DC V and I sources d/dvalue, MOSFETs d/dW,
 * diodes d/dIs; PULSE sources 0. probe is a node index. The adjoint
 * differentiates each step's exact root, which the reference Newton loop
 * reaches only within its solver's factorization error: the DC OP runs the
 * reference loop on s and then Newton on the exact Jacobian (a Jacobian
 * stamp whose right-hand side is consistent with it, factored by SparseLU
 * with zero-diagonal unknowns last); every TRAP step, in the forward and in
 * each recomputation alike, runs that exact-Jacobian Newton from x_{k-1}
 * (the reference loop first if it does not converge). x_final (x_cap) gets
 * x_N. info: polish_iters = exact-Jacobian Newton iterations, forward_iters =
 * reference Newton iterations; status 0 ok, 1 DC OP failed, 2 a step failed
 * (no gradient). */
typedef struct rs_tadj_info {
    int32_t status, checkpoints, max_refine, polish_iters;
    int64_t forward_iters, recompute_iters, recomputed_steps, adjoint_solves;
    int64_t inexact_steps; /* steps (forward and recomputation) whose exact root was out of
                              reach: the reference loop's point is kept there */
    double forward_ms, recompute_ms, backward_ms; /* host wall clock of the three phases */
} rs_tadj_info;
int32_t rs_tran_adjoint(rs_circuit *c, rs_solver *s, const rs_newton_opts *o, double dt,
                        int64_t nsteps, int64_t probe, int64_t every, double *dp,
                        double *x_final, int64_t x_cap, int64_t *x_len, rs_tadj_info *info);
/* 1 (default): a DC OP whose solver starts from scratch (the circuit has aux
 * rows and the solver its default configuration, no stepping) runs as one
 * device-resident launch, with results identical to the host-driven loop;
 * afterwards the solver keeps its counts but no factorization. 0: every
 * Newton iteration is driven from the host. */
void rs_circuit_set_device_newton(rs_circuit *c, int32_t on);
/* Initial guess of the next rs_tran's DC operating point: x[0..n) over the
 * zeros engine.rs:340-348 starts from (the netlist's .ic values by node);
 * n <= node_count, n = 0 clears it. */
int32_t rs_circuit_set_initial_guess(rs_circuit *c, const double *x, int64_t n);
/* TransientState per device (stamp.rs:699-718), device-index keyed: each
 * array holds the circuit's ndev entries. */
int32_t rs_circuit_set_cap_state(rs_circuit *c, const double *cap_v, const double *cap_i,
                                 const uint8_t *has_v, const uint8_t *has_i);
/* One build(x) of the engine closure on the device: ax (unmerged, nnz) and
 * rhs (n) exactly as SparseBuilder::finalize and the RHS vector hold them.
 * *nnz and *n receive the sizes; when they exceed ax_cap / rhs_cap nothing is
 * copied and RS_ERR_INVALID_MATRIX is returned (with BSIM devices the
 * pattern depends on x). rs_circuit_pattern afterwards gives this pattern. */
int32_t rs_stamp(rs_circuit *c, int32_t mode, const double *x, int64_t x_len, double t, double dt,
                 double gmin, double source_scale, double *ax_out, int64_t ax_cap,
                 double *rhs_out, int64_t rhs_cap, int64_t *nnz, int64_t *n);
/* Engine::run_dc_result (engine.rs:249-311) with run_newton_with_stepping.
 * x_out receives the solution (n entries) when Newton converged, otherwise
 * the iterate the loop ended with (node_count entries: the last damped x of
 * newton.rs:93-107); *x_len is the count written either way. */
int32_t rs_dc_op(rs_circuit *c, rs_solver *s, const rs_newton_opts *o, double *x_out, int64_t x_cap,
                 int64_t *x_len, rs_newton_result *res);
/* Engine::run_tran_result_with_params (engine.rs:333-632): BE + TRAP solves,
 * Milne LTE, PI step control, PULSE breakpoints. Stores up to max_points
 * (time, x[x_stride]) pairs. */
int32_t rs_tran(rs_circuit *c, rs_solver *s, const rs_newton_opts *o, double tstep, double tstop,
                double tstart, double tmax, int64_t max_points, double *times, double *sols,
                int64_t x_stride, rs_tran_result *res);
/* DC sweep summary (RunResult of engine.rs:684-820); status: 0 Converged,
 * 1 MaxIters, 2 Failed (solver failure), 3 source index out of range
 * ("DC sweep source not found"). */
typedef struct rs_sweep_result {
    int32_t status, iterations, factor_calls, pad;
    int64_t npoints;  /* sweep values (RunResult.sweep_values)               */
    int64_t nsolved;  /* points with a solution (RunResult.sweep_solutions)  */
    double fail_value; /* sweep value that failed to converge (status 1, 2)  */
} rs_sweep_result;
/* Sweep values of Engine::run_dc_sweep_result (engine.rs:715-741): points
 * start + i*step up to stop, plus stop itself when the last point falls more
 * than half a step short. Returns the count; writes up to cap values. */
int64_t rs_dc_sweep_points(double start, double stop, double step, double *values, int64_t cap);
/* Engine::run_dc_sweep_result (engine.rs:684-820): for each sweep value, sets
 * device `source`'s value (a PULSE source becomes a plain DC source, as the
 * reference replaces the value text) and runs run_newton_with_stepping from
 * the previous point's solution on one solver, prepared once. Stops at the
 * first point that does not converge. The device keeps the last swept value,
 * as the reference's instance does. Stores up to max_points values and
 * solutions (x[x_stride], the full MNA vector of each converged point). */
int32_t rs_dc_sweep(rs_circuit *c, rs_solver *s, const rs_newton_opts *o, int64_t source,
                    double start, double stop, double step, int64_t max_points, double *values,
                    double *sols, int64_t x_stride, rs_sweep_result *res);
/* Fixed-step trapezoidal transient (no reference counterpart): DC OP, then
 * one TRAP Newton solve per step of dt with the accepted-step state update.
 * status: 0 ok, 1 DC OP failed, 2 a step failed. */
int32_t rs_tran_fixed(rs_circuit *c, rs_solver *s, const rs_newton_opts *o, double dt,
                      int64_t nsteps, double *x_final, int64_t x_cap, int64_t *x_len,
                      int32_t *newton_iters, int32_t *status);

/* Continues a fixed-step trapezoidal transient (no reference counterpart):
 * nsteps TRAP steps of dt from time t0, Newton from x_io (*x_len entries,
 * node_count .. x_cap) and the circuit's current capacitor history (left by
 * rs_tran_fixed or a previous call), with the accepted-step update after
 * each step. A run split into calls takes the same steps, iterations and
 * values as one rs_tran_fixed call, on any solver (a BbdSolver's SparseLU
 * fallback included). x_io and *x_len receive the last state. status: 0 ok,
 * 2 a step failed. */
int32_t rs_tran_steps(rs_circuit *c, rs_solver *s, const rs_newton_opts *o, double t0, double dt,
                      int64_t nsteps, double *x_io, int64_t x_cap, int64_t *x_len,
                      int32_t *newton_iters, int32_t *status);

/* ---- batched DC operating points (no reference counterpart: the reference
 * runs one Engine per circuit). Instance k owns devices
 * dev[dev_ptr[k] .. dev_ptr[k+1]) over node_count[k] nodes. One launch runs
 * Engine::run_dc_result (engine.rs:249-311) for every instance with its own
 * fresh solver state: the same Newton iterations, pivots and results as one
 * rs_dc_op per instance. kind is an rs_solver_kind. ----------------------- */
typedef struct rs_batch rs_batch;
rs_batch *rs_batch_create(const rs_device *dev, const int64_t *dev_ptr, const int64_t *node_count,
                          int64_t ninst, int32_t kind);
void rs_batch_destroy(rs_batch *b);
int32_t rs_batch_last_error(const rs_batch *b, char *buf, int64_t len);
/* out[5] = {instances, distinct patterns, largest n, persistent grid size,
 *           doubles of L/U values kept in LDS per instance (0: in HBM)} */
int32_t rs_batch_info(const rs_batch *b, int64_t *out);
/* Enqueues the batched Newton solve on `stream` (hipStream_t, NULL = default). */
int32_t rs_batch_dc_op(rs_batch *b, const rs_newton_opts *o, void *stream);
/* Copies results: res[k] (pad = number of pivoting factors) and, at
 * x_out + x_ptr[k] (x_ptr has ninst + 1 entries, instance k owns its n
 * entries), the solution of a converged instance or, otherwise, the iterate
 * its Newton loop ended with (node_count entries, the aux entries zero). */
int32_t rs_batch_results(rs_batch *b, rs_newton_result *res, double *x_out, int64_t *x_ptr);

/* Per-phase shader-clock cycles summed over the wavefronts of later
 * rs_batch_dc_op calls (0 device evaluation, 1 ax/RHS gathers, 2 RHS sums,
 * 3 first factor, 4 refactor, 5 scatter-slot map, 6 solve, 7 Newton update
 * and bookkeeping). on != 0 clears and enables; out has 8 entries. No
 * reference counterpart: measurement only. */
int32_t rs_batch_phase_profile(rs_batch *b, int32_t on);
int32_t rs_batch_phase_cycles(rs_batch *b, int64_t *out);

/* ---- measurement (no reference counterpart) -------------------------------
 * Kernel timing with HIP events recorded on each kernel's launch stream.
 * Index: 0 refactor, 1 solve, 2 device evaluation, 3 MNA assembly,
 * 4 batched Newton (k_batch_dcop), 5 each k_sn_gather launch of a
 * supernodal factor (also inside index 0), 6 a BbdSolver's Schur product
 * (k_bbd_schur, or the MFMA tiles with packing and chunk sum). ms_total and
 * counts hold 7 entries.
 * Enabling (or disabling) clears the totals. */
void rs_profile_enable(int32_t on);
void rs_profile_get(double *ms_total, int64_t *counts);
/* Supernodal factor and solve (no reference counterpart: a schedule of the
 * same arithmetic). mode 0: off; 1 (default): SparseLU patterns (and a
 * BbdSolver's SparseLU fallback) of n >= 20,000; 2: every SparseLU pattern.
 * Such a pattern is factored on the device from its first factor on: every
 * entry of L and U receives the reference's terms in the reference's order
 * (sparse_lu.rs numeric_factor), grouped into dense supernode panels that
 * the levels of the supernodal elimination tree gather and factor in
 * parallel; the solves run the same levels. A pattern that is not
 * structurally symmetric (MOSFET stamps) is planned on the symmetric closure
 * of its fill, whose extra entries stay exact zeros and change no result.
 * rs_tran_adjoint's own solvers follow the mode of its solver argument.
 * Takes effect at the next analysis. rs_supernodal_info: out[9] =
 * {active, supernodes, levels, contributions, panel doubles, device bytes,
 * multiply-subtracts per factor, widest supernode, the multiply-subtracts
 * of descendant contributions (the k_sn_gather launches)}. */
void rs_set_supernodal(rs_solver *h, int32_t mode);
int32_t rs_supernodal_info(const rs_solver *h, int64_t *out);
/* Checking aid (no reference counterpart): with on != 0, every FP64 device
 * buffer the library allocates afterwards starts filled with NaN, so a value
 * read before it is written changes the result visibly instead of reading
 * whatever the allocator's memory held. The environment variable
 * RSGPU_POISON=1 sets it for the whole process. */
void rs_set_poison(int32_t on);

/* ---- AC small-signal sweep (Engine::run_ac_result, engine.rs:829-970) ------
 * The reference solves one frequency after another with ComplexDenseSolver
 * (complex_solver.rs:50-136); here every frequency of the sweep is one
 * workgroup of one launch (chunked so the dense matrices stay within 512 MiB).
 * rs_ac_frequencies: generate_frequency_sweep (engine.rs:1002-1044), sweep 0
 * dec, 1 oct, 2 lin; returns the count (-1: unknown sweep), writes up to cap.
 * rs_ac_create: the devices the AC stamps read (stamp.rs:1525-1747: R, C, V,
 * I, diode, MOSFET; a V or I source whose value does not parse is still
 * stamped here), node_count as in rs_circuit_create, the BSIM model table
 * (sa and sb are taken as 0, as evaluate_mos passes them), and each device's
 * AC magnitude and phase in degrees (0 for none; read for V and I only).
 * rs_ac_run: the sweep at the DC solution x_dc (nx entries); x_out[f][i]
 * (cap_n complex entries per frequency, (re, im) pairs) and magph[f][i] =
 * (20 log10|v| or -600 below 1e-30, arg v in degrees) for the node_count
 * nodes (engine.rs:932-946); *n_out = the complex MNA size; *nsolved = the
 * frequencies before the first whose solve fails (the reference stops there
 * and reports Failed). rs_ac_timing: out[5] = {ms of the last sweep's solve
 * launches, launches, n, dense entries, contributions}. */
typedef struct rs_ac rs_ac;
int64_t rs_ac_frequencies(int32_t sweep, int64_t points, double fstart, double fstop, double *out,
                          int64_t cap);
rs_ac *rs_ac_create(const rs_device *dev, int64_t ndev, int64_t node_count,
                    const rs_bsim_model *models, int64_t nmodels, const double *ac_mag,
                    const double *ac_phase);
int32_t rs_ac_run(rs_ac *h, const double *x_dc, int64_t nx, const double *freqs, int64_t nfreq,
                  double *x_out, double *magph, int64_t cap_n, int64_t *n_out, int64_t *nsolved);
int32_t rs_ac_timing(const rs_ac *h, double *out);
int32_t rs_ac_last_error(const rs_ac *h, char *buf, int64_t len);
void rs_ac_destroy(rs_ac *h);

#ifdef __cplusplus
}
#endif
#endif
