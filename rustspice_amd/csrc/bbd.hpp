// bbd.hpp - bordered-block-diagonal decomposition and the BbdSolver /
// DenseSolver back ends of rs_solver.
//
// Partition and decomposition follow crates/sim-core/src/bbd.rs (greedy BFS
// level-set bisection, promotion of cross-block nodes to the border); they are
// integer work and run on the host. Numerics follow bbd_solver.rs and
// solver.rs:595-724 and run on the GPU:
//   - every diagonal block is a SparseLU matrix (no pivoting), so its plan has
//     identity pivots and one k_refactor launch factors all blocks at once;
//   - Y = B_kk^-1 R_k(:, j) for all (k, j) pairs is one multi-RHS k_solve
//     launch;
//   - the Schur complement S = B_T - sum_k C_k Y is assembled entry-parallel,
//     each entry's updates in the reference's (block, column, entry) order;
//   - S is factored by a dense partial-pivoting LU (border <= 64) or by a
//     SparseLU plan (larger borders);
//   - the solve is the reference's block / border / back-substitution
//     sequence with the same zero skips.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "plan.hpp"

struct rs_solver;

namespace rsgpu {

constexpr int64_t kDenseSchurMax = 64;  // DENSE_SCHUR_THRESHOLD (bbd_solver.rs:85)

struct BbdDecomp {
    int64_t n = 0, num_blocks = 1, border_start = 0, border_size = 0;
    std::vector<int64_t> perm, inv_perm, block_ptr;  // perm[new] = old
};

// GreedyBisectionPartitioner::partition (bbd.rs:103-171); -1 marks border.
int64_t bbd_partition(int64_t n, const int64_t* ap, const int64_t* ai, int64_t num_blocks,
                      std::vector<int64_t>& assign);
// bbd_decompose (bbd.rs:355-485)
void bbd_decompose(int64_t n, const int64_t* ap, const int64_t* ai, int64_t num_blocks,
                   BbdDecomp& d);

struct BbdImpl;    // defined in bbd.cpp
struct DenseImpl;  // defined in bbd.cpp

BbdImpl* bbd_create();
void bbd_destroy(BbdImpl* b);
DenseImpl* dense_create();
void dense_destroy(DenseImpl* d);

// LinearSolver steps for rs_solver kinds RS_SOLVER_BBD and RS_SOLVER_DENSE.
// They report failures through the solver's error slot.
int32_t bbd_analyze(::rs_solver* h, const int64_t* ap, int64_t apl, const int64_t* ai, int64_t ail);
int32_t bbd_factor(::rs_solver* h, const int64_t* ap, int64_t apl, const int64_t* ai, int64_t ail,
                   const double* ax);
int32_t bbd_solve(::rs_solver* h, double* rhs);
void bbd_reset(::rs_solver* h);
void bbd_set_target(::rs_solver* h, int64_t num_blocks);
int32_t bbd_info(const ::rs_solver* h, int64_t* out);
// Dense Schur product on FP64 MFMA tiles: mode 0 off, 1 where the packed
// operands are dense enough, 2 always (dense Schur only).
void bbd_set_mfma(::rs_solver* h, int32_t mode);
int32_t bbd_mfma_info(const ::rs_solver* h, int64_t* out);
// The same on device-resident values: a SparseLU fallback factors and solves
// them on the device; a real decomposition copies them to the host first.
int32_t bbd_factor_dev(::rs_solver* h, const int64_t* ap, int64_t apl, const int64_t* ai, int64_t ail,
                       const double* ax_dev, hipStream_t s);
int32_t bbd_solve_dev(::rs_solver* h, const double* rhs_dev, double* x_dev, hipStream_t s);
// the supernodal mode handed to the SparseLU fallback (rs_set_supernodal)
void bbd_set_supernodal(::rs_solver* h, int mode);
// the SparseLU fallback solver, when the decomposition collapsed (else null)
const ::rs_solver* bbd_fallback_solver(const ::rs_solver* h);
int32_t dense_factor(::rs_solver* h, const int64_t* ap, int64_t apl, const int64_t* ai, int64_t ail,
                     const double* ax);
int32_t dense_solve(::rs_solver* h, double* rhs);

// kernels (bbd_kernels.hip)
hipError_t launch_gather(double* dst, const double* src, const int32_t* idx, int64_t m, hipStream_t s);
hipError_t launch_gather_sum(double* dst, const double* src, const int32_t* ptr, const int32_t* idx,
                             int64_t m, hipStream_t s);
hipError_t launch_bbd_pair_rhs(double* y, const int32_t* pair_off, const int32_t* pair_ptr,
                               const int32_t* ent_row, const double* ent_val, int64_t npairs,
                               hipStream_t s);
hipError_t launch_bbd_schur(double* out, const int32_t* cell_row, const int32_t* cell_col,
                            const double* base, const int32_t* crow_ptr, const int32_t* c_blk,
                            const int32_t* c_col, const double* c_val, const int32_t* pair_of,
                            const int32_t* pair_off, const double* y, int64_t bs, int64_t ncells,
                            hipStream_t s);
// S = base - Cp Yp on FP64 MFMA tiles (packing, split-K tiles, ordered sum
// of the chunks); kp and kc multiples of 16, nchunk * kc >= kp.
hipError_t launch_bbd_schur_mfma(double* out, const double* base, double* cp, double* yp,
                                 double* part, const double* ax, const int32_t* cptr,
                                 const int32_t* cidx, const int32_t* ymap, const double* y,
                                 int32_t bs, int64_t kp, int32_t nchunk, int64_t kc,
                                 hipStream_t s);
hipError_t launch_bbd_border_rhs(double* bb, const double* work, int64_t b0,
                                 const int32_t* crow_ptr, const int32_t* c_blk,
                                 const int32_t* c_col, const double* c_val,
                                 const int32_t* blk_start, const double* z, int64_t bs,
                                 hipStream_t s);
hipError_t launch_bbd_rprod(double* w, const int32_t* rrow_ptr, const int32_t* r_col,
                            const double* r_val, const double* xt, int64_t m, hipStream_t s);
hipError_t launch_permute(double* dst, const double* src, const int32_t* idx, int64_t n,
                          hipStream_t s);
hipError_t launch_bbd_combine(double* work, const double* z, const double* w, const double* xt,
                              int64_t b0, int64_t n, hipStream_t s);
hipError_t launch_dense_build(double* lu, const int32_t* ap, const int32_t* ai, const double* ax,
                              int64_t n, hipStream_t s);
hipError_t launch_dense_lu(double* lu, int32_t* piv, int32_t* status, int64_t n, int schur,
                           hipStream_t s);
hipError_t launch_dense_solve(const double* lu, const int32_t* piv, double* rhs, double* b,
                              int32_t* status, int64_t n, int schur, hipStream_t s);

}  // namespace rsgpu
