// symbolic.hpp - host-side integer analysis for the GPU solver.
//
// Everything here is integer work that runs once per sparsity pattern and
// must reproduce the reference's orderings exactly:
//   amd_order      - crates/sim-core/src/amd.rs:143-411
//   btf_decompose  - crates/sim-core/src/btf.rs:143-550
//   block_symbolic - crates/sim-core/src/native_klu.rs:863-1014
// plus the first (pivoting) numeric factor of native_klu.rs:1194-1382, whose
// pivot choices are sequential and data dependent, so it stays on the host.
#pragma once
#include <cstdint>
#include <vector>

namespace rsgpu {

using i64 = int64_t;

struct AmdResult {
    std::vector<i64> perm;      // perm[old] = new
    std::vector<i64> inv_perm;  // inv_perm[new] = old
};
AmdResult amd_order(i64 n, const i64* ap, const i64* ai);

struct Btf {
    std::vector<i64> row_perm, col_perm, row_perm_inv, col_perm_inv, block_ptr;
    i64 num_blocks = 0, structural_rank = 0, num_singletons = 0, max_block_size = 0;
};
Btf btf_decompose(i64 n, const i64* ap, const i64* ai);

// L/U pattern of one diagonal block in its symmetric AMD order.
struct BlockSymbolic {
    i64 n = 0;
    std::vector<i64> col_perm, col_perm_inv;
    std::vector<i64> lcp, lri;  // L: strictly lower, rows ascending
    std::vector<i64> ucp, uri;  // U: rows ascending, diagonal last
};
void block_symbolic(i64 n, const i64* ap, const i64* ai, bool natural, BlockSymbolic& s,
                    bool zero_diag_last = false);

struct BlockNumeric {
    std::vector<double> lx, ux, udiag;
    std::vector<i64> row_perm, row_perm_inv;
};
// Reference factor_block (native_klu.rs:1194-1382) on one block given as CSC.
// refactor=true reuses old's pivots (no pivot search). sparselu: SparseLuSolver
// semantics for a 1x1 matrix (plain 1e-14 clamp, no singleton rule).
void factor_block_host(const BlockSymbolic& s, const i64* ap, const i64* ai, const double* ax,
                       const BlockNumeric* old, bool refactor, double pivot_tol,
                       BlockNumeric& out, bool sparselu = false);

}  // namespace rsgpu
