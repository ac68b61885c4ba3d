// plan.hpp - flat device plan of one factorization (pattern, pivots,
// scatter map, level schedules) and its builder.
//
// Coordinates: a "position" g in [0, n) is block_start + pivot step inside a
// BTF diagonal block (the reference's per-block permuted index). L and U of
// all blocks live in one CSC over positions; off-diagonal BTF entries are
// kept unmerged, per row, in the reference's solve order
// (native_klu.rs:1501-1523).
#pragma once
#include <cstdint>
#include <vector>

#include "symbolic.hpp"

namespace rsgpu {

// Offsets (in elements) into a shared int32 pool and double pools. The same
// struct is read by the kernels, so it is plain data.
struct PlanDev {
    int32_t n, nblk, nflev, flags;
    int32_t nl, nu, nsc, nod;
    int32_t o_lcp, o_lri, o_ucp, o_uri, o_colflag, o_flev_ptr, o_flev_cols;
    int32_t o_sc_ptr, o_sc_src, o_sc_cnt, o_sc_dst;
    int32_t o_perm_in, o_perm_out;
    int32_t o_lrp, o_lrc, o_lrv, o_urp, o_urc, o_urv;
    int32_t o_blk_ptr, o_lsl_blk, o_lsl_ptr, o_lsl_rows, o_usl_blk, o_usl_ptr, o_usl_rows;
    int32_t o_odp, o_odax, o_odsrc, o_blk_has_od;
    int32_t o_lsl_nlong, o_usl_nlong;  // per solve level: its first rows are long (a wavefront each)
    // flat pipelined solve (sq_nlev == 0: none): every solve level in
    // execution order as int4 {s0, k0, div, 0} (nlev + 1 of them, the last a
    // sentinel) over entry slots s (column sq_col, value offset sq_src from
    // o_lx) and row records k (int4 {row, slot begin, slot end, 0})
    int32_t sq_nlev, sq_maxent, o_sq_lev, o_sq_row, o_sq_col, o_sq_src;
    // right-looking refactor schedule (rl_nph == 0: none); value indices are
    // relative to o_lx with lx | pad | ux | pad | ud contiguous
    int32_t rl_nph, rl_nv;
    int32_t o_rl_ini_src, o_rl_ini_cnt, o_rl_div_ptr, o_rl_div, o_rl_piv_ptr, o_rl_piv;
    int32_t o_rl_task_ptr, o_rl_task_nlong, o_rl_task_tgt, o_rl_task_beg, o_rl_pair;
    int32_t o_rl_sgl_ptr, o_rl_sgl, o_rl_piv_a;
    int64_t o_lx, o_ux, o_ud;  // value pool
    int64_t o_lxr, o_uxr;      // row-ordered copies of lx / ux (flags & PLAN_CSR)
    int64_t o_ax;              // ax pool (unmerged CSC values as handed in)
    int64_t o_vec;             // rhs/x pool (n entries)
    int64_t o_work;            // factor work vectors, one per wavefront (global path)
    int64_t o_tvec;            // solve vector (global path)
};

enum { COL_SINGLETON = 1 };
enum { PLAN_CSR = 1 };

// A solve row with at least this many off-diagonal entries is a long row.
constexpr int32_t kSolveLongRow = 32;

struct HostFactor {
    i64 n = 0;
    bool has_btf = false;
    bool sparselu = false;  // SparseLuSolver: no singleton rule for a 1x1 matrix
    Btf btf;
    std::vector<BlockSymbolic> sym;
    std::vector<BlockNumeric> num;
    i64 nnz_l = 0, nnz_u = 0, noffdiag = 0;
};

// Host-side plan arrays (int32) before upload.
struct PlanHost {
    int32_t n = 0, nblk = 0, nflev = 0, nl = 0, nu = 0, nsc = 0, nod = 0;
    std::vector<int32_t> lcp, lri, ucp, uri, colflag, flev_ptr, flev_cols;
    std::vector<int32_t> sc_ptr, sc_src, sc_cnt, sc_dst;
    std::vector<int32_t> perm_in, perm_out;
    std::vector<int32_t> lrp, lrc, lrv, urp, urc, urv;
    std::vector<int32_t> blk_ptr, lsl_blk, lsl_ptr, lsl_rows, usl_blk, usl_ptr, usl_rows;
    std::vector<int32_t> lsl_nlong, usl_nlong;
    std::vector<int32_t> odp, odax, odsrc, blk_has_od;
    // level statistics
    int32_t max_flev_width = 0, n_lsl = 0, n_usl = 0;
    // algorithmic work of one refactor / one solve (counted on the pattern)
    double refactor_flops = 0, solve_flops = 0;
};

void build_plan(const HostFactor& f, const i64* ap, const i64* ai, PlanHost& p);

// Level records of the pipelined solve (PlanDev::sq_*), appended to pool at
// 16-byte aligned offsets; none when a block has BTF off-diagonal entries.
void append_level_records(const PlanHost& p, std::vector<int32_t>& pool, PlanDev& d);

// The same records for the transposed solve J^T y = e of a single-block plan,
// in a pool of their own with the permutations swapped: U^T rows (the columns
// of U above the diagonal, on the refactor's column levels, divided by the
// diagonal), then L^T rows (the columns of L, on their own levels). d is the
// plan's descriptor with perm and record offsets into that pool.
bool build_transposed_records(const PlanHost& p, std::vector<int32_t>& pool, PlanDev& d);

// Levels of L^T (unit upper triangular) for the level-set transposed solve of
// a single-block plan: column k waits for every row of L(:, k). ptr holds
// nlev + 1 entries, cols the columns level by level.
void lt_levels(const PlanHost& p, int64_t n, std::vector<int32_t>& ptr, std::vector<int32_t>& cols);

// Gather L/U values of all blocks into position-space arrays (lx[nl], ux[nu],
// ud[n]) matching PlanHost's lri/uri order.
void gather_values(const HostFactor& f, std::vector<double>& lx, std::vector<double>& ux,
                   std::vector<double>& ud);

// Plan of a pattern factored with identity pivots (kind 0: NativeKLU with
// BTF, 1: SparseLU): its scatter targets, solve permutation and off-diagonal
// row keys are pre-pivot positions, to be mapped through the pivots found
// later (row_perm_inv per block, as positions).
void build_identity_plan(i64 n, const i64* ap, const i64* ai, int kind, PlanHost& ph);

// Append one plan's int arrays to a pool and fill the int offsets of d.
void append_plan(const PlanHost& p, std::vector<int32_t>& pool, PlanDev& d);

// Right-looking schedule of the static-pivot refactor. Pivot j runs in phase
// lev(j) = 1 + max lev over the columns it waits for (U(:,j) and the columns
// j' with L(j,j') != 0). Every update L(r,j)*U(j,g) of an entry (r,g) is
// applied in phase max(lev of the entry's pivots so far, in the column's U
// order), so each entry sees the reference's updates in the reference's order
// while the entries of one phase are updated in parallel. Value indices are
// relative to lx with lx | pad | ux | pad | ud contiguous.
constexpr int32_t kRlLongTask = 24;  // a task with this many updates takes a wavefront
// One-update records go to the device packed in 8 bytes: target, L and U
// value slots of kRlSlotBits each (the schedule is used only below 2^21 slots).
constexpr int32_t kRlSlotBits = 21;

struct RlHost {
    int32_t nph = 0, nv = 0, ntask = 0, npair = 0;
    std::vector<int32_t> ini_src, ini_cnt;  // per value slot: run of ax entries summed into it
    std::vector<int32_t> div_ptr, div;      // per phase: (L slot, diagonal slot) pairs
    std::vector<int32_t> piv_ptr, piv;      // per phase: pivot columns
    std::vector<int32_t> piv_a;             // per pivot: its diagonal slot, bit 31 a singleton column
    std::vector<int32_t> task_ptr, task_nlong, task_tgt, task_beg;  // long tasks first per phase
    std::vector<int32_t> pair;              // (L slot, U slot) pairs of each task, in order
    std::vector<int32_t> sgl_ptr, sgl;      // per phase: one-update tasks as (target, L, U, 0); packed on upload
    int32_t nsgl = 0;
};

// false when the plan has no such schedule (a column's U order that is not
// topological, or an update outside the column's pattern); the column kernel
// then stays in use.
bool build_rl_plan(const PlanHost& p, RlHost& r);
void append_rl(const RlHost& r, std::vector<int32_t>& pool, PlanDev& d);

}  // namespace rsgpu

// Host analysis and first factor with its plan (rs_host_factor_create).
struct rs_host_factor {
    rsgpu::HostFactor f;
    rsgpu::PlanHost p;
};
