// batch.hpp - many independent DC operating points in one persistent launch.
//
// Each instance (e.g. a Monte-Carlo sample of a circuit) runs the whole
// Engine::run_dc_result Newton loop (engine.rs:249-311, newton.rs:65-185) on
// one wavefront: stamping, the first factor with threshold pivoting
// (native_klu.rs:1194-1382), refactors with the reused pivots, solves, the
// convergence test and damping. Instances are pulled from a work queue in
// descending size order. Instances with the same DC sparsity pattern share
// one host analysis (BTF + AMD + L/U pattern), as one NativeKluSolver would
// reuse it for an unchanged pattern.
//
// Position space: block b of the BTF occupies positions [bs_b, be_b); inside
// it, position bs_b + k is pivot step k of that block. "pre" is the position
// a matrix row has before row pivoting (bs + col_perm[local row]); the
// instance's RPI[pre] is its current pivot position (row_perm_inv).
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/rs_circuit.h"

namespace rsgpu {

struct BPat {
    int32_t n, node_count, nnz, nblk;
    int32_t nl, nu, max_pad0, max_pad1;
    int32_t o_ax_src, o_rhs_ptr, o_rhs_src;
    int32_t o_sc_ptr, o_sc_src, o_sc_cnt, o_sc_pre;
    int32_t o_lcp, o_lri, o_ucp, o_uri, o_colflag;
    int32_t o_pin, o_pout, o_lrp, o_lrc, o_lrv, o_urp, o_urc, o_urv;
    int32_t o_blk_ptr, o_lsl_blk, o_lsl_ptr, o_lsl_rows, o_usl_blk, o_usl_ptr, o_usl_rows;
    int32_t o_odp, o_odax, o_odsrc, o_blk_has_od;
    // per-column update program: for each U(j,g), j ascending, every L(:,j)
    // entry as (j, row, index into L values)
    int32_t o_up_ptr, o_up_j, o_up_row, o_up_l;
    int32_t nrhs, o_sc_col;
    // step programs (sp.hpp): op-word offsets into the program pool, chunk counts
    int64_t o_spf, o_sps;
    int32_t spf_chunks, sps_chunks, sp_ok, t_base, od_base, nslots;
    // gather records (int4, 16-byte aligned): per scatter run and per RHS row
    // {count, slot0, slot1, slot2} when count <= 3, else {count, first, 0, 0}
    // with the row/run listed for the cooperative long-sum path
    int32_t o_screc, o_lrun, n_lrun, o_rhsrec, o_lrow, n_lrow;
};

struct BInst {
    int32_t pat, ndev;
    int64_t o_dev;  // first device record
    int64_t o_sv, o_ax, o_lx, o_ux, o_ud, o_x, o_rv;
    int64_t o_ss;  // scatter slots for the step-program refactor (int scratch)
};

struct BResult {
    int32_t converged, iterations, reason, full_factors;
    double final_norm;
};

struct BOpts {  // rs_newton_opts as kernel argument
    int32_t max_iters, gmin_steps, source_steps;
    int32_t pad;
    double abs_tol, rel_tol, gmin, damping, damping_min, pivot_tol;
};

}  // namespace rsgpu
