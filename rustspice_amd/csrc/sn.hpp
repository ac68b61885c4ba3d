// sn.hpp - supernodal factor and solve of one large SparseLU matrix spread
// over the whole GPU (the RC power grid's n = 1,048,833 TRAN matrix).
//
// The reference factor (sparse_lu.rs numeric_factor, restated in
// native_klu.rs:1194-1382 with pivot_tol 0) computes every entry (i, k) of
// L and U as
//     x = 0.0 + A(i, k)  (duplicates in CSC order)
//     x = x - L(i, j) * U(j, k)   for j ascending, skipped where U(j, k) == 0
//     U(k, k) = x clamped to +-1e-14,  L(i, k) = x / U(k, k)
// Any schedule that gives each entry the same terms in the same order
// reproduces it bit for bit. This one groups the pivot columns into
// fundamental supernodes J = [j0, j1) (consecutive columns whose L patterns
// nest: L(:, c) = {c+1 .. j1-1} u R_J) and keeps each supernode's values as
// dense panels:
//   LP_J: (w + h) x w, column-major: rows J then R_J, columns J (the
//         diagonal block holds L below and U on and above its diagonal);
//   UP_J: w x h, column-major: U(J, R_J).
// A pattern that is structurally symmetric (L(:, k) and U(k, :) hold the
// same positions, as an MNA matrix does) makes every U row of J span R_J.
//
// Fan-in order per supernode J, level by level on the supernodal
// elimination tree (J waits for every descendant):
//   1. gather: for each descendant J' whose R_J' meets J, in ascending order,
//      subtract L(rows, J') U(J', cols) from the entries of LP_J and UP_J it
//      touches (each entry's j' ascending: the J' ranges are disjoint and
//      increasing);
//   2. factor: right-looking over J's own columns t = 0..w-1 (pivot clamp,
//      L column divided by it, rank-one update of the later columns of LP_J
//      and the later rows of UP_J, skipping zero multipliers).
// Every entry thus receives its descendants' terms first (all smaller
// indices), then J's own, each in ascending order.
//
// The solves follow the reference's column sweeps (solve_block_internal):
// forward y[i] -= L(i, j) y[j] for j ascending; backward x[k] = y[k] / U(k,k),
// then y[i] -= U(i, k) x[k] for k descending. Fan-in per supernode again:
// forward, J's rows gather its descendants' columns (pairs ascending, then
// J's own columns); backward, root first, J's rows gather R_J's x (columns
// descending), then J's own columns descending.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "symbolic.hpp"

namespace rsgpu {

// Target tiles of the gather: kSnTile x kSnTile entries of LP_J or UP_J.
constexpr int32_t kSnTile = 64;
// An entry that this many contributions touch is summed on its own.
constexpr int32_t kSnHotRecords = 256;

// A gather record as the kernel reads it: the entries it touches in its
// tile as a set of rows and a set of columns (bit r of rowmask: tile row r;
// the k-th set bit is the record's k-th row, in order), and where its
// operands start: L(row k, s) = v[lbase + s * lds + k], U(s, column k) =
// v[ubase + k * ws + s], s in [0, ws).
struct SnRec {
    int64_t lbase, ubase;
    uint64_t rowmask, colmask;
    int32_t ws, lds;
};

// A stretch [q0, q1) of one long ordered sum (a hot entry's terms, a long
// forward row's terms), folded on a side stream once its terms are final.
struct SnPart {
    int64_t q0, q1;
    int32_t idx, pad;  // hot entry, or y position
};

struct SnHost {
    int32_t n = 0, ns = 0, nlev = 0;
    int64_t nvals = 0;               // doubles in the panel pool
    std::vector<int32_t> j0;         // ns + 1: supernode column ranges
    std::vector<int32_t> rptr, rows; // R_J (sorted) of each supernode
    std::vector<int64_t> oL, oU;     // panel offsets in the value pool
    std::vector<int32_t> parent;     // supernodal elimination tree (-1: root)
    std::vector<int32_t> lev_ptr, lev_sn;  // supernodes by level, ascending inside a level
    // contributions to each target J, ascending J': J' (src), the range
    // [a0, a1) of R_J' that falls in J, and rel[] (from rel_off) giving, for
    // every t in [a0, |R_J'|), the row of LP_J holding R_J'[t]
    std::vector<int32_t> cptr, csrc, ca0, ca1, crel;
    std::vector<int64_t> crel_off;
    std::vector<int32_t> rel;
    // gather tiles per level: {J, kind (0 LP, 1 UP), r0, c0} as int4, each
    // with its records (tile_rec[t] .. tile_rec[t+1]): the contributions that
    // touch the tile, ascending, as {pair, t0, t1, u0, u1}: the pair's rows
    // t in [t0, t1) of R_J' and columns u in [u0, u1) of R_J' that land in
    // the tile (LP: rows rel[t], columns R_J'[u] - j0; UP: rows R_J'[u] - j0,
    // columns rel[t] - w)
    std::vector<int32_t> glev_ptr, gitems;
    std::vector<int64_t> tile_rec;
    std::vector<int32_t> rec;  // 5 per record
    // entries with at least kSnHotRecords contributions (the ground node's
    // diagonal): taken out of their tiles and summed by a wavefront each from
    // a flat list of (L slot, U slot) terms in the reference's order
    std::vector<int64_t> hot_slot, hot_ptr, hot_terms;  // hot_terms: 2 per term
    std::vector<int32_t> hot_lev;                       // level of each hot entry
    std::vector<int32_t> hot_tlev;                      // level of each term's source supernode
    // the records in the kernel's form (rec's order), and per gather item
    // the hot entries inside its tile as row | column << 8
    std::vector<SnRec> frec;
    std::vector<int32_t> tile_hot_ptr, tile_hot;
    // per gather item and quarter of its columns (a wavefront's share), the
    // records that touch the quarter, in order: qrec[qrec_ptr[4 it + q] ..]
    std::vector<int64_t> qrec_ptr;
    std::vector<int32_t> qrec;
    // forward-solve rows: for every position p, the supernodes S (ascending)
    // with p in R_S and p's index t in R_S
    std::vector<int64_t> lrow_ptr;
    std::vector<int32_t> lrow_sn, lrow_t;
    // scatter of A: every run of duplicate CSC entries -> panel slot
    std::vector<int32_t> sc_src, sc_cnt;
    std::vector<int64_t> sc_slot;
    // position -> original index (the solve's permutations)
    std::vector<int32_t> perm;       // perm[original] = position
    // statistics
    bool symmetric = true;           // false: the plan holds the pattern's symmetric closure
    int64_t npairs = 0, max_w = 0, max_h = 0;
    double updates = 0.0;            // multiply-subtracts of one factor (pattern count)
    double gather_updates = 0.0;     // the part applied by descendant contributions
};

// Supernodes at least this wide are factored by the blocked multi-workgroup
// path (panel of kSnPanel columns, U block row, trailing update tiles).
constexpr int32_t kSnBigW = 16;
constexpr int32_t kSnPanel = 32;

// Device view of a plan (plain pointers, passed by value to the kernels).
struct SnPtrs {
    const int32_t *j0, *rptr, *rows, *csrc, *ca0, *ca1, *rel, *perm, *lev_sn;
    const SnRec* frec;                // tile records
    const int32_t *tile_hot_ptr, *tile_hot;
    const int64_t* qrec_ptr;
    const int32_t* qrec;
    const int4* gitems;
    const int64_t *oL, *oU, *crel_off, *tile_rec, *hot_slot;
    const int64_t *lrow_ptr;          // forward-solve row lists: per position
    const int32_t *lrow_sn, *lrow_t;  // (S, t): L(p, S) = LP_S row w_S + t
    int32_t nhot, n;
    double* v;   // panels
    double* ud;  // pivots (n)
};

// Builds the supernodal plan of a single-block SparseLU symbolic factor
// (positions = pivot steps, identity row pivots). An unsymmetric pattern is
// planned on its symmetric closure (explicit zeros that change no value).
// transpose: the plan factors A^T from A's values (the same closure and
// order; the scatter swaps every entry's row and column). Returns false,
// with a reason, when the pattern exceeds the 32-bit index ranges.
bool build_sn_plan(const BlockSymbolic& s, const i64* ap, const i64* ai, SnHost& h,
                   std::string* why = nullptr, int32_t hot_min = kSnHotRecords, bool transpose = false);

// Host execution of the plan in the device kernels' order: factor of ax,
// then the solve of rhs into x; lx/ux/ud (optional) receive the factor in
// the symbolic's CSC order. For checking the schedule against the reference
// arithmetic on the CPU.
void sn_host_factor(const SnHost& h, const double* ax, std::vector<double>& vals,
                    std::vector<double>& ud);
void sn_host_solve(const SnHost& h, const std::vector<double>& vals, const std::vector<double>& ud,
                   const double* rhs, double* x);
void sn_to_csc(const SnHost& h, const BlockSymbolic& s, const std::vector<double>& vals,
               std::vector<double>& lx, std::vector<double>& ux);

}  // namespace rsgpu

namespace rsgpu {

// Device state of one supernodal plan: the plan arrays, the panels, the
// pivots and the solve vector, with the per-level launch lists.
struct SnImpl;
SnImpl* sn_create(const BlockSymbolic& s, const i64* ap, const i64* ai, std::string* why,
                  bool transpose = false);
void sn_destroy(SnImpl* p);
// factor of the unmerged CSC values ax_dev (device memory, nnz entries)
hipError_t sn_factor(SnImpl* p, const double* ax_dev, hipStream_t s);
// x = A^-1 rhs (device vectors of n entries; rhs and x may alias). y: a
// device vector of n entries for the permuted solution in place of the
// handle's own (nullptr: the handle's).
hipError_t sn_solve(SnImpl* p, const double* rhs_dev, double* x_dev, hipStream_t s, double* y = nullptr);
// factor values in the symbolic's CSC order (host arrays)
hipError_t sn_values(SnImpl* p, double* lx, double* ux, double* ud);
// pivots (device, n entries)
const double* sn_pivots(const SnImpl* p);
const SnHost& sn_plan(const SnImpl* p);
const BlockSymbolic& sn_symbolic(const SnImpl* p);
int64_t sn_device_bytes(const SnImpl* p);

}  // namespace rsgpu
