// sn_kernels.hip - supernodal factor and solve on the device (see sn.hpp for
// the schedule and why it reproduces the reference's arithmetic).
//
// Panels: LP(r, c) = v[oL[J] + c * (w + h) + r], UP(r, c) = v[oU[J] + c * w + r].
// Every entry is updated by a single thread at a time, in the reference's
// order: a thread runs an entry's terms in ascending order and a workgroup
// barrier separates two records (or two pivot steps) that may touch the same
// entry. Built without FMA contraction: every term is a rounded product and
// a rounded difference, as in the reference.
#include <hip/hip_runtime.h>

#include "fold.hpp"
#include "sn.hpp"

namespace {

using rsgpu::SnPtrs;
constexpr int kT = 256;
constexpr int kTile = rsgpu::kSnTile;
constexpr int kPanel = rsgpu::kSnPanel;

__device__ __forceinline__ double clamp_pivot(double d) {
    // SparseLU: |d| < 1e-14 -> +-1e-14 (sign of d, +0 and -0 give +1e-14)
    if (fabs(d) < 1e-14) d = d >= 0.0 ? 1e-14 : -1e-14;
    return d;
}

__global__ __launch_bounds__(kT) void k_sn_zero(double* __restrict__ v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT) v[i] = 0.0;
}

// one thread per run of duplicate CSC entries: 0.0 + the run in CSC order
__global__ __launch_bounds__(kT) void k_sn_scatter(const int32_t* __restrict__ src, const int32_t* __restrict__ cnt,
                                                   const int64_t* __restrict__ slot, int64_t nruns,
                                                   const double* __restrict__ ax, double* __restrict__ v) {
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= nruns) return;
    double x = 0.0;
    const int32_t s = src[i], c = cnt[i];
    for (int32_t q = 0; q < c; q++) x += ax[s + q];
    v[slot[i]] = x;
}

// One workgroup per gather tile, no barrier: wavefront w owns the tile's
// columns 16w .. 16w + 15 and lane r its row r, so the workgroup's 256 lanes
// hold the 64 x 64 tile in registers, 16 entries each. Each wavefront runs
// through the records that touch its columns, in order, and applies the
// terms of its own entries: an entry therefore takes its records in order
// and, inside a record, its terms s = 0 .. ws - 1 in order, skipping zero
// multipliers, as the reference does. The record's row and column masks give
// each lane its L row and each column its U column by rank.
//
// The work is a stream of chunks (a record's terms s0 .. s0 + 7): the record
// fields of 64 records at a time sit one per lane (read out with readlane),
// and the operands of the next chunk are loaded while the current one is
// applied, L one value per lane and term, U (the same for every row) one
// value per lane for 8 columns x 8 terms.
constexpr int kCols = kTile / 4;  // columns per wavefront
constexpr int kSb = 8;            // terms per chunk

__device__ __forceinline__ uint64_t bits_below(int b) { return b == 0 ? 0ull : (~0ull >> (64 - b)); }

__device__ __forceinline__ int64_t rl64(int64_t v, int k) {
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)v, k);
    const int hi = __builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), k);
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ double rlf64(double v, int k) {
    return __longlong_as_double((long long)rl64((int64_t)__double_as_longlong(v), k));
}

struct GRec {  // one record, wavefront-uniform
    int64_t lbase, ubase;
    uint64_t rowmask;
    uint32_t cm;  // this wavefront's 16 columns of colmask
    int crank0, ws, lds;
};

struct GOps {  // one chunk's operands
    double l[kSb];
    double u0, u1;  // U(s0 + (lane & 7), column (lane >> 3) [+ 8])
};

__device__ __forceinline__ void gather_load(const double* __restrict__ v, const GRec& R, int s0, int lane,
                                            uint64_t below, GOps& o) {
    const bool mine = (R.rowmask >> lane) & 1;
    const double* L = v + R.lbase + __popcll(R.rowmask & below);
#pragma unroll
    for (int j = 0; j < kSb; j++) o.l[j] = (mine && s0 + j < R.ws) ? L[(int64_t)(s0 + j) * R.lds] : 0.0;
    const int s = s0 + (lane & 7), c = lane >> 3;
    const bool a0 = ((R.cm >> c) & 1) && s < R.ws, a1 = ((R.cm >> (c + 8)) & 1) && s < R.ws;
    const int k0 = R.crank0 + __popc(R.cm & ((1u << c) - 1u));
    const int k1 = R.crank0 + __popc(R.cm & ((1u << (c + 8)) - 1u));
    o.u0 = a0 ? v[R.ubase + (int64_t)k0 * R.ws + s] : 0.0;
    o.u1 = a1 ? v[R.ubase + (int64_t)k1 * R.ws + s] : 0.0;
}

__device__ __forceinline__ void gather_apply(const GRec& R, int s0, const GOps& o, bool mine, uint32_t skip,
                                             double (&acc)[kCols]) {
    const int sb = min(kSb, R.ws - s0);
#pragma unroll
    for (int k = 0; k < kCols; k++) {
        if (!((R.cm >> k) & 1)) continue;
        const bool upd = mine && !((skip >> k) & 1);
#pragma unroll
        for (int j = 0; j < kSb; j++) {
            if (j < sb) {
                const double uu = rlf64(k < 8 ? o.u0 : o.u1, (k & 7) * 8 + j);
                if (upd && uu != 0.0) acc[k] = acc[k] - o.l[j] * uu;
            }
        }
    }
}

template <int kDepth>
__device__ __forceinline__ void gather_tile(const SnPtrs& p, int it) {
    const int4 g = p.gitems[it];
    const int J = g.x, kind = g.y, r0 = g.z, c0 = g.w;
    const int jj0 = p.j0[J], w = p.j0[J + 1] - jj0, hh = p.rptr[J + 1] - p.rptr[J];
    const int W = w + hh;
    const int nr = kind == 0 ? min(kTile, W - r0) : min(kTile, w - r0);
    const int nc = kind == 0 ? min(kTile, w - c0) : min(kTile, hh - c0);
    const int ld = kind == 0 ? W : w;
    double* base = p.v + (kind == 0 ? p.oL[J] : p.oU[J]) + (int64_t)c0 * ld + r0;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int cb = wv * kCols;
    double acc[kCols];
    // entries summed on their own elsewhere (the ground node's diagonal),
    // which this kernel neither reads nor writes: another stream folds them
    uint32_t skip = 0;
    for (int q = p.tile_hot_ptr[it]; q < p.tile_hot_ptr[it + 1]; q++) {
        const int e = p.tile_hot[q], hr = e & 255, hc = e >> 8;
        if (hr == lane && hc >= cb && hc < cb + kCols) skip |= 1u << (hc - cb);
    }
#pragma unroll
    for (int k = 0; k < kCols; k++)
        acc[k] = (lane < nr && cb + k < nc && !((skip >> k) & 1)) ? base[(int64_t)(cb + k) * ld + lane] : 0.0;
    const uint64_t below = bits_below(lane), cbelow = bits_below(cb);
    const int64_t qb = p.qrec_ptr[4 * it + wv], qe = p.qrec_ptr[4 * it + wv + 1];
    for (int64_t b0 = qb; b0 < qe; b0 += 64) {
        // fields of up to 64 records, one per lane
        const int nb = (int)min((int64_t)64, qe - b0);
        int64_t f_lb = 0, f_ub = 0;
        uint64_t f_rm = 0, f_cm = 0;
        int f_ws = 0, f_lds = 0;
        if (lane < nb) {
            const rsgpu::SnRec& r = p.frec[p.qrec[b0 + lane]];
            f_lb = r.lbase;
            f_ub = r.ubase;
            f_rm = r.rowmask;
            f_cm = r.colmask;
            f_ws = r.ws;
            f_lds = r.lds;
        }
        auto rec = [&](int k) {
            GRec R;
            R.lbase = rl64(f_lb, k);
            R.ubase = rl64(f_ub, k);
            R.rowmask = (uint64_t)rl64((int64_t)f_rm, k);
            const uint64_t cmk = (uint64_t)rl64((int64_t)f_cm, k);
            R.cm = (uint32_t)(cmk >> cb) & 0xFFFFu;
            R.crank0 = __popcll(cmk & cbelow);
            R.ws = __builtin_amdgcn_readlane(f_ws, k);
            R.lds = __builtin_amdgcn_readlane(f_lds, k);
            return R;
        };
        if constexpr (kDepth <= 2) {
            // the next chunk's operands load while the current one is applied
            GRec cur = rec(0);
            GOps oc, on;
            gather_load(p.v, cur, 0, lane, below, oc);
            int k = 0, s0 = 0;
            while (true) {
                // the next chunk: the rest of this record, else the next record
                GRec nxt = cur;
                int nk = k, ns0 = s0 + kSb;
                const bool more = ns0 < cur.ws;
                if (!more) {
                    nk = k + 1;
                    ns0 = 0;
                    if (nk < nb) nxt = rec(nk);
                }
                const bool have = more || nk < nb;
                if (have) gather_load(p.v, nxt, ns0, lane, below, on);
                gather_apply(cur, s0, oc, (cur.rowmask >> lane) & 1, skip, acc);
                if (!have) break;
                cur = nxt;
                k = nk;
                s0 = ns0;
                oc = on;
            }
        } else {
            // a ring of kDepth chunks in flight: slot i holds chunk kDepth *
            // round + i; a chunk is applied, then its slot is refilled with
            // the chunk kDepth ahead, so the loads of the next kDepth - 1
            // chunks overlap each apply (for chains of short records, where
            // one chunk ahead leaves the load latency exposed)
            int lk = 0, ls0 = 0;  // the last chunk issued: record lk, terms ls0 ..
            GRec lr = rec(0);
            auto advance = [&]() -> bool {
                if (lk >= nb) return false;
                ls0 += kSb;
                if (ls0 < lr.ws) return true;
                lk++;
                ls0 = 0;
                if (lk >= nb) return false;
                lr = rec(lk);
                return true;
            };
            GRec rq[kDepth];
            GOps oq[kDepth];
            int sq[kDepth];
            bool vq[kDepth];
#pragma unroll
            for (int i = 0; i < kDepth; i++) {
                vq[i] = i == 0 ? true : advance();
                if (vq[i]) {
                    rq[i] = lr;
                    sq[i] = ls0;
                    gather_load(p.v, rq[i], sq[i], lane, below, oq[i]);
                }
            }
            bool go = true;
            while (go) {
#pragma unroll
                for (int i = 0; i < kDepth; i++) {
                    if (!vq[i]) {
                        go = false;
                        break;
                    }
                    gather_apply(rq[i], sq[i], oq[i], (rq[i].rowmask >> lane) & 1, skip, acc);
                    vq[i] = advance();
                    if (vq[i]) {
                        rq[i] = lr;
                        sq[i] = ls0;
                        gather_load(p.v, rq[i], sq[i], lane, below, oq[i]);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kCols; k++)
        if (lane < nr && cb + k < nc && !((skip >> k) & 1)) base[(int64_t)(cb + k) * ld + lane] = acc[k];
}

__global__ __launch_bounds__(kT) void k_sn_gather(SnPtrs p, int item0) { gather_tile<2>(p, item0 + blockIdx.x); }
// levels with few tiles (the top of the tree): long chains of short records
// on few wavefronts, so occupancy matters less than chunks in flight
__global__ __launch_bounds__(kT) void k_sn_gather_deep(SnPtrs p, int item0) {
    gather_tile<4>(p, item0 + blockIdx.x);
}

// ---- ordered long sums (fold_long, fold.hpp) ------------------------------
using rsgpu::fold_long;

// a run of many duplicate CSC entries (the ground diagonal): 0.0 + the run
__global__ __launch_bounds__(kT) void k_sn_scatter_long(const int32_t* __restrict__ src,
                                                        const int32_t* __restrict__ cnt,
                                                        const int64_t* __restrict__ slot,
                                                        const double* __restrict__ ax, double* __restrict__ v) {
    const int r = blockIdx.x;
    const int64_t s0 = src[r];
    const double x = fold_long<false>(0.0, s0, s0 + cnt[r], [&](int64_t q) { return ax[q]; });
    if (threadIdx.x == 0) v[slot[r]] = x;
}

// an entry summed on its own: x - L * U over a stretch [q0, q1) of its flat
// term list, continuing from the value the previous stretch left; zero
// multipliers skipped (a skipped term is +0.0, which leaves x unchanged)
__global__ __launch_bounds__(kT) void k_sn_hot_parts(SnPtrs p, const rsgpu::SnPart* __restrict__ parts,
                                                     const int64_t* __restrict__ terms) {
    const rsgpu::SnPart pt = parts[blockIdx.x];
    double* v = p.v;
    const int64_t slot = p.hot_slot[pt.idx];
    const double x = fold_long<true>(v[slot], pt.q0, pt.q1, [&](int64_t q) {
        const double u = v[terms[2 * q + 1]];
        return u != 0.0 ? v[terms[2 * q]] * u : 0.0;
    });
    if (threadIdx.x == 0) v[slot] = x;
}

// a stretch [q0, q1) of a long forward row's terms: y[row] - L(row, j) y[j]
__global__ __launch_bounds__(kT) void k_sn_fwd_parts(SnPtrs p, const rsgpu::SnPart* __restrict__ parts,
                                                     const int64_t* __restrict__ lslot,
                                                     const int32_t* __restrict__ yidx, double* __restrict__ y) {
    const rsgpu::SnPart pt = parts[blockIdx.x];
    const double* v = p.v;
    const double x = fold_long<true>(y[pt.idx], pt.q0, pt.q1, [&](int64_t q) { return v[lslot[q]] * y[yidx[q]]; });
    if (threadIdx.x == 0) y[pt.idx] = x;
}

// forward solve of a row with a long list: y[p] - L(p, j) y[j], j ascending
__global__ __launch_bounds__(kT) void k_sn_fwd_long(SnPtrs p, const int32_t* __restrict__ rows,
                                                    const int64_t* __restrict__ tptr,
                                                    const int64_t* __restrict__ lslot,
                                                    const int32_t* __restrict__ yidx, double* __restrict__ y) {
    const int r = blockIdx.x;
    const double* v = p.v;
    const double x = fold_long<true>(y[rows[r]], tptr[r], tptr[r + 1],
                                     [&](int64_t q) { return v[lslot[q]] * y[yidx[q]]; });
    if (threadIdx.x == 0) y[rows[r]] = x;
}

// One workgroup per supernode narrower than kSnBigW: right-looking over its
// columns (pivot clamp, L column, rank-one update of later LP columns and
// later UP rows), a barrier between phases.
__global__ __launch_bounds__(kT) void k_sn_factor(SnPtrs p, const int32_t* __restrict__ list) {
    __shared__ double sd;
    const int J = list[blockIdx.x];
    const int jj0 = p.j0[J], w = p.j0[J + 1] - jj0, hh = p.rptr[J + 1] - p.rptr[J];
    const int W = w + hh;
    double* LP = p.v + p.oL[J];
    double* UP = p.v + p.oU[J];
    for (int t = 0; t < w; t++) {
        if (threadIdx.x == 0) {
            const double d = clamp_pivot(LP[(int64_t)t * W + t]);
            LP[(int64_t)t * W + t] = d;
            p.ud[jj0 + t] = d;
            sd = d;
        }
        __syncthreads();
        const double d = sd;
        for (int r = t + 1 + threadIdx.x; r < W; r += kT) LP[(int64_t)t * W + r] = LP[(int64_t)t * W + r] / d;
        __syncthreads();
        const int m = W - t - 1, mw = w - t - 1;
        for (int e = threadIdx.x; e < m * mw; e += kT) {
            const int r = t + 1 + e % m, c = t + 1 + e / m;
            const double u = LP[(int64_t)c * W + t];
            if (u != 0.0) LP[(int64_t)c * W + r] -= LP[(int64_t)t * W + r] * u;
        }
        for (int e = threadIdx.x; e < mw * hh; e += kT) {
            const int r = t + 1 + e % mw, c = e / mw;
            const double u = UP[(int64_t)c * w + t];
            if (u != 0.0) UP[(int64_t)c * w + r] -= LP[(int64_t)t * W + r] * u;
        }
        __syncthreads();
    }
}

// ---- wide supernodes: blocked right-looking over panels of kPanel columns --
// (i + ii) one launch per panel step. Every workgroup loads the panel's
// diagonal block D = LP([t0, t1), [t0, t1)) into LDS and factors it there
// (the same operations in every workgroup, so the same bits); workgroup 0
// writes D and the pivots back. Then each workgroup takes a chunk of
//  - rows r >= t1 of the panel's columns (one thread per row): for t in order,
//    L(r, t) = x[t] / d_t, then x[c] -= L(r, t) U(t, c) for the later panel
//    columns c, skipping zero U(t, c) - the terms each entry receives in the
//    column-by-column factor, in the same order;
//  - columns right of the panel in LP and every UP column (one thread per
//    column): for t in order, x[r] -= L(r, t) x[t] for the later panel rows r,
//    skipping zero x[t].
constexpr int kRowsPerBlock = kT;

__device__ void panel_diag(double* D, const double* LP, int W, int t0, int nb) {
    // D[c * kPanel + r], factored right-looking with the pivot clamp
    for (int e = threadIdx.x; e < kPanel * kPanel; e += kT) {
        const int r = e % kPanel, c = e / kPanel;
        D[e] = (r < nb && c < nb) ? LP[(int64_t)(t0 + c) * W + t0 + r] : 0.0;
    }
    __syncthreads();
    for (int t = 0; t < nb; t++) {
        if (threadIdx.x == 0) D[t * kPanel + t] = clamp_pivot(D[t * kPanel + t]);
        __syncthreads();
        const double d = D[t * kPanel + t];
        if (threadIdx.x > t && threadIdx.x < nb) D[t * kPanel + threadIdx.x] = D[t * kPanel + threadIdx.x] / d;
        __syncthreads();
        const int m = nb - t - 1;
        for (int e = threadIdx.x; e < m * m; e += kT) {
            const int r = t + 1 + e % m, c = t + 1 + e / m;
            const double u = D[c * kPanel + t];
            if (u != 0.0) D[c * kPanel + r] -= D[t * kPanel + r] * u;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kT) void k_sn_panel(SnPtrs p, const int32_t* __restrict__ list, int t0) {
    __shared__ double D[kPanel * kPanel];
    const int J = list[blockIdx.y];
    const int jj0 = p.j0[J], w = p.j0[J + 1] - jj0, hh = p.rptr[J + 1] - p.rptr[J];
    const int W = w + hh, t1 = min(t0 + kPanel, w), nb = t1 - t0;
    const int nrow = W - t1, ncol = (w - t1) + hh;
    const int nrb = (nrow + kRowsPerBlock - 1) / kRowsPerBlock, ncb = (ncol + kT - 1) / kT;
    if ((int)blockIdx.x >= max(nrb + ncb, 1)) return;  // the launch covers the widest supernode
    double* LP = p.v + p.oL[J];
    panel_diag(D, LP, W, t0, nb);  // D stays unwritten here: other workgroups still read it
    const int b = blockIdx.x;
    if (b < nrb) {  // rows below the diagonal block
        const int r = t1 + b * kRowsPerBlock + threadIdx.x;
        if (r >= W) return;
        double x[kPanel];
#pragma unroll
        for (int c = 0; c < kPanel; c++) x[c] = c < nb ? LP[(int64_t)(t0 + c) * W + r] : 0.0;
#pragma unroll
        for (int t = 0; t < kPanel; t++) {
            if (t < nb) {
                const double l = x[t] / D[t * kPanel + t];
                x[t] = l;
#pragma unroll
                for (int c = t + 1; c < kPanel; c++) {
                    const double u = D[c * kPanel + t];
                    if (c < nb && u != 0.0) x[c] -= l * u;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < kPanel; c++)
            if (c < nb) LP[(int64_t)(t0 + c) * W + r] = x[c];
    } else if (b < nrb + ncb) {  // columns right of the panel (U block row)
        const int k = (b - nrb) * kT + threadIdx.x;
        if (k >= ncol) return;
        const int nlp = w - t1;
        double* U = k < nlp ? LP + (int64_t)(t1 + k) * W : p.v + p.oU[J] + (int64_t)(k - nlp) * w;
        double x[kPanel];
#pragma unroll
        for (int r = 0; r < kPanel; r++) x[r] = r < nb ? U[t0 + r] : 0.0;
#pragma unroll
        for (int t = 0; t < kPanel; t++) {
            const double u = x[t];
            if (t >= nb || u == 0.0) continue;
#pragma unroll
            for (int r = t + 1; r < kPanel; r++)
                if (r < nb) x[r] -= D[t * kPanel + r] * u;
        }
#pragma unroll
        for (int r = 0; r < kPanel; r++)
            if (r < nb) U[t0 + r] = x[r];
    }
}

// the factored diagonal block and its pivots, written after every workgroup
// of the panel step has read the original block (by one extra workgroup of
// the trailing update, which does not read the diagonal block)
__device__ void panel_diag_store(double* D, SnPtrs& p, double* LP, int W, int jj0, int t0, int nb) {
    panel_diag(D, LP, W, t0, nb);
    for (int e = threadIdx.x; e < nb * nb; e += kT) {
        const int r = e % nb, c = e / nb;
        LP[(int64_t)(t0 + c) * W + t0 + r] = D[c * kPanel + r];
    }
    if (threadIdx.x < nb) p.ud[jj0 + t0 + threadIdx.x] = D[threadIdx.x * kPanel + threadIdx.x];
}

// (iii) trailing update of the tiles right of and below the panel: LP rows
// [t1, W) x columns [t1, w) and UP rows [t1, w) x columns [0, h); the panel's
// L rows and U columns are staged in LDS, each thread owns 16 entries and
// subtracts their kPanel terms in order.
__global__ __launch_bounds__(kT) void k_sn_trail(SnPtrs p, const int32_t* __restrict__ list, int t0) {
    __shared__ double A[kTile * kPanel];  // A[t * kTile + r]
    __shared__ double B[kTile * kPanel];  // B[c * kPanel + t]
    const int J = list[blockIdx.y];
    const int jj0 = p.j0[J], w = p.j0[J + 1] - jj0, hh = p.rptr[J + 1] - p.rptr[J];
    const int W = w + hh, t1 = min(t0 + kPanel, w), nb = t1 - t0;
    const int nrl = (W - t1 + kTile - 1) / kTile, ncl = (w - t1 + kTile - 1) / kTile;
    const int nru = (w - t1 + kTile - 1) / kTile, ncu = (hh + kTile - 1) / kTile;
    const int ntiles = nrl * ncl + nru * ncu;
    if ((int)blockIdx.x == ntiles) {  // the extra workgroup: the diagonal block and pivots
        panel_diag_store(A, p, p.v + p.oL[J], W, jj0, t0, nb);
        return;
    }
    if ((int)blockIdx.x > ntiles) return;  // the launch covers the widest supernode
    int b = blockIdx.x, kind, r0, c0;
    if (b < nrl * ncl) {
        kind = 0; r0 = t1 + (b % nrl) * kTile; c0 = t1 + (b / nrl) * kTile;
    } else {
        b -= nrl * ncl;
        kind = 1; r0 = t1 + (b % nru) * kTile; c0 = (b / nru) * kTile;
    }
    const double* LP = p.v + p.oL[J];
    const int rmax = kind == 0 ? W : w, cmax = kind == 0 ? w : hh;
    const int nr = min(kTile, rmax - r0), nc = min(kTile, cmax - c0);
    double* C = kind == 0 ? p.v + p.oL[J] : p.v + p.oU[J];
    const int ldc = kind == 0 ? W : w;
    for (int e = threadIdx.x; e < kTile * kPanel; e += kT) {
        const int r = e % kTile, t = e / kTile;
        A[e] = (r < nr && t < nb) ? LP[(int64_t)(t0 + t) * W + r0 + r] : 0.0;
        const int c = e / kPanel, tt = e % kPanel;
        B[e] = (c < nc && tt < nb) ? C[(int64_t)(c0 + c) * ldc + t0 + tt] : 0.0;
    }
    __syncthreads();
    // lane r holds row r of the wavefront's 16 columns in registers; the
    // panel's terms t go in order, each applied to all 16 entries (16
    // independent chains, one LDS read of L per term)
    const int r = threadIdx.x & (kTile - 1), cg = threadIdx.x / kTile;
    if (r >= nr) return;
    constexpr int kC = kTile / 4;
    const int cb = cg * kC;
    double x[kC];
#pragma unroll
    for (int m = 0; m < kC; m++) x[m] = cb + m < nc ? C[(int64_t)(c0 + cb + m) * ldc + r0 + r] : 0.0;
    for (int t = 0; t < nb; t++) {
        const double a = A[t * kTile + r];
#pragma unroll
        for (int m = 0; m < kC; m++) {
            const double u = B[(cb + m) * kPanel + t];  // zero beyond nc: skipped
            if (u != 0.0) x[m] -= a * u;
        }
    }
#pragma unroll
    for (int m = 0; m < kC; m++)
        if (cb + m < nc) C[(int64_t)(c0 + cb + m) * ldc + r0 + r] = x[m];
}

// ---- solve ----------------------------------------------------------------
__global__ __launch_bounds__(kT) void k_sn_perm_in(const int32_t* __restrict__ perm, const double* __restrict__ b,
                                                   double* __restrict__ y, int n) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i < n) y[perm[i]] = b[i];
}
__global__ __launch_bounds__(kT) void k_sn_perm_out(const int32_t* __restrict__ perm, const double* __restrict__ y,
                                                    double* __restrict__ x, int n) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i < n) x[i] = y[perm[i]];
}

// forward, rows of one level: y[p] -= L(p, j) y[j] over the descendants'
// columns in ascending order (the row list of p), one thread per row
__global__ __launch_bounds__(kT) void k_sn_fwd_gather(SnPtrs p, const int32_t* __restrict__ rows, int nrows,
                                                      double* __restrict__ y) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= nrows) return;
    const int pos = rows[i];
    double x = y[pos];
    for (int64_t q = p.lrow_ptr[pos]; q < p.lrow_ptr[pos + 1]; q++) {
        const int S = p.lrow_sn[q], t = p.lrow_t[q];
        const int jS = p.j0[S], wS = p.j0[S + 1] - jS, hS = p.rptr[S + 1] - p.rptr[S];
        const double* L = p.v + p.oL[S] + wS + t;
        const int64_t ld = wS + hS;
        int s = 0;
        for (; s + 8 <= wS; s += 8) {  // the operands of 8 terms in flight, then the ordered subtractions
            double l[8], yv[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                l[k] = L[(int64_t)(s + k) * ld];
                yv[k] = y[jS + s + k];
            }
#pragma unroll
            for (int k = 0; k < 8; k++) x -= l[k] * yv[k];
        }
        for (; s < wS; s++) x -= L[(int64_t)s * ld] * y[jS + s];
    }
    y[pos] = x;
}

// forward inside each supernode of the list (w > 1): column sweep
__global__ __launch_bounds__(kT) void k_sn_fwd_inner(SnPtrs p, const int32_t* __restrict__ list,
                                                     double* __restrict__ y) {
    const int J = list[blockIdx.x];
    const int jj0 = p.j0[J], w = p.j0[J + 1] - jj0, hh = p.rptr[J + 1] - p.rptr[J];
    const int W = w + hh;
    const double* LP = p.v + p.oL[J];
    for (int t = 0; t < w - 1; t++) {
        const double yt = y[jj0 + t];
        for (int r = t + 1 + threadIdx.x; r < w; r += kT) y[jj0 + r] -= LP[(int64_t)t * W + r] * yt;
        __syncthreads();
    }
}

// backward, rows of one level: y[p] -= U(p, k) x[k] over R_J descending; a
// one-column supernode then divides by its pivot
__global__ __launch_bounds__(kT) void k_sn_bwd_gather(SnPtrs p, const int32_t* __restrict__ rows,
                                                      const int32_t* __restrict__ sn_of, int nrows,
                                                      double* __restrict__ y) {
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= nrows) return;
    const int pos = rows[i], J = sn_of[pos];
    const int jj0 = p.j0[J], w = p.j0[J + 1] - jj0, hh = p.rptr[J + 1] - p.rptr[J];
    const double* U = p.v + p.oU[J] + (pos - jj0);
    const int32_t* R = p.rows + p.rptr[J];
    double x = y[pos];
    int c = hh - 1;
    for (; c >= 7; c -= 8) {  // the operands of 8 terms in flight, then the ordered subtractions
        int idx[8];
        double u[8], yv[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            idx[k] = R[c - k];
            u[k] = U[(int64_t)(c - k) * w];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) yv[k] = y[idx[k]];
#pragma unroll
        for (int k = 0; k < 8; k++) x -= u[k] * yv[k];
    }
    for (; c >= 0; c--) x -= U[(int64_t)c * w] * y[R[c]];
    if (w == 1) x /= p.ud[pos];
    y[pos] = x;
}

// backward inside each supernode of the list (w > 1): x[t] = y[t] / U(t,t),
// then the rows above, columns descending
__global__ __launch_bounds__(kT) void k_sn_bwd_inner(SnPtrs p, const int32_t* __restrict__ list,
                                                     double* __restrict__ y) {
    __shared__ double sx;
    const int J = list[blockIdx.x];
    const int jj0 = p.j0[J], w = p.j0[J + 1] - jj0, hh = p.rptr[J + 1] - p.rptr[J];
    const int W = w + hh;
    const double* LP = p.v + p.oL[J];
    for (int t = w - 1; t >= 0; t--) {
        if (threadIdx.x == 0) {
            const double xt = y[jj0 + t] / p.ud[jj0 + t];
            y[jj0 + t] = xt;
            sx = xt;
        }
        __syncthreads();
        const double xt = sx;
        for (int r = threadIdx.x; r < t; r += kT) y[jj0 + r] -= LP[(int64_t)t * W + r] * xt;
        __syncthreads();
    }
}

// The same two sweeps with the supernode's y in LDS and kIB columns per
// barrier pair: wavefront 0 finishes the block's own rows (lane r - t0 holds
// row r, the finished value of a row passed to the others by shuffle), then
// every thread applies the block's columns to its rows outside the block.
// Each row still takes its terms one column at a time in the order above;
// a row's kIB multipliers of the block are loaded into registers before its
// first subtraction, so the loads overlap instead of waiting one by one.
constexpr int kIB = 16;

__global__ __launch_bounds__(kT) void k_sn_fwd_inner_lds(SnPtrs p, const int32_t* __restrict__ list,
                                                         double* __restrict__ y) {
    extern __shared__ double ys[];
    const int tid = threadIdx.x;
    const int J = list[blockIdx.x];
    const int jj0 = p.j0[J], w = p.j0[J + 1] - jj0, hh = p.rptr[J + 1] - p.rptr[J];
    const int W = w + hh;
    const double* LP = p.v + p.oL[J];
    for (int r = tid; r < w; r += kT) ys[r] = y[jj0 + r];
    __syncthreads();
    for (int t0 = 0; t0 < w; t0 += kIB) {
        const int t1 = min(t0 + kIB, w);
        if (tid < 64) {
            const int r = t0 + tid;
            double l[kIB];
#pragma unroll
            for (int k = 0; k < kIB; k++) l[k] = (t0 + k < r && r < t1) ? LP[(int64_t)(t0 + k) * W + r] : 0.0;
            double yr = r < t1 ? ys[r] : 0.0;
#pragma unroll
            for (int k = 0; k < kIB - 1; k++) {  // t = t0 + k < t1 - 1 is uniform
                const int t = t0 + k;
                if (t < t1 - 1) {
                    const double yt = __shfl(yr, k);
                    if (r > t && r < t1) yr -= l[k] * yt;
                }
            }
            if (r < t1) ys[r] = yr;
        }
        __syncthreads();
        if (t1 - t0 == kIB) {
            for (int r = t1 + tid; r < w; r += kT) {
                double l[kIB];
#pragma unroll
                for (int k = 0; k < kIB; k++) l[k] = LP[(int64_t)(t0 + k) * W + r];
                double yr = ys[r];
#pragma unroll
                for (int k = 0; k < kIB; k++) yr -= l[k] * ys[t0 + k];
                ys[r] = yr;
            }
        } else {
            for (int r = t1 + tid; r < w; r += kT) {
                double yr = ys[r];
                for (int t = t0; t < t1; t++) yr -= LP[(int64_t)t * W + r] * ys[t];
                ys[r] = yr;
            }
        }
        __syncthreads();
    }
    for (int r = tid; r < w; r += kT) y[jj0 + r] = ys[r];
}

__global__ __launch_bounds__(kT) void k_sn_bwd_inner_lds(SnPtrs p, const int32_t* __restrict__ list,
                                                         double* __restrict__ y) {
    extern __shared__ double ys[];
    const int tid = threadIdx.x;
    const int J = list[blockIdx.x];
    const int jj0 = p.j0[J], w = p.j0[J + 1] - jj0, hh = p.rptr[J + 1] - p.rptr[J];
    const int W = w + hh;
    const double* LP = p.v + p.oL[J];
    for (int r = tid; r < w; r += kT) ys[r] = y[jj0 + r];
    __syncthreads();
    for (int t1 = w; t1 > 0; t1 -= kIB) {
        const int t0 = max(t1 - kIB, 0);
        if (tid < 64) {
            const int r = t0 + tid;
            double l[kIB];
#pragma unroll
            for (int k = 0; k < kIB; k++) l[k] = (r < t0 + k && t0 + k < t1) ? LP[(int64_t)(t0 + k) * W + r] : 0.0;
            const double d = r < t1 ? p.ud[jj0 + r] : 1.0;
            double yr = r < t1 ? ys[r] : 0.0;
#pragma unroll
            for (int k = kIB - 1; k >= 0; k--) {  // t = t0 + k < t1 is uniform
                const int t = t0 + k;
                if (t < t1) {
                    if (r == t) yr = yr / d;
                    const double xt = __shfl(yr, k);
                    if (r < t) yr -= l[k] * xt;
                }
            }
            if (r < t1) ys[r] = yr;
        }
        __syncthreads();
        if (t1 - t0 == kIB) {
            for (int r = tid; r < t0; r += kT) {
                double l[kIB];
#pragma unroll
                for (int k = 0; k < kIB; k++) l[k] = LP[(int64_t)(t0 + k) * W + r];
                double yr = ys[r];
#pragma unroll
                for (int k = kIB - 1; k >= 0; k--) yr -= l[k] * ys[t0 + k];
                ys[r] = yr;
            }
        } else {
            for (int r = tid; r < t0; r += kT) {
                double yr = ys[r];
                for (int t = t1 - 1; t >= t0; t--) yr -= LP[(int64_t)t * W + r] * ys[t];
                ys[r] = yr;
            }
        }
        __syncthreads();
    }
    for (int r = tid; r < w; r += kT) y[jj0 + r] = ys[r];
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

}  // namespace

namespace rsgpu {

hipError_t sn_launch_zero(double* v, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned g = (unsigned)std::min<int64_t>((n + kT - 1) / kT, 65536);
    hipLaunchKernelGGL(k_sn_zero, dim3(g), dim3(kT), 0, s, v, n);
    return hipGetLastError();
}
hipError_t sn_launch_scatter(const int32_t* src, const int32_t* cnt, const int64_t* slot, int64_t nruns,
                             const double* ax, double* v, hipStream_t s) {
    if (nruns <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sn_scatter, dim3(blocks_for(nruns)), dim3(kT), 0, s, src, cnt, slot, nruns, ax, v);
    return hipGetLastError();
}
// levels with fewer gather tiles than this take the deep-ring gather
constexpr int kDeepItems = 512;

hipError_t sn_launch_gather(const SnPtrs& p, int item0, int nitems, hipStream_t s) {
    if (nitems <= 0) return hipSuccess;
    if (nitems < kDeepItems)
        hipLaunchKernelGGL(k_sn_gather_deep, dim3(nitems), dim3(kT), 0, s, p, item0);
    else
        hipLaunchKernelGGL(k_sn_gather, dim3(nitems), dim3(kT), 0, s, p, item0);
    return hipGetLastError();
}
hipError_t sn_launch_hot_parts(const SnPtrs& p, const rsgpu::SnPart* parts, int nparts, const int64_t* terms,
                               hipStream_t s) {
    if (nparts <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sn_hot_parts, dim3(nparts), dim3(kT), 0, s, p, parts, terms);
    return hipGetLastError();
}
hipError_t sn_launch_fwd_parts(const SnPtrs& p, const rsgpu::SnPart* parts, int nparts, const int64_t* lslot,
                               const int32_t* yidx, double* y, hipStream_t s) {
    if (nparts <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sn_fwd_parts, dim3(nparts), dim3(kT), 0, s, p, parts, lslot, yidx, y);
    return hipGetLastError();
}
hipError_t sn_launch_scatter_long(const int32_t* src, const int32_t* cnt, const int64_t* slot, int nruns,
                                  const double* ax, double* v, hipStream_t s) {
    if (nruns <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sn_scatter_long, dim3(nruns), dim3(kT), 0, s, src, cnt, slot, ax, v);
    return hipGetLastError();
}
hipError_t sn_launch_fwd_long(const SnPtrs& p, const int32_t* rows, int nrows, const int64_t* tptr,
                              const int64_t* lslot, const int32_t* yidx, double* y, hipStream_t s) {
    if (nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sn_fwd_long, dim3(nrows), dim3(kT), 0, s, p, rows, tptr, lslot, yidx, y);
    return hipGetLastError();
}
hipError_t sn_launch_factor(const SnPtrs& p, const int32_t* list, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sn_factor, dim3(n), dim3(kT), 0, s, p, list);
    return hipGetLastError();
}
// Wide supernodes of one level, widest first (list on the device, w and h on
// the host in the same order): panel step t0 runs for the first k of them,
// those wider than t0; a launch's grid covers its widest member.
hipError_t sn_launch_big(const SnPtrs& p, const int32_t* list, const int32_t* w, const int32_t* hh, int nbig,
                         hipStream_t s) {
    for (int t0 = 0; nbig > 0 && t0 < w[0]; t0 += kPanel) {
        int k = 0, nt = 0;
        for (; k < nbig && w[k] > t0; k++) {
            const int t1 = std::min(t0 + kPanel, w[k]), W = w[k] + hh[k];
            const int nrl = (W - t1 + kTile - 1) / kTile, ncl = (w[k] - t1 + kTile - 1) / kTile;
            const int nru = (w[k] - t1 + kTile - 1) / kTile, ncu = (hh[k] + kTile - 1) / kTile;
            nt = std::max(nt, nrl * ncl + nru * ncu);
        }
        int nblk = 1;
        for (int j = 0; j < k; j++) {
            const int t1 = std::min(t0 + kPanel, w[j]);
            const int nrow = w[j] + hh[j] - t1, nc = (w[j] - t1) + hh[j];
            nblk = std::max(nblk, (nrow + kRowsPerBlock - 1) / kRowsPerBlock + (nc + kT - 1) / kT);
        }
        hipLaunchKernelGGL(k_sn_panel, dim3(nblk, k), dim3(kT), 0, s, p, list, t0);
        hipLaunchKernelGGL(k_sn_trail, dim3(nt + 1, k), dim3(kT), 0, s, p, list, t0);  // + the diagonal block
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
hipError_t sn_launch_perm(const int32_t* perm, const double* in, double* out, int n, bool to_pos, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (to_pos)
        hipLaunchKernelGGL(k_sn_perm_in, dim3(blocks_for(n)), dim3(kT), 0, s, perm, in, out, n);
    else
        hipLaunchKernelGGL(k_sn_perm_out, dim3(blocks_for(n)), dim3(kT), 0, s, perm, in, out, n);
    return hipGetLastError();
}
// inner sweeps keep the supernode's y in LDS up to this width
constexpr int kInnerLdsMax = 8192;

hipError_t sn_launch_fwd(const SnPtrs& p, const int32_t* rows, int nrows, const int32_t* inner, int ninner,
                         int maxw, double* y, hipStream_t s) {
    if (nrows > 0) hipLaunchKernelGGL(k_sn_fwd_gather, dim3(blocks_for(nrows)), dim3(kT), 0, s, p, rows, nrows, y);
    if (ninner > 0) {
        if (maxw <= kInnerLdsMax)
            hipLaunchKernelGGL(k_sn_fwd_inner_lds, dim3(ninner), dim3(kT), (size_t)maxw * sizeof(double), s, p,
                               inner, y);
        else
            hipLaunchKernelGGL(k_sn_fwd_inner, dim3(ninner), dim3(kT), 0, s, p, inner, y);
    }
    return hipGetLastError();
}
hipError_t sn_launch_bwd(const SnPtrs& p, const int32_t* rows, const int32_t* sn_of, int nrows,
                         const int32_t* inner, int ninner, int maxw, double* y, hipStream_t s) {
    if (nrows > 0)
        hipLaunchKernelGGL(k_sn_bwd_gather, dim3(blocks_for(nrows)), dim3(kT), 0, s, p, rows, sn_of, nrows, y);
    if (ninner > 0) {
        if (maxw <= kInnerLdsMax)
            hipLaunchKernelGGL(k_sn_bwd_inner_lds, dim3(ninner), dim3(kT), (size_t)maxw * sizeof(double), s, p,
                               inner, y);
        else
            hipLaunchKernelGGL(k_sn_bwd_inner, dim3(ninner), dim3(kT), 0, s, p, inner, y);
    }
    return hipGetLastError();
}

}  // namespace rsgpu
