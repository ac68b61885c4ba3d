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

// One workgroup per gather tile: the tile is held in LDS, its records are
// applied in order (a barrier between records), each entry of a record by one
// thread, its terms ascending with zero multipliers skipped.
__global__ __launch_bounds__(kT) void k_sn_gather(SnPtrs p, int item0) {
    __shared__ double T[kTile * kTile];
    const int it = item0 + blockIdx.x;
    const int4 g = p.gitems[it];
    const int J = g.x, kind = g.y, r0 = g.z, c0 = g.w;
    const int jj0 = p.j0[J], w = p.j0[J + 1] - jj0, hh = p.rptr[J + 1] - p.rptr[J];
    const int W = w + hh;
    const int nr = kind == 0 ? min(kTile, W - r0) : min(kTile, w - r0);
    const int nc = kind == 0 ? min(kTile, w - c0) : min(kTile, hh - c0);
    const int ld = kind == 0 ? W : w;
    double* base = p.v + (kind == 0 ? p.oL[J] : p.oU[J]) + (int64_t)c0 * ld + r0;
    for (int e = threadIdx.x; e < kTile * kTile; e += kT) {
        const int r = e & (kTile - 1), c = e / kTile;
        T[e] = (r < nr && c < nc) ? base[(int64_t)c * ld + r] : 0.0;
    }
    __syncthreads();
    const int64_t base_off = base - p.v;
    const int64_t rb = p.tile_rec[it], re = p.tile_rec[it + 1];
    for (int64_t ri = rb; ri < re; ri++) {
        const int32_t* rc = p.rec + 5 * ri;
        const int pi = rc[0], t0 = rc[1], t1 = rc[2], u0 = rc[3], u1 = rc[4];
        const int S = p.csrc[pi], a0 = p.ca0[pi];
        const int wS = p.j0[S + 1] - p.j0[S], hS = p.rptr[S + 1] - p.rptr[S];
        const int64_t ldS = wS + hS;
        const int32_t* rl = p.rel + p.crel_off[pi] - a0;
        const int32_t* RS = p.rows + p.rptr[S];
        const double* LS = p.v + p.oL[S] + wS;  // LS[s * ldS + row]: L_S(R_S[row], s)
        const double* US = p.v + p.oU[S];       // US[col * wS + s]:  U_S(s, R_S[col])
        const int mt = t1 - t0, cnt = mt * (u1 - u0);
        for (int e = threadIdx.x; e < cnt; e += kT) {
            const int t = t0 + e % mt, u = u0 + e / mt;
            int tr, tc, lrow, ucol;
            if (kind == 0) {
                tr = rl[t] - r0; tc = RS[u] - jj0 - c0; lrow = t; ucol = u;
            } else {
                tr = RS[u] - jj0 - r0; tc = rl[t] - w - c0; lrow = u; ucol = t;
            }
            // entries summed on their own elsewhere (the ground node's diagonal)
            bool hot = false;
            for (int k = 0; k < p.nhot; k++) hot |= p.hot_slot[k] == base_off + (int64_t)tc * ld + tr;
            if (hot) continue;
            double x = T[tc * kTile + tr];
            const double* Lr = LS + lrow;
            const double* Uc = US + (int64_t)ucol * wS;
            for (int s = 0; s < wS; s++) {
                const double uu = Uc[s];
                if (uu != 0.0) x -= Lr[(int64_t)s * ldS] * uu;
            }
            T[tc * kTile + tr] = x;
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < kTile * kTile; e += kT) {
        const int r = e & (kTile - 1), c = e / kTile;
        if (r < nr && c < nc) base[(int64_t)c * ld + r] = T[e];
    }
}

// One wavefront per entry summed on its own: lanes load 64 terms at a time
// (the next 64 while lane 0 subtracts the current ones in order).
__global__ __launch_bounds__(64) void k_sn_hot(SnPtrs p, const int64_t* __restrict__ hot_ptr,
                                               const int64_t* __restrict__ terms, int e) {
    __shared__ double SL[64], SU[64];
    const int lane = threadIdx.x;
    const int64_t q0 = hot_ptr[e], q1 = hot_ptr[e + 1];
    double x = lane == 0 ? p.v[p.hot_slot[e]] : 0.0;
    double nl = 0.0, nu = 0.0;
    if (q0 + lane < q1) {
        nl = p.v[terms[2 * (q0 + lane)]];
        nu = p.v[terms[2 * (q0 + lane) + 1]];
    }
    for (int64_t b = q0; b < q1; b += 64) {
        const int c = (int)min((int64_t)64, q1 - b);
        SL[lane] = nl;
        SU[lane] = nu;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        if (b + 64 + lane < q1) {
            nl = p.v[terms[2 * (b + 64 + lane)]];
            nu = p.v[terms[2 * (b + 64 + lane) + 1]];
        }
        if (lane == 0)
            for (int i = 0; i < c; i++) {
                const double uu = SU[i];
                if (uu != 0.0) x -= SL[i] * uu;
            }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0) p.v[p.hot_slot[e]] = x;
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
// (i) the panel [t0, t1) over all rows [t0, W), one workgroup
__global__ __launch_bounds__(kT) void k_sn_panel(SnPtrs p, int J, int t0) {
    __shared__ double sd;
    const int jj0 = p.j0[J], w = p.j0[J + 1] - jj0, hh = p.rptr[J + 1] - p.rptr[J];
    const int W = w + hh, t1 = min(t0 + kPanel, w);
    double* LP = p.v + p.oL[J];
    for (int t = t0; t < t1; t++) {
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
        const int m = W - t - 1, mw = t1 - t - 1;
        for (int e = threadIdx.x; e < m * mw; e += kT) {
            const int r = t + 1 + e % m, c = t + 1 + e / m;
            const double u = LP[(int64_t)c * W + t];
            if (u != 0.0) LP[(int64_t)c * W + r] -= LP[(int64_t)t * W + r] * u;
        }
        __syncthreads();
    }
}

// (ii) the U block row: rows [t0, t1) of the later LP columns and of every
// UP column, one thread per column, the panel's pivots in order
__global__ __launch_bounds__(kT) void k_sn_urow(SnPtrs p, int J, int t0) {
    const int jj0 = p.j0[J], w = p.j0[J + 1] - jj0, hh = p.rptr[J + 1] - p.rptr[J];
    const int W = w + hh, t1 = min(t0 + kPanel, w);
    const int k = blockIdx.x * kT + threadIdx.x;
    const int nlp = w - t1;
    if (k >= nlp + hh) return;
    const double* LP = p.v + p.oL[J];
    double* U = k < nlp ? p.v + p.oL[J] + (int64_t)(t1 + k) * W : p.v + p.oU[J] + (int64_t)(k - nlp) * w;
    for (int t = t0; t < t1; t++) {
        const double u = U[t];
        if (u == 0.0) continue;
        for (int r = t + 1; r < t1; r++) U[r] -= LP[(int64_t)t * W + r] * u;
    }
}

// (iii) trailing update of the tiles right of and below the panel: LP rows
// [t1, W) x columns [t1, w) and UP rows [t1, w) x columns [0, h); the panel's
// L rows and U columns are staged in LDS, each thread owns 16 entries and
// subtracts their kPanel terms in order.
__global__ __launch_bounds__(kT) void k_sn_trail(SnPtrs p, int J, int t0) {
    __shared__ double A[kTile * kPanel];  // A[t * kTile + r]
    __shared__ double B[kTile * kPanel];  // B[c * kPanel + t]
    const int jj0 = p.j0[J], w = p.j0[J + 1] - jj0, hh = p.rptr[J + 1] - p.rptr[J];
    const int W = w + hh, t1 = min(t0 + kPanel, w), nb = t1 - t0;
    const int nrl = (W - t1 + kTile - 1) / kTile, ncl = (w - t1 + kTile - 1) / kTile;
    const int nru = (w - t1 + kTile - 1) / kTile;
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
    const int r = threadIdx.x & (kTile - 1), cg = threadIdx.x / kTile;
    if (r >= nr) return;
    for (int m = 0; m < kTile / 4; m++) {
        const int c = cg * (kTile / 4) + m;
        if (c >= nc) break;
        double x = C[(int64_t)(c0 + c) * ldc + r0 + r];
        for (int t = 0; t < nb; t++) {
            const double u = B[c * kPanel + t];
            if (u != 0.0) x -= A[t * kTile + r] * u;
        }
        C[(int64_t)(c0 + c) * ldc + r0 + r] = x;
    }
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
        for (int s = 0; s < wS; s++) x -= L[(int64_t)s * ld] * y[jS + s];
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
    for (int c = hh - 1; c >= 0; c--) x -= U[(int64_t)c * w] * y[R[c]];
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
hipError_t sn_launch_gather(const SnPtrs& p, int item0, int nitems, hipStream_t s) {
    if (nitems <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sn_gather, dim3(nitems), dim3(kT), 0, s, p, item0);
    return hipGetLastError();
}
hipError_t sn_launch_hot(const SnPtrs& p, const int64_t* hot_ptr, const int64_t* terms, int e, hipStream_t s) {
    hipLaunchKernelGGL(k_sn_hot, dim3(1), dim3(64), 0, s, p, hot_ptr, terms, e);
    return hipGetLastError();
}
hipError_t sn_launch_factor(const SnPtrs& p, const int32_t* list, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sn_factor, dim3(n), dim3(kT), 0, s, p, list);
    return hipGetLastError();
}
hipError_t sn_launch_big(const SnPtrs& p, int J, int w, int hh, hipStream_t s) {
    const int W = w + hh;
    for (int t0 = 0; t0 < w; t0 += kPanel) {
        const int t1 = std::min(t0 + kPanel, w);
        hipLaunchKernelGGL(k_sn_panel, dim3(1), dim3(kT), 0, s, p, J, t0);
        const int ncol = (w - t1) + hh;
        if (ncol > 0) hipLaunchKernelGGL(k_sn_urow, dim3(blocks_for(ncol)), dim3(kT), 0, s, p, J, t0);
        const int nrl = (W - t1 + kTile - 1) / kTile, ncl = (w - t1 + kTile - 1) / kTile;
        const int nru = (w - t1 + kTile - 1) / kTile, ncu = (hh + kTile - 1) / kTile;
        const int nt = nrl * ncl + nru * ncu;
        if (nt > 0) hipLaunchKernelGGL(k_sn_trail, dim3(nt), dim3(kT), 0, s, p, J, t0);
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
hipError_t sn_launch_fwd(const SnPtrs& p, const int32_t* rows, int nrows, const int32_t* inner, int ninner,
                         double* y, hipStream_t s) {
    if (nrows > 0) hipLaunchKernelGGL(k_sn_fwd_gather, dim3(blocks_for(nrows)), dim3(kT), 0, s, p, rows, nrows, y);
    if (ninner > 0) hipLaunchKernelGGL(k_sn_fwd_inner, dim3(ninner), dim3(kT), 0, s, p, inner, y);
    return hipGetLastError();
}
hipError_t sn_launch_bwd(const SnPtrs& p, const int32_t* rows, const int32_t* sn_of, int nrows,
                         const int32_t* inner, int ninner, double* y, hipStream_t s) {
    if (nrows > 0)
        hipLaunchKernelGGL(k_sn_bwd_gather, dim3(blocks_for(nrows)), dim3(kT), 0, s, p, rows, sn_of, nrows, y);
    if (ninner > 0) hipLaunchKernelGGL(k_sn_bwd_inner, dim3(ninner), dim3(kT), 0, s, p, inner, y);
    return hipGetLastError();
}

}  // namespace rsgpu
