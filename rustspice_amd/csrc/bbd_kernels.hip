// bbd_kernels.hip - device kernels of the BbdSolver and DenseSolver back ends
// (bbd.hpp). All sums keep the reference's operand order per output entry;
// parallelism is across entries only. The one exception is the opt-in dense
// Schur product on MFMA tiles (k_bbd_schur_mfma), which sums per tile.
#include <hip/hip_runtime.h>

#include "bbd.hpp"

namespace {

constexpr int kT = 256;

inline dim3 grid_for(int64_t m) { return dim3((unsigned)((m + kT - 1) / kT)); }

__global__ void k_gather(double* __restrict__ dst, const double* __restrict__ src,
                         const int32_t* __restrict__ idx, int64_t m) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) dst[i] = src[idx[i]];
}

// dst[i] = 0.0 + src[idx[ptr[i]]] + ... (duplicates summed in CSC order)
__global__ void k_gather_sum(double* __restrict__ dst, const double* __restrict__ src,
                             const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                             int64_t m) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double acc = 0.0;
    for (int32_t q = ptr[i]; q < ptr[i + 1]; ++q) acc += src[idx[q]];
    dst[i] = acc;
}

// Right-hand sides of the Schur solves: pair p = (block k, border column j)
// gets R_k(:, j) written by assignment in CSC order, so a repeated row keeps
// its last value (bbd_solver.rs:786-789). y must be zero on entry.
__global__ void k_bbd_pair_rhs(double* __restrict__ y, const int32_t* __restrict__ pair_off,
                               const int32_t* __restrict__ pair_ptr,
                               const int32_t* __restrict__ ent_row,
                               const double* __restrict__ ent_val, int64_t npairs) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    double* yp = y + pair_off[p];
    for (int32_t q = pair_ptr[p]; q < pair_ptr[p + 1]; ++q) yp[ent_row[q]] = ent_val[q];
}

// One Schur entry (r, j): base minus C_k(r, col) * Y_kj(col) over the C
// entries of border row r in (block, column, entry) order, skipping zero y
// (bbd_solver.rs:790-806, 858-876).
__global__ void k_bbd_schur(double* __restrict__ out, const int32_t* __restrict__ cell_row,
                            const int32_t* __restrict__ cell_col, const double* __restrict__ base,
                            const int32_t* __restrict__ crow_ptr, const int32_t* __restrict__ c_blk,
                            const int32_t* __restrict__ c_col, const double* __restrict__ c_val,
                            const int32_t* __restrict__ pair_of,
                            const int32_t* __restrict__ pair_off, const double* __restrict__ y,
                            int64_t bs, int64_t ncells) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncells) return;
    const int32_t r = cell_row[i], j = cell_col[i];
    double acc = base[i];
    for (int32_t e = crow_ptr[r]; e < crow_ptr[r + 1]; ++e) {
        const int32_t p = pair_of[(int64_t)c_blk[e] * bs + j];
        if (p < 0) continue;
        const double yv = y[pair_off[p] + c_col[e]];
        if (yv == 0.0) continue;
        acc = acc - c_val[e] * yv;
    }
    out[i] = acc;
}

// ---- Dense Schur product on FP64 MFMA tiles --------------------------------
// S = B_T - Cp * Yp, where the K "coupling columns" are the (block, local
// column) pairs that hold a C entry. Cp (K x bsp) holds, at (c, r), the sum
// of border row r's C entries in coupling column c; Yp (K x bsp) holds, at
// (c, j), Y_kj(col) of that column (0 where the pair (k, j) has no R entry).
// Both are k-major with a row stride of bsp = border size rounded up to 16,
// zero in the padding, and K is padded to a multiple of 16. The product's
// terms are summed per 16x16 tile and K chunk by the matrix cores, then the
// chunks in chunk order: deterministic, but not the reference's term order.
__global__ void k_bbd_pack(double* __restrict__ cp, double* __restrict__ yp,
                           const double* __restrict__ ax, const int32_t* __restrict__ cptr,
                           const int32_t* __restrict__ cidx, const int32_t* __restrict__ ymap,
                           const double* __restrict__ y, int64_t m) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    double s = 0.0;
    for (int32_t q = cptr[t]; q < cptr[t + 1]; ++q) s += ax[cidx[q]];
    cp[t] = s;
    const int32_t o = ymap[t];
    yp[t] = o >= 0 ? y[o] : 0.0;
}

typedef double dx4 __attribute__((ext_vector_type(4)));

// Grid (tiles, chunks), 256 threads. Workgroup (t, s) multiplies tile t =
// (tr, tc) of the output over K rows [s*kc, min(kp, (s+1)*kc)); wavefront w
// takes the 4-row steps w*4, w*4+16, ... of that range. kc and kp are
// multiples of 16, so every step's 4 rows lie below kp. v_mfma_f64_16x16x4f64
// operands: lane l holds A(i = l%16, k = l/16) and B(k = l/16, j = l%16); the
// result's element v of lane l is D(l/16 + 4*v, l%16) (rows interleaved by
// four, unlike the f32 forms). The four
// wavefronts' tiles are added in wavefront order into part[s][t][16x16].
__global__ __launch_bounds__(256) void k_bbd_schur_mfma(double* __restrict__ part,
                                                        const double* __restrict__ cp,
                                                        const double* __restrict__ yp, int32_t nt,
                                                        int32_t bsp, int64_t kp, int64_t kc) {
    __shared__ double red[4][256];
    const int32_t t = blockIdx.x, tr = t / nt, tc = t % nt;
    const int32_t w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int64_t kb = (int64_t)blockIdx.y * kc;
    const int64_t ke = kb + kc < kp ? kb + kc : kp;
    const int32_t ri = tr * 16 + (l & 15), cj = tc * 16 + (l & 15);
    dx4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int64_t k0 = kb + w * 4; k0 < ke; k0 += 16) {
        const int64_t k = k0 + (l >> 4);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(cp[k * bsp + ri], yp[k * bsp + cj], acc, 0, 0, 0);
    }
    for (int v = 0; v < 4; ++v) red[w][((l >> 4) + 4 * v) * 16 + (l & 15)] = acc[v];
    __syncthreads();
    const int32_t e = threadIdx.x;
    const double s = ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
    part[((int64_t)blockIdx.y * gridDim.x + t) * 256 + e] = s;
}

// out[r * bs + j] = base - sum over chunks (in chunk order) of the partial
// products of cell (r, j).
__global__ void k_bbd_schur_mfma_out(double* __restrict__ out, const double* __restrict__ base,
                                     const double* __restrict__ part, int32_t bs, int32_t nt,
                                     int32_t nchunk) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bs * bs) return;
    const int32_t r = i / bs, j = i % bs;
    const int32_t t = (r >> 4) * nt + (j >> 4), e = (r & 15) * 16 + (j & 15);
    const int64_t stride = (int64_t)nt * nt * 256;
    double acc = 0.0;
    for (int32_t c = 0; c < nchunk; ++c) acc += part[c * stride + (int64_t)t * 256 + e];
    out[i] = base[i] - acc;
}

// Border right-hand side b_T - sum_k C_k z_k (matvec_add_neg,
// bbd_solver.rs:988-1000), one border row per thread.
__global__ void k_bbd_border_rhs(double* __restrict__ bb, const double* __restrict__ work,
                                 int64_t b0, const int32_t* __restrict__ crow_ptr,
                                 const int32_t* __restrict__ c_blk, const int32_t* __restrict__ c_col,
                                 const double* __restrict__ c_val,
                                 const int32_t* __restrict__ blk_start,
                                 const double* __restrict__ z, int64_t bs) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= bs) return;
    double acc = work[b0 + r];
    for (int32_t e = crow_ptr[r]; e < crow_ptr[r + 1]; ++e) {
        const double xc = z[blk_start[c_blk[e]] + c_col[e]];
        if (xc == 0.0) continue;
        acc = acc - c_val[e] * xc;
    }
    bb[r] = acc;
}

// w = R_k x_T for every block row (matvec_add, bbd_solver.rs:68-81): row
// entries in (border column, entry) order, zero x skipped, from 0.0.
__global__ void k_bbd_rprod(double* __restrict__ w, const int32_t* __restrict__ rrow_ptr,
                            const int32_t* __restrict__ r_col, const double* __restrict__ r_val,
                            const double* __restrict__ xt, int64_t m) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double acc = 0.0;
    for (int32_t e = rrow_ptr[i]; e < rrow_ptr[i + 1]; ++e) {
        const double xc = xt[r_col[e]];
        if (xc == 0.0) continue;
        acc = acc + r_val[e] * xc;
    }
    w[i] = acc;
}

__global__ void k_permute(double* __restrict__ dst, const double* __restrict__ src,
                          const int32_t* __restrict__ idx, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

__global__ void k_bbd_combine(double* __restrict__ work, const double* __restrict__ z,
                              const double* __restrict__ w, const double* __restrict__ xt,
                              int64_t b0, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    work[i] = i < b0 ? z[i] - w[i] : xt[i - b0];
}

// DenseSolver::build_dense (solver.rs:619-637): one column per thread, so a
// cell's duplicates are added in CSC order. lu must be zero on entry.
__global__ void k_dense_build(double* __restrict__ lu, const int32_t* __restrict__ ap,
                              const int32_t* __restrict__ ai, const double* __restrict__ ax,
                              int64_t n) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    for (int32_t q = ap[c]; q < ap[c + 1]; ++q) {
        const int32_t r = ai[q];
        if (r < n) lu[(int64_t)r * n + c] += ax[q];
    }
}

// Dense LU with partial pivoting, one workgroup (solver.rs:640-672 with
// schur == 0, bbd_solver.rs:112-156 with schur != 0). The pivot search is the
// reference's sequential scan (first strictly larger value wins); the row
// swap, the multipliers and the rank-1 update are entry-parallel and every
// entry is updated once per step k, in k order.
__global__ __launch_bounds__(kT) void k_dense_lu(double* __restrict__ lu, int32_t* __restrict__ piv,
                                                 int32_t* __restrict__ status, int64_t n,
                                                 int schur) {
    __shared__ int64_t s_p, s_kfail;
    __shared__ int s_fail;
    const int t = threadIdx.x;
    for (int64_t i = t; i < n; i += kT) piv[i] = (int32_t)i;
    if (t == 0) s_fail = 0;
    __syncthreads();
    for (int64_t k = 0; k < n; ++k) {
        if (t == 0) {
            int64_t p = k;
            double mx = fabs(lu[k * n + k]);
            for (int64_t i = k + 1; i < n; ++i) {
                const double v = fabs(lu[i * n + k]);
                if (v > mx) {
                    mx = v;
                    p = i;
                }
            }
            if (schur ? (mx < 1e-30) : (mx == 0.0)) {
                s_fail = schur ? 4 : 2;
                s_kfail = k;
            }
            s_p = p;
        }
        __syncthreads();
        if (s_fail) break;
        const int64_t p = s_p;
        if (p != k) {
            for (int64_t j = t; j < n; j += kT) {
                const double a = lu[k * n + j];
                lu[k * n + j] = lu[p * n + j];
                lu[p * n + j] = a;
            }
            if (t == 0) {
                const int32_t a = piv[k];
                piv[k] = piv[p];
                piv[p] = a;
            }
        }
        __syncthreads();
        const double pv = lu[k * n + k];
        for (int64_t i = k + 1 + t; i < n; i += kT) lu[i * n + k] = lu[i * n + k] / pv;
        __syncthreads();
        const int64_t m = n - k - 1;
        for (int64_t e = t; e < m * m; e += kT) {
            const int64_t i = k + 1 + e / m, j = k + 1 + e % m;
            lu[i * n + j] = lu[i * n + j] - lu[i * n + k] * lu[k * n + j];
        }
        __syncthreads();
    }
    if (t == 0) {  // status[0]: 0, 2 (FactorFailed) or 4 (SingularMatrix); status[1]: the pivot step
        status[0] = s_fail;
        status[1] = s_fail ? (int32_t)s_kfail : -1;
    }
}

// Forward and backward substitution of the dense factor (solver.rs:698-721,
// bbd_solver.rs:159-196), sequential in one lane like the reference.
__global__ void k_dense_solve(const double* __restrict__ lu, const int32_t* __restrict__ piv,
                              double* __restrict__ rhs, double* __restrict__ b,
                              int32_t* __restrict__ status, int64_t n, int schur) {
    if (threadIdx.x != 0) return;
    for (int64_t i = 0; i < n; ++i) b[i] = rhs[piv[i]];
    for (int64_t i = 0; i < n; ++i) {
        double s = b[i];
        for (int64_t j = 0; j < i; ++j) s = s - lu[i * n + j] * b[j];
        b[i] = s;
    }
    int st = 0;
    for (int64_t i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int64_t j = i + 1; j < n; ++j) s = s - lu[i * n + j] * rhs[j];
        const double dg = lu[i * n + i];
        if (schur ? (fabs(dg) < 1e-30) : (dg == 0.0)) {
            st = 3;
            break;
        }
        rhs[i] = s / dg;
    }
    *status = st;
}

}  // namespace

namespace rsgpu {

hipError_t launch_gather(double* dst, const double* src, const int32_t* idx, int64_t m,
                         hipStream_t s) {
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather, grid_for(m), dim3(kT), 0, s, dst, src, idx, m);
    return hipGetLastError();
}

hipError_t launch_gather_sum(double* dst, const double* src, const int32_t* ptr, const int32_t* idx,
                             int64_t m, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_sum, grid_for(m), dim3(kT), 0, s, dst, src, ptr, idx, m);
    return hipGetLastError();
}

hipError_t launch_bbd_pair_rhs(double* y, const int32_t* pair_off, const int32_t* pair_ptr,
                               const int32_t* ent_row, const double* ent_val, int64_t npairs,
                               hipStream_t s) {
    if (npairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bbd_pair_rhs, grid_for(npairs), dim3(kT), 0, s, y, pair_off, pair_ptr,
                       ent_row, ent_val, npairs);
    return hipGetLastError();
}

hipError_t launch_bbd_schur(double* out, const int32_t* cell_row, const int32_t* cell_col,
                            const double* base, const int32_t* crow_ptr, const int32_t* c_blk,
                            const int32_t* c_col, const double* c_val, const int32_t* pair_of,
                            const int32_t* pair_off, const double* y, int64_t bs, int64_t ncells,
                            hipStream_t s) {
    if (ncells <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bbd_schur, grid_for(ncells), dim3(kT), 0, s, out, cell_row, cell_col, base,
                       crow_ptr, c_blk, c_col, c_val, pair_of, pair_off, y, bs, ncells);
    return hipGetLastError();
}

hipError_t launch_bbd_schur_mfma(double* out, const double* base, double* cp, double* yp,
                                 double* part, const double* ax, const int32_t* cptr,
                                 const int32_t* cidx, const int32_t* ymap, const double* y,
                                 int32_t bs, int64_t kp, int32_t nchunk, int64_t kc,
                                 hipStream_t s) {
    if (bs <= 0) return hipSuccess;
    const int32_t bsp = (bs + 15) / 16 * 16, nt = bsp / 16;
    if (kp <= 0 || kp % 16 || kc % 16 || nchunk < 1 || (int64_t)nchunk * kc < kp)
        return hipErrorInvalidValue;
    const int64_t m = kp * bsp;
    hipLaunchKernelGGL(k_bbd_pack, grid_for(m), dim3(kT), 0, s, cp, yp, ax, cptr, cidx, ymap, y, m);
    hipLaunchKernelGGL(k_bbd_schur_mfma, dim3(nt * nt, nchunk), dim3(256), 0, s, part, cp, yp, nt,
                       bsp, kp, kc);
    hipLaunchKernelGGL(k_bbd_schur_mfma_out, grid_for((int64_t)bs * bs), dim3(kT), 0, s, out, base,
                       part, bs, nt, nchunk);
    return hipGetLastError();
}

hipError_t launch_bbd_border_rhs(double* bb, const double* work, int64_t b0,
                                 const int32_t* crow_ptr, const int32_t* c_blk,
                                 const int32_t* c_col, const double* c_val,
                                 const int32_t* blk_start, const double* z, int64_t bs,
                                 hipStream_t s) {
    if (bs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bbd_border_rhs, grid_for(bs), dim3(kT), 0, s, bb, work, b0, crow_ptr, c_blk,
                       c_col, c_val, blk_start, z, bs);
    return hipGetLastError();
}

hipError_t launch_bbd_rprod(double* w, const int32_t* rrow_ptr, const int32_t* r_col,
                            const double* r_val, const double* xt, int64_t m, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bbd_rprod, grid_for(m), dim3(kT), 0, s, w, rrow_ptr, r_col, r_val, xt, m);
    return hipGetLastError();
}

hipError_t launch_permute(double* dst, const double* src, const int32_t* idx, int64_t n,
                          hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_permute, grid_for(n), dim3(kT), 0, s, dst, src, idx, n);
    return hipGetLastError();
}

hipError_t launch_bbd_combine(double* work, const double* z, const double* w, const double* xt,
                              int64_t b0, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_bbd_combine, grid_for(n), dim3(kT), 0, s, work, z, w, xt, b0, n);
    return hipGetLastError();
}

hipError_t launch_dense_build(double* lu, const int32_t* ap, const int32_t* ai, const double* ax,
                              int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_dense_build, grid_for(n), dim3(kT), 0, s, lu, ap, ai, ax, n);
    return hipGetLastError();
}

hipError_t launch_dense_lu(double* lu, int32_t* piv, int32_t* status, int64_t n, int schur,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_dense_lu, dim3(1), dim3(kT), 0, s, lu, piv, status, n, schur);
    return hipGetLastError();
}

hipError_t launch_dense_solve(const double* lu, const int32_t* piv, double* rhs, double* b,
                              int32_t* status, int64_t n, int schur, hipStream_t s) {
    hipLaunchKernelGGL(k_dense_solve, dim3(1), dim3(64), 0, s, lu, piv, rhs, b, status, n, schur);
    return hipGetLastError();
}

}  // namespace rsgpu
