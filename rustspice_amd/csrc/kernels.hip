// kernels.hip - CDNA4 (gfx950) kernels for the static-pattern sparse LU.
//
// k_refactor: numeric refactorization on a fixed pattern and pivot sequence
//   (native_klu.rs:1261-1373 with reused pivots; sparse_lu.rs:451-534 is the
//   same loop with identity pivots). One workgroup per matrix; columns run in
//   dependency levels, one wavefront per column. Each wavefront owns a dense
//   work vector (LDS when it fits). Per column the U entries are applied in
//   ascending row order, each lane owning distinct rows of L(:,j), so every
//   work[row] sees the reference's updates in the reference's order: results
//   are bit-identical to the sequential loop (no FMA contraction).
// k_solve: P, L (unit), U, P^T sweeps of native_klu.rs:1442-1488 and the BTF
//   block back-substitution of :1491-1551, row-oriented so each unknown is a
//   sequential dot product in the reference's update order, rows of one
//   dependency level in parallel.
#include <hip/hip_runtime.h>

#include "plan.hpp"

using rsgpu::PlanDev;

namespace {

constexpr int kWave = 64;

__device__ __forceinline__ void wave_sync() {
    // Orders this wave's LDS/global accesses before the following ones
    // (lanes exchange data through memory between steps).
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

template <bool kLds>
__global__ __launch_bounds__(1024) void k_refactor(const PlanDev* __restrict__ descs,
                                                  const int32_t* __restrict__ ip,
                                                  const double* __restrict__ axpool,
                                                  double* __restrict__ vpool,
                                                  double* __restrict__ wpool) {
    extern __shared__ double smem[];
    const PlanDev d = descs[blockIdx.x];
    const int n = d.n;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    double* work = kLds ? smem + (size_t)wid * n : wpool + d.o_work + (size_t)wid * n;
    if (kLds) {
        for (int i = threadIdx.x; i < nw * n; i += blockDim.x) smem[i] = 0.0;
        __syncthreads();
    }
    const int32_t* lcp = ip + d.o_lcp;
    const int32_t* lri = ip + d.o_lri;
    const int32_t* ucp = ip + d.o_ucp;
    const int32_t* uri = ip + d.o_uri;
    const int32_t* colflag = ip + d.o_colflag;
    const int32_t* flev_ptr = ip + d.o_flev_ptr;
    const int32_t* flev_cols = ip + d.o_flev_cols;
    const int32_t* sc_ptr = ip + d.o_sc_ptr;
    const int32_t* sc_src = ip + d.o_sc_src;
    const int32_t* sc_cnt = ip + d.o_sc_cnt;
    const int32_t* sc_dst = ip + d.o_sc_dst;
    const double* ax = axpool + d.o_ax;
    double* lx = vpool + d.o_lx;
    double* ux = vpool + d.o_ux;
    double* ud = vpool + d.o_ud;

    for (int lev = 0; lev < d.nflev; ++lev) {
        const int c0 = flev_ptr[lev], c1 = flev_ptr[lev + 1];
        for (int c = c0 + wid; c < c1; c += nw) {
            const int g = flev_cols[c];
            // scatter A(:,g): duplicates summed in CSC order starting from 0.0
            for (int s = sc_ptr[g] + lane; s < sc_ptr[g + 1]; s += kWave) {
                double v = 0.0;
                const int src = sc_src[s], cnt = sc_cnt[s];
                for (int q = 0; q < cnt; ++q) v += ax[src + q];
                work[sc_dst[s]] = v;
            }
            wave_sync();
            const int us = ucp[g], ue = ucp[g + 1] - 1;
            for (int u = us; u < ue; ++u) {
                const int j = uri[u];
                const double ujk = work[j];
                if (ujk != 0.0) {
                    for (int l = lcp[j] + lane; l < lcp[j + 1]; l += kWave) {
                        const int row = lri[l];
                        work[row] = work[row] - lx[l] * ujk;
                    }
                }
                wave_sync();
            }
            for (int u = us + lane; u < ue; u += kWave) ux[u] = work[uri[u]];
            double diag = work[g];
            if (colflag[g] & rsgpu::COL_SINGLETON) {
                if (fabs(diag) < 1e-300) diag = 1e-14;
            } else if (fabs(diag) < 1e-14) {
                diag = diag >= 0.0 ? 1e-14 : -1e-14;
            }
            if (lane == 0) {
                ux[ue] = diag;
                ud[g] = diag;
            }
            for (int l = lcp[g] + lane; l < lcp[g + 1]; l += kWave) lx[l] = work[lri[l]] / diag;
            wave_sync();
            // clear touched entries (keeps the all-zero invariant of work)
            for (int s = sc_ptr[g] + lane; s < sc_ptr[g + 1]; s += kWave) work[sc_dst[s]] = 0.0;
            for (int l = lcp[g] + lane; l < lcp[g + 1]; l += kWave) work[lri[l]] = 0.0;
            for (int u = us + lane; u < ue; u += kWave) work[uri[u]] = 0.0;
            if (lane == 0) work[g] = 0.0;
            wave_sync();
        }
        __syncthreads();
    }
}

template <bool kLds>
__global__ __launch_bounds__(1024) void k_solve(const PlanDev* __restrict__ descs,
                                               const int32_t* __restrict__ ip,
                                               const double* __restrict__ axpool,
                                               const double* __restrict__ vpool,
                                               const double* __restrict__ rhspool,
                                               double* __restrict__ xpool,
                                               double* __restrict__ wpool) {
    extern __shared__ double smem[];
    const PlanDev d = descs[blockIdx.x];
    const int n = d.n;
    double* T = kLds ? smem : wpool + d.o_tvec;
    const int32_t* perm_in = ip + d.o_perm_in;
    const int32_t* perm_out = ip + d.o_perm_out;
    const int32_t* lrp = ip + d.o_lrp;
    const int32_t* lrc = ip + d.o_lrc;
    const int32_t* lrv = ip + d.o_lrv;
    const int32_t* urp = ip + d.o_urp;
    const int32_t* urc = ip + d.o_urc;
    const int32_t* urv = ip + d.o_urv;
    const int32_t* lsl_blk = ip + d.o_lsl_blk;
    const int32_t* lsl_ptr = ip + d.o_lsl_ptr;
    const int32_t* lsl_rows = ip + d.o_lsl_rows;
    const int32_t* usl_blk = ip + d.o_usl_blk;
    const int32_t* usl_ptr = ip + d.o_usl_ptr;
    const int32_t* usl_rows = ip + d.o_usl_rows;
    const int32_t* blk_ptr = ip + d.o_blk_ptr;
    const int32_t* odp = ip + d.o_odp;
    const int32_t* odax = ip + d.o_odax;
    const int32_t* odsrc = ip + d.o_odsrc;
    const int32_t* has_od = ip + d.o_blk_has_od;
    const double* ax = axpool + d.o_ax;
    const double* lx = vpool + d.o_lx;
    const double* ux = vpool + d.o_ux;
    const double* ud = vpool + d.o_ud;
    const double* rhs = rhspool + d.o_vec;
    double* x = xpool + d.o_vec;

    for (int i = threadIdx.x; i < n; i += blockDim.x) T[perm_in[i]] = rhs[i];
    __syncthreads();
    for (int b = d.nblk - 1; b >= 0; --b) {
        if (has_od[b]) {
            for (int r = blk_ptr[b] + threadIdx.x; r < blk_ptr[b + 1]; r += blockDim.x) {
                const int e0 = odp[r], e1 = odp[r + 1];
                if (e0 == e1) continue;
                double t = T[r];
                for (int e = e0; e < e1; ++e) t = t - ax[odax[e]] * T[odsrc[e]];
                T[r] = t;
            }
            __syncthreads();
        }
        for (int lev = lsl_blk[b]; lev < lsl_blk[b + 1]; ++lev) {
            for (int q = lsl_ptr[lev] + threadIdx.x; q < lsl_ptr[lev + 1]; q += blockDim.x) {
                const int r = lsl_rows[q];
                const int e0 = lrp[r], e1 = lrp[r + 1];
                if (e0 == e1) continue;
                double t = T[r];
                for (int e = e0; e < e1; ++e) t = t - lx[lrv[e]] * T[lrc[e]];
                T[r] = t;
            }
            __syncthreads();
        }
        for (int lev = usl_blk[b]; lev < usl_blk[b + 1]; ++lev) {
            for (int q = usl_ptr[lev] + threadIdx.x; q < usl_ptr[lev + 1]; q += blockDim.x) {
                const int r = usl_rows[q];
                double t = T[r];
                for (int e = urp[r]; e < urp[r + 1]; ++e) t = t - ux[urv[e]] * T[urc[e]];
                T[r] = t / ud[r];
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = T[perm_out[i]];
}

}  // namespace

namespace rsgpu {

hipError_t launch_refactor(const PlanDev* descs, int nmat, const int32_t* ip, const double* ax,
                           double* vals, double* wpool, int max_n, bool lds, int threads,
                           hipStream_t stream) {
    size_t shm = lds ? (size_t)(threads / kWave) * max_n * sizeof(double) : 0;
    if (lds)
        hipLaunchKernelGGL(k_refactor<true>, dim3(nmat), dim3(threads), shm, stream, descs, ip, ax,
                           vals, wpool);
    else
        hipLaunchKernelGGL(k_refactor<false>, dim3(nmat), dim3(threads), 0, stream, descs, ip, ax,
                           vals, wpool);
    return hipGetLastError();
}

hipError_t launch_solve(const PlanDev* descs, int nmat, const int32_t* ip, const double* ax,
                        const double* vals, const double* rhs, double* x, double* wpool, int max_n,
                        bool lds, int threads, hipStream_t stream) {
    size_t shm = lds ? (size_t)max_n * sizeof(double) : 0;
    if (lds)
        hipLaunchKernelGGL(k_solve<true>, dim3(nmat), dim3(threads), shm, stream, descs, ip, ax, vals,
                           rhs, x, wpool);
    else
        hipLaunchKernelGGL(k_solve<false>, dim3(nmat), dim3(threads), 0, stream, descs, ip, ax, vals,
                           rhs, x, wpool);
    return hipGetLastError();
}

}  // namespace rsgpu
