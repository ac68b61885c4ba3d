// stamp_kernels.hip - device evaluation and MNA assembly (K1/K2).
//
// k_eval: one thread per device instance, writing that device's stamp values
//   (kSlotsPerDevice doubles) in the order stamp.rs adds them:
//   R stamp.rs:242-259, I :261-276/:331-349, V :278-325, C (BE) :759-787,
//   C (TRAP) :871-915, D :351-387, M level 1 :389-553 with
//   crates/sim-devices/src/bsim/evaluate.rs:285-356; PULSE waveform.rs:82-113.
// k_assemble: ax[q] = value of the stamp that produced CSC entry q (the
//   reference's unmerged array, bit for bit); rhs[i] = 0.0 + its
//   contributions in insertion order (mna.rs:465-469).
// k_cap_state: capacitor history after an accepted step (stamp.rs:572-668).
#include <hip/hip_runtime.h>

#include "device_eval.hpp"
#include "fold.hpp"
#include "stamp.hpp"

namespace {

using rsgpu::kSlotsPerDevice;
using rsgpu::StampArgs;
using rsgpu::xget;

__global__ __launch_bounds__(256) void k_eval(const rs_device* __restrict__ dev, int ndev,
                                              const double* __restrict__ x, int xl, StampArgs a,
                                              const double* __restrict__ cap_v,
                                              const double* __restrict__ cap_i,
                                              const uint8_t* __restrict__ has_v,
                                              const uint8_t* __restrict__ has_i,
                                              double* __restrict__ sv,
                                              const int32_t* __restrict__ sbase,
                                              uint8_t* __restrict__ gmask) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ndev) return;
    if (a.skip && *a.skip) return;
    if (a.t_dev) a.t = *a.t_dev;
    if (a.dt_dev) a.dt = *a.dt_dev;
    if (a.mode_dev) {
        a.mode = *a.mode_dev;
        if (a.mode == rsgpu::MODE_TRAN_BE && a.be_cap_v) {
            cap_v = a.be_cap_v;
            cap_i = a.be_cap_i;
            has_v = a.be_has_v;
            has_i = a.be_has_i;
        }
    }
    const rs_device v = dev[d];
    if (sbase && rsgpu::is_bsim(v)) {
        gmask[d] = rsgpu::eval_bsim(v, x, xl, a, sv + sbase[d]);
        return;
    }
    rsgpu::eval_device(v, d, x, xl, a, cap_v, cap_i, has_v, has_i, sv + (size_t)d * kSlotsPerDevice);
}

// Model evaluation for arrays of bias points (BSIM levels).
__global__ __launch_bounds__(256) void k_mos_eval(const rs_bsim_model* __restrict__ models,
                                                  const int32_t* __restrict__ mi,
                                                  const int32_t* __restrict__ lv,
                                                  const double* __restrict__ in, int64_t n,
                                                  double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* q = in + 7 * i;  // w, l, vd, vg, vs, vb, temp
    const rsgpu::BsimOut r = rsgpu::bsim_eval(models[mi[i]], lv[i], q[0], q[1], q[2], q[3], q[4],
                                              q[5], q[6]);
    double* o = out + 13 * i;
    o[0] = r.ids; o[1] = r.gm; o[2] = r.gds; o[3] = r.gmbs; o[4] = r.ieq; o[5] = (double)r.region;
    o[6] = r.vth; o[7] = r.isub; o[8] = r.gsub; o[9] = r.igs; o[10] = r.igd; o[11] = r.gigs;
    o[12] = r.gigd;
}

__device__ __forceinline__ void lane_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// RHS rows with more contributions than this are summed by the whole
// wavefront: lanes gather 64 values at a time, lane 0 adds them in order.
constexpr int kLongRhs = 16;

// ax[q] from its stamp slot (thread q); rhs[i] = 0.0 + its contributions in
// insertion order (thread i, or the wavefront holding row i when the row is
// long, e.g. the ground row of a large mesh). The additions are the same in
// both cases, so the result does not depend on which path summed a row.
__global__ __launch_bounds__(256) void k_assemble(const int32_t* __restrict__ ax_src, int nnz,
                                                  const int32_t* __restrict__ rhs_ptr,
                                                  const int32_t* __restrict__ rhs_src, int n,
                                                  const double* __restrict__ sv,
                                                  double* __restrict__ ax,
                                                  double* __restrict__ rhs,
                                                  const int32_t* __restrict__ skip) {
    __shared__ double stage[256];
    if (skip && *skip) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    if (t < nnz) {
        const int s = ax_src[t];
        ax[t] = s < 0 ? 1.0 : sv[s];
    }
    if (t - lane >= n) return;  // wavefront-uniform: no row in this wavefront
    int q0 = 0, q1 = 0;
    if (t < n) {
        q0 = rhs_ptr[t];
        q1 = rhs_ptr[t + 1];
    }
    if (q1 - q0 > rsgpu::kFoldRhs) q1 = q0 - 1;  // a row for k_assemble_long: no work here
    const bool lng = q1 - q0 > kLongRhs;
    if (t < n && !lng && q1 >= q0) {
        double r = 0.0;
        for (int q = q0; q < q1; ++q) r += sv[rhs_src[q]];
        rhs[t] = r;
    }
    double* S = stage + (threadIdx.x - lane);
    for (unsigned long long m = __ballot(lng); m; m &= m - 1) {
        const int l = __ffsll((long long)m) - 1;
        const int b0 = __shfl(q0, l), b1 = __shfl(q1, l);
        double acc = 0.0;
        // double-buffered: the next 64 values are gathered into registers
        // while lane 0 adds the current ones
        double nxt = b0 + lane < b1 ? sv[rhs_src[b0 + lane]] : 0.0;
        for (int b = b0; b < b1; b += 64) {
            const int c = min(64, b1 - b);
            if (lane < c) S[lane] = nxt;
            lane_sync();
            if (b + 64 + lane < b1) nxt = sv[rhs_src[b + 64 + lane]];
            if (lane == 0) {  // reads issued eight at a time, the additions in order
                int i = 0;
                for (; i + 8 <= c; i += 8) {
                    double p[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) p[j] = S[i + j];
#pragma unroll
                    for (int j = 0; j < 8; ++j) acc += p[j];
                }
                for (; i < c; ++i) acc += S[i];
            }
            lane_sync();
        }
        if (lane == 0) rhs[t - lane + l] = acc;
    }
}

// RHS rows with more than kFoldRhs contributions (the ground row: one per
// capacitor and source), one workgroup each: 0.0 + the contributions in
// insertion order through the workgroup's ordered fold.
__global__ __launch_bounds__(256) void k_assemble_long(const int32_t* __restrict__ rows,
                                                       const int32_t* __restrict__ rhs_ptr,
                                                       const int32_t* __restrict__ rhs_src,
                                                       const double* __restrict__ sv,
                                                       double* __restrict__ rhs,
                                                       const int32_t* __restrict__ skip) {
    if (skip && *skip) return;
    const int r = rows[blockIdx.x];
    const double x = rsgpu::fold_long<false>(0.0, rhs_ptr[r], rhs_ptr[r + 1],
                                             [&](int64_t q) { return sv[rhs_src[q]]; });
    if (threadIdx.x == 0) rhs[r] = x;
}

// full = 1: update_transient_state_full (v and i = C (v - v_prev) / dt);
// full = 0: update_transient_state (v only).
__global__ __launch_bounds__(256) void k_cap_state(const rs_device* __restrict__ dev, int ndev,
                                                   const double* __restrict__ x, int xl, double dt,
                                                   int full, double* __restrict__ cap_v,
                                                   double* __restrict__ cap_i,
                                                   uint8_t* __restrict__ has_v,
                                                   uint8_t* __restrict__ has_i,
                                                   const int32_t* __restrict__ gate,
                                                   const double* __restrict__ dt_dev) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ndev || dev[d].kind != RS_DEV_C) return;
    if (gate && !*gate) return;
    if (dt_dev) dt = *dt_dev;
    const double v = xget(x, xl, dev[d].node[0]) - xget(x, xl, dev[d].node[1]);
    if (full) {
        const double vp = has_v[d] ? cap_v[d] : v;
        cap_i[d] = dev[d].value * (v - vp) / dt;
        has_i[d] = 1;
    }
    cap_v[d] = v;
    has_v[d] = 1;
}

}  // namespace

namespace rsgpu {

hipError_t launch_mos_eval(const rs_bsim_model* models, const int32_t* mi, const int32_t* lv,
                           const double* in, int64_t n, double* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mos_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, models, mi,
                       lv, in, n, out);
    return hipGetLastError();
}

hipError_t launch_eval(const rs_device* dev, int ndev, const double* x, int xl, StampArgs a,
                       const double* cap_v, const double* cap_i, const uint8_t* has_v,
                       const uint8_t* has_i, double* sv, hipStream_t s, const int32_t* sbase,
                       uint8_t* gmask) {
    if (ndev <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_eval, dim3((ndev + 255) / 256), dim3(256), 0, s, dev, ndev, x, xl, a,
                       cap_v, cap_i, has_v, has_i, sv, sbase, gmask);
    return hipGetLastError();
}

hipError_t launch_assemble(const int32_t* ax_src, int nnz, const int32_t* rhs_ptr,
                           const int32_t* rhs_src, int n, const double* sv, double* ax, double* rhs,
                           const int32_t* long_rows, int nlong, hipStream_t s, const int32_t* skip) {
    int m = nnz > n ? nnz : n;
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_assemble, dim3((m + 255) / 256), dim3(256), 0, s, ax_src, nnz, rhs_ptr,
                       rhs_src, n, sv, ax, rhs, skip);
    if (nlong > 0)
        hipLaunchKernelGGL(k_assemble_long, dim3(nlong), dim3(256), 0, s, long_rows, rhs_ptr, rhs_src, sv, rhs,
                           skip);
    return hipGetLastError();
}

hipError_t launch_assemble_long(const int32_t* rhs_ptr, const int32_t* rhs_src, const double* sv, double* rhs,
                                const int32_t* long_rows, int nlong, hipStream_t s) {
    if (nlong <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_assemble_long, dim3(nlong), dim3(256), 0, s, long_rows, rhs_ptr, rhs_src, sv, rhs,
                       (const int32_t*)nullptr);
    return hipGetLastError();
}

hipError_t launch_cap_state(const rs_device* dev, int ndev, const double* x, int xl, double dt,
                            int full, double* cap_v, double* cap_i, uint8_t* has_v,
                            uint8_t* has_i, hipStream_t s, const int32_t* gate,
                            const double* dt_dev) {
    if (ndev <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_cap_state, dim3((ndev + 255) / 256), dim3(256), 0, s, dev, ndev, x, xl,
                       dt, full, cap_v, cap_i, has_v, has_i, gate, dt_dev);
    return hipGetLastError();
}

// ---- one Newton iteration's update on the device (run_newton of the engine,
// newton.rs:65-117) for graph-captured transient steps --------------------------
// state: done, converged, iterations, pad; then damping, previous |dx|.
__global__ __launch_bounds__(256) void k_newton_begin(double* __restrict__ t_dev, double dt,
                                                      int32_t* __restrict__ st, double damping) {
    if (threadIdx.x == 0) {
        *t_dev = *t_dev + dt;
        st[0] = 0;
        st[1] = 0;
        st[2] = 0;
        double* sd = reinterpret_cast<double*>(st + 4);
        sd[0] = damping;
        sd[1] = 1.7976931348623157e308;
    }
}

// x <- x_new when every |dx_i| <= abs_tol + rel_tol |x_new_i| (converged),
// else the damped step x + damping (x_new - x), with damping halved (not
// below damping_min) when |dx| grew; done after max_iters. One workgroup;
// |dx| is summed in index order by one thread, as the host loop sums it.
__global__ __launch_bounds__(256) void k_newton_update(int n, double* __restrict__ x,
                                                       const double* __restrict__ xn, double abs_tol,
                                                       double rel_tol, double damping_min,
                                                       int max_iters, int32_t* __restrict__ st) {
    __shared__ int bad;
    __shared__ double dmp;
    if (st[0]) return;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double d = xn[i] - x[i];
        if (!(fabs(d) <= abs_tol + rel_tol * fabs(xn[i]))) bad = 1;
    }
    __syncthreads();
    double* sd = reinterpret_cast<double*>(st + 4);
    if (!bad) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = xn[i];
        if (threadIdx.x == 0) {
            st[2] += 1;
            st[1] = 1;
            st[0] = 1;
        }
        return;
    }
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) {
            const double d = xn[i] - x[i];
            s += d * d;
        }
        const double dxn = sqrt(s);
        double damping = sd[0];
        if (dxn > sd[1] && damping > damping_min) {
            const double d2 = damping * 0.5;
            damping = d2 >= damping_min ? d2 : damping_min;
        }
        sd[0] = damping;
        sd[1] = dxn;
        dmp = damping;
        st[2] += 1;
        if (st[2] >= max_iters) st[0] = 1;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = x[i] + dmp * (xn[i] - x[i]);
}

hipError_t launch_newton_begin(double* t_dev, double dt, int32_t* st, double damping, hipStream_t s) {
    hipLaunchKernelGGL(k_newton_begin, dim3(1), dim3(64), 0, s, t_dev, dt, st, damping);
    return hipGetLastError();
}

hipError_t launch_newton_update(int n, double* x, const double* xn, double abs_tol, double rel_tol,
                                double damping_min, int max_iters, int32_t* st, hipStream_t s) {
    hipLaunchKernelGGL(k_newton_update, dim3(1), dim3(256), 0, s, n, x, xn, abs_tol, rel_tol,
                       damping_min, max_iters, st);
    return hipGetLastError();
}

}  // namespace rsgpu
