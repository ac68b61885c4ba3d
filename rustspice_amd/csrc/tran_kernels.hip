// tran_kernels.hip - the adaptive transient's controller on the device
// (tran_ctl.hpp). Every decision restates the host loop of rs_tran, which
// follows engine.rs:463-602, analysis.rs:147-200 (Milne LTE) and :500-742
// (PI controller, emergency reduction), waveform.rs:567-638 (breakpoints and
// settling), operation for operation.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "ddpow.hpp"
#include "stamp.hpp"
#include "tran_ctl.hpp"

namespace {

using rsgpu::TranCtl;
using rsgpu::TranParams;

__device__ __forceinline__ double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
__device__ __forceinline__ double maxd(double a, double b) { return a < b ? b : a; }  // std::max

// first breakpoint > t (binary search over the ascending list)
__device__ int first_after(const double* bp, int nbp, double t) {
    int lo = 0, hi = nbp;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (bp[mid] > t) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// BreakpointManager::limit_dt (waveform.rs:567-595)
__device__ double limit_dt(const TranParams& p, double tc, double pdt, double mm) {
    const int i = first_after(p.bp, p.nbp, tc);
    if (i < p.nbp) {
        const double ttb = p.bp[i] - tc;
        if (pdt >= ttb) return ttb >= mm ? ttb : mm;
        if (pdt > 0.9 * ttb) return ttb;
        const double rem = ttb - pdt;
        if (rem < 0.1 * pdt && rem < mm * 10.0) return ttb;
    }
    return pdt;
}

// update_settling (waveform.rs:612-626): a breakpoint in (t0, t1] starts 5 settling steps
__device__ void update_settling(const TranParams& p, TranCtl* c, double t0, double t1) {
    const int i = first_after(p.bp, p.nbp, t0);
    if (i < p.nbp && p.bp[i] <= t1) c->settle = 5;
    if (c->settle > 0) c->settle--;
}

__device__ void newton_reset(TranCtl* c, double damping) {
    c->nw_done = 0;
    c->nw_conv = 0;
    c->nw_iters = 0;
    c->nw_damping = damping;
    c->nw_prev = DBL_MAX;
}

// a BE or TRAP Newton solve that did not converge (engine.rs:488-497, 521-530)
__device__ void newton_reject(TranCtl* c, const TranParams& p) {
    c->consec++;
    c->rejected++;
    if (c->consec >= 10) {
        c->status = 2;
        c->phase = rsgpu::kPhaseFinished;
        return;
    }
    c->dt = maxd(c->dt * 0.25, p.min_dt);
    c->phase = rsgpu::kPhaseBegin;
}

constexpr double kKp = 0.7 / 2.0, kKi = 0.4 / 2.0;  // PiControllerConfig::for_order(2)

// which powers the verdict needs: bit 0 (1/e_n)^k_p, bit 1 (1/e_prev)^k_i,
// bit 2 (1/err)^0.5 (the LTE's suggested factor, used on a rejection)
__device__ int32_t powers_needed(const TranCtl* c) {
    const bool accept = c->lte_ok && !(c->err > 1.0);
    if (accept) return 1 | (c->has_prev ? 2 : 0);
    return (c->ctl_consec + 1 < 3 && c->err > 1e-10) ? 4 : 0;
}

__device__ void power_bases(const TranCtl* c, double b[3]) {
    const double en = c->err > 1e-10 ? c->err : 1e-10;
    const double pe = c->prev_err > 1e-10 ? c->prev_err : 1e-10;
    b[0] = 1.0 / en;
    b[1] = 1.0 / pe;
    b[2] = 1.0 / c->err;
}

// AdaptiveStepController::process_lte and the engine's accept / reject
// (analysis.rs:720-742, engine.rs:545-601), given the powers
__device__ void verdict(TranCtl* c, const TranParams& p, const double pw[3]) {
    const bool accept = c->lte_ok && !(c->err > 1.0);
    const double dt = c->dt;
    if (accept) {
        const double pf = pw[0];
        const double inf = c->has_prev ? pw[1] : 1.0;
        const double fac = clampd(pf * inf * 0.9, 0.1, 2.0);
        c->has_prev = 1;
        c->prev_err = c->err;
        c->ctl_consec = 0;
        const double dt_new = clampd(dt * fac, p.min_dt, p.max_dt);
        const double tn = c->t + dt;
        update_settling(p, c, c->t, tn);
        c->accept_now = 1;
        c->cp_acc = 1;
        c->npoints++;
        c->t = tn;
        c->accepted++;
        c->consec = 0;
        c->dt = clampd(dt_new, p.min_dt, p.max_dt);
        c->phase = c->t < p.tstop ? rsgpu::kPhaseBegin : rsgpu::kPhaseFinished;
        return;
    }
    c->ctl_consec++;
    double dt_new;
    if (c->ctl_consec >= 3) {
        dt_new = c->ctl_consec <= 5 ? dt * 0.25 : (c->ctl_consec <= 10 ? dt * 0.1 : p.min_dt);
    } else {
        const double f = c->err > 1e-10 ? clampd(pw[2] * 0.9, 0.1, 2.0) : 2.0;
        dt_new = clampd(dt * f, p.min_dt, p.max_dt);
    }
    c->rejected++;
    c->consec++;
    if (c->consec >= 10) {
        c->status = 2;
        c->phase = rsgpu::kPhaseFinished;
        return;
    }
    c->dt = clampd(dt_new, p.min_dt, p.max_dt);
    c->phase = rsgpu::kPhaseBegin;
}

__global__ void k_tctl_pre(TranCtl* __restrict__ c, TranParams p) {
    if (threadIdx.x != 0) return;
    c->accept_now = 0;
    c->cp_begin = 0;
    c->cp_be = 0;
    c->cp_acc = 0;
    const int32_t ph = c->phase;
    if (ph == rsgpu::kPhaseBe || ph == rsgpu::kPhaseTrap) return;
    if (ph != rsgpu::kPhaseBegin) {  // finished, waiting for the host, or resuming
        c->nw_done = 1;
        return;
    }
    if (!(c->t < p.tstop)) {
        c->phase = rsgpu::kPhaseFinished;
        c->nw_done = 1;
        return;
    }
    double dt = limit_dt(p, c->t, c->dt, p.min_dt);
    if (c->settle > 0) dt = maxd(dt * 0.1, p.min_dt);
    if (c->t + dt > p.tstop) dt = p.tstop - c->t;
    c->dt = dt;
    c->dt_step = dt;
    c->t_target = c->t + dt;
    c->phase = rsgpu::kPhaseBe;
    c->mode = rsgpu::MODE_TRAN_BE;
    newton_reset(c, p.damping);
    c->cp_begin = 1;  // the BE solve starts from the accepted x
}

// After the slot's Newton iteration (one workgroup of 256 threads).
__global__ __launch_bounds__(256) void k_tctl_post(TranCtl* __restrict__ c, TranParams p,
                                                   const double* __restrict__ x_be,
                                                   const double* __restrict__ x) {
    __shared__ double s_best[256];
    __shared__ int s_go;
    const int tid = threadIdx.x;
    if (tid == 0) {
        c->cp_begin = 0;  // the slot's start-of-attempt copy has been made
        const int32_t ph = c->phase;
        s_go = 0;
        if (ph == rsgpu::kPhaseResume) {
            verdict(c, p, c->pw_out);
            c->pw_need = 0;
        } else if ((ph == rsgpu::kPhaseBe || ph == rsgpu::kPhaseTrap) && c->nw_done) {
            c->iters_total += c->nw_iters;
            if (!c->nw_conv) {
                newton_reject(c, p);
            } else if (ph == rsgpu::kPhaseBe) {
                c->cp_be = 1;     // x_be = x
                c->cp_begin = 1;  // the TRAP solve starts from the accepted x again
                c->phase = rsgpu::kPhaseTrap;
                c->mode = rsgpu::MODE_TRAN_TRAP;
                newton_reset(c, p.damping);
            } else {
                s_go = 1;  // TRAP converged: LTE and verdict
            }
        }
    }
    __syncthreads();
    if (!s_go) return;
    // estimate_lte_milne (analysis.rs:147-200): the largest ratio (a maximum,
    // so any reduction order gives the same value; NaN ratios never win)
    double best = 0.0;
    for (int i = tid; i < p.n; i += blockDim.x) {
        const double be = x_be[i], tr = x[i];
        if (fabs(be) < 1e-30 && fabs(tr) < 1e-30) continue;
        const double lte = fabs(tr - be) / 3.0;
        const double ab = fabs(be), at = fabs(tr);
        const double tol = p.abs_tol + p.rel_tol * (ab >= at ? ab : at);
        const double ratio = tol > 0.0 ? lte / tol : 0.0;
        if (ratio > best) best = ratio;
    }
    s_best[tid] = best;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w && s_best[tid + w] > s_best[tid]) s_best[tid] = s_best[tid + w];
        __syncthreads();
    }
    if (tid != 0) return;
    c->err = s_best[0];
    c->lte_ok = c->err <= 1.0;
    const int32_t need = powers_needed(c);
    double b[3], pw[3] = {1.0, 1.0, 1.0};
    power_bases(c, b);
    const double ex[3] = {kKp, kKi, 0.5};
    bool any_near = false;
    for (int j = 0; j < 3; j++)
        if (need & (1 << j)) {
            bool near = false;
            pw[j] = rsgpu::pow_dd(b[j], ex[j], &near, 0.02);
            any_near = any_near || near;
        }
    if (any_near) {  // the host evaluates the powers with libm, then the verdict resumes
        for (int j = 0; j < 3; j++) c->pw_in[j] = b[j];
        c->pw_need = need;
        c->phase = rsgpu::kPhaseHostPow;
        return;
    }
    verdict(c, p, pw);
}

__global__ void k_gated_copy(double* __restrict__ dst, const double* __restrict__ src, int64_t n,
                             const int32_t* __restrict__ gate) {
    if (!*gate) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_tstore(const TranCtl* __restrict__ c, const double* __restrict__ x, double* __restrict__ xa,
                         double* __restrict__ ring_t, double* __restrict__ ring_x, int64_t n) {
    if (!c->cp_acc) return;
    const int64_t slot = (c->npoints - 1) % rsgpu::kTranRing;
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 == 0) ring_t[slot] = c->t;
    for (int64_t i = i0; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        xa[i] = x[i];
        ring_x[slot * n + i] = x[i];
    }
}

inline unsigned copy_grid(int64_t n) {
    const int64_t g = (n + 255) / 256;
    return (unsigned)(g < 1024 ? (g > 0 ? g : 1) : 1024);
}

}  // namespace

namespace rsgpu {

hipError_t launch_tctl_pre(TranCtl* c, TranParams p, hipStream_t s) {
    hipLaunchKernelGGL(k_tctl_pre, dim3(1), dim3(64), 0, s, c, p);
    return hipGetLastError();
}

hipError_t launch_tctl_post(TranCtl* c, TranParams p, const double* x_be, const double* x, hipStream_t s) {
    hipLaunchKernelGGL(k_tctl_post, dim3(1), dim3(256), 0, s, c, p, x_be, x);
    return hipGetLastError();
}

hipError_t launch_gated_copy(double* dst, const double* src, int64_t n, const int32_t* gate, hipStream_t s) {
    hipLaunchKernelGGL(k_gated_copy, dim3(copy_grid(n)), dim3(256), 0, s, dst, src, n, gate);
    return hipGetLastError();
}

hipError_t launch_tstore(const TranCtl* c, const double* x, double* xa, double* ring_t, double* ring_x,
                         int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_tstore, dim3(copy_grid(n)), dim3(256), 0, s, c, x, xa, ring_t, ring_x, n);
    return hipGetLastError();
}

}  // namespace rsgpu
