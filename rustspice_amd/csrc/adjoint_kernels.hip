// adjoint_kernels.hip - DC sensitivities by the adjoint method (no reference
// counterpart; the forward path it differentiates is engine.rs's DC OP).
//
// At a converged operating point F(x*, p) = 0, with J = dF/dx(x*):
//   J^T lambda = e_probe,   dV_probe/dp = -lambda^T dF/dp.
// k_solve_t solves J^T lambda = e with the stored L/U of J (a single-block
// plan: J = Pout U^-1 L^-1 Pin inverted, so J^-T = Pin^T L^-T U^-T Pout^T);
// k_adjoint_contract forms dF/dp per device in closed form and contracts it
// with lambda, one device per thread.
#include <hip/hip_runtime.h>

#include "../../include/rs_circuit.h"
#include "device_eval.hpp"
#include "plan.hpp"

namespace {

constexpr int kT = 256, kWave = 64;

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// One workgroup. U^T is lower triangular: column k of U gives row k of U^T,
// so y_k = (t_k - sum_j U(j,k) y_j) / U(k,k) in the refactor's column levels
// (column k waits for every j in U(:,k)). L^T is unit upper: z_k = y_k -
// sum_i L(i,k) z_i over the L^T levels (column k waits for every row of
// L(:,k)). One wavefront per column, lanes over its entries.
__global__ __launch_bounds__(kT) void k_solve_t(const rsgpu::PlanDev* __restrict__ pd,
                                               const int32_t* __restrict__ ip,
                                               const double* __restrict__ vals,
                                               const double* __restrict__ e,
                                               const int32_t* __restrict__ lt_ptr,
                                               const int32_t* __restrict__ lt_cols, int n_lt,
                                               double* __restrict__ T, double* __restrict__ lam) {
    const rsgpu::PlanDev d = *pd;
    const int n = d.n, lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    const int32_t* perm_in = ip + d.o_perm_in;
    const int32_t* perm_out = ip + d.o_perm_out;
    const int32_t* lcp = ip + d.o_lcp;
    const int32_t* lri = ip + d.o_lri;
    const int32_t* ucp = ip + d.o_ucp;
    const int32_t* uri = ip + d.o_uri;
    const int32_t* flev_ptr = ip + d.o_flev_ptr;
    const int32_t* flev_cols = ip + d.o_flev_cols;
    const double* lx = vals + d.o_lx;
    const double* ux = vals + d.o_ux;
    const double* ud = vals + d.o_ud;
    for (int i = threadIdx.x; i < n; i += blockDim.x) T[perm_out[i]] = e[i];
    __syncthreads();
    for (int lev = 0; lev < d.nflev; ++lev) {
        for (int q = flev_ptr[lev] + wid; q < flev_ptr[lev + 1]; q += nw) {
            const int k = flev_cols[q];
            double s = 0.0;
            for (int u = ucp[k] + lane; u < ucp[k + 1] - 1; u += kWave) s += ux[u] * T[uri[u]];
            s = wave_sum(s);
            if (lane == 0) T[k] = (T[k] - s) / ud[k];
        }
        __syncthreads();
    }
    for (int lev = 0; lev < n_lt; ++lev) {
        for (int q = lt_ptr[lev] + wid; q < lt_ptr[lev + 1]; q += nw) {
            const int k = lt_cols[q];
            double s = 0.0;
            for (int l = lcp[k] + lane; l < lcp[k + 1]; l += kWave) s += lx[l] * T[lri[l]];
            s = wave_sum(s);
            if (lane == 0) T[k] = T[k] - s;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) lam[i] = T[perm_in[i]];
}

__device__ __forceinline__ double at(const double* v, int xl, int i) { return i < xl ? v[i] : 0.0; }

// r = b - A x with A the unmerged CSC through its row map (rp: row pointers,
// rq: CSC entry of each row entry, in column order); x beyond xl reads 0.
// One wavefront per row, lanes over its entries and a butterfly sum, so a
// dense row (the ground row of a large mesh) costs a few loads per lane
// instead of a serial walk. Deterministic: the summation tree is fixed.
__global__ __launch_bounds__(kT) void k_resid_n(const int32_t* __restrict__ rp,
                                               const int32_t* __restrict__ rc,
                                               const int32_t* __restrict__ rq,
                                               const double* __restrict__ ax,
                                               const double* __restrict__ b,
                                               const double* __restrict__ x, int xl,
                                               double* __restrict__ r, int n) {
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (i >= n) return;
    double s = 0.0;
    for (int e = rp[i] + lane; e < rp[i + 1]; e += kWave) s += ax[rq[e]] * at(x, xl, rc[e]);
    s = wave_sum(s);
    if (lane == 0) r[i] = b[i] - s;
}

// r = e - J^T lam with J the unmerged CSC (one wavefront per column): the
// residual of the transposed system for iterative refinement.
__global__ __launch_bounds__(kT) void k_resid_t(const int32_t* __restrict__ ap,
                                               const int32_t* __restrict__ ai,
                                               const double* __restrict__ ax,
                                               const double* __restrict__ e,
                                               const double* __restrict__ lam,
                                               double* __restrict__ r, int n) {
    const int j = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (j >= n) return;
    double s = 0.0;
    for (int q = ap[j] + lane; q < ap[j + 1]; q += kWave) {
        const int row = ai[q];
        if (row < n) s += ax[q] * lam[row];
    }
    s = wave_sum(s);
    if (lane == 0) r[j] = e[j] - s;
}

// lam += delta; out[0] = max |delta|, out[1] = max |lam| (one workgroup)
__global__ __launch_bounds__(kT) void k_refine_update(double* __restrict__ lam,
                                                      const double* __restrict__ delta, int n,
                                                      double* __restrict__ out) {
    __shared__ double sd[kT], sl[kT];
    double md = 0.0, ml = 0.0;
    for (int i = threadIdx.x; i < n; i += kT) {
        const double v = lam[i] + delta[i];
        lam[i] = v;
        md = fmax(md, fabs(delta[i]));
        ml = fmax(ml, fabs(v));
    }
    sd[threadIdx.x] = md;
    sl[threadIdx.x] = ml;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sd[threadIdx.x] = fmax(sd[threadIdx.x], sd[threadIdx.x + w]);
            sl[threadIdx.x] = fmax(sl[threadIdx.x], sl[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sd[0];
        out[1] = sl[0];
    }
}

// dV_probe/dp per device: R -> d/dR, V and I sources -> d/dvalue,
// MOSFET -> d/dW (level 1: the stamped current is linear in W), diode -> d/dIs,
// capacitor -> 0 (open in DC).
// -lambda^T dF/dp of one step, per device (one thread each), written or
// added to out (acc). Sources scale with the source-stepping factor of the
// solve; PULSE sources contribute nothing (their stamp does not use the
// value, and DC does not stamp them). Capacitors: in a TRAP step k the
// current is I = (2C/dt)(v_k - v_{k-1}) - i_{k-1}, with the stored current
// i_{k-1} = C (v_{k-1} - v_{k-2}) / dt from the step before (0 after the DC
// OP: xp2 null), so dI/dC = (2/dt)(v_k - v_{k-1}) - (v_{k-1} - v_{k-2})/dt.
__global__ void k_adjoint_contract(const rs_device* __restrict__ dev, int ndev,
                                   const int32_t* __restrict__ aux_of, const double* __restrict__ x,
                                   int xl, const double* __restrict__ lam, double gmin, double scale,
                                   int mode, const double* __restrict__ xp,
                                   const double* __restrict__ xp2, double dt, int acc,
                                   double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ndev) return;
    const rs_device v = dev[k];
    const int a = v.node[0], b = v.node[1];
    double r = 0.0;
    switch (v.kind) {
    case RS_DEV_R: {
        const double vab = at(x, xl, a) - at(x, xl, b);
        r = vab * (lam[a] - lam[b]) / (v.value * v.value);
        break;
    }
    case RS_DEV_V:
        if (!(v.flags & RS_F_WAVEFORM)) r = aux_of[k] >= 0 ? lam[aux_of[k]] * scale : 0.0;
        break;
    case RS_DEV_I:
        if (!(v.flags & RS_F_WAVEFORM)) r = -(lam[a] - lam[b]) * scale;
        break;
    case RS_DEV_C:
        if (mode == rsgpu::MODE_TRAN_TRAP && xp) {
            const double vk = x[a] - x[b], v1 = xp[a] - xp[b];
            double didc = (2.0 / dt) * (vk - v1);
            if (xp2) didc -= (v1 - (xp2[a] - xp2[b])) / dt;
            r = -(lam[a] - lam[b]) * didc;
        }
        break;
    case RS_DEV_D: {  // F_a += id(Is), id linear in Is
        const double vt = 0.02585 * v.p[1];
        const double vd = at(x, xl, a) - at(x, xl, b);
        const double id_per_is = exp(vd / vt) - 1.0;
        r = -(lam[a] - lam[b]) * id_per_is;
        break;
    }
    case RS_DEV_M: {  // F_d += ids_s, F_s -= ids_s
        // The Newton stamp at x* reproduces the current: gds (vd - vs) +
        // gm (vg - vs) + ieq = ids_s (gds is never below the gmin floor here).
        double o[rsgpu::kSlotsPerDevice];
        const rsgpu::StampArgs sa{0.0, 0.0, gmin, 1.0, rsgpu::MODE_DC, 0};
        rsgpu::eval_device(v, k, x, xl, sa, nullptr, nullptr, nullptr, nullptr, o);
        const double vd = at(x, xl, v.node[0]), vg = at(x, xl, v.node[1]), vs = at(x, xl, v.node[2]);
        const double ids = o[0] * (vd - vs) + o[4] * (vg - vs) + o[9];
        r = -(lam[v.node[0]] - lam[v.node[2]]) * ids / v.p[RS_MP_W];
        break;
    }
    default:
        break;
    }
    out[k] = acc ? out[k] + r : r;
}

// Right-hand side of the adjoint equation of TRAP step k (or of the DC OP,
// k = 0): rhs = e_probe [k = N] - sum_j (dF_j/dx_k)^T lam_j over the two
// later steps, whose capacitor currents read v_k:
//   dI_{k+1}/dv_k = -2C/dt - C/dt [k >= 1],   dI_{k+2}/dv_k = C/dt.
// One wavefront per unknown, lanes over the capacitors of its node
// (cap_ptr/cap_dev, the node's sign in bit 31) and a fixed summation tree.
__global__ __launch_bounds__(kT) void k_tadj_rhs(const rs_device* __restrict__ dev,
                                                const int32_t* __restrict__ cap_ptr,
                                                const int32_t* __restrict__ cap_dev,
                                                int node_count, const double* __restrict__ lam1,
                                                const double* __restrict__ lam2, double dt,
                                                int k_ge1, int probe, int n,
                                                double* __restrict__ rhs) {
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (i >= n) return;
    double s = 0.0;
    if (i < node_count) {
        for (int q = cap_ptr[i] + lane; q < cap_ptr[i + 1]; q += kWave) {
            const int w = cap_dev[q];
            const rs_device& v = dev[w & 0x7FFFFFFF];
            const int a = v.node[0], b = v.node[1];
            const double cdt = v.value / dt;
            double coef = 0.0;
            if (lam1) coef += (2.0 * cdt + (k_ge1 ? cdt : 0.0)) * (lam1[a] - lam1[b]);
            if (lam2) coef -= cdt * (lam2[a] - lam2[b]);
            s += (w < 0) ? -coef : coef;
        }
    }
    s = wave_sum(s);
    if (lane == 0) rhs[i] = (i == probe) ? s + 1.0 : s;
}

}  // namespace

namespace rsgpu {

// workgroups of kT threads holding one wavefront per item
static unsigned wave_blocks(int items) {
    const long long waves_per_block = kT / kWave;
    return (unsigned)((items + waves_per_block - 1) / waves_per_block);
}

hipError_t launch_solve_t(const PlanDev* pd, const int32_t* ip, const double* vals, const double* e,
                          const int32_t* lt_ptr, const int32_t* lt_cols, int n_lt, double* T,
                          double* lam, hipStream_t s) {
    hipLaunchKernelGGL(k_solve_t, dim3(1), dim3(kT), 0, s, pd, ip, vals, e, lt_ptr, lt_cols, n_lt,
                       T, lam);
    return hipGetLastError();
}

hipError_t launch_resid_t(const int32_t* ap, const int32_t* ai, const double* ax, const double* e,
                          const double* lam, double* r, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_resid_t, dim3(wave_blocks(n)), dim3(kT), 0, s, ap, ai, ax, e, lam, r, n);
    return hipGetLastError();
}

hipError_t launch_resid_n(const int32_t* rp, const int32_t* rc, const int32_t* rq, const double* ax,
                          const double* b, const double* x, int xl, double* r, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_resid_n, dim3(wave_blocks(n)), dim3(kT), 0, s, rp, rc, rq, ax, b, x, xl, r, n);
    return hipGetLastError();
}

hipError_t launch_refine_update(double* lam, const double* delta, int n, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_refine_update, dim3(1), dim3(kT), 0, s, lam, delta, n, out);
    return hipGetLastError();
}

hipError_t launch_adjoint_contract(const rs_device* dev, int ndev, const int32_t* aux_of,
                                   const double* x, int xl, const double* lam, double gmin,
                                   double scale, int mode, const double* xp, const double* xp2,
                                   double dt, int acc, double* out, hipStream_t s) {
    if (ndev <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_adjoint_contract, dim3((ndev + kT - 1) / kT), dim3(kT), 0, s, dev, ndev,
                       aux_of, x, xl, lam, gmin, scale, mode, xp, xp2, dt, acc, out);
    return hipGetLastError();
}

hipError_t launch_tadj_rhs(const rs_device* dev, const int32_t* cap_ptr, const int32_t* cap_dev,
                           int node_count, const double* lam1, const double* lam2, double dt,
                           int k_ge1, int probe, int n, double* rhs, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tadj_rhs, dim3(wave_blocks(n)), dim3(kT), 0, s, dev, cap_ptr, cap_dev,
                       node_count, lam1, lam2, dt, k_ge1, probe, n, rhs);
    return hipGetLastError();
}

}  // namespace rsgpu
