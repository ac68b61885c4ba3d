// ac_kernels.hip - the AC sweep's device evaluation and per-frequency dense
// complex solve (ac.hpp).
#include <hip/hip_runtime.h>

#include "ac.hpp"
#include "device_eval.hpp"

namespace {

using rsgpu::AcDev;

constexpr int kT = 256;
constexpr double kPi = 3.14159265358979323846;  // std::f64::consts::PI

struct Cx {
    double re, im;
};

__device__ __forceinline__ Cx cadd(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cx csub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Cx cmul(Cx a, Cx b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ Cx cdiv(Cx a, Cx b) {
    const double ns = b.re * b.re + b.im * b.im;
    const double re = a.re * b.re + a.im * b.im;
    const double im = a.im * b.re - a.re * b.im;
    return {re / ns, im / ns};
}
__device__ __forceinline__ double cnorm(Cx a) { return hypot(a.re, a.im); }
__device__ __forceinline__ Cx ld(const double* p, int64_t i) { return {p[2 * i], p[2 * i + 1]}; }
__device__ __forceinline__ void st(double* p, int64_t i, Cx v) {
    p[2 * i] = v.re;
    p[2 * i + 1] = v.im;
}

// stamp.rs:1525-1747: the values the AC stamps read from each device at the
// DC solution (gmin 1e-12 throughout, as stamp_diode_ac / stamp_mos_ac fix it)
__global__ void k_ac_eval(const rs_device* __restrict__ dev, int64_t ndev,
                          const rs_bsim_model* __restrict__ models, const double* __restrict__ x,
                          int xl, double* __restrict__ vals, uint8_t* __restrict__ mask) {
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= ndev) return;
    const rs_device v = dev[d];
    double o[32];
    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    uint8_t m = 0;
    rsgpu::StampArgs a{};
    a.mode = rsgpu::MODE_DC;
    a.gmin = 1e-12;
    a.source_scale = 1.0;
    a.models = models;
    switch (v.kind) {
    case RS_DEV_R:
    case RS_DEV_D:
        rsgpu::eval_device(v, (int)d, x, xl, a, nullptr, nullptr, nullptr, nullptr, o);
        g0 = o[0];
        break;
    case RS_DEV_C:
        g0 = v.value;
        break;
    case RS_DEV_M:
        if ((int)v.p[RS_MP_LEVEL] == 1) {
            rsgpu::eval_device(v, (int)d, x, xl, a, nullptr, nullptr, nullptr, nullptr, o);
            g0 = o[0];  // gds floored at gmin
            g1 = o[4];  // gm; level 1 has gmbs = 0, never stamped
        } else {
            m = rsgpu::eval_bsim(v, x, xl, a, o) & 1;
            g0 = o[0];
            g1 = o[4];
            g2 = o[8];
        }
        break;
    default:
        break;
    }
    vals[3 * d] = g0;
    vals[3 * d + 1] = g1;
    vals[3 * d + 2] = g2;
    mask[d] = m;
}

// One workgroup per frequency: assemble, ComplexDenseSolver::solve
// (complex_solver.rs:50-136), solution left in x.
__global__ __launch_bounds__(kT) void k_ac_solve(AcDev p, int64_t f0, double* __restrict__ a_all,
                                                 double* __restrict__ x_all,
                                                 int32_t* __restrict__ status) {
    __shared__ double s_val[kT];
    __shared__ int s_idx[kT];
    __shared__ int s_piv;
    __shared__ int s_fail;
    const int tid = threadIdx.x;
    const int64_t n = p.n;
    const int64_t f = f0 + blockIdx.x;
    double* A = a_all + (int64_t)blockIdx.x * n * n * 2;
    double* X = x_all + f * n * 2;
    const double omega = 2.0 * kPi * p.freqs[f];
    for (int64_t q = tid; q < n * n; q += kT) st(A, q, Cx{0.0, 0.0});
    for (int64_t i = tid; i < n; i += kT) st(X, i, ld(p.rhs, i));
    __syncthreads();
    // each dense entry: 0 + its contributions in insertion order
    for (int32_t e = tid; e < p.nent; e += kT) {
        Cx acc{0.0, 0.0};
        for (int32_t c = p.ent_ptr[e]; c < p.ent_ptr[e + 1]; ++c) {
            const int32_t w = p.con[c];
            const double b = p.vals[w >> 2];
            Cx t;
            if (w & 1) {
                const double y = omega * b;
                t = Cx{0.0, (w & 2) ? -y : y};
            } else {
                t = Cx{(w & 2) ? -b : b, 0.0};
            }
            acc = cadd(acc, t);
        }
        st(A, p.ent_pos[e], acc);
    }
    if (tid == 0) s_fail = 0;
    __syncthreads();
    for (int64_t k = 0; k < n; ++k) {
        // first row of the largest modulus in column k (NaN moduli never win)
        double best = -1.0;
        int bi = (int)n;
        for (int64_t i = k + tid; i < n; i += kT) {
            const double v = cnorm(ld(A, i * n + k));
            if (v > best) { best = v; bi = (int)i; }
        }
        s_val[tid] = best;
        s_idx[tid] = bi;
        __syncthreads();
        for (int w = kT / 2; w > 0; w >>= 1) {
            if (tid < w) {
                const double v2 = s_val[tid + w];
                const int i2 = s_idx[tid + w];
                if (v2 > s_val[tid] || (v2 == s_val[tid] && i2 < s_idx[tid])) {
                    s_val[tid] = v2;
                    s_idx[tid] = i2;
                }
            }
            __syncthreads();
        }
        if (tid == 0) {
            const double dk = cnorm(ld(A, k * n + k));
            s_piv = (dk != dk) ? (int)k : s_idx[0];  // a NaN diagonal is never replaced
        }
        __syncthreads();
        const int64_t r = s_piv;
        if (r != k) {
            for (int64_t j = tid; j < n; j += kT) {
                const Cx t = ld(A, k * n + j);
                st(A, k * n + j, ld(A, r * n + j));
                st(A, r * n + j, t);
            }
            if (tid == 0) {
                const Cx t = ld(X, k);
                st(X, k, ld(X, r));
                st(X, r, t);
            }
        }
        __syncthreads();
        const Cx piv = ld(A, k * n + k);
        if (cnorm(piv) < 1e-30) {
            if (tid == 0) s_fail = 1;
            break;
        }
        for (int64_t i = k + 1 + tid; i < n; i += kT) st(A, i * n + k, cdiv(ld(A, i * n + k), piv));
        __syncthreads();
        const int64_t m = n - k - 1;
        for (int64_t t = tid; t < m * m; t += kT) {
            const int64_t i = k + 1 + t / m, j = k + 1 + t % m;
            st(A, i * n + j, csub(ld(A, i * n + j), cmul(ld(A, i * n + k), ld(A, k * n + j))));
        }
        const Cx xk = ld(X, k);
        for (int64_t i = k + 1 + tid; i < n; i += kT) st(X, i, csub(ld(X, i), cmul(ld(A, i * n + k), xk)));
        __syncthreads();
    }
    __syncthreads();
    if (!s_fail) {
        for (int64_t i = n - 1; i >= 0; --i) {
            // row i's products a[i][j] * x[j] in place, then the ordered subtractions
            for (int64_t j = i + 1 + tid; j < n; j += kT) st(A, i * n + j, cmul(ld(A, i * n + j), ld(X, j)));
            __syncthreads();
            if (tid == 0) {
                Cx xi = ld(X, i);
                for (int64_t j = i + 1; j < n; ++j) xi = csub(xi, ld(A, i * n + j));
                const Cx piv = ld(A, i * n + i);
                if (cnorm(piv) < 1e-30) s_fail = 1;
                else st(X, i, cdiv(xi, piv));
            }
            __syncthreads();
            if (s_fail) break;
        }
    }
    if (tid == 0) status[f] = s_fail ? 0 : 1;
}

}  // namespace

namespace rsgpu {

hipError_t launch_ac_eval(const rs_device* dev, int64_t ndev, const rs_bsim_model* models,
                          const double* x_dc, int64_t nx, double* vals, uint8_t* mask,
                          hipStream_t s) {
    if (ndev <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ac_eval, dim3((unsigned)((ndev + kT - 1) / kT)), dim3(kT), 0, s, dev, ndev,
                       models, x_dc, (int)nx, vals, mask);
    return hipGetLastError();
}

hipError_t launch_ac_solve(const AcDev& p, int64_t f0, int64_t nf, double* a, double* x,
                           int32_t* status, hipStream_t s) {
    if (nf <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ac_solve, dim3((unsigned)nf), dim3(kT), 0, s, p, f0, a, x, status);
    return hipGetLastError();
}

}  // namespace rsgpu
