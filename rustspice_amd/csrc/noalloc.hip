// noalloc.hip - rs_solve_device_noalloc: the device solve of rs_solve_device
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops). Its kernels are the ones rs_solve_device
// launches; the solve vector they need besides the inputs and outputs (the
// permuted solution of the supernodal solve, the global solve vector of the
// level-set solve) lives in a file-scope device array instead of the
// handle's own buffer, so the call makes no allocation of any kind.
#include <hip/hip_runtime.h>

#include <mutex>

#include "solver_impl.hpp"

namespace {

// 2^21 doubles (16 MiB) per GPU, held until the process ends. Every call on
// a GPU shares that GPU's copy: calls that overlap on one GPU race.
constexpr int64_t kScratchDoubles = int64_t(1) << 21;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

// The device address of this GPU's copy of `scratch` (host code must not
// use the array's name as a device pointer), looked up once per device with
// that device current.
hipError_t scratch_ptr(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        if ((e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) != hipSuccess) return e;
        g_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_ptr[dev];
    return hipSuccess;
}

}  // namespace

extern "C" {

int64_t rs_noalloc_max_n(void) { return kScratchDoubles; }

int32_t rs_solve_device_noalloc(rs_solver* h, const double* rhs_dev, double* x_dev, void* stream) {
    double* tmp = nullptr;
    const hipError_t e = scratch_ptr(&tmp);
    return rsgpu::solver_solve_noalloc(h, rhs_dev, x_dev, (hipStream_t)stream, tmp, kScratchDoubles, e);
}

}  // extern "C"
