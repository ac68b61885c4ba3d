// solver_impl.hpp - internal state of rs_solver shared by the C ABI
// (solver.cpp) and the Newton/TRAN drivers (engine.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <type_traits>
#include <vector>

#include "../../include/rsgpu.h"
#include "plan.hpp"
#include "sp.hpp"

namespace rsgpu {
// mode: bit 0 LDS, bit 1 right-looking schedule (rl_nv value slots)
hipError_t launch_refactor(const PlanDev* descs, int nmat, const int32_t* ip, const double* ax,
                           double* vals, double* wpool, int max_n, int mode, int threads,
                           hipStream_t stream, int rl_nv = 0, const int32_t* skip = nullptr);
// tvec (one matrix, global path only): the solve vector in place of
// wpool + o_tvec
hipError_t launch_solve(const PlanDev* descs, int nmat, const int32_t* ip, const double* ax,
                        const double* vals, const double* rhs, double* x, double* wpool, int max_n,
                        bool lds, int threads, hipStream_t stream, const int32_t* skip = nullptr,
                        double* tvec = nullptr);
size_t device_max_lds();
// A^T y = e of a single-block plan on its L^T levels (adjoint_kernels.hip)
hipError_t launch_solve_t(const PlanDev* pd, const int32_t* ip, const double* vals, const double* e,
                          const int32_t* lt_ptr, const int32_t* lt_cols, int n_lt, double* T,
                          double* lam, hipStream_t s);
// solve mode 3: flat pipelined level sweeps of one matrix (PlanDev::sq_nlev > 0)
size_t solve_lp_lds(int n, int maxent, int nlev);
hipError_t launch_solve_lp(const PlanDev* descs, const int32_t* ip, const double* vals, const double* rhs,
                           double* x, int n, int maxent, int nlev, int threads, hipStream_t stream,
                           const int32_t* skip = nullptr);
hipError_t launch_sp_refactor(const SpSolverDev* sd, const PlanDev* pd, const int32_t* ip,
                              const int32_t* sip, const SpOpWord* words, const double* ax,
                              double* vals, int nslots, int threads, hipStream_t s,
                              const int32_t* skip = nullptr);
hipError_t launch_sp_solve(const SpSolverDev* sd, const PlanDev* pd, const int32_t* ip,
                           const int32_t* sip, const SpOpWord* words, const double* ax,
                           const double* vals, const double* rhs, double* x, int nslots, int threads,
                           hipStream_t s, const int32_t* skip = nullptr);
size_t sp_solver_static_lds();
hipError_t launch_growth(const double* ax, int nnz, const double* ux, int nu, double* out, hipStream_t s);

struct BbdImpl;
struct DenseImpl;
struct SnImpl;
// SparseLU matrices at least this large with a structurally symmetric
// pattern use the supernodal factor and solve (rs_set_supernodal mode 1).
constexpr int64_t kSnAutoN = 20000;

enum { ST_EMPTY = 0, ST_ANALYZED = 1, ST_FACTORED = 2 };
constexpr int kThreads = 256;
constexpr int kMaxThreads = 1024;  // largest launch the autotuner tries (16 wavefronts)

// RSGPU_POISON=1 (a checking aid): fresh FP64 buffers are filled with NaN, so
// a value read before it is written shows up instead of reading whatever the
// allocator's memory held.
bool poison_fp64_buffers();

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t cnt) {
        if (cnt <= n && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void**)&p, (cnt ? cnt : 1) * sizeof(T));
        if (e == hipSuccess && std::is_same<T, double>::value && poison_fp64_buffers())
            e = hipMemset(p, 0xFF, (cnt ? cnt : 1) * sizeof(T));
        if (e == hipSuccess) n = cnt;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
};

// Factor a pattern whose values are already in device memory: refactor on
// the GPU when the pattern is the factored one, else a host first factor.
int32_t solver_factor_dev(struct ::rs_solver* h, const i64* ap, i64 apl, const i64* ai, i64 ail,
                          const double* ax_dev, hipStream_t s);
int32_t solver_solve_dev(struct ::rs_solver* h, const double* rhs_dev, double* x_dev,
                         hipStream_t s);
// The symbolic analysis of a pattern when it is not the analyzed one (no
// factor); each analysis advances rs_solver::analyses.
int32_t solver_analyze_dev(struct ::rs_solver* h, const i64* ap, i64 apl, const i64* ai, i64 ail);
}  // namespace rsgpu

using rsgpu::i64;

struct rs_solver {
    int kind = RS_SOLVER_NATIVEKLU;
    rsgpu::BbdImpl* bbd = nullptr;      // kind RS_SOLVER_BBD
    rsgpu::DenseImpl* dense = nullptr;  // kind RS_SOLVER_DENSE
    rsgpu::SnImpl* sn = nullptr;        // supernodal factor and solve of the analyzed pattern
    int sn_mode = 1;                    // 0 off, 1 large SparseLU patterns, 2 every SparseLU pattern
    i64 n = 0;
    double pivot_tol = 0.001;
    int ordering = 0;
    bool use_btf = true;
    i64 btf_min_blocks = 2;
    int state = rsgpu::ST_EMPTY;
    std::vector<i64> last_ap, last_ai;
    rsgpu::HostFactor hf;
    rsgpu::PlanHost ph;
    rsgpu::PlanDev pd{};
    bool plan_on_device = false;
    i64 factor_count = 0, refactor_count = 0;
    int64_t analyses = 0;  // symbolic analyses run (a new one replaces sn)
    double rcond = 0.0;
    bool last_parallel = false;
    i64 threads_hint = 0;
    std::string err;
    size_t plan_bytes = 0;
    int ref_mode = 0;  // bit 0 LDS, bit 1 right-looking schedule, 4 step program
    int sol_mode = 0;  // 0 solve vector in global memory, 1 in LDS, 2 step program, 3 pipelined levels
    // launch shapes of k_refactor / k_solve: chosen by timing every candidate
    // on the first device refactor and solve of a pattern (autotune), or set
    int ref_threads = rsgpu::kThreads, sol_threads = rsgpu::kThreads;
    bool autotune = true, tuned_ref = false, tuned_sol = false, forced = false;
    float tune_ref_us = 0.0f, tune_sol_us = 0.0f;  // best times found
    uint64_t tune_key = 0;                         // pattern key of the tuning cache
    uint64_t rl_key = 0;                           // plan the cached right-looking schedule belongs to
    bool rl_ok = false;
    rsgpu::RlHost rl;
    uint64_t sp_key = 0;                           // plan the cached step programs belong to
    bool sp_ok = false;                            // step programs built and on the device
    rsgpu::SpSolverDev spd{};
    std::vector<rsgpu::SpOpWord> sp_words;
    std::vector<int32_t> sp_ints;
    rsgpu::DevBuf<rsgpu::SpSolverDev> d_spd;
    rsgpu::DevBuf<rsgpu::SpOpWord> d_spw;
    rsgpu::DevBuf<int32_t> d_spi;
    rsgpu::DevBuf<double> d_tune;                  // solve output while tuning
    rsgpu::DevBuf<rsgpu::PlanDev> d_desc;
    rsgpu::DevBuf<int32_t> d_ip;
    rsgpu::DevBuf<double> d_vals, d_ax, d_rhs, d_x, d_work;
    std::vector<double> h_ud;
    bool rcond_stale = false;
    // pivot-growth check of device refactors (0: off, the reference's behaviour)
    double growth_limit = 0.0;
    int64_t growth_fallbacks = 0;  // refactors replaced by a fresh pivoting factor
    bool refactor_fell_back = false;
    double last_growth = 0.0;
    rsgpu::DevBuf<double> d_growth;
    // transposed solve: L^T levels of the plan (built for factor_count ts_factors)
    int64_t err_pivot = -1;  // the pivot of the last RS_ERR_SINGULAR
    int64_t ts_factors = -1;
    int ts_nlev = 0;
    rsgpu::DevBuf<int32_t> ts_ptr, ts_cols;
    rsgpu::DevBuf<double> ts_t;
};

namespace rsgpu {
// The device solve of rs_solve_device without allocation or tuning: the
// kernels' solve vector is tmp (a device array of tmp_len doubles), the
// launch shape the handle holds. tmp_err: the error of finding tmp.
// A^T x = rhs with the stored factor (single-block level-set plans).
int32_t solver_tsolve_dev(rs_solver* h, const double* rhs_dev, double* x_dev, hipStream_t s);
int32_t solver_solve_noalloc(rs_solver* h, const double* rhs_dev, double* x_dev, hipStream_t s, double* tmp,
                             int64_t tmp_len, hipError_t tmp_err);
bool solver_graph_ready(const rs_solver* h);  // false while the growth check is on
hipError_t solver_graph_iteration(rs_solver* h, const double* ax_dev, const double* rhs_dev, double* x_dev,
                                  hipStream_t s, const int32_t* skip);
}  // namespace rsgpu
