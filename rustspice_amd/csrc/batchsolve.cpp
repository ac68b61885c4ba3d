// batchsolve.cpp - many independent matrices (patterns may differ) refactored
// in one launch and solved in another, one workgroup per matrix: the
// single-matrix column refactor and level solve kernels over an array of plan
// descriptors. Each matrix keeps its own reference semantics: NativeKLU's
// first factor picks its pivots on the host and every batched refactor reuses
// them (native_klu.rs:1244-1255), SparseLU refactors with identity pivots
// (sparse_lu.rs:440-537); values are those of a single solver bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rsgpu.h"
#include "../../include/rsgpu_host.h"
#include "plan.hpp"
#include "solver_impl.hpp"

using namespace rsgpu;

struct rs_solver_batch {
    int32_t kind = RS_SOLVER_NATIVEKLU;
    int64_t count = 0, max_n = 0, tot_n = 0, tot_nnz = 0;
    std::vector<int64_t> n, nnz, ax_off, vec_off, o_lx, o_ux, o_ud, nl, nu;
    DevBuf<PlanDev> descs;
    DevBuf<int32_t> ip;
    DevBuf<double> vals, work, ax;
    int ref_mode = 0, threads = kThreads;
    bool sol_lds = false;
    std::string err;
};

namespace {

int32_t bfail(rs_solver_batch* b, int32_t code, const std::string& m) {
    b->err = m;
    return code;
}

int32_t bhip(rs_solver_batch* b, hipError_t e, const char* what) {
    return bfail(b, RS_ERR_DEVICE, std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

}  // namespace

extern "C" {

rs_solver_batch* rs_solver_batch_create(int32_t kind, int64_t count, const int64_t* n, const int64_t* ap_cat,
                                        const int64_t* ai_cat, const double* ax_cat) {
    rs_solver_batch* b = new rs_solver_batch();
    b->kind = kind;
    if (kind != RS_SOLVER_NATIVEKLU && kind != RS_SOLVER_SPARSELU) {
        bfail(b, RS_ERR_INVALID_MATRIX, "Invalid matrix: batched solvers are NativeKLU or SparseLU");
        return b;
    }
    if (count < 1) {
        bfail(b, RS_ERR_INVALID_MATRIX, "Invalid matrix: empty batch");
        return b;
    }
    b->count = count;
    std::vector<int32_t> pool;
    std::vector<PlanDev> descs(count);
    std::vector<double> vals;
    int64_t apo = 0, wp = 0;
    for (int64_t m = 0; m < count; m++) {
        const int64_t nm = n[m];
        const int64_t* ap = ap_cat + apo;
        if (nm < 1 || ap[0] != 0) {
            bfail(b, RS_ERR_INVALID_MATRIX, "Invalid matrix: matrix " + std::to_string(m) + " has a bad column pointer");
            b->count = 0;
            return b;
        }
        for (int64_t j = 0; j < nm; j++)
            if (ap[j + 1] < ap[j]) {
                bfail(b, RS_ERR_INVALID_MATRIX, "Invalid matrix: matrix " + std::to_string(m) + " has a bad column pointer");
                b->count = 0;
                return b;
            }
        const int64_t nz = ap[nm];
        const int64_t* ai = ai_cat + b->tot_nnz;
        const double* ax = ax_cat + b->tot_nnz;
        rs_host_factor* hf = rs_host_factor_create(kind == RS_SOLVER_SPARSELU ? 1 : 0, nm, ap, ai, ax, 0.001, 0, 1);
        PlanDev d{};
        append_plan(hf->p, pool, d);
        std::vector<double> lx, ux, ud;
        gather_values(hf->f, lx, ux, ud);
        d.o_lx = (int64_t)vals.size();
        d.o_ux = d.o_lx + hf->p.nl + 1;
        d.o_ud = d.o_ux + hf->p.nu + 1;
        vals.resize(d.o_ud + nm + 1, 0.0);
        std::copy(lx.begin(), lx.end(), vals.begin() + d.o_lx);
        std::copy(ux.begin(), ux.end(), vals.begin() + d.o_ux);
        std::copy(ud.begin(), ud.end(), vals.begin() + d.o_ud);
        d.o_ax = b->tot_nnz;
        d.o_vec = b->tot_n;
        d.o_work = wp;
        wp += (int64_t)(kThreads / 64) * nm;
        d.o_tvec = wp;
        wp += nm;
        d.flags = 0;
        descs[m] = d;
        b->n.push_back(nm);
        b->nnz.push_back(nz);
        b->ax_off.push_back(b->tot_nnz);
        b->vec_off.push_back(b->tot_n);
        b->o_lx.push_back(d.o_lx);
        b->o_ux.push_back(d.o_ux);
        b->o_ud.push_back(d.o_ud);
        b->nl.push_back(hf->p.nl);
        b->nu.push_back(hf->p.nu);
        b->max_n = std::max(b->max_n, nm);
        b->tot_n += nm;
        b->tot_nnz += nz;
        apo += nm + 1;
        rs_host_factor_destroy(hf);
    }
    // work vectors in LDS when every matrix's fit (the kernels size the LDS
    // by the largest matrix)
    const size_t lds = device_max_lds();
    b->ref_mode = (size_t)(kThreads / 64) * b->max_n * sizeof(double) <= lds ? 1 : 0;
    b->sol_lds = (size_t)kThreads * sizeof(double) + (size_t)b->max_n * sizeof(double) <= lds;
    hipError_t e;
    if ((e = b->descs.ensure(count)) || (e = b->ip.ensure(std::max<size_t>(pool.size(), 1))) ||
        (e = b->vals.ensure(std::max<size_t>(vals.size(), 1))) || (e = b->work.ensure(std::max<int64_t>(wp, 1))) ||
        (e = b->ax.ensure(std::max<int64_t>(b->tot_nnz, 1))) ||
        (e = hipMemcpy(b->descs.p, descs.data(), count * sizeof(PlanDev), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(b->ip.p, pool.data(), pool.size() * 4, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(b->vals.p, vals.data(), vals.size() * 8, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(b->ax.p, ax_cat, b->tot_nnz * 8, hipMemcpyHostToDevice)) ||
        (e = hipMemset(b->work.p, 0, std::max<int64_t>(wp, 1) * 8)))  // the column kernel's all-zero work vectors
        bhip(b, e, "batched solver upload");
    return b;
}

void rs_solver_batch_destroy(rs_solver_batch* b) { delete b; }

int32_t rs_solver_batch_last_error(const rs_solver_batch* b, char* buf, int64_t len) {
    if (len > 0) std::snprintf(buf, (size_t)len, "%s", b->err.c_str());
    return (int32_t)b->err.size();
}

int32_t rs_solver_batch_layout(const rs_solver_batch* b, int64_t* sizes, int64_t* ax_off, int64_t* vec_off) {
    if (!b->err.empty() || b->count == 0) return RS_ERR_INVALID_MATRIX;
    sizes[0] = b->count;
    sizes[1] = b->tot_nnz;
    sizes[2] = b->tot_n;
    if (ax_off) std::copy(b->ax_off.begin(), b->ax_off.end(), ax_off);
    if (vec_off) std::copy(b->vec_off.begin(), b->vec_off.end(), vec_off);
    return RS_OK;
}

int32_t rs_solver_batch_refactor(rs_solver_batch* b, const double* ax_dev, void* stream) {
    if (!b->err.empty() && b->count == 0) return RS_ERR_INVALID_MATRIX;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    // the solve reads BTF off-diagonal values from the batch's own copy
    if ((e = hipMemcpyAsync(b->ax.p, ax_dev, b->tot_nnz * 8, hipMemcpyDeviceToDevice, s)) ||
        (e = launch_refactor(b->descs.p, (int)b->count, b->ip.p, b->ax.p, b->vals.p, b->work.p, (int)b->max_n,
                             b->ref_mode, b->threads, s)))
        return bhip(b, e, "batched refactor");
    return RS_OK;
}

int32_t rs_solver_batch_solve(rs_solver_batch* b, const double* rhs_dev, double* x_dev, void* stream) {
    if (!b->err.empty() && b->count == 0) return RS_ERR_INVALID_MATRIX;
    hipError_t e = launch_solve(b->descs.p, (int)b->count, b->ip.p, b->ax.p, b->vals.p, rhs_dev, x_dev, b->work.p,
                                (int)b->max_n, b->sol_lds, b->threads, (hipStream_t)stream);
    return e == hipSuccess ? RS_OK : bhip(b, e, "batched solve");
}

int32_t rs_solver_batch_factor_values(rs_solver_batch* b, int64_t m, double* lx, double* ux, double* udiag) {
    if (m < 0 || m >= b->count) return bfail(b, RS_ERR_INVALID_MATRIX, "Invalid matrix: no such matrix");
    hipError_t e;
    if ((e = hipDeviceSynchronize()) ||
        (lx && b->nl[m] && (e = hipMemcpy(lx, b->vals.p + b->o_lx[m], b->nl[m] * 8, hipMemcpyDeviceToHost))) ||
        (ux && (e = hipMemcpy(ux, b->vals.p + b->o_ux[m], b->nu[m] * 8, hipMemcpyDeviceToHost))) ||
        (udiag && (e = hipMemcpy(udiag, b->vals.p + b->o_ud[m], b->n[m] * 8, hipMemcpyDeviceToHost))))
        return bhip(b, e, "batched factor values");
    return RS_OK;
}

}  // extern "C"
