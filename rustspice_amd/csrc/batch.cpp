// batch.cpp - rs_batch: host preparation and launch of the batched DC kernel
// (see batch.hpp for the execution model).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/rsgpu.h"
#include "batch.hpp"
#include "plan.hpp"
#include "profile.hpp"
#include "sp.hpp"
#include "solver_impl.hpp"
#include "stamp.hpp"

namespace rsgpu {
hipError_t launch_batch_dcop(const BPat* pats, const BInst* insts, const int32_t* order, int ninst,
                             int* counter, const int32_t* ip, const SpOpWord* prog,
                             const rs_device* devs, double* scr, int32_t* iscr, BResult* res,
                             const BOpts& o, int nmax, int lumax, int grid, int waves, hipStream_t s,
                             unsigned long long* phase_cyc);
int batch_blocks_per_cu(int nmax, int lumax, int waves);
size_t batch_lds_bytes(int nmax, int lumax, int waves);
size_t batch_static_lds();
}  // namespace rsgpu

using namespace rsgpu;

struct rs_batch {
    int kind = RS_SOLVER_NATIVEKLU;
    int64_t ninst = 0;
    int nmax = 0, npat = 0, lumax = 0;
    bool vals_lds = false;
    std::vector<BInst> inst;
    std::vector<int64_t> x_off;   // output x offset per instance
    std::vector<int32_t> n_of;    // matrix size per instance
    DevBuf<BPat> d_pats;
    DevBuf<BInst> d_inst;
    DevBuf<int32_t> d_order, d_ip, d_iscr;
    DevBuf<SpOpWord> d_prog;
    DevBuf<unsigned long long> d_phase;  // per-phase cycle totals when profiling
    bool phase_on = false;
    DevBuf<int> d_counter;
    DevBuf<rs_device> d_dev;
    DevBuf<double> d_scr;
    DevBuf<BResult> d_res;
    std::string err;
    int grid = 0;
    int waves = 1;  // wavefronts per instance
    double last_ms = 0.0;
};

namespace {

int32_t bfail(rs_batch* b, int32_t code, const std::string& m) {
    b->err = m;
    return code;
}
int32_t bhip(rs_batch* b, hipError_t e, const char* what) {
    b->err = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
    return RS_ERR_DEVICE;
}

int32_t put16(std::vector<int32_t>& pool, const std::vector<int32_t>& v) {  // 16-byte aligned
    while (pool.size() % 4) pool.push_back(0);
    int32_t off = (int32_t)pool.size();
    pool.insert(pool.end(), v.begin(), v.end());
    pool.push_back(0);
    return off;
}

int32_t put(std::vector<int32_t>& pool, const std::vector<int32_t>& v) {
    int32_t off = (int32_t)pool.size();
    pool.insert(pool.end(), v.begin(), v.end());
    pool.push_back(0);
    return off;
}

}  // namespace

extern "C" {

rs_batch* rs_batch_create(const rs_device* dev, const int64_t* dev_ptr, const int64_t* node_count,
                          int64_t ninst, int32_t kind) {
    rs_batch* b = new rs_batch();
    b->kind = kind;
    b->ninst = ninst;
    std::vector<int32_t> pool;
    std::vector<SpOpWord> prog;
    std::vector<BPat> pats;
    std::map<std::vector<int64_t>, int> pat_id;  // key: node_count, n, ap, ai, maps
    b->inst.resize(ninst);
    b->x_off.resize(ninst + 1);
    b->n_of.resize(ninst);
    int64_t scr = 0, iscr = 0;
    std::vector<int32_t> ph_nsc;  // scatter runs per pattern
    try {
        for (int64_t k = 0; k < ninst; k++) {
            const rs_device* d = dev + dev_ptr[k];
            const int64_t nd = dev_ptr[k + 1] - dev_ptr[k];
            for (int64_t i = 0; i < nd; i++)
                if (is_bsim(d[i]))
                    throw std::runtime_error("BSIM devices (MOSFET level != 1) are not supported by "
                                             "the batched kernel; use the circuit path");
            StampPattern sp;
            build_stamp_pattern(d, nd, node_count[k], MODE_DC, sp);
            std::vector<int64_t> key;
            key.reserve(4 + sp.ap.size() + 2 * sp.ai.size() + sp.rhs_ptr.size() + sp.rhs_src.size());
            key.push_back(node_count[k]);
            key.push_back(sp.n);
            key.push_back(nd);
            key.insert(key.end(), sp.ap.begin(), sp.ap.end());
            key.insert(key.end(), sp.ai.begin(), sp.ai.end());
            key.insert(key.end(), sp.ax_src.begin(), sp.ax_src.end());
            key.insert(key.end(), sp.rhs_ptr.begin(), sp.rhs_ptr.end());
            key.insert(key.end(), sp.rhs_src.begin(), sp.rhs_src.end());
            auto it = pat_id.find(key);
            int pid;
            if (it == pat_id.end()) {
                PlanHost ph;
                build_identity_plan(sp.n, sp.ap.data(), sp.ai.data(), kind, ph);
                BPat P{};
                P.n = (int32_t)sp.n;
                P.node_count = (int32_t)node_count[k];
                P.nnz = (int32_t)sp.ai.size();
                P.nblk = ph.nblk;
                P.nl = ph.nl;
                P.nu = ph.nu;
                std::vector<int32_t> axs(sp.ax_src.begin(), sp.ax_src.end());
                P.o_ax_src = put(pool, axs);
                P.o_rhs_ptr = put(pool, sp.rhs_ptr);
                P.o_rhs_src = put(pool, sp.rhs_src);
                P.o_sc_ptr = put(pool, ph.sc_ptr);
                P.o_sc_src = put(pool, ph.sc_src);
                P.o_sc_cnt = put(pool, ph.sc_cnt);
                P.o_sc_pre = put(pool, ph.sc_dst);
                P.o_lcp = put(pool, ph.lcp);
                P.o_lri = put(pool, ph.lri);
                P.o_ucp = put(pool, ph.ucp);
                P.o_uri = put(pool, ph.uri);
                P.o_colflag = put(pool, ph.colflag);
                P.o_pin = put(pool, ph.perm_in);
                P.o_pout = put(pool, ph.perm_out);
                P.o_lrp = put(pool, ph.lrp);
                P.o_lrc = put(pool, ph.lrc);
                P.o_lrv = put(pool, ph.lrv);
                P.o_urp = put(pool, ph.urp);
                P.o_urc = put(pool, ph.urc);
                P.o_urv = put(pool, ph.urv);
                P.o_blk_ptr = put(pool, ph.blk_ptr);
                P.o_lsl_blk = put(pool, ph.lsl_blk);
                P.o_lsl_ptr = put(pool, ph.lsl_ptr);
                P.o_lsl_rows = put(pool, ph.lsl_rows);
                P.o_usl_blk = put(pool, ph.usl_blk);
                P.o_usl_ptr = put(pool, ph.usl_ptr);
                P.o_usl_rows = put(pool, ph.usl_rows);
                P.o_odp = put(pool, ph.odp);
                P.o_odax = put(pool, ph.odax);
                P.o_odsrc = put(pool, ph.odsrc);
                P.o_blk_has_od = put(pool, ph.blk_has_od);
                {  // update program (pairs ordered as the left-looking loop visits them)
                    std::vector<int32_t> up_ptr(ph.n + 1, 0), up_j, up_row, up_l;
                    for (int32_t g = 0; g < ph.n; g++) {
                        for (int32_t u = ph.ucp[g]; u < ph.ucp[g + 1] - 1; u++) {
                            const int32_t j = ph.uri[u];
                            for (int32_t l = ph.lcp[j]; l < ph.lcp[j + 1]; l++) {
                                up_j.push_back(j);
                                up_row.push_back(ph.lri[l]);
                                up_l.push_back(l);
                            }
                        }
                        up_ptr[g + 1] = (int32_t)up_j.size();
                    }
                    P.o_up_ptr = put(pool, up_ptr);
                    P.o_up_j = put(pool, up_j);
                    P.o_up_row = put(pool, up_row);
                    P.o_up_l = put(pool, up_l);
                }
                P.nrhs = (int32_t)sp.rhs_src.size();
                {  // gather records (see BPat)
                    const int32_t nsc = ph.sc_ptr[ph.n];
                    std::vector<int32_t> rec(4 * (size_t)std::max(nsc, 1), 0), lrun, rrec(4 * (size_t)ph.n, 0), lrow;
                    for (int32_t q = 0; q < nsc; q++) {
                        const int32_t cnt = ph.sc_cnt[q], src = ph.sc_src[q];
                        rec[4 * q] = cnt;
                        if (cnt <= 3) {
                            for (int32_t k = 0; k < cnt; k++) rec[4 * q + 1 + k] = axs[src + k];
                        } else {
                            rec[4 * q + 1] = src;
                            lrun.push_back(q);
                        }
                    }
                    for (int32_t i = 0; i < ph.n; i++) {
                        const int32_t k0 = sp.rhs_ptr[i], cnt = sp.rhs_ptr[i + 1] - k0;
                        rrec[4 * i] = cnt;
                        if (cnt <= 3) {
                            for (int32_t k = 0; k < cnt; k++) rrec[4 * i + 1 + k] = sp.rhs_src[k0 + k];
                        } else {
                            rrec[4 * i + 1] = k0;
                            lrow.push_back(i);
                        }
                    }
                    P.o_screc = put16(pool, rec);
                    P.o_lrun = put(pool, lrun);
                    P.n_lrun = (int32_t)lrun.size();
                    P.o_rhsrec = put16(pool, rrec);
                    P.o_lrow = put(pool, lrow);
                    P.n_lrow = (int32_t)lrow.size();
                }
                {  // column of every scatter run
                    std::vector<int32_t> col(ph.sc_ptr[ph.n]);
                    for (int32_t g = 0; g < ph.n; g++)
                        for (int32_t q = ph.sc_ptr[g]; q < ph.sc_ptr[g + 1]; q++) col[q] = g;
                    P.o_sc_col = put(pool, col);
                }
                {  // step programs of the refactor and the solve
                    SpProgram fac, sol;
                    std::string why;
                    const SpLayout L = sp_layout(ph);
                    P.sp_ok = build_sp_programs(ph, fac, sol, &why) ? 1 : 0;
                    P.t_base = L.t_base;
                    P.od_base = L.od_base;
                    P.nslots = L.nslots;
                    if (P.sp_ok) {
                        P.o_spf = (int64_t)prog.size();
                        prog.insert(prog.end(), fac.words.begin(), fac.words.end());
                        P.spf_chunks = fac.nsteps_pad / kSpChunk - 2;
                        P.o_sps = (int64_t)prog.size();
                        prog.insert(prog.end(), sol.words.begin(), sol.words.end());
                        P.sps_chunks = sol.nsteps_pad / kSpChunk - 2;
                    }
                }
                if (pool.size() > (size_t)INT32_MAX) throw std::runtime_error("pattern pool too large");
                pid = (int)pats.size();
                pats.push_back(P);
                ph_nsc.push_back(ph.sc_ptr[ph.n]);
                pat_id.emplace(std::move(key), pid);
            } else {
                pid = it->second;
            }
            const BPat& P = pats[pid];
            BInst& I = b->inst[k];
            I.pat = pid;
            I.ndev = (int32_t)nd;
            I.o_dev = dev_ptr[k];
            I.o_sv = scr;
            scr += nd * kSlotsPerDevice;
            I.o_ax = scr;
            scr += P.nnz;
            I.o_lx = scr;
            scr += P.nl;
            I.o_ux = scr;
            scr += P.nu;
            I.o_ud = scr;
            scr += P.n;
            I.o_x = scr;
            scr += P.n;
            I.o_rv = scr;
            scr += P.nrhs;
            I.o_ss = iscr;
            iscr += ph_nsc[pid];
            b->n_of[k] = P.n;
            b->nmax = std::max(b->nmax, P.n);
            b->lumax = std::max(b->lumax, P.nslots + 1);  // + the dummy slot
        }
    } catch (const std::exception& ex) {
        b->err = std::string("Invalid matrix: ") + ex.what();
        return b;
    }
    b->npat = (int)pats.size();
    // queue order: largest instances first (longest-processing-time first)
    std::vector<int32_t> order(ninst);
    for (int64_t k = 0; k < ninst; k++) order[k] = (int32_t)k;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t x, int32_t y) { return b->n_of[x] > b->n_of[y]; });
    hipError_t e;
    int64_t ndev_total = dev_ptr[ninst];
    if ((e = b->d_pats.ensure(pats.size())) || (e = b->d_inst.ensure(ninst)) ||
        (e = b->d_order.ensure(ninst)) || (e = b->d_ip.ensure(pool.size())) ||
        (e = b->d_counter.ensure(1)) || (e = b->d_dev.ensure(ndev_total)) ||
        (e = b->d_scr.ensure(scr)) || (e = b->d_res.ensure(ninst)) ||
        (e = b->d_iscr.ensure(std::max<int64_t>(iscr, 1))) ||
        (e = b->d_prog.ensure(std::max<size_t>(prog.size(), 1)))) {
        bhip(b, e, "hipMalloc(batch)");
        return b;
    }
    if ((e = hipMemcpy(b->d_pats.p, pats.data(), pats.size() * sizeof(BPat), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(b->d_inst.p, b->inst.data(), ninst * sizeof(BInst), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(b->d_order.p, order.data(), ninst * sizeof(int32_t), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(b->d_ip.p, pool.data(), pool.size() * sizeof(int32_t), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(b->d_dev.p, dev, ndev_total * sizeof(rs_device), hipMemcpyHostToDevice)) ||
        (e = hipMemset(b->d_scr.p, 0, scr * sizeof(double))) ||
        (e = hipMemcpy(b->d_prog.p, prog.data(), prog.size() * sizeof(SpOpWord), hipMemcpyHostToDevice))) {
        bhip(b, e, "hipMemcpy(batch)");
        return b;
    }
    int dev_id = 0, cus = 0;
    (void)hipGetDevice(&dev_id);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id);
    // L/U values in LDS when they fit: faster per instance, fewer instances
    // resident per CU. Kept when every instance still gets a resident
    // wavefront, or when at least 4 fit per CU (waves per CU capped at 16).
    int lds_max = 0;
    (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev_id);
    const size_t kStatic = 0;  // the kernel has no static LDS
    // Four wavefronts per instance when every instance has a CU of its own
    // (the data-parallel phases spread over them), else one.
    b->waves = ninst <= (int64_t)std::max(cus, 1) ? 4 : 1;
    if (batch_lds_bytes(b->nmax, 0, b->waves) + kStatic > (size_t)lds_max) {
        char m[160];
        std::snprintf(m, sizeof m, "instance of %d unknowns exceeds the %d bytes of LDS per workgroup",
                      b->nmax, lds_max);
        b->err = m;
        return b;
    }
    const bool lds_fit = batch_lds_bytes(b->nmax, b->lumax, b->waves) + kStatic <= (size_t)lds_max;
    int per_cu_g = std::min(16, batch_blocks_per_cu(b->nmax, 0, b->waves));
    int per_cu_l = lds_fit ? std::min(16, batch_blocks_per_cu(b->nmax, b->lumax, b->waves)) : 0;
    // (step programs need the values in LDS and address them absolutely)
    b->vals_lds = lds_fit && batch_static_lds() == 0 &&
                  (per_cu_l >= 4 || ninst <= (int64_t)std::max(cus, 1) * per_cu_l);
    int per_cu = b->vals_lds ? per_cu_l : per_cu_g;
    b->grid = (int)std::min<int64_t>(ninst, (int64_t)std::max(cus, 1) * per_cu);
    if (b->grid < 1) b->grid = 1;
    return b;
}

void rs_batch_destroy(rs_batch* b) {
    if (!b) return;
    b->d_pats.release();
    b->d_inst.release();
    b->d_order.release();
    b->d_ip.release();
    b->d_counter.release();
    b->d_dev.release();
    b->d_scr.release();
    b->d_res.release();
    b->d_iscr.release();
    b->d_prog.release();
    b->d_phase.release();
    delete b;
}

int32_t rs_batch_last_error(const rs_batch* b, char* buf, int64_t len) {
    if (len > 0) std::snprintf(buf, (size_t)len, "%s", b->err.c_str());
    return (int32_t)b->err.size();
}

int32_t rs_batch_info(const rs_batch* b, int64_t* out) {
    out[0] = b->ninst;
    out[1] = b->npat;
    out[2] = b->nmax;
    out[3] = b->grid;
    out[4] = b->vals_lds ? b->lumax : 0;
    return b->err.empty() ? RS_OK : RS_ERR_DEVICE;
}

int32_t rs_batch_dc_op(rs_batch* b, const rs_newton_opts* o, void* stream) {
    if (!b->err.empty()) return RS_ERR_DEVICE;
    BOpts bo{o->max_iters, o->gmin_steps, o->source_steps, 0,
             o->abs_tol, o->rel_tol, o->gmin, o->damping, o->damping_min,
             b->kind == RS_SOLVER_SPARSELU ? 0.0 : 0.001};
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipMemsetAsync(b->d_counter.p, 0, sizeof(int), s);
    if (e) return bhip(b, e, "hipMemset(counter)");
    const int ps = prof_begin(PK_BATCH, s);
    e = launch_batch_dcop(b->d_pats.p, b->d_inst.p, b->d_order.p, (int)b->ninst, b->d_counter.p,
                          b->d_ip.p, b->d_prog.p, b->d_dev.p, b->d_scr.p, b->d_iscr.p, b->d_res.p, bo, b->nmax,
                          b->vals_lds ? b->lumax : 0, b->grid, b->waves, s,
                          b->phase_on ? b->d_phase.p : nullptr);
    prof_end(ps, s);
    if (e) return bhip(b, e, "k_batch_dcop");
    return RS_OK;
}

int32_t rs_batch_phase_profile(rs_batch* b, int32_t on) {
    b->phase_on = on != 0;
    if (!b->phase_on) return RS_OK;
    hipError_t e;
    if ((e = b->d_phase.ensure(8)) || (e = hipMemset(b->d_phase.p, 0, 8 * sizeof(unsigned long long))))
        return bhip(b, e, "phase profile");
    return RS_OK;
}

int32_t rs_batch_phase_cycles(rs_batch* b, int64_t* out) {
    if (!b->d_phase.p) return RS_ERR_DEVICE;
    hipError_t e = hipMemcpy(out, b->d_phase.p, 8 * sizeof(int64_t), hipMemcpyDeviceToHost);
    return e ? bhip(b, e, "phase cycles") : RS_OK;
}

int32_t rs_batch_results(rs_batch* b, rs_newton_result* res, double* x_out, int64_t* x_ptr) {
    std::vector<BResult> r(b->ninst);
    hipError_t e = hipMemcpy(r.data(), b->d_res.p, b->ninst * sizeof(BResult), hipMemcpyDeviceToHost);
    if (e) return bhip(b, e, "hipMemcpy(results)");
    int64_t off = 0;
    for (int64_t k = 0; k < b->ninst; k++) {
        if (res) {
            res[k].converged = r[k].converged;
            res[k].iterations = r[k].iterations;
            res[k].reason = r[k].reason;
            res[k].pad = r[k].full_factors;
            res[k].final_norm = r[k].final_norm;
        }
        if (x_ptr) x_ptr[k] = off;
        if (x_out) {
            e = hipMemcpy(x_out + off, b->d_scr.p + b->inst[k].o_x, b->n_of[k] * sizeof(double),
                          hipMemcpyDeviceToHost);
            if (e) return bhip(b, e, "hipMemcpy(x)");
        }
        off += b->n_of[k];
    }
    if (x_ptr) x_ptr[b->ninst] = off;
    return RS_OK;
}

}  // extern "C"
