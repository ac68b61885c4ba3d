// engine.cpp - rs_circuit: device-resident stamping plus the DC/TRAN Newton
// drivers of crates/sim-core/src/engine.rs over the GPU solver.
//
// Per Newton iteration (newton.rs:65-117) everything on the device: stamp
// evaluation and assembly into the unmerged CSC, refactor (or the first,
// pivoting factor of a new pattern on the host), solve. The host keeps the
// control flow that decides convergence, damping, LTE acceptance and the next
// time step (newton.rs:93-107, engine.rs:463-602, analysis.rs:147-742,
// waveform.rs:459-645) on the same doubles the reference computes, so
// iteration counts and accepted steps follow the reference exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <functional>
#include <vector>

#include "../../include/rsgpu.h"
#include "solver_impl.hpp"
#include "stamp.hpp"
#include "fold.hpp"
#include "profile.hpp"
#include "tran_ctl.hpp"
#include "sn.hpp"

namespace rsgpu {
hipError_t launch_eval(const rs_device* dev, int ndev, const double* x, int xl, StampArgs a,
                       const double* cap_v, const double* cap_i, const uint8_t* has_v,
                       const uint8_t* has_i, double* sv, hipStream_t s, const int32_t* sbase,
                       uint8_t* gmask);
hipError_t launch_mos_eval(const rs_bsim_model* models, const int32_t* mi, const int32_t* lv,
                           const double* in, int64_t n, double* out, hipStream_t s);
hipError_t launch_assemble(const int32_t* ax_src, int nnz, const int32_t* rhs_ptr,
                           const int32_t* rhs_src, int n, const double* sv, double* ax, double* rhs,
                           const int32_t* long_rows, int nlong, hipStream_t s, const int32_t* skip = nullptr);
hipError_t launch_assemble_long(const int32_t* rhs_ptr, const int32_t* rhs_src, const double* sv, double* rhs,
                                const int32_t* long_rows, int nlong, hipStream_t s);
hipError_t launch_cap_state(const rs_device* dev, int ndev, const double* x, int xl, double dt,
                            int full, double* cap_v, double* cap_i, uint8_t* has_v,
                            uint8_t* has_i, hipStream_t s, const int32_t* gate = nullptr,
                            const double* dt_dev = nullptr);
hipError_t launch_newton_begin(double* t_dev, double dt, int32_t* st, double damping, hipStream_t s);
hipError_t launch_newton_update(int n, double* x, const double* xn, double abs_tol, double rel_tol,
                                double damping_min, int max_iters, int32_t* st, hipStream_t s);
hipError_t launch_adjoint_contract(const rs_device* dev, int ndev, const int32_t* aux_of,
                                   const double* x, int xl, const double* lam, double gmin,
                                   double scale, int mode, const double* xp, const double* xp2,
                                   double dt, int acc, double* out, hipStream_t s);
hipError_t launch_tadj_rhs(const rs_device* dev, const int32_t* cap_ptr, const int32_t* cap_dev,
                           int node_count, const double* lam1, const double* lam2, double dt,
                           int k_ge1, int probe, int n, double* rhs, hipStream_t s);
hipError_t launch_resid_t(const int32_t* ap, const int32_t* ai, const double* ax, const double* e,
                          const double* lam, double* r, int n, hipStream_t s);
hipError_t launch_refine_update(double* lam, const double* delta, int n, double* out, hipStream_t s);
hipError_t launch_resid_n(const int32_t* rp, const int32_t* rc, const int32_t* rq, const double* ax,
                          const double* b, const double* x, int xl, double* r, int n, hipStream_t s);
}  // namespace rsgpu

using namespace rsgpu;

extern "C" {
rs_batch* rs_batch_create(const rs_device* dev, const int64_t* dev_ptr, const int64_t* node_count,
                          int64_t ninst, int32_t kind);
void rs_batch_destroy(rs_batch* b);
int32_t rs_batch_dc_op(rs_batch* b, const rs_newton_opts* o, void* stream);
int32_t rs_batch_results(rs_batch* b, rs_newton_result* res, double* x_out, int64_t* x_ptr);
int32_t rs_batch_last_error(const rs_batch* b, char* buf, int64_t len);
}

namespace {
struct DevPattern {
    bool built = false;
    std::vector<uint8_t> gmask;  // BSIM group masks the pattern was built for
    StampPattern p;
    DevBuf<int32_t> ax_src, rhs_ptr, rhs_src;
    DevBuf<int32_t> long_rows;  // RHS rows assembled by a workgroup each (kFoldRhs)
    int nlong = 0;
};
}  // namespace

struct rs_circuit {
    std::vector<rs_device> dev;
    i64 node_count = 0;
    DevBuf<rs_device> d_dev;
    DevPattern pat[2];  // 0: DC, 1: TRAN (BE and TRAP share the pattern)
    DevBuf<double> d_sv, d_ax, d_rhs, d_x, d_xnew, d_cap_v, d_cap_i;
    DevBuf<uint8_t> d_has_v, d_has_i;
    // side stream for the long RHS rows of a deferred stamp (default priority)
    hipStream_t side = nullptr;
    hipEvent_t ev_sv = nullptr, ev_rhs = nullptr;
    bool rhs_pending = false;
    std::string err;
    int64_t stamp_calls = 0, newton_iters = 0, factor_calls = 0;
    bool device_newton = true;  // whole DC OP in one launch when equivalent
    rs_batch* dcb[2] = {nullptr, nullptr};  // one-instance batch per solver kind
    int64_t device_dc_runs = 0;
    // adjoint: per pattern (0 DC, 1 TRAN) a SparseLU factor of the exact
    // Jacobian with its L^T levels; shared vectors
    struct AdjCtx {
        rs_solver* s = nullptr;
        int64_t factors = -1;  // factor_count when the L^T levels were built
        int n_lt = 0;
        DevBuf<int32_t> aux, ltp, ltc, jap, jai;
        // flat pipelined transposed solve (k_solve_lp on the transposed
        // records) when they fit the LDS; k_solve_t otherwise
        bool t_ok = false;
        int t_nlev = 0, t_maxent = 0;
        DevBuf<int32_t> tip;
        DevBuf<PlanDev> tdesc;
        bool have_rows = false;        // row map of the stamped pattern (polish residuals)
        DevBuf<int32_t> rp, rc, rq;
        // a pattern the solver factors supernodally: J^T factored on its own
        // (the plan of A^T on the solver's symbolic factor) for the solves
        // with J^T, rebuilt when the solver analyzes a new pattern
        SnImpl* snT = nullptr;
        int64_t snT_analysis = -1;
    } adj[2];
    int adj_sn_mode = 1;  // supernodal setting of the adjoint's solvers (the TRAN adjoint's caller's)
    int64_t polish_iters = 0;
    // fixed-step transient with each TRAP step as a captured hipGraph
    bool use_graph = false;
    int64_t graph_steps = 0, graph_launches = 0;
    int64_t tran_graph_runs = 0, tran_host_pows = 0;  // adaptive TRAN on the device controller
    DevBuf<double> d_e, d_t, d_lam, d_dp, d_r, d_del, d_nrm;
    int32_t adj_refine = 0;  // refinement steps of the last adjoint solve (largest over a TRAN)
    // BSIM devices: slot bases, model table, group masks of the last evaluation
    int64_t nbsim = 0;
    std::vector<int32_t> sbase;
    std::vector<uint8_t> gmask;
    DevBuf<int32_t> d_sbase;
    DevBuf<uint8_t> d_gmask;
    DevBuf<rs_bsim_model> d_models;
    int64_t nmodels = 0;
    int64_t pattern_switches = 0;  // patterns rebuilt because a BSIM group switched
    std::vector<double> ic;        // rs_tran's initial guess (.ic), empty: zeros
};

namespace {

int32_t cfail(rs_circuit* c, int32_t code, const std::string& m) {
    c->err = m;
    return code;
}
int32_t chip(rs_circuit* c, hipError_t e, const char* what) {
    c->err = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
    return RS_ERR_DEVICE;
}
int32_t from_solver(rs_circuit* c, rs_solver* s, int32_t rc) {
    c->err = s->err;
    return rc;
}

int32_t get_pattern(rs_circuit* c, int mode, DevPattern*& out) {
    int k = mode == MODE_DC ? 0 : 1;
    DevPattern& dp = c->pat[k];
    if (dp.built && c->nbsim && dp.gmask != c->gmask) {  // a BSIM group switched
        dp.built = false;
        c->pattern_switches++;
    }
    if (!dp.built) {
        try {
            build_stamp_pattern(c->dev.data(), (i64)c->dev.size(), c->node_count,
                                k == 0 ? MODE_DC : MODE_TRAN_BE, dp.p,
                                c->nbsim ? c->sbase.data() : nullptr,
                                c->nbsim ? c->gmask.data() : nullptr);
            dp.gmask = c->gmask;
        } catch (const std::exception& ex) {
            return cfail(c, RS_ERR_INVALID_MATRIX, std::string("Invalid matrix: ") + ex.what());
        }
        hipError_t e;
        if ((e = dp.ax_src.ensure(dp.p.ax_src.size())) ||
            (e = dp.rhs_ptr.ensure(dp.p.rhs_ptr.size())) ||
            (e = dp.rhs_src.ensure(dp.p.rhs_src.size())))
            return chip(c, e, "hipMalloc(pattern)");
        if ((e = hipMemcpy(dp.ax_src.p, dp.p.ax_src.data(), dp.p.ax_src.size() * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(dp.rhs_ptr.p, dp.p.rhs_ptr.data(), dp.p.rhs_ptr.size() * 4, hipMemcpyHostToDevice)) ||
            (!dp.p.rhs_src.empty() &&
             (e = hipMemcpy(dp.rhs_src.p, dp.p.rhs_src.data(), dp.p.rhs_src.size() * 4, hipMemcpyHostToDevice))))
            return chip(c, e, "hipMemcpy(pattern)");
        {
            std::vector<int32_t> lr;
            for (size_t r = 0; r + 1 < dp.p.rhs_ptr.size(); r++)
                if (dp.p.rhs_ptr[r + 1] - dp.p.rhs_ptr[r] > kFoldRhs) lr.push_back((int32_t)r);
            dp.nlong = (int)lr.size();
            if ((e = dp.long_rows.ensure(std::max<size_t>(lr.size(), 1))) ||
                (!lr.empty() && (e = hipMemcpy(dp.long_rows.p, lr.data(), lr.size() * 4, hipMemcpyHostToDevice))))
                return chip(c, e, "hipMemcpy(long rows)");
        }
        size_t nmax = (size_t)dp.p.n;
        if ((e = c->d_ax.ensure(std::max(c->d_ax.n, dp.p.ai.size()))) ||
            (e = c->d_rhs.ensure(std::max(c->d_rhs.n, nmax))) ||
            (e = c->d_xnew.ensure(std::max(c->d_xnew.n, nmax))))
            return chip(c, e, "hipMalloc(vectors)");
        if (c->d_x.n < nmax) {
            // x may already hold the caller's iterate: a larger buffer keeps it
            DevBuf<double> nx;
            if ((e = nx.ensure(nmax)) ||
                (c->d_x.p && (e = hipMemcpy(nx.p, c->d_x.p, c->d_x.n * sizeof(double), hipMemcpyDeviceToDevice))))
                return chip(c, e, "hipMalloc(x)");
            std::swap(c->d_x.p, nx.p);
            std::swap(c->d_x.n, nx.n);
        }
        dp.built = true;
    }
    out = &dp;
    return RS_OK;
}

// Stamp on the device from a device-resident x.
int32_t stamp_dev(rs_circuit* c, int mode, const double* x_dev, i64 xl, double t, double dt,
                  double gmin, double scale, DevPattern** pout, hipStream_t s, int32_t flags = 0) {
    DevPattern* dp;
    int32_t rc;
    if (c->nbsim && c->nmodels == 0)
        return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: BSIM devices without a model table");
    if (c->nbsim && (flags & kStampJacobian))
        return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: no exact Jacobian for BSIM devices");
    const bool defer = (flags & kStampDeferLongRhs) != 0;
    flags &= ~kStampDeferLongRhs;
    StampArgs a{t, dt, gmin, scale, mode, flags, c->d_models.p};
    int ps = prof_begin(PK_EVAL, s);
    hipError_t e = launch_eval(c->d_dev.p, (int)c->dev.size(), x_dev, (int)xl, a, c->d_cap_v.p,
                               c->d_cap_i.p, c->d_has_v.p, c->d_has_i.p, c->d_sv.p, s,
                               c->nbsim ? c->d_sbase.p : nullptr, c->d_gmask.p);
    prof_end(ps, s);
    if (e) return chip(c, e, "k_eval");
    if (c->nbsim) {  // the stamp pattern depends on which BSIM groups are stamped
        e = hipMemcpyAsync(c->gmask.data(), c->d_gmask.p, c->gmask.size(), hipMemcpyDeviceToHost, s);
        if (!e) e = hipStreamSynchronize(s);
        if (e) return chip(c, e, "hipMemcpy(group masks)");
    }
    if ((rc = get_pattern(c, mode, dp))) return rc;
    ps = prof_begin(PK_ASSEMBLE, s);
    const bool side = defer && dp->nlong > 0;
    e = launch_assemble(dp->ax_src.p, (int)dp->p.ai.size(), dp->rhs_ptr.p, dp->rhs_src.p,
                        (int)dp->p.n, c->d_sv.p, c->d_ax.p, c->d_rhs.p, dp->long_rows.p, side ? 0 : dp->nlong, s);
    prof_end(ps, s);
    if (e) return chip(c, e, "k_assemble");
    if (side) {  // the long rows overlap the factor; rhs_join() before the solve
        if (!c->side &&
            ((e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking)) ||
             (e = hipEventCreateWithFlags(&c->ev_sv, hipEventDisableTiming)) ||
             (e = hipEventCreateWithFlags(&c->ev_rhs, hipEventDisableTiming))))
            return chip(c, e, "side stream");
        if ((e = hipEventRecord(c->ev_sv, s)) || (e = hipStreamWaitEvent(c->side, c->ev_sv, 0)) ||
            (e = launch_assemble_long(dp->rhs_ptr.p, dp->rhs_src.p, c->d_sv.p, c->d_rhs.p, dp->long_rows.p,
                                      dp->nlong, c->side)) ||
            (e = hipEventRecord(c->ev_rhs, c->side)))
            return chip(c, e, "k_assemble_long");
        c->rhs_pending = true;
    }
    c->stamp_calls++;
    if (pout) *pout = dp;
    return RS_OK;
}

// the stream waits for the RHS rows a deferred stamp left on the side stream
int32_t rhs_join(rs_circuit* c, hipStream_t s) {
    if (!c->rhs_pending) return RS_OK;
    c->rhs_pending = false;
    const hipError_t e = hipStreamWaitEvent(s, c->ev_rhs, 0);
    return e ? chip(c, e, "rhs join") : RS_OK;
}

double norm2(const double* v, i64 n) {
    double s = 0.0;
    for (i64 i = 0; i < n; i++) s += v[i] * v[i];
    return std::sqrt(s);
}

// run_newton (newton.rs:65-117)
int32_t run_newton(rs_circuit* c, rs_solver* sol, const rs_newton_opts* o, std::vector<double>& x,
                   int mode, double t, double dt, double gmin, double scale,
                   rs_newton_result* r, int32_t stamp_flags = 0) {
    hipStream_t s = nullptr;
    r->converged = 0;
    r->iterations = 0;
    r->reason = 1;
    r->final_norm = 0.0;
    double damping = o->damping;
    double prev = 1.7976931348623157e308;
    std::vector<double> xn, dx;
    {
        DevPattern* dp0;  // allocates the pattern's device vectors before x is uploaded
        int32_t rc0 = get_pattern(c, mode, dp0);
        if (rc0) return rc0;
        hipError_t e0 = c->d_x.ensure(std::max(c->d_x.n, x.size()));
        if (e0) return chip(c, e0, "hipMalloc(x)");
    }
    for (int it = 0; it < o->max_iters; it++) {
        r->iterations = it + 1;
        c->newton_iters++;
        hipError_t e = hipMemcpyAsync(c->d_x.p, x.data(), x.size() * 8, hipMemcpyHostToDevice, s);
        if (e) return chip(c, e, "hipMemcpy(x)");
        DevPattern* dp;
        int32_t rc = stamp_dev(c, mode, c->d_x.p, (i64)x.size(), t, dt, gmin, scale, &dp, s,
                               stamp_flags | kStampDeferLongRhs);
        if (rc) return rc;
        const i64 n = dp->p.n;
        rs_prepare(sol, n);
        rc = solver_factor_dev(sol, dp->p.ap.data(), n + 1, dp->p.ai.data(), (i64)dp->p.ai.size(),
                               c->d_ax.p, s);
        c->factor_calls++;
        const int32_t rj = rhs_join(c, s);  // also on a failed factor: nothing is left behind
        if (rj) return rj;
        if (rc == RS_OK) rc = solver_solve_dev(sol, c->d_rhs.p, c->d_xnew.p, s);
        if (rc != RS_OK) {
            if (rc == RS_ERR_DEVICE) return from_solver(c, sol, rc);
            r->reason = 2;  // SolverFailure
            c->err = sol->err;
            break;
        }
        xn.resize(n);
        e = hipMemcpyAsync(xn.data(), c->d_xnew.p, n * 8, hipMemcpyDeviceToHost, s);
        if (!e) e = hipStreamSynchronize(s);
        if (e) return chip(c, e, "hipMemcpy(x_new)");
        const i64 nd = std::min<i64>(n, (i64)x.size());
        dx.resize(nd);
        for (i64 i = 0; i < nd; i++) dx[i] = xn[i] - x[i];
        const double dxn = norm2(dx.data(), nd);
        r->final_norm = norm2(xn.data(), n);
        bool conv = true;
        for (i64 i = 0; i < nd; i++)
            if (!(std::fabs(dx[i]) <= o->abs_tol + o->rel_tol * std::fabs(xn[i]))) {
                conv = false;
                break;
            }
        if (conv) {
            x = xn;
            r->converged = 1;
            r->reason = 0;
            break;
        }
        if (dxn > prev && damping > o->damping_min) {
            double d2 = damping * 0.5;
            damping = d2 >= o->damping_min ? d2 : o->damping_min;
        }
        prev = dxn;
        for (i64 i = 0; i < nd; i++) x[i] = x[i] + damping * (xn[i] - x[i]);
    }
    return RS_OK;
}

// run_newton_with_stepping (newton.rs:142-185)
int32_t run_newton_stepping(rs_circuit* c, rs_solver* sol, const rs_newton_opts* o,
                            std::vector<double>& x, int mode, double t, double dt,
                            rs_newton_result* out, int32_t* iters_total, double* gmin_used = nullptr,
                            double* scale_used = nullptr) {
    double gstart = o->gmin * 1e3;
    if (gstart < 1e-6) gstart = 1e-6;
    rs_newton_result last{0, 0, 1, 0, 0.0};
    for (int gi = 0; gi <= o->gmin_steps; gi++) {
        double gmin = o->gmin_steps == 0
                          ? o->gmin
                          : gstart + (o->gmin - gstart) * ((double)gi / (double)o->gmin_steps);
        for (int si = 0; si <= o->source_steps; si++) {
            double scale = o->source_steps == 0 ? 1.0 : (double)si / (double)o->source_steps;
            rs_newton_result r;
            int32_t rc = run_newton(c, sol, o, x, mode, t, dt, gmin, scale, &r);
            if (rc) return rc;
            if (iters_total) *iters_total += r.iterations;
            if (r.converged) {
                *out = r;
                if (gmin_used) *gmin_used = gmin;
                if (scale_used) *scale_used = scale;
                return RS_OK;
            }
            last = r;
        }
    }
    *out = last;
    return RS_OK;
}

double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

struct Breakpoints {  // BreakpointManager (waveform.rs:380-645)
    std::vector<double> t;
    int settle = 0, steps = 5;
    double factor = 0.1;
    void add(double time) {
        if (time < 0.0) return;
        for (double v : t)
            if (v == time) return;
        t.insert(std::upper_bound(t.begin(), t.end(), time), time);
    }
    void pulse(const double* p, double tstop) {
        double td = p[RS_PP_TD], tr = p[RS_PP_TR], tf = p[RS_PP_TF], pw = p[RS_PP_PW], per = p[RS_PP_PER];
        double tt = td;
        int64_t cnt = 0;
        if (td > 0.0) add(td);
        while (tt <= tstop && cnt < 10000) {
            add(tt);
            double th = tt + tr;
            if (th > tstop) break;
            add(th);
            double tfa = th + pw;
            if (tfa > tstop) break;
            add(tfa);
            double tl = tfa + tf;
            if (tl > tstop) break;
            add(tl);
            if (per > 0.0) {
                tt = td + per * (double)(cnt + 1);
                cnt++;
            } else {
                break;
            }
        }
    }
    double limit_dt(double tc, double pdt, double mm) const {
        for (double b : t)
            if (b > tc) {
                double ttb = b - tc;
                if (pdt >= ttb) return ttb >= mm ? ttb : mm;
                if (pdt > 0.9 * ttb) return ttb;
                double rem = ttb - pdt;
                if (rem < 0.1 * pdt && rem < mm * 10.0) return ttb;
                break;
            }
        return pdt;
    }
    void update(double t0, double t1) {
        for (double b : t)
            if (b > t0 && b <= t1) {
                settle = steps;
                break;
            }
        if (settle > 0) settle--;
    }
};

void lte_milne(const std::vector<double>& xb, const std::vector<double>& xt, double abs_tol,
               double rel_tol, double out[5]) {  // analysis.rs:147-200
    double mr = 0.0, raw = 0.0;
    i64 mn = 0, n = (i64)std::min(xb.size(), xt.size());
    for (i64 i = 0; i < n; i++) {
        double be = xb[i], tr = xt[i];
        if (std::fabs(be) < 1e-30 && std::fabs(tr) < 1e-30) continue;
        double lte = std::fabs(tr - be) / 3.0;
        double ab = std::fabs(be), at = std::fabs(tr);
        double tol = abs_tol + rel_tol * (ab >= at ? ab : at);
        double ratio = tol > 0.0 ? lte / tol : 0.0;
        if (ratio > mr) {
            mr = ratio;
            mn = i;
            raw = lte;
        }
    }
    double f = 2.0;
    if (mr > 1e-10) f = clampd(std::pow(1.0 / mr, 0.5) * 0.9, 0.1, 2.0);
    out[0] = mr;
    out[1] = (double)mn;
    out[2] = mr <= 1.0 ? 1.0 : 0.0;
    out[3] = f;
    out[4] = raw;
}

// DC OP as one device-resident launch (the batched Newton kernel on a single
// instance). Equivalent to run_newton_with_stepping on the host path when the
// solver starts from scratch: the circuit has aux rows (so prepare(n) resets
// the solver, native_klu.rs:1604-1610) and the solver uses the default
// configuration. Afterwards the handle keeps the reference's factor/refactor
// counts but no factorization; callers that continue on the same pattern go
// through the host path (see can_run_device_dc).
constexpr int64_t kDeviceDcMaxN = 2048;

bool can_run_device_dc(rs_circuit* c, rs_solver* sol) {
    if (!c->device_newton || c->nbsim) return false;  // BSIM patterns change with the bias
    DevPattern* dp;
    if (get_pattern(c, MODE_DC, dp)) return false;
    if (dp->p.n == c->node_count) return false;
    // one wavefront per instance: large meshes run faster on the level-scheduled
    // multi-wave kernels, and W/X/XN/RPI/RP must fit in LDS
    if (dp->p.n > kDeviceDcMaxN) return false;
    if (sol->kind == RS_SOLVER_NATIVEKLU &&
        !(sol->pivot_tol == 0.001 && sol->ordering == 0 && sol->use_btf))
        return false;
    return true;
}

int32_t device_dc(rs_circuit* c, rs_solver* sol, const rs_newton_opts* o, std::vector<double>& x,
                  rs_newton_result* r, int32_t* iters_total) {
    const int k = sol->kind == RS_SOLVER_SPARSELU ? 1 : 0;
    if (!c->dcb[k]) {
        int64_t ptr[2] = {0, (int64_t)c->dev.size()};
        int64_t nc = c->node_count;
        c->dcb[k] = rs_batch_create(c->dev.data(), ptr, &nc, 1, sol->kind);
        char buf[256];
        if (rs_batch_last_error(c->dcb[k], buf, sizeof buf) > 0) return cfail(c, RS_ERR_DEVICE, buf);
    }
    int32_t rc = rs_batch_dc_op(c->dcb[k], o, nullptr);
    DevPattern* dp;
    get_pattern(c, MODE_DC, dp);
    std::vector<double> xo(dp->p.n);
    int64_t xp[2];
    if (!rc) rc = rs_batch_results(c->dcb[k], r, xo.data(), xp);
    if (rc) {
        char buf[256];
        rs_batch_last_error(c->dcb[k], buf, sizeof buf);
        return cfail(c, rc, buf);
    }
    const int32_t full = r->pad;
    r->pad = 0;
    if (r->converged)
        x = xo;
    else  // the last damped iterate, node_count entries (newton.rs:93-107)
        x.assign(xo.begin(), xo.begin() + std::min<size_t>(xo.size(), (size_t)c->node_count));
    // iterations of earlier stepping passes are not reported by the kernel;
    // the reference engine runs one pass (gmin_steps = source_steps = 0)
    if (iters_total) *iters_total += r->iterations;
    c->newton_iters += r->iterations;
    c->factor_calls += r->iterations;
    c->device_dc_runs++;
    rs_prepare(sol, dp->p.n);
    rs_reset_pattern(sol);
    sol->factor_count = full;
    sol->refactor_count = r->iterations - full;
    return RS_OK;
}

int32_t reset_cap_state(rs_circuit* c) {
    hipError_t e;
    size_t nd = c->dev.size();
    if ((e = hipMemset(c->d_has_v.p, 0, nd)) || (e = hipMemset(c->d_has_i.p, 0, nd)) ||
        (e = hipMemset(c->d_cap_v.p, 0, nd * 8)) || (e = hipMemset(c->d_cap_i.p, 0, nd * 8)))
        return chip(c, e, "hipMemset(state)");
    return RS_OK;
}

// Capacitor history update from a host x (uploaded to d_x).
int32_t cap_update(rs_circuit* c, const std::vector<double>& x, double dt, int full,
                   double* cv, double* ci, uint8_t* hv, uint8_t* hi) {
    hipError_t e = c->d_x.ensure(std::max(c->d_x.n, x.size()));
    if (e) return chip(c, e, "hipMalloc(x)");
    e = hipMemcpy(c->d_x.p, x.data(), x.size() * 8, hipMemcpyHostToDevice);
    if (e) return chip(c, e, "hipMemcpy(x)");
    e = launch_cap_state(c->d_dev.p, (int)c->dev.size(), c->d_x.p, (int)x.size(), dt, full, cv, ci,
                         hv, hi, nullptr);
    if (e) return chip(c, e, "k_cap_state");
    return RS_OK;
}

// Adjoint factor: the exact Jacobian of the stamped currents at x (mode DC
// or TRAN) in c->d_ax, factored by the mode's SparseLU solver (a device
// refactor when the pattern is unchanged), and the L^T levels and J's int32
// pattern of that factorization.
int32_t adj_factor(rs_circuit* c, int mode, const double* x_dev, i64 xl, double t, double dt,
                   double gmin) {
    DevPattern* dp;
    int32_t rc;
    if ((rc = stamp_dev(c, mode, x_dev, xl, t, dt, gmin, 1.0, &dp, nullptr, kStampJacobian)))
        return rc;
    const i64 n = dp->p.n, nd = (i64)c->dev.size();
    rs_circuit::AdjCtx& A = c->adj[mode == MODE_DC ? 0 : 1];
    if (!A.s) {
        // SparseLU in one block (the transposed solve needs one), AMD order
        // with the unknowns that have no diagonal entry (sources' aux rows)
        // last: their pivots then carry the Schur complement of the nodes
        // instead of a structural zero clamped to 1e-14
        A.s = rs_solver_create(RS_SOLVER_SPARSELU, n);
        A.s->ordering = 4;
        rs_set_supernodal(A.s, c->adj_sn_mode);
    }
    // SparseLU picks no pivots (pivot_tol 0), so a device refactor on the
    // stored symbolic analysis computes what a fresh factor would
    rs_prepare(A.s, n);
    rs_solver* a = A.s;
    hipError_t e;
    if ((rc = solver_analyze_dev(a, dp->p.ap.data(), n + 1, dp->p.ai.data(), (i64)dp->p.ai.size())))
        return from_solver(c, a, rc);
    if (a->sn) {  // J^T factored supernodally from J's values; J's own factor is not needed
        if (!A.snT || A.snT_analysis != a->analyses) {
            sn_destroy(A.snT);
            std::string why;
            A.snT = sn_create(sn_symbolic(a->sn), dp->p.ap.data(), dp->p.ai.data(), &why, true);
            if (!A.snT) return cfail(c, RS_ERR_DEVICE, "adjoint transposed supernodal plan: " + why);
            A.snT_analysis = a->analyses;
            // J's int32 pattern for the refinement's residual, and every
            // device's aux unknown for the contraction
            std::vector<int32_t> jap(n + 1), jai(std::max<size_t>(dp->p.ai.size(), 1)), aux(std::max<i64>(nd, 1));
            for (i64 i = 0; i <= n; i++) jap[i] = (int32_t)dp->p.ap[i];
            for (size_t q = 0; q < dp->p.ai.size(); q++) jai[q] = (int32_t)dp->p.ai[q];
            for (i64 k = 0; k < nd; k++) aux[k] = dp->p.aux_of[k];
            if ((e = A.jap.ensure(n + 1)) || (e = A.jai.ensure(jai.size())) || (e = A.aux.ensure(aux.size())) ||
                (e = hipMemcpy(A.jap.p, jap.data(), jap.size() * 4, hipMemcpyHostToDevice)) ||
                (e = hipMemcpy(A.jai.p, jai.data(), jai.size() * 4, hipMemcpyHostToDevice)) ||
                (e = hipMemcpy(A.aux.p, aux.data(), aux.size() * 4, hipMemcpyHostToDevice)))
                return chip(c, e, "adjoint pattern");
        }
        const int ps = prof_begin(PK_REFACTOR, nullptr);
        e = sn_factor(A.snT, c->d_ax.p, nullptr);
        prof_end(ps, nullptr);
        if (e) return chip(c, e, "adjoint transposed supernodal factor");
        return RS_OK;
    }
    if ((rc = solver_factor_dev(a, dp->p.ap.data(), n + 1, dp->p.ai.data(), (i64)dp->p.ai.size(),
                                c->d_ax.p, nullptr)))
        return from_solver(c, a, rc);
    if (a->pd.nblk != 1) return cfail(c, RS_ERR_SOLVE_FAILED, "adjoint needs a single-block factor");
    if (A.factors != a->factor_count || !A.ltp.p) {  // L^T levels of this factorization
        std::vector<int32_t> ptr, cols;
        lt_levels(a->ph, n, ptr, cols);
        const int32_t nlev = (int32_t)ptr.size() - 1;
        std::vector<int32_t> aux(nd);
        for (i64 k = 0; k < nd; k++) aux[k] = dp->p.aux_of[k];
        if ((e = A.ltp.ensure(ptr.size())) || (e = A.ltc.ensure(std::max<size_t>(cols.size(), 1))) ||
            (e = A.aux.ensure(std::max<i64>(nd, 1))) ||
            (e = hipMemcpy(A.ltp.p, ptr.data(), ptr.size() * 4, hipMemcpyHostToDevice)) ||
            (!cols.empty() && (e = hipMemcpy(A.ltc.p, cols.data(), cols.size() * 4, hipMemcpyHostToDevice))) ||
            (nd && (e = hipMemcpy(A.aux.p, aux.data(), nd * 4, hipMemcpyHostToDevice))))
            return chip(c, e, "adjoint levels");
        std::vector<int32_t> jap(n + 1), jai(std::max<size_t>(dp->p.ai.size(), 1));
        for (i64 i = 0; i <= n; i++) jap[i] = (int32_t)dp->p.ap[i];
        for (size_t q = 0; q < dp->p.ai.size(); q++) jai[q] = (int32_t)dp->p.ai[q];
        if ((e = A.jap.ensure(n + 1)) || (e = A.jai.ensure(jai.size())) ||
            (e = hipMemcpy(A.jap.p, jap.data(), jap.size() * 4, hipMemcpyHostToDevice)) ||
            (e = hipMemcpy(A.jai.p, jai.data(), jai.size() * 4, hipMemcpyHostToDevice)))
            return chip(c, e, "adjoint pattern");
        A.n_lt = nlev;
        std::vector<int32_t> tpool;
        PlanDev td = a->pd;
        A.t_ok = build_transposed_records(a->ph, tpool, td) &&
                 solve_lp_lds((int)n, td.sq_maxent, td.sq_nlev) <= device_max_lds();
        if (A.t_ok) {
            td.o_vec = 0;
            A.t_nlev = td.sq_nlev;
            A.t_maxent = td.sq_maxent;
            if ((e = A.tip.ensure(tpool.size())) || (e = A.tdesc.ensure(1)) ||
                (e = hipMemcpy(A.tip.p, tpool.data(), tpool.size() * 4, hipMemcpyHostToDevice)) ||
                (e = hipMemcpy(A.tdesc.p, &td, sizeof td, hipMemcpyHostToDevice)))
                return chip(c, e, "adjoint transposed records");
        }
        A.factors = a->factor_count;
    }
    return RS_OK;
}

// J^T lam = rhs with the factor of the last adj_factor of this mode (J still
// in c->d_ax), refined against the exact J^T (a pivot below SparseLU's 1e-14
// clamp would otherwise leave its rank-one error in lambda).
int32_t adj_solve(rs_circuit* c, int mode, const double* rhs_dev, double* lam_dev, int max_refine = 12) {
    rs_circuit::AdjCtx& A = c->adj[mode == MODE_DC ? 0 : 1];
    rs_solver* a = A.s;
    const i64 n = a->n;
    hipStream_t s = nullptr;
    hipError_t e;
    auto tsolve = [&](const double* b, double* y) -> hipError_t {
        if (a->sn && A.snT) {
            const int ps = prof_begin(PK_SOLVE, s);
            hipError_t er = sn_solve(A.snT, b, y, s);
            prof_end(ps, s);
            return er;
        }
        if (A.t_ok)
            return launch_solve_lp(A.tdesc.p, A.tip.p, a->d_vals.p, b, y, (int)n, A.t_maxent, A.t_nlev, 256, s);
        return launch_solve_t(a->d_desc.p, a->d_ip.p, a->d_vals.p, b, A.ltp.p, A.ltc.p, A.n_lt, c->d_t.p, y, s);
    };
    if ((e = c->d_t.ensure(n)) || (e = c->d_r.ensure(n)) ||
        (e = c->d_del.ensure(n)) || (e = c->d_nrm.ensure(2)) || (e = tsolve(rhs_dev, lam_dev)))
        return chip(c, e, "adjoint solve");
    double prev = 1.7976931348623157e308;
    for (int it = 0; it < max_refine; ++it) {
        double nrm[2] = {0.0, 1.0};
        const bool last = it + 1 == max_refine && max_refine < 12;  // a fixed count: no read-back
        if ((e = launch_resid_t(A.jap.p, A.jai.p, c->d_ax.p, rhs_dev, lam_dev, c->d_r.p, (int)n, s)) ||
            (e = tsolve(c->d_r.p, c->d_del.p)) ||
            (e = launch_refine_update(lam_dev, c->d_del.p, (int)n, c->d_nrm.p, s)) ||
            (!last && (e = hipMemcpy(nrm, c->d_nrm.p, sizeof nrm, hipMemcpyDeviceToHost))))
            return chip(c, e, "adjoint refinement");
        if (last) {
            c->adj_refine = std::max(c->adj_refine, it + 1);
            break;
        }
        c->adj_refine = std::max(c->adj_refine, it + 1);
        // converged, or stalled at the round-off of the residual (also stops on NaN)
        if (!(nrm[0] > 1e-14 * nrm[1]) || (nrm[0] <= 1e-9 * nrm[1] && nrm[0] >= 0.5 * prev)) break;
        prev = nrm[0];
    }
    return RS_OK;
}

// Newton on the exact Jacobian in correction form: the Jacobian stamp's
// right-hand side is consistent with its entries (J x - rhs is the stamped
// current), so dx = J^{-1} (rhs - J x) is a Newton step of the stamped
// currents with their exact derivative; x converges quadratically to the
// step's exact root, and the residual form keeps the solve's round-off
// relative to the shrinking residual. The factor is the mode's adjoint solver
// (SparseLU, zero-diagonal unknowns last: no clamped pivot). After the first
// step the factor is reused (chord steps: the residual is always the exact
// one) while each correction is at most a tenth of the one before, and
// refactored at x otherwise. Stops when every |dx| <= 1e-12 + 1e-10 |x| (the
// update then leaves an error of a small fraction of |dx|), or when a step on
// a fresh factor stalls at round-off below 1e-8 |x|.
int32_t exact_newton(rs_circuit* c, std::vector<double>& x, int mode, double t, double dt,
                     double gmin, double scale, bool* ok) {
    rs_circuit::AdjCtx& A = c->adj[mode == MODE_DC ? 0 : 1];
    hipError_t e;
    int32_t rc;
    *ok = false;
    double prev = 1.7976931348623157e308;
    bool refactor = true;
    {
        // the pattern's device vectors first: building them may grow d_x,
        // which must not happen between the upload of x and its last read
        DevPattern* dp0;
        if ((rc = get_pattern(c, mode, dp0))) return rc;
    }
    for (int it = 0; it < 50; ++it) {
        const i64 xl = (i64)x.size();
        if ((e = c->d_x.ensure(std::max(c->d_x.n, x.size()))) ||
            (e = hipMemcpy(c->d_x.p, x.data(), xl * 8, hipMemcpyHostToDevice)))
            return chip(c, e, "exact Newton x");
        DevPattern* dp;
        if ((rc = stamp_dev(c, mode, c->d_x.p, xl, t, dt, gmin, scale, &dp, nullptr, kStampJacobian)))
            return rc;
        const i64 n = dp->p.n;
        if (!A.s) {
            A.s = rs_solver_create(RS_SOLVER_SPARSELU, n);
            A.s->ordering = 4;
            rs_set_supernodal(A.s, c->adj_sn_mode);
        }
        if (!A.have_rows) {  // row map of the pattern, columns ascending within a row
            std::vector<int32_t> rp(n + 1, 0), rcol(std::max<size_t>(dp->p.ai.size(), 1)), rq(rcol.size());
            for (size_t q = 0; q < dp->p.ai.size(); q++)
                if (dp->p.ai[q] < n) rp[dp->p.ai[q] + 1]++;
            for (i64 i = 0; i < n; i++) rp[i + 1] += rp[i];
            std::vector<int32_t> f(rp.begin(), rp.end() - 1);
            for (i64 j = 0; j < n; j++)
                for (i64 q = dp->p.ap[j]; q < dp->p.ap[j + 1]; q++)
                    if (dp->p.ai[q] < n) {
                        const int32_t at = f[dp->p.ai[q]]++;
                        rcol[at] = (int32_t)j;
                        rq[at] = (int32_t)q;
                    }
            if ((e = A.rp.ensure(n + 1)) || (e = A.rc.ensure(rcol.size())) || (e = A.rq.ensure(rq.size())) ||
                (e = hipMemcpy(A.rp.p, rp.data(), (n + 1) * 4, hipMemcpyHostToDevice)) ||
                (e = hipMemcpy(A.rc.p, rcol.data(), rcol.size() * 4, hipMemcpyHostToDevice)) ||
                (e = hipMemcpy(A.rq.p, rq.data(), rq.size() * 4, hipMemcpyHostToDevice)))
                return chip(c, e, "exact Newton rows");
            A.have_rows = true;
        }
        if ((e = c->d_r.ensure(n)) || (e = c->d_del.ensure(n)) ||
            (e = launch_resid_n(A.rp.p, A.rc.p, A.rq.p, c->d_ax.p, c->d_rhs.p, c->d_x.p, (int)xl, c->d_r.p,
                                (int)n, nullptr)))
            return chip(c, e, "exact Newton residual");
        const bool full = refactor || A.s->state != ST_FACTORED || A.s->n != n;
        if (full) {
            rs_prepare(A.s, n);
            if ((rc = solver_factor_dev(A.s, dp->p.ap.data(), n + 1, dp->p.ai.data(), (i64)dp->p.ai.size(),
                                        c->d_ax.p, nullptr)))
                return from_solver(c, A.s, rc);
        }
        if ((rc = solver_solve_dev(A.s, c->d_r.p, c->d_del.p, nullptr))) return from_solver(c, A.s, rc);
        std::vector<double> dx(n);
        if ((e = hipMemcpy(dx.data(), c->d_del.p, n * 8, hipMemcpyDeviceToHost))) return chip(c, e, "exact Newton dx");
        x.resize(n, 0.0);
        bool conv = true;
        double rel = 0.0;
        for (i64 i = 0; i < n; i++) {
            x[i] += dx[i];
            const double ad = std::fabs(dx[i]);
            if (!(ad <= 1e-12 + 1e-10 * std::fabs(x[i]))) conv = false;
            rel = std::max(rel, ad / (1e-12 + std::fabs(x[i])));
        }
        c->polish_iters++;
        if (!(rel == rel)) return RS_OK;  // NaN: not converged
        // a stall counts only on a step with a fresh factor (a chord step
        // that contracts slowly is followed by one)
        if (conv || (full && rel <= 1e-8 && rel >= 0.5 * prev)) {
            *ok = true;
            return RS_OK;
        }
        refactor = !(rel <= 0.1 * prev);
        prev = rel;
    }
    return RS_OK;
}

// Steps of the fixed-step TRAP transient from time t as one captured graph
// each: kGraphIters Newton iterations of device evaluation, assembly,
// refactor, solve and k_newton_update (the host loop's convergence test and
// damping), every kernel returning at once after the update has set done;
// then the accepted-step capacitor history, gated on convergence. The host
// reads the 32-byte state once per graph launch and relaunches while a step
// has neither converged nor spent max_iters. Results equal the host-driven
// loop's bit for bit. *done: steps completed (-1: not eligible, nothing run).
int32_t tran_fixed_graph(rs_circuit* c, rs_solver* sol, const rs_newton_opts* o, std::vector<double>& x,
                         double t, double dt, int64_t nsteps, int64_t* done, int32_t* iters, int32_t* status) {
    constexpr int kGraphIters = 2;
    *done = -1;
    DevPattern* dp;
    int32_t rc;
    if (c->nbsim || o->gmin_steps != 0 || o->source_steps != 0 || !solver_graph_ready(sol)) return RS_OK;
    if ((rc = get_pattern(c, MODE_TRAN_TRAP, dp))) return rc;
    const i64 n = dp->p.n;
    if ((i64)x.size() != n || sol->n != n || sol->last_ap != dp->p.ap || sol->last_ai != dp->p.ai)
        return RS_OK;  // the solver holds another pattern
    hipError_t e;
    DevBuf<double> d_t;
    DevBuf<int32_t> d_st;  // done, converged, iterations, pad, damping, previous |dx|
    hipStream_t s = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto cleanup = [&]() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (s) (void)hipStreamDestroy(s);
    };
    if ((e = d_t.ensure(1)) || (e = d_st.ensure(8)) || (e = hipStreamCreate(&s)) ||
        (e = hipMemcpy(d_t.p, &t, 8, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(c->d_x.p, x.data(), n * 8, hipMemcpyHostToDevice))) {
        cleanup();
        return chip(c, e, "graph setup");
    }
    int32_t* st = d_st.p;
    StampArgs a{0.0, dt, o->gmin, 1.0, MODE_TRAN_TRAP, 0, c->d_models.p, d_t.p, st};
    if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal))) {
        cleanup();
        return chip(c, e, "hipStreamBeginCapture");
    }
    for (int m = 0; m < kGraphIters && !e; m++) {
        if ((e = launch_eval(c->d_dev.p, (int)c->dev.size(), c->d_x.p, (int)n, a, c->d_cap_v.p, c->d_cap_i.p,
                             c->d_has_v.p, c->d_has_i.p, c->d_sv.p, s, nullptr, c->d_gmask.p)) ||
            (e = launch_assemble(dp->ax_src.p, (int)dp->p.ai.size(), dp->rhs_ptr.p, dp->rhs_src.p, (int)n,
                                 c->d_sv.p, c->d_ax.p, c->d_rhs.p, dp->long_rows.p, dp->nlong, s, st)) ||
            (e = solver_graph_iteration(sol, c->d_ax.p, c->d_rhs.p, c->d_xnew.p, s, st)) ||
            (e = launch_newton_update((int)n, c->d_x.p, c->d_xnew.p, o->abs_tol, o->rel_tol, o->damping_min,
                                      o->max_iters, st, s)))
            break;
    }
    if (!e)
        e = launch_cap_state(c->d_dev.p, (int)c->dev.size(), c->d_x.p, (int)n, dt, 1, c->d_cap_v.p,
                             c->d_cap_i.p, c->d_has_v.p, c->d_has_i.p, s, st + 1);
    hipError_t e2 = hipStreamEndCapture(s, &graph);
    if (e || (e = e2) || (e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0))) {
        cleanup();
        return chip(c, e, "graph capture");
    }
    int64_t k = 0;
    for (; k < nsteps; k++) {
        int32_t h[8];
        if ((e = launch_newton_begin(d_t.p, dt, st, o->damping, s))) break;
        do {
            if ((e = hipGraphLaunch(exec, s)) || (e = hipMemcpyAsync(h, st, sizeof h, hipMemcpyDeviceToHost, s)) ||
                (e = hipStreamSynchronize(s)))
                break;
            c->graph_launches++;
        } while (!h[0]);
        if (e) break;
        *iters += h[2];
        c->newton_iters += h[2];
        c->factor_calls += h[2];
        c->stamp_calls += h[2];
        sol->refactor_count += h[2];
        c->graph_steps++;
        if (!h[1]) {
            *status = 2;
            break;
        }
    }
    if (!e) e = hipMemcpy(x.data(), c->d_x.p, n * 8, hipMemcpyDeviceToHost);
    cleanup();
    if (e) return chip(c, e, "graph step");
    *done = k;
    return RS_OK;
}

// Loop state of rs_tran handed to and back from the device controller.
struct TranLoop {
    double t, dt, prev_err;
    int32_t has_prev, ctl_consec, consec, settle, status;
    int32_t accepted, rejected, iters;
    int64_t np;
};

// The rest of rs_tran's adaptive loop (engine.rs:463-602) as captured graphs
// of kTranSlots Newton-iteration slots with the step controller on the device
// (tran_ctl.hpp). Starts from an accepted step: x holds the TRAN solution, the
// capacitor histories are current (TRAP in the circuit's arrays, BE in be_*)
// and the solver holds the TRAN pattern's factor. The host reads the state
// block and the new points once per graph launch, and evaluates the
// controller's powers with libm when the device finds them within 0.02 ulp of
// a rounding midpoint. *ran = false (nothing changed) when the solver or the
// circuit cannot run this way.
int32_t tran_adaptive_graph(rs_circuit* c, rs_solver* sol, const rs_newton_opts* o, std::vector<double>& x,
                            TranLoop& L, const std::vector<double>& bps, double tstop, double min_dt,
                            double max_dt, double* be_v, double* be_i, uint8_t* be_hv, uint8_t* be_hi,
                            const std::function<void(double, const double*)>& store, bool* ran) {
    *ran = false;
    DevPattern* dp;
    int32_t rc;
    if (c->nbsim || o->gmin_steps != 0 || o->source_steps != 0 || !solver_graph_ready(sol)) return RS_OK;
    if ((rc = get_pattern(c, MODE_TRAN_TRAP, dp))) return rc;
    const i64 n = dp->p.n;
    if ((i64)x.size() != n || sol->n != n || sol->last_ap != dp->p.ap || sol->last_ai != dp->p.ai)
        return RS_OK;
    hipError_t e;
    DevBuf<TranCtl> d_ctl;
    DevBuf<double> d_xa, d_xbe, d_bp, d_rt, d_rx;
    hipStream_t s = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto cleanup = [&]() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (s) (void)hipStreamDestroy(s);
    };
    TranCtl h{};
    h.t = L.t;
    h.dt = L.dt;
    h.prev_err = L.prev_err;
    h.has_prev = L.has_prev;
    h.ctl_consec = L.ctl_consec;
    h.consec = L.consec;
    h.settle = L.settle;
    h.accepted = L.accepted;
    h.rejected = L.rejected;
    h.npoints = L.np;
    h.nw_done = 1;
    h.phase = L.t < tstop ? kPhaseBegin : kPhaseFinished;
    h.mode = MODE_TRAN_BE;
    if ((e = d_ctl.ensure(1)) || (e = d_xa.ensure(n)) || (e = d_xbe.ensure(n)) ||
        (e = d_bp.ensure(std::max<size_t>(bps.size(), 1))) || (e = d_rt.ensure(kTranRing)) ||
        (e = d_rx.ensure((size_t)kTranRing * n)) || (e = c->d_x.ensure(std::max<size_t>(c->d_x.n, n))) ||
        (e = hipStreamCreate(&s)) ||
        (e = hipMemcpy(d_ctl.p, &h, sizeof h, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(d_xa.p, x.data(), n * 8, hipMemcpyHostToDevice)) ||
        (!bps.empty() && (e = hipMemcpy(d_bp.p, bps.data(), bps.size() * 8, hipMemcpyHostToDevice)))) {
        cleanup();
        return chip(c, e, "adaptive graph setup");
    }
    TranParams P{tstop, min_dt, max_dt, o->abs_tol, o->rel_tol, o->damping, d_bp.p, (int32_t)bps.size(), (int32_t)n};
    TranCtl* ctl = d_ctl.p;
    int32_t* st = reinterpret_cast<int32_t*>(ctl);  // the Newton state block leads the struct
    StampArgs a{0.0, 0.0, o->gmin, 1.0, MODE_TRAN_TRAP, 0, c->d_models.p, &ctl->t_target, st};
    a.mode_dev = &ctl->mode;
    a.dt_dev = &ctl->dt_step;
    a.be_cap_v = be_v;
    a.be_cap_i = be_i;
    a.be_has_v = be_hv;
    a.be_has_i = be_hi;
    if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal))) {
        cleanup();
        return chip(c, e, "hipStreamBeginCapture");
    }
    const int nd = (int)c->dev.size();
    for (int m = 0; m < kTranSlots && !e; m++) {
        if ((e = launch_tctl_pre(ctl, P, s)) || (e = launch_gated_copy(c->d_x.p, d_xa.p, n, &ctl->cp_begin, s)) ||
            (e = launch_eval(c->d_dev.p, nd, c->d_x.p, (int)n, a, c->d_cap_v.p, c->d_cap_i.p, c->d_has_v.p,
                             c->d_has_i.p, c->d_sv.p, s, nullptr, c->d_gmask.p)) ||
            (e = launch_assemble(dp->ax_src.p, (int)dp->p.ai.size(), dp->rhs_ptr.p, dp->rhs_src.p, (int)n,
                                 c->d_sv.p, c->d_ax.p, c->d_rhs.p, dp->long_rows.p, dp->nlong, s, st)) ||
            (e = solver_graph_iteration(sol, c->d_ax.p, c->d_rhs.p, c->d_xnew.p, s, st)) ||
            (e = launch_newton_update((int)n, c->d_x.p, c->d_xnew.p, o->abs_tol, o->rel_tol, o->damping_min,
                                      o->max_iters, st, s)) ||
            (e = launch_tctl_post(ctl, P, d_xbe.p, c->d_x.p, s)) ||
            (e = launch_gated_copy(d_xbe.p, c->d_x.p, n, &ctl->cp_be, s)) ||
            (e = launch_gated_copy(c->d_x.p, d_xa.p, n, &ctl->cp_begin, s)) ||
            (e = launch_tstore(ctl, c->d_x.p, d_xa.p, d_rt.p, d_rx.p, n, s)) ||
            (e = launch_cap_state(c->d_dev.p, nd, d_xa.p, (int)n, 0.0, 1, c->d_cap_v.p, c->d_cap_i.p,
                                  c->d_has_v.p, c->d_has_i.p, s, &ctl->accept_now, &ctl->dt_step)) ||
            (e = launch_cap_state(c->d_dev.p, nd, d_xa.p, (int)n, 0.0, 0, be_v, be_i, be_hv, be_hi, s,
                                  &ctl->accept_now, &ctl->dt_step)))
            break;
    }
    hipError_t e2 = hipStreamEndCapture(s, &graph);
    if (e || (e = e2) || (e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0))) {
        cleanup();
        return chip(c, e, "adaptive graph capture");
    }
    *ran = true;
    std::vector<double> rx((size_t)n);
    int64_t np_host = L.np;
    const int64_t fc0 = c->factor_calls;
    int32_t it0 = 0;
    for (;;) {
        if ((e = hipGraphLaunch(exec, s)) || (e = hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, s)) ||
            (e = hipStreamSynchronize(s)))
            break;
        c->graph_launches++;
        for (; np_host < h.npoints; np_host++) {  // new points, oldest first
            const int64_t slot = np_host % kTranRing;
            double tt;
            if ((e = hipMemcpy(&tt, d_rt.p + slot, 8, hipMemcpyDeviceToHost)) ||
                (e = hipMemcpy(rx.data(), d_rx.p + slot * n, n * 8, hipMemcpyDeviceToHost)))
                break;
            store(tt, rx.data());
        }
        if (e) break;
        if (h.phase == kPhaseHostPow) {  // libm's pow where the device's power is near a midpoint
            const double ex[3] = {0.7 / 2.0, 0.4 / 2.0, 0.5};
            for (int j = 0; j < 3; j++)
                if (h.pw_need & (1 << j)) h.pw_out[j] = std::pow(h.pw_in[j], ex[j]);
            h.phase = kPhaseResume;
            c->tran_host_pows++;
            if ((e = hipMemcpy(&ctl->pw_out[0], h.pw_out, sizeof h.pw_out, hipMemcpyHostToDevice)) ||
                (e = hipMemcpy(&ctl->phase, &h.phase, 4, hipMemcpyHostToDevice)))
                break;
            continue;
        }
        if (h.phase == kPhaseFinished) break;
    }
    if (!e) e = hipMemcpy(x.data(), d_xa.p, n * 8, hipMemcpyDeviceToHost);
    cleanup();
    if (e) return chip(c, e, "adaptive graph step");
    (void)it0;
    const int32_t its = h.iters_total;
    c->newton_iters += its;
    c->stamp_calls += its;
    sol->refactor_count += its;
    c->factor_calls = fc0 + its;
    L.t = h.t;
    L.dt = h.dt;
    L.prev_err = h.prev_err;
    L.has_prev = h.has_prev;
    L.ctl_consec = h.ctl_consec;
    L.consec = h.consec;
    L.settle = h.settle;
    L.status = h.status;
    L.accepted = h.accepted;
    L.rejected = h.rejected;
    L.iters += its;
    L.np = h.npoints;
    return RS_OK;
}

}  // namespace

extern "C" {

rs_circuit* rs_circuit_create(const rs_device* dev, int64_t ndev, int64_t node_count) {
    rs_circuit* c = new rs_circuit();
    c->dev.assign(dev, dev + ndev);
    c->node_count = node_count;
    hipError_t e;
    size_t nd = (size_t)ndev;
    i64 nslots = nd * kSlotsPerDevice;
    try {
        nslots = stamp_slot_bases(dev, ndev, c->sbase);
    } catch (const std::exception& ex) {
        c->err = ex.what();
        return c;
    }
    for (i64 d = 0; d < ndev; d++) c->nbsim += is_bsim(dev[d]) ? 1 : 0;
    c->gmask.assign(c->nbsim ? nd : 0, 0);
    if (c->nbsim &&
        ((e = c->d_sbase.ensure(nd)) || (e = c->d_gmask.ensure(nd)) ||
         (e = hipMemcpy(c->d_sbase.p, c->sbase.data(), nd * 4, hipMemcpyHostToDevice)) ||
         (e = hipMemset(c->d_gmask.p, 0, nd)))) {
        c->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        return c;
    }
    if ((e = c->d_dev.ensure(nd)) || (e = c->d_sv.ensure((size_t)nslots)) ||
        (e = c->d_cap_v.ensure(nd)) || (e = c->d_cap_i.ensure(nd)) || (e = c->d_has_v.ensure(nd)) ||
        (e = c->d_has_i.ensure(nd))) {
        c->err = std::string("hipMalloc: ") + hipGetErrorString(e);
        return c;
    }
    if (nd && (e = hipMemcpy(c->d_dev.p, dev, nd * sizeof(rs_device), hipMemcpyHostToDevice)))
        c->err = std::string("hipMemcpy: ") + hipGetErrorString(e);
    else if (nd && (e = hipMemset(c->d_sv.p, 0, (size_t)nslots * 8)))
        c->err = std::string("hipMemset: ") + hipGetErrorString(e);
    else
        reset_cap_state(c);
    return c;
}

void rs_circuit_destroy(rs_circuit* c) {
    if (!c) return;
    if (c->side) {
        (void)hipStreamSynchronize(c->side);
        (void)hipEventDestroy(c->ev_sv);
        (void)hipEventDestroy(c->ev_rhs);
        (void)hipStreamDestroy(c->side);
    }
    for (auto& a : c->adj) {
        sn_destroy(a.snT);
        if (a.s) rs_solver_destroy(a.s);
    }
    for (auto*& b : c->dcb)
        if (b) rs_batch_destroy(b);
    for (auto& p : c->pat) {
        p.ax_src.release();
        p.rhs_ptr.release();
        p.rhs_src.release();
        p.long_rows.release();
    }
    c->d_dev.release();
    c->d_sv.release();
    c->d_ax.release();
    c->d_rhs.release();
    c->d_x.release();
    c->d_xnew.release();
    c->d_cap_v.release();
    c->d_cap_i.release();
    c->d_has_v.release();
    c->d_has_i.release();
    c->d_sbase.release();
    c->d_gmask.release();
    c->d_models.release();
    delete c;
}

void rs_circuit_set_device_newton(rs_circuit* c, int32_t on) { c->device_newton = on != 0; }

int32_t rs_mos_eval(const rs_bsim_model* models, int64_t nmodels, const int32_t* model,
                    const int32_t* level, const double* bias, int64_t n, double* out) {
    if (n <= 0) return RS_OK;
    for (int64_t i = 0; i < n; i++)
        if (model[i] < 0 || model[i] >= nmodels || level[i] == 1) return RS_ERR_INVALID_MATRIX;
    DevBuf<rs_bsim_model> dm;
    DevBuf<int32_t> dmi, dlv;
    DevBuf<double> din, dout;
    hipError_t e;
    if ((e = dm.ensure((size_t)nmodels)) || (e = dmi.ensure((size_t)n)) || (e = dlv.ensure((size_t)n)) ||
        (e = din.ensure((size_t)n * 7)) || (e = dout.ensure((size_t)n * 13)) ||
        (e = hipMemcpy(dm.p, models, (size_t)nmodels * sizeof(rs_bsim_model), hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(dmi.p, model, (size_t)n * 4, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(dlv.p, level, (size_t)n * 4, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(din.p, bias, (size_t)n * 7 * 8, hipMemcpyHostToDevice)) ||
        (e = launch_mos_eval(dm.p, dmi.p, dlv.p, din.p, n, dout.p, nullptr)) ||
        (e = hipMemcpy(out, dout.p, (size_t)n * 13 * 8, hipMemcpyDeviceToHost)))
        return RS_ERR_DEVICE;
    return RS_OK;
}

int32_t rs_circuit_set_bsim_models(rs_circuit* c, const rs_bsim_model* models, int64_t n) {
    for (const auto& d : c->dev)
        if (is_bsim(d) && !(d.p[RS_MP_MODEL] >= 0.0 && d.p[RS_MP_MODEL] < (double)n))
            return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: BSIM model index out of range");
    hipError_t e = c->d_models.ensure((size_t)std::max<int64_t>(n, 1));
    if (!e && n) e = hipMemcpy(c->d_models.p, models, (size_t)n * sizeof(rs_bsim_model), hipMemcpyHostToDevice);
    if (e) return chip(c, e, "hipMemcpy(models)");
    c->nmodels = n;
    return RS_OK;
}

int32_t rs_circuit_last_error(const rs_circuit* c, char* buf, int64_t len) {
    if (len > 0) std::snprintf(buf, (size_t)len, "%s", c->err.c_str());
    return (int32_t)c->err.size();
}

int32_t rs_circuit_pattern(rs_circuit* c, int32_t mode, int64_t* n, int64_t* nnz, int64_t* ap,
                           int64_t* ai) {
    DevPattern* dp;
    int32_t rc = get_pattern(c, mode, dp);
    if (rc) return rc;
    if (n) *n = dp->p.n;
    if (nnz) *nnz = (int64_t)dp->p.ai.size();
    if (ap) std::memcpy(ap, dp->p.ap.data(), dp->p.ap.size() * 8);
    if (ai && !dp->p.ai.empty()) std::memcpy(ai, dp->p.ai.data(), dp->p.ai.size() * 8);
    return RS_OK;
}

int32_t rs_circuit_set_cap_state(rs_circuit* c, const double* cap_v, const double* cap_i,
                                 const uint8_t* has_v, const uint8_t* has_i) {
    size_t nd = c->dev.size();
    hipError_t e;
    if ((e = hipMemcpy(c->d_cap_v.p, cap_v, nd * 8, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(c->d_cap_i.p, cap_i, nd * 8, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(c->d_has_v.p, has_v, nd, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(c->d_has_i.p, has_i, nd, hipMemcpyHostToDevice)))
        return chip(c, e, "hipMemcpy(state)");
    return RS_OK;
}

int32_t rs_stamp(rs_circuit* c, int32_t mode, const double* x, int64_t x_len, double t, double dt,
                 double gmin, double source_scale, double* ax_out, int64_t ax_cap,
                 double* rhs_out, int64_t rhs_cap, int64_t* nnz_out, int64_t* n_out) {
    DevPattern* dp;
    int32_t rc = get_pattern(c, mode, dp);
    if (rc) return rc;
    hipError_t e = c->d_x.ensure(std::max<size_t>(c->d_x.n, (size_t)x_len));
    if (e) return chip(c, e, "hipMalloc(x)");
    e = hipMemcpy(c->d_x.p, x, x_len * 8, hipMemcpyHostToDevice);
    if (e) return chip(c, e, "hipMemcpy(x)");
    rc = stamp_dev(c, mode, c->d_x.p, x_len, t, dt, gmin, source_scale, &dp, nullptr);
    if (rc) return rc;
    if (nnz_out) *nnz_out = (int64_t)dp->p.ai.size();
    if (n_out) *n_out = dp->p.n;
    if ((int64_t)dp->p.ai.size() > ax_cap || dp->p.n > rhs_cap)
        return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: stamp output buffers too small");
    if ((e = hipMemcpy(ax_out, c->d_ax.p, dp->p.ai.size() * 8, hipMemcpyDeviceToHost)) ||
        (e = hipMemcpy(rhs_out, c->d_rhs.p, dp->p.n * 8, hipMemcpyDeviceToHost)))
        return chip(c, e, "hipMemcpy(out)");
    return RS_OK;
}

int32_t rs_stamp_device(rs_circuit* c, int32_t mode, const double* x_dev, int64_t x_len, double t, double dt,
                        double gmin, double source_scale, double* ax_dev, int64_t ax_cap, double* rhs_dev,
                        int64_t rhs_cap, int64_t* nnz_out, int64_t* n_out, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    DevPattern* dp;
    int32_t rc = get_pattern(c, mode, dp);
    if (rc) return rc;
    rc = stamp_dev(c, mode, x_dev, x_len, t, dt, gmin, source_scale, &dp, s);
    if (rc) return rc;
    if (nnz_out) *nnz_out = (int64_t)dp->p.ai.size();
    if (n_out) *n_out = dp->p.n;
    if ((int64_t)dp->p.ai.size() > ax_cap || dp->p.n > rhs_cap)
        return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: stamp output buffers too small");
    hipError_t e;
    if ((e = hipMemcpyAsync(ax_dev, c->d_ax.p, dp->p.ai.size() * 8, hipMemcpyDeviceToDevice, s)) ||
        (e = hipMemcpyAsync(rhs_dev, c->d_rhs.p, dp->p.n * 8, hipMemcpyDeviceToDevice, s)))
        return chip(c, e, "hipMemcpy(out)");
    return RS_OK;
}

int32_t rs_dc_op(rs_circuit* c, rs_solver* sol, const rs_newton_opts* o, double* x_out,
                 int64_t x_cap, int64_t* x_len, rs_newton_result* res) {
    std::vector<double> x(c->node_count, 0.0);
    rs_prepare(sol, c->node_count);  // engine.rs:254
    rs_newton_result r{};
    int32_t rc = (can_run_device_dc(c, sol) && o->gmin_steps == 0 && o->source_steps == 0)
                     ? device_dc(c, sol, o, x, &r, nullptr)
                     : run_newton_stepping(c, sol, o, x, MODE_DC, 0.0, 0.0, &r, nullptr);
    if (rc) return rc;
    *res = r;
    if (x_out) std::memcpy(x_out, x.data(), std::min<size_t>(x.size(), (size_t)x_cap) * 8);
    if (x_len) *x_len = (int64_t)x.size();
    return RS_OK;
}

int32_t rs_dc_adjoint(rs_circuit* c, const double* x, int64_t x_len, int64_t probe, double gmin,
                      double* dp_out, double* lambda_out) {
    if (c->nbsim) return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: the adjoint has no BSIM derivatives");
    DevPattern* dp;
    int32_t rc = get_pattern(c, MODE_DC, dp);
    if (rc) return rc;
    const i64 n = dp->p.n, nd = (i64)c->dev.size();
    if (probe < 0 || probe >= n) return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: probe out of range");
    if (x_len < c->node_count) return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: x shorter than node_count");
    const i64 xl = std::min<i64>(x_len, n);
    hipStream_t s = nullptr;
    hipError_t e;
    if ((e = hipMemcpy(c->d_x.p, x, xl * sizeof(double), hipMemcpyHostToDevice)))
        return chip(c, e, "hipMemcpy(x)");
    c->adj_refine = 0;
    if ((rc = adj_factor(c, MODE_DC, c->d_x.p, xl, 0.0, 0.0, gmin)))
        return rc;
    std::vector<double> ev(n, 0.0);
    ev[probe] = 1.0;
    if ((e = c->d_e.ensure(n)) || (e = c->d_lam.ensure(n)) || (e = c->d_dp.ensure(std::max<i64>(nd, 1))) ||
        (e = hipMemcpy(c->d_e.p, ev.data(), n * sizeof(double), hipMemcpyHostToDevice)))
        return chip(c, e, "adjoint rhs");
    if ((rc = adj_solve(c, MODE_DC, c->d_e.p, c->d_lam.p)))
        return rc;
    if (nd && !c->adj[0].aux.p) return cfail(c, RS_ERR_DEVICE, "adjoint: no aux map for the contraction");
    if ((e = launch_adjoint_contract(c->d_dev.p, (int)nd, c->adj[0].aux.p, c->d_x.p, (int)xl, c->d_lam.p,
                                     gmin, 1.0, MODE_DC, nullptr, nullptr, 0.0, 0, c->d_dp.p, s)) ||
        (nd && (e = hipMemcpy(dp_out, c->d_dp.p, nd * sizeof(double), hipMemcpyDeviceToHost))) ||
        (lambda_out && (e = hipMemcpy(lambda_out, c->d_lam.p, n * sizeof(double), hipMemcpyDeviceToHost))))
        return chip(c, e, "adjoint");
    return RS_OK;
}

int32_t rs_circuit_set_device(rs_circuit* c, int64_t k, const rs_device* d) {
    if (k < 0 || k >= (int64_t)c->dev.size() || !d)
        return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: device index out of range");
    const rs_device& o = c->dev[k];
    bool same = d->kind == o.kind && d->flags == o.flags;
    for (int q = 0; q < 4; q++) same = same && d->node[q] == o.node[q];
    if (is_bsim(o) || is_bsim(*d)) same = false;
    if (!same)
        return cfail(c, RS_ERR_INVALID_MATRIX,
                     "Invalid matrix: a device update must keep its kind, nodes, flags and level-1 model");
    c->dev[k] = *d;
    hipError_t e = hipMemcpy(c->d_dev.p + k, d, sizeof(rs_device), hipMemcpyHostToDevice);
    if (e) return chip(c, e, "hipMemcpy(device)");
    for (auto*& b : c->dcb)  // the one-instance DC batches hold copies of the devices
        if (b) {
            rs_batch_destroy(b);
            b = nullptr;
        }
    return RS_OK;
}

int32_t rs_circuit_set_initial_guess(rs_circuit* c, const double* x, int64_t n) {
    if (n < 0 || n > c->node_count || (n > 0 && !x))
        return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: initial guess longer than the node count");
    c->ic.assign(x, x + n);
    return RS_OK;
}

int32_t rs_tran(rs_circuit* c, rs_solver* sol, const rs_newton_opts* o, double tstep, double tstop,
                double tstart, double tmax, int64_t max_points, double* times, double* sols,
                int64_t x_stride, rs_tran_result* res) {
    std::memset(res, 0, sizeof(*res));
    const i64 nd = (i64)c->dev.size();
    std::vector<double> x(c->node_count, 0.0);
    for (size_t q = 0; q < c->ic.size() && (i64)q < c->node_count; q++) x[q] = c->ic[q];
    int32_t rc = reset_cap_state(c);
    if (rc) return rc;
    // BE and TRAP keep separate histories (engine.rs:351-355)
    DevBuf<double> bv, bi;
    DevBuf<uint8_t> bhv, bhi;
    hipError_t e;
    if ((e = bv.ensure(nd)) || (e = bi.ensure(nd)) || (e = bhv.ensure(nd)) || (e = bhi.ensure(nd)))
        return chip(c, e, "hipMalloc(state)");
    if ((e = hipMemset(bhv.p, 0, nd)) || (e = hipMemset(bhi.p, 0, nd)) ||
        (e = hipMemset(bv.p, 0, nd * 8)) || (e = hipMemset(bi.p, 0, nd * 8)))
        return chip(c, e, "hipMemset(state)");
    rs_prepare(sol, c->node_count);
    const double min_dt = std::max(tstep * 1e-6, 1e-15);
    const double max_dt = std::min(tmax, tstop / 10.0);
    const double abs_tol = o->abs_tol, rel_tol = o->rel_tol;
    const double k_p = 0.7 / 2.0, k_i = 0.4 / 2.0;
    bool has_prev = false;
    double prev_err = 0.0;
    int ctl_consec = 0;
    Breakpoints bp;
    for (const auto& d : c->dev)
        if ((d.kind == RS_DEV_V || d.kind == RS_DEV_I) && (d.flags & RS_F_WAVEFORM))
            bp.pulse(d.p, tstop);
    double t = tstart, dt = std::min(tstep, max_dt);
    int consec = 0;
    int32_t iters = 0;
    rs_newton_result dcr{};
    int64_t fc0 = c->factor_calls;
    // the one-launch DC OP starts from zeros: not with an initial guess
    bool dev_dc = can_run_device_dc(c, sol) && o->gmin_steps == 0 && o->source_steps == 0 && c->ic.empty();
    if (dev_dc) {  // the TRAN pattern must differ from DC's, else TRAN would refactor on DC pivots
        DevPattern *pd, *pt;
        if ((rc = get_pattern(c, MODE_DC, pd)) || (rc = get_pattern(c, MODE_TRAN_BE, pt))) return rc;
        dev_dc = !(pd->p.ap == pt->p.ap && pd->p.ai == pt->p.ai);
    }
    rc = dev_dc ? device_dc(c, sol, o, x, &dcr, &iters)
                : run_newton_stepping(c, sol, o, x, MODE_DC, 0.0, 0.0, &dcr, &iters);
    if (rc) return rc;
    res->dc_iters = dcr.iterations;
    if (!dcr.converged) {
        res->status = 2;
        res->factor_calls = (int32_t)(c->factor_calls - fc0);
        return RS_OK;
    }
    int64_t np = 0;
    auto store = [&](double tt) {
        if (np < max_points) {
            if (times) times[np] = tt;
            if (sols)
                for (i64 q = 0; q < x_stride; q++) sols[np * x_stride + q] = q < (i64)x.size() ? x[q] : 0.0;
        }
        np++;
    };
    store(tstart);
    auto store_ptr = [&](double tt, const double* xp) {
        if (np < max_points) {
            if (times) times[np] = tt;
            if (sols)
                for (i64 q = 0; q < x_stride; q++) sols[np * x_stride + q] = q < (i64)x.size() ? xp[q] : 0.0;
        }
        np++;
    };
    // update_transient_state into both histories (engine.rs:448-449)
    if ((rc = cap_update(c, x, 0.0, 0, bv.p, bi.p, bhv.p, bhi.p)) ||
        (rc = cap_update(c, x, 0.0, 0, c->d_cap_v.p, c->d_cap_i.p, c->d_has_v.p, c->d_has_i.p)))
        return rc;
    int status = 0;
    bool graph_tried = false;
    while (t < tstop) {
        // after the first accepted step (the solver then holds the TRAN
        // pattern's factor), the rest of the loop on the device controller
        if (c->use_graph && !graph_tried && res->accepted > 0) {
            graph_tried = true;
            TranLoop L{t, dt, prev_err, has_prev ? 1 : 0, ctl_consec, consec, bp.settle, 0,
                       res->accepted, res->rejected, 0, np};
            bool ran = false;
            rc = tran_adaptive_graph(c, sol, o, x, L, bp.t, tstop, min_dt, max_dt, bv.p, bi.p, bhv.p, bhi.p,
                                     store_ptr, &ran);
            if (rc) return rc;
            if (ran) {
                c->tran_graph_runs++;
                t = L.t;
                dt = L.dt;
                status = L.status;
                iters += L.iters;
                res->accepted = L.accepted;
                res->rejected = L.rejected;
                break;
            }
        }
        dt = bp.limit_dt(t, dt, min_dt);
        if (bp.settle > 0) dt = std::max(dt * bp.factor, min_dt);
        if (t + dt > tstop) dt = tstop - t;
        const double tt = t + dt;
        // BE solve on the BE history
        std::swap(bv.p, c->d_cap_v.p);
        std::swap(bi.p, c->d_cap_i.p);
        std::swap(bhv.p, c->d_has_v.p);
        std::swap(bhi.p, c->d_has_i.p);
        std::vector<double> xb = x;
        rs_newton_result rb{};
        rc = run_newton_stepping(c, sol, o, xb, MODE_TRAN_BE, tt, dt, &rb, &iters);
        std::swap(bv.p, c->d_cap_v.p);
        std::swap(bi.p, c->d_cap_i.p);
        std::swap(bhv.p, c->d_has_v.p);
        std::swap(bhi.p, c->d_has_i.p);
        if (rc) return rc;
        if (!rb.converged) {
            consec++;
            res->rejected++;
            if (consec >= 10) {
                status = 2;
                break;
            }
            dt = std::max(dt * 0.25, min_dt);
            continue;
        }
        std::vector<double> xt = x;
        rs_newton_result rt{};
        rc = run_newton_stepping(c, sol, o, xt, MODE_TRAN_TRAP, tt, dt, &rt, &iters);
        if (rc) return rc;
        if (!rt.converged) {
            consec++;
            res->rejected++;
            if (consec >= 10) {
                status = 2;
                break;
            }
            dt = std::max(dt * 0.25, min_dt);
            continue;
        }
        double lte[5];
        lte_milne(xb, xt, abs_tol, rel_tol, lte);
        const double err = lte[0];
        const bool accept = lte[2] != 0.0 && !(err > 1.0);
        double dt_new;
        if (accept) {  // AdaptiveStepController::process_lte (analysis.rs:720-742)
            double en = err > 1e-10 ? err : 1e-10;
            double pf = std::pow(1.0 / en, k_p);
            double inf = 1.0;
            if (has_prev) inf = std::pow(1.0 / (prev_err > 1e-10 ? prev_err : 1e-10), k_i);
            double fac = clampd(pf * inf * 0.9, 0.1, 2.0);
            has_prev = true;
            prev_err = err;
            ctl_consec = 0;
            dt_new = clampd(dt * fac, min_dt, max_dt);
        } else {
            ctl_consec++;
            if (ctl_consec >= 3)
                dt_new = ctl_consec <= 5 ? dt * 0.25 : (ctl_consec <= 10 ? dt * 0.1 : min_dt);
            else
                dt_new = clampd(dt * lte[3], min_dt, max_dt);
        }
        if (accept) {
            const double tn = t + dt;
            bp.update(t, tn);
            x = xt;
            if ((rc = cap_update(c, x, dt, 1, c->d_cap_v.p, c->d_cap_i.p, c->d_has_v.p, c->d_has_i.p)) ||
                (rc = cap_update(c, x, dt, 0, bv.p, bi.p, bhv.p, bhi.p)))
                return rc;
            store(tn);
            t = tn;
            res->accepted++;
            consec = 0;
            dt = clampd(dt_new, min_dt, max_dt);
        } else {
            res->rejected++;
            consec++;
            if (consec >= 10) {
                status = 2;
                break;
            }
            dt = clampd(dt_new, min_dt, max_dt);
        }
    }
    res->status = status;
    res->npoints = np;
    res->newton_iters_total = iters;
    res->factor_calls = (int32_t)(c->factor_calls - fc0);
    bv.release();
    bi.release();
    bhv.release();
    bhi.release();
    return RS_OK;
}

int64_t rs_dc_sweep_points(double start, double stop, double step, double* values, int64_t cap) {
    std::vector<double> v;
    const double st = stop >= start ? std::fabs(step) : -std::fabs(step);
    if (std::fabs(st) < 1e-15) {
        v.push_back(start);
    } else {
        const double npt = std::floor(std::fabs((stop - start) / st)) + 1.0;
        if (!(npt >= 1.0 && npt <= 1e8)) return -1;  // non-finite or absurd ranges
        const int64_t np = (int64_t)npt;
        v.reserve((size_t)np + 1);
        for (int64_t i = 0; i < np; i++) v.push_back(start + (double)i * st);
        const double last = v.back();
        if (std::fabs(last - stop) > 1e-12 && v.size() < 10000 &&
            std::fabs(last - stop) / std::fabs(st) > 0.5)
            v.push_back(stop);
    }
    if (values)
        for (int64_t i = 0; i < (int64_t)v.size() && i < cap; i++) values[i] = v[i];
    return (int64_t)v.size();
}

int32_t rs_dc_sweep(rs_circuit* c, rs_solver* sol, const rs_newton_opts* o, int64_t source,
                    double start, double stop, double step, int64_t max_points, double* values,
                    double* sols, int64_t x_stride, rs_sweep_result* res) {
    std::memset(res, 0, sizeof(*res));
    if (source < 0 || source >= (int64_t)c->dev.size()) {
        res->status = 3;
        return cfail(c, RS_OK, "DC sweep source not found");
    }
    const int64_t np = rs_dc_sweep_points(start, stop, step, nullptr, 0);
    if (np < 0) return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: DC sweep range is not finite");
    std::vector<double> pts((size_t)np);
    rs_dc_sweep_points(start, stop, step, pts.data(), np);
    res->npoints = np;
    if (values)
        for (int64_t i = 0; i < np && i < max_points; i++) values[i] = pts[i];
    rs_device& d = c->dev[source];
    if (d.flags & RS_F_WAVEFORM) {  // the DC pattern changes (a PULSE source is not stamped in DC)
        d.flags &= ~RS_F_WAVEFORM;
        c->pat[0].built = false;
        c->pat[1].built = false;
        for (auto*& b : c->dcb)
            if (b) {
                rs_batch_destroy(b);
                b = nullptr;
            }
    }
    std::vector<double> x(c->node_count, 0.0);
    rs_prepare(sol, c->node_count);
    const int64_t fc0 = c->factor_calls;
    int32_t iters = 0, rc = RS_OK;
    for (int64_t k = 0; k < np; k++) {
        d.value = pts[k];
        hipError_t e = hipMemcpy(c->d_dev.p + source, &d, sizeof(rs_device), hipMemcpyHostToDevice);
        if (e) return chip(c, e, "hipMemcpy(device)");
        rs_newton_result r{};
        rc = run_newton_stepping(c, sol, o, x, MODE_DC, 0.0, 0.0, &r, nullptr);
        if (rc) return rc;
        iters += r.iterations;  // the returned pass only (engine.rs:771)
        if (r.converged) {
            if (sols && res->nsolved < max_points)
                for (int64_t q = 0; q < x_stride; q++)
                    sols[res->nsolved * x_stride + q] = q < (int64_t)x.size() ? x[q] : 0.0;
            res->nsolved++;
        } else {
            res->status = r.reason == 2 ? 2 : 1;
            res->fail_value = pts[k];
            break;
        }
    }
    res->iterations = iters;
    res->factor_calls = (int32_t)(c->factor_calls - fc0);
    return RS_OK;
}

int32_t rs_tran_fixed(rs_circuit* c, rs_solver* sol, const rs_newton_opts* o, double dt,
                      int64_t nsteps, double* x_final, int64_t x_cap, int64_t* x_len,
                      int32_t* newton_iters, int32_t* status_out) {
    std::vector<double> x(c->node_count, 0.0);
    int32_t rc = reset_cap_state(c);
    if (rc) return rc;
    rs_prepare(sol, c->node_count);
    int32_t it = 0;
    rs_newton_result r{};
    rc = run_newton_stepping(c, sol, o, x, MODE_DC, 0.0, 0.0, &r, &it);
    if (rc) return rc;
    int32_t status = r.converged ? 0 : 1;
    if (!status) {
        rc = cap_update(c, x, 0.0, 0, c->d_cap_v.p, c->d_cap_i.p, c->d_has_v.p, c->d_has_i.p);
        if (rc) return rc;
        double t = 0.0;
        for (int64_t k = 0; k < nsteps; k++) {
            double tt = t + dt;
            if (k >= 1 && c->use_graph) {  // the remaining steps as captured graphs
                int64_t done = 0;
                rc = tran_fixed_graph(c, sol, o, x, t, dt, nsteps - k, &done, &it, &status);
                if (rc) return rc;
                if (done >= 0) break;  // the graph path ran (or stopped at a failed step)
            }
            rc = run_newton_stepping(c, sol, o, x, MODE_TRAN_TRAP, tt, dt, &r, &it);
            if (rc) return rc;
            if (!r.converged) {
                status = 2;
                break;
            }
            rc = cap_update(c, x, dt, 1, c->d_cap_v.p, c->d_cap_i.p, c->d_has_v.p, c->d_has_i.p);
            if (rc) return rc;
            t = tt;
        }
    }
    if (x_final) std::memcpy(x_final, x.data(), std::min<size_t>(x.size(), (size_t)x_cap) * 8);
    if (x_len) *x_len = (int64_t)x.size();
    if (newton_iters) *newton_iters = it;
    if (status_out) *status_out = status;
    return RS_OK;
}

int32_t rs_tran_steps(rs_circuit* c, rs_solver* sol, const rs_newton_opts* o, double t0, double dt,
                      int64_t nsteps, double* x_io, int64_t x_cap, int64_t* x_len, int32_t* newton_iters,
                      int32_t* status_out) {
    if (!x_len || *x_len < c->node_count || *x_len > x_cap)
        return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: x must hold node_count..x_cap entries");
    std::vector<double> x(x_io, x_io + *x_len);
    int32_t it = 0, status = 0;
    rs_newton_result r{};
    double t = t0;
    for (int64_t k = 0; k < nsteps; k++) {
        const double tt = t + dt;
        int32_t rc = run_newton_stepping(c, sol, o, x, MODE_TRAN_TRAP, tt, dt, &r, &it);
        if (rc) return rc;
        if (!r.converged) {
            status = 2;
            break;
        }
        rc = cap_update(c, x, dt, 1, c->d_cap_v.p, c->d_cap_i.p, c->d_has_v.p, c->d_has_i.p);
        if (rc) return rc;
        t = tt;
    }
    std::memcpy(x_io, x.data(), std::min<size_t>(x.size(), (size_t)x_cap) * 8);
    *x_len = (int64_t)std::min<size_t>(x.size(), (size_t)x_cap);
    if (newton_iters) *newton_iters = it;
    if (status_out) *status_out = status;
    return RS_OK;
}

void rs_circuit_set_tran_graph(rs_circuit* c, int32_t on) { c->use_graph = on != 0; }

int64_t rs_circuit_graph_steps(const rs_circuit* c) { return c->graph_steps; }

int32_t rs_circuit_tran_graph_info(const rs_circuit* c, int64_t* out) {
    out[0] = c->tran_graph_runs;
    out[1] = c->graph_launches;
    out[2] = c->tran_host_pows;
    return RS_OK;
}

int32_t rs_tran_adjoint(rs_circuit* c, rs_solver* sol, const rs_newton_opts* o, double dt,
                        int64_t nsteps, int64_t probe, int64_t every, double* dp_out,
                        double* x_final, int64_t x_cap, int64_t* x_len, rs_tadj_info* info) {
    std::memset(info, 0, sizeof(*info));
    if (c->nbsim) return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: the adjoint has no BSIM derivatives");
    if (probe < 0 || probe >= c->node_count)
        return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: probe must be a node");
    if (nsteps < 0 || !(dt > 0.0)) return cfail(c, RS_ERR_INVALID_MATRIX, "Invalid matrix: bad dt or step count");
    const bool keep_all = every < 1;  // every state kept: no recomputation
    if (keep_all) every = nsteps + 1;
    // the adjoint's own solvers follow the caller's supernodal setting
    c->adj_sn_mode = sol->sn_mode;
    for (auto& a : c->adj)
        if (a.s && a.s->sn_mode != sol->sn_mode) {
            rs_set_supernodal(a.s, sol->sn_mode);
            rs_reset_pattern(a.s);
        }
    const i64 nd = (i64)c->dev.size(), nc = c->node_count;
    hipError_t e;
    int32_t rc;
    // a checkpoint: x_k, x_{k-1} and the capacitor history after step k
    struct Ckpt {
        int64_t k;
        std::vector<double> x, xm1, cv, ci;
        std::vector<uint8_t> hv, hi;
    };
    std::vector<Ckpt> ck;
    auto save = [&](int64_t k, const std::vector<double>& x, const std::vector<double>& xm1) -> int32_t {
        Ckpt p{k, x, xm1, std::vector<double>(nd), std::vector<double>(nd), std::vector<uint8_t>(nd),
               std::vector<uint8_t>(nd)};
        if (nd && ((e = hipMemcpy(p.cv.data(), c->d_cap_v.p, nd * 8, hipMemcpyDeviceToHost)) ||
                   (e = hipMemcpy(p.ci.data(), c->d_cap_i.p, nd * 8, hipMemcpyDeviceToHost)) ||
                   (e = hipMemcpy(p.hv.data(), c->d_has_v.p, nd, hipMemcpyDeviceToHost)) ||
                   (e = hipMemcpy(p.hi.data(), c->d_has_i.p, nd, hipMemcpyDeviceToHost))))
            return chip(c, e, "checkpoint");
        ck.push_back(std::move(p));
        info->checkpoints++;
        return RS_OK;
    };
    auto restore = [&](const Ckpt& p) -> int32_t {
        if (nd && ((e = hipMemcpy(c->d_cap_v.p, p.cv.data(), nd * 8, hipMemcpyHostToDevice)) ||
                   (e = hipMemcpy(c->d_cap_i.p, p.ci.data(), nd * 8, hipMemcpyHostToDevice)) ||
                   (e = hipMemcpy(c->d_has_v.p, p.hv.data(), nd, hipMemcpyHostToDevice)) ||
                   (e = hipMemcpy(c->d_has_i.p, p.hi.data(), nd, hipMemcpyHostToDevice))))
            return chip(c, e, "restore checkpoint");
        return RS_OK;
    };
    using clk = std::chrono::steady_clock;
    auto ms_since = [](clk::time_point a) {
        return std::chrono::duration<double, std::milli>(clk::now() - a).count();
    };
    const clk::time_point t_fwd = clk::now();
    // ---- forward: rs_tran_fixed, keeping a checkpoint every `every` steps
    std::vector<double> tk(nsteps + 1, 0.0), gm(nsteps + 1, o->gmin), sc(nsteps + 1, 1.0);
    std::vector<double> x(nc, 0.0), xm1;
    if ((rc = reset_cap_state(c))) return rc;
    rs_prepare(sol, nc);
    rs_newton_result r{};
    int32_t it = 0;
    c->polish_iters = 0;
    if ((rc = run_newton_stepping(c, sol, o, x, MODE_DC, 0.0, 0.0, &r, &it, &gm[0], &sc[0]))) return rc;
    if (!r.converged) {
        info->status = 1;
        info->forward_iters = it;
        return RS_OK;
    }
    bool ok = false;
    if ((rc = exact_newton(c, x, MODE_DC, 0.0, 0.0, gm[0], sc[0], &ok))) return rc;
    if (!ok) {
        info->status = 1;
        info->forward_iters = it;
        return RS_OK;
    }
    if ((rc = cap_update(c, x, 0.0, 0, c->d_cap_v.p, c->d_cap_i.p, c->d_has_v.p, c->d_has_i.p))) return rc;
    if ((rc = save(0, x, {}))) return rc;
    std::vector<std::vector<double>> xall;  // keep_all: x_0 .. x_nsteps
    if (keep_all) {
        xall.reserve(nsteps + 1);
        xall.push_back(x);
    }
    // one TRAP step to its exact root: exact-Jacobian Newton from x_{k-1};
    // should that not converge, the reference loop (with its stepping) first,
    // on the adjoint's own factor (SparseLU with the unknowns that have no
    // diagonal entry last: no pivot clamped to 1e-14, which on large meshes
    // keeps the caller's SparseLU from converging at all). The recomputation
    // runs the same function, so it reproduces x_k exactly.
    int64_t inexact = 0;
    auto tran_step = [&](std::vector<double>& xv, int64_t k, double tt, int32_t* iters,
                         bool* conv) -> int32_t {
        std::vector<double> x0 = xv;
        bool okk = false;
        int32_t rcc = exact_newton(c, xv, MODE_TRAN_TRAP, tt, dt, gm[k], sc[k], &okk);
        if (rcc) return rcc;
        if (!okk) {
            xv = x0;
            rs_newton_result rr{};
            rs_solver* fb = c->adj[1].s ? c->adj[1].s : sol;
            if ((rcc = run_newton_stepping(c, fb, o, xv, MODE_TRAN_TRAP, tt, dt, &rr, iters, &gm[k], &sc[k])))
                return rcc;
            if (!rr.converged) {
                *conv = false;
                return RS_OK;
            }
            const std::vector<double> xref = xv;
            if ((rcc = exact_newton(c, xv, MODE_TRAN_TRAP, tt, dt, gm[k], sc[k], &okk))) return rcc;
            if (!okk) {  // no exact root within reach (the level-1 current is not
                xv = xref;  // smooth at vds = 0): the reference loop's point
                inexact++;
            }
        }
        *conv = true;
        return RS_OK;
    };
    double t = 0.0;
    for (int64_t k = 1; k <= nsteps; k++) {
        const double tt = t + dt;
        xm1 = x;
        if ((rc = tran_step(x, k, tt, &it, &ok))) return rc;
        if (!ok) {
            info->status = 2;
            info->forward_iters = it;
            info->polish_iters = (int32_t)c->polish_iters;
            info->inexact_steps = inexact;
            return RS_OK;
        }
        if ((rc = cap_update(c, x, dt, 1, c->d_cap_v.p, c->d_cap_i.p, c->d_has_v.p, c->d_has_i.p)))
            return rc;
        t = tt;
        tk[k] = t;
        if (k % every == 0 && k < nsteps && (rc = save(k, x, xm1))) return rc;
        if (keep_all) xall.push_back(x);
    }
    if (keep_all) info->checkpoints = (int32_t)xall.size();
    info->forward_iters = it;
    info->forward_ms = ms_since(t_fwd);
    if (x_final) std::memcpy(x_final, x.data(), std::min<size_t>(x.size(), (size_t)x_cap) * 8);
    if (x_len) *x_len = (int64_t)x.size();
    // ---- backward, segment by segment from the last checkpoint
    // capacitors of each node, sign in bit 31 for the second terminal
    std::vector<int32_t> cptr(nc + 1, 0), cdev;
    for (i64 d = 0; d < nd; d++)
        if (c->dev[d].kind == RS_DEV_C)
            for (int tn = 0; tn < 2; tn++)
                if (c->dev[d].node[tn] < nc) cptr[c->dev[d].node[tn] + 1]++;
    for (i64 i = 0; i < nc; i++) cptr[i + 1] += cptr[i];
    cdev.resize(std::max<int32_t>(cptr[nc], 1));
    {
        std::vector<int32_t> f(cptr.begin(), cptr.end() - 1);
        for (i64 d = 0; d < nd; d++)
            if (c->dev[d].kind == RS_DEV_C)
                for (int tn = 0; tn < 2; tn++) {
                    const int64_t nn = c->dev[d].node[tn];
                    if (nn < nc) cdev[f[nn]++] = (int32_t)d | (tn == 1 ? (int32_t)0x80000000u : 0);
                }
    }
    DevBuf<int32_t> d_cptr, d_cdev;
    DevBuf<double> lamb[3], xk, xk1, xk2, dpa;
    i64 nmax = (i64)x.size() + 8;
    DevPattern* pdc;
    if ((rc = get_pattern(c, MODE_DC, pdc))) return rc;
    nmax = std::max<i64>(nmax, pdc->p.n);
    if ((e = d_cptr.ensure(nc + 1)) || (e = d_cdev.ensure(cdev.size())) ||
        (e = hipMemcpy(d_cptr.p, cptr.data(), (nc + 1) * 4, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(d_cdev.p, cdev.data(), cdev.size() * 4, hipMemcpyHostToDevice)) ||
        (e = lamb[0].ensure(nmax)) || (e = lamb[1].ensure(nmax)) || (e = lamb[2].ensure(nmax)) ||
        (e = xk.ensure(nmax)) || (e = xk1.ensure(nmax)) || (e = xk2.ensure(nmax)) ||
        (e = dpa.ensure(std::max<i64>(nd, 1))) || (e = c->d_e.ensure(nmax)) ||
        (e = hipMemset(dpa.p, 0, std::max<i64>(nd, 1) * 8)))
        return chip(c, e, "adjoint buffers");
    c->adj_refine = 0;
    const double* l1 = nullptr;  // lambda_{k+1}
    const double* l2 = nullptr;  // lambda_{k+2}
    int cur = 0;
    auto up = [&](DevBuf<double>& b, const std::vector<double>& v) -> hipError_t {
        return v.empty() ? hipSuccess : hipMemcpy(b.p, v.data(), v.size() * 8, hipMemcpyHostToDevice);
    };
    for (int64_t ci = (int64_t)ck.size() - 1; ci >= 0; ci--) {
        const Ckpt& p = ck[ci];
        const int64_t k0 = p.k, k1 = ci + 1 < (int64_t)ck.size() ? ck[ci + 1].k : nsteps;
        std::vector<std::vector<double>> xs(k1 - k0 + 1);
        xs[0] = p.x;
        if (keep_all) xs = std::move(xall);  // one segment, every state stored
        if ((rc = restore(p))) return rc;
        std::vector<double> xr = p.x;
        int32_t rit = 0;
        const clk::time_point t_rec = clk::now();
        for (int64_t k = k0 + 1; k <= k1 && !keep_all; k++) {  // recompute the segment
            bool okr = false;
            if ((rc = tran_step(xr, k, tk[k], &rit, &okr))) return rc;
            if (!okr) return cfail(c, RS_ERR_SOLVE_FAILED, "recomputed step did not converge");
            if ((rc = cap_update(c, xr, dt, 1, c->d_cap_v.p, c->d_cap_i.p, c->d_has_v.p, c->d_has_i.p)))
                return rc;
            xs[k - k0] = xr;
            info->recomputed_steps++;
        }
        info->recompute_iters += rit;
        info->recompute_ms += ms_since(t_rec);
        const clk::time_point t_bwd = clk::now();
        for (int64_t k = k1; k >= (k0 == 0 ? 0 : k0 + 1); k--) {
            const int mode = k == 0 ? MODE_DC : MODE_TRAN_TRAP;
            const std::vector<double>& xv = xs[k - k0];
            const std::vector<double>* xv1 = k >= 1 ? (k - 1 >= k0 ? &xs[k - 1 - k0] : &p.xm1) : nullptr;
            const std::vector<double>* xv2 =
                k >= 2 ? (k - 2 >= k0 ? &xs[k - 2 - k0] : (k - 2 == k0 - 1 ? &p.xm1 : nullptr)) : nullptr;
            if (k >= 2 && !xv2) return cfail(c, RS_ERR_SOLVE_FAILED, "adjoint checkpoint lacks x_{k-2}");
            if ((e = up(xk, xv)) || (xv1 && (e = up(xk1, *xv1))) || (xv2 && (e = up(xk2, *xv2))))
                return chip(c, e, "adjoint x");
            if ((rc = adj_factor(c, mode, xk.p, (i64)xv.size(), tk[k], dt, gm[k]))) return rc;
            rs_circuit::AdjCtx& A = c->adj[mode == MODE_DC ? 0 : 1];
            const i64 n = A.s->n;
            if ((e = launch_tadj_rhs(c->d_dev.p, d_cptr.p, d_cdev.p, (int)nc, l1, l2, dt, k >= 1 ? 1 : 0,
                                     k == nsteps ? (int)probe : -1, (int)n, c->d_e.p, nullptr)))
                return chip(c, e, "adjoint rhs");
            double* lam = lamb[cur].p;
            // the adjoint factor has no clamped pivot: one refinement step
            if ((rc = adj_solve(c, mode, c->d_e.p, lam, 1))) return rc;
            if (nd && !A.aux.p) return cfail(c, RS_ERR_DEVICE, "adjoint: no aux map for the contraction");
            if ((e = launch_adjoint_contract(c->d_dev.p, (int)nd, A.aux.p, xk.p, (int)xv.size(), lam, gm[k],
                                             sc[k], mode, xv1 ? xk1.p : nullptr, xv2 ? xk2.p : nullptr, dt, 1,
                                             dpa.p, nullptr)))
                return chip(c, e, "adjoint contraction");
            info->adjoint_solves++;
            l2 = l1;
            l1 = lam;
            cur = (cur + 1) % 3;
        }
        if ((e = hipDeviceSynchronize())) return chip(c, e, "adjoint segment");
        info->backward_ms += ms_since(t_bwd);
    }
    if (nd && (e = hipMemcpy(dp_out, dpa.p, nd * 8, hipMemcpyDeviceToHost))) return chip(c, e, "adjoint dp");
    info->max_refine = c->adj_refine;
    info->polish_iters = (int32_t)c->polish_iters;
    info->inexact_steps = inexact;
    return RS_OK;
}

}  // extern "C"
