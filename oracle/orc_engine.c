/*
 * orc_engine.c - Newton loop and DC/TRAN drivers. TEST INFRASTRUCTURE (oracle).
 *
 * crates/sim-core/src/newton.rs:65-201 (run_newton, run_newton_with_stepping,
 * check_convergence, apply_damping), engine.rs:249-311 (DC OP) and :333-632
 * (TRAN), analysis.rs:147-200 (Milne LTE) and :381-742 (PI controller),
 * waveform.rs:459-645 (PULSE breakpoints, limit_dt, settling).
 */
#include "orc.h"
#include "orc_util.h"

typedef struct {
    const rs_device *dev;
    int64_t ndev, node_count;
    int mode;
    double t, dt;
    const orc_tstate *st;
} build_ctx;

/* LinearSolver dispatch (solver.rs:141-166) over the two oracle solvers */
static int sv_bbd(const void *s) { return *(const uint32_t *)s == ORC_MAGIC_BBD; }
static void sv_prepare(void *s, int64_t n) {
    if (sv_bbd(s)) orc_bbd_prepare((orc_bbd *)s, n); else orc_klu_prepare((orc_klu *)s, n);
}
static int sv_analyze(void *s, int64_t apl, const int64_t *ap, int64_t ail, const int64_t *ai) {
    return sv_bbd(s) ? orc_bbd_analyze((orc_bbd *)s, apl, ap, ail, ai)
                     : orc_klu_analyze((orc_klu *)s, apl, ap, ail, ai);
}
static int sv_factor(void *s, int64_t apl, const int64_t *ap, int64_t ail, const int64_t *ai,
                     const double *ax) {
    return sv_bbd(s) ? orc_bbd_factor((orc_bbd *)s, apl, ap, ail, ai, ax)
                     : orc_klu_factor((orc_klu *)s, apl, ap, ail, ai, ax);
}
static int sv_solve(void *s, int64_t n, double *rhs) {
    return sv_bbd(s) ? orc_bbd_solve((orc_bbd *)s, n, rhs) : orc_klu_solve((orc_klu *)s, n, rhs);
}

static double norm2(const double *v, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += v[i] * v[i];
    return sqrt(s);
}

/* run_newton (newton.rs:65-117). *x is a malloc'd vector of *xl entries; on
 * convergence it is replaced by the full solution (newton.rs:98). */
static orc_newton_result run_newton(const rs_newton_opts *o, double **x, int64_t *xl,
                                    const build_ctx *bc, double gmin, double source_scale,
                                    void *solver, int32_t *factor_calls) {
    orc_newton_result r = {0, 0, 1, 0.0};
    double damping = o->damping;
    double prev_dx_norm = 1.7976931348623157e308;
    for (int it = 0; it < o->max_iters; it++) {
        r.iterations = it + 1;
        orc_mna m;
        orc_build_mna(bc->dev, bc->ndev, bc->node_count, *x, *xl, bc->mode, bc->t, bc->dt,
                      bc->st, gmin, source_scale, &m);
        sv_prepare(solver, m.n);
        int err = sv_analyze(solver, m.n + 1, m.ap, m.nnz, m.ai);
        if (!err) { err = sv_factor(solver, m.n + 1, m.ap, m.nnz, m.ai, m.ax); if (factor_calls) (*factor_calls)++; }
        if (!err) err = sv_solve(solver, m.n, m.rhs);
        if (err) { r.reason = 2; orc_mna_free(&m); break; }
        double *xn = m.rhs;
        int64_t nn = m.n;
        int64_t nd = nn < *xl ? nn : *xl;
        double *dx = (double *)orc_xmalloc((size_t)(nd ? nd : 1) * sizeof(double));
        for (int64_t i = 0; i < nd; i++) dx[i] = xn[i] - (*x)[i];
        double dxn = norm2(dx, nd);
        r.final_norm = norm2(xn, nn);
        int conv = 1;
        for (int64_t i = 0; i < nd; i++) {
            if (!(fabs(dx[i]) <= o->abs_tol + o->rel_tol * fabs(xn[i]))) { conv = 0; break; }
        }
        free(dx);
        if (conv) {
            free(*x);
            *x = (double *)orc_xmalloc((size_t)(nn ? nn : 1) * sizeof(double));
            memcpy(*x, xn, (size_t)nn * sizeof(double));
            *xl = nn;
            r.converged = 1; r.reason = 0;
            orc_mna_free(&m);
            break;
        }
        if (dxn > prev_dx_norm && damping > o->damping_min) {
            double d2 = damping * 0.5;
            damping = d2 >= o->damping_min ? d2 : o->damping_min;
        }
        prev_dx_norm = dxn;
        for (int64_t i = 0; i < nd; i++) (*x)[i] = (*x)[i] + damping * (xn[i] - (*x)[i]);
        orc_mna_free(&m);
    }
    return r;
}

/* run_newton_with_stepping (newton.rs:142-185) */
static orc_newton_result run_newton_stepping(const rs_newton_opts *o, double **x, int64_t *xl,
                                             const build_ctx *bc, void *solver,
                                             int32_t *factor_calls, int32_t *iters_total) {
    double gstart = o->gmin * 1e3; if (gstart < 1e-6) gstart = 1e-6;
    orc_newton_result last = {0, 0, 1, 0.0};
    for (int gi = 0; gi <= o->gmin_steps; gi++) {
        double gmin = o->gmin_steps == 0 ? o->gmin
                      : gstart + (o->gmin - gstart) * ((double)gi / (double)o->gmin_steps);
        for (int si = 0; si <= o->source_steps; si++) {
            double scale = o->source_steps == 0 ? 1.0 : (double)si / (double)o->source_steps;
            orc_newton_result r = run_newton(o, x, xl, bc, gmin, scale, solver, factor_calls);
            if (iters_total) *iters_total += r.iterations;
            if (r.converged) return r;
            last = r;
        }
    }
    return last;
}

int orc_dc_op(const rs_device *dev, int64_t ndev, int64_t node_count, const rs_newton_opts *o,
              void *solver, double *x_out, int64_t x_cap, int64_t *x_len,
              orc_newton_result *res) {
    int64_t xl = node_count;
    double *x = (double *)orc_xcalloc((size_t)(xl ? xl : 1), sizeof(double));
    sv_prepare(solver, node_count);
    build_ctx bc = {dev, ndev, node_count, ORC_MODE_DC, 0.0, 0.0, NULL};
    *res = run_newton_stepping(o, &x, &xl, &bc, solver, NULL, NULL);
    int64_t cpy = xl < x_cap ? xl : x_cap;
    if (x_out) memcpy(x_out, x, (size_t)cpy * sizeof(double));
    *x_len = cpy;
    free(x);
    return res->converged ? 0 : 1;
}

void orc_lte_milne(const double *xb, const double *xt, int64_t n, double abs_tol,
                   double rel_tol, double out[5]) {
    double max_ratio = 0.0, raw = 0.0;
    int64_t max_node = 0;
    for (int64_t i = 0; i < n; i++) {
        double be = xb[i], tr = xt[i];
        if (fabs(be) < 1e-30 && fabs(tr) < 1e-30) continue;
        double lte = fabs(tr - be) / 3.0;
        double mb = fabs(be), mt = fabs(tr);
        double tol = abs_tol + rel_tol * (mb >= mt ? mb : mt);
        double ratio = tol > 0.0 ? lte / tol : 0.0;
        if (ratio > max_ratio) { max_ratio = ratio; max_node = i; raw = lte; }
    }
    double f;
    if (max_ratio > 1e-10) {
        f = pow(1.0 / max_ratio, 0.5) * 0.9;
        if (f < 0.1) f = 0.1;
        if (f > 2.0) f = 2.0;
    } else {
        f = 2.0;
    }
    out[0] = max_ratio; out[1] = (double)max_node; out[2] = max_ratio <= 1.0;
    out[3] = f; out[4] = raw;
}

static double clampd(double x, double lo, double hi) {
    if (x < lo) return lo;
    if (x > hi) return hi;
    return x;
}

/* BreakpointManager (waveform.rs:380-645) on a sorted unique time list */
typedef struct { double *t; int64_t n, cap; int settle, settle_steps; double factor; } bpm;

static void bp_add(bpm *b, double time) {
    if (time < 0.0) return;
    for (int64_t i = 0; i < b->n; i++) if (b->t[i] == time) return;
    if (b->n == b->cap) { b->cap = b->cap ? b->cap * 2 : 16; b->t = (double *)realloc(b->t, (size_t)b->cap * sizeof(double)); }
    int64_t i = b->n++;
    while (i > 0 && b->t[i - 1] > time) { b->t[i] = b->t[i - 1]; i--; }
    b->t[i] = time;
}
static void bp_pulse(bpm *b, const double *p, double tstop) { /* waveform.rs:472-516 */
    double td = p[RS_PP_TD], tr = p[RS_PP_TR], tf = p[RS_PP_TF], pw = p[RS_PP_PW], per = p[RS_PP_PER];
    double t = td;
    int64_t cnt = 0;
    if (td > 0.0) bp_add(b, td);
    while (t <= tstop && cnt < 10000) {
        bp_add(b, t);
        double th = t + tr; if (th > tstop) break; bp_add(b, th);
        double tfa = th + pw; if (tfa > tstop) break; bp_add(b, tfa);
        double tl = tfa + tf; if (tl > tstop) break; bp_add(b, tl);
        if (per > 0.0) { t = td + per * (double)(cnt + 1); cnt++; }
        else break;
    }
}
static int bp_next(const bpm *b, double t, double *out) {
    for (int64_t i = 0; i < b->n; i++) if (b->t[i] > t) { *out = b->t[i]; return 1; }
    return 0;
}
static double bp_limit_dt(const bpm *b, double t, double pdt, double min_margin) {
    double nb;
    if (bp_next(b, t, &nb)) {
        double ttb = nb - t;
        if (pdt >= ttb) return ttb >= min_margin ? ttb : min_margin;
        if (pdt > 0.9 * ttb) return ttb;
        double rem = ttb - pdt;
        if (rem < 0.1 * pdt && rem < min_margin * 10.0) return ttb;
    }
    return pdt;
}
static int bp_crossed(const bpm *b, double t0, double t1) {
    for (int64_t i = 0; i < b->n; i++) if (b->t[i] > t0 && b->t[i] <= t1) return 1;
    return 0;
}
static void bp_update_settling(bpm *b, double t0, double t1) {
    if (bp_crossed(b, t0, t1)) b->settle = b->settle_steps;
    if (b->settle > 0) b->settle--;
}

static void tstate_init(orc_tstate *s, int64_t ndev, int method) {
    s->cap_v = (double *)orc_xcalloc((size_t)(ndev ? ndev : 1), sizeof(double));
    s->cap_i = (double *)orc_xcalloc((size_t)(ndev ? ndev : 1), sizeof(double));
    s->has_v = (uint8_t *)orc_xcalloc((size_t)(ndev ? ndev : 1), 1);
    s->has_i = (uint8_t *)orc_xcalloc((size_t)(ndev ? ndev : 1), 1);
    s->method = method;
}
static void tstate_free(orc_tstate *s) { free(s->cap_v); free(s->cap_i); free(s->has_v); free(s->has_i); }

/* update_transient_state (stamp.rs:572-602) */
static void tstate_update(const rs_device *dev, int64_t ndev, const double *x, int64_t xl,
                          orc_tstate *s) {
    for (int64_t i = 0; i < ndev; i++) {
        if (dev[i].kind != RS_DEV_C) continue;
        int64_t a = dev[i].node[0], b = dev[i].node[1];
        double va = a < xl ? x[a] : 0.0, vb = b < xl ? x[b] : 0.0;
        s->cap_v[i] = va - vb; s->has_v[i] = 1;
    }
}
/* update_transient_state_full (stamp.rs:615-668) */
static void tstate_update_full(const rs_device *dev, int64_t ndev, const double *x, int64_t xl,
                               double dt, orc_tstate *s) {
    for (int64_t i = 0; i < ndev; i++) {
        if (dev[i].kind != RS_DEV_C) continue;
        int64_t a = dev[i].node[0], b = dev[i].node[1];
        double va = a < xl ? x[a] : 0.0, vb = b < xl ? x[b] : 0.0;
        double v = va - vb;
        double c = dev[i].value;
        double vp = s->has_v[i] ? s->cap_v[i] : v;
        double cur = c * (v - vp) / dt;
        s->cap_v[i] = v; s->has_v[i] = 1;
        s->cap_i[i] = cur; s->has_i[i] = 1;
    }
}

/* .ic values applied to x before the initial DC OP (engine.rs:343-348) */
static double *g_ic = NULL;
static int64_t g_nic = 0;

void orc_set_initial_guess(const double *x, int64_t n) {
    free(g_ic);
    g_ic = NULL;
    g_nic = 0;
    if (n > 0) {
        g_ic = (double *)orc_xcalloc((size_t)n, sizeof(double));
        memcpy(g_ic, x, (size_t)n * sizeof(double));
        g_nic = n;
    }
}

int orc_tran(const rs_device *dev, int64_t ndev, int64_t node_count, const rs_newton_opts *o,
             double tstep, double tstop, double tstart, double tmax, void *solver,
             int64_t max_points, double *times, double *sols, int64_t x_stride,
             orc_tran_result *res) {
    memset(res, 0, sizeof(*res));
    int64_t xl = node_count;
    double *x = (double *)orc_xcalloc((size_t)(xl ? xl : 1), sizeof(double));
    for (int64_t q = 0; q < g_nic && q < node_count; q++) x[q] = g_ic[q];
    orc_tstate sbe, str;
    tstate_init(&sbe, ndev, ORC_BE);
    tstate_init(&str, ndev, ORC_TRAP);
    sv_prepare(solver, node_count);
    double min_dt = tstep * 1e-6; if (min_dt < 1e-15) min_dt = 1e-15;
    double max_dt = tmax < tstop / 10.0 ? tmax : tstop / 10.0;
    double abs_tol = o->abs_tol, rel_tol = o->rel_tol;
    /* PI controller state (analysis.rs:381-634, order 2) */
    double k_p = 0.7 / 2.0, k_i = 0.4 / 2.0;
    int has_prev_err = 0; double prev_err = 0.0;
    int ctl_consec = 0;
    bpm bp = {0}; bp.settle_steps = 5; bp.factor = 0.1;
    for (int64_t i = 0; i < ndev; i++)
        if ((dev[i].kind == RS_DEV_V || dev[i].kind == RS_DEV_I) && (dev[i].flags & RS_F_WAVEFORM))
            bp_pulse(&bp, dev[i].p, tstop);
    double t = tstart, dt = tstep < max_dt ? tstep : max_dt;
    int consec = 0;
    build_ctx bc = {dev, ndev, node_count, ORC_MODE_DC, 0.0, 0.0, NULL};
    int32_t iters_total = 0, fcalls = 0;
    orc_newton_result dcr = run_newton_stepping(o, &x, &xl, &bc, solver, &fcalls, &iters_total);
    res->dc_iters = dcr.iterations;
    if (!dcr.converged) {
        res->status = 2; res->npoints = 0;
        res->factor_calls = fcalls; res->newton_iters_total = iters_total;
        free(x); tstate_free(&sbe); tstate_free(&str); free(bp.t);
        return 1;
    }
    int64_t np = 0;
#define STORE_POINT(tt)                                                                      \
    do {                                                                                     \
        if (np < max_points) {                                                               \
            if (times) times[np] = (tt);                                                     \
            if (sols) {                                                                      \
                for (int64_t q = 0; q < x_stride; q++) sols[np * x_stride + q] = q < xl ? x[q] : 0.0; \
            }                                                                                \
        }                                                                                    \
        np++;                                                                                \
    } while (0)
    STORE_POINT(tstart);
    tstate_update(dev, ndev, x, xl, &sbe);
    tstate_update(dev, ndev, x, xl, &str);
    int status = 0;
    while (t < tstop) {
        dt = bp_limit_dt(&bp, t, dt, min_dt);
        if (bp.settle > 0) { double sd = dt * bp.factor; dt = sd >= min_dt ? sd : min_dt; }
        if (t + dt > tstop) dt = tstop - t;
        double tt = t + dt;
        /* BE solve */
        int64_t xbl = xl;
        double *xb = (double *)orc_xmalloc((size_t)(xl ? xl : 1) * sizeof(double));
        memcpy(xb, x, (size_t)xl * sizeof(double));
        sbe.method = ORC_BE;
        build_ctx bbe = {dev, ndev, node_count, ORC_MODE_TRAN, tt, dt, &sbe};
        orc_newton_result rb = run_newton_stepping(o, &xb, &xbl, &bbe, solver, &fcalls, &iters_total);
        if (!rb.converged) {
            free(xb);
            consec++; res->rejected++;
            if (consec >= 10) { status = 2; break; }
            double d4 = dt * 0.25; dt = d4 >= min_dt ? d4 : min_dt;
            continue;
        }
        int64_t xtl = xl;
        double *xt = (double *)orc_xmalloc((size_t)(xl ? xl : 1) * sizeof(double));
        memcpy(xt, x, (size_t)xl * sizeof(double));
        str.method = ORC_TRAP;
        build_ctx btr = {dev, ndev, node_count, ORC_MODE_TRAN, tt, dt, &str};
        orc_newton_result rt = run_newton_stepping(o, &xt, &xtl, &btr, solver, &fcalls, &iters_total);
        if (!rt.converged) {
            free(xb); free(xt);
            consec++; res->rejected++;
            if (consec >= 10) { status = 2; break; }
            double d4 = dt * 0.25; dt = d4 >= min_dt ? d4 : min_dt;
            continue;
        }
        int64_t nl = xbl < xtl ? xbl : xtl;
        double lte[5];
        orc_lte_milne(xb, xt, nl, abs_tol, rel_tol, lte);
        /* AdaptiveStepController::process_lte (analysis.rs:720-742) */
        double err = lte[0];
        int accept = lte[2] != 0.0 && !(err > 1.0);
        double dt_new;
        if (accept) {
            double en = err > 1e-10 ? err : 1e-10;
            double pf = pow(1.0 / en, k_p);
            double inf = 1.0;
            if (has_prev_err) { double ep = prev_err > 1e-10 ? prev_err : 1e-10; inf = pow(1.0 / ep, k_i); }
            double rawf = pf * inf * 0.9;
            double fac = clampd(rawf, 0.1, 2.0);
            double sug = dt * fac;
            has_prev_err = 1; prev_err = err; ctl_consec = 0;
            dt_new = clampd(sug, min_dt, max_dt);
        } else {
            ctl_consec++;
            if (ctl_consec >= 3) {
                if (ctl_consec <= 5) dt_new = dt * 0.25;
                else if (ctl_consec <= 10) dt_new = dt * 0.1;
                else dt_new = min_dt;
            } else {
                dt_new = clampd(dt * lte[3], min_dt, max_dt);
            }
        }
        free(xb);
        if (accept) {
            double tn = t + dt;
            bp_update_settling(&bp, t, tn);
            free(x);
            x = xt; xl = xtl;
            tstate_update_full(dev, ndev, x, xl, dt, &str);
            tstate_update(dev, ndev, x, xl, &sbe);
            STORE_POINT(tn);
            t = tn;
            res->accepted++;
            consec = 0;
            dt = clampd(dt_new, min_dt, max_dt);
        } else {
            free(xt);
            res->rejected++;
            consec++;
            if (consec >= 10) { status = 2; break; }
            dt = clampd(dt_new, min_dt, max_dt);
        }
    }
#undef STORE_POINT
    res->status = status;
    res->npoints = np;
    res->newton_iters_total = iters_total;
    res->factor_calls = fcalls;
    free(x); tstate_free(&sbe); tstate_free(&str); free(bp.t);
    return status == 0 ? 0 : 1;
}

int orc_tran_fixed(const rs_device *dev, int64_t ndev, int64_t node_count, const rs_newton_opts *o,
                   double dt, int64_t nsteps, void *solver, double *x_final, int64_t x_cap,
                   int64_t *x_len, int32_t *newton_iters_total) {
    int64_t xl = node_count;
    double *x = (double *)orc_xcalloc((size_t)(xl ? xl : 1), sizeof(double));
    orc_tstate str;
    tstate_init(&str, ndev, ORC_TRAP);
    sv_prepare(solver, node_count);
    build_ctx bc = {dev, ndev, node_count, ORC_MODE_DC, 0.0, 0.0, NULL};
    int32_t it = 0;
    orc_newton_result r = run_newton_stepping(o, &x, &xl, &bc, solver, NULL, &it);
    int rc = r.converged ? 0 : 1;
    if (!rc) {
        tstate_update(dev, ndev, x, xl, &str);
        double t = 0.0;
        for (int64_t s = 0; s < nsteps; s++) {
            double tt = t + dt;
            build_ctx btr = {dev, ndev, node_count, ORC_MODE_TRAN, tt, dt, &str};
            r = run_newton_stepping(o, &x, &xl, &btr, solver, NULL, &it);
            if (!r.converged) { rc = 2; break; }
            tstate_update_full(dev, ndev, x, xl, dt, &str);
            t = tt;
        }
    }
    int64_t cpy = xl < x_cap ? xl : x_cap;
    if (x_final) memcpy(x_final, x, (size_t)cpy * sizeof(double));
    *x_len = xl;
    if (newton_iters_total) *newton_iters_total = it;
    free(x); tstate_free(&str);
    return rc;
}

/* engine.rs:715-741: n = floor(|range/step|) + 1 points start + i*step, then
 * stop appended when the last point misses it by more than half a step (and
 * fewer than 10000 points were generated). A step below 1e-15 gives {start}. */
int64_t orc_dc_sweep_points(double start, double stop, double step, double *values,
                            int64_t cap) {
    double st = stop >= start ? fabs(step) : -fabs(step);
    int64_t cnt = 0;
    double last = start;
    if (fabs(st) < 1e-15) {
        if (values && cap > 0) values[0] = start;
        return 1;
    }
    int64_t np = (int64_t)floor(fabs((stop - start) / st)) + 1;
    for (int64_t i = 0; i < np; i++) {
        last = start + (double)i * st;
        if (values && cnt < cap) values[cnt] = last;
        cnt++;
    }
    if (fabs(last - stop) > 1e-12 && cnt < 10000 && fabs(last - stop) / fabs(st) > 0.5) {
        if (values && cnt < cap) values[cnt] = stop;
        cnt++;
    }
    return cnt;
}

/* engine.rs:684-820: one solver prepared with node_count, x carried from
 * point to point, the swept value written into the instance (a PULSE source
 * parses as a plain number afterwards), stop at the first failure. */
int orc_dc_sweep(const rs_device *dev, int64_t ndev, int64_t node_count,
                 const rs_newton_opts *o, int64_t source, double start, double stop,
                 double step, void *solver, int64_t max_points, double *sols,
                 int64_t x_stride, orc_sweep_result *res) {
    memset(res, 0, sizeof(*res));
    if (source < 0 || source >= ndev) { res->status = 3; return 3; }
    int64_t np = orc_dc_sweep_points(start, stop, step, NULL, 0);
    double *pts = (double *)orc_xmalloc((size_t)np * sizeof(double));
    orc_dc_sweep_points(start, stop, step, pts, np);
    res->npoints = np;
    rs_device *d = (rs_device *)orc_xmalloc((size_t)(ndev ? ndev : 1) * sizeof(rs_device));
    memcpy(d, dev, (size_t)ndev * sizeof(rs_device));
    d[source].flags &= ~RS_F_WAVEFORM;
    int64_t xl = node_count;
    double *x = (double *)orc_xcalloc((size_t)(xl ? xl : 1), sizeof(double));
    sv_prepare(solver, node_count);
    build_ctx bc = {d, ndev, node_count, ORC_MODE_DC, 0.0, 0.0, NULL};
    for (int64_t k = 0; k < np; k++) {
        d[source].value = pts[k];
        orc_newton_result r = run_newton_stepping(o, &x, &xl, &bc, solver, NULL, NULL);
        res->iterations += r.iterations;
        if (r.converged) {
            if (sols && res->nsolved < max_points)
                for (int64_t q = 0; q < x_stride; q++)
                    sols[res->nsolved * x_stride + q] = q < xl ? x[q] : 0.0;
            res->nsolved++;
        } else {
            res->status = r.reason == 2 ? 2 : 1;
            break;
        }
    }
    free(x); free(d); free(pts);
    return res->status;
}
