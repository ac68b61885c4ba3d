/*
 * orc_mna.c - MNA assembly and device stamps. TEST INFRASTRUCTURE (oracle).
 *
 * Follows crates/sim-core/src/mna.rs (SparseBuilder insert/finalize,
 * StampContext, MnaBuilder), crates/sim-core/src/stamp.rs (R/I/V/C/D/M stamps,
 * DC and TRAN), crates/sim-devices/src/bsim/evaluate.rs:285-356 (level 1) and
 * crates/sim-core/src/waveform.rs:82-113 (PULSE).
 */
#include "orc.h"
#include "orc_util.h"

typedef struct { int64_t row; double v; } entry;
typedef struct { entry *a; int64_t len, cap; } evec;

typedef struct {
    int64_t n, node_count;
    evec *cols;
    int64_t ncols_cap;
    double *rhs;
    int64_t *aux_dev; /* aux id -> device index (name key) */
    int64_t naux, aux_cap;
    double gmin, source_scale;
} builder;

static void b_resize(builder *b, int64_t new_n) { /* mna.rs:235-256 */
    if (new_n <= b->n) return;
    if (new_n > b->ncols_cap) {
        int64_t cap = b->ncols_cap ? b->ncols_cap : 16;
        while (cap < new_n) cap *= 2;
        b->cols = (evec *)realloc(b->cols, (size_t)cap * sizeof(evec));
        b->rhs = (double *)realloc(b->rhs, (size_t)cap * sizeof(double));
        for (int64_t i = b->ncols_cap; i < cap; i++) b->cols[i] = (evec){0};
        b->ncols_cap = cap;
    }
    for (int64_t i = b->n; i < new_n; i++) { b->cols[i].len = 0; b->rhs[i] = 0.0; }
    b->n = new_n;
}

/* SparseBuilder::insert(col, row, v) (mna.rs:151-168) */
static void b_insert(builder *b, int64_t col, int64_t row, double v) {
    if (col >= b->n || row >= b->n) return;
    evec *c = &b->cols[col];
    if (c->len == c->cap) {
        c->cap = c->cap ? c->cap * 2 : 8;
        c->a = (entry *)realloc(c->a, (size_t)c->cap * sizeof(entry));
    }
    c->a[c->len++] = (entry){row, v};
}
/* StampContext::add(i, j, v) = insert(col=j, row=i) (mna.rs:462-464) */
static void ctx_add(builder *b, int64_t i, int64_t j, double v) { b_insert(b, j, i, v); }
static void ctx_add_rhs(builder *b, int64_t i, double v) {
    if (i >= 0 && i < b->n) b->rhs[i] += v;
}
/* StampContext::allocate_aux (mna.rs:472-480), keyed by device index */
static int64_t ctx_alloc_aux(builder *b, int64_t dev) {
    for (int64_t k = 0; k < b->naux; k++)
        if (b->aux_dev[k] == dev) return b->node_count + k;
    if (b->naux == b->aux_cap) {
        b->aux_cap = b->aux_cap ? b->aux_cap * 2 : 8;
        b->aux_dev = (int64_t *)realloc(b->aux_dev, (size_t)b->aux_cap * sizeof(int64_t));
    }
    b->aux_dev[b->naux++] = dev;
    b_resize(b, b->node_count + b->naux);
    return b->node_count + b->naux - 1;
}

double orc_pulse_eval(const double p[7], double t) { /* waveform.rs:82-113 */
    double v1 = p[0], v2 = p[1], td = p[2], tr = p[3], tf = p[4], pw = p[5], per = p[6];
    if (t < td) return v1;
    double tr_ = per > 0.0 ? fmod(t - td, per) : t - td;
    if (tr_ < tr) return v1 + (v2 - v1) * tr_ / tr;
    if (tr_ < tr + pw) return v2;
    if (tr_ < tr + pw + tf) {
        double tfall = tr_ - tr - pw;
        return v2 - (v2 - v1) * tfall / tf;
    }
    return v1;
}

/* Planted mistakes for the tests that show a parity check can fail (0: the
 * reference's behaviour). 1: PMOS current sign flipped in orc_level1. */
static int g_mutation = 0;
void orc_set_mutation(int m) { g_mutation = m; }

/* evaluate_level1_dc (bsim/evaluate.rs:285-356) with beta from
 * bsim/mod.rs:441-455: lambda fixed at 0.02, beta = u0*1e-4*cox, cox =
 * EPSILON_OX/tox with EPSILON_OX = 3.9*8.854e-12 (params.rs:10, 460-462). */
void orc_level1(const rs_device *d, double vd, double vg, double vs, double vb,
                double out[5]) {
    (void)vb;
    const double GMIN = 1e-12;
    const double EPS_OX = 3.9 * 8.854e-12;
    int is_pmos = (d->flags & RS_F_PMOS) != 0;
    double vth0 = d->p[RS_MP_VTH0];
    double lambda = 0.02;
    double cox = EPS_OX / d->p[RS_MP_TOX];
    double beta = d->p[RS_MP_U0] * 1e-4 * cox;
    double w = d->p[RS_MP_W], l = d->p[RS_MP_L];
    double vdi, vgi, vsi, sign;
    if (is_pmos) { vdi = -vs; vgi = -vg; vsi = -vd; sign = g_mutation == 1 ? 1.0 : -1.0; }
    else { vdi = vd; vgi = vg; vsi = vs; sign = 1.0; }
    double vgs = vgi - vsi, vds = vdi - vsi;
    if (vds < 0.0) { vds = -vds; vgs = vgi - vdi; }
    double vth = is_pmos ? -fabs(vth0) : fabs(vth0);
    double beta_eff = beta * w / l;
    double ids, gm, gds;
    if (vgs <= vth) {
        ids = 0.0; gm = 0.0; gds = GMIN;
    } else if (vds < vgs - vth) {
        ids = beta_eff * ((vgs - vth) * vds - 0.5 * vds * vds);
        gm = beta_eff * vds;
        gds = beta_eff * ((vgs - vth) - vds);
        if (!(gds >= GMIN)) gds = GMIN; /* f64::max: NaN-free inputs here */
    } else {
        double vov = vgs - vth;
        ids = 0.5 * beta_eff * (vov * vov) * (1.0 + lambda * vds);
        gm = beta_eff * (vgs - vth) * (1.0 + lambda * vds);
        gds = 0.5 * beta_eff * (vov * vov) * lambda;
        if (!(gds >= GMIN)) gds = GMIN;
    }
    double ids_s = ids * sign;
    double vgs_o = vg - vs, vds_o = vd - vs;
    double ieq = ids_s - gm * vgs_o - gds * vds_o;
    out[0] = ids_s; out[1] = gm; out[2] = gds; out[3] = 0.0; out[4] = ieq;
}

static double xget(const double *x, int64_t xl, int64_t i) {
    return (i >= 0 && i < xl) ? x[i] : 0.0;
}

static int cmp_entry_row(const void *p, const void *q) {
    const entry *a = (const entry *)p, *b = (const entry *)q;
    return (a->row > b->row) - (a->row < b->row);
}

/* stable sort by row (Rust slice::sort_by_key is stable): insertion-ordered
 * merge sort on (row, original position) */
static void stable_sort(entry *a, int64_t n) {
    if (n < 2) return;
    /* insertion sort for short columns, merge sort otherwise */
    if (n <= 32) {
        for (int64_t i = 1; i < n; i++) {
            entry e = a[i];
            int64_t j = i - 1;
            while (j >= 0 && a[j].row > e.row) { a[j + 1] = a[j]; j--; }
            a[j + 1] = e;
        }
        return;
    }
    entry *tmp = (entry *)orc_xmalloc((size_t)n * sizeof(entry));
    int64_t h = n / 2;
    stable_sort(a, h);
    stable_sort(a + h, n - h);
    int64_t i = 0, j = h, k = 0;
    while (i < h && j < n) tmp[k++] = (a[j].row < a[i].row) ? a[j++] : a[i++];
    while (i < h) tmp[k++] = a[i++];
    while (j < n) tmp[k++] = a[j++];
    memcpy(a, tmp, (size_t)n * sizeof(entry));
    free(tmp);
    (void)cmp_entry_row;
}

void orc_build_mna(const rs_device *dev, int64_t ndev, int64_t node_count, const double *x,
                   int64_t xl, int mode, double t, double dt, const orc_tstate *st,
                   double gmin_ctx, double source_scale, orc_mna *out) {
    builder b = {0};
    b.node_count = node_count;
    b.gmin = gmin_ctx; b.source_scale = source_scale;
    b_resize(&b, node_count);
    for (int64_t di = 0; di < ndev; di++) {
        const rs_device *d = &dev[di];
        int64_t a = d->node[0], c = d->node[1];
        switch (d->kind) {
        case RS_DEV_R: { /* stamp.rs:242-259 */
            double g = 1.0 / d->value;
            ctx_add(&b, a, a, g); ctx_add(&b, c, c, g);
            ctx_add(&b, a, c, -g); ctx_add(&b, c, a, -g);
            break;
        }
        case RS_DEV_I: { /* stamp.rs:261-276 (DC), :331-349 (at_time) */
            double v;
            if (mode == ORC_MODE_DC) {
                if (d->flags & RS_F_WAVEFORM) break; /* MissingValue */
                v = d->value;
            } else {
                v = (d->flags & RS_F_WAVEFORM) ? orc_pulse_eval(d->p, t) : d->value;
            }
            v = v * b.source_scale;
            ctx_add_rhs(&b, a, -v); ctx_add_rhs(&b, c, v);
            break;
        }
        case RS_DEV_V: { /* stamp.rs:278-300 (DC), :303-325 (at_time) */
            double v;
            if (mode == ORC_MODE_DC) {
                if (d->flags & RS_F_WAVEFORM) break; /* MissingValue before aux */
                v = d->value;
            } else {
                v = (d->flags & RS_F_WAVEFORM) ? orc_pulse_eval(d->p, t) : d->value;
            }
            v = v * b.source_scale;
            int64_t k = ctx_alloc_aux(&b, di);
            ctx_add(&b, a, k, 1.0); ctx_add(&b, c, k, -1.0);
            ctx_add(&b, k, a, 1.0); ctx_add(&b, k, c, -1.0);
            ctx_add_rhs(&b, k, v);
            break;
        }
        case RS_DEV_C: {
            if (mode == ORC_MODE_DC) break; /* open in DC (stamp.rs:137) */
            double cv = d->value;
            double vprev = (st && st->has_v[di]) ? st->cap_v[di] : 0.0;
            if (st->method == ORC_BE) { /* stamp.rs:759-787 */
                double g = cv / dt;
                double ieq = g * vprev;
                ctx_add(&b, a, a, g); ctx_add(&b, c, c, g);
                ctx_add(&b, a, c, -g); ctx_add(&b, c, a, -g);
                ctx_add_rhs(&b, a, -ieq); ctx_add_rhs(&b, c, ieq);
            } else { /* stamp.rs:871-915 */
                double g = 2.0 * cv / dt;
                double iprev = (st && st->has_i[di]) ? st->cap_i[di] : 0.0;
                double ieq = g * vprev + iprev;
                ctx_add(&b, a, a, g); ctx_add(&b, c, c, g);
                ctx_add(&b, a, c, -g); ctx_add(&b, c, a, -g);
                ctx_add_rhs(&b, a, ieq); ctx_add_rhs(&b, c, -ieq);
            }
            break;
        }
        case RS_DEV_D: { /* stamp.rs:351-387 */
            double gmin = b.gmin > 0.0 ? b.gmin : 1e-12;
            double isat = d->p[0], emis = d->p[1];
            double vt = 0.02585 * emis;
            double vdd = xget(x, xl, a) - xget(x, xl, c);
            double e = exp(vdd / vt);
            double id = isat * (e - 1.0);
            double gd = (isat / vt) * e;
            double g = gd >= gmin ? gd : gmin;
            double ieq = id - gd * vdd;
            ctx_add(&b, a, a, g); ctx_add(&b, c, c, g);
            ctx_add(&b, a, c, -g); ctx_add(&b, c, a, -g);
            ctx_add_rhs(&b, a, -ieq); ctx_add_rhs(&b, c, ieq);
            break;
        }
        case RS_DEV_M: { /* stamp.rs:389-553 */
            int64_t dr = d->node[0], ga = d->node[1], so = d->node[2], bu = d->node[3];
            double gmin = b.gmin > 0.0 ? b.gmin : 1e-12;
            double vd = xget(x, xl, dr), vg = xget(x, xl, ga), vs = xget(x, xl, so), vb = xget(x, xl, bu);
            int level = (int)d->p[RS_MP_LEVEL];
            double o[13] = {0};
            if (level == 1) {
                orc_level1(d, vd, vg, vs, vb, o);
            } else {
                const rs_bsim_model *mm = orc_bsim_model((int64_t)d->p[RS_MP_MODEL]);
                if (!mm) break; /* no model: the caller did not provide one */
                orc_bsim_eval(mm, level, d->p[RS_MP_W], d->p[RS_MP_L], vd, vg, vs, vb,
                              d->p[RS_MP_TEMP], o);
            }
            double gm = o[1], gds = o[2] >= gmin ? o[2] : gmin, gmbs = o[3], ieq = o[4];
            ctx_add(&b, dr, dr, gds); ctx_add(&b, so, so, gds);
            ctx_add(&b, dr, so, -gds); ctx_add(&b, so, dr, -gds);
            ctx_add(&b, dr, ga, gm); ctx_add(&b, dr, so, -gm);
            ctx_add(&b, so, ga, -gm); ctx_add(&b, so, so, gm);
            if (fabs(gmbs) > gmin * 0.01) {
                ctx_add(&b, dr, bu, gmbs); ctx_add(&b, dr, so, -gmbs);
                ctx_add(&b, so, bu, -gmbs); ctx_add(&b, so, so, gmbs);
            }
            ctx_add_rhs(&b, dr, -ieq); ctx_add_rhs(&b, so, ieq);
            if (level == 54) { /* BSIM4 substrate and gate currents (stamp.rs:466-509) */
                double isub = o[7], gsub = o[8], igs = o[9], igd = o[10], gigs = o[11], gigd = o[12];
                if (fabs(isub) > gmin && gsub > gmin * 0.01) {
                    ctx_add(&b, dr, dr, gsub); ctx_add(&b, bu, bu, gsub);
                    ctx_add(&b, dr, bu, -gsub); ctx_add(&b, bu, dr, -gsub);
                    double isub_eq = isub - gsub * (vd - vb);
                    ctx_add_rhs(&b, dr, -isub_eq); ctx_add_rhs(&b, bu, isub_eq);
                }
                if (fabs(igs) > gmin && gigs > gmin * 0.01) {
                    ctx_add(&b, ga, ga, gigs); ctx_add(&b, so, so, gigs);
                    ctx_add(&b, ga, so, -gigs); ctx_add(&b, so, ga, -gigs);
                    double igs_eq = igs - gigs * (vg - vs);
                    ctx_add_rhs(&b, ga, -igs_eq); ctx_add_rhs(&b, so, igs_eq);
                }
                if (fabs(igd) > gmin && gigd > gmin * 0.01) {
                    ctx_add(&b, ga, ga, gigd); ctx_add(&b, dr, dr, gigd);
                    ctx_add(&b, ga, dr, -gigd); ctx_add(&b, dr, ga, -gigd);
                    double igd_eq = igd - gigd * (vg - vd);
                    ctx_add_rhs(&b, ga, -igd_eq); ctx_add_rhs(&b, dr, igd_eq);
                }
            }
            break;
        }
        default: break;
        }
    }
    /* ground pin (engine.rs:279-280) then finalize (mna.rs:320-347) */
    b_insert(&b, 0, 0, 1.0);
    int64_t n = b.n, nnz = 0;
    for (int64_t cidx = 0; cidx < n; cidx++) nnz += b.cols[cidx].len;
    out->n = n; out->nnz = nnz;
    out->ap = (int64_t *)orc_xmalloc((size_t)(n + 1) * sizeof(int64_t));
    out->ai = (int64_t *)orc_xmalloc((size_t)(nnz ? nnz : 1) * sizeof(int64_t));
    out->ax = (double *)orc_xmalloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    out->rhs = (double *)orc_xmalloc((size_t)(n ? n : 1) * sizeof(double));
    int64_t p = 0;
    out->ap[0] = 0;
    for (int64_t cidx = 0; cidx < n; cidx++) {
        evec *cv = &b.cols[cidx];
        stable_sort(cv->a, cv->len);
        for (int64_t q = 0; q < cv->len; q++) { out->ai[p] = cv->a[q].row; out->ax[p] = cv->a[q].v; p++; }
        out->ap[cidx + 1] = p;
    }
    memcpy(out->rhs, b.rhs, (size_t)n * sizeof(double));
    for (int64_t i = 0; i < b.ncols_cap; i++) free(b.cols[i].a);
    free(b.cols); free(b.rhs); free(b.aux_dev);
}

void orc_mna_free(orc_mna *m) {
    free(m->ap); free(m->ai); free(m->ax); free(m->rhs);
    memset(m, 0, sizeof(*m));
}
