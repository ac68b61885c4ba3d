/*
 * orc_ac.c - CPU oracle of the AC (small-signal) analysis. TEST
 * INFRASTRUCTURE ONLY (see orc.h).
 *
 * Restates, in the reference's order of operations:
 *   - generate_frequency_sweep (crates/sim-core/src/engine.rs:1002-1044);
 *   - the AC stamps at the DC operating point (stamp.rs:1525-1747: R, C, V,
 *     I, diode, MOSFET through evaluate_mos) into ComplexSparseBuilder
 *     (complex_mna.rs:10-59: per-column insertion lists, stable sort by row);
 *   - the engine's ground fix (engine.rs:882-891: every entry of the ground
 *     row dropped, (0,0) = 1, rhs[0] = 0);
 *   - ComplexDenseSolver::solve (complex_solver.rs:50-136): dense row-major
 *     LU with partial pivoting on |z| (hypot), singular below 1e-30;
 *   - the (dB, degrees) conversion of engine.rs:932-946.
 * Complex arithmetic follows num-complex 0.4.6 (Cargo.lock:770-772), the
 * crate the reference uses; it is not vendored, so its published formulas
 * are restated: (a+ib)(c+id) = (ac-bd) + i(ad+bc); (a+ib)/(c+id) =
 * ((ac+bd) + i(bc-ad)) / (c*c+d*d); |z| = hypot(re, im); arg z = atan2(im, re);
 * from_polar(r, t) = (r cos t, r sin t).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"
#include "orc_util.h"

/* std::f64::consts::PI */
static const double ORC_PI = 3.14159265358979323846;

typedef struct { double re, im; } cplx;

static cplx c_add(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static cplx c_sub(cplx a, cplx b) { cplx r = {a.re - b.re, a.im - b.im}; return r; }
static cplx c_mul(cplx a, cplx b) {
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static cplx c_div(cplx a, cplx b) {
    double ns = b.re * b.re + b.im * b.im;
    double re = a.re * b.re + a.im * b.im;
    double im = a.im * b.re - a.re * b.im;
    cplx r = {re / ns, im / ns};
    return r;
}
static double c_norm(cplx a) { return hypot(a.re, a.im); }

/* engine.rs:1002-1044. sweep: 0 dec, 1 oct, 2 lin. */
int64_t orc_ac_frequencies(int sweep, int64_t points, double fstart, double fstop, double *out,
                           int64_t cap) {
    int64_t total = 0;
    if (sweep == 0 || sweep == 1) {
        double span = sweep == 0 ? log10(fstop) - log10(fstart) : log2(fstop / fstart);
        double c = ceil(span * (double)points);
        /* `as usize` saturates: NaN and negatives give 0 */
        total = (c != c || c <= 0.0) ? 1 : (int64_t)c + 1;
        double den = (double)(total - 1 > 1 ? total - 1 : 1);
        for (int64_t i = 0; i < total; i++) {
            double f;
            if (sweep == 0) {
                double lf = log10(fstart) + (double)i * span / den;
                f = pow(10.0, lf);
            } else {
                f = fstart * pow(2.0, (double)i * span / den);
            }
            if (i < cap) out[i] = f;
        }
    } else {
        total = points;
        double den = (double)(points - 1 > 1 ? points - 1 : 1);
        double step = (fstop - fstart) / den;
        for (int64_t i = 0; i < total; i++)
            if (i < cap) out[i] = fstart + (double)i * step;
    }
    return total;
}

typedef struct { int64_t row, col, ord; cplx v; } entry;

typedef struct {
    entry *e;
    int64_t n, cap;
} elist;

static void push(elist *l, int64_t row, int64_t col, cplx v) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 64;
        l->e = (entry *)realloc(l->e, (size_t)l->cap * sizeof(entry));
        if (!l->e) abort();
    }
    entry x = {row, col, l->n, v};
    l->e[l->n++] = x;
}
static void add_real(elist *l, int64_t i, int64_t j, double v) { cplx z = {v, 0.0}; push(l, i, j, z); }
static void add_imag(elist *l, int64_t i, int64_t j, double v) { cplx z = {0.0, v}; push(l, i, j, z); }

static double xat(const double *x, int64_t nx, int64_t i) { return i < nx ? x[i] : 0.0; }

/* stamp.rs:1525-1747 for one device; aux rows are numbered node_count + k in
 * the order the V sources stamp. */
static void stamp_ac(const rs_device *d, int64_t di, int64_t node_count, int64_t *naux,
                     const double *x, int64_t nx, double omega, const double *ac_mag,
                     const double *ac_phase, elist *l, cplx *rhs) {
    const int64_t a = d->node[0], b = d->node[1];
    switch (d->kind) {
    case RS_DEV_R: {
        double g = 1.0 / d->value;
        add_real(l, a, a, g); add_real(l, b, b, g); add_real(l, a, b, -g); add_real(l, b, a, -g);
        break;
    }
    case RS_DEV_C: {
        double y = omega * d->value;
        add_imag(l, a, a, y); add_imag(l, b, b, y); add_imag(l, a, b, -y); add_imag(l, b, a, -y);
        break;
    }
    case RS_DEV_V: {
        int64_t k = node_count + (*naux)++;
        add_real(l, a, k, 1.0); add_real(l, b, k, -1.0);
        add_real(l, k, a, 1.0); add_real(l, k, b, -1.0);
        double th = ac_phase[di] * ORC_PI / 180.0;
        cplx v = {ac_mag[di] * cos(th), ac_mag[di] * sin(th)};
        if (rhs) rhs[k] = c_add(rhs[k], v);
        break;
    }
    case RS_DEV_I: {
        double th = ac_phase[di] * ORC_PI / 180.0;
        cplx v = {ac_mag[di] * cos(th), ac_mag[di] * sin(th)};
        cplx nv = {-v.re, -v.im};
        if (rhs) {
            rhs[a] = c_add(rhs[a], nv);
            rhs[b] = c_add(rhs[b], v);
        }
        break;
    }
    case RS_DEV_D: {
        double vt = 0.02585 * d->p[1];
        double vd = xat(x, nx, a) - xat(x, nx, b);
        double e = exp(vd / vt);
        double gd = (d->p[0] / vt) * e;
        double g = (gd >= 1e-12 || gd != gd) ? gd : 1e-12; /* gd.max(gmin) */
        add_real(l, a, a, g); add_real(l, b, b, g); add_real(l, a, b, -g); add_real(l, b, a, -g);
        break;
    }
    case RS_DEV_M: {
        const int64_t dr = d->node[0], ga = d->node[1], so = d->node[2], bu = d->node[3];
        double vd = xat(x, nx, dr), vg = xat(x, nx, ga), vs = xat(x, nx, so), vb = xat(x, nx, bu);
        double gm, gds, gmbs;
        if ((int)d->p[RS_MP_LEVEL] == 1) {
            double o[5];
            orc_level1(d, vd, vg, vs, vb, o);
            gm = o[1]; gds = o[2]; gmbs = o[3];
        } else {
            /* stamp_mos_ac calls evaluate_mos, which runs BSIM4 with sa = sb = 0
             * (bsim/mod.rs:463-469) whatever the instance gives */
            double o[13];
            rs_bsim_model m = *orc_bsim_model((int64_t)d->p[RS_MP_MODEL]);
            m.sa = 0.0;
            m.sb = 0.0;
            orc_bsim_eval(&m, (int)d->p[RS_MP_LEVEL], d->p[RS_MP_W], d->p[RS_MP_L], vd, vg, vs, vb,
                          d->p[RS_MP_TEMP], o);
            gm = o[1]; gds = o[2]; gmbs = o[3];
        }
        if (!(gds >= 1e-12 || gds != gds)) gds = 1e-12;
        add_real(l, dr, dr, gds); add_real(l, so, so, gds); add_real(l, dr, so, -gds); add_real(l, so, dr, -gds);
        add_real(l, dr, ga, gm); add_real(l, dr, so, -gm); add_real(l, so, ga, -gm); add_real(l, so, so, gm);
        if (fabs(gmbs) > 1e-12 * 0.01) {
            add_real(l, dr, bu, gmbs); add_real(l, dr, so, -gmbs);
            add_real(l, so, bu, -gmbs); add_real(l, so, so, gmbs);
        }
        break;
    }
    default:
        break;
    }
}

/* complex_solver.rs:50-136; returns 1 when solved. a is n*n row-major. */
static int dense_solve(int64_t n, cplx *a, cplx *x) {
    for (int64_t k = 0; k < n; k++) {
        double mv = c_norm(a[k * n + k]);
        int64_t mr = k;
        for (int64_t i = k + 1; i < n; i++) {
            double v = c_norm(a[i * n + k]);
            if (v > mv) { mv = v; mr = i; }
        }
        if (mr != k) {
            for (int64_t j = 0; j < n; j++) {
                cplx t = a[k * n + j]; a[k * n + j] = a[mr * n + j]; a[mr * n + j] = t;
            }
            cplx t = x[k]; x[k] = x[mr]; x[mr] = t;
        }
        cplx piv = a[k * n + k];
        if (c_norm(piv) < 1e-30) return 0;
        for (int64_t i = k + 1; i < n; i++) {
            cplx f = c_div(a[i * n + k], piv);
            a[i * n + k] = f;
            for (int64_t j = k + 1; j < n; j++) a[i * n + j] = c_sub(a[i * n + j], c_mul(f, a[k * n + j]));
            x[i] = c_sub(x[i], c_mul(f, x[k]));
        }
    }
    for (int64_t i = n - 1; i >= 0; i--) {
        for (int64_t j = i + 1; j < n; j++) x[i] = c_sub(x[i], c_mul(a[i * n + j], x[j]));
        cplx piv = a[i * n + i];
        if (c_norm(piv) < 1e-30) return 0;
        x[i] = c_div(x[i], piv);
    }
    return 1;
}

static int by_col_row(const void *p, const void *q) {
    const entry *a = (const entry *)p, *b = (const entry *)q;
    if (a->col != b->col) return a->col < b->col ? -1 : 1;
    if (a->row != b->row) return a->row < b->row ? -1 : 1;
    return a->ord < b->ord ? -1 : (a->ord > b->ord);
}

/* Engine::run_ac_result (engine.rs:864-949) after its DC operating point.
 * xout[f * 2 * cap_n + 2 i (+1)]: the complex solution; magph[f * 2 *
 * node_count + 2 i (+1)]: (dB, degrees) of node i. Returns the count of
 * frequencies solved (the reference stops at the first failed solve). */
int64_t orc_ac(const rs_device *dev, int64_t ndev, int64_t node_count, const double *ac_mag,
               const double *ac_phase, const double *x_dc, int64_t nx, const double *freqs,
               int64_t nfreq, double *xout, double *magph, int64_t cap_n, int64_t *n_out) {
    int64_t naux = 0;
    for (int64_t d = 0; d < ndev; d++) naux += dev[d].kind == RS_DEV_V;
    const int64_t n = node_count + naux;
    *n_out = n;
    cplx *a = (cplx *)orc_xcalloc((size_t)(n * n > 0 ? n * n : 1), sizeof(cplx));
    cplx *x = (cplx *)orc_xcalloc((size_t)(n > 0 ? n : 1), sizeof(cplx));
    elist l = {0};
    int64_t solved = 0;
    for (int64_t f = 0; f < nfreq; f++) {
        const double omega = 2.0 * ORC_PI * freqs[f];
        l.n = 0;
        for (int64_t i = 0; i < n; i++) { x[i].re = 0.0; x[i].im = 0.0; }
        int64_t k = 0;
        for (int64_t d = 0; d < ndev; d++)
            stamp_ac(&dev[d], d, node_count, &k, x_dc, nx, omega, ac_mag, ac_phase, &l, x);
        /* ground fix: drop row 0, then (0,0) = 1 (engine.rs:882-891) */
        int64_t m = 0;
        for (int64_t q = 0; q < l.n; q++)
            if (l.e[q].row != 0) l.e[m++] = l.e[q];
        l.n = m;
        add_real(&l, 0, 0, 1.0);
        x[0].re = 0.0; x[0].im = 0.0;
        /* stable sort by (col, row) (complex_mna.rs:43-47): the insertion
         * index breaks ties, so an entry's terms stay in stamp order */
        for (int64_t q = 0; q < l.n; q++) l.e[q].ord = q;
        entry *s = (entry *)orc_xcalloc((size_t)(l.n ? l.n : 1), sizeof(entry));
        memcpy(s, l.e, (size_t)l.n * sizeof(entry));
        qsort(s, (size_t)l.n, sizeof(entry), by_col_row);
        for (int64_t q = 0; q < n * n; q++) { a[q].re = 0.0; a[q].im = 0.0; }
        for (int64_t q = 0; q < l.n; q++) {
            cplx *t = &a[s[q].row * n + s[q].col];
            *t = c_add(*t, s[q].v);
        }
        free(s);
        if (!dense_solve(n, a, x)) break;
        for (int64_t i = 0; i < n && i < cap_n; i++) {
            xout[(f * cap_n + i) * 2] = x[i].re;
            xout[(f * cap_n + i) * 2 + 1] = x[i].im;
        }
        for (int64_t i = 0; i < node_count; i++) {
            double mag = c_norm(x[i]);
            double db = mag > 1e-30 ? 20.0 * log10(mag) : -600.0;
            double ph = atan2(x[i].im, x[i].re) * 180.0 / ORC_PI;
            magph[(f * node_count + i) * 2] = db;
            magph[(f * node_count + i) * 2 + 1] = ph;
        }
        solved++;
    }
    free(l.e);
    free(a);
    free(x);
    return solved;
}
