/*
 * orc_klu.c - restatement of NativeKluSolver, crates/sim-core/src/native_klu.rs.
 * TEST INFRASTRUCTURE (oracle). Sequential path only; the reference's
 * column-parallel refactor (native_klu.rs:402-576) computes the same numbers
 * (each column's updates are applied in the same ascending-j order).
 */
#include "orc.h"
#include "orc_util.h"

typedef struct {
    int64_t n;
    int64_t *col_perm, *col_perm_inv;
    int64_t *l_col_ptr, *l_row_idx, *u_col_ptr, *u_row_idx;
} symf;

typedef struct {
    double *l_values, *u_values, *u_diag;
    int64_t *row_perm, *row_perm_inv;
    int64_t nl, nu, n;
} numf;

typedef struct {
    int64_t block_row, block_col, row_start, col_start, nrows, ncols;
    int64_t *col_ptr, *row_idx;
    double *values;
    int64_t nnz;
} offd;

enum { ST_EMPTY = 0, ST_ANALYZED = 1, ST_FACTORED = 2 };

struct orc_klu {
    uint32_t magic; /* ORC_MAGIC_KLU: the engine dispatches on it */
    int64_t n;
    double pivot_tol;
    int ordering, use_btf;
    int64_t btf_min_blocks;
    int has_btf;
    int sparselu; /* SparseLuSolver semantics: no pivoting, no BTF, no singleton rule */
    orc_btf btf;
    int64_t nblk;
    symf *sym;
    numf *num;
    int64_t nnum;
    offd *off;
    int64_t noff;
    int64_t *last_ap, *last_ai;
    int64_t last_ap_len, last_ai_len;
    int state;
    int64_t factor_count, refactor_count, nnz_l, nnz_u, nblocks, noffdiag;
    double rcond;
    double *work;
};

static void symf_free(symf *s) {
    free(s->col_perm); free(s->col_perm_inv); free(s->l_col_ptr);
    free(s->l_row_idx); free(s->u_col_ptr); free(s->u_row_idx);
}
static void numf_free(numf *f) {
    free(f->l_values); free(f->u_values); free(f->u_diag);
    free(f->row_perm); free(f->row_perm_inv);
}
static void clear_factors(orc_klu *s) {
    for (int64_t b = 0; b < s->nblk; b++) symf_free(&s->sym[b]);
    free(s->sym); s->sym = NULL; s->nblk = 0;
    for (int64_t b = 0; b < s->nnum; b++) numf_free(&s->num[b]);
    free(s->num); s->num = NULL; s->nnum = 0;
    for (int64_t b = 0; b < s->noff; b++) {
        free(s->off[b].col_ptr); free(s->off[b].row_idx); free(s->off[b].values);
    }
    free(s->off); s->off = NULL; s->noff = 0;
}

orc_klu *orc_klu_new(int64_t n) {
    orc_klu *s = (orc_klu *)orc_xcalloc(1, sizeof(orc_klu));
    s->magic = ORC_MAGIC_KLU;
    s->n = n;
    s->pivot_tol = 0.001; s->ordering = 0; s->use_btf = 1; s->btf_min_blocks = 2;
    s->state = ST_EMPTY;
    s->work = (double *)orc_xcalloc((size_t)(n ? n : 1), sizeof(double));
    return s;
}

/* SparseLuSolver (sparse_lu.rs:302-618) is this solver with pivot_tol 0, no
 * BTF and the plain 1e-14 clamp also for a 1x1 matrix (:493-499; it has no
 * singleton fast path). */
void orc_klu_set_sparselu(orc_klu *s) {
    s->pivot_tol = 0.0; s->ordering = 0; s->use_btf = 0; s->sparselu = 1;
}

void orc_klu_set_config(orc_klu *s, double pivot_tol, int ordering, int use_btf) {
    s->pivot_tol = pivot_tol; s->ordering = ordering; s->use_btf = use_btf;
}

void orc_klu_reset_pattern(orc_klu *s) { /* native_klu.rs:1663-1672 */
    s->state = ST_EMPTY;
    if (s->has_btf) orc_btf_free(&s->btf);
    s->has_btf = 0;
    clear_factors(s);
    free(s->last_ap); free(s->last_ai); s->last_ap = s->last_ai = NULL;
    s->last_ap_len = s->last_ai_len = 0;
    s->factor_count = s->refactor_count = s->nnz_l = s->nnz_u = s->nblocks = s->noffdiag = 0;
    s->rcond = 0.0;
}

void orc_klu_free(orc_klu *s) {
    if (!s) return;
    orc_klu_reset_pattern(s);
    free(s->work);
    free(s);
}

void orc_klu_prepare(orc_klu *s, int64_t n) { /* native_klu.rs:1604-1610 */
    if (n != s->n) {
        orc_klu_reset_pattern(s);
        s->n = n;
        free(s->work);
        s->work = (double *)orc_xcalloc((size_t)(n ? n : 1), sizeof(double));
    }
}

static int pattern_matches(const orc_klu *s, int64_t apl, const int64_t *ap, int64_t ail,
                           const int64_t *ai) {
    if (s->last_ap_len != apl || s->last_ai_len != ail) return 0;
    if (apl && memcmp(s->last_ap, ap, (size_t)apl * sizeof(int64_t))) return 0;
    if (ail && memcmp(s->last_ai, ai, (size_t)ail * sizeof(int64_t))) return 0;
    return 1;
}

/* symbolic_for_block (native_klu.rs:863-1014) */
static void symbolic_for_block(const orc_klu *s, int64_t n, const int64_t *ap,
                               const int64_t *ai, symf *out) {
    memset(out, 0, sizeof(*out));
    out->n = n;
    out->col_perm = (int64_t *)orc_xmalloc((size_t)(n ? n : 1) * sizeof(int64_t));
    out->col_perm_inv = (int64_t *)orc_xmalloc((size_t)(n ? n : 1) * sizeof(int64_t));
    out->l_col_ptr = (int64_t *)orc_xcalloc((size_t)n + 1, sizeof(int64_t));
    out->u_col_ptr = (int64_t *)orc_xcalloc((size_t)n + 1, sizeof(int64_t));
    if (n == 0) { out->l_row_idx = orc_xmalloc(8); out->u_row_idx = orc_xmalloc(8); return; }
    if (s->ordering == 3) {
        for (int64_t i = 0; i < n; i++) out->col_perm[i] = out->col_perm_inv[i] = i;
    } else {
        orc_amd_order(n, ap, ai, out->col_perm, out->col_perm_inv);
    }
    /* permuted CSC of P*A*P^T (rows mapped, order within column kept) */
    int64_t *pap = (int64_t *)orc_xcalloc((size_t)n + 1, sizeof(int64_t));
    ivec pai = {0};
    for (int64_t nc = 0; nc < n; nc++) {
        int64_t oc = out->col_perm_inv[nc];
        for (int64_t idx = ap[oc]; idx < ap[oc + 1]; idx++) {
            int64_t orow = ai[idx];
            if (orow < n) iv_push(&pai, out->col_perm[orow]);
        }
        pap[nc + 1] = pai.len;
    }
    ivec lri = {0}, uri = {0}, pattern = {0}, ur = {0}, lr = {0};
    int64_t *mark = (int64_t *)orc_xcalloc((size_t)n, sizeof(int64_t));
    int64_t cur = 0;
    int64_t *st_node = (int64_t *)orc_xmalloc((size_t)(2 * n + 2) * sizeof(int64_t) * 8);
    uint8_t *st_first = (uint8_t *)orc_xmalloc((size_t)(2 * n + 2) * 8);
    int64_t st_cap = (2 * n + 2) * 8;
    for (int64_t k = 0; k < n; k++) {
        cur++;
        pattern.len = 0;
        int64_t sp = 0;
        for (int64_t idx = pap[k]; idx < pap[k + 1]; idx++) {
            int64_t row = pai.a[idx];
            if (row != k && mark[row] != cur) {
                mark[row] = cur;
                if (row < k) { st_node[sp] = row; st_first[sp] = 1; sp++; }
                else iv_push(&pattern, row);
            }
        }
        while (sp > 0) {
            sp--;
            int64_t node = st_node[sp];
            int first = st_first[sp];
            if (first) {
                st_node[sp] = node; st_first[sp] = 0; sp++;
                for (int64_t li = out->l_col_ptr[node]; li < out->l_col_ptr[node + 1]; li++) {
                    int64_t child = lri.a[li];
                    if (mark[child] != cur) {
                        mark[child] = cur;
                        if (child < k) {
                            if (sp >= st_cap) {
                                st_cap *= 2;
                                st_node = (int64_t *)realloc(st_node, (size_t)st_cap * sizeof(int64_t));
                                st_first = (uint8_t *)realloc(st_first, (size_t)st_cap);
                            }
                            st_node[sp] = child; st_first[sp] = 1; sp++;
                        } else {
                            iv_push(&pattern, child);
                        }
                    }
                }
            }
            if (!first) iv_push(&pattern, node);
        }
        ur.len = 0; lr.len = 0;
        for (int64_t q = 0; q < pattern.len; q++) {
            int64_t row = pattern.a[q];
            if (row < k) iv_push(&ur, row);
            else if (row > k) iv_push(&lr, row);
        }
        qsort(ur.a, (size_t)ur.len, sizeof(int64_t), orc_cmp_i64);
        qsort(lr.a, (size_t)lr.len, sizeof(int64_t), orc_cmp_i64);
        out->u_col_ptr[k] = uri.len;
        for (int64_t q = 0; q < ur.len; q++) iv_push(&uri, ur.a[q]);
        iv_push(&uri, k);
        out->u_col_ptr[k + 1] = uri.len;
        out->l_col_ptr[k] = lri.len;
        for (int64_t q = 0; q < lr.len; q++) iv_push(&lri, lr.a[q]);
        out->l_col_ptr[k + 1] = lri.len;
    }
    out->l_row_idx = lri.len ? lri.a : (int64_t *)orc_xmalloc(8);
    if (!lri.len) free(lri.a);
    out->u_row_idx = uri.a;
    free(pap); iv_free(&pai); iv_free(&pattern); iv_free(&ur); iv_free(&lr);
    free(mark); free(st_node); free(st_first);
}

/* extract_block_pattern / extract_block_values (native_klu.rs:764-797, 1151-1188) */
static void extract_block(const orc_klu *s, const int64_t *ap, const int64_t *ai,
                          const double *ax, int64_t bs, int64_t be, int64_t **bap,
                          int64_t **bai, double **bax) {
    int64_t n = s->n, sz = be - bs;
    *bap = (int64_t *)orc_xcalloc((size_t)sz + 1, sizeof(int64_t));
    ivec bi = {0};
    int64_t cap = 16, cnt = 0;
    double *bx = ax ? (double *)orc_xmalloc((size_t)cap * sizeof(double)) : NULL;
    for (int64_t lc = 0; lc < sz; lc++) {
        int64_t gco = s->btf.col_perm[bs + lc];
        for (int64_t idx = ap[gco]; idx < ap[gco + 1]; idx++) {
            int64_t gro = ai[idx];
            if (gro >= n) continue;
            int64_t grn = s->btf.row_perm_inv[gro];
            if (grn >= bs && grn < be) {
                iv_push(&bi, grn - bs);
                if (ax) {
                    if (cnt == cap) { cap *= 2; bx = (double *)realloc(bx, (size_t)cap * sizeof(double)); }
                    bx[cnt] = ax[idx];
                }
                cnt++;
            }
        }
        (*bap)[lc + 1] = bi.len;
    }
    *bai = bi.a ? bi.a : (int64_t *)orc_xmalloc(8);
    if (bax) *bax = bx;
}

/* extract_off_diag_patterns (native_klu.rs:800-857) */
static void extract_offdiag(orc_klu *s, const int64_t *ap, const int64_t *ai) {
    int64_t n = s->n, nb = s->btf.num_blocks, total = 0;
    s->off = NULL; s->noff = 0;
    int64_t cap = 0;
    for (int64_t br = 0; br < nb; br++) {
        int64_t rs = s->btf.block_ptr[br], re = s->btf.block_ptr[br + 1];
        for (int64_t bc = br + 1; bc < nb; bc++) {
            int64_t cs = s->btf.block_ptr[bc], ce = s->btf.block_ptr[bc + 1];
            int64_t ncols = ce - cs;
            int64_t *cp = (int64_t *)orc_xcalloc((size_t)ncols + 1, sizeof(int64_t));
            ivec ri = {0};
            for (int64_t lc = 0; lc < ncols; lc++) {
                int64_t gco = s->btf.col_perm[cs + lc];
                for (int64_t idx = ap[gco]; idx < ap[gco + 1]; idx++) {
                    int64_t gro = ai[idx];
                    if (gro >= n) continue;
                    int64_t grn = s->btf.row_perm_inv[gro];
                    if (grn >= rs && grn < re) iv_push(&ri, grn - rs);
                }
                cp[lc + 1] = ri.len;
            }
            if (ri.len) {
                if (s->noff == cap) {
                    cap = cap ? cap * 2 : 8;
                    s->off = (offd *)realloc(s->off, (size_t)cap * sizeof(offd));
                }
                offd *o = &s->off[s->noff++];
                o->block_row = br; o->block_col = bc; o->row_start = rs; o->col_start = cs;
                o->nrows = re - rs; o->ncols = ncols; o->col_ptr = cp; o->row_idx = ri.a;
                o->nnz = ri.len;
                o->values = (double *)orc_xcalloc((size_t)ri.len, sizeof(double));
                total += ri.len;
            } else {
                free(cp); iv_free(&ri);
            }
        }
    }
    s->noffdiag = total;
}

/* do_analyze (native_klu.rs:704-761) */
static int do_analyze(orc_klu *s, int64_t apl, const int64_t *ap, int64_t ail, const int64_t *ai) {
    int64_t n = s->n;
    int64_t fc = s->factor_count, rc = s->refactor_count;
    if (s->has_btf) orc_btf_free(&s->btf);
    s->has_btf = 0;
    clear_factors(s);
    s->nnz_l = s->nnz_u = s->nblocks = s->noffdiag = 0;
    s->rcond = 0.0;
    s->factor_count = fc; s->refactor_count = rc;
    int use_btf = s->use_btf && n >= 4;
    if (use_btf) {
        orc_btf_decompose(n, ap, ai, &s->btf);
        if (s->btf.num_blocks >= s->btf_min_blocks && s->btf.max_block_size < n) s->has_btf = 1;
        else orc_btf_free(&s->btf);
    }
    if (s->has_btf) {
        s->nblocks = s->btf.num_blocks;
        s->nblk = s->btf.num_blocks;
        s->sym = (symf *)orc_xcalloc((size_t)s->nblk, sizeof(symf));
        for (int64_t k = 0; k < s->nblk; k++) {
            int64_t bs = s->btf.block_ptr[k], be = s->btf.block_ptr[k + 1];
            int64_t *bap, *bai;
            extract_block(s, ap, ai, NULL, bs, be, &bap, &bai, NULL);
            symbolic_for_block(s, be - bs, bap, bai, &s->sym[k]);
            free(bap); free(bai);
        }
        extract_offdiag(s, ap, ai);
    } else {
        s->nblocks = 1;
        s->nblk = 1;
        s->sym = (symf *)orc_xcalloc(1, sizeof(symf));
        symbolic_for_block(s, n, ap, ai, &s->sym[0]);
    }
    s->state = ST_ANALYZED;
    free(s->last_ap); free(s->last_ai);
    s->last_ap = (int64_t *)orc_xmalloc((size_t)(apl ? apl : 1) * sizeof(int64_t));
    s->last_ai = (int64_t *)orc_xmalloc((size_t)(ail ? ail : 1) * sizeof(int64_t));
    memcpy(s->last_ap, ap, (size_t)apl * sizeof(int64_t));
    memcpy(s->last_ai, ai, (size_t)ail * sizeof(int64_t));
    s->last_ap_len = apl; s->last_ai_len = ail;
    for (int64_t b = 0; b < s->nblk; b++) {
        s->nnz_l += s->sym[b].l_col_ptr[s->sym[b].n];
        s->nnz_u += s->sym[b].u_col_ptr[s->sym[b].n];
    }
    return 0;
}

/* factor_block (native_klu.rs:1194-1382) */
static void factor_block(const orc_klu *s, const symf *sym, const int64_t *ap, const int64_t *ai,
                         const double *ax, const numf *old, int refactor, numf *out) {
    int64_t n = sym->n;
    memset(out, 0, sizeof(*out));
    out->n = n;
    if (n == 0) {
        out->l_values = orc_xmalloc(8); out->u_values = orc_xmalloc(8); out->u_diag = orc_xmalloc(8);
        out->row_perm = orc_xmalloc(8); out->row_perm_inv = orc_xmalloc(8);
        return;
    }
    if (n == 1 && !s->sparselu) {
        int64_t oc = sym->col_perm_inv[0];
        double diag = 0.0;
        for (int64_t idx = ap[oc]; idx < ap[oc + 1]; idx++) {
            int64_t row = ai[idx];
            if (row < n && sym->col_perm[row] == 0) diag += ax[idx];
        }
        if (fabs(diag) < 1e-300) diag = 1e-14;
        out->l_values = orc_xmalloc(8);
        out->u_values = (double *)orc_xmalloc(sizeof(double)); out->u_values[0] = diag;
        out->u_diag = (double *)orc_xmalloc(sizeof(double)); out->u_diag[0] = diag;
        out->row_perm = (int64_t *)orc_xcalloc(1, sizeof(int64_t));
        out->row_perm_inv = (int64_t *)orc_xcalloc(1, sizeof(int64_t));
        out->nu = 1;
        return;
    }
    int64_t nl = sym->l_col_ptr[n], nu = sym->u_col_ptr[n];
    out->nl = nl; out->nu = nu;
    double *lv = (double *)orc_xcalloc((size_t)(nl ? nl : 1), sizeof(double));
    double *uv = (double *)orc_xcalloc((size_t)nu, sizeof(double));
    double *ud = (double *)orc_xcalloc((size_t)n, sizeof(double));
    int64_t *rp = (int64_t *)orc_xmalloc((size_t)n * sizeof(int64_t));
    int64_t *rpi = (int64_t *)orc_xmalloc((size_t)n * sizeof(int64_t));
    int use_old = refactor && old != NULL;
    if (use_old) {
        memcpy(rp, old->row_perm, (size_t)n * sizeof(int64_t));
        memcpy(rpi, old->row_perm_inv, (size_t)n * sizeof(int64_t));
    } else {
        for (int64_t i = 0; i < n; i++) rp[i] = rpi[i] = i;
    }
    double *work = (double *)orc_xcalloc((size_t)n, sizeof(double));
    for (int64_t k = 0; k < n; k++) {
        int64_t oc = sym->col_perm_inv[k];
        int64_t st = ap[oc], en = ap[oc + 1];
        for (int64_t idx = st; idx < en; idx++) {
            int64_t orow = ai[idx];
            if (orow < n) work[rpi[sym->col_perm[orow]]] += ax[idx];
        }
        int64_t us = sym->u_col_ptr[k], ue = sym->u_col_ptr[k + 1] - 1;
        for (int64_t ui = us; ui < ue; ui++) {
            int64_t j = sym->u_row_idx[ui];
            double ujk = work[j];
            if (ujk != 0.0) {
                for (int64_t li = sym->l_col_ptr[j]; li < sym->l_col_ptr[j + 1]; li++)
                    work[sym->l_row_idx[li]] -= lv[li] * ujk;
            }
        }
        if (!use_old) {
            int64_t ls = sym->l_col_ptr[k], le = sym->l_col_ptr[k + 1];
            double dabs = fabs(work[k]), maxabs = dabs;
            int64_t maxrow = k;
            for (int64_t li = ls; li < le; li++) {
                int64_t row = sym->l_row_idx[li];
                double v = fabs(work[row]);
                if (v > maxabs) { maxabs = v; maxrow = row; }
            }
            if (maxrow != k && dabs < s->pivot_tol * maxabs) {
                double t = work[k]; work[k] = work[maxrow]; work[maxrow] = t;
                int64_t ok = rp[k], om = rp[maxrow];
                rp[k] = om; rp[maxrow] = ok;
                rpi[ok] = maxrow; rpi[om] = k;
            }
        }
        for (int64_t ui = us; ui < ue; ui++) uv[ui] = work[sym->u_row_idx[ui]];
        double diag = work[k];
        if (fabs(diag) < 1e-14) diag = diag >= 0.0 ? 1e-14 : -1e-14;
        uv[sym->u_col_ptr[k + 1] - 1] = diag;
        ud[k] = diag;
        int64_t ls = sym->l_col_ptr[k], le = sym->l_col_ptr[k + 1];
        for (int64_t li = ls; li < le; li++) lv[li] = work[sym->l_row_idx[li]] / diag;
        for (int64_t idx = st; idx < en; idx++) {
            int64_t orow = ai[idx];
            if (orow < n) work[rpi[sym->col_perm[orow]]] = 0.0;
        }
        for (int64_t li = ls; li < le; li++) work[sym->l_row_idx[li]] = 0.0;
        for (int64_t ui = us; ui < ue; ui++) work[sym->u_row_idx[ui]] = 0.0;
        work[k] = 0.0;
    }
    free(work);
    out->l_values = lv; out->u_values = uv; out->u_diag = ud;
    out->row_perm = rp; out->row_perm_inv = rpi;
}

static void compute_rcond(orc_klu *s) { /* native_klu.rs:1406-1427 */
    double mn = 1.7976931348623157e308, mx = 0.0;
    for (int64_t b = 0; b < s->nnum; b++)
        for (int64_t i = 0; i < s->num[b].n; i++) {
            double a = fabs(s->num[b].u_diag[i]);
            if (a < mn) mn = a;
            if (a > mx) mx = a;
        }
    s->rcond = mx > 0.0 ? mn / mx : 0.0;
}

/* do_factor (native_klu.rs:1020-1148), sequential build */
static int do_factor(orc_klu *s, int64_t apl, const int64_t *ap, int64_t ail, const int64_t *ai,
                     const double *ax) {
    int pattern_same = s->state == ST_FACTORED && pattern_matches(s, apl, ap, ail, ai);
    numf *nn = (numf *)orc_xcalloc((size_t)s->nblk, sizeof(numf));
    if (s->has_btf) {
        for (int64_t k = 0; k < s->nblk; k++) {
            int64_t bs = s->btf.block_ptr[k], be = s->btf.block_ptr[k + 1];
            int64_t *bap, *bai; double *bax;
            extract_block(s, ap, ai, ax, bs, be, &bap, &bai, &bax);
            const numf *old = (pattern_same && k < s->nnum) ? &s->num[k] : NULL;
            factor_block(s, &s->sym[k], bap, bai, bax, old, pattern_same, &nn[k]);
            free(bap); free(bai); free(bax);
        }
    } else {
        const numf *old = (pattern_same && s->nnum > 0) ? &s->num[0] : NULL;
        factor_block(s, &s->sym[0], ap, ai, ax, old, pattern_same, &nn[0]);
    }
    for (int64_t b = 0; b < s->nnum; b++) numf_free(&s->num[b]);
    free(s->num);
    s->num = nn; s->nnum = s->nblk;
    if (s->has_btf) {
        int64_t n = s->n;
        for (int64_t q = 0; q < s->noff; q++) {
            offd *o = &s->off[q];
            int64_t rs = o->row_start, re = rs + o->nrows, cnt = 0;
            for (int64_t lc = 0; lc < o->ncols; lc++) {
                int64_t gco = s->btf.col_perm[o->col_start + lc];
                for (int64_t idx = ap[gco]; idx < ap[gco + 1]; idx++) {
                    int64_t gro = ai[idx];
                    if (gro >= n) continue;
                    int64_t grn = s->btf.row_perm_inv[gro];
                    if (grn >= rs && grn < re) o->values[cnt++] = ax[idx];
                }
            }
        }
    }
    if (pattern_same && s->state == ST_FACTORED) s->refactor_count++;
    else s->factor_count++;
    compute_rcond(s);
    s->state = ST_FACTORED;
    return 0;
}

int orc_klu_analyze(orc_klu *s, int64_t apl, const int64_t *ap, int64_t ail, const int64_t *ai) {
    if (s->state != ST_EMPTY && pattern_matches(s, apl, ap, ail, ai)) return 0;
    if (apl != s->n + 1) return 6; /* InvalidMatrix */
    return do_analyze(s, apl, ap, ail, ai);
}

int orc_klu_factor(orc_klu *s, int64_t apl, const int64_t *ap, int64_t ail, const int64_t *ai,
                   const double *ax) {
    if (s->state == ST_EMPTY || !pattern_matches(s, apl, ap, ail, ai)) {
        int r = do_analyze(s, apl, ap, ail, ai);
        if (r) return r;
    }
    return do_factor(s, apl, ap, ail, ai, ax);
}

static void solve_block_internal(const symf *sym, const numf *num, double *rhs) {
    int64_t n = sym->n;
    double *temp = (double *)orc_xcalloc((size_t)(n ? n : 1), sizeof(double));
    for (int64_t i = 0; i < n; i++) temp[num->row_perm_inv[sym->col_perm[i]]] = rhs[i];
    for (int64_t k = 0; k < n; k++) {
        double zk = temp[k];
        for (int64_t li = sym->l_col_ptr[k]; li < sym->l_col_ptr[k + 1]; li++)
            temp[sym->l_row_idx[li]] -= num->l_values[li] * zk;
    }
    for (int64_t k = n - 1; k >= 0; k--) {
        temp[k] /= num->u_diag[k];
        double xk = temp[k];
        for (int64_t ui = sym->u_col_ptr[k]; ui < sym->u_col_ptr[k + 1] - 1; ui++)
            temp[sym->u_row_idx[ui]] -= num->u_values[ui] * xk;
    }
    for (int64_t i = 0; i < n; i++) rhs[i] = temp[sym->col_perm[i]];
    free(temp);
}

int orc_klu_solve(orc_klu *s, int64_t n, double *rhs) { /* native_klu.rs:1643-1657 */
    if (s->state != ST_FACTORED) return 3; /* SolveFailed */
    if (n != s->n) return 6;
    if (!s->has_btf) {
        solve_block_internal(&s->sym[0], &s->num[0], rhs);
        return 0;
    }
    /* solve_btf (native_klu.rs:1491-1551) */
    double *w = s->work;
    for (int64_t r = 0; r < n; r++) w[s->btf.row_perm_inv[r]] = rhs[r];
    for (int64_t k = s->btf.num_blocks - 1; k >= 0; k--) {
        int64_t bs = s->btf.block_ptr[k], be = s->btf.block_ptr[k + 1], sz = be - bs;
        for (int64_t q = 0; q < s->noff; q++) {
            const offd *o = &s->off[q];
            if (o->block_row != k) continue;
            for (int64_t lc = 0; lc < o->ncols; lc++) {
                double xj = w[o->col_start + lc];
                for (int64_t idx = o->col_ptr[lc]; idx < o->col_ptr[lc + 1]; idx++)
                    w[o->row_start + o->row_idx[idx]] -= o->values[idx] * xj;
            }
        }
        if (sz == 1) w[bs] /= s->num[k].u_diag[0];
        else solve_block_internal(&s->sym[k], &s->num[k], w + bs);
    }
    for (int64_t c = 0; c < n; c++) rhs[c] = w[s->btf.col_perm_inv[c]];
    return 0;
}

void orc_klu_stats(const orc_klu *s, int64_t out[6], double *rcond) {
    out[0] = s->factor_count; out[1] = s->refactor_count; out[2] = s->nnz_l;
    out[3] = s->nnz_u; out[4] = s->nblocks; out[5] = s->noffdiag;
    if (rcond) *rcond = s->rcond;
}

int64_t orc_klu_block_n(const orc_klu *s, int64_t b) {
    return (b >= 0 && b < s->nblk) ? s->sym[b].n : -1;
}
void orc_klu_block_sizes(const orc_klu *s, int64_t b, int64_t *nl, int64_t *nu) {
    *nl = s->sym[b].l_col_ptr[s->sym[b].n];
    *nu = s->sym[b].u_col_ptr[s->sym[b].n];
}
void orc_klu_block_get(const orc_klu *s, int64_t b, int64_t *col_perm, int64_t *row_perm,
                       int64_t *l_col_ptr, int64_t *l_row_idx, int64_t *u_col_ptr,
                       int64_t *u_row_idx, double *l_values, double *u_values, double *u_diag) {
    const symf *y = &s->sym[b];
    int64_t n = y->n, nl = y->l_col_ptr[n], nu = y->u_col_ptr[n];
    if (col_perm) memcpy(col_perm, y->col_perm, (size_t)n * 8);
    if (l_col_ptr) memcpy(l_col_ptr, y->l_col_ptr, (size_t)(n + 1) * 8);
    if (l_row_idx) memcpy(l_row_idx, y->l_row_idx, (size_t)nl * 8);
    if (u_col_ptr) memcpy(u_col_ptr, y->u_col_ptr, (size_t)(n + 1) * 8);
    if (u_row_idx) memcpy(u_row_idx, y->u_row_idx, (size_t)nu * 8);
    if (b < s->nnum) {
        const numf *f = &s->num[b];
        if (row_perm) memcpy(row_perm, f->row_perm, (size_t)n * 8);
        if (l_values) memcpy(l_values, f->l_values, (size_t)nl * 8);
        if (u_values) memcpy(u_values, f->u_values, (size_t)nu * 8);
        if (u_diag) memcpy(u_diag, f->u_diag, (size_t)n * 8);
    }
}
int orc_klu_btf_get(const orc_klu *s, int64_t *row_perm, int64_t *col_perm, int64_t *block_ptr) {
    if (!s->has_btf) return 0;
    memcpy(row_perm, s->btf.row_perm, (size_t)s->n * 8);
    memcpy(col_perm, s->btf.col_perm, (size_t)s->n * 8);
    memcpy(block_ptr, s->btf.block_ptr, (size_t)(s->btf.num_blocks + 1) * 8);
    return (int)s->btf.num_blocks;
}
