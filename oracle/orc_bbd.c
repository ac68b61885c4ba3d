/*
 * orc_bbd.c - TEST INFRASTRUCTURE ONLY (see orc.h). Restatement of
 *   crates/sim-core/src/bbd.rs        (GreedyBisectionPartitioner, bfs_bisect,
 *                                      bbd_decompose)
 *   crates/sim-core/src/bbd_solver.rs (BbdSolver, DenseSchurSolver, Schur
 *                                      pattern, factor_bbd, solve_bbd)
 *   crates/sim-core/src/solver.rs     (DenseSolver, :595-724)
 * with the reference's loop orders, tie-breaks and quirks (R_k scatter by
 * assignment, zero-skip in the sparse products, 1e-30 singular checks).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"
#include "orc_util.h"

#define BORDER ((int64_t)-1)

/* ---- symmetric adjacency without self loops (bbd.rs:297-321) ------------ */
typedef struct {
    int64_t *ptr, *idx;
} adjl;

static int cmp_i64(const void *a, const void *b) {
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return x < y ? -1 : x > y;
}

static adjl build_adj(int64_t n, const int64_t *ap, const int64_t *ai) {
    int64_t *cnt = (int64_t *)orc_xcalloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t c = 0; c < n; c++)
        for (int64_t q = ap[c]; q < ap[c + 1]; q++) {
            int64_t r = ai[q];
            if (r < n && r != c) { cnt[c]++; cnt[r]++; }
        }
    adjl a;
    a.ptr = (int64_t *)orc_xcalloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) a.ptr[i + 1] = a.ptr[i] + cnt[i];
    a.idx = (int64_t *)orc_xmalloc((size_t)(a.ptr[n] ? a.ptr[n] : 1) * sizeof(int64_t));
    int64_t *fill = (int64_t *)orc_xcalloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t c = 0; c < n; c++)
        for (int64_t q = ap[c]; q < ap[c + 1]; q++) {
            int64_t r = ai[q];
            if (r < n && r != c) {
                a.idx[a.ptr[c] + fill[c]++] = r;
                a.idx[a.ptr[r] + fill[r]++] = c;
            }
        }
    /* sort + dedup, compacting in place */
    int64_t w = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t s = a.ptr[i], e = a.ptr[i + 1];
        qsort(a.idx + s, (size_t)(e - s), sizeof(int64_t), cmp_i64);
        int64_t start = w;
        for (int64_t q = s; q < e; q++)
            if (q == s || a.idx[q] != a.idx[q - 1]) a.idx[w++] = a.idx[q];
        a.ptr[i] = start;
    }
    /* rebuild ptr: entries of node i are [ptr[i], ptr[i+1]) after compaction */
    int64_t *np = (int64_t *)orc_xmalloc(((size_t)n + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) np[i] = a.ptr[i];
    np[n] = w;
    free(a.ptr);
    a.ptr = np;
    free(cnt);
    free(fill);
    return a;
}

/* ---- bfs_bisect (bbd.rs:185-295): group per node, 0=A 1=B 2=Cut -------- */
enum { GA = 0, GB = 1, GCUT = 2, GNONE = 3 };

static void bfs_bisect(const int64_t *nodes, int64_t m, const adjl *adj, int64_t n, uint8_t *grp,
                       uint8_t *inset) {
    /* grp[] is GNONE outside the node set on entry and on exit */
    if (m == 0) return;
    for (int64_t i = 0; i < m; i++) inset[nodes[i]] = 1;
    if (m <= 3) {
        for (int64_t i = 0; i < m; i++) grp[nodes[i]] = GA;
        for (int64_t i = 0; i < m; i++) inset[nodes[i]] = 0;
        return;
    }
    /* min_by_key: the first node of minimum in-set degree */
    int64_t start = nodes[0], best = -1;
    for (int64_t i = 0; i < m; i++) {
        int64_t v = nodes[i], d = 0;
        for (int64_t q = adj->ptr[v]; q < adj->ptr[v + 1]; q++) d += inset[adj->idx[q]];
        if (best < 0 || d < best) { best = d; start = v; }
    }
    /* BFS level sets */
    int64_t *order = (int64_t *)orc_xmalloc((size_t)m * sizeof(int64_t));
    int64_t *lvl_start = (int64_t *)orc_xmalloc(((size_t)m + 2) * sizeof(int64_t));
    uint8_t *vis = (uint8_t *)orc_xcalloc((size_t)n, 1);
    int64_t head = 0, tail = 0, nl = 0;
    order[tail++] = start;
    vis[start] = 1;
    while (head < tail) {
        int64_t lsz = tail - head;
        lvl_start[nl++] = head;
        for (int64_t t = 0; t < lsz; t++) {
            int64_t v = order[head++];
            for (int64_t q = adj->ptr[v]; q < adj->ptr[v + 1]; q++) {
                int64_t u = adj->idx[q];
                if (inset[u] && !vis[u]) { vis[u] = 1; order[tail++] = u; }
            }
        }
    }
    /* disconnected nodes: singleton levels in node order */
    for (int64_t i = 0; i < m; i++)
        if (!vis[nodes[i]]) { lvl_start[nl++] = tail; order[tail++] = nodes[i]; vis[nodes[i]] = 1; }
    lvl_start[nl] = tail;
    int64_t half = m / 2, count = 0, split = nl;
    for (int64_t l = 0; l < nl; l++) {
        count += lvl_start[l + 1] - lvl_start[l];
        if (count >= half) { split = l + 1; break; }
    }
    for (int64_t l = 0; l < nl; l++)
        for (int64_t q = lvl_start[l]; q < lvl_start[l + 1]; q++) grp[order[q]] = l < split ? GA : GB;
    /* cut: nodes with a neighbour (in the set) of the other group */
    uint8_t *cut = (uint8_t *)orc_xcalloc((size_t)m, 1);
    for (int64_t i = 0; i < m; i++) {
        int64_t v = nodes[i];
        for (int64_t q = adj->ptr[v]; q < adj->ptr[v + 1]; q++) {
            int64_t u = adj->idx[q];
            if (inset[u] && grp[u] != grp[v]) { cut[i] = 1; break; }
        }
    }
    for (int64_t i = 0; i < m; i++)
        if (cut[i]) grp[nodes[i]] = GCUT;
    for (int64_t i = 0; i < m; i++) inset[nodes[i]] = 0;
    free(cut); free(order); free(lvl_start); free(vis);
}

/* GreedyBisectionPartitioner::partition (bbd.rs:103-171) */
int64_t orc_bbd_partition(int64_t n, const int64_t *ap, const int64_t *ai, int64_t num_blocks,
                          int64_t *assign) {
    if (n == 0 || num_blocks <= 1) {
        for (int64_t i = 0; i < n; i++) assign[i] = 0;
        return 1;
    }
    adjl adj = build_adj(n, ap, ai);
    for (int64_t i = 0; i < n; i++) assign[i] = 0;
    int64_t cur = 1;
    int64_t *sizes = (int64_t *)orc_xcalloc((size_t)num_blocks + 1, sizeof(int64_t));
    int64_t *nodes = (int64_t *)orc_xmalloc((size_t)n * sizeof(int64_t));
    uint8_t *grp = (uint8_t *)orc_xmalloc((size_t)n);
    uint8_t *inset = (uint8_t *)orc_xcalloc((size_t)n, 1);
    memset(grp, GNONE, (size_t)n);
    while (cur < num_blocks) {
        for (int64_t b = 0; b < cur; b++) sizes[b] = 0;
        for (int64_t i = 0; i < n; i++)
            if (assign[i] != BORDER && assign[i] < cur) sizes[assign[i]]++;
        /* max_by_key: the last block of maximum size */
        int64_t tgt = 0;
        for (int64_t b = 0; b < cur; b++)
            if (sizes[b] >= sizes[tgt]) tgt = b;
        if (sizes[tgt] < 4) break;
        int64_t m = 0;
        for (int64_t i = 0; i < n; i++)
            if (assign[i] == tgt) nodes[m++] = i;
        bfs_bisect(nodes, m, &adj, n, grp, inset);
        for (int64_t i = 0; i < m; i++) {
            int64_t v = nodes[i];
            if (grp[v] == GB) assign[v] = cur;
            else if (grp[v] == GCUT) assign[v] = BORDER;
            grp[v] = GNONE;
        }
        cur++;
    }
    free(sizes); free(nodes); free(grp); free(inset); free(adj.ptr); free(adj.idx);
    return cur;
}

/* bbd_decompose (bbd.rs:355-485) */
void orc_bbd_decompose(int64_t n, const int64_t *ap, const int64_t *ai, int64_t num_blocks,
                       orc_bbd_decomp *d) {
    memset(d, 0, sizeof(*d));
    d->n = n;
    int64_t *assign = (int64_t *)orc_xmalloc(((size_t)n + 1) * sizeof(int64_t));
    int64_t nb = orc_bbd_partition(n, ap, ai, num_blocks, assign);
    for (int64_t c = 0; c < n; c++) { /* promote columns touching another block */
        if (assign[c] == BORDER) continue;
        int64_t mine = assign[c];
        for (int64_t q = ap[c]; q < ap[c + 1]; q++) {
            int64_t r = ai[q];
            if (r < n && r != c) {
                int64_t rb = assign[r];
                if (rb != BORDER && rb != mine) { assign[c] = BORDER; break; }
            }
        }
    }
    for (int64_t c = 0; c < n; c++) /* promote rows touched from another block */
        for (int64_t q = ap[c]; q < ap[c + 1]; q++) {
            int64_t r = ai[q];
            if (r < n && r != c) {
                if (assign[r] == BORDER) continue;
                int64_t cb = assign[c];
                if (cb != BORDER && cb != assign[r]) assign[r] = BORDER;
            }
        }
    int64_t *bsz = (int64_t *)orc_xcalloc((size_t)nb + 1, sizeof(int64_t));
    int64_t nborder = 0;
    for (int64_t i = 0; i < n; i++) {
        if (assign[i] == BORDER) nborder++;
        else bsz[assign[i]]++;
    }
    int64_t nonempty = 0;
    for (int64_t b = 0; b < nb; b++) nonempty += bsz[b] > 0;
    d->perm = (int64_t *)orc_xmalloc(((size_t)n + 1) * sizeof(int64_t));
    d->inv_perm = (int64_t *)orc_xmalloc(((size_t)n + 1) * sizeof(int64_t));
    if (nonempty <= 1) { /* trivial: one block, no border */
        int64_t w = 0;
        if (nonempty == 1) {
            for (int64_t b = 0; b < nb; b++)
                if (bsz[b])
                    for (int64_t i = 0; i < n; i++)
                        if (assign[i] == b) d->perm[w++] = i;
            for (int64_t i = 0; i < n; i++)
                if (assign[i] == BORDER) d->perm[w++] = i;
        } else {
            for (int64_t i = 0; i < n; i++) d->perm[w++] = i;
        }
        d->num_blocks = 1;
        d->block_ptr = (int64_t *)orc_xmalloc(2 * sizeof(int64_t));
        d->block_ptr[0] = 0;
        d->block_ptr[1] = n;
        d->border_start = n;
        d->border_size = 0;
    } else {
        d->num_blocks = nonempty;
        d->block_ptr = (int64_t *)orc_xmalloc(((size_t)nonempty + 1) * sizeof(int64_t));
        int64_t w = 0, k = 0;
        for (int64_t b = 0; b < nb; b++) {
            if (!bsz[b]) continue;
            d->block_ptr[k++] = w;
            for (int64_t i = 0; i < n; i++)
                if (assign[i] == b) d->perm[w++] = i;
        }
        d->block_ptr[k] = w;
        d->border_start = w;
        for (int64_t i = 0; i < n; i++)
            if (assign[i] == BORDER) d->perm[w++] = i;
        d->border_size = n - d->border_start;
    }
    for (int64_t i = 0; i < n; i++) d->inv_perm[d->perm[i]] = i;
    free(assign); free(bsz);
}

void orc_bbd_decomp_free(orc_bbd_decomp *d) {
    free(d->perm); free(d->inv_perm); free(d->block_ptr);
    memset(d, 0, sizeof(*d));
}

/* ---- dense LU with partial pivoting (solver.rs:640-672, bbd_solver.rs:112-156)
 * tiny: 0 -> DenseSolver (max == 0 fails with FactorFailed),
 *       1 -> DenseSchurSolver (max < 1e-30 fails with SingularMatrix{k}) */
int orc_dense_factor(int64_t n, double *lu, int64_t *piv, int schur) {
    for (int64_t k = 0; k < n; k++) piv[k] = k;
    for (int64_t k = 0; k < n; k++) {
        int64_t p = k;
        double mx = fabs(lu[k * n + k]);
        for (int64_t i = k + 1; i < n; i++) {
            double v = fabs(lu[i * n + k]);
            if (v > mx) { mx = v; p = i; }
        }
        if (schur ? (mx < 1e-30) : (mx == 0.0)) return schur ? 4 : 2;
        if (p != k) {
            for (int64_t j = 0; j < n; j++) {
                double t = lu[k * n + j]; lu[k * n + j] = lu[p * n + j]; lu[p * n + j] = t;
            }
            int64_t t = piv[k]; piv[k] = piv[p]; piv[p] = t;
        }
        double pv = lu[k * n + k];
        for (int64_t i = k + 1; i < n; i++) {
            double f = lu[i * n + k] / pv;
            lu[i * n + k] = f;
            for (int64_t j = k + 1; j < n; j++) lu[i * n + j] -= f * lu[k * n + j];
        }
    }
    return 0;
}

/* solve (solver.rs:698-721, bbd_solver.rs:159-196) */
int orc_dense_solve(int64_t n, const double *lu, const int64_t *piv, double *rhs, int schur) {
    if (n == 0) return 0;
    double *b = (double *)orc_xmalloc((size_t)n * sizeof(double));
    for (int64_t i = 0; i < n; i++) b[i] = rhs[piv[i]];
    for (int64_t i = 0; i < n; i++) {
        double s = b[i];
        for (int64_t j = 0; j < i; j++) s -= lu[i * n + j] * b[j];
        b[i] = s;
    }
    for (int64_t i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int64_t j = i + 1; j < n; j++) s -= lu[i * n + j] * rhs[j];
        double dg = lu[i * n + i];
        if (schur ? (fabs(dg) < 1e-30) : (dg == 0.0)) { free(b); return 3; }
        rhs[i] = s / dg;
    }
    free(b);
    return 0;
}

/* DenseSolver::build_dense (solver.rs:619-637): row-major, += in CSC order */
int orc_dense_build(int64_t n, int64_t apl, const int64_t *ap, const int64_t *ai, const double *ax,
                    double *lu) {
    if (apl != n + 1) return 1; /* AnalyzeFailed */
    memset(lu, 0, (size_t)(n * n) * sizeof(double));
    for (int64_t c = 0; c < n; c++)
        for (int64_t q = ap[c]; q < ap[c + 1]; q++) {
            int64_t r = ai[q];
            if (r < n) lu[r * n + c] += ax[q];
        }
    return 0;
}

/* ---- BbdSolver (bbd_solver.rs:231-1094) ---------------------------------- */
typedef struct {
    int64_t ncols, nnz;
    int64_t *col_ptr, *row_idx;
    double *values;
} sblk;

struct orc_bbd {
    uint32_t magic; /* ORC_MAGIC_BBD: the engine dispatches on it */
    int64_t n, target;
    int analyzed, factored, has_decomp;
    orc_bbd_decomp d;
    orc_klu **blk;
    sblk *C, *R;
    int dense;              /* Schur kind: 1 dense, 0 sparse */
    double *bt, *S, *lu;    /* dense path */
    int64_t *piv;
    int dense_ok;
    orc_klu *ssolver;       /* sparse path */
    int64_t *s_ap, *s_ai;
    double *s_ax;
    orc_klu *fallback;
    int64_t *last_ap, *last_ai, last_apl, last_ail;
    int64_t factor_count;
};

static void sblk_free(sblk *b) { free(b->col_ptr); free(b->row_idx); free(b->values); }

static void bbd_clear(orc_bbd *s) {
    if (s->has_decomp) {
        for (int64_t k = 0; k < s->d.num_blocks; k++) {
            orc_klu_free(s->blk[k]);
            sblk_free(&s->C[k]);
            sblk_free(&s->R[k]);
        }
        free(s->blk); free(s->C); free(s->R);
        orc_bbd_decomp_free(&s->d);
    }
    s->blk = NULL; s->C = s->R = NULL;
    free(s->bt); free(s->S); free(s->lu); free(s->piv);
    s->bt = s->S = s->lu = NULL; s->piv = NULL;
    if (s->ssolver) orc_klu_free(s->ssolver);
    s->ssolver = NULL;
    free(s->s_ap); free(s->s_ai); free(s->s_ax);
    s->s_ap = s->s_ai = NULL; s->s_ax = NULL;
    if (s->fallback) orc_klu_free(s->fallback);
    s->fallback = NULL;
    s->has_decomp = 0;
    s->dense_ok = 0;
}

orc_bbd *orc_bbd_new(int64_t n, int64_t num_blocks) {
    orc_bbd *s = (orc_bbd *)orc_xcalloc(1, sizeof(orc_bbd));
    s->magic = ORC_MAGIC_BBD;
    s->n = n;
    s->target = num_blocks > 2 ? num_blocks : 2;
    return s;
}

void orc_bbd_reset_pattern(orc_bbd *s) { /* :1079-1089 */
    s->analyzed = 0; s->factored = 0;
    free(s->last_ap); free(s->last_ai);
    s->last_ap = s->last_ai = NULL; s->last_apl = s->last_ail = 0;
    bbd_clear(s);
}

void orc_bbd_free(orc_bbd *s) {
    if (!s) return;
    orc_bbd_reset_pattern(s);
    free(s);
}

void orc_bbd_prepare(orc_bbd *s, int64_t n) { /* :1000-1006 */
    if (n != s->n) { orc_bbd_reset_pattern(s); s->n = n; }
}

static orc_klu *new_sparselu(int64_t n) {
    orc_klu *k = orc_klu_new(n);
    orc_klu_set_sparselu(k);
    orc_klu_prepare(k, n);
    return k;
}

/* entries of permuted column nc, as (new_row, CSC index) in CSC order */
#define FOR_COL(nc, body)                                              \
    do {                                                               \
        int64_t oc_ = s->d.perm[(nc)];                                 \
        for (int64_t q = ap[oc_]; q < ap[oc_ + 1]; q++) {              \
            int64_t orow_ = ai[q];                                     \
            if (orow_ >= n) continue;                                  \
            int64_t nr = s->d.inv_perm[orow_];                         \
            body                                                       \
        }                                                              \
    } while (0)

static int bbd_setup_fallback(orc_bbd *s, int64_t apl, const int64_t *ap, int64_t ail,
                              const int64_t *ai) { /* :492-499 */
    bbd_clear(s);
    s->fallback = new_sparselu(s->n);
    return orc_klu_analyze(s->fallback, apl, ap, ail, ai);
}

/* analyze_bbd + extract_block_patterns + compute_schur_pattern (:330-623) */
static int bbd_analyze(orc_bbd *s, int64_t apl, const int64_t *ap, int64_t ail, const int64_t *ai) {
    const int64_t n = s->n;
    bbd_clear(s);
    orc_bbd_decomp d;
    orc_bbd_decompose(n, ap, ai, s->target, &d);
    if (d.num_blocks <= 1 || d.border_size == 0 || (double)d.border_size / (double)n > 0.5) {
        orc_bbd_decomp_free(&d);
        return bbd_setup_fallback(s, apl, ap, ail, ai);
    }
    s->d = d;
    s->has_decomp = 1;
    const int64_t nb = d.num_blocks, bs = d.border_size, b0 = d.border_start;
    s->blk = (orc_klu **)orc_xcalloc((size_t)nb, sizeof(orc_klu *));
    s->C = (sblk *)orc_xcalloc((size_t)nb, sizeof(sblk));
    s->R = (sblk *)orc_xcalloc((size_t)nb, sizeof(sblk));
    int rc = 0;
    for (int64_t k = 0; k < nb && !rc; k++) {
        const int64_t st = d.block_ptr[k], en = d.block_ptr[k + 1], sz = en - st;
        s->blk[k] = new_sparselu(sz);
        int64_t *bap = (int64_t *)orc_xcalloc((size_t)sz + 1, sizeof(int64_t));
        int64_t *bai = (int64_t *)orc_xmalloc((size_t)(ap[n] + 1) * sizeof(int64_t));
        int64_t cnt = 0;
        for (int64_t lc = 0; lc < sz; lc++) {
            FOR_COL(st + lc, if (nr >= st && nr < en) bai[cnt++] = nr - st;);
            bap[lc + 1] = cnt;
        }
        rc = orc_klu_analyze(s->blk[k], sz + 1, bap, cnt, bai);
        free(bap); free(bai);
        /* C_k: border rows x block cols, R_k: block rows x border cols */
        sblk *c = &s->C[k], *r = &s->R[k];
        c->ncols = sz;
        c->col_ptr = (int64_t *)orc_xcalloc((size_t)sz + 1, sizeof(int64_t));
        c->row_idx = (int64_t *)orc_xmalloc((size_t)(ap[n] + 1) * sizeof(int64_t));
        for (int64_t lc = 0; lc < sz; lc++) {
            FOR_COL(st + lc, if (nr >= b0) c->row_idx[c->nnz++] = nr - b0;);
            c->col_ptr[lc + 1] = c->nnz;
        }
        r->ncols = bs;
        r->col_ptr = (int64_t *)orc_xcalloc((size_t)bs + 1, sizeof(int64_t));
        r->row_idx = (int64_t *)orc_xmalloc((size_t)(ap[n] + 1) * sizeof(int64_t));
        for (int64_t lc = 0; lc < bs; lc++) {
            FOR_COL(b0 + lc, if (nr >= st && nr < en) r->row_idx[r->nnz++] = nr - st;);
            r->col_ptr[lc + 1] = r->nnz;
        }
        c->values = (double *)orc_xmalloc((size_t)(c->nnz + 1) * sizeof(double));
        r->values = (double *)orc_xmalloc((size_t)(r->nnz + 1) * sizeof(double));
    }
    if (rc) return rc;
    if (bs <= ORC_DENSE_SCHUR_MAX) {
        s->dense = 1;
        s->bt = (double *)orc_xcalloc((size_t)(bs * bs), sizeof(double));
        s->S = (double *)orc_xcalloc((size_t)(bs * bs), sizeof(double));
        s->lu = (double *)orc_xcalloc((size_t)(bs * bs), sizeof(double));
        s->piv = (int64_t *)orc_xcalloc((size_t)bs, sizeof(int64_t));
    } else { /* compute_schur_pattern (:330-408) */
        s->dense = 0;
        uint8_t *crow = (uint8_t *)orc_xcalloc((size_t)(nb * bs), 1);
        uint8_t *rcol = (uint8_t *)orc_xcalloc((size_t)(nb * bs), 1);
        for (int64_t k = 0; k < nb; k++) {
            for (int64_t q = 0; q < s->C[k].nnz; q++) crow[k * bs + s->C[k].row_idx[q]] = 1;
            for (int64_t j = 0; j < bs; j++)
                rcol[k * bs + j] = s->R[k].col_ptr[j] < s->R[k].col_ptr[j + 1];
        }
        s->s_ap = (int64_t *)orc_xcalloc((size_t)bs + 1, sizeof(int64_t));
        int64_t cap = bs * 4 + 16, nz = 0;
        s->s_ai = (int64_t *)orc_xmalloc((size_t)cap * sizeof(int64_t));
        uint8_t *mark = (uint8_t *)orc_xmalloc((size_t)bs);
        for (int64_t j = 0; j < bs; j++) {
            memset(mark, 0, (size_t)bs);
            FOR_COL(b0 + j, if (nr >= b0) mark[nr - b0] = 1;);
            for (int64_t k = 0; k < nb; k++)
                if (rcol[k * bs + j])
                    for (int64_t i = 0; i < bs; i++)
                        if (crow[k * bs + i]) mark[i] = 1;
            for (int64_t i = 0; i < bs; i++)
                if (mark[i]) {
                    if (nz == cap) { cap *= 2; s->s_ai = (int64_t *)realloc(s->s_ai, (size_t)cap * sizeof(int64_t)); }
                    s->s_ai[nz++] = i;
                }
            s->s_ap[j + 1] = nz;
        }
        s->s_ax = (double *)orc_xcalloc((size_t)nz + 1, sizeof(double));
        s->ssolver = new_sparselu(bs);
        rc = orc_klu_analyze(s->ssolver, bs + 1, s->s_ap, nz, s->s_ai);
        free(crow); free(rcol); free(mark);
    }
    return rc;
}

int orc_bbd_analyze(orc_bbd *s, int64_t apl, const int64_t *ap, int64_t ail, const int64_t *ai) {
    if (s->analyzed && apl == s->last_apl && ail == s->last_ail &&
        !memcmp(ap, s->last_ap, (size_t)apl * 8) && !memcmp(ai, s->last_ai, (size_t)ail * 8))
        return 0;
    if (apl != s->n + 1) return 6;
    s->analyzed = 0;
    s->factored = 0;
    int rc = bbd_analyze(s, apl, ap, ail, ai);
    if (rc) return rc;
    free(s->last_ap); free(s->last_ai);
    s->last_ap = (int64_t *)orc_xmalloc((size_t)(apl + 1) * 8);
    s->last_ai = (int64_t *)orc_xmalloc((size_t)(ail + 1) * 8);
    memcpy(s->last_ap, ap, (size_t)apl * 8);
    memcpy(s->last_ai, ai, (size_t)ail * 8);
    s->last_apl = apl; s->last_ail = ail;
    s->analyzed = 1;
    return 0;
}

static int64_t bsearch_row(const int64_t *rows, int64_t m, int64_t key) {
    int64_t lo = 0, hi = m;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (rows[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < m && rows[lo] == key) ? lo : -1;
}

/* factor_bbd (:626-904) */
static int bbd_factor(orc_bbd *s, const int64_t *ap, const int64_t *ai, const double *ax) {
    const int64_t n = s->n, nb = s->d.num_blocks, bs = s->d.border_size, b0 = s->d.border_start;
    int rc = 0;
    int64_t maxb = 0;
    for (int64_t k = 0; k < nb; k++) {
        const int64_t st = s->d.block_ptr[k], en = s->d.block_ptr[k + 1], sz = en - st;
        if (sz > maxb) maxb = sz;
        int64_t *bap = (int64_t *)orc_xcalloc((size_t)sz + 1, sizeof(int64_t));
        int64_t *bai = (int64_t *)orc_xmalloc((size_t)(ap[n] + 1) * sizeof(int64_t));
        double *bax = (double *)orc_xmalloc((size_t)(ap[n] + 1) * sizeof(double));
        int64_t cnt = 0;
        for (int64_t lc = 0; lc < sz; lc++) {
            FOR_COL(st + lc, if (nr >= st && nr < en) { bai[cnt] = nr - st; bax[cnt++] = ax[q]; });
            bap[lc + 1] = cnt;
        }
        rc = orc_klu_factor(s->blk[k], sz + 1, bap, cnt, bai, bax);
        free(bap); free(bai); free(bax);
        if (rc) return rc;
    }
    for (int64_t k = 0; k < nb; k++) { /* C_k and R_k values */
        const int64_t st = s->d.block_ptr[k], en = s->d.block_ptr[k + 1], sz = en - st;
        int64_t w = 0;
        for (int64_t lc = 0; lc < sz; lc++) FOR_COL(st + lc, if (nr >= b0) s->C[k].values[w++] = ax[q];);
        w = 0;
        for (int64_t lc = 0; lc < bs; lc++)
            FOR_COL(b0 + lc, if (nr >= st && nr < en) s->R[k].values[w++] = ax[q];);
    }
    double *work = (double *)orc_xmalloc((size_t)(maxb + 1) * sizeof(double));
    if (s->dense) {
        memset(s->bt, 0, (size_t)(bs * bs) * sizeof(double));
        for (int64_t lc = 0; lc < bs; lc++)
            FOR_COL(b0 + lc, if (nr >= b0) s->bt[(nr - b0) * bs + lc] += ax[q];);
        memcpy(s->S, s->bt, (size_t)(bs * bs) * sizeof(double));
    } else {
        memset(s->s_ax, 0, (size_t)s->s_ap[bs] * sizeof(double));
        for (int64_t lc = 0; lc < bs; lc++) {
            const int64_t cs = s->s_ap[lc], ce = s->s_ap[lc + 1];
            FOR_COL(b0 + lc, if (nr >= b0) {
                int64_t pos = bsearch_row(s->s_ai + cs, ce - cs, nr - b0);
                if (pos >= 0) s->s_ax[cs + pos] += ax[q];
            });
        }
    }
    for (int64_t k = 0; k < nb && !rc; k++) { /* S -= C_k B_kk^-1 R_k */
        const int64_t sz = s->d.block_ptr[k + 1] - s->d.block_ptr[k];
        const sblk *R = &s->R[k], *C = &s->C[k];
        for (int64_t j = 0; j < bs && !rc; j++) {
            if (R->col_ptr[j] == R->col_ptr[j + 1]) continue;
            memset(work, 0, (size_t)sz * sizeof(double));
            for (int64_t q = R->col_ptr[j]; q < R->col_ptr[j + 1]; q++) work[R->row_idx[q]] = R->values[q];
            rc = orc_klu_solve(s->blk[k], sz, work);
            if (rc) break;
            for (int64_t col = 0; col < sz; col++) {
                double y = work[col];
                if (y == 0.0) continue;
                for (int64_t q = C->col_ptr[col]; q < C->col_ptr[col + 1]; q++) {
                    int64_t br = C->row_idx[q];
                    if (s->dense) {
                        s->S[br * bs + j] -= C->values[q] * y;
                    } else {
                        const int64_t cs = s->s_ap[j], ce = s->s_ap[j + 1];
                        int64_t pos = bsearch_row(s->s_ai + cs, ce - cs, br);
                        if (pos >= 0) s->s_ax[cs + pos] -= C->values[q] * y;
                    }
                }
            }
        }
    }
    free(work);
    if (rc) return rc;
    if (s->dense) {
        memcpy(s->lu, s->S, (size_t)(bs * bs) * sizeof(double));
        rc = orc_dense_factor(bs, s->lu, s->piv, 1);
        s->dense_ok = rc == 0;
        return rc;
    }
    return orc_klu_factor(s->ssolver, bs + 1, s->s_ap, s->s_ap[bs], s->s_ai, s->s_ax);
}

int orc_bbd_factor(orc_bbd *s, int64_t apl, const int64_t *ap, int64_t ail, const int64_t *ai,
                   const double *ax) { /* :1035-1052 */
    int rc;
    if (!s->analyzed || apl != s->last_apl || ail != s->last_ail ||
        memcmp(ap, s->last_ap, (size_t)apl * 8) || memcmp(ai, s->last_ai, (size_t)ail * 8)) {
        if ((rc = orc_bbd_analyze(s, apl, ap, ail, ai))) return rc;
    }
    if (s->has_decomp) rc = bbd_factor(s, ap, ai, ax);
    else if (s->fallback) rc = orc_klu_factor(s->fallback, apl, ap, ail, ai, ax);
    else rc = 2;
    if (rc) return rc;
    s->factored = 1;
    s->factor_count++;
    return 0;
}

/* solve_bbd (:907-985) */
int orc_bbd_solve(orc_bbd *s, int64_t n, double *rhs) {
    if (!s->factored) return 3;
    if (n != s->n) return 6;
    if (!s->has_decomp) return s->fallback ? orc_klu_solve(s->fallback, n, rhs) : 3;
    const int64_t nb = s->d.num_blocks, bs = s->d.border_size, b0 = s->d.border_start;
    double *work = (double *)orc_xmalloc((size_t)(n + 1) * sizeof(double));
    double *z = (double *)orc_xmalloc((size_t)(n + 1) * sizeof(double));
    double *w = (double *)orc_xmalloc((size_t)(n + 1) * sizeof(double));
    double *bb = (double *)orc_xmalloc((size_t)(bs + 1) * sizeof(double));
    int rc = 0;
    for (int64_t i = 0; i < n; i++) work[i] = rhs[s->d.perm[i]];
    for (int64_t k = 0; k < nb && !rc; k++) {
        const int64_t st = s->d.block_ptr[k], sz = s->d.block_ptr[k + 1] - st;
        memcpy(z + st, work + st, (size_t)sz * sizeof(double));
        rc = orc_klu_solve(s->blk[k], sz, z + st);
    }
    if (rc) goto out;
    memcpy(bb, work + b0, (size_t)bs * sizeof(double));
    for (int64_t k = 0; k < nb; k++) { /* matvec_add_neg (:988-1000) */
        const int64_t st = s->d.block_ptr[k];
        const sblk *C = &s->C[k];
        for (int64_t col = 0; col < C->ncols; col++) {
            double xc = z[st + col];
            if (xc == 0.0) continue;
            for (int64_t q = C->col_ptr[col]; q < C->col_ptr[col + 1]; q++)
                bb[C->row_idx[q]] -= C->values[q] * xc;
        }
    }
    rc = s->dense ? orc_dense_solve(bs, s->lu, s->piv, bb, 1) : orc_klu_solve(s->ssolver, bs, bb);
    if (rc) goto out;
    for (int64_t k = 0; k < nb && !rc; k++) {
        const int64_t st = s->d.block_ptr[k], sz = s->d.block_ptr[k + 1] - st;
        const sblk *R = &s->R[k];
        for (int64_t i = 0; i < sz; i++) w[i] = 0.0;
        for (int64_t j = 0; j < bs; j++) { /* matvec_add (:68-81) */
            double xc = bb[j];
            if (xc == 0.0) continue;
            for (int64_t q = R->col_ptr[j]; q < R->col_ptr[j + 1]; q++) w[R->row_idx[q]] += R->values[q] * xc;
        }
        rc = orc_klu_solve(s->blk[k], sz, w);
        for (int64_t i = 0; i < sz; i++) work[st + i] = z[st + i] - w[i];
    }
    if (rc) goto out;
    for (int64_t i = 0; i < bs; i++) work[b0 + i] = bb[i];
    for (int64_t i = 0; i < n; i++) rhs[i] = work[s->d.inv_perm[i]];
out:
    free(work); free(z); free(w); free(bb);
    return rc;
}

int orc_bbd_is_bbd(const orc_bbd *s) { return s->has_decomp; }

int orc_bbd_info(const orc_bbd *s, int64_t *out) {
    out[0] = s->has_decomp ? s->d.num_blocks : 1;
    out[1] = s->has_decomp ? s->d.border_size : 0;
    out[2] = s->has_decomp ? s->dense : -1;
    out[3] = s->factor_count;
    return s->has_decomp;
}
