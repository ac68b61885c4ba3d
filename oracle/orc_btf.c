/*
 * orc_btf.c - restatement of crates/sim-core/src/btf.rs:143-550.
 * TEST INFRASTRUCTURE (oracle).
 *
 * The reference's recursive DFS (augment_path, strong_connect) is written here
 * with explicit stacks that visit neighbours in the same order, and its
 * per-row `vec![false; n]` visited array is a generation-stamped array; both
 * give the same matching, SCC order and permutations.
 */
#include "orc.h"
#include "orc_util.h"

typedef struct { int64_t *ptr, *idx; } adjl; /* CSR-like lists */

/* build_row_adjacency (btf.rs:226-241): row -> columns in CSC order, with
 * duplicate entries kept */
static adjl row_adjacency(int64_t n, const int64_t *ap, const int64_t *ai) {
    adjl r;
    r.ptr = (int64_t *)orc_xcalloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t c = 0; c < n; c++)
        for (int64_t k = ap[c]; k < ap[c + 1]; k++)
            if (ai[k] < n) r.ptr[ai[k] + 1]++;
    for (int64_t i = 0; i < n; i++) r.ptr[i + 1] += r.ptr[i];
    r.idx = (int64_t *)orc_xmalloc((size_t)(r.ptr[n] ? r.ptr[n] : 1) * sizeof(int64_t));
    int64_t *fill = (int64_t *)orc_xmalloc((size_t)n * sizeof(int64_t));
    memcpy(fill, r.ptr, (size_t)n * sizeof(int64_t));
    for (int64_t c = 0; c < n; c++)
        for (int64_t k = ap[c]; k < ap[c + 1]; k++)
            if (ai[k] < n) r.idx[fill[ai[k]]++] = c;
    free(fill);
    return r;
}

/* maximum_transversal + augment_path (btf.rs:202-280) */
static void max_transversal(int64_t n, const adjl *r2c, int64_t *col_match,
                            int64_t *row_match) {
    for (int64_t i = 0; i < n; i++) col_match[i] = row_match[i] = -1;
    int64_t *visited = (int64_t *)orc_xcalloc((size_t)n, sizeof(int64_t));
    int64_t *fr_row = (int64_t *)orc_xmalloc((size_t)(n + 1) * sizeof(int64_t));
    int64_t *fr_pos = (int64_t *)orc_xmalloc((size_t)(n + 1) * sizeof(int64_t));
    int64_t *fr_col = (int64_t *)orc_xmalloc((size_t)(n + 1) * sizeof(int64_t));
    for (int64_t row = 0; row < n; row++) {
        int64_t gen = row + 1; /* fresh visited set per starting row */
        int64_t d = 0;
        fr_row[0] = row; fr_pos[0] = r2c->ptr[row]; fr_col[0] = -1;
        d = 1;
        int child_ret = -1; /* -1: none, 0: false, 1: true */
        while (d > 0) {
            int64_t f = d - 1;
            if (child_ret == 1) {
                col_match[fr_col[f]] = fr_row[f];
                row_match[fr_row[f]] = fr_col[f];
                d--; child_ret = 1;
                continue;
            }
            child_ret = -1;
            int pushed = 0, done = 0;
            while (fr_pos[f] < r2c->ptr[fr_row[f] + 1]) {
                int64_t col = r2c->idx[fr_pos[f]++];
                if (visited[col] == gen) continue;
                visited[col] = gen;
                if (col_match[col] == -1) {
                    col_match[col] = fr_row[f];
                    row_match[fr_row[f]] = col;
                    done = 1;
                    break;
                }
                fr_col[f] = col;
                int64_t mr = col_match[col];
                fr_row[d] = mr; fr_pos[d] = r2c->ptr[mr]; fr_col[d] = -1;
                d++;
                pushed = 1;
                break;
            }
            if (pushed) continue;
            d--;
            child_ret = done ? 1 : 0;
        }
    }
    free(visited); free(fr_row); free(fr_pos); free(fr_col);
}

/* tarjan_scc / strong_connect (btf.rs:388-447), iterative */
static int64_t tarjan(int64_t n, const adjl *g, int64_t *scc_id_order,
                      int64_t *scc_ptr) {
    int64_t *index = (int64_t *)orc_xmalloc((size_t)n * sizeof(int64_t));
    int64_t *low = (int64_t *)orc_xmalloc((size_t)n * sizeof(int64_t));
    uint8_t *on = (uint8_t *)orc_xcalloc((size_t)n, 1);
    int64_t *stk = (int64_t *)orc_xmalloc((size_t)n * sizeof(int64_t));
    int64_t *cs_v = (int64_t *)orc_xmalloc((size_t)n * sizeof(int64_t));
    int64_t *cs_p = (int64_t *)orc_xmalloc((size_t)n * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) index[i] = -1;
    int64_t counter = 0, sp = 0, nscc = 0, out = 0;
    scc_ptr[0] = 0;
    for (int64_t v0 = 0; v0 < n; v0++) {
        if (index[v0] != -1) continue;
        int64_t d = 0;
        cs_v[0] = v0; cs_p[0] = g->ptr[v0]; d = 1;
        index[v0] = low[v0] = counter++;
        stk[sp++] = v0; on[v0] = 1;
        while (d > 0) {
            int64_t v = cs_v[d - 1];
            if (cs_p[d - 1] < g->ptr[v + 1]) {
                int64_t w = g->idx[cs_p[d - 1]++];
                if (index[w] == -1) {
                    index[w] = low[w] = counter++;
                    stk[sp++] = w; on[w] = 1;
                    cs_v[d] = w; cs_p[d] = g->ptr[w]; d++;
                } else if (on[w]) {
                    if (index[w] < low[v]) low[v] = index[w];
                }
                continue;
            }
            /* v finished */
            if (low[v] == index[v]) {
                for (;;) {
                    int64_t w = stk[--sp];
                    on[w] = 0;
                    scc_id_order[out++] = w;
                    if (w == v) break;
                }
                scc_ptr[++nscc] = out;
            }
            d--;
            if (d > 0) {
                int64_t parent = cs_v[d - 1];
                if (low[v] < low[parent]) low[parent] = low[v];
            }
        }
    }
    free(index); free(low); free(on); free(stk); free(cs_v); free(cs_p);
    return nscc;
}

void orc_btf_decompose(int64_t n, const int64_t *ap, const int64_t *ai, orc_btf *o) {
    memset(o, 0, sizeof(*o));
    o->n = n;
    o->row_perm = (int64_t *)orc_xmalloc((size_t)(n ? n : 1) * sizeof(int64_t));
    o->col_perm = (int64_t *)orc_xmalloc((size_t)(n ? n : 1) * sizeof(int64_t));
    o->row_perm_inv = (int64_t *)orc_xmalloc((size_t)(n ? n : 1) * sizeof(int64_t));
    o->col_perm_inv = (int64_t *)orc_xmalloc((size_t)(n ? n : 1) * sizeof(int64_t));
    o->block_ptr = (int64_t *)orc_xcalloc((size_t)n + 2, sizeof(int64_t));
    if (n == 0) return;
    adjl r2c = row_adjacency(n, ap, ai);
    int64_t *cm = (int64_t *)orc_xmalloc((size_t)n * sizeof(int64_t));
    int64_t *rm = (int64_t *)orc_xmalloc((size_t)n * sizeof(int64_t));
    max_transversal(n, &r2c, cm, rm);
    int64_t rank = 0;
    for (int64_t j = 0; j < n; j++) rank += cm[j] != -1;
    /* build_directed_graph (btf.rs:296-350): edge col_i -> col_j */
    adjl g;
    g.ptr = (int64_t *)orc_xcalloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t j = 0; j < n; j++) {
        if (cm[j] == -1) continue;
        int64_t r = cm[j];
        for (int64_t k = r2c.ptr[r]; k < r2c.ptr[r + 1]; k++)
            if (r2c.idx[k] != j) g.ptr[r2c.idx[k] + 1]++;
    }
    for (int64_t i = 0; i < n; i++) g.ptr[i + 1] += g.ptr[i];
    g.idx = (int64_t *)orc_xmalloc((size_t)(g.ptr[n] ? g.ptr[n] : 1) * sizeof(int64_t));
    int64_t *fill = (int64_t *)orc_xmalloc((size_t)n * sizeof(int64_t));
    memcpy(fill, g.ptr, (size_t)n * sizeof(int64_t));
    for (int64_t j = 0; j < n; j++) {
        if (cm[j] == -1) continue;
        int64_t r = cm[j];
        for (int64_t k = r2c.ptr[r]; k < r2c.ptr[r + 1]; k++) {
            int64_t ci = r2c.idx[k];
            if (ci != j) g.idx[fill[ci]++] = j;
        }
    }
    free(fill);
    int64_t *order = (int64_t *)orc_xmalloc((size_t)n * sizeof(int64_t));
    int64_t *sptr = (int64_t *)orc_xmalloc((size_t)(n + 1) * sizeof(int64_t));
    int64_t nscc = tarjan(n, &g, order, sptr);
    /* build_permutations (btf.rs:457-550): SCCs reversed */
    uint8_t *used = (uint8_t *)orc_xcalloc((size_t)n, 1);
    int64_t np = 0, nb = 0, nsing = 0, maxb = 0;
    o->block_ptr[0] = 0;
    for (int64_t s = nscc - 1; s >= 0; s--) {
        int64_t cnt = 0;
        for (int64_t q = sptr[s]; q < sptr[s + 1]; q++) {
            int64_t col = order[q];
            if (cm[col] == -1) continue;
            o->col_perm[np] = col;
            o->row_perm[np] = cm[col];
            used[cm[col]] = 1;
            np++; cnt++;
        }
        if (cnt == 0) continue;
        o->block_ptr[++nb] = np;
        if (cnt == 1) nsing++;
        if (cnt > maxb) maxb = cnt;
    }
    for (int64_t col = 0; col < n; col++) {
        if (cm[col] != -1) continue;
        o->col_perm[np] = col;
        for (int64_t row = 0; row < n; row++) {
            if (!used[row]) { o->row_perm[np] = row; used[row] = 1; break; }
        }
        np++;
        o->block_ptr[++nb] = np;
        nsing++;
    }
    o->num_blocks = nb;
    for (int64_t i = 0; i < n; i++) {
        o->row_perm_inv[o->row_perm[i]] = i;
        o->col_perm_inv[o->col_perm[i]] = i;
    }
    o->structural_rank = rank;
    o->num_singletons = nsing;
    o->max_block_size = maxb;
    free(r2c.ptr); free(r2c.idx); free(g.ptr); free(g.idx);
    free(cm); free(rm); free(order); free(sptr); free(used);
}

void orc_btf_free(orc_btf *b) {
    free(b->row_perm); free(b->col_perm); free(b->row_perm_inv);
    free(b->col_perm_inv); free(b->block_ptr);
    memset(b, 0, sizeof(*b));
}
