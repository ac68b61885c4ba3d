/*
 * orc_amd.c - literal restatement of crates/sim-core/src/amd.rs:143-411.
 * TEST INFRASTRUCTURE (oracle). Keeps the reference's data structures
 * (adjacency lists with linear `contains`, a (degree,node) min-heap with stale
 * re-push, marker-based degree recount) so the output permutation is the
 * reference's, including on ties.
 */
#include "orc.h"
#include "orc_util.h"

typedef struct { int64_t deg, node; } hent;
typedef struct { hent *a; int64_t len, cap; } heap_t;

static int hless(hent x, hent y) {
    return x.deg < y.deg || (x.deg == y.deg && x.node < y.node);
}
static void hpush(heap_t *h, hent e) {
    if (h->len == h->cap) {
        h->cap = h->cap ? h->cap * 2 : 16;
        h->a = (hent *)realloc(h->a, (size_t)h->cap * sizeof(hent));
    }
    int64_t i = h->len++;
    h->a[i] = e;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (!hless(h->a[i], h->a[p])) break;
        hent t = h->a[i]; h->a[i] = h->a[p]; h->a[p] = t; i = p;
    }
}
static int hpop(heap_t *h, hent *out) {
    if (h->len == 0) return 0;
    *out = h->a[0];
    h->a[0] = h->a[--h->len];
    int64_t i = 0;
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, m = i;
        if (l < h->len && hless(h->a[l], h->a[m])) m = l;
        if (r < h->len && hless(h->a[r], h->a[m])) m = r;
        if (m == i) break;
        hent t = h->a[i]; h->a[i] = h->a[m]; h->a[m] = t; i = m;
    }
    return 1;
}

/* status: -1 = Element, >=0 = Variable(degree) (amd.rs:164-170) */
void orc_amd_order(int64_t n, const int64_t *ap, const int64_t *ai, int64_t *perm,
                   int64_t *inv_perm) {
    if (n == 0) return;
    ivec *adj = (ivec *)orc_xcalloc((size_t)n, sizeof(ivec));
    /* amd.rs:210-224 */
    for (int64_t col = 0; col < n; col++) {
        for (int64_t idx = ap[col]; idx < ap[col + 1]; idx++) {
            int64_t row = ai[idx];
            if (row < n && row != col) {
                if (!iv_contains(&adj[col], row)) iv_push(&adj[col], row);
                if (!iv_contains(&adj[row], col)) iv_push(&adj[row], col);
            }
        }
    }
    int64_t *status = (int64_t *)orc_xmalloc((size_t)n * sizeof(int64_t));
    heap_t heap = {0};
    for (int64_t i = 0; i < n; i++) {
        status[i] = adj[i].len;
        hpush(&heap, (hent){adj[i].len, i});
    }
    int64_t *marker = (int64_t *)orc_xcalloc((size_t)n, sizeof(int64_t));
    int64_t cur_mark = 0, num_elim = 0;
    ivec nvars = {0}, nelems = {0};

    while (num_elim < n) {
        /* find_min_degree_variable (amd.rs:268-288) */
        int64_t p = -1;
        hent e;
        while (hpop(&heap, &e)) {
            if (status[e.node] < 0) continue;
            if (e.deg != status[e.node]) {
                hpush(&heap, (hent){status[e.node], e.node});
                continue;
            }
            p = e.node;
            break;
        }
        if (p < 0) break;
        /* eliminate (amd.rs:291-340) */
        perm[p] = num_elim;
        inv_perm[num_elim] = p;
        num_elim++;
        cur_mark++;
        int64_t mark = cur_mark;
        nvars.len = 0; nelems.len = 0;
        for (int64_t q = 0; q < adj[p].len; q++) {
            int64_t a = adj[p].a[q];
            if (status[a] >= 0) {
                if (marker[a] != mark) { marker[a] = mark; iv_push(&nvars, a); }
            } else {
                iv_push(&nelems, a);
            }
        }
        for (int64_t q = 0; q < nelems.len; q++) {
            int64_t el = nelems.a[q];
            for (int64_t r = 0; r < adj[el].len; r++) {
                int64_t a = adj[el].a[r];
                if (status[a] >= 0 && a != p && marker[a] != mark) {
                    marker[a] = mark; iv_push(&nvars, a);
                }
            }
        }
        status[p] = -1;
        adj[p].len = 0;
        for (int64_t q = 0; q < nvars.len; q++) iv_push(&adj[p], nvars.a[q]);
        /* absorb_element (amd.rs:343-369) */
        for (int64_t q = 0; q < nelems.len; q++) {
            int64_t src = nelems.a[q], dst = p;
            if (src == dst) continue;
            ivec src_adj = adj[src];
            adj[src] = (ivec){0};
            for (int64_t r = 0; r < src_adj.len; r++) {
                int64_t var = src_adj.a[r];
                if (status[var] >= 0) {
                    if (!iv_contains(&adj[dst], var)) iv_push(&adj[dst], var);
                    for (int64_t s = 0; s < adj[var].len; s++) {
                        if (adj[var].a[s] == src) { adj[var].a[s] = dst; break; }
                    }
                }
            }
            iv_free(&src_adj);
        }
        /* update_degree (amd.rs:372-410) */
        for (int64_t q = 0; q < nvars.len; q++) {
            int64_t var = nvars.a[q];
            if (status[var] < 0) continue;
            cur_mark++;
            int64_t m2 = cur_mark;
            marker[var] = m2;
            int64_t degree = 0;
            for (int64_t r = 0; r < adj[var].len; r++) {
                int64_t a = adj[var].a[r];
                if (status[a] >= 0) {
                    if (marker[a] != m2) { marker[a] = m2; degree++; }
                } else {
                    for (int64_t s = 0; s < adj[a].len; s++) {
                        int64_t ea = adj[a].a[s];
                        if (status[ea] >= 0 && marker[ea] != m2) { marker[ea] = m2; degree++; }
                    }
                }
            }
            status[var] = degree;
            hpush(&heap, (hent){degree, var});
        }
    }
    for (int64_t i = 0; i < n; i++) iv_free(&adj[i]);
    free(adj); free(status); free(marker); free(heap.a);
    iv_free(&nvars); iv_free(&nelems);
}
