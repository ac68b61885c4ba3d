// symbolic.cpp - host integer analysis (AMD, BTF, L/U patterns) and the
// first pivoting factor. See symbolic.hpp for the reference files followed.
#include "symbolic.hpp"

#include <algorithm>
#include <cmath>
#include <queue>
#include <utility>

namespace rsgpu {

// ---------------------------------------------------------------------------
// AMD (amd.rs:143-411).
//
// The reference eliminates, at every step, the live variable with the
// smallest (degree, index), where degree is the number of distinct live
// variables reachable from it directly or through an eliminated element, and
// only the pivot's neighbours get their degree recomputed. Its output is a
// function of those sets alone, so this version keeps the same quotient graph
// with cheaper bookkeeping: static variable lists, element lists that are
// appended to instead of rewritten in place, absorbed elements skipped when
// scanned, and a shortcut for a variable whose own live neighbours already
// cover every other live variable (the ground node of an MNA matrix), whose
// degree is then nlive-1 without a scan.
//
// The degree of a variable v next to the new element L_p is
//   |L_p \ {v}| + |(A_v u the other elements of v) \ L_p \ {v}|,
// so L_p is counted once instead of scanned for each of its members, and an
// element whose variables all lie in L_p (|L_e \ L_p| = 0, counted once per
// pivot and element) adds nothing and is not scanned. When only one other
// element e1 reaches outside L_p, its contribution is |L_e1 \ L_p| and a
// live neighbour of v outside L_p counts unless e1 is among its elements, so
// nothing is scanned element-wide at all. Together these keep the large
// separator cliques of a grid from being rescanned for every one of their
// variables. Non-absorbed elements hold live variables only: eliminating a
// variable absorbs every element that contains it.
// ---------------------------------------------------------------------------
AmdResult amd_order(i64 n, const i64* ap, const i64* ai) {
    AmdResult r;
    r.perm.assign(n, 0);
    r.inv_perm.assign(n, 0);
    if (n == 0) return r;
    // symmetric variable adjacency without self loops or duplicates
    std::vector<std::pair<i64, i64>> edges;
    edges.reserve((size_t)ap[n] * 2);
    for (i64 c = 0; c < n; c++)
        for (i64 k = ap[c]; k < ap[c + 1]; k++) {
            i64 rw = ai[k];
            if (rw < n && rw != c) {
                edges.emplace_back(c, rw);
                edges.emplace_back(rw, c);
            }
        }
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    std::vector<i64> vptr(n + 1, 0), vadj(edges.size());
    for (auto& e : edges) vptr[e.first + 1]++;
    for (i64 i = 0; i < n; i++) vptr[i + 1] += vptr[i];
    for (size_t q = 0; q < edges.size(); q++) vadj[q] = edges[q].second;
    edges.clear();
    edges.shrink_to_fit();

    std::vector<i64> status(n);  // -1 element, else degree
    std::vector<i64> live_nbrs(n);
    std::vector<std::vector<i64>> eadj(n), evars(n);
    std::vector<uint8_t> absorbed(n, 0);
    std::vector<i64> marker(n, 0);   // per-degree dedup stamps
    std::vector<i64> in_lp(n, -1);   // == p: the variable belongs to L_p
    std::vector<i64> ecnt_at(n, -1), ecnt(n, 0);  // |L_e \ L_p| of pivot ecnt_at[e]
    i64 cur = 0, nlive = n;
    using E = std::pair<i64, i64>;
    std::priority_queue<E, std::vector<E>, std::greater<E>> heap;
    for (i64 i = 0; i < n; i++) {
        status[i] = vptr[i + 1] - vptr[i];
        live_nbrs[i] = status[i];
        heap.emplace(status[i], i);
    }
    std::vector<i64> nvars;
    for (i64 step = 0; step < n; step++) {
        i64 p = -1;
        while (!heap.empty()) {
            E e = heap.top();
            heap.pop();
            if (status[e.second] < 0 || status[e.second] != e.first) continue;
            p = e.second;
            break;
        }
        if (p < 0) break;
        r.perm[p] = step;
        r.inv_perm[step] = p;
        i64 mark = ++cur;
        nvars.clear();
        marker[p] = mark;
        for (i64 q = vptr[p]; q < vptr[p + 1]; q++) {
            i64 a = vadj[q];
            if (status[a] >= 0 && marker[a] != mark) { marker[a] = mark; nvars.push_back(a); }
        }
        for (i64 e : eadj[p]) {
            if (absorbed[e]) continue;
            for (i64 a : evars[e])
                if (status[a] >= 0 && marker[a] != mark) { marker[a] = mark; nvars.push_back(a); }
        }
        status[p] = -1;
        nlive--;
        for (i64 e : eadj[p]) {
            if (absorbed[e]) continue;
            absorbed[e] = 1;
            std::vector<i64>().swap(evars[e]);
        }
        std::vector<i64>().swap(eadj[p]);
        evars[p] = nvars;
        for (i64 q = vptr[p]; q < vptr[p + 1]; q++) live_nbrs[vadj[q]]--;
        for (i64 v : nvars) eadj[v].push_back(p);
        // L_p membership, and |L_e \ L_p| once per element next to a member
        // of L_p whose degree is computed below (the live neighbours of a
        // variable adjacent to every live variable need no element lists)
        for (i64 v : nvars) in_lp[v] = p;
        for (i64 v : nvars) {
            if (live_nbrs[v] == nlive - 1) continue;
            auto& el = eadj[v];
            size_t w = 0;
            for (size_t q = 0; q < el.size(); q++) {
                i64 e = el[q];
                if (absorbed[e]) continue;
                el[w++] = e;
                if (e == p || ecnt_at[e] == p) continue;
                ecnt_at[e] = p;
                i64 outside = 0;
                for (i64 a : evars[e]) outside += in_lp[a] != p;
                ecnt[e] = outside;
            }
            el.resize(w);
        }
        const i64 lp_size = (i64)nvars.size();
        for (i64 v : nvars) {
            i64 deg;
            if (live_nbrs[v] == nlive - 1) {
                deg = nlive - 1;
            } else {
                deg = lp_size - 1;  // L_p without v
                // the elements of v that reach outside L_p
                i64 e1 = -1, nout = 0;
                for (i64 e : eadj[v])  // pruned above
                    if (e != p && ecnt[e] != 0) { e1 = e; nout++; }
                if (nout <= 1) {
                    // one such element: |L_e1 \ L_p| is known, and a live
                    // neighbour outside L_p counts unless it is in L_e1
                    if (nout == 1) deg += ecnt[e1];
                    for (i64 q = vptr[v]; q < vptr[v + 1]; q++) {
                        i64 a = vadj[q];
                        if (status[a] < 0 || in_lp[a] == p) continue;
                        bool in_e1 = false;
                        if (nout == 1)
                            for (i64 e : eadj[a])
                                if (e == e1) { in_e1 = true; break; }
                        deg += !in_e1;
                    }
                } else {
                    i64 m2 = ++cur;
                    marker[v] = m2;
                    for (i64 q = vptr[v]; q < vptr[v + 1]; q++) {
                        i64 a = vadj[q];
                        if (status[a] >= 0 && in_lp[a] != p && marker[a] != m2) { marker[a] = m2; deg++; }
                    }
                    for (i64 e : eadj[v]) {
                        if (e == p || ecnt[e] == 0) continue;  // L_p itself, or inside it
                        for (i64 a : evars[e])
                            if (in_lp[a] != p && marker[a] != m2) { marker[a] = m2; deg++; }
                    }
                }
            }
            status[v] = deg;
            heap.emplace(deg, v);
        }
    }
    return r;
}

// ---------------------------------------------------------------------------
// BTF (btf.rs:143-550): DFS augmenting-path matching over rows 0..n-1 with
// columns in CSC order, Tarjan SCC over columns, blocks in reverse Tarjan
// order, unmatched columns as trailing singletons paired with the lowest
// unused rows. Recursion is replaced by explicit stacks that visit neighbours
// in the same order; the per-row visited array is generation stamped.
// ---------------------------------------------------------------------------
Btf btf_decompose(i64 n, const i64* ap, const i64* ai) {
    Btf b;
    b.block_ptr.push_back(0);
    if (n == 0) return b;
    std::vector<i64> rptr(n + 1, 0);
    for (i64 c = 0; c < n; c++)
        for (i64 k = ap[c]; k < ap[c + 1]; k++)
            if (ai[k] < n) rptr[ai[k] + 1]++;
    for (i64 i = 0; i < n; i++) rptr[i + 1] += rptr[i];
    std::vector<i64> rcols(rptr[n]);
    {
        std::vector<i64> f(rptr.begin(), rptr.end() - 1);
        for (i64 c = 0; c < n; c++)
            for (i64 k = ap[c]; k < ap[c + 1]; k++)
                if (ai[k] < n) rcols[f[ai[k]]++] = c;
    }
    std::vector<i64> cm(n, -1), rm(n, -1), vis(n, 0);
    struct Fr { i64 row, pos, col; };
    std::vector<Fr> st;
    st.reserve(64);
    for (i64 row = 0; row < n; row++) {
        i64 gen = row + 1;
        st.clear();
        st.push_back({row, rptr[row], -1});
        int ret = -1;
        while (!st.empty()) {
            Fr& f = st.back();
            if (ret == 1) {
                cm[f.col] = f.row;
                rm[f.row] = f.col;
                st.pop_back();
                continue;
            }
            ret = -1;
            bool pushed = false, found = false;
            while (f.pos < rptr[f.row + 1]) {
                i64 col = rcols[f.pos++];
                if (vis[col] == gen) continue;
                vis[col] = gen;
                if (cm[col] == -1) {
                    cm[col] = f.row;
                    rm[f.row] = col;
                    found = true;
                    break;
                }
                f.col = col;
                i64 mr = cm[col];
                st.push_back({mr, rptr[mr], -1});
                pushed = true;
                break;
            }
            if (pushed) continue;
            st.pop_back();
            ret = found ? 1 : 0;
        }
    }
    i64 rank = 0;
    for (i64 j = 0; j < n; j++) rank += cm[j] != -1;
    // directed graph col_i -> col_j when A(match[j], i) != 0 (duplicates kept)
    std::vector<i64> gptr(n + 1, 0);
    for (i64 j = 0; j < n; j++) {
        if (cm[j] < 0) continue;
        i64 r = cm[j];
        for (i64 k = rptr[r]; k < rptr[r + 1]; k++)
            if (rcols[k] != j) gptr[rcols[k] + 1]++;
    }
    for (i64 i = 0; i < n; i++) gptr[i + 1] += gptr[i];
    std::vector<i64> gadj(gptr[n]);
    {
        std::vector<i64> f(gptr.begin(), gptr.end() - 1);
        for (i64 j = 0; j < n; j++) {
            if (cm[j] < 0) continue;
            i64 r = cm[j];
            for (i64 k = rptr[r]; k < rptr[r + 1]; k++)
                if (rcols[k] != j) gadj[f[rcols[k]]++] = j;
        }
    }
    // Tarjan
    std::vector<i64> idx(n, -1), low(n, 0), stk, order, sptr{0};
    std::vector<uint8_t> on(n, 0);
    std::vector<std::pair<i64, i64>> cs;
    i64 counter = 0;
    for (i64 v0 = 0; v0 < n; v0++) {
        if (idx[v0] != -1) continue;
        idx[v0] = low[v0] = counter++;
        stk.push_back(v0);
        on[v0] = 1;
        cs.push_back({v0, gptr[v0]});
        while (!cs.empty()) {
            i64 v = cs.back().first;
            i64& pos = cs.back().second;
            if (pos < gptr[v + 1]) {
                i64 w = gadj[pos++];
                if (idx[w] == -1) {
                    idx[w] = low[w] = counter++;
                    stk.push_back(w);
                    on[w] = 1;
                    cs.push_back({w, gptr[w]});
                } else if (on[w]) {
                    low[v] = std::min(low[v], idx[w]);
                }
                continue;
            }
            if (low[v] == idx[v]) {
                for (;;) {
                    i64 w = stk.back();
                    stk.pop_back();
                    on[w] = 0;
                    order.push_back(w);
                    if (w == v) break;
                }
                sptr.push_back((i64)order.size());
            }
            cs.pop_back();
            if (!cs.empty()) {
                i64 par = cs.back().first;
                low[par] = std::min(low[par], low[v]);
            }
        }
    }
    i64 nscc = (i64)sptr.size() - 1;
    std::vector<uint8_t> used(n, 0);
    for (i64 s = nscc - 1; s >= 0; s--) {
        i64 cnt = 0;
        for (i64 q = sptr[s]; q < sptr[s + 1]; q++) {
            i64 col = order[q];
            if (cm[col] < 0) continue;
            b.col_perm.push_back(col);
            b.row_perm.push_back(cm[col]);
            used[cm[col]] = 1;
            cnt++;
        }
        if (!cnt) continue;
        b.block_ptr.push_back((i64)b.col_perm.size());
        if (cnt == 1) b.num_singletons++;
        b.max_block_size = std::max(b.max_block_size, cnt);
    }
    i64 next_free = 0;
    for (i64 col = 0; col < n; col++) {
        if (cm[col] >= 0) continue;
        b.col_perm.push_back(col);
        while (next_free < n && used[next_free]) next_free++;
        if (next_free < n) {
            b.row_perm.push_back(next_free);
            used[next_free] = 1;
        }
        b.block_ptr.push_back((i64)b.col_perm.size());
        b.num_singletons++;
    }
    b.num_blocks = (i64)b.block_ptr.size() - 1;
    b.row_perm_inv.assign(n, 0);
    b.col_perm_inv.assign(n, 0);
    for (i64 i = 0; i < n; i++) {
        b.row_perm_inv[b.row_perm[i]] = i;
        b.col_perm_inv[b.col_perm[i]] = i;
    }
    b.structural_rank = rank;
    return b;
}

// ---------------------------------------------------------------------------
// L/U pattern (native_klu.rs:863-1014): the pattern of column k of P*A*P^T is
// the set reachable from its off-diagonal rows through the L columns of rows
// < k; rows < k form U (ascending, then the diagonal), rows > k form L.
// ---------------------------------------------------------------------------
void block_symbolic(i64 n, const i64* ap, const i64* ai, bool natural, BlockSymbolic& s,
                    bool zero_diag_last) {
    s.n = n;
    s.lcp.assign(n + 1, 0);
    s.ucp.assign(n + 1, 0);
    s.lri.clear();
    s.uri.clear();
    if (n == 0) return;
    if (natural) {
        s.col_perm.resize(n);
        s.col_perm_inv.resize(n);
        for (i64 i = 0; i < n; i++) s.col_perm[i] = s.col_perm_inv[i] = i;
    } else {
        AmdResult a = amd_order(n, ap, ai);
        s.col_perm = std::move(a.perm);
        s.col_perm_inv = std::move(a.inv_perm);
    }
    if (zero_diag_last) {  // unknowns without a diagonal entry after the others, order kept
        std::vector<char> has(n, 0);
        for (i64 c = 0; c < n; c++)
            for (i64 q = ap[c]; q < ap[c + 1]; q++)
                if (ai[q] == c) has[c] = 1;
        std::vector<i64> inv;
        inv.reserve(n);
        for (int pass = 1; pass >= 0; pass--)
            for (i64 k = 0; k < n; k++)
                if (has[s.col_perm_inv[k]] == pass) inv.push_back(s.col_perm_inv[k]);
        s.col_perm_inv = inv;
        for (i64 k = 0; k < n; k++) s.col_perm[inv[k]] = k;
    }
    std::vector<i64> mark(n, -1), stack, pat, ur, lr;
    for (i64 k = 0; k < n; k++) {
        pat.clear();
        stack.clear();
        i64 oc = s.col_perm_inv[k];
        for (i64 q = ap[oc]; q < ap[oc + 1]; q++) {
            i64 orow = ai[q];
            if (orow >= n) continue;
            i64 row = s.col_perm[orow];
            if (row == k || mark[row] == k) continue;
            mark[row] = k;
            pat.push_back(row);
            if (row < k) stack.push_back(row);
        }
        while (!stack.empty()) {
            i64 node = stack.back();
            stack.pop_back();
            for (i64 q = s.lcp[node]; q < s.lcp[node + 1]; q++) {
                i64 ch = s.lri[q];
                if (mark[ch] == k) continue;
                mark[ch] = k;
                pat.push_back(ch);
                if (ch < k) stack.push_back(ch);
            }
        }
        ur.clear();
        lr.clear();
        for (i64 row : pat) {
            if (row < k) ur.push_back(row);
            else if (row > k) lr.push_back(row);
        }
        std::sort(ur.begin(), ur.end());
        std::sort(lr.begin(), lr.end());
        s.ucp[k] = (i64)s.uri.size();
        s.uri.insert(s.uri.end(), ur.begin(), ur.end());
        s.uri.push_back(k);
        s.ucp[k + 1] = (i64)s.uri.size();
        s.lcp[k] = (i64)s.lri.size();
        s.lri.insert(s.lri.end(), lr.begin(), lr.end());
        s.lcp[k + 1] = (i64)s.lri.size();
    }
}

// ---------------------------------------------------------------------------
// factor_block (native_klu.rs:1194-1382). Same operation order as the
// reference so results are bit-identical (host code is built without FMA
// contraction).
// ---------------------------------------------------------------------------
void factor_block_host(const BlockSymbolic& s, const i64* ap, const i64* ai, const double* ax,
                       const BlockNumeric* old, bool refactor, double pivot_tol,
                       BlockNumeric& out, bool sparselu) {
    const i64 n = s.n;
    out = BlockNumeric();
    if (n == 0) return;
    if (n == 1 && !sparselu) {
        i64 oc = s.col_perm_inv[0];
        double diag = 0.0;
        for (i64 q = ap[oc]; q < ap[oc + 1]; q++) {
            i64 row = ai[q];
            if (row < n && s.col_perm[row] == 0) diag += ax[q];
        }
        if (std::fabs(diag) < 1e-300) diag = 1e-14;
        out.ux = {diag};
        out.udiag = {diag};
        out.row_perm = {0};
        out.row_perm_inv = {0};
        return;
    }
    out.lx.assign(s.lri.size(), 0.0);
    out.ux.assign(s.uri.size(), 0.0);
    out.udiag.assign(n, 0.0);
    const bool use_old = refactor && old != nullptr;
    if (use_old) {
        out.row_perm = old->row_perm;
        out.row_perm_inv = old->row_perm_inv;
    } else {
        out.row_perm.resize(n);
        out.row_perm_inv.resize(n);
        for (i64 i = 0; i < n; i++) out.row_perm[i] = out.row_perm_inv[i] = i;
    }
    std::vector<i64>& rp = out.row_perm;
    std::vector<i64>& rpi = out.row_perm_inv;
    std::vector<double> work(n, 0.0);
    for (i64 k = 0; k < n; k++) {
        i64 oc = s.col_perm_inv[k];
        i64 st = ap[oc], en = ap[oc + 1];
        for (i64 q = st; q < en; q++) {
            i64 orow = ai[q];
            if (orow < n) work[rpi[s.col_perm[orow]]] += ax[q];
        }
        i64 us = s.ucp[k], ue = s.ucp[k + 1] - 1;
        for (i64 u = us; u < ue; u++) {
            i64 j = s.uri[u];
            double ujk = work[j];
            if (ujk != 0.0)
                for (i64 l = s.lcp[j]; l < s.lcp[j + 1]; l++) work[s.lri[l]] -= out.lx[l] * ujk;
        }
        i64 ls = s.lcp[k], le = s.lcp[k + 1];
        if (!use_old) {
            double dabs = std::fabs(work[k]), mx = dabs;
            i64 mrow = k;
            for (i64 l = ls; l < le; l++) {
                double v = std::fabs(work[s.lri[l]]);
                if (v > mx) { mx = v; mrow = s.lri[l]; }
            }
            if (mrow != k && dabs < pivot_tol * mx) {
                std::swap(work[k], work[mrow]);
                i64 ok = rp[k], om = rp[mrow];
                rp[k] = om;
                rp[mrow] = ok;
                rpi[ok] = mrow;
                rpi[om] = k;
            }
        }
        for (i64 u = us; u < ue; u++) out.ux[u] = work[s.uri[u]];
        double diag = work[k];
        if (std::fabs(diag) < 1e-14) diag = diag >= 0.0 ? 1e-14 : -1e-14;
        out.ux[s.ucp[k + 1] - 1] = diag;
        out.udiag[k] = diag;
        for (i64 l = ls; l < le; l++) out.lx[l] = work[s.lri[l]] / diag;
        for (i64 q = st; q < en; q++) {
            i64 orow = ai[q];
            if (orow < n) work[rpi[s.col_perm[orow]]] = 0.0;
        }
        for (i64 l = ls; l < le; l++) work[s.lri[l]] = 0.0;
        for (i64 u = us; u < ue; u++) work[s.uri[u]] = 0.0;
        work[k] = 0.0;
    }
}

}  // namespace rsgpu
