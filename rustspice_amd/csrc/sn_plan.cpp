// sn_plan.cpp - host side of the supernodal factor (see sn.hpp): supernodes,
// panel layout, the ordered descendant contributions of every supernode, the
// levels of the supernodal elimination tree, the gather work items and the
// scatter of A; plus a host execution of the same schedule for checking it
// against the reference arithmetic on the CPU.
#include <algorithm>
#include <array>
#include <cmath>
#include <limits>

#include "sn.hpp"

namespace rsgpu {

namespace {

// index of r in the sorted range [b, e), or -1
int64_t find_sorted(const int32_t* b, const int32_t* e, int32_t r) {
    const int32_t* it = std::lower_bound(b, e, r);
    return (it != e && *it == r) ? (int64_t)(it - b) : -1;
}

bool fail(std::string* why, const char* m) {
    if (why) *why = m;
    return false;
}

// Whether the strictly upper part of U's row k is L(:, k)'s pattern for
// every k (U's rows collected column by column come out ascending, like lri).
bool structurally_symmetric(const BlockSymbolic& s) {
    const i64 n = s.n;
    std::vector<i64> cnt(n + 1, 0);
    for (i64 c = 0; c < n; c++)
        for (i64 q = s.ucp[c]; q < s.ucp[c + 1] - 1; q++) cnt[s.uri[q] + 1]++;
    for (i64 k = 0; k < n; k++)
        if (cnt[k + 1] != s.lcp[k + 1] - s.lcp[k]) return false;
    for (i64 k = 0; k < n; k++) cnt[k + 1] += cnt[k];
    std::vector<i64> fill(cnt.begin(), cnt.end() - 1);
    std::vector<int32_t> ucols((size_t)cnt[n]);
    for (i64 c = 0; c < n; c++)
        for (i64 q = s.ucp[c]; q < s.ucp[c + 1] - 1; q++) ucols[fill[s.uri[q]]++] = (int32_t)c;
    for (i64 k = 0; k < n; k++)
        for (i64 q = s.lcp[k], t = cnt[k]; q < s.lcp[k + 1]; q++, t++)
            if (ucols[t] != (int32_t)s.lri[q]) return false;
    return true;
}

// The symmetric closure of an unsymmetric fill pattern: the Cholesky fill of
// pattern(L + U^T) in the same order, column by column (a column's rows are
// its own entries and those of its elimination-tree children, less itself).
// It holds L and U^T (both lie in the Cholesky fill of A + A^T), so the
// plan's panels hold every entry of the factor plus explicit zeros. Those
// change no value: an entry outside the factor's pattern stays exactly +0.0
// through the fan-in (0.0 - (+-0) = +0.0) and contributes +-0 terms, and the
// reference's running sum is never -0.0 (it starts as 0.0 + A(i, k), and
// an exact cancellation rounds to +0.0), so subtracting +-0 leaves it
// unchanged; a zero U(j, k) is skipped as the reference skips it.
void symmetric_closure(const BlockSymbolic& s, BlockSymbolic& o) {
    const i64 n = s.n;
    // lower part of L + U^T by column
    std::vector<i64> bp(n + 1, 0);
    for (i64 k = 0; k < n; k++) bp[k + 1] += s.lcp[k + 1] - s.lcp[k];
    for (i64 c = 0; c < n; c++)
        for (i64 q = s.ucp[c]; q < s.ucp[c + 1] - 1; q++) bp[s.uri[q] + 1]++;
    for (i64 k = 0; k < n; k++) bp[k + 1] += bp[k];
    std::vector<int32_t> bi((size_t)bp[n]);
    {
        std::vector<i64> f(bp.begin(), bp.end() - 1);
        for (i64 k = 0; k < n; k++)
            for (i64 q = s.lcp[k]; q < s.lcp[k + 1]; q++) bi[f[k]++] = (int32_t)s.lri[q];
        for (i64 c = 0; c < n; c++)
            for (i64 q = s.ucp[c]; q < s.ucp[c + 1] - 1; q++) bi[f[s.uri[q]]++] = (int32_t)c;
    }
    o.n = n;
    o.col_perm = s.col_perm;
    o.col_perm_inv = s.col_perm_inv;
    o.lcp.assign(n + 1, 0);
    o.lri.clear();
    std::vector<int32_t> head(n, -1), next(n, -1), mark(n, -1), col;
    for (i64 k = 0; k < n; k++) {
        col.clear();
        auto add = [&](int32_t r) {
            if (r > k && mark[r] != k) {
                mark[r] = (int32_t)k;
                col.push_back(r);
            }
        };
        for (i64 q = bp[k]; q < bp[k + 1]; q++) add(bi[q]);
        for (int32_t c = head[k]; c >= 0; c = next[c])
            for (i64 q = o.lcp[c]; q < o.lcp[c + 1]; q++) add((int32_t)o.lri[q]);
        std::sort(col.begin(), col.end());
        for (int32_t r : col) o.lri.push_back(r);
        o.lcp[k + 1] = (i64)o.lri.size();
        if (!col.empty()) {  // elimination-tree parent: the first row below k
            next[k] = head[col[0]];
            head[col[0]] = (int32_t)k;
        }
    }
    // U = L^T with the diagonal last; rows come out ascending
    o.ucp.assign(n + 1, 0);
    for (int32_t r : o.lri) o.ucp[r + 1]++;
    for (i64 c = 0; c < n; c++) o.ucp[c + 1] += o.ucp[c] + 1;
    o.uri.assign((size_t)o.ucp[n], 0);
    std::vector<i64> f(o.ucp.begin(), o.ucp.end() - 1);
    for (i64 k = 0; k < n; k++)
        for (i64 q = o.lcp[k]; q < o.lcp[k + 1]; q++) o.uri[f[o.lri[q]]++] = k;
    for (i64 c = 0; c < n; c++) o.uri[f[c]] = c;
}

}  // namespace

bool build_sn_plan(const BlockSymbolic& s0, const i64* ap, const i64* ai, SnHost& h,
                   std::string* why, int32_t hot_min, bool transpose) {
    h = SnHost();
    const i64 n = s0.n;
    if (n <= 0) return fail(why, "empty matrix");
    if (n >= (i64)std::numeric_limits<int32_t>::max() / 2) return fail(why, "n exceeds 32-bit indices");
    h.n = (int32_t)n;

    // 1. a structurally symmetric pattern (every U row of a supernode spans
    //    its R_J): the symbolic's own, or its symmetric closure
    BlockSymbolic closure;
    h.symmetric = structurally_symmetric(s0);
    if (!h.symmetric) symmetric_closure(s0, closure);
    const BlockSymbolic& s = h.symmetric ? s0 : closure;
    const std::vector<i64>& lcp = s.lcp;
    const std::vector<i64>& lri = s.lri;
    const std::vector<i64>& ucp = s.ucp;
    const std::vector<i64>& uri = s.uri;

    // 2. fundamental supernodes: column c + 1 continues c when L(:, c) is
    //    {c + 1} u L(:, c + 1) and c is c + 1's only child
    std::vector<int32_t> nchild(n, 0);
    for (i64 c = 0; c < n; c++)
        if (lcp[c + 1] > lcp[c]) nchild[lri[lcp[c]]]++;
    h.j0.push_back(0);
    for (i64 c = 1; c < n; c++) {
        const i64 lp = lcp[c] - lcp[c - 1];
        const bool cont = lp > 0 && lri[lcp[c - 1]] == c && lp == lcp[c + 1] - lcp[c] + 1 && nchild[c] == 1;
        if (!cont) h.j0.push_back((int32_t)c);
    }
    h.j0.push_back((int32_t)n);
    const int32_t ns = (int32_t)h.j0.size() - 1;
    h.ns = ns;
    std::vector<int32_t> sn_of(n);
    for (int32_t J = 0; J < ns; J++)
        for (int32_t c = h.j0[J]; c < h.j0[J + 1]; c++) sn_of[c] = J;

    // 3. R_J (the L rows of J's last column) and the panel layout
    h.rptr.assign(ns + 1, 0);
    for (int32_t J = 0; J < ns; J++) {
        const i64 lc = h.j0[J + 1] - 1;
        h.rptr[J + 1] = h.rptr[J] + (int32_t)(lcp[lc + 1] - lcp[lc]);
    }
    h.rows.resize(h.rptr[ns]);
    h.oL.resize(ns);
    h.oU.resize(ns);
    int64_t off = 0;
    for (int32_t J = 0; J < ns; J++) {
        const i64 lc = h.j0[J + 1] - 1;
        for (i64 q = lcp[lc]; q < lcp[lc + 1]; q++) h.rows[h.rptr[J] + (q - lcp[lc])] = (int32_t)lri[q];
        const int64_t w = h.j0[J + 1] - h.j0[J], hh = h.rptr[J + 1] - h.rptr[J];
        h.max_w = std::max(h.max_w, w);
        h.max_h = std::max(h.max_h, hh);
        h.oL[J] = off;
        off += (w + hh) * w;
        h.oU[J] = off;
        off += w * hh;
    }
    h.nvals = off;

    // 4. supernodal elimination tree and its levels
    h.parent.assign(ns, -1);
    std::vector<int32_t> level(ns, 0);
    for (int32_t J = 0; J < ns; J++) {
        if (h.rptr[J + 1] > h.rptr[J]) {
            const int32_t P = sn_of[h.rows[h.rptr[J]]];
            if (P <= J) return fail(why, "supernodal tree is not ascending");
            h.parent[J] = P;
            level[P] = std::max(level[P], level[J] + 1);
        }
    }
    h.nlev = ns ? *std::max_element(level.begin(), level.end()) + 1 : 0;
    h.lev_ptr.assign(h.nlev + 1, 0);
    for (int32_t J = 0; J < ns; J++) h.lev_ptr[level[J] + 1]++;
    for (int32_t l = 0; l < h.nlev; l++) h.lev_ptr[l + 1] += h.lev_ptr[l];
    h.lev_sn.resize(ns);
    {
        std::vector<int32_t> fill(h.lev_ptr.begin(), h.lev_ptr.end() - 1);
        for (int32_t J = 0; J < ns; J++) h.lev_sn[fill[level[J]]++] = J;
    }

    // 5. contributions: R_J' splits into runs of rows by the supernode they
    //    fall in; each run makes a pair (J' -> J), listed per target in
    //    ascending J' (J' is visited in ascending order)
    std::vector<int32_t> npair(ns + 1, 0);
    std::vector<int64_t> nrel(ns + 1, 0);
    for (int32_t Jp = 0; Jp < ns; Jp++) {
        const int32_t b = h.rptr[Jp], e = h.rptr[Jp + 1];
        for (int32_t t = b; t < e;) {
            const int32_t J = sn_of[h.rows[t]];
            int32_t t2 = t + 1;
            while (t2 < e && sn_of[h.rows[t2]] == J) t2++;
            npair[J + 1]++;
            nrel[J + 1] += e - t;
            t = t2;
        }
    }
    h.cptr.assign(ns + 1, 0);
    std::vector<int64_t> rbase(ns + 1, 0);
    for (int32_t J = 0; J < ns; J++) {
        h.cptr[J + 1] = h.cptr[J] + npair[J + 1];
        rbase[J + 1] = rbase[J] + nrel[J + 1];
    }
    h.npairs = h.cptr[ns];
    if (rbase[ns] >= (int64_t)std::numeric_limits<int32_t>::max())
        return fail(why, "relative row index table exceeds 32-bit offsets");
    h.csrc.resize(h.npairs);
    h.ca0.resize(h.npairs);
    h.ca1.resize(h.npairs);
    h.crel_off.resize(h.npairs);
    h.rel.resize(rbase[ns]);
    {
        std::vector<int32_t> pf(h.cptr.begin(), h.cptr.end() - 1);
        std::vector<int64_t> rf(rbase.begin(), rbase.end() - 1);
        for (int32_t Jp = 0; Jp < ns; Jp++) {
            const int32_t b = h.rptr[Jp], e = h.rptr[Jp + 1];
            for (int32_t t = b; t < e;) {
                const int32_t J = sn_of[h.rows[t]];
                int32_t t2 = t + 1;
                while (t2 < e && sn_of[h.rows[t2]] == J) t2++;
                const int32_t pi = pf[J]++;
                h.csrc[pi] = Jp;
                h.ca0[pi] = t - b;
                h.ca1[pi] = t2 - b;
                h.crel_off[pi] = rf[J];
                const int32_t jj0 = h.j0[J], jj1 = h.j0[J + 1], w = jj1 - jj0;
                const int32_t* R = h.rows.data() + h.rptr[J];
                const int32_t* RE = h.rows.data() + h.rptr[J + 1];
                for (int32_t u = t; u < e; u++) {
                    const int32_t r = h.rows[u];
                    int64_t rr;
                    if (r < jj1) {
                        rr = r - jj0;
                    } else {
                        const int64_t x = find_sorted(R, RE, r);
                        if (x < 0) return fail(why, "a descendant row lies outside its ancestor's panel");
                        rr = w + x;
                    }
                    h.rel[rf[J]++] = (int32_t)rr;
                }
                t = t2;
            }
        }
    }

    // 6. gather tiles and their records. A position that lies in the R of at
    //    least hot_min supernodes (the ground node) has its diagonal
    //    entry summed on its own: its records are dropped from the tile
    //    unless they touch another entry of it, and its terms listed flat.
    //    In a supernode that at least 4 hot_min descendants update (the
    //    root), every entry in the row or the column of such a position
    //    lying in its panels is summed on its own too: all those
    //    descendants hold the position in their R, so the tiles holding its
    //    row or column would otherwise chain all of their records.
    std::vector<int32_t> occ(n, 0);
    for (int32_t r : h.rows) occ[r]++;
    auto hot_pos = [&](int32_t p) { return occ[p] >= hot_min; };
    std::vector<std::vector<int32_t>> hlines(ns);  // hot positions whose lines are hot in J
    for (int32_t J = 0; J < ns; J++) {
        if (h.cptr[J + 1] - h.cptr[J] < 4 * (int64_t)hot_min) continue;
        for (int32_t c = h.j0[J]; c < h.j0[J + 1]; c++)
            if (hot_pos(c)) hlines[J].push_back(c);
        for (int32_t q = h.rptr[J]; q < h.rptr[J + 1]; q++)
            if (hot_pos(h.rows[q])) hlines[J].push_back(h.rows[q]);
        std::sort(hlines[J].begin(), hlines[J].end());
    }
    auto in_lines = [&](int32_t J, int32_t pos) {
        return std::binary_search(hlines[J].begin(), hlines[J].end(), pos);
    };
    std::vector<int64_t> tile_base(ns + 1, 0);
    std::vector<int32_t> nrt(ns), nct(ns), nur(ns), nuc(ns);
    for (int32_t J = 0; J < ns; J++) {
        const int32_t w = h.j0[J + 1] - h.j0[J], hh = h.rptr[J + 1] - h.rptr[J];
        nrt[J] = (w + hh + kSnTile - 1) / kSnTile;
        nct[J] = (w + kSnTile - 1) / kSnTile;
        nur[J] = (w + kSnTile - 1) / kSnTile;
        nuc[J] = (hh + kSnTile - 1) / kSnTile;
        const bool any = h.cptr[J + 1] > h.cptr[J];
        tile_base[J + 1] = tile_base[J] + (any ? (int64_t)nrt[J] * nct[J] + (int64_t)nur[J] * nuc[J] : 0);
    }
    const int64_t ntiles = tile_base[ns];
    if (ntiles >= (int64_t)std::numeric_limits<int32_t>::max()) return fail(why, "too many gather tiles");
    // whether every entry (R_S[i], R_S[k]), i in [i0, i1), k in [k0, k1), of
    // J's panels is summed elsewhere
    auto all_hot = [&](int32_t J, const int32_t* RS, int32_t i0, int32_t i1, int32_t k0, int32_t k1) {
        if (hlines[J].empty())  // only the diagonal of a hot position of J's own columns
            return i1 - i0 == 1 && k1 - k0 == 1 && RS[i0] == RS[k0] && hot_pos(RS[i0]) &&
                   RS[i0] >= h.j0[J] && RS[i0] < h.j0[J + 1];
        bool rows = true, cols = true;
        for (int32_t i = i0; i < i1 && rows; i++) rows = in_lines(J, RS[i]);
        for (int32_t k = k0; k < k1 && cols; k++) cols = in_lines(J, RS[k]);
        return rows || cols;
    };
    // enumerate the records of one target J, calling emit(tile, pair, t0, t1, u0, u1)
    auto records = [&](int32_t J, auto&& emit) {
        const int32_t jj0 = h.j0[J], w = h.j0[J + 1] - jj0;
        for (int32_t pi = h.cptr[J]; pi < h.cptr[J + 1]; pi++) {
            const int32_t S = h.csrc[pi], a0 = h.ca0[pi], a1 = h.ca1[pi];
            const int32_t hS = h.rptr[S + 1] - h.rptr[S];
            const int32_t* RS = h.rows.data() + h.rptr[S];
            const int32_t* rl = h.rel.data() + h.crel_off[pi] - a0;  // rl[t], t in [a0, hS)
            // LP: rows t in [a0, hS) by row tile, columns u in [a0, a1) by column tile
            for (int32_t u0 = a0; u0 < a1;) {
                const int32_t ct = (RS[u0] - jj0) / kSnTile;
                int32_t u1 = u0 + 1;
                while (u1 < a1 && (RS[u1] - jj0) / kSnTile == ct) u1++;
                for (int32_t t0 = a0; t0 < hS;) {
                    const int32_t rt = rl[t0] / kSnTile;
                    int32_t t1 = t0 + 1;
                    while (t1 < hS && rl[t1] / kSnTile == rt) t1++;
                    // a record whose entries are all summed elsewhere (a hot
                    // diagonal, or lines of hot positions) is dropped
                    const bool only_hot = all_hot(J, RS, t0, t1, u0, u1);
                    if (!only_hot) emit(tile_base[J] + (int64_t)ct * nrt[J] + rt, pi, t0, t1, u0, u1);
                    t0 = t1;
                }
                u0 = u1;
            }
            // UP: rows u in [a0, a1) by row tile, columns t in [a1, hS) by column tile
            for (int32_t u0 = a0; u0 < a1;) {
                const int32_t rt = (RS[u0] - jj0) / kSnTile;
                int32_t u1 = u0 + 1;
                while (u1 < a1 && (RS[u1] - jj0) / kSnTile == rt) u1++;
                for (int32_t t0 = a1; t0 < hS;) {
                    const int32_t ct = (rl[t0] - w) / kSnTile;
                    int32_t t1 = t0 + 1;
                    while (t1 < hS && (rl[t1] - w) / kSnTile == ct) t1++;
                    if (!all_hot(J, RS, u0, u1, t0, t1))  // rows R_S(u), columns R_S(t)
                        emit(tile_base[J] + (int64_t)nrt[J] * nct[J] + (int64_t)ct * nur[J] + rt, pi, t0, t1, u0, u1);
                    t0 = t1;
                }
                u0 = u1;
            }
        }
    };
    h.tile_rec.assign(ntiles + 1, 0);
    for (int32_t J = 0; J < ns; J++)
        records(J, [&](int64_t tile, int32_t, int32_t, int32_t, int32_t, int32_t) { h.tile_rec[tile + 1]++; });
    for (int64_t t = 0; t < ntiles; t++) h.tile_rec[t + 1] += h.tile_rec[t];
    h.rec.resize((size_t)h.tile_rec[ntiles] * 5);
    {
        std::vector<int64_t> fill(h.tile_rec.begin(), h.tile_rec.end() - 1);
        for (int32_t J = 0; J < ns; J++)
            records(J, [&](int64_t tile, int32_t pi, int32_t t0, int32_t t1, int32_t u0, int32_t u1) {
                int32_t* r = h.rec.data() + 5 * fill[tile]++;
                r[0] = pi; r[1] = t0; r[2] = t1; r[3] = u0; r[4] = u1;
            });
    }
    // tiles in level order (a tile is {J, kind, r0, c0}); tile ids of J are
    // consecutive, so the level's tiles are the ids of its supernodes
    h.glev_ptr.assign(h.nlev + 1, 0);
    h.gitems.clear();
    std::vector<int64_t> order;  // tile id of each gather item
    for (int32_t l = 0; l < h.nlev; l++) {
        for (int32_t q = h.lev_ptr[l]; q < h.lev_ptr[l + 1]; q++) {
            const int32_t J = h.lev_sn[q];
            for (int64_t t = tile_base[J]; t < tile_base[J + 1]; t++) {
                const int64_t loc = t - tile_base[J];
                const int64_t nlp = (int64_t)nrt[J] * nct[J];
                int32_t kind, r0, c0;
                if (loc < nlp) {
                    kind = 0; c0 = (int32_t)(loc / nrt[J]) * kSnTile; r0 = (int32_t)(loc % nrt[J]) * kSnTile;
                } else {
                    const int64_t u = loc - nlp;
                    kind = 1; c0 = (int32_t)(u / nur[J]) * kSnTile; r0 = (int32_t)(u % nur[J]) * kSnTile;
                }
                h.gitems.push_back(J);
                h.gitems.push_back(kind);
                h.gitems.push_back(r0);
                h.gitems.push_back(c0);
                order.push_back(t);
            }
        }
        h.glev_ptr[l + 1] = (int32_t)(h.gitems.size() / 4);
    }
    // records re-indexed in gather-item order
    {
        std::vector<int64_t> tr(order.size() + 1, 0);
        std::vector<int32_t> rec2;
        rec2.reserve(h.rec.size());
        for (size_t i = 0; i < order.size(); i++) {
            const int64_t t = order[i];
            rec2.insert(rec2.end(), h.rec.begin() + 5 * h.tile_rec[t], h.rec.begin() + 5 * h.tile_rec[t + 1]);
            tr[i + 1] = tr[i] + (h.tile_rec[t + 1] - h.tile_rec[t]);
        }
        h.rec.swap(rec2);
        h.tile_rec.swap(tr);
    }
    // hot entries: every contribution (S ascending, s ascending) as the
    // slots of L_S(row, s) and U_S(s, column); hot_at: (J, kind, row, col)
    // of each in J's panels (kind 0 LP, 1 UP)
    std::vector<std::array<int32_t, 4>> hot_at;
    auto add_hot = [&](int32_t J, int32_t rowpos, int32_t colpos) {
        const int32_t jj0 = h.j0[J], w = h.j0[J + 1] - jj0, hh = h.rptr[J + 1] - h.rptr[J];
        const int32_t* RJ = h.rows.data() + h.rptr[J];
        const bool rin = rowpos >= jj0 && rowpos < jj0 + w, cin = colpos >= jj0 && colpos < jj0 + w;
        int64_t slot;
        std::array<int32_t, 4> at;
        if (cin) {  // LP(row, column)
            const int32_t rr = rin ? rowpos - jj0 : w + (int32_t)find_sorted(RJ, RJ + hh, rowpos);
            slot = h.oL[J] + (int64_t)(colpos - jj0) * (w + hh) + rr;
            at = {J, 0, rr, colpos - jj0};
        } else {  // UP(row, index of the column in R_J)
            const int32_t tc = (int32_t)find_sorted(RJ, RJ + hh, colpos);
            slot = h.oU[J] + (int64_t)tc * w + (rowpos - jj0);
            at = {J, 1, rowpos - jj0, tc};
        }
        const size_t t0 = h.hot_terms.size();
        const int64_t p0 = (int64_t)t0 / 2;
        for (int32_t pi = h.cptr[J]; pi < h.cptr[J + 1]; pi++) {
            const int32_t S = h.csrc[pi];
            const int32_t wS = h.j0[S + 1] - h.j0[S], hS = h.rptr[S + 1] - h.rptr[S];
            const int32_t* RS = h.rows.data() + h.rptr[S];
            const int64_t tr = find_sorted(RS, RS + hS, rowpos), tcS = find_sorted(RS, RS + hS, colpos);
            if (tr < 0 || tcS < 0) continue;
            for (int32_t s2 = 0; s2 < wS; s2++) {
                h.hot_terms.push_back(h.oL[S] + (int64_t)s2 * (wS + hS) + wS + tr);  // L_S(row, s)
                h.hot_terms.push_back(h.oU[S] + tcS * wS + s2);                      // U_S(s, column)
                h.hot_tlev.push_back(level[S]);
            }
        }
        if (h.hot_terms.size() == t0 && !(rowpos == colpos && hot_pos(rowpos))) return;  // no terms: not hot
        h.hot_slot.push_back(slot);
        h.hot_lev.push_back(level[J]);
        h.hot_ptr.push_back(p0);
        hot_at.push_back(at);
    };
    for (int32_t J = 0; J < ns; J++) {
        const int32_t jj0 = h.j0[J], w = h.j0[J + 1] - jj0, hh = h.rptr[J + 1] - h.rptr[J];
        const int32_t* RJ = h.rows.data() + h.rptr[J];
        if (hlines[J].empty()) {  // the diagonals of hot positions among J's columns
            for (int32_t p = jj0; p < jj0 + w; p++)
                if (hot_pos(p)) add_hot(J, p, p);
            continue;
        }
        // every entry in a row or column of a line position, once
        std::vector<std::pair<int32_t, int32_t>> ents;
        for (int32_t p : hlines[J]) {
            const bool pin = p >= jj0 && p < jj0 + w;
            for (int32_t c = jj0; c < jj0 + w; c++) {  // row p and column p within J's columns
                ents.push_back({p, c});
                ents.push_back({c, p});
            }
            if (pin)
                for (int32_t q = 0; q < hh; q++) {  // p's U row and L column over R_J
                    ents.push_back({p, RJ[q]});
                    ents.push_back({RJ[q], p});
                }
        }
        std::sort(ents.begin(), ents.end());
        ents.erase(std::unique(ents.begin(), ents.end()), ents.end());
        for (const auto& e : ents) {
            const bool rin = e.first >= jj0 && e.first < jj0 + w, cin = e.second >= jj0 && e.second < jj0 + w;
            if (!rin && !cin) continue;  // an ancestor's entry
            add_hot(J, e.first, e.second);
        }
    }
    h.hot_ptr.push_back((int64_t)h.hot_terms.size() / 2);
    // records in the kernel's form, and the hot entries of every tile
    {
        const int64_t nrec = h.tile_rec.back();
        h.frec.resize((size_t)nrec);
        h.tile_hot_ptr.assign(h.gitems.size() / 4 + 1, 0);
        // hot entries by supernode
        std::vector<std::vector<int32_t>> hot_of(ns);
        for (size_t e = 0; e < hot_at.size(); e++) hot_of[hot_at[e][0]].push_back((int32_t)e);
        for (size_t it = 0; it < h.gitems.size() / 4; it++) {
            const int32_t J = h.gitems[4 * it], kind = h.gitems[4 * it + 1], r0 = h.gitems[4 * it + 2],
                          c0 = h.gitems[4 * it + 3];
            const int32_t jj0 = h.j0[J], w = h.j0[J + 1] - jj0;
            for (int64_t ri = h.tile_rec[it]; ri < h.tile_rec[it + 1]; ri++) {
                const int32_t* r = h.rec.data() + 5 * ri;
                const int32_t pi = r[0], t0 = r[1], t1 = r[2], u0 = r[3], u1 = r[4];
                const int32_t S = h.csrc[pi], a0 = h.ca0[pi];
                const int32_t wS = h.j0[S + 1] - h.j0[S], hS = h.rptr[S + 1] - h.rptr[S];
                const int32_t* RS = h.rows.data() + h.rptr[S];
                const int32_t* rl = h.rel.data() + h.crel_off[pi] - a0;
                SnRec& q = h.frec[(size_t)ri];
                q.rowmask = q.colmask = 0;
                // LP tiles: rows from R_S(t), columns from R_S(u); UP tiles the other way round
                const int32_t lfirst = kind == 0 ? t0 : u0, ufirst = kind == 0 ? u0 : t0;
                for (int32_t t = t0; t < t1; t++) {
                    const int32_t b = kind == 0 ? rl[t] - r0 : rl[t] - w - c0;
                    (kind == 0 ? q.rowmask : q.colmask) |= 1ull << b;
                }
                for (int32_t u = u0; u < u1; u++) {
                    const int32_t b = kind == 0 ? RS[u] - jj0 - c0 : RS[u] - jj0 - r0;
                    (kind == 0 ? q.colmask : q.rowmask) |= 1ull << b;
                }
                q.lbase = h.oL[S] + wS + lfirst;
                q.ubase = h.oU[S] + (int64_t)ufirst * wS;
                q.ws = wS;
                q.lds = wS + hS;
            }
            for (int32_t e : hot_of[J]) {
                const auto& a = hot_at[e];
                if (a[1] == kind && a[2] >= r0 && a[2] < r0 + kSnTile && a[3] >= c0 && a[3] < c0 + kSnTile)
                    h.tile_hot.push_back((a[2] - r0) | ((a[3] - c0) << 8));
            }
            h.tile_hot_ptr[it + 1] = (int32_t)h.tile_hot.size();
        }
        h.qrec_ptr.assign(h.gitems.size() + 1, 0);  // 4 quarters per item
        for (size_t it = 0; it < h.gitems.size() / 4; it++)
            for (int32_t qq = 0; qq < 4; qq++) {
                for (int64_t ri = h.tile_rec[it]; ri < h.tile_rec[it + 1]; ri++)
                    if ((h.frec[(size_t)ri].colmask >> (16 * qq)) & 0xFFFFull) {
                        if (ri > std::numeric_limits<int32_t>::max()) return fail(why, "too many gather records");
                        h.qrec.push_back((int32_t)ri);
                    }
                h.qrec_ptr[4 * it + qq + 1] = (int64_t)h.qrec.size();
            }
    }

    // 7. scatter of A: runs of duplicate rows of each column, in the column
    //    order of the pivot steps, onto the panel slot of (position, step)
    auto slot_of = [&](int32_t p, int32_t k) -> int64_t {
        const int32_t J = sn_of[k];
        const int32_t jj0 = h.j0[J], jj1 = h.j0[J + 1], w = jj1 - jj0;
        const int32_t hh = h.rptr[J + 1] - h.rptr[J];
        if (p >= jj0) {
            int64_t rr;
            if (p < jj1) {
                rr = p - jj0;
            } else {
                rr = find_sorted(h.rows.data() + h.rptr[J], h.rows.data() + h.rptr[J + 1], p);
                if (rr < 0) return -1;
                rr += w;
            }
            return h.oL[J] + (int64_t)(k - jj0) * (w + hh) + rr;
        }
        const int32_t Jp = sn_of[p];
        const int32_t wp = h.j0[Jp + 1] - h.j0[Jp];
        const int64_t c = find_sorted(h.rows.data() + h.rptr[Jp], h.rows.data() + h.rptr[Jp + 1], k);
        if (c < 0) return -1;
        return h.oU[Jp] + c * wp + (p - h.j0[Jp]);
    };
    for (i64 k = 0; k < n; k++) {
        const i64 oc = s.col_perm_inv[k];
        i64 q = ap[oc];
        while (q < ap[oc + 1]) {
            const i64 orow = ai[q];
            i64 q2 = q + 1;
            while (q2 < ap[oc + 1] && ai[q2] == orow) q2++;
            if (orow < n) {
                // A(orow, oc) at (position of orow, k); for A^T at (k, position of orow)
                const int32_t pr = (int32_t)s.col_perm[orow];
                const int64_t sl = transpose ? slot_of((int32_t)k, pr) : slot_of(pr, (int32_t)k);
                if (sl >= 0) {
                    if (q > std::numeric_limits<int32_t>::max() - 1) return fail(why, "nnz exceeds 32-bit offsets");
                    h.sc_src.push_back((int32_t)q);
                    h.sc_cnt.push_back((int32_t)(q2 - q));
                    h.sc_slot.push_back(sl);
                }
            }
            q = q2;
        }
    }
    h.perm.resize(n);
    for (i64 i = 0; i < n; i++) h.perm[i] = (int32_t)s.col_perm[i];
    h.lrow_ptr.assign(n + 1, 0);
    for (int32_t r : h.rows) h.lrow_ptr[r + 1]++;
    for (i64 i = 0; i < n; i++) h.lrow_ptr[i + 1] += h.lrow_ptr[i];
    h.lrow_sn.resize(h.rows.size());
    h.lrow_t.resize(h.rows.size());
    {
        std::vector<int64_t> fill(h.lrow_ptr.begin(), h.lrow_ptr.end() - 1);
        for (int32_t S = 0; S < ns; S++)
            for (int32_t t = 0; t < h.rptr[S + 1] - h.rptr[S]; t++) {
                const int64_t at = fill[h.rows[h.rptr[S] + t]]++;
                h.lrow_sn[at] = S;
                h.lrow_t[at] = t;
            }
    }

    // 8. algorithmic work: every U(j, k), j < k, applies L(:, j); those with
    //    j and k in different supernodes are the gather's
    double upd = 0.0, gupd = 0.0;
    {
        std::vector<int32_t> sn_of(n);
        for (int32_t S = 0; S < ns; S++)
            for (i64 j = h.j0[S]; j < h.j0[S + 1]; j++) sn_of[j] = S;
        for (i64 k = 0; k < n; k++)
            for (i64 q = ucp[k]; q < ucp[k + 1] - 1; q++) {
                const double c = (double)(lcp[uri[q] + 1] - lcp[uri[q]]);
                upd += c;
                if (sn_of[uri[q]] != sn_of[k]) gupd += c;
            }
    }
    h.updates = upd;
    h.gather_updates = gupd;
    return true;
}

// ---- host execution of the schedule ------------------------------------------

namespace {

struct Panel {
    const SnHost& h;
    int32_t J, jj0, w, hh;
    int64_t oL, oU;
    Panel(const SnHost& hs, int32_t j)
        : h(hs), J(j), jj0(hs.j0[j]), w(hs.j0[j + 1] - hs.j0[j]), hh(hs.rptr[j + 1] - hs.rptr[j]),
          oL(hs.oL[j]), oU(hs.oU[j]) {}
    int64_t lp(int64_t r, int64_t c) const { return oL + c * (w + hh) + r; }
    int64_t up(int64_t r, int64_t c) const { return oU + c * w + r; }
};

}  // namespace

void sn_host_factor(const SnHost& h, const double* ax, std::vector<double>& v, std::vector<double>& ud) {
    v.assign((size_t)h.nvals, 0.0);
    ud.assign(h.n, 0.0);
    for (size_t s = 0; s < h.sc_src.size(); s++) {
        double x = 0.0;
        for (int32_t q = 0; q < h.sc_cnt[s]; q++) x += ax[h.sc_src[s] + q];
        v[h.sc_slot[s]] = x;
    }
    for (int32_t l = 0; l < h.nlev; l++) {
        // gather: every tile applies its records in order; each record is a
        // block of entries, each entry's terms in ascending order
        for (int32_t it = h.glev_ptr[l]; it < h.glev_ptr[l + 1]; it++) {
            const Panel P(h, h.gitems[4 * it]);
            const int32_t kind = h.gitems[4 * it + 1], r0 = h.gitems[4 * it + 2], c0 = h.gitems[4 * it + 3];
            const int64_t base = kind == 0 ? P.lp(r0, c0) : P.up(r0, c0);
            const int64_t ld = kind == 0 ? P.w + P.hh : P.w;
            for (int64_t ri = h.tile_rec[it]; ri < h.tile_rec[it + 1]; ri++) {
                const SnRec& q = h.frec[(size_t)ri];
                int32_t i = 0;
                for (int32_t r = 0; r < kSnTile; r++) {
                    if (!((q.rowmask >> r) & 1)) continue;
                    int32_t k = 0;
                    for (int32_t c = 0; c < kSnTile; c++) {
                        if (!((q.colmask >> c) & 1)) continue;
                        bool hot = false;
                        for (int32_t e = h.tile_hot_ptr[it]; e < h.tile_hot_ptr[it + 1]; e++)
                            hot |= h.tile_hot[e] == (r | (c << 8));
                        const int64_t slot = base + (int64_t)c * ld + r;
                        if (!hot) {
                            double x = v[slot];
                            for (int32_t s2 = 0; s2 < q.ws; s2++) {
                                const double uu = v[q.ubase + (int64_t)k * q.ws + s2];
                                if (uu != 0.0) x -= v[q.lbase + (int64_t)s2 * q.lds + i] * uu;
                            }
                            v[slot] = x;
                        }
                        k++;
                    }
                    i++;
                }
            }
        }
        for (size_t e = 0; e < h.hot_slot.size(); e++) {
            if (h.hot_lev[e] != l) continue;
            double x = v[h.hot_slot[e]];
            for (int64_t q = h.hot_ptr[e]; q < h.hot_ptr[e + 1]; q++) {
                const double uu = v[h.hot_terms[2 * q + 1]];
                if (uu != 0.0) x -= v[h.hot_terms[2 * q]] * uu;
            }
            v[h.hot_slot[e]] = x;
        }
        for (int32_t q = h.lev_ptr[l]; q < h.lev_ptr[l + 1]; q++) {
            const Panel P(h, h.lev_sn[q]);
            // J's own columns, right-looking
            const int32_t W = P.w + P.hh;
            for (int32_t t = 0; t < P.w; t++) {
                double d = v[P.lp(t, t)];
                if (std::fabs(d) < 1e-14) d = d >= 0.0 ? 1e-14 : -1e-14;
                v[P.lp(t, t)] = d;
                ud[P.jj0 + t] = d;
                for (int32_t r = t + 1; r < W; r++) v[P.lp(r, t)] = v[P.lp(r, t)] / d;
                for (int32_t c = t + 1; c < P.w; c++) {
                    const double u = v[P.lp(t, c)];
                    if (u == 0.0) continue;
                    for (int32_t r = t + 1; r < W; r++) v[P.lp(r, c)] -= v[P.lp(r, t)] * u;
                }
                for (int32_t c = 0; c < P.hh; c++) {
                    const double u = v[P.up(t, c)];
                    if (u == 0.0) continue;
                    for (int32_t r = t + 1; r < P.w; r++) v[P.up(r, c)] -= v[P.lp(r, t)] * u;
                }
            }
        }
    }
}

void sn_host_solve(const SnHost& h, const std::vector<double>& v, const std::vector<double>& ud,
                   const double* rhs, double* xout) {
    std::vector<double> y(h.n);
    for (int32_t i = 0; i < h.n; i++) y[h.perm[i]] = rhs[i];
    for (int32_t l = 0; l < h.nlev; l++)
        for (int32_t q = h.lev_ptr[l]; q < h.lev_ptr[l + 1]; q++) {
            const Panel P(h, h.lev_sn[q]);
            for (int32_t pi = h.cptr[P.J]; pi < h.cptr[P.J + 1]; pi++) {
                const Panel S(h, h.csrc[pi]);
                for (int32_t ci = h.ca0[pi]; ci < h.ca1[pi]; ci++) {
                    const int32_t row = h.rows[h.rptr[S.J] + ci];
                    double x = y[row];
                    for (int32_t t = 0; t < S.w; t++) x -= v[S.lp(S.w + ci, t)] * y[S.jj0 + t];
                    y[row] = x;
                }
            }
            for (int32_t t = 0; t < P.w; t++)
                for (int32_t r = t + 1; r < P.w; r++) y[P.jj0 + r] -= v[P.lp(r, t)] * y[P.jj0 + t];
        }
    for (int32_t l = h.nlev - 1; l >= 0; l--)
        for (int32_t q = h.lev_ptr[l]; q < h.lev_ptr[l + 1]; q++) {
            const Panel P(h, h.lev_sn[q]);
            const int32_t* R = h.rows.data() + h.rptr[P.J];
            for (int32_t r = 0; r < P.w; r++) {
                double x = y[P.jj0 + r];
                for (int32_t c = P.hh - 1; c >= 0; c--) x -= v[P.up(r, c)] * y[R[c]];
                y[P.jj0 + r] = x;
            }
            for (int32_t t = P.w - 1; t >= 0; t--) {
                y[P.jj0 + t] /= ud[P.jj0 + t];
                const double xt = y[P.jj0 + t];
                for (int32_t r = 0; r < t; r++) y[P.jj0 + r] -= v[P.lp(r, t)] * xt;
            }
        }
    for (int32_t i = 0; i < h.n; i++) xout[i] = y[h.perm[i]];
}

void sn_to_csc(const SnHost& h, const BlockSymbolic& s, const std::vector<double>& v,
               std::vector<double>& lx, std::vector<double>& ux) {
    const i64 n = s.n;
    std::vector<int32_t> sn_of(n);
    for (int32_t J = 0; J < h.ns; J++)
        for (int32_t c = h.j0[J]; c < h.j0[J + 1]; c++) sn_of[c] = J;
    lx.assign(s.lri.size(), 0.0);
    ux.assign(s.uri.size(), 0.0);
    for (i64 k = 0; k < n; k++) {
        const Panel P(h, sn_of[k]);
        const int32_t* R = h.rows.data() + h.rptr[P.J];
        for (i64 q = s.lcp[k]; q < s.lcp[k + 1]; q++) {
            const int32_t r = (int32_t)s.lri[q];
            const int64_t rr = r < P.jj0 + P.w ? r - P.jj0 : P.w + find_sorted(R, R + P.hh, r);
            lx[q] = v[P.lp(rr, k - P.jj0)];
        }
        for (i64 q = s.ucp[k]; q < s.ucp[k + 1]; q++) {
            const int32_t r = (int32_t)s.uri[q];
            if (r >= P.jj0) {
                ux[q] = v[P.lp(r - P.jj0, k - P.jj0)];
            } else {
                const Panel S(h, sn_of[r]);
                const int32_t* RS = h.rows.data() + h.rptr[S.J];
                ux[q] = v[S.up(r - S.jj0, find_sorted(RS, RS + S.hh, (int32_t)k))];
            }
        }
    }
}

}  // namespace rsgpu
