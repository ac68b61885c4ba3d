// plan.cpp - builds the flat device plan of a factorization; see plan.hpp.
#include "plan.hpp"

#include <algorithm>
#include <stdexcept>

namespace rsgpu {

namespace {

void group_by_level(const std::vector<int32_t>& lvl, const std::vector<int32_t>& items,
                    std::vector<int32_t>& ptr, std::vector<int32_t>& out) {
    int32_t nlev = 0;
    for (int32_t it : items) nlev = std::max(nlev, lvl[it] + 1);
    std::vector<int32_t> cnt(nlev + 1, 0);
    for (int32_t it : items) cnt[lvl[it] + 1]++;
    for (int32_t l = 0; l < nlev; l++) cnt[l + 1] += cnt[l];
    size_t base = out.size();
    out.resize(base + items.size());
    std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
    for (int32_t it : items) out[base + fill[lvl[it]]++] = it;
    int32_t pbase = ptr.empty() ? 0 : ptr.back();
    if (ptr.empty()) ptr.push_back(0);
    for (int32_t l = 0; l < nlev; l++) ptr.push_back(pbase + cnt[l + 1]);
    (void)pbase;
}

int32_t chk32(i64 v) {
    if (v < 0 || v > INT32_MAX) throw std::runtime_error("plan index exceeds int32 range");
    return (int32_t)v;
}

}  // namespace

void build_plan(const HostFactor& f, const i64* ap, const i64* ai, PlanHost& p) {
    const i64 n = f.n;
    p = PlanHost();
    p.n = chk32(n);
    const i64 nb = (i64)f.sym.size();
    p.nblk = (int32_t)nb;
    std::vector<i64> bstart(nb + 1, 0);
    if (f.has_btf) {
        for (i64 b = 0; b <= nb; b++) bstart[b] = f.btf.block_ptr[b];
    } else {
        bstart[1] = n;
    }
    for (i64 b = 0; b <= nb; b++) p.blk_ptr.push_back(chk32(bstart[b]));
    // block of each position
    std::vector<int32_t> blk_of(n, 0);
    for (i64 b = 0; b < nb; b++)
        for (i64 g = bstart[b]; g < bstart[b + 1]; g++) blk_of[g] = (int32_t)b;

    // 1. position-space L/U CSC
    p.lcp.assign(n + 1, 0);
    p.ucp.assign(n + 1, 0);
    p.colflag.assign(n, 0);
    for (i64 b = 0; b < nb; b++) {
        const BlockSymbolic& s = f.sym[b];
        const i64 bs = bstart[b];
        const bool singleton = (s.n == 1) && !f.sparselu;
        for (i64 k = 0; k < s.n; k++) {
            i64 g = bs + k;
            if (singleton) {
                p.colflag[g] = COL_SINGLETON;
                p.uri.push_back(chk32(g));  // diagonal only
            } else {
                for (i64 q = s.lcp[k]; q < s.lcp[k + 1]; q++) p.lri.push_back(chk32(bs + s.lri[q]));
                for (i64 q = s.ucp[k]; q < s.ucp[k + 1]; q++) p.uri.push_back(chk32(bs + s.uri[q]));
            }
            p.lcp[g + 1] = (int32_t)p.lri.size();
            p.ucp[g + 1] = (int32_t)p.uri.size();
        }
    }
    p.nl = (int32_t)p.lri.size();
    p.nu = (int32_t)p.uri.size();

    // 2. factor level schedule: column g waits for every j in U(:,g), j < g
    std::vector<int32_t> flev(n, 0), cols(n);
    for (i64 g = 0; g < n; g++) {
        int32_t lv = 0;
        for (int32_t q = p.ucp[g]; q < p.ucp[g + 1] - 1; q++) lv = std::max(lv, flev[p.uri[q]] + 1);
        flev[g] = lv;
        cols[g] = (int32_t)g;
    }
    group_by_level(flev, cols, p.flev_ptr, p.flev_cols);
    p.nflev = (int32_t)p.flev_ptr.size() - 1;
    for (int32_t l = 0; l < p.nflev; l++)
        p.max_flev_width = std::max(p.max_flev_width, p.flev_ptr[l + 1] - p.flev_ptr[l]);

    // 3. scatter map: runs of duplicate (row, col) entries -> target position,
    //    dropping targets outside the column's pattern (their values are
    //    cleared unread by the reference).
    std::vector<int32_t> inpat(n, -1);
    p.sc_ptr.assign(n + 1, 0);
    for (i64 b = 0; b < nb; b++) {
        const BlockSymbolic& s = f.sym[b];
        const BlockNumeric& m = f.num[b];
        const i64 bs = bstart[b], be = bstart[b + 1];
        for (i64 k = 0; k < s.n; k++) {
            const i64 g = bs + k;
            for (int32_t q = p.lcp[g]; q < p.lcp[g + 1]; q++) inpat[p.lri[q]] = (int32_t)g;
            for (int32_t q = p.ucp[g]; q < p.ucp[g + 1]; q++) inpat[p.uri[q]] = (int32_t)g;
            i64 oc;  // original column
            if (f.has_btf) {
                i64 lc = s.col_perm_inv[k];
                oc = f.btf.col_perm[bs + lc];
            } else {
                oc = s.col_perm_inv[k];
            }
            i64 q = ap[oc];
            while (q < ap[oc + 1]) {
                i64 orow = ai[q];
                i64 q2 = q + 1;
                while (q2 < ap[oc + 1] && ai[q2] == orow) q2++;
                if (orow < n) {
                    i64 lr = -1;
                    if (f.has_btf) {
                        i64 grn = f.btf.row_perm_inv[orow];
                        if (grn >= bs && grn < be) lr = grn - bs;
                    } else {
                        lr = orow;
                    }
                    if (lr >= 0) {
                        i64 tgt = bs + m.row_perm_inv[s.col_perm[lr]];
                        if (inpat[tgt] == (int32_t)g) {
                            p.sc_src.push_back(chk32(q));
                            p.sc_cnt.push_back((int32_t)(q2 - q));
                            p.sc_dst.push_back((int32_t)tgt);
                        }
                    }
                }
                q = q2;
            }
            p.sc_ptr[g + 1] = (int32_t)p.sc_src.size();
        }
    }
    p.nsc = (int32_t)p.sc_src.size();

    // 4. solve permutations
    p.perm_in.assign(n, 0);
    p.perm_out.assign(n, 0);
    for (i64 i = 0; i < n; i++) {
        if (f.has_btf) {
            i64 r = f.btf.row_perm_inv[i];
            i64 b = blk_of[r];
            const BlockSymbolic& s = f.sym[b];
            const BlockNumeric& m = f.num[b];
            i64 bs = bstart[b];
            p.perm_in[i] = (int32_t)(bs + m.row_perm_inv[s.col_perm[r - bs]]);
            i64 c = f.btf.col_perm_inv[i];
            i64 b2 = blk_of[c];
            p.perm_out[i] = (int32_t)(bstart[b2] + f.sym[b2].col_perm[c - bstart[b2]]);
        } else {
            const BlockSymbolic& s = f.sym[0];
            const BlockNumeric& m = f.num[0];
            p.perm_in[i] = (int32_t)m.row_perm_inv[s.col_perm[i]];
            p.perm_out[i] = (int32_t)s.col_perm[i];
        }
    }

    // 5. row-oriented L (cols ascending) and U without diagonal (cols descending)
    p.lrp.assign(n + 1, 0);
    p.urp.assign(n + 1, 0);
    for (int32_t q = 0; q < p.nl; q++) p.lrp[p.lri[q] + 1]++;
    for (i64 g = 0; g < n; g++)
        for (int32_t q = p.ucp[g]; q < p.ucp[g + 1] - 1; q++) p.urp[p.uri[q] + 1]++;
    for (i64 i = 0; i < n; i++) {
        p.lrp[i + 1] += p.lrp[i];
        p.urp[i + 1] += p.urp[i];
    }
    p.lrc.resize(p.lrp[n]);
    p.lrv.resize(p.lrp[n]);
    p.urc.resize(p.urp[n]);
    p.urv.resize(p.urp[n]);
    {
        std::vector<int32_t> fl(p.lrp.begin(), p.lrp.end() - 1);
        for (i64 g = 0; g < n; g++)
            for (int32_t q = p.lcp[g]; q < p.lcp[g + 1]; q++) {
                int32_t r = p.lri[q];
                p.lrc[fl[r]] = (int32_t)g;
                p.lrv[fl[r]++] = q;
            }
        std::vector<int32_t> fu(p.urp.begin(), p.urp.end() - 1);
        for (i64 g = n - 1; g >= 0; g--)
            for (int32_t q = p.ucp[g]; q < p.ucp[g + 1] - 1; q++) {
                int32_t r = p.uri[q];
                p.urc[fu[r]] = (int32_t)g;
                p.urv[fu[r]++] = q;
            }
    }

    // 6. per-block level schedules of the two triangular solves
    std::vector<int32_t> llev(n, 0), ulev(n, 0), rows;
    for (i64 b = 0; b < nb; b++) {
        const i64 bs = bstart[b], be = bstart[b + 1];
        rows.clear();
        for (i64 r = bs; r < be; r++) {
            int32_t lv = 0;
            for (int32_t q = p.lrp[r]; q < p.lrp[r + 1]; q++) lv = std::max(lv, llev[p.lrc[q]] + 1);
            llev[r] = lv;
            rows.push_back((int32_t)r);
        }
        p.lsl_blk.push_back(p.lsl_ptr.empty() ? 0 : (int32_t)p.lsl_ptr.size() - 1);
        group_by_level(llev, rows, p.lsl_ptr, p.lsl_rows);
        rows.clear();
        for (i64 r = be - 1; r >= bs; r--) {
            int32_t lv = 0;
            for (int32_t q = p.urp[r]; q < p.urp[r + 1]; q++) lv = std::max(lv, ulev[p.urc[q]] + 1);
            ulev[r] = lv;
        }
        for (i64 r = bs; r < be; r++) rows.push_back((int32_t)r);
        p.usl_blk.push_back(p.usl_ptr.empty() ? 0 : (int32_t)p.usl_ptr.size() - 1);
        group_by_level(ulev, rows, p.usl_ptr, p.usl_rows);
    }
    p.lsl_blk.push_back((int32_t)p.lsl_ptr.size() - 1);
    p.usl_blk.push_back((int32_t)p.usl_ptr.size() - 1);
    p.n_lsl = (int32_t)p.lsl_ptr.size() - 1;
    p.n_usl = (int32_t)p.usl_ptr.size() - 1;
    // rows of a level are independent: long rows go first and are solved by a
    // whole wavefront (products in parallel, the ordered sum on one lane)
    auto long_first = [&](const std::vector<int32_t>& ptr, std::vector<int32_t>& rows,
                          const std::vector<int32_t>& rp, std::vector<int32_t>& nlong) {
        nlong.assign(ptr.size() ? ptr.size() - 1 : 0, 0);
        for (size_t l = 0; l + 1 < ptr.size(); l++) {
            auto b = rows.begin() + ptr[l], e = rows.begin() + ptr[l + 1];
            auto mid = std::stable_partition(b, e, [&](int32_t r) { return rp[r + 1] - rp[r] >= kSolveLongRow; });
            nlong[l] = (int32_t)(mid - b);
        }
    };
    long_first(p.lsl_ptr, p.lsl_rows, p.lrp, p.lsl_nlong);
    long_first(p.usl_ptr, p.usl_rows, p.urp, p.usl_nlong);

    // 7. off-diagonal BTF entries per row position, reference order
    p.odp.assign(n + 1, 0);
    p.blk_has_od.assign(nb, 0);
    if (f.has_btf) {
        struct OD { int32_t pos, ax, src; };
        std::vector<OD> ods;
        for (i64 br = 0; br < nb; br++) {
            const i64 rs = bstart[br], re = bstart[br + 1];
            for (i64 bc = br + 1; bc < nb; bc++) {
                const i64 cs = bstart[bc], ce = bstart[bc + 1];
                for (i64 lc = 0; lc < ce - cs; lc++) {
                    i64 gco = f.btf.col_perm[cs + lc];
                    int32_t src = (int32_t)(cs + f.sym[bc].col_perm[lc]);
                    for (i64 q = ap[gco]; q < ap[gco + 1]; q++) {
                        i64 gro = ai[q];
                        if (gro >= n) continue;
                        i64 grn = f.btf.row_perm_inv[gro];
                        if (grn >= rs && grn < re) {
                            int32_t pos = (int32_t)(rs + f.num[br].row_perm_inv[f.sym[br].col_perm[grn - rs]]);
                            ods.push_back({pos, chk32(q), src});
                            p.blk_has_od[br] = 1;
                        }
                    }
                }
            }
        }
        for (auto& o : ods) p.odp[o.pos + 1]++;
        for (i64 i = 0; i < n; i++) p.odp[i + 1] += p.odp[i];
        p.odax.resize(ods.size());
        p.odsrc.resize(ods.size());
        std::vector<int32_t> fo(p.odp.begin(), p.odp.end() - 1);
        for (auto& o : ods) {  // stable: keeps reference order per row
            p.odax[fo[o.pos]] = o.ax;
            p.odsrc[fo[o.pos]++] = o.src;
        }
    }
    p.nod = (int32_t)p.odax.size();

    // 8. algorithmic work
    double ff = 0;
    for (i64 g = 0; g < n; g++) {
        for (int32_t q = p.ucp[g]; q < p.ucp[g + 1] - 1; q++) {
            int32_t j = p.uri[q];
            ff += 2.0 * (p.lcp[j + 1] - p.lcp[j]);
        }
        ff += p.lcp[g + 1] - p.lcp[g];
    }
    p.refactor_flops = ff;
    p.solve_flops = 2.0 * ((double)p.nl + (double)p.nu - (double)n) + (double)n;
}

void gather_values(const HostFactor& f, std::vector<double>& lx, std::vector<double>& ux,
                   std::vector<double>& ud) {
    lx.clear();
    ux.clear();
    ud.clear();
    for (size_t b = 0; b < f.sym.size(); b++) {
        const BlockNumeric& m = f.num[b];
        if (f.sym[b].n == 1) {
            ux.push_back(m.udiag[0]);
            ud.push_back(m.udiag[0]);
            continue;
        }
        lx.insert(lx.end(), m.lx.begin(), m.lx.end());
        ux.insert(ux.end(), m.ux.begin(), m.ux.end());
        ud.insert(ud.end(), m.udiag.begin(), m.udiag.end());
    }
}

void append_plan(const PlanHost& p, std::vector<int32_t>& pool, PlanDev& d) {
    auto put = [&](const std::vector<int32_t>& v) -> int32_t {
        size_t off = pool.size();
        if (off + v.size() > (size_t)INT32_MAX) throw std::runtime_error("plan pool too large");
        pool.insert(pool.end(), v.begin(), v.end());
        pool.push_back(0);  // keeps every offset valid for empty arrays
        return (int32_t)off;
    };
    d.n = p.n;
    d.nblk = p.nblk;
    d.nflev = p.nflev;
    d.nl = p.nl;
    d.nu = p.nu;
    d.nsc = p.nsc;
    d.nod = p.nod;
    d.o_lcp = put(p.lcp);
    d.o_lri = put(p.lri);
    d.o_ucp = put(p.ucp);
    d.o_uri = put(p.uri);
    d.o_colflag = put(p.colflag);
    d.o_flev_ptr = put(p.flev_ptr);
    d.o_flev_cols = put(p.flev_cols);
    d.o_sc_ptr = put(p.sc_ptr);
    d.o_sc_src = put(p.sc_src);
    d.o_sc_cnt = put(p.sc_cnt);
    d.o_sc_dst = put(p.sc_dst);
    d.o_perm_in = put(p.perm_in);
    d.o_perm_out = put(p.perm_out);
    d.o_lrp = put(p.lrp);
    d.o_lrc = put(p.lrc);
    d.o_lrv = put(p.lrv);
    d.o_urp = put(p.urp);
    d.o_urc = put(p.urc);
    d.o_urv = put(p.urv);
    d.o_blk_ptr = put(p.blk_ptr);
    d.o_lsl_blk = put(p.lsl_blk);
    d.o_lsl_ptr = put(p.lsl_ptr);
    d.o_lsl_rows = put(p.lsl_rows);
    d.o_usl_blk = put(p.usl_blk);
    d.o_usl_ptr = put(p.usl_ptr);
    d.o_usl_rows = put(p.usl_rows);
    d.o_odp = put(p.odp);
    d.o_odax = put(p.odax);
    d.o_odsrc = put(p.odsrc);
    d.o_blk_has_od = put(p.blk_has_od);
    d.o_lsl_nlong = put(p.lsl_nlong);
    d.o_usl_nlong = put(p.usl_nlong);
}

// Every solve level in execution order (blocks from the last; per block its
// L levels, then its U levels) as int4 {s0, k0, div, 0} plus a sentinel: the
// level's entries are the slots [s0, s0') and its rows the records [k0, k0').
// Each row record is int4 {row, slot begin, slot end, 0}; a row's slots hold
// its entries in the row-ordered sweep's order (lrp/urp), each with its
// column (sq_col) and its value's offset from o_lx (sq_src: L position q, or
// nl + 1 + U position q). L rows without entries leave T unchanged and get
// no record; L levels left empty are dropped.
void append_level_records(const PlanHost& p, std::vector<int32_t>& pool, PlanDev& d) {
    d.sq_nlev = d.sq_maxent = 0;
    d.o_sq_lev = d.o_sq_row = d.o_sq_col = d.o_sq_src = 0;
    for (int32_t v : p.blk_has_od)
        if (v) return;
    std::vector<int32_t> lev, row, col, src;
    int32_t maxent = 0;
    auto add = [&](const std::vector<int32_t>& ptr, const std::vector<int32_t>& rows,
                   const std::vector<int32_t>& rp, const std::vector<int32_t>& rc,
                   const std::vector<int32_t>& rv, int32_t l, int32_t div) {
        const int32_t s0 = (int32_t)col.size(), k0 = (int32_t)(row.size() / 4);
        for (int32_t q = ptr[l]; q < ptr[l + 1]; q++) {
            const int32_t r = rows[q];
            if (!div && rp[r] == rp[r + 1]) continue;
            const int32_t sb = (int32_t)col.size();
            for (int32_t e = rp[r]; e < rp[r + 1]; e++) {
                col.push_back(rc[e]);
                src.push_back(div ? p.nl + 1 + rv[e] : rv[e]);
            }
            row.insert(row.end(), {r, sb, (int32_t)col.size(), 0});
        }
        if ((int32_t)(row.size() / 4) == k0) return;
        lev.insert(lev.end(), {s0, k0, div, 0});
        maxent = std::max(maxent, (int32_t)col.size() - s0);
    };
    for (int32_t b = p.nblk - 1; b >= 0; b--) {
        for (int32_t l = p.lsl_blk[b]; l < p.lsl_blk[b + 1]; l++)
            add(p.lsl_ptr, p.lsl_rows, p.lrp, p.lrc, p.lrv, l, 0);
        for (int32_t l = p.usl_blk[b]; l < p.usl_blk[b + 1]; l++)
            add(p.usl_ptr, p.usl_rows, p.urp, p.urc, p.urv, l, 1);
    }
    if (lev.empty()) return;
    const int32_t nlev = (int32_t)(lev.size() / 4);
    lev.insert(lev.end(), {(int32_t)col.size(), (int32_t)(row.size() / 4), 0, 0});
    auto put16 = [&](const std::vector<int32_t>& v) -> int32_t {
        while (pool.size() & 3) pool.push_back(0);
        const size_t off = pool.size();
        if (off + v.size() + 4 > (size_t)INT32_MAX) throw std::runtime_error("plan pool too large");
        pool.insert(pool.end(), v.begin(), v.end());
        pool.insert(pool.end(), {0, 0, 0, 0});  // index 0 stays valid for empty arrays
        return (int32_t)off;
    };
    d.o_sq_lev = put16(lev);
    d.o_sq_row = put16(row);
    d.o_sq_col = put16(col);
    d.o_sq_src = put16(src);
    d.sq_nlev = nlev;
    d.sq_maxent = maxent;
}

void lt_levels(const PlanHost& p, int64_t n, std::vector<int32_t>& ptr, std::vector<int32_t>& cols) {
    std::vector<int32_t> lev(n, 0);
    int32_t nlev = 0;
    for (int64_t k = n - 1; k >= 0; k--) {
        int32_t lv = 0;
        for (int32_t q = p.lcp[k]; q < p.lcp[k + 1]; q++) lv = std::max(lv, lev[p.lri[q]] + 1);
        lev[k] = lv;
        nlev = std::max(nlev, lv + 1);
    }
    ptr.assign(nlev + 1, 0);
    for (int64_t k = 0; k < n; k++) ptr[lev[k] + 1]++;
    for (int32_t l = 0; l < nlev; l++) ptr[l + 1] += ptr[l];
    cols.resize(n);
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t k = 0; k < n; k++) cols[fill[lev[k]]++] = (int32_t)k;
}

bool build_transposed_records(const PlanHost& p, std::vector<int32_t>& pool, PlanDev& d) {
    if (p.nblk != 1) return false;
    const int32_t n = p.n;
    pool.clear();
    auto put16 = [&](const std::vector<int32_t>& v) -> int32_t {
        while (pool.size() & 3) pool.push_back(0);
        const size_t off = pool.size();
        if (off + v.size() + 4 > (size_t)INT32_MAX) throw std::runtime_error("plan pool too large");
        pool.insert(pool.end(), v.begin(), v.end());
        pool.insert(pool.end(), {0, 0, 0, 0});
        return (int32_t)off;
    };
    std::vector<int32_t> lev, row, col, src;
    int32_t maxent = 0;
    auto close_level = [&](int32_t s0, int32_t k0, int32_t div) {
        if ((int32_t)(row.size() / 4) == k0) return;
        lev.insert(lev.end(), {s0, k0, div, 0});
        maxent = std::max(maxent, (int32_t)col.size() - s0);
    };
    // U^T: y_k = (t_k - sum_j U(j,k) y_j) / U(k,k) on the column levels
    for (int32_t l = 0; l < p.nflev; l++) {
        const int32_t s0 = (int32_t)col.size(), k0 = (int32_t)(row.size() / 4);
        for (int32_t q = p.flev_ptr[l]; q < p.flev_ptr[l + 1]; q++) {
            const int32_t k = p.flev_cols[q], sb = (int32_t)col.size();
            for (int32_t u = p.ucp[k]; u < p.ucp[k + 1] - 1; u++) {
                col.push_back(p.uri[u]);
                src.push_back(p.nl + 1 + u);
            }
            row.insert(row.end(), {k, sb, (int32_t)col.size(), 0});
        }
        close_level(s0, k0, 1);
    }
    // L^T: z_k = y_k - sum_i L(i,k) z_i, column k after every row of L(:,k)
    std::vector<int32_t> llev(n, 0);
    int32_t nl_lev = 0;
    for (int32_t k = n - 1; k >= 0; k--) {
        int32_t lv = 0;
        for (int32_t q = p.lcp[k]; q < p.lcp[k + 1]; q++) lv = std::max(lv, llev[p.lri[q]] + 1);
        llev[k] = lv;
        nl_lev = std::max(nl_lev, lv + 1);
    }
    std::vector<std::vector<int32_t>> by(nl_lev);
    for (int32_t k = 0; k < n; k++)
        if (p.lcp[k + 1] > p.lcp[k]) by[llev[k]].push_back(k);
    for (int32_t l = 0; l < nl_lev; l++) {
        const int32_t s0 = (int32_t)col.size(), k0 = (int32_t)(row.size() / 4);
        for (int32_t k : by[l]) {
            const int32_t sb = (int32_t)col.size();
            for (int32_t q = p.lcp[k]; q < p.lcp[k + 1]; q++) {
                col.push_back(p.lri[q]);
                src.push_back(q);
            }
            row.insert(row.end(), {k, sb, (int32_t)col.size(), 0});
        }
        close_level(s0, k0, 0);
    }
    if (lev.empty()) return false;
    const int32_t nlev = (int32_t)(lev.size() / 4);
    lev.insert(lev.end(), {(int32_t)col.size(), (int32_t)(row.size() / 4), 0, 0});
    d.o_perm_in = put16(p.perm_out);  // T[perm_out[i]] = e[i]
    d.o_perm_out = put16(p.perm_in);  // y[i] = T[perm_in[i]]
    d.o_sq_lev = put16(lev);
    d.o_sq_row = put16(row);
    d.o_sq_col = put16(col);
    d.o_sq_src = put16(src);
    d.sq_nlev = nlev;
    d.sq_maxent = maxent;
    return true;
}

// Pattern-level analysis: symbolic factor with identity pivots so that the
// plan's scatter targets, solve permutation and off-diagonal row keys are the
// pre-pivot positions the kernel maps through its own RPI.
void build_identity_plan(i64 n, const i64* ap, const i64* ai, int kind, PlanHost& ph) {
    HostFactor f;
    f.n = n;
    f.sparselu = kind == 1;
    if (kind == 0 && n >= 4) {
        f.btf = btf_decompose(n, ap, ai);
        f.has_btf = f.btf.num_blocks >= 2 && f.btf.max_block_size < n;
    }
    if (f.has_btf) {
        f.sym.resize(f.btf.num_blocks);
        for (i64 b = 0; b < f.btf.num_blocks; b++) {
            i64 bs = f.btf.block_ptr[b], be = f.btf.block_ptr[b + 1], sz = be - bs;
            std::vector<i64> bap(sz + 1, 0), bai;
            for (i64 lc = 0; lc < sz; lc++) {
                i64 gco = f.btf.col_perm[bs + lc];
                for (i64 q = ap[gco]; q < ap[gco + 1]; q++) {
                    i64 gro = ai[q];
                    if (gro >= n) continue;
                    i64 grn = f.btf.row_perm_inv[gro];
                    if (grn >= bs && grn < be) bai.push_back(grn - bs);
                }
                bap[lc + 1] = (i64)bai.size();
            }
            block_symbolic(sz, bap.data(), bai.data(), false, f.sym[b]);
        }
    } else {
        f.sym.resize(1);
        block_symbolic(n, ap, ai, false, f.sym[0]);
    }
    f.num.resize(f.sym.size());
    for (size_t b = 0; b < f.sym.size(); b++) {
        i64 m = f.sym[b].n;
        f.num[b].row_perm.resize(m);
        f.num[b].row_perm_inv.resize(m);
        for (i64 i = 0; i < m; i++) f.num[b].row_perm[i] = f.num[b].row_perm_inv[i] = i;
    }
    build_plan(f, ap, ai, ph);
}


}  // namespace rsgpu
