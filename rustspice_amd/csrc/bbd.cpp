// bbd.cpp - BBD decomposition (host) and the BbdSolver / DenseSolver back ends
// of rs_solver (see bbd.hpp).
#include "bbd.hpp"

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>

#include "../../include/rsgpu.h"
#include "profile.hpp"
#include "solver_impl.hpp"

namespace rsgpu {

// ---------------------------------------------------------------------------
// Partition and decomposition (bbd.rs). Integer work; results equal the
// reference's: BFS from the first node of least in-block degree, level sets,
// half split, both sides of the cut to the border, largest block (last on
// ties) bisected next, then cross-block nodes promoted column- and row-wise.
// ---------------------------------------------------------------------------
namespace {

struct Adj {
    std::vector<int64_t> ptr, idx;
};

Adj symmetric_adjacency(int64_t n, const int64_t* ap, const int64_t* ai) {  // bbd.rs:297-321
    std::vector<std::vector<int64_t>> l(n);
    for (int64_t c = 0; c < n; c++)
        for (int64_t q = ap[c]; q < ap[c + 1]; q++) {
            const int64_t r = ai[q];
            if (r < n && r != c) {
                l[c].push_back(r);
                l[r].push_back(c);
            }
        }
    Adj a;
    a.ptr.assign(n + 1, 0);
    for (int64_t i = 0; i < n; i++) {
        std::sort(l[i].begin(), l[i].end());
        l[i].erase(std::unique(l[i].begin(), l[i].end()), l[i].end());
        a.ptr[i + 1] = a.ptr[i] + (int64_t)l[i].size();
    }
    a.idx.reserve(a.ptr[n]);
    for (int64_t i = 0; i < n; i++) a.idx.insert(a.idx.end(), l[i].begin(), l[i].end());
    return a;
}

enum : uint8_t { G_A = 0, G_B = 1, G_CUT = 2 };

// bfs_bisect (bbd.rs:185-295) on the node list `nodes` (ascending).
void bfs_bisect(const std::vector<int64_t>& nodes, const Adj& adj, std::vector<uint8_t>& inset,
                std::vector<uint8_t>& grp, std::vector<uint8_t>& seen) {
    const int64_t m = (int64_t)nodes.size();
    for (int64_t v : nodes) inset[v] = 1;
    if (m <= 3) {
        for (int64_t v : nodes) grp[v] = G_A;
    } else {
        int64_t start = nodes[0], best = -1;
        for (int64_t v : nodes) {
            int64_t d = 0;
            for (int64_t q = adj.ptr[v]; q < adj.ptr[v + 1]; q++) d += inset[adj.idx[q]];
            if (best < 0 || d < best) {
                best = d;
                start = v;
            }
        }
        std::vector<int64_t> order;
        std::vector<int64_t> lvl;  // level start offsets into order
        order.reserve(m);
        order.push_back(start);
        seen[start] = 1;
        size_t head = 0;
        while (head < order.size()) {
            const size_t end = order.size();
            lvl.push_back((int64_t)head);
            for (; head < end; head++) {
                const int64_t v = order[head];
                for (int64_t q = adj.ptr[v]; q < adj.ptr[v + 1]; q++) {
                    const int64_t u = adj.idx[q];
                    if (inset[u] && !seen[u]) {
                        seen[u] = 1;
                        order.push_back(u);
                    }
                }
            }
        }
        for (int64_t v : nodes)
            if (!seen[v]) {
                lvl.push_back((int64_t)order.size());
                order.push_back(v);
                seen[v] = 1;
            }
        lvl.push_back((int64_t)order.size());
        const int64_t nl = (int64_t)lvl.size() - 1, half = m / 2;
        int64_t split = nl, count = 0;
        for (int64_t l = 0; l < nl; l++) {
            count += lvl[l + 1] - lvl[l];
            if (count >= half) {
                split = l + 1;
                break;
            }
        }
        for (int64_t l = 0; l < nl; l++)
            for (int64_t q = lvl[l]; q < lvl[l + 1]; q++) grp[order[q]] = l < split ? G_A : G_B;
        std::vector<int64_t> cut;
        for (int64_t v : nodes)
            for (int64_t q = adj.ptr[v]; q < adj.ptr[v + 1]; q++) {
                const int64_t u = adj.idx[q];
                if (inset[u] && grp[u] != grp[v]) {
                    cut.push_back(v);
                    break;
                }
            }
        for (int64_t v : cut) grp[v] = G_CUT;
        for (int64_t v : order) seen[v] = 0;
    }
    for (int64_t v : nodes) inset[v] = 0;
}

}  // namespace

int64_t bbd_partition(int64_t n, const int64_t* ap, const int64_t* ai, int64_t num_blocks,
                      std::vector<int64_t>& assign) {
    assign.assign(n, 0);
    if (n == 0 || num_blocks <= 1) return 1;
    const Adj adj = symmetric_adjacency(n, ap, ai);
    std::vector<uint8_t> inset(n, 0), grp(n, G_A), seen(n, 0);
    int64_t cur = 1;
    std::vector<int64_t> sizes, nodes;
    while (cur < num_blocks) {
        sizes.assign(cur, 0);
        for (int64_t i = 0; i < n; i++)
            if (assign[i] >= 0 && assign[i] < cur) sizes[assign[i]]++;
        int64_t tgt = 0;
        for (int64_t b = 0; b < cur; b++)
            if (sizes[b] >= sizes[tgt]) tgt = b;  // max_by_key: last maximum
        if (sizes[tgt] < 4) break;
        nodes.clear();
        for (int64_t i = 0; i < n; i++)
            if (assign[i] == tgt) nodes.push_back(i);
        bfs_bisect(nodes, adj, inset, grp, seen);
        for (int64_t v : nodes) {
            if (grp[v] == G_B) assign[v] = cur;
            else if (grp[v] == G_CUT) assign[v] = -1;
        }
        cur++;
    }
    return cur;
}

void bbd_decompose(int64_t n, const int64_t* ap, const int64_t* ai, int64_t num_blocks,
                   BbdDecomp& d) {
    d = BbdDecomp();
    d.n = n;
    std::vector<int64_t> a;
    const int64_t nb = bbd_partition(n, ap, ai, num_blocks, a);
    for (int64_t c = 0; c < n; c++) {  // columns that reach another block
        if (a[c] < 0) continue;
        for (int64_t q = ap[c]; q < ap[c + 1]; q++) {
            const int64_t r = ai[q];
            if (r < n && r != c && a[r] >= 0 && a[r] != a[c]) {
                a[c] = -1;
                break;
            }
        }
    }
    for (int64_t c = 0; c < n; c++)  // rows reached from another block
        for (int64_t q = ap[c]; q < ap[c + 1]; q++) {
            const int64_t r = ai[q];
            if (r < n && r != c && a[r] >= 0 && a[c] >= 0 && a[c] != a[r]) a[r] = -1;
        }
    std::vector<int64_t> bsz(nb, 0);
    for (int64_t i = 0; i < n; i++)
        if (a[i] >= 0) bsz[a[i]]++;
    int64_t nonempty = 0;
    for (int64_t b = 0; b < nb; b++) nonempty += bsz[b] > 0;
    d.perm.reserve(n);
    if (nonempty <= 1) {
        if (nonempty == 1) {
            for (int64_t i = 0; i < n; i++)
                if (a[i] >= 0) d.perm.push_back(i);
            for (int64_t i = 0; i < n; i++)
                if (a[i] < 0) d.perm.push_back(i);
        } else {
            for (int64_t i = 0; i < n; i++) d.perm.push_back(i);
        }
        d.num_blocks = 1;
        d.block_ptr = {0, n};
        d.border_start = n;
        d.border_size = 0;
    } else {
        std::vector<std::vector<int64_t>> members(nb);
        for (int64_t i = 0; i < n; i++)
            if (a[i] >= 0) members[a[i]].push_back(i);
        for (int64_t b = 0; b < nb; b++) {
            if (members[b].empty()) continue;
            d.block_ptr.push_back((int64_t)d.perm.size());
            d.perm.insert(d.perm.end(), members[b].begin(), members[b].end());
        }
        d.block_ptr.push_back((int64_t)d.perm.size());
        d.num_blocks = nonempty;
        d.border_start = (int64_t)d.perm.size();
        for (int64_t i = 0; i < n; i++)
            if (a[i] < 0) d.perm.push_back(i);
        d.border_size = n - d.border_start;
    }
    d.inv_perm.assign(n, 0);
    for (int64_t i = 0; i < n; i++) d.inv_perm[d.perm[i]] = i;
}

// ---------------------------------------------------------------------------
// BbdSolver back end
// ---------------------------------------------------------------------------
struct BbdImpl {
    int64_t target = 4;  // SolverType::Bbd hard-wires 4 blocks (solver.rs:169-200)
    int sn_mode = 1;     // supernodal mode of the SparseLU fallback
    bool analyzed = false, factored = false, has_decomp = false, dense = true;
    BbdDecomp d;
    rs_solver* fallback = nullptr;  // SparseLU on the whole matrix
    rs_solver* schur = nullptr;     // SparseLU on a large Schur complement
    int64_t nnz = 0, npairs = 0, ncells = 0, max_blk = 0;
    int64_t n_bmap = 0, n_cidx = 0, n_ridx = 0, n_pidx = 0;  // gather lengths
    std::vector<int64_t> s_ap, s_ai;
    bool lds_ref = false, lds_sol = false;
    // dense Schur product on MFMA tiles: packed K x bsp operands
    int32_t mfma_mode = 0;
    bool mfma_dense = false, mfma_used = false;
    int64_t mfma_k = 0, mfma_kp = 0, mfma_kc = 0, mfma_permille = 0;
    int32_t mfma_chunks = 0;
    DevBuf<int32_t> d_mcptr, d_mcidx, d_mymap;
    DevBuf<double> d_mcp, d_myp, d_mpart;
    DevBuf<PlanDev> d_fac, d_blk, d_pair;
    DevBuf<int32_t> d_ip, d_bmap, d_perm, d_iperm, d_blkstart;
    DevBuf<int32_t> d_crow, d_cblk, d_ccol, d_cidx, d_rrow, d_rcol, d_ridx;
    DevBuf<int32_t> d_poff, d_pptr, d_prow, d_pidx, d_pof, d_cellr, d_cellc, d_btp, d_bti;
    DevBuf<int32_t> d_piv, d_status;
    DevBuf<double> d_ax, d_bax, d_vals, d_wpool, d_cval, d_rval, d_pval, d_y, d_y2, d_base,
        d_s, d_rhs, d_work, d_z, d_w, d_wz, d_bb, d_xt, d_tmp;
};

BbdImpl* bbd_create() { return new BbdImpl(); }
void bbd_destroy(BbdImpl* b) {
    if (!b) return;
    if (b->fallback) rs_solver_destroy(b->fallback);
    if (b->schur) rs_solver_destroy(b->schur);
    delete b;
}

namespace {

// column pointers monotone within ai, row indices non-negative
bool csc_ok(int64_t n, const int64_t* ap, const int64_t* ai, int64_t ail) {
    if (ap[0] != 0) return false;
    for (int64_t c = 0; c < n; c++)
        if (ap[c + 1] < ap[c] || ap[c + 1] > ail) return false;
    for (int64_t q = 0; q < ap[n]; q++)
        if (ai[q] < 0) return false;
    return true;
}

int32_t bfail(rs_solver* h, int32_t code, const std::string& m) {
    h->err = m;
    return code;
}
int32_t bhip(rs_solver* h, hipError_t e, const char* what) {
    h->err = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
    return RS_ERR_DEVICE;
}

template <class T>
hipError_t put(DevBuf<T>& b, const std::vector<T>& v) {
    hipError_t e = b.ensure(std::max<size_t>(v.size(), 1));
    if (e || v.empty()) return e;
    return hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

int32_t c32(int64_t v) {
    if (v < 0 || v > INT32_MAX) throw std::runtime_error("index exceeds int32");
    return (int32_t)v;
}

// entries of permuted column nc: (new row, CSC index), CSC order
template <class F>
void for_col(const BbdDecomp& d, const int64_t* ap, const int64_t* ai, int64_t nc, F f) {
    const int64_t oc = d.perm[nc];
    for (int64_t q = ap[oc]; q < ap[oc + 1]; q++) {
        if (ai[q] >= d.n) continue;
        f(d.inv_perm[ai[q]], q);
    }
}

void clear_impl(BbdImpl* b) {
    if (b->fallback) rs_solver_destroy(b->fallback);
    if (b->schur) rs_solver_destroy(b->schur);
    b->fallback = b->schur = nullptr;
    b->has_decomp = b->factored = false;
}

int32_t setup_fallback(rs_solver* h, const int64_t* ap, int64_t apl, const int64_t* ai, int64_t ail) {
    BbdImpl* b = h->bbd;
    clear_impl(b);
    b->fallback = rs_solver_create(RS_SOLVER_SPARSELU, h->n);
    rs_set_supernodal(b->fallback, b->sn_mode);
    rs_prepare(b->fallback, h->n);
    int32_t rc = rs_analyze(b->fallback, ap, apl, ai, ail);
    if (rc) {
        char buf[512];
        rs_last_error(b->fallback, buf, sizeof buf);
        return bfail(h, rc, buf);
    }
    return RS_OK;
}

int32_t build_bbd(rs_solver* h, const int64_t* ap, int64_t apl, const int64_t* ai, int64_t ail) {
    BbdImpl* b = h->bbd;
    const int64_t n = h->n;
    BbdDecomp d;
    bbd_decompose(n, ap, ai, b->target, d);
    if (d.num_blocks <= 1 || d.border_size == 0 || (double)d.border_size / (double)n > 0.5)
        return setup_fallback(h, ap, apl, ai, ail);  // bbd_solver.rs:420-432
    clear_impl(b);
    b->d = d;
    const int64_t nb = d.num_blocks, bs = d.border_size, b0 = d.border_start;
    b->nnz = ap[n];
    // ---- diagonal blocks: identity-pivot SparseLU plans in one pool -------
    std::vector<int32_t> pool, bmap;
    std::vector<PlanDev> fac(nb), blk(nb);
    std::vector<PlanHost> ph(nb);
    int64_t vals = 0, bax = 0, wp = 0;
    b->max_blk = 0;
    for (int64_t k = 0; k < nb; k++) {
        const int64_t st = d.block_ptr[k], en = d.block_ptr[k + 1], sz = en - st;
        b->max_blk = std::max(b->max_blk, sz);
        std::vector<int64_t> bap(sz + 1, 0), bai;
        const int64_t m0 = (int64_t)bmap.size();
        for (int64_t lc = 0; lc < sz; lc++) {  // extract_block_patterns (:501-542)
            for_col(d, ap, ai, st + lc, [&](int64_t nr, int64_t q) {
                if (nr >= st && nr < en) {
                    bai.push_back(nr - st);
                    bmap.push_back(c32(q));
                }
            });
            bap[lc + 1] = (int64_t)bai.size();
        }
        build_identity_plan(sz, bap.data(), bai.data(), RS_SOLVER_SPARSELU, ph[k]);
        PlanDev p{};
        append_plan(ph[k], pool, p);
        p.o_lx = vals;
        p.o_ux = p.o_lx + ph[k].nl + 1;
        p.o_ud = p.o_ux + ph[k].nu + 1;
        vals = p.o_ud + sz + 1;
        p.o_ax = bax;
        bax += (int64_t)bmap.size() - m0;
        p.o_work = wp;
        wp += (int64_t)(kThreads / 64) * sz;
        p.o_tvec = wp;
        wp += sz;
        p.o_vec = st;  // block rows of the permuted vectors
        fac[k] = p;
        blk[k] = p;
    }
    // ---- C_k by border row, R_k by block row ------------------------------
    std::vector<int32_t> crow(bs + 1, 0), cblk, ccol, cidx;
    {
        std::vector<std::vector<std::array<int64_t, 3>>> rows(bs);  // (k, col, q)
        for (int64_t k = 0; k < nb; k++) {
            const int64_t st = d.block_ptr[k], sz = d.block_ptr[k + 1] - st;
            for (int64_t lc = 0; lc < sz; lc++)
                for_col(d, ap, ai, st + lc, [&](int64_t nr, int64_t q) {
                    if (nr >= b0) rows[nr - b0].push_back({k, lc, q});
                });
        }
        for (int64_t r = 0; r < bs; r++) {  // already in (k, col, entry) order
            for (auto& e : rows[r]) {
                cblk.push_back(c32(e[0]));
                ccol.push_back(c32(e[1]));
                cidx.push_back(c32(e[2]));
            }
            crow[r + 1] = c32((int64_t)cblk.size());
        }
    }
    std::vector<int32_t> rrow(b0 + 1, 0), rcol, ridx;
    std::vector<int32_t> poff, pptr{0}, prow, pidx, pof(nb * bs, -1);
    std::vector<PlanDev> pairs;
    int64_t ytot = 0;
    {
        std::vector<std::vector<std::array<int64_t, 2>>> rows(b0);  // (border col, q)
        for (int64_t k = 0; k < nb; k++) {
            const int64_t st = d.block_ptr[k], en = d.block_ptr[k + 1], sz = en - st;
            for (int64_t j = 0; j < bs; j++) {
                const size_t e0 = prow.size();
                for_col(d, ap, ai, b0 + j, [&](int64_t nr, int64_t q) {
                    if (nr >= st && nr < en) {
                        rows[nr].push_back({j, q});
                        prow.push_back(c32(nr - st));
                        pidx.push_back(c32(q));
                    }
                });
                if (prow.size() == e0) continue;
                pof[k * bs + j] = c32((int64_t)poff.size());
                poff.push_back(c32(ytot));
                pptr.push_back(c32((int64_t)prow.size()));
                PlanDev p = blk[k];
                p.o_vec = ytot;
                p.o_tvec = wp;
                wp += sz;
                pairs.push_back(p);
                ytot += sz;
            }
        }
        for (int64_t i = 0; i < b0; i++) {
            for (auto& e : rows[i]) {
                rcol.push_back(c32(e[0]));
                ridx.push_back(c32(e[1]));
            }
            rrow[i + 1] = c32((int64_t)rcol.size());
        }
    }
    b->npairs = (int64_t)pairs.size();
    // ---- Schur cells: dense bs x bs, or the sparse pattern (:330-408) -----
    std::vector<int32_t> cellr, cellc, btp{0}, bti;
    b->dense = bs <= kDenseSchurMax;
    {
        std::vector<std::vector<int64_t>> bt(bs * (b->dense ? bs : 0));
        std::vector<std::vector<std::pair<int64_t, int64_t>>> btc(bs);  // per column: (row, q)
        for (int64_t j = 0; j < bs; j++)
            for_col(d, ap, ai, b0 + j, [&](int64_t nr, int64_t q) {
                if (nr >= b0) btc[j].push_back({nr - b0, q});
            });
        if (b->dense) {
            for (int64_t r = 0; r < bs; r++)
                for (int64_t j = 0; j < bs; j++) {
                    cellr.push_back(c32(r));
                    cellc.push_back(c32(j));
                }
            for (int64_t j = 0; j < bs; j++)
                for (auto& e : btc[j]) bt[e.first * bs + j].push_back(e.second);
            for (auto& v : bt) {
                for (int64_t q : v) bti.push_back(c32(q));
                btp.push_back(c32((int64_t)bti.size()));
            }
        } else {
            std::vector<uint8_t> crows(nb * bs, 0), rcols(nb * bs, 0), mark(bs);
            for (int64_t r = 0; r < bs; r++)
                for (int32_t e = crow[r]; e < crow[r + 1]; e++) crows[cblk[e] * bs + r] = 1;
            for (int64_t k = 0; k < nb; k++)
                for (int64_t j = 0; j < bs; j++) rcols[k * bs + j] = pof[k * bs + j] >= 0;
            b->s_ap.assign(bs + 1, 0);
            b->s_ai.clear();
            for (int64_t j = 0; j < bs; j++) {
                std::fill(mark.begin(), mark.end(), 0);
                for (auto& e : btc[j]) mark[e.first] = 1;
                for (int64_t k = 0; k < nb; k++)
                    if (rcols[k * bs + j])
                        for (int64_t i = 0; i < bs; i++)
                            if (crows[k * bs + i]) mark[i] = 1;
                const size_t c0 = b->s_ai.size();
                for (int64_t i = 0; i < bs; i++)
                    if (mark[i]) b->s_ai.push_back(i);
                b->s_ap[j + 1] = (int64_t)b->s_ai.size();
                for (size_t p = c0; p < b->s_ai.size(); p++) {
                    const int64_t r = b->s_ai[p];
                    cellr.push_back(c32(r));
                    cellc.push_back(c32(j));
                    for (auto& e : btc[j])  // B_T entries of (r, j) in CSC order
                        if (e.first == r) bti.push_back(c32(e.second));
                    btp.push_back(c32((int64_t)bti.size()));
                }
            }
            b->schur = rs_solver_create(RS_SOLVER_SPARSELU, bs);
            rs_prepare(b->schur, bs);
            int32_t rc = rs_analyze(b->schur, b->s_ap.data(), bs + 1, b->s_ai.data(),
                                    (int64_t)b->s_ai.size());
            if (rc) {
                char buf[512];
                rs_last_error(b->schur, buf, sizeof buf);
                return bfail(h, rc, buf);
            }
        }
    }
    // ---- packed operands of the MFMA Schur product (dense Schur only) ------
    std::vector<int32_t> mcptr, mcidx, mymap;
    b->mfma_k = b->mfma_kp = b->mfma_kc = b->mfma_permille = 0;
    b->mfma_chunks = 0;
    b->mfma_dense = false;
    if (b->dense && bs > 0) {
        std::vector<std::vector<int32_t>> cid(nb);  // coupling column id of (k, col)
        for (int64_t k = 0; k < nb; k++) cid[k].assign(d.block_ptr[k + 1] - d.block_ptr[k], -1);
        for (size_t e = 0; e < cblk.size(); e++) cid[cblk[e]][ccol[e]] = 0;
        std::vector<std::array<int32_t, 2>> ccols;  // (k, col) in (k, col) order
        for (int64_t k = 0; k < nb; k++)
            for (size_t c = 0; c < cid[k].size(); c++)
                if (cid[k][c] == 0) {
                    cid[k][c] = c32((int64_t)ccols.size());
                    ccols.push_back({c32(k), c32((int64_t)c)});
                }
        const int64_t K = (int64_t)ccols.size(), bsp = (bs + 15) / 16 * 16;
        if (K > 0) {
            const int64_t kp = (K + 15) / 16 * 16, m = kp * bsp;
            std::vector<std::vector<int32_t>> cell(m);  // C entries' ax indices per packed cell
            for (int64_t r = 0; r < bs; r++)
                for (int32_t e = crow[r]; e < crow[r + 1]; e++)
                    cell[(int64_t)cid[cblk[e]][ccol[e]] * bsp + r].push_back(cidx[e]);
            mcptr.assign(m + 1, 0);
            int64_t cnz = 0, ynz = 0;
            for (int64_t t = 0; t < m; t++) {
                cnz += !cell[t].empty();
                mcidx.insert(mcidx.end(), cell[t].begin(), cell[t].end());
                mcptr[t + 1] = c32((int64_t)mcidx.size());
            }
            mymap.assign(m, -1);
            for (int64_t c = 0; c < K; c++)
                for (int64_t j = 0; j < bs; j++) {
                    const int32_t p = pof[ccols[c][0] * bs + j];
                    if (p < 0) continue;
                    mymap[c * bsp + j] = c32((int64_t)poff[p] + ccols[c][1]);
                    ynz++;
                }
            const double dc = (double)cnz / (double)(K * bs), dy = (double)ynz / (double)(K * bs);
            b->mfma_k = K;
            b->mfma_kp = kp;
            b->mfma_permille = (int64_t)(1000.0 * dc);
            b->mfma_dense = K >= 64 && dc >= 0.25 && dy >= 0.25;
            // split K so the tiles fill the chip: about 256 workgroups in all
            const int64_t nt2 = (bsp / 16) * (bsp / 16);
            int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(kp / 64, (256 + nt2 - 1) / nt2));
            b->mfma_kc = (kp / 16 + chunks - 1) / chunks * 16;
            b->mfma_chunks = (int32_t)((kp + b->mfma_kc - 1) / b->mfma_kc);
        }
    }
    b->ncells = (int64_t)cellr.size();
    b->n_bmap = (int64_t)bmap.size();
    b->n_cidx = (int64_t)cidx.size();
    b->n_ridx = (int64_t)ridx.size();
    b->n_pidx = (int64_t)pidx.size();
    std::vector<int32_t> perm(n), iperm(n), bstart(nb);
    for (int64_t i = 0; i < n; i++) {
        perm[i] = c32(d.perm[i]);
        iperm[i] = c32(d.inv_perm[i]);
    }
    for (int64_t k = 0; k < nb; k++) bstart[k] = c32(d.block_ptr[k]);
    const size_t maxlds = device_max_lds();
    b->lds_ref = (size_t)(kThreads / 64) * b->max_blk * sizeof(double) <= maxlds;
    b->lds_sol = ((size_t)b->max_blk + kThreads) * sizeof(double) <= maxlds;  // + ordered-sum staging
    hipError_t e;
    if ((e = put(b->d_fac, fac)) || (e = put(b->d_blk, blk)) || (e = put(b->d_pair, pairs)) ||
        (e = put(b->d_ip, pool)) || (e = put(b->d_bmap, bmap)) || (e = put(b->d_perm, perm)) ||
        (e = put(b->d_iperm, iperm)) || (e = put(b->d_blkstart, bstart)) ||
        (e = put(b->d_crow, crow)) || (e = put(b->d_cblk, cblk)) || (e = put(b->d_ccol, ccol)) ||
        (e = put(b->d_cidx, cidx)) || (e = put(b->d_rrow, rrow)) || (e = put(b->d_rcol, rcol)) ||
        (e = put(b->d_ridx, ridx)) || (e = put(b->d_poff, poff)) || (e = put(b->d_pptr, pptr)) ||
        (e = put(b->d_prow, prow)) || (e = put(b->d_pidx, pidx)) || (e = put(b->d_pof, pof)) ||
        (e = put(b->d_cellr, cellr)) || (e = put(b->d_cellc, cellc)) || (e = put(b->d_btp, btp)) ||
        (e = put(b->d_bti, bti)) || (e = put(b->d_mcptr, mcptr)) ||
        (e = put(b->d_mcidx, mcidx)) || (e = put(b->d_mymap, mymap)))
        return bhip(h, e, "BBD plan upload");
    if (b->mfma_kp > 0) {
        const int64_t bsp = (bs + 15) / 16 * 16, nt = bsp / 16;
        if ((e = b->d_mcp.ensure(b->mfma_kp * bsp)) || (e = b->d_myp.ensure(b->mfma_kp * bsp)) ||
            (e = b->d_mpart.ensure((size_t)b->mfma_chunks * nt * nt * 256)))
            return bhip(h, e, "BBD MFMA operands");
    }
    if ((e = b->d_ax.ensure(b->nnz + 1)) || (e = b->d_bax.ensure(bax + 1)) ||
        (e = b->d_vals.ensure(vals + 1)) || (e = b->d_wpool.ensure(wp + 1)) ||
        (e = b->d_cval.ensure(cidx.size() + 1)) || (e = b->d_rval.ensure(ridx.size() + 1)) ||
        (e = b->d_pval.ensure(pidx.size() + 1)) || (e = b->d_y.ensure(ytot + 1)) ||
        (e = b->d_y2.ensure(ytot + 1)) || (e = b->d_base.ensure(b->ncells + 1)) ||
        (e = b->d_s.ensure(b->ncells + 1)) || (e = b->d_rhs.ensure(n + 1)) ||
        (e = b->d_work.ensure(n + 1)) || (e = b->d_z.ensure(n + 1)) || (e = b->d_w.ensure(n + 1)) ||
        (e = b->d_wz.ensure(n + 1)) || (e = b->d_bb.ensure(bs + 1)) || (e = b->d_xt.ensure(bs + 1)) ||
        (e = b->d_tmp.ensure(bs + 1)) || (e = b->d_piv.ensure(bs + 1)) ||
        (e = b->d_status.ensure(2)))
        return bhip(h, e, "BBD buffers");
    if ((e = hipMemset(b->d_wpool.p, 0, (wp + 1) * sizeof(double))))
        return bhip(h, e, "BBD work");
    b->has_decomp = true;
    return RS_OK;
}

}  // namespace

int32_t bbd_analyze(rs_solver* h, const int64_t* ap, int64_t apl, const int64_t* ai, int64_t ail) {
    BbdImpl* b = h->bbd;
    if (b->analyzed && h->last_ap.size() == (size_t)apl && h->last_ai.size() == (size_t)ail &&
        std::equal(ap, ap + apl, h->last_ap.begin()) && std::equal(ai, ai + ail, h->last_ai.begin()))
        return RS_OK;
    if (apl != h->n + 1) {
        char m[160];
        std::snprintf(m, sizeof m, "Invalid matrix: Column pointer length %lld != expected %lld",
                      (long long)apl, (long long)(h->n + 1));
        return bfail(h, RS_ERR_INVALID_MATRIX, m);
    }
    if (!csc_ok(h->n, ap, ai, ail))
        return bfail(h, RS_ERR_INVALID_MATRIX, "Invalid matrix: column pointers are not monotone");
    b->analyzed = b->factored = false;
    int32_t rc;
    try {
        rc = build_bbd(h, ap, apl, ai, ail);
    } catch (const std::exception& ex) {
        return bfail(h, RS_ERR_INVALID_MATRIX, std::string("Invalid matrix: ") + ex.what());
    }
    if (rc) return rc;
    h->last_ap.assign(ap, ap + apl);
    h->last_ai.assign(ai, ai + ail);
    b->analyzed = true;
    return RS_OK;
}

int32_t bbd_factor(rs_solver* h, const int64_t* ap, int64_t apl, const int64_t* ai, int64_t ail,
                   const double* ax) {
    BbdImpl* b = h->bbd;
    int32_t rc = bbd_analyze(h, ap, apl, ai, ail);
    if (rc) return rc;
    if (!b->has_decomp) {
        rc = rs_factor(b->fallback, ap, apl, ai, ail, ax, ap[h->n]);
        if (rc) {
            char buf[512];
            rs_last_error(b->fallback, buf, sizeof buf);
            return bfail(h, rc, buf);
        }
    } else {
        const int64_t nb = b->d.num_blocks, bs = b->d.border_size;
        hipError_t e;
        if ((e = hipMemcpy(b->d_ax.p, ax, b->nnz * sizeof(double), hipMemcpyHostToDevice)))
            return bhip(h, e, "hipMemcpy(ax)");
        hipStream_t s = nullptr;
        // block values, all diagonal blocks factored in one launch
        if ((e = launch_gather(b->d_bax.p, b->d_ax.p, b->d_bmap.p, b->n_bmap, s)) ||
            (e = launch_refactor(b->d_fac.p, (int)nb, b->d_ip.p, b->d_bax.p, b->d_vals.p,
                                 b->d_wpool.p, (int)b->max_blk, b->lds_ref, kThreads, s)))
            return bhip(h, e, "BBD block factor");
        // C, R and pair values; Y = B_kk^-1 R_k(:, j) for every pair in one launch
        if ((e = launch_gather(b->d_cval.p, b->d_ax.p, b->d_cidx.p, b->n_cidx, s)) ||
            (e = launch_gather(b->d_rval.p, b->d_ax.p, b->d_ridx.p, b->n_ridx, s)) ||
            (e = launch_gather(b->d_pval.p, b->d_ax.p, b->d_pidx.p, b->n_pidx, s)))
            return bhip(h, e, "BBD C and R values");
        if (b->npairs) {
            if ((e = hipMemsetAsync(b->d_y.p, 0, b->d_y.n * sizeof(double), s)) ||
                (e = launch_bbd_pair_rhs(b->d_y.p, b->d_poff.p, b->d_pptr.p, b->d_prow.p, b->d_pval.p,
                                         b->npairs, s)) ||
                (e = launch_solve(b->d_pair.p, (int)b->npairs, b->d_ip.p, b->d_bax.p, b->d_vals.p,
                                  b->d_y.p, b->d_y2.p, b->d_wpool.p, (int)b->max_blk, b->lds_sol,
                                  kThreads, s)))
                return bhip(h, e, "BBD Schur solves");
        }
        b->mfma_used = b->dense && b->mfma_kp > 0 &&
                       (b->mfma_mode == 2 || (b->mfma_mode == 1 && b->mfma_dense));
        if ((e = launch_gather_sum(b->d_base.p, b->d_ax.p, b->d_btp.p, b->d_bti.p, b->ncells, s)))
            return bhip(h, e, "BBD Schur assembly");
        const int ps = prof_begin(PK_BBD_SCHUR, s);
        if (b->mfma_used)
            e = launch_bbd_schur_mfma(b->d_s.p, b->d_base.p, b->d_mcp.p, b->d_myp.p, b->d_mpart.p,
                                      b->d_ax.p, b->d_mcptr.p, b->d_mcidx.p, b->d_mymap.p,
                                      b->d_y2.p, (int32_t)bs, b->mfma_kp, b->mfma_chunks,
                                      b->mfma_kc, s);
        else
            e = launch_bbd_schur(b->d_s.p, b->d_cellr.p, b->d_cellc.p, b->d_base.p, b->d_crow.p,
                                 b->d_cblk.p, b->d_ccol.p, b->d_cval.p, b->d_pof.p, b->d_poff.p,
                                 b->d_y2.p, bs, b->ncells, s);
        prof_end(ps, s);
        if (e) return bhip(h, e, "BBD Schur assembly");
        if (b->dense) {
            int32_t st[2] = {0, -1};
            if ((e = launch_dense_lu(b->d_s.p, b->d_piv.p, b->d_status.p, bs, 1, s)) ||
                (e = hipMemcpy(st, b->d_status.p, sizeof st, hipMemcpyDeviceToHost)))
                return bhip(h, e, "BBD dense Schur LU");
            if (st[0]) {  // DenseSchurSolver::factor (bbd_solver.rs:138-140)
                b->factored = false;
                h->err_pivot = st[1];
                return bfail(h, RS_ERR_SINGULAR,
                             "Singular matrix: zero pivot at row/column " + std::to_string(st[1]));
            }
        } else {
            rc = solver_factor_dev(b->schur, b->s_ap.data(), bs + 1, b->s_ai.data(),
                                   (int64_t)b->s_ai.size(), b->d_s.p, s);
            if (rc) return bfail(h, rc, b->schur->err);
        }
    }
    b->factored = true;
    h->factor_count++;
    h->state = ST_FACTORED;
    return RS_OK;
}

int32_t bbd_solve(rs_solver* h, double* rhs) {
    BbdImpl* b = h->bbd;
    if (!b->factored) return bfail(h, RS_ERR_SOLVE_FAILED, "Solve step failed");
    if (!b->has_decomp) {
        int32_t rc = rs_solve(b->fallback, rhs, h->n);
        if (rc) {
            char buf[512];
            rs_last_error(b->fallback, buf, sizeof buf);
            return bfail(h, rc, buf);
        }
        return RS_OK;
    }
    const int64_t n = h->n, nb = b->d.num_blocks, bs = b->d.border_size, b0 = b->d.border_start;
    hipStream_t s = nullptr;
    hipError_t e;
    if ((e = hipMemcpy(b->d_rhs.p, rhs, n * sizeof(double), hipMemcpyHostToDevice)) ||
        (e = launch_permute(b->d_work.p, b->d_rhs.p, b->d_perm.p, n, s)) ||
        (e = launch_solve(b->d_blk.p, (int)nb, b->d_ip.p, b->d_bax.p, b->d_vals.p, b->d_work.p,
                          b->d_z.p, b->d_wpool.p, (int)b->max_blk, b->lds_sol, kThreads, s)) ||
        (e = launch_bbd_border_rhs(b->d_bb.p, b->d_work.p, b0, b->d_crow.p, b->d_cblk.p,
                                   b->d_ccol.p, b->d_cval.p, b->d_blkstart.p, b->d_z.p, bs, s)))
        return bhip(h, e, "BBD solve (blocks)");
    if (b->dense) {
        int32_t st = 0;
        if ((e = launch_dense_solve(b->d_s.p, b->d_piv.p, b->d_bb.p, b->d_tmp.p, b->d_status.p, bs, 1, s)) ||
            (e = hipMemcpyAsync(b->d_xt.p, b->d_bb.p, bs * sizeof(double), hipMemcpyDeviceToDevice, s)) ||
            (e = hipMemcpy(&st, b->d_status.p, sizeof st, hipMemcpyDeviceToHost)))
            return bhip(h, e, "BBD dense Schur solve");
        if (st) return bfail(h, RS_ERR_SOLVE_FAILED, "Solve step failed");
    } else {
        int32_t rc = solver_solve_dev(b->schur, b->d_bb.p, b->d_xt.p, s);
        if (rc) return bfail(h, rc, b->schur->err);
    }
    if ((e = launch_bbd_rprod(b->d_w.p, b->d_rrow.p, b->d_rcol.p, b->d_rval.p, b->d_xt.p, b0, s)) ||
        (e = launch_solve(b->d_blk.p, (int)nb, b->d_ip.p, b->d_bax.p, b->d_vals.p, b->d_w.p,
                          b->d_wz.p, b->d_wpool.p, (int)b->max_blk, b->lds_sol, kThreads, s)) ||
        (e = launch_bbd_combine(b->d_work.p, b->d_z.p, b->d_wz.p, b->d_xt.p, b0, n, s)) ||
        (e = launch_permute(b->d_rhs.p, b->d_work.p, b->d_iperm.p, n, s)) ||
        (e = hipMemcpy(rhs, b->d_rhs.p, n * sizeof(double), hipMemcpyDeviceToHost)))
        return bhip(h, e, "BBD solve (back-substitution)");
    return RS_OK;
}

int32_t bbd_factor_dev(rs_solver* h, const int64_t* ap, int64_t apl, const int64_t* ai, int64_t ail,
                       const double* ax_dev, hipStream_t s) {
    BbdImpl* b = h->bbd;
    int32_t rc = bbd_analyze(h, ap, apl, ai, ail);
    if (rc) return rc;
    if (b->has_decomp) {  // the decomposition's launches read host-handed values
        std::vector<double> ax((size_t)ap[h->n]);
        hipError_t e = ax.empty() ? hipSuccess
                                  : hipMemcpyAsync(ax.data(), ax_dev, ax.size() * sizeof(double),
                                                   hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return bhip(h, e, "hipMemcpy(ax)");
        return bbd_factor(h, ap, apl, ai, ail, ax.data());
    }
    rc = solver_factor_dev(b->fallback, ap, apl, ai, ail, ax_dev, s);
    if (rc) return bfail(h, rc, b->fallback->err);
    b->factored = true;
    h->factor_count++;
    h->state = ST_FACTORED;
    return RS_OK;
}

int32_t bbd_solve_dev(rs_solver* h, const double* rhs_dev, double* x_dev, hipStream_t s) {
    BbdImpl* b = h->bbd;
    if (!b->factored) return bfail(h, RS_ERR_SOLVE_FAILED, "Solve step failed");
    if (!b->has_decomp) {
        int32_t rc = solver_solve_dev(b->fallback, rhs_dev, x_dev, s);
        return rc ? bfail(h, rc, b->fallback->err) : RS_OK;
    }
    std::vector<double> r(h->n);
    hipError_t e = hipMemcpyAsync(r.data(), rhs_dev, h->n * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return bhip(h, e, "hipMemcpy(rhs)");
    int32_t rc = bbd_solve(h, r.data());
    if (rc) return rc;
    e = hipMemcpyAsync(x_dev, r.data(), h->n * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? RS_OK : bhip(h, e, "hipMemcpy(x)");
}

void bbd_set_supernodal(rs_solver* h, int mode) {
    h->bbd->sn_mode = mode;
    if (h->bbd->fallback) rs_set_supernodal(h->bbd->fallback, mode);
}

const rs_solver* bbd_fallback_solver(const rs_solver* h) {
    return h->bbd && !h->bbd->has_decomp ? h->bbd->fallback : nullptr;
}

void bbd_reset(rs_solver* h) {
    BbdImpl* b = h->bbd;
    clear_impl(b);
    b->analyzed = false;
    h->last_ap.clear();
    h->last_ai.clear();
}

void bbd_set_target(rs_solver* h, int64_t num_blocks) {
    h->bbd->target = num_blocks > 2 ? num_blocks : 2;  // BbdSolver::new (:281)
    bbd_reset(h);
}

int32_t bbd_info(const rs_solver* h, int64_t* out) {
    const BbdImpl* b = h->bbd;
    out[0] = b->has_decomp ? b->d.num_blocks : 1;
    out[1] = b->has_decomp ? b->d.border_size : 0;
    out[2] = b->has_decomp ? (b->dense ? 1 : 0) : -1;
    out[3] = h->factor_count;
    return RS_OK;
}

void bbd_set_mfma(rs_solver* h, int32_t mode) { h->bbd->mfma_mode = mode < 0 ? 0 : mode > 2 ? 2 : mode; }

int32_t bbd_mfma_info(const rs_solver* h, int64_t* out) {
    const BbdImpl* b = h->bbd;
    out[0] = b->mfma_mode;
    out[1] = b->has_decomp && b->mfma_used ? 1 : 0;
    out[2] = b->has_decomp ? b->mfma_k : 0;
    out[3] = b->has_decomp ? b->mfma_permille : 0;
    return RS_OK;
}

// ---------------------------------------------------------------------------
// DenseSolver back end (solver.rs:595-724)
// ---------------------------------------------------------------------------
struct DenseImpl {
    bool factored = false;
    DevBuf<int32_t> d_ap, d_ai, d_piv, d_status;
    DevBuf<double> d_ax, d_lu, d_rhs, d_b;
};

DenseImpl* dense_create() { return new DenseImpl(); }
void dense_destroy(DenseImpl* d) { delete d; }

int32_t dense_factor(rs_solver* h, const int64_t* ap, int64_t apl, const int64_t* ai, int64_t ail,
                     const double* ax) {
    DenseImpl* d = h->dense;
    const int64_t n = h->n;
    d->factored = false;
    if (apl != n + 1) return bfail(h, RS_ERR_ANALYZE_FAILED, "Analysis failed");
    if (!csc_ok(n, ap, ai, ail))
        return bfail(h, RS_ERR_INVALID_MATRIX, "Invalid matrix: column pointers are not monotone");
    const int64_t nnz = ap[n];
    std::vector<int32_t> ap32(n + 1), ai32(std::max<int64_t>(nnz, 1));
    try {
        for (int64_t i = 0; i <= n; i++) ap32[i] = c32(ap[i]);
        for (int64_t q = 0; q < nnz; q++) ai32[q] = c32(std::min<int64_t>(ai[q], n));
    } catch (const std::exception& ex) {
        return bfail(h, RS_ERR_INVALID_MATRIX, std::string("Invalid matrix: ") + ex.what());
    }
    hipError_t e;
    int32_t st = 0;
    if ((e = put(d->d_ap, ap32)) || (e = put(d->d_ai, ai32)) ||
        (e = d->d_ax.ensure(nnz + 1)) || (e = d->d_lu.ensure(n * n + 1)) ||
        (e = d->d_piv.ensure(n + 1)) || (e = d->d_status.ensure(2)) ||
        (e = d->d_rhs.ensure(n + 1)) || (e = d->d_b.ensure(n + 1)) ||
        (e = hipMemcpy(d->d_ax.p, ax, nnz * sizeof(double), hipMemcpyHostToDevice)) ||
        (e = hipMemset(d->d_lu.p, 0, (n * n + 1) * sizeof(double))) ||
        (e = launch_dense_build(d->d_lu.p, d->d_ap.p, d->d_ai.p, d->d_ax.p, n, nullptr)) ||
        (e = launch_dense_lu(d->d_lu.p, d->d_piv.p, d->d_status.p, n, 0, nullptr)) ||
        (e = hipMemcpy(&st, d->d_status.p, sizeof st, hipMemcpyDeviceToHost)))
        return bhip(h, e, "dense factor");
    if (st) return bfail(h, RS_ERR_FACTOR_FAILED, "Factorization failed");
    d->factored = true;
    h->factor_count++;
    h->state = ST_FACTORED;
    return RS_OK;
}

int32_t dense_solve(rs_solver* h, double* rhs) {
    DenseImpl* d = h->dense;
    const int64_t n = h->n;
    if (!d->factored) return bfail(h, RS_ERR_SOLVE_FAILED, "Solve step failed");
    hipError_t e;
    int32_t st = 0;
    if ((e = hipMemcpy(d->d_rhs.p, rhs, n * sizeof(double), hipMemcpyHostToDevice)) ||
        (e = launch_dense_solve(d->d_lu.p, d->d_piv.p, d->d_rhs.p, d->d_b.p, d->d_status.p, n, 0, nullptr)) ||
        (e = hipMemcpy(&st, d->d_status.p, sizeof st, hipMemcpyDeviceToHost)) ||
        (e = hipMemcpy(rhs, d->d_rhs.p, n * sizeof(double), hipMemcpyDeviceToHost)))
        return bhip(h, e, "dense solve");
    if (st) return bfail(h, RS_ERR_SOLVE_FAILED, "Solve step failed");
    return RS_OK;
}

}  // namespace rsgpu
