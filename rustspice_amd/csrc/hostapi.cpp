// hostapi.cpp - C entry points for the host-side integer analysis (AMD, BTF,
// L/U patterns) and the first pivoting factor, so they can be checked on a
// machine without a GPU. They touch no device memory.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rsgpu.h"
#include "../../include/rsgpu_host.h"
#include "bbd.hpp"
#include "ddpow.hpp"
#include "plan.hpp"
#include "sn.hpp"
#include "sp.hpp"
#include "stamp.hpp"

using namespace rsgpu;


extern "C" {

void rs_host_amd_order(int64_t n, const int64_t* ap, const int64_t* ai, int64_t* perm,
                       int64_t* inv_perm) {
    AmdResult r = amd_order(n, ap, ai);
    if (n) {
        std::memcpy(perm, r.perm.data(), n * sizeof(int64_t));
        std::memcpy(inv_perm, r.inv_perm.data(), n * sizeof(int64_t));
    }
}

int64_t rs_host_btf(int64_t n, const int64_t* ap, const int64_t* ai, int64_t* row_perm,
                    int64_t* col_perm, int64_t* block_ptr, int64_t* structural_rank) {
    Btf b = btf_decompose(n, ap, ai);
    if (n) {
        std::memcpy(row_perm, b.row_perm.data(), n * sizeof(int64_t));
        std::memcpy(col_perm, b.col_perm.data(), n * sizeof(int64_t));
    }
    std::memcpy(block_ptr, b.block_ptr.data(), b.block_ptr.size() * sizeof(int64_t));
    if (structural_rank) *structural_rank = b.structural_rank;
    return b.num_blocks;
}

rs_host_factor* rs_host_factor_create(int32_t kind, int64_t n, const int64_t* ap,
                                      const int64_t* ai, const double* ax, double pivot_tol,
                                      int32_t ordering, int32_t use_btf) {
    rs_host_factor* h = new rs_host_factor();
    HostFactor& f = h->f;
    f.n = n;
    const bool natural = ordering == 3;
    const bool sparselu = kind == 1;
    const double tol = sparselu ? 0.0 : pivot_tol;
    if (!sparselu && use_btf && n >= 4) {
        f.btf = btf_decompose(n, ap, ai);
        f.has_btf = f.btf.num_blocks >= 2 && f.btf.max_block_size < n;
    }
    if (f.has_btf) {
        const i64 nb = f.btf.num_blocks;
        f.sym.resize(nb);
        f.num.resize(nb);
        for (i64 b = 0; b < nb; b++) {
            i64 bs = f.btf.block_ptr[b], be = f.btf.block_ptr[b + 1], sz = be - bs;
            std::vector<i64> bap(sz + 1, 0), bai;
            std::vector<double> bax;
            for (i64 lc = 0; lc < sz; lc++) {
                i64 gco = f.btf.col_perm[bs + lc];
                for (i64 q = ap[gco]; q < ap[gco + 1]; q++) {
                    i64 gro = ai[q];
                    if (gro >= n) continue;
                    i64 grn = f.btf.row_perm_inv[gro];
                    if (grn >= bs && grn < be) {
                        bai.push_back(grn - bs);
                        bax.push_back(ax[q]);
                    }
                }
                bap[lc + 1] = (i64)bai.size();
            }
            block_symbolic(sz, bap.data(), bai.data(), natural, f.sym[b]);
            factor_block_host(f.sym[b], bap.data(), bai.data(), bax.data(), nullptr, false, tol,
                              f.num[b]);
        }
    } else {
        f.sym.resize(1);
        f.num.resize(1);
        block_symbolic(n, ap, ai, natural, f.sym[0], ordering == 4);
        f.sparselu = sparselu;
        factor_block_host(f.sym[0], ap, ai, ax, nullptr, false, tol, f.num[0], f.sparselu);
    }
    build_plan(f, ap, ai, h->p);
    return h;
}

void rs_host_factor_destroy(rs_host_factor* h) { delete h; }

int64_t rs_host_factor_nblocks(const rs_host_factor* h) { return (int64_t)h->f.sym.size(); }

void rs_host_factor_block_sizes(const rs_host_factor* h, int64_t b, int64_t* n, int64_t* nl,
                                int64_t* nu) {
    const BlockSymbolic& s = h->f.sym[b];
    *n = s.n;
    *nl = (int64_t)s.lri.size();
    *nu = (int64_t)s.uri.size();
}

void rs_host_factor_block_get(const rs_host_factor* h, int64_t b, int64_t* col_perm,
                              int64_t* row_perm, int64_t* lcp, int64_t* lri, int64_t* ucp,
                              int64_t* uri, double* lx, double* ux, double* udiag) {
    const BlockSymbolic& s = h->f.sym[b];
    const BlockNumeric& m = h->f.num[b];
    const i64 n = s.n;
    auto cp = [](int64_t* dst, const std::vector<i64>& v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(int64_t));
    };
    auto cpd = [](double* dst, const std::vector<double>& v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(double));
    };
    cp(col_perm, s.col_perm);
    cp(row_perm, m.row_perm);
    cp(lcp, s.lcp);
    cp(lri, s.lri);
    cp(ucp, s.ucp);
    cp(uri, s.uri);
    if (n == 1) {
        if (ux) ux[0] = m.ux[0];
        if (udiag) udiag[0] = m.udiag[0];
        return;
    }
    cpd(lx, m.lx);
    cpd(ux, m.ux);
    cpd(udiag, m.udiag);
}

void rs_host_factor_plan_stats(const rs_host_factor* h, int64_t* out) {
    const PlanHost& p = h->p;
    out[0] = p.nflev;
    out[1] = p.max_flev_width;
    out[2] = p.n_lsl;
    out[3] = p.n_usl;
    out[4] = p.nsc;
    out[5] = p.nod;
    out[6] = p.nl;
    out[7] = p.nu;
}

int32_t rs_host_stamp_pattern(const rs_device* dev, int64_t ndev, int64_t node_count,
                              int32_t mode, int64_t* n, int64_t* nnz, int64_t* ap, int64_t* ai) {
    StampPattern p;
    try {
        build_stamp_pattern(dev, ndev, node_count, mode == 0 ? MODE_DC : MODE_TRAN_BE, p);
    } catch (...) {
        return 6;
    }
    if (n) *n = p.n;
    if (nnz) *nnz = (int64_t)p.ai.size();
    if (ap) std::memcpy(ap, p.ap.data(), p.ap.size() * sizeof(int64_t));
    if (ai && !p.ai.empty()) std::memcpy(ai, p.ai.data(), p.ai.size() * sizeof(int64_t));
    return 0;
}

int32_t rs_host_sp_selftest(int32_t kind, int64_t n, const int64_t* ap, const int64_t* ai,
                            const double* ax0, const double* ax1, const double* rhs, double* x,
                            int64_t* stats) {
    rs_host_factor* h = nullptr;
    try {
        // kind bit 8: programs of the solver's own plan (positions after the
        // pivots, RPI the identity) instead of the identity-pivot plan
        const bool own = (kind & 0x100) != 0;
        kind &= 0xFF;
        h = rs_host_factor_create(kind, n, ap, ai, ax0, 0.001, 0, 1);
        PlanHost ph;
        if (own)
            ph = h->p;
        else
            build_identity_plan(n, ap, ai, kind, ph);
        SpProgram fac, sol;
        std::string why;
        if (!build_sp_programs(ph, fac, sol, &why)) {
            delete h;
            return RS_ERR_INVALID_MATRIX;
        }
        const SpLayout L = sp_layout(ph);
        std::vector<int32_t> rpi(n);
        for (int32_t b = 0; b < ph.nblk; b++) {
            const int32_t bs = ph.blk_ptr[b], be = ph.blk_ptr[b + 1];
            for (int32_t i = 0; i < be - bs; i++)
                rpi[bs + i] = own ? bs + i : bs + (int32_t)h->f.num[b].row_perm_inv[i];
        }
        std::vector<int32_t> ss;
        sp_scatter_slots(ph, rpi.data(), ss);
        std::vector<double> v((size_t)L.nslots + 1, 0.0);
        for (int32_t s = 0; s < (int32_t)ss.size(); s++) {
            if (ss[s] < 0) continue;
            double acc = 0.0;
            for (int32_t q = 0; q < ph.sc_cnt[s]; q++) acc += ax1[ph.sc_src[s] + q];
            v[ss[s]] = acc;
        }
        sp_emulate(fac, L, v.data(), rpi.data());
        for (int64_t i = 0; i < n; i++) v[L.t_base + rpi[ph.perm_in[i]]] = rhs[i];
        for (int32_t e = 0; e < ph.nod; e++) v[L.od_base + e] = ax1[ph.odax[e]];
        sp_emulate(sol, L, v.data(), rpi.data());
        for (int64_t i = 0; i < n; i++) x[i] = v[L.t_base + ph.perm_out[i]];
        if (stats) {
            stats[0] = fac.nsteps;
            stats[1] = fac.nops;
            stats[2] = sol.nsteps;
            stats[3] = sol.nops;
            stats[4] = fac.nslow;
            stats[5] = sol.nslow;
            stats[6] = fac.nfwd;
            stats[7] = sol.nfwd;
        }
        delete h;
        return RS_OK;
    } catch (const std::exception&) {
        delete h;
        return RS_ERR_INVALID_MATRIX;
    }
}

int32_t rs_host_rl_selftest(int32_t kind, int64_t n, const int64_t* ap, const int64_t* ai,
                            const double* ax0, const double* ax1, double* lx, double* ux, double* ud,
                            int64_t* stats) {
    rs_host_factor* h = nullptr;
    try {
        h = rs_host_factor_create(kind, n, ap, ai, ax0, 0.001, 0, 1);
        const PlanHost& p = h->p;
        RlHost r;
        if (!build_rl_plan(p, r)) {
            delete h;
            return RS_ERR_INVALID_MATRIX;
        }
        {  // the device form of the one-update records decodes to the records
            std::vector<int32_t> pool;
            PlanDev d{};
            append_rl(r, pool, d);
            const uint64_t mask = (1ull << kRlSlotBits) - 1;
            for (int32_t k = 0; k < r.nsgl; k++) {
                const uint64_t w = (uint64_t)(uint32_t)pool.at(d.o_rl_sgl + 2 * (size_t)k) |
                                   ((uint64_t)(uint32_t)pool.at(d.o_rl_sgl + 2 * (size_t)k + 1) << 32);
                if ((int32_t)(w & mask) != r.sgl[4 * (size_t)k] ||
                    (int32_t)((w >> kRlSlotBits) & mask) != r.sgl[4 * (size_t)k + 1] ||
                    (int32_t)(w >> (2 * kRlSlotBits)) != r.sgl[4 * (size_t)k + 2])
                    throw std::runtime_error("packed record");
            }
        }
        // the schedule, with every phase's independent work run in reverse
        // order: the values must not depend on the order inside a phase
        std::vector<double> V((size_t)r.nv, 0.0);
        const int32_t ux0 = p.nl + 1;
        for (int32_t a = 0; a < r.nv; a++) {
            double v = 0.0;
            for (int32_t q = 0; q < r.ini_cnt[a]; q++) v += ax1[r.ini_src[a] + q];
            V[a] = v;
        }
        auto clampd = [](double d, bool single) {
            if (single) return std::fabs(d) < 1e-300 ? 1e-14 : d;
            if (std::fabs(d) < 1e-14) return d >= 0.0 ? 1e-14 : -1e-14;
            return d;
        };
        std::vector<double> udv(n, 0.0);
        int64_t nlong = 0;
        for (int32_t ph = 0; ph < r.nph; ph++) {
            for (int32_t k = r.div_ptr[ph + 1] - 1; k >= r.div_ptr[ph]; k--)
                V[r.div[2 * k]] = V[r.div[2 * k]] / clampd(V[r.div[2 * k + 1]], false);
            for (int32_t k = r.piv_ptr[ph + 1] - 1; k >= r.piv_ptr[ph]; k--) {
                // the kernel's record: diagonal slot, singleton flag in bit 31
                const int32_t j = r.piv[k], pa = r.piv_a[k];
                const int32_t a = pa & 0x7FFFFFFF;
                if (a != ux0 + p.ucp[j + 1] - 1 || (pa < 0) != ((p.colflag[j] & COL_SINGLETON) != 0))
                    throw std::runtime_error("pivot record");
                V[a] = clampd(V[a], pa < 0);
                udv[j] = V[a];
            }
            nlong += r.task_nlong[ph];
            for (int32_t k = r.sgl_ptr[ph + 1] - 1; k >= r.sgl_ptr[ph]; k--) {
                const int32_t* g = &r.sgl[4 * (size_t)k];
                const double uu = V[g[2]];
                if (uu != 0.0) V[g[0]] = V[g[0]] - V[g[1]] * uu;
            }
            for (int32_t k = r.task_ptr[ph + 1] - 1; k >= r.task_ptr[ph]; k--) {
                double v = V[r.task_tgt[k]];
                for (int32_t q = r.task_beg[k]; q < r.task_beg[k + 1]; q++) {
                    const double uu = V[r.pair[2 * q + 1]];
                    if (uu != 0.0) v = v - V[r.pair[2 * q]] * uu;
                }
                V[r.task_tgt[k]] = v;
            }
        }
        if (p.nl) std::memcpy(lx, V.data(), p.nl * sizeof(double));
        if (p.nu) std::memcpy(ux, V.data() + ux0, p.nu * sizeof(double));
        if (n) std::memcpy(ud, udv.data(), n * sizeof(double));
        if (stats) {
            stats[0] = r.nph;
            stats[1] = (int64_t)r.ntask + r.nsgl;
            stats[2] = r.npair;
            stats[3] = nlong;
            stats[4] = p.nflev;
            stats[5] = p.nl;
            stats[6] = p.nu;
            stats[7] = r.nv;
        }
        delete h;
        return RS_OK;
    } catch (const std::exception&) {
        delete h;
        return RS_ERR_INVALID_MATRIX;
    }
}

int32_t rs_host_lp_selftest(int32_t kind, int64_t n, const int64_t* ap, const int64_t* ai, const double* ax,
                            const double* rhs, double* x, int64_t* stats) {
    rs_host_factor* h = nullptr;
    try {
        // kind bits 8..15: the ordering (rs_solver_set_config); bit 16: transposed
        h = rs_host_factor_create(kind & 0xFF, n, ap, ai, ax, 0.001, (kind >> 8) & 0xFF, 1);
        const PlanHost& p = h->p;
        std::vector<int32_t> pool;
        PlanDev d{};
        const bool transposed = (kind >> 16) & 1;  // J^T x = rhs through the transposed records
        std::vector<int32_t> pin = p.perm_in, pout = p.perm_out;
        if (transposed) {
            if (!build_transposed_records(p, pool, d)) {
                delete h;
                return RS_ERR_INVALID_MATRIX;
            }
            pin.assign(pool.begin() + d.o_perm_in, pool.begin() + d.o_perm_in + n);
            pout.assign(pool.begin() + d.o_perm_out, pool.begin() + d.o_perm_out + n);
        } else {
            append_level_records(p, pool, d);
        }
        if (d.sq_nlev == 0) {
            delete h;
            return RS_ERR_INVALID_MATRIX;
        }
        // value layout of the solver: lx | pad | ux | pad, offsets from lx
        std::vector<double> lx, ux, ud;
        gather_values(h->f, lx, ux, ud);
        std::vector<double> vals((size_t)p.nl + 1 + p.nu + 1, 0.0);
        std::copy(lx.begin(), lx.end(), vals.begin());
        std::copy(ux.begin(), ux.end(), vals.begin() + p.nl + 1);
        auto at = [&](int32_t off, size_t i) { return pool.at((size_t)off + i); };
        std::vector<double> T(n), P((size_t)d.sq_maxent);
        for (int64_t i = 0; i < n; i++) T.at(pin[i]) = rhs[i];
        int64_t slots = 0;
        for (int32_t l = 0; l < d.sq_nlev; l++) {
            const int32_t s0 = at(d.o_sq_lev, 4 * l), k0 = at(d.o_sq_lev, 4 * l + 1), div = at(d.o_sq_lev, 4 * l + 2);
            const int32_t s1 = at(d.o_sq_lev, 4 * l + 4), k1 = at(d.o_sq_lev, 4 * l + 5);
            if (s1 - s0 > d.sq_maxent || k1 <= k0) throw std::runtime_error("bad level");
            // the products in reverse order, then the rows in reverse order:
            // neither phase may depend on the order inside a level
            for (int32_t sl = s1 - 1; sl >= s0; sl--)
                P.at(sl - s0) = vals.at(at(d.o_sq_src, sl)) * T.at(at(d.o_sq_col, sl));
            for (int32_t k = k1 - 1; k >= k0; k--) {
                const int32_t r = at(d.o_sq_row, 4 * k), sb = at(d.o_sq_row, 4 * k + 1), se = at(d.o_sq_row, 4 * k + 2);
                if (sb < s0 || se > s1 || sb > se) throw std::runtime_error("bad row record");
                double t = T.at(r);
                for (int32_t sl = sb; sl < se; sl++) t = t - P.at(sl - s0);
                T.at(r) = div ? t / ud.at(r) : t;
            }
            slots += s1 - s0;
        }
        for (int64_t i = 0; i < n; i++) x[i] = T.at(pout[i]);
        if (stats) {
            stats[0] = d.sq_nlev;
            stats[1] = d.sq_maxent;
            stats[2] = slots;
            stats[3] = at(d.o_sq_lev, 4 * d.sq_nlev + 1);  // row records
            stats[4] = p.n_lsl + p.n_usl;                   // levels of the row sweeps
        }
        delete h;
        return RS_OK;
    } catch (const std::exception&) {
        delete h;
        return RS_ERR_INVALID_MATRIX;
    }
}

int32_t rs_host_sn_selftest(int64_t n, const int64_t* ap, const int64_t* ai, const double* ax,
                            const double* rhs, double* x, double* lx, double* ux, double* ud,
                            int64_t hot_min, int64_t flags, int64_t* stats) {
    try {
        BlockSymbolic s;
        block_symbolic(n, ap, ai, false, s);
        SnHost h;
        const bool transpose = (flags & 1) != 0;
        if (!build_sn_plan(s, ap, ai, h, nullptr, hot_min > 0 ? (int32_t)hot_min : kSnHotRecords, transpose))
            return RS_ERR_INVALID_MATRIX;
        if (stats) {
            stats[0] = h.ns;
            stats[1] = h.nlev;
            stats[2] = h.npairs;
            stats[3] = h.nvals;
            stats[4] = h.max_w;
            stats[5] = h.max_h;
            stats[6] = (int64_t)h.updates;
            stats[7] = (int64_t)(h.gitems.size() / 4);
            stats[8] = (int64_t)s.lri.size();
            stats[9] = (int64_t)s.uri.size();
            stats[10] = (int64_t)h.hot_slot.size();
            stats[11] = h.tile_rec.empty() ? 0 : h.tile_rec.back();
            stats[12] = h.symmetric ? 1 : 0;
        }
        if (!ax) return RS_OK;  // plan statistics only
        std::vector<double> v, d;
        sn_host_factor(h, ax, v, d);
        if (rhs && x) sn_host_solve(h, v, d, rhs, x);
        if (lx && ux && !transpose) {
            std::vector<double> l, u;
            sn_to_csc(h, s, v, l, u);
            std::copy(l.begin(), l.end(), lx);
            std::copy(u.begin(), u.end(), ux);
        }
        if (ud) std::copy(d.begin(), d.end(), ud);
        return RS_OK;
    } catch (const std::exception&) {
        return RS_ERR_INVALID_MATRIX;
    }
}

int64_t rs_host_bbd_decompose(int64_t n, const int64_t* ap, const int64_t* ai, int64_t num_blocks,
                              int64_t* perm, int64_t* block_ptr, int64_t* border_start) {
    BbdDecomp d;
    bbd_decompose(n, ap, ai, num_blocks, d);
    if (n) std::memcpy(perm, d.perm.data(), n * sizeof(int64_t));
    std::memcpy(block_ptr, d.block_ptr.data(), d.block_ptr.size() * sizeof(int64_t));
    if (border_start) *border_start = d.border_start;
    return d.num_blocks;
}

int32_t rs_host_pow_dd(const double* x, const double* y, int64_t n, double zone, double* out, uint8_t* near) {
    for (int64_t i = 0; i < n; i++) {
        bool nr = false;
        out[i] = rsgpu::pow_dd(x[i], y[i], &nr, zone);
        if (near) near[i] = nr ? 1 : 0;
    }
    return RS_OK;
}

}  // extern "C"
