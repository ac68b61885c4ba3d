// sn.cpp - device state and launch sequence of the supernodal factor and
// solve (plan and schedule in sn.hpp / sn_plan.cpp, kernels in
// sn_kernels.hip).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "profile.hpp"
#include "sn.hpp"
#include "solver_impl.hpp"

namespace rsgpu {

hipError_t sn_launch_zero(double* v, int64_t n, hipStream_t s);
hipError_t sn_launch_scatter(const int32_t* src, const int32_t* cnt, const int64_t* slot, int64_t nruns,
                             const double* ax, double* v, hipStream_t s);
hipError_t sn_launch_gather(const SnPtrs& p, int item0, int nitems, hipStream_t s);
hipError_t sn_launch_hot_parts(const SnPtrs& p, const SnPart* parts, int nparts, const int64_t* terms,
                               hipStream_t s);
hipError_t sn_launch_fwd_parts(const SnPtrs& p, const SnPart* parts, int nparts, const int64_t* lslot,
                               const int32_t* yidx, double* y, hipStream_t s);
hipError_t sn_launch_scatter_long(const int32_t* src, const int32_t* cnt, const int64_t* slot, int nruns,
                                  const double* ax, double* v, hipStream_t s);
hipError_t sn_launch_fwd_long(const SnPtrs& p, const int32_t* rows, int nrows, const int64_t* tptr,
                              const int64_t* lslot, const int32_t* yidx, double* y, hipStream_t s);
hipError_t sn_launch_factor(const SnPtrs& p, const int32_t* list, int n, hipStream_t s);
hipError_t sn_launch_big(const SnPtrs& p, const int32_t* list, const int32_t* w, const int32_t* hh, int nbig,
                         hipStream_t s);
hipError_t sn_launch_perm(const int32_t* perm, const double* in, double* out, int n, bool to_pos, hipStream_t s);
hipError_t sn_launch_fwd(const SnPtrs& p, const int32_t* rows, int nrows, const int32_t* inner, int ninner,
                         int maxw, double* y, hipStream_t s);
hipError_t sn_launch_bwd(const SnPtrs& p, const int32_t* rows, const int32_t* sn_of, int nrows,
                         const int32_t* inner, int ninner, int maxw, double* y, hipStream_t s);

// Runs of duplicates and forward-solve rows with more terms than these are
// folded by a workgroup each (ordered long sums).
constexpr int32_t kSnLongRun = 64;
constexpr int64_t kSnLongRow = 256;
// forward rows with more terms than this are folded on the side stream
constexpr int64_t kSnStreamRow = 65536;

struct SnImpl {
    SnHost h;
    BlockSymbolic sym;
    // scatter: short runs (a thread each) and long runs (a workgroup each)
    DevBuf<int32_t> ss_src, ss_cnt, sl_src, sl_cnt;
    DevBuf<int64_t> ss_slot, sl_slot;
    int64_t n_short = 0, n_long = 0;
    // forward solve: per level the short rows [fr[l], fr[l+1]) of fwd_rows and
    // the long rows [fl[l], fl[l+1]) of long_rows with their flat terms
    DevBuf<int32_t> fwd_rows, long_rows, yidx;
    DevBuf<int64_t> long_tptr, lslot;
    std::vector<int32_t> fr, fl;
    DevBuf<int32_t> j0, rptr, rows, csrc, ca0, ca1, rel, perm, lev_sn, tile_hot_ptr, tile_hot, qrec, lrow_sn, lrow_t, sc_src, sc_cnt,
        sn_of, lev_rows, lists;
    DevBuf<int4> gitems;
    DevBuf<int64_t> oL, oU, crel_off, tile_rec, hot_slot, hot_ptr, hot_terms, lrow_ptr, sc_slot, qrec_ptr;
    DevBuf<SnRec> frec;
    DevBuf<double> v, ud, y;
    SnPtrs ptr{};
    // per level: rows [lr[l], lr[l+1]) of lev_rows; small-factor list
    // [fs[l], fs[l+1]) and inner list [in[l], in[l+1]) of lists; wide supernodes
    std::vector<int32_t> lr, fs, in;
    std::vector<int32_t> inw;  // per level: the widest supernode of its inner list
    std::vector<std::vector<int32_t>> big, big_w, big_h;  // per level, widest first
    std::vector<int32_t> big_off;                        // their lists' offsets in `lists`
    int64_t bytes = 0;
    // long ordered sums on a side stream (the ground node's diagonal and
    // forward row): their terms folded level by level as they become final,
    // in order, while the main stream factors or solves the next levels.
    // Factor: the hot entries' long scatter runs, then parts fparts
    // [fpp[l], fpp[l+1]) after level l; the main stream joins before
    // factoring a level whose fjoin is set. Solve: sparts likewise.
    hipStream_t side = nullptr;
    std::vector<hipEvent_t> ev;  // 4 nlev + 2: main/side per level, factor and solve
    DevBuf<int32_t> hl_src, hl_cnt;
    DevBuf<int64_t> hl_slot;
    int64_t n_hot_long = 0;
    DevBuf<SnPart> fparts, sparts;
    std::vector<int32_t> fpp, spp;
    std::vector<char> fjoin, sjoin;
    DevBuf<int64_t> st_lslot;
    DevBuf<int32_t> st_yidx;
    ~SnImpl() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        if (side) (void)hipStreamDestroy(side);
    }
};

namespace {

template <class T>
hipError_t up(SnImpl* p, DevBuf<T>& b, const std::vector<T>& v) {
    hipError_t e = b.ensure(std::max<size_t>(v.size(), 1));
    if (e || v.empty()) return e;
    p->bytes += (int64_t)(v.size() * sizeof(T));
    return hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

}  // namespace

SnImpl* sn_create(const BlockSymbolic& s, const i64* ap, const i64* ai, std::string* why, bool transpose) {
    SnImpl* p = new SnImpl();
    const char* vb = std::getenv("RSGPU_VERBOSE");
    const bool verbose = vb && vb[0] == '1';
    const auto t0 = std::chrono::steady_clock::now();
    if (!build_sn_plan(s, ap, ai, p->h, why, kSnHotRecords, transpose)) {
        delete p;
        return nullptr;
    }
    if (verbose)
        std::fprintf(stderr, "rsgpu: supernodal plan n=%d supernodes=%d levels=%d records=%lld in %.1f s\n", p->h.n,
                     p->h.ns, p->h.nlev, (long long)(p->h.tile_rec.empty() ? 0 : p->h.tile_rec.back()),
                     std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    p->sym = s;
    SnHost& h = p->h;
    // launch lists
    std::vector<int32_t> sn_of(h.n), lev_rows, lists_small, lists_inner;
    for (int32_t J = 0; J < h.ns; J++)
        for (int32_t c = h.j0[J]; c < h.j0[J + 1]; c++) sn_of[c] = J;
    p->lr.assign(1, 0);
    p->fs.assign(1, 0);
    p->in.assign(1, 0);
    p->inw.assign(h.nlev, 0);
    p->big.resize(h.nlev);
    for (int32_t l = 0; l < h.nlev; l++) {
        for (int32_t q = h.lev_ptr[l]; q < h.lev_ptr[l + 1]; q++) {
            const int32_t J = h.lev_sn[q], w = h.j0[J + 1] - h.j0[J];
            for (int32_t c = h.j0[J]; c < h.j0[J + 1]; c++) lev_rows.push_back(c);
            if (w >= kSnBigW) p->big[l].push_back(J);
            else lists_small.push_back(J);
            if (w > 1) {
                lists_inner.push_back(J);
                p->inw[l] = std::max(p->inw[l], w);
            }
        }
        p->lr.push_back((int32_t)lev_rows.size());
        p->fs.push_back((int32_t)lists_small.size());
        p->in.push_back((int32_t)lists_inner.size());
    }
    // scatter runs by length
    std::vector<int32_t> ss_src, ss_cnt, sl_src, sl_cnt, hl_src, hl_cnt;
    std::vector<int64_t> ss_slot, sl_slot, hl_slot;
    std::vector<int64_t> hots(h.hot_slot.begin(), h.hot_slot.end());
    std::sort(hots.begin(), hots.end());
    for (size_t r = 0; r < h.sc_src.size(); r++) {
        const bool lng = h.sc_cnt[r] > kSnLongRun;
        const bool hot = std::binary_search(hots.begin(), hots.end(), h.sc_slot[r]);
        // a hot entry's run: the side stream sums it (the gather never touches the slot)
        auto& src = hot ? hl_src : lng ? sl_src : ss_src;
        auto& cnt = hot ? hl_cnt : lng ? sl_cnt : ss_cnt;
        auto& slot = hot ? hl_slot : lng ? sl_slot : ss_slot;
        src.push_back(h.sc_src[r]);
        cnt.push_back(h.sc_cnt[r]);
        slot.push_back(h.sc_slot[r]);
    }
    p->n_short = (int64_t)ss_src.size();
    p->n_long = (int64_t)sl_src.size();
    p->n_hot_long = (int64_t)hl_src.size();
    // levels of the supernodes, for the side stream's parts
    std::vector<int32_t> lev_of(h.ns);
    for (int32_t l = 0; l < h.nlev; l++)
        for (int32_t q = h.lev_ptr[l]; q < h.lev_ptr[l + 1]; q++) lev_of[h.lev_sn[q]] = l;
    // factor parts: every hot entry's terms, cut where a term's source is
    // still to be factored
    std::vector<std::vector<SnPart>> fpl(h.nlev);
    p->fjoin.assign(h.nlev, 0);
    for (size_t e = 0; e < h.hot_slot.size(); e++) {
        const int64_t t1 = h.hot_ptr[e + 1];
        int64_t q = h.hot_ptr[e];
        const int32_t L = h.hot_lev[e];
        for (int32_t l = 0; l < L && q < t1; l++) {
            int64_t b = q;
            while (b < t1 && h.hot_tlev[b] <= l) b++;
            if (b > q) fpl[l].push_back(SnPart{q, b, (int32_t)e, 0});
            q = b;
        }
        if (q < t1) fpl[std::max(L - 1, 0)].push_back(SnPart{q, t1, (int32_t)e, 0});  // not expected
        p->fjoin[L] = 1;
    }
    // forward rows by the length of their term lists
    std::vector<int32_t> fwd_rows, long_rows, yidx, st_yidx;
    std::vector<int64_t> long_tptr(1, 0), lslot, st_lslot;
    std::vector<std::vector<SnPart>> spl(h.nlev);
    p->sjoin.assign(h.nlev, 0);
    p->fr.assign(1, 0);
    p->fl.assign(1, 0);
    for (int32_t l = 0; l < h.nlev; l++) {
        for (int32_t q = p->lr[l]; q < p->lr[l + 1]; q++) {
            const int32_t pos = lev_rows[q];
            int64_t terms = 0;
            for (int64_t e2 = h.lrow_ptr[pos]; e2 < h.lrow_ptr[pos + 1]; e2++)
                terms += h.j0[h.lrow_sn[e2] + 1] - h.j0[h.lrow_sn[e2]];
            if (terms <= kSnLongRow) {
                fwd_rows.push_back(pos);
                continue;
            }
            if (terms > kSnStreamRow) {  // folded on the side stream, level by level
                std::vector<int32_t> tl;
                const int64_t t0 = (int64_t)st_lslot.size();
                for (int64_t e2 = h.lrow_ptr[pos]; e2 < h.lrow_ptr[pos + 1]; e2++) {
                    const int32_t S = h.lrow_sn[e2], t = h.lrow_t[e2];
                    const int32_t wS = h.j0[S + 1] - h.j0[S], hS = h.rptr[S + 1] - h.rptr[S];
                    for (int32_t s2 = 0; s2 < wS; s2++) {
                        st_lslot.push_back(h.oL[S] + (int64_t)s2 * (wS + hS) + wS + t);
                        st_yidx.push_back(h.j0[S] + s2);
                        tl.push_back(lev_of[S]);
                    }
                }
                const int64_t t1 = (int64_t)st_lslot.size();
                int64_t q = t0;
                for (int32_t l2 = 0; l2 < l && q < t1; l2++) {
                    int64_t b = q;
                    while (b < t1 && tl[b - t0] <= l2) b++;
                    if (b > q) spl[l2].push_back(SnPart{q, b, pos, 0});
                    q = b;
                }
                if (q < t1) spl[std::max(l - 1, 0)].push_back(SnPart{q, t1, pos, 0});  // not expected
                p->sjoin[l] = 1;
                continue;
            }
            long_rows.push_back(pos);
            for (int64_t e2 = h.lrow_ptr[pos]; e2 < h.lrow_ptr[pos + 1]; e2++) {
                const int32_t S = h.lrow_sn[e2], t = h.lrow_t[e2];
                const int32_t wS = h.j0[S + 1] - h.j0[S], hS = h.rptr[S + 1] - h.rptr[S];
                for (int32_t s2 = 0; s2 < wS; s2++) {
                    lslot.push_back(h.oL[S] + (int64_t)s2 * (wS + hS) + wS + t);
                    yidx.push_back(h.j0[S] + s2);
                }
            }
            long_tptr.push_back((int64_t)lslot.size());
        }
        p->fr.push_back((int32_t)fwd_rows.size());
        p->fl.push_back((int32_t)long_rows.size());
    }
    std::vector<SnPart> fparts, sparts;
    p->fpp.assign(1, 0);
    p->spp.assign(1, 0);
    for (int32_t l = 0; l < h.nlev; l++) {
        fparts.insert(fparts.end(), fpl[l].begin(), fpl[l].end());
        sparts.insert(sparts.end(), spl[l].begin(), spl[l].end());
        p->fpp.push_back((int32_t)fparts.size());
        p->spp.push_back((int32_t)sparts.size());
    }
    const int32_t inner_base = (int32_t)lists_small.size();
    std::vector<int32_t> lists = lists_small;
    lists.insert(lists.end(), lists_inner.begin(), lists_inner.end());
    for (auto& x : p->in) x += inner_base;
    p->big_w.resize(h.nlev);
    p->big_h.resize(h.nlev);
    p->big_off.assign(h.nlev, 0);
    for (int32_t l = 0; l < h.nlev; l++) {
        auto& b = p->big[l];
        std::stable_sort(b.begin(), b.end(), [&](int32_t a, int32_t c) {
            return h.j0[a + 1] - h.j0[a] > h.j0[c + 1] - h.j0[c];
        });
        p->big_off[l] = (int32_t)lists.size();
        for (int32_t J : b) {
            lists.push_back(J);
            p->big_w[l].push_back(h.j0[J + 1] - h.j0[J]);
            p->big_h[l].push_back(h.rptr[J + 1] - h.rptr[J]);
        }
    }
    std::vector<int4> gi(h.gitems.size() / 4);
    std::memcpy(gi.data(), h.gitems.data(), gi.size() * sizeof(int4));
    hipError_t e;
    if ((e = up(p, p->j0, h.j0)) || (e = up(p, p->rptr, h.rptr)) || (e = up(p, p->rows, h.rows)) ||
        (e = up(p, p->csrc, h.csrc)) || (e = up(p, p->ca0, h.ca0)) || (e = up(p, p->ca1, h.ca1)) ||
        (e = up(p, p->rel, h.rel)) || (e = up(p, p->perm, h.perm)) || (e = up(p, p->lev_sn, h.lev_sn)) ||
        (e = up(p, p->frec, h.frec)) || (e = up(p, p->tile_hot_ptr, h.tile_hot_ptr)) ||
        (e = up(p, p->qrec_ptr, h.qrec_ptr)) || (e = up(p, p->qrec, h.qrec)) ||
        (e = up(p, p->tile_hot, h.tile_hot)) || (e = up(p, p->lrow_sn, h.lrow_sn)) || (e = up(p, p->lrow_t, h.lrow_t)) ||
        (e = up(p, p->sc_src, h.sc_src)) || (e = up(p, p->sc_cnt, h.sc_cnt)) || (e = up(p, p->sn_of, sn_of)) ||
        (e = up(p, p->lev_rows, lev_rows)) || (e = up(p, p->lists, lists)) || (e = up(p, p->gitems, gi)) ||
        (e = up(p, p->oL, h.oL)) || (e = up(p, p->oU, h.oU)) || (e = up(p, p->crel_off, h.crel_off)) ||
        (e = up(p, p->tile_rec, h.tile_rec)) || (e = up(p, p->hot_slot, h.hot_slot)) ||
        (e = up(p, p->hot_ptr, h.hot_ptr)) || (e = up(p, p->hot_terms, h.hot_terms)) ||
        (e = up(p, p->lrow_ptr, h.lrow_ptr)) || (e = up(p, p->sc_slot, h.sc_slot)) ||
        (e = up(p, p->ss_src, ss_src)) || (e = up(p, p->ss_cnt, ss_cnt)) || (e = up(p, p->ss_slot, ss_slot)) ||
        (e = up(p, p->sl_src, sl_src)) || (e = up(p, p->sl_cnt, sl_cnt)) || (e = up(p, p->sl_slot, sl_slot)) ||
        (e = up(p, p->fwd_rows, fwd_rows)) || (e = up(p, p->long_rows, long_rows)) || (e = up(p, p->yidx, yidx)) ||
        (e = up(p, p->long_tptr, long_tptr)) || (e = up(p, p->lslot, lslot)) ||
        (e = up(p, p->hl_src, hl_src)) || (e = up(p, p->hl_cnt, hl_cnt)) || (e = up(p, p->hl_slot, hl_slot)) ||
        (e = up(p, p->fparts, fparts)) || (e = up(p, p->sparts, sparts)) ||
        (e = up(p, p->st_lslot, st_lslot)) || (e = up(p, p->st_yidx, st_yidx)) ||
        (e = p->v.ensure((size_t)std::max<int64_t>(h.nvals, 1))) || (e = p->ud.ensure(h.n)) ||
        (e = p->y.ensure(h.n))) {
        if (why) *why = std::string("device upload: ") + hipGetErrorString(e);
        delete p;
        return nullptr;
    }
    p->bytes += (h.nvals + 2 * (int64_t)h.n) * (int64_t)sizeof(double);
    SnPtrs& q = p->ptr;
    q.j0 = p->j0.p; q.rptr = p->rptr.p; q.rows = p->rows.p; q.csrc = p->csrc.p; q.ca0 = p->ca0.p;
    q.ca1 = p->ca1.p; q.rel = p->rel.p; q.perm = p->perm.p; q.lev_sn = p->lev_sn.p; q.frec = p->frec.p;
    q.tile_hot_ptr = p->tile_hot_ptr.p; q.tile_hot = p->tile_hot.p; q.qrec_ptr = p->qrec_ptr.p; q.qrec = p->qrec.p;
    q.gitems = p->gitems.p; q.oL = p->oL.p; q.oU = p->oU.p; q.crel_off = p->crel_off.p;
    q.tile_rec = p->tile_rec.p; q.hot_slot = p->hot_slot.p; q.lrow_ptr = p->lrow_ptr.p;
    q.lrow_sn = p->lrow_sn.p; q.lrow_t = p->lrow_t.p; q.nhot = (int32_t)h.hot_slot.size(); q.n = h.n;
    q.v = p->v.p; q.ud = p->ud.p;
    if (!fparts.empty() || !sparts.empty() || p->n_hot_long) {
        // the side stream, default priority; events without timing
        if ((e = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking))) {
            if (why) *why = std::string("side stream: ") + hipGetErrorString(e);
            delete p;
            return nullptr;
        }
        p->ev.assign(4 * (size_t)h.nlev + 2, nullptr);
        for (auto& x : p->ev)
            if ((e = hipEventCreateWithFlags(&x, hipEventDisableTiming))) {
                if (why) *why = std::string("events: ") + hipGetErrorString(e);
                delete p;
                return nullptr;
            }
    }
    return p;
}

void sn_destroy(SnImpl* p) { delete p; }

hipError_t sn_factor(SnImpl* p, const double* ax, hipStream_t s) {
    const SnHost& h = p->h;
    hipError_t e;
    if ((e = sn_launch_zero(p->v.p, h.nvals, s)) ||
        (e = sn_launch_scatter(p->ss_src.p, p->ss_cnt.p, p->ss_slot.p, p->n_short, ax, p->v.p, s)) ||
        (e = sn_launch_scatter_long(p->sl_src.p, p->sl_cnt.p, p->sl_slot.p, (int)p->n_long, ax, p->v.p, s)))
        return e;
    const int32_t L = h.nlev;
    hipEvent_t* ev_main = p->ev.empty() ? nullptr : p->ev.data();   // [0, L): after level l (main)
    hipEvent_t* ev_side = p->ev.empty() ? nullptr : p->ev.data() + L;  // [L, 2L): after level l's parts (side)
    if (p->side) {
        // the hot entries' long runs start at once on the side stream
        if ((e = hipEventRecord(p->ev[4 * L], s)) || (e = hipStreamWaitEvent(p->side, p->ev[4 * L], 0)) ||
            (e = sn_launch_scatter_long(p->hl_src.p, p->hl_cnt.p, p->hl_slot.p, (int)p->n_hot_long, ax, p->v.p,
                                        p->side)))
            return e;
    }
    for (int32_t l = 0; l < L; l++) {
        const int ng = h.glev_ptr[l + 1] - h.glev_ptr[l];
        const int ps = ng > 0 ? prof_begin(PK_SN_GATHER, s) : -1;
        e = sn_launch_gather(p->ptr, h.glev_ptr[l], ng, s);
        prof_end(ps, s);
        if (e) return e;
        if (p->fjoin[l] && l > 0 && (e = hipStreamWaitEvent(s, ev_side[l - 1], 0))) return e;  // hot entries done
        if ((e = sn_launch_factor(p->ptr, p->lists.p + p->fs[l], p->fs[l + 1] - p->fs[l], s))) return e;
        if (!p->big[l].empty() &&
            (e = sn_launch_big(p->ptr, p->lists.p + p->big_off[l], p->big_w[l].data(), p->big_h[l].data(),
                               (int)p->big[l].size(), s)))
            return e;
        if (p->side) {
            const int np = p->fpp[l + 1] - p->fpp[l];
            if (np > 0 && ((e = hipEventRecord(ev_main[l], s)) || (e = hipStreamWaitEvent(p->side, ev_main[l], 0)) ||
                           (e = sn_launch_hot_parts(p->ptr, p->fparts.p + p->fpp[l], np, p->hot_terms.p, p->side))))
                return e;
            if ((e = hipEventRecord(ev_side[l], p->side))) return e;
        }
    }
    return hipSuccess;
}

hipError_t sn_solve(SnImpl* p, const double* rhs, double* x, hipStream_t s, double* y) {
    const SnHost& h = p->h;
    if (!y) y = p->y.p;
    hipError_t e;
    if ((e = sn_launch_perm(p->perm.p, rhs, y, h.n, true, s))) return e;
    const int32_t L = h.nlev;
    const bool st = p->side && p->spp[L] > 0;
    hipEvent_t* ev_main = st ? p->ev.data() + 2 * L : nullptr;
    hipEvent_t* ev_side = st ? p->ev.data() + 3 * L : nullptr;
    for (int32_t l = 0; l < L; l++) {
        if (st && p->sjoin[l] && l > 0 && (e = hipStreamWaitEvent(s, ev_side[l - 1], 0))) return e;
        if ((e = sn_launch_fwd_long(p->ptr, p->long_rows.p + p->fl[l], p->fl[l + 1] - p->fl[l],
                                    p->long_tptr.p + p->fl[l], p->lslot.p, p->yidx.p, y, s)) ||
            (e = sn_launch_fwd(p->ptr, p->fwd_rows.p + p->fr[l], p->fr[l + 1] - p->fr[l], p->lists.p + p->in[l],
                               p->in[l + 1] - p->in[l], p->inw[l], y, s)))
            return e;
        if (st) {
            const int np = p->spp[l + 1] - p->spp[l];
            if (np > 0 && ((e = hipEventRecord(ev_main[l], s)) || (e = hipStreamWaitEvent(p->side, ev_main[l], 0)) ||
                           (e = sn_launch_fwd_parts(p->ptr, p->sparts.p + p->spp[l], np, p->st_lslot.p,
                                                    p->st_yidx.p, y, p->side))))
                return e;
            if ((e = hipEventRecord(ev_side[l], p->side))) return e;
        }
    }
    for (int32_t l = h.nlev - 1; l >= 0; l--)
        if ((e = sn_launch_bwd(p->ptr, p->lev_rows.p + p->lr[l], p->sn_of.p, p->lr[l + 1] - p->lr[l],
                               p->lists.p + p->in[l], p->in[l + 1] - p->in[l], p->inw[l], y, s)))
            return e;
    return sn_launch_perm(p->perm.p, y, x, h.n, false, s);
}

hipError_t sn_values(SnImpl* p, double* lx, double* ux, double* ud) {
    std::vector<double> v((size_t)p->h.nvals);
    hipError_t e;
    if ((e = hipDeviceSynchronize()) ||
        (!v.empty() && (e = hipMemcpy(v.data(), p->v.p, v.size() * sizeof(double), hipMemcpyDeviceToHost))) ||
        (ud && (e = hipMemcpy(ud, p->ud.p, p->h.n * sizeof(double), hipMemcpyDeviceToHost))))
        return e;
    std::vector<double> l, u;
    sn_to_csc(p->h, p->sym, v, l, u);
    if (lx) std::copy(l.begin(), l.end(), lx);
    if (ux) std::copy(u.begin(), u.end(), ux);
    return hipSuccess;
}

const double* sn_pivots(const SnImpl* p) { return p->ud.p; }
const SnHost& sn_plan(const SnImpl* p) { return p->h; }
const BlockSymbolic& sn_symbolic(const SnImpl* p) { return p->sym; }
int64_t sn_device_bytes(const SnImpl* p) { return p->bytes; }

}  // namespace rsgpu
