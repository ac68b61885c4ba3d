// rlplan.cpp - right-looking schedule of the static-pivot refactor (see
// plan.hpp). The reference computes column g left-looking
// (native_klu.rs:1261-1373): work = A(:,g); for each j of U(:,g) in order,
// work[r] -= L(r,j) * work[j] for r in L(:,j) when work[j] != 0. Entry (r,g)
// thus receives its updates in U(:,g) order, from final values of L(r,j) and
// U(j,g). The schedule below keeps exactly that per-entry order.
#include <algorithm>
#include <stdexcept>

#include "plan.hpp"

namespace rsgpu {

bool build_rl_plan(const PlanHost& p, RlHost& r) {
    r = RlHost();
    const int32_t n = p.n, nl = p.nl, nu = p.nu;
    const int64_t nv = (int64_t)nl + 1 + nu + 1;
    if (nv + n >= INT32_MAX) return false;
    const int32_t ux0 = nl + 1;
    r.nv = (int32_t)nv;
    // 1. phases of the pivots
    std::vector<int32_t> lev(n, 0);
    int32_t nph = 0;
    for (int32_t g = 0; g < n; g++) {
        int32_t lv = 0;
        for (int32_t q = p.ucp[g]; q < p.ucp[g + 1] - 1; q++) lv = std::max(lv, lev[p.uri[q]] + 1);
        for (int32_t e = p.lrp[g]; e < p.lrp[g + 1]; e++) lv = std::max(lv, lev[p.lrc[e]] + 1);
        lev[g] = lv;
        nph = std::max(nph, lv + 1);
    }
    if (n == 0) nph = 0;
    r.nph = nph;
    // 2. per column: the value slot of each pattern row, the scatter of A,
    //    and the ordered updates grouped into (entry, phase) tasks
    std::vector<int32_t> mark(n, -1), slot(n, 0), upos(n, -1), phase(n, 0), cur(n, -1);
    r.ini_src.assign(nv, 0);
    r.ini_cnt.assign(nv, 0);
    struct Task { int32_t ph, tgt, beg, cnt; };
    std::vector<Task> tasks;
    std::vector<int32_t> tpairs;  // (l, u) in generation order, regrouped per column
    struct Up { int32_t row, ph, l, u; };
    std::vector<Up> ups;
    std::vector<int32_t> up_task, col_off;
    for (int32_t g = 0; g < n; g++) {
        const int32_t us = p.ucp[g], ue = p.ucp[g + 1] - 1;
        for (int32_t q = us; q <= ue; q++) {
            const int32_t row = p.uri[q];
            mark[row] = g;
            slot[row] = ux0 + q;
            upos[row] = q;
        }
        for (int32_t q = p.lcp[g]; q < p.lcp[g + 1]; q++) {
            const int32_t row = p.lri[q];
            mark[row] = g;
            slot[row] = q;
            upos[row] = -1;
        }
        for (int32_t s = p.sc_ptr[g]; s < p.sc_ptr[g + 1]; s++) {
            const int32_t row = p.sc_dst[s];
            if (mark[row] != g) return false;
            r.ini_src[slot[row]] = p.sc_src[s];
            r.ini_cnt[slot[row]] = p.sc_cnt[s];
        }
        ups.clear();
        for (int32_t q = us; q < ue; q++) {
            const int32_t j = p.uri[q];
            for (int32_t l = p.lcp[j]; l < p.lcp[j + 1]; l++) {
                const int32_t row = p.lri[l];
                if (mark[row] != g) return false;            // update outside the pattern
                if (upos[row] >= 0 && upos[row] < q) return false;  // U order not topological
                ups.push_back({row, 0, l, ux0 + q});
            }
        }
        // phases: running max per entry, in generation order
        for (const Up& u : ups) phase[u.row] = 0;
        for (Up& u : ups) {
            const int32_t j = p.uri[u.u - ux0];
            u.ph = std::max(phase[u.row], lev[j]);
            phase[u.row] = u.ph;
        }
        // tasks: runs of equal (entry, phase); pairs regrouped per task
        const int32_t t0 = (int32_t)tasks.size();
        for (const Up& u : ups) cur[u.row] = -1;
        up_task.resize(ups.size());
        for (size_t k = 0; k < ups.size(); k++) {
            const Up& u = ups[k];
            int32_t t = cur[u.row];
            if (t < 0 || tasks[t].ph != u.ph) {
                t = (int32_t)tasks.size();
                tasks.push_back({u.ph, slot[u.row], 0, 0});
                cur[u.row] = t;
            }
            tasks[t].cnt++;
            up_task[k] = t;
        }
        int32_t off = (int32_t)(tpairs.size() / 2);
        for (int32_t t = t0; t < (int32_t)tasks.size(); t++) {
            tasks[t].beg = off;
            off += tasks[t].cnt;
        }
        if ((int64_t)off * 2 >= INT32_MAX) return false;
        tpairs.resize((size_t)off * 2);
        col_off.assign(tasks.size() - t0, 0);
        for (size_t k = 0; k < ups.size(); k++) {
            const int32_t t = up_task[k];
            const int32_t at = tasks[t].beg + col_off[t - t0]++;
            tpairs[2 * (size_t)at] = ups[k].l;
            tpairs[2 * (size_t)at + 1] = ups[k].u;
        }
    }
    // 3. divisions and pivots by phase
    r.div_ptr.assign(nph + 1, 0);
    r.piv_ptr.assign(nph + 1, 0);
    for (int32_t j = 0; j < n; j++) {
        r.div_ptr[lev[j] + 1] += p.lcp[j + 1] - p.lcp[j];
        r.piv_ptr[lev[j] + 1]++;
    }
    for (int32_t k = 0; k < nph; k++) {
        r.div_ptr[k + 1] += r.div_ptr[k];
        r.piv_ptr[k + 1] += r.piv_ptr[k];
    }
    r.div.assign(2 * (size_t)r.div_ptr[nph], 0);
    r.piv.assign(r.piv_ptr[nph], 0);
    r.piv_a.assign(r.piv_ptr[nph], 0);
    {
        std::vector<int32_t> fd(r.div_ptr.begin(), r.div_ptr.end() - 1);
        std::vector<int32_t> fp(r.piv_ptr.begin(), r.piv_ptr.end() - 1);
        for (int32_t j = 0; j < n; j++) {
            const int32_t dslot = ux0 + p.ucp[j + 1] - 1;
            for (int32_t l = p.lcp[j]; l < p.lcp[j + 1]; l++) {
                const int32_t at = fd[lev[j]]++;
                r.div[2 * (size_t)at] = l;
                r.div[2 * (size_t)at + 1] = dslot;
            }
            r.piv_a[fp[lev[j]]] = dslot | ((p.colflag[j] & COL_SINGLETON) ? (int32_t)0x80000000u : 0);
            r.piv[fp[lev[j]]++] = j;
        }
    }
    // 4. tasks by phase: one-update tasks as records; the others with their
    //    pairs, long tasks first
    r.task_ptr.assign(nph + 1, 0);
    r.task_nlong.assign(nph, 0);
    r.sgl_ptr.assign(nph + 1, 0);
    for (const Task& t : tasks) {
        if (t.cnt == 1) {
            r.sgl_ptr[t.ph + 1]++;
            continue;
        }
        r.task_ptr[t.ph + 1]++;
        if (t.cnt >= kRlLongTask) r.task_nlong[t.ph]++;
    }
    for (int32_t k = 0; k < nph; k++) {
        r.task_ptr[k + 1] += r.task_ptr[k];
        r.sgl_ptr[k + 1] += r.sgl_ptr[k];
    }
    const int32_t ntask = r.task_ptr[nph];
    r.ntask = ntask;
    r.nsgl = r.sgl_ptr[nph];
    std::vector<int32_t> order(ntask);
    r.sgl.assign(4 * (size_t)r.nsgl, 0);
    {
        std::vector<int32_t> fl(nph), fs(nph), fg(r.sgl_ptr.begin(), r.sgl_ptr.end() - 1);
        for (int32_t k = 0; k < nph; k++) {
            fl[k] = r.task_ptr[k];
            fs[k] = r.task_ptr[k] + r.task_nlong[k];
        }
        for (int32_t t = 0; t < (int32_t)tasks.size(); t++) {
            const Task& x = tasks[t];
            if (x.cnt == 1) {
                const size_t at = 4 * (size_t)fg[x.ph]++;
                r.sgl[at] = x.tgt;
                r.sgl[at + 1] = tpairs[2 * (size_t)x.beg];
                r.sgl[at + 2] = tpairs[2 * (size_t)x.beg + 1];
                continue;
            }
            order[x.cnt >= kRlLongTask ? fl[x.ph]++ : fs[x.ph]++] = t;
        }
    }
    r.task_tgt.resize(ntask);
    r.task_beg.resize(ntask + 1);
    int64_t npair = 0;
    for (int32_t k = 0; k < ntask; k++) npair += tasks[order[k]].cnt;
    r.pair.resize(2 * (size_t)npair);
    int32_t at = 0;
    for (int32_t k = 0; k < ntask; k++) {
        const Task& x = tasks[order[k]];
        r.task_tgt[k] = x.tgt;
        r.task_beg[k] = at;
        std::copy(tpairs.begin() + 2 * (size_t)x.beg, tpairs.begin() + 2 * ((size_t)x.beg + x.cnt),
                  r.pair.begin() + 2 * (size_t)at);
        at += x.cnt;
    }
    r.task_beg[ntask] = at;
    r.npair = at + r.nsgl;
    return true;
}

void append_rl(const RlHost& r, std::vector<int32_t>& pool, PlanDev& d) {
    auto put = [&](const std::vector<int32_t>& v, bool even) -> int32_t {
        if (even && (pool.size() & 1)) pool.push_back(0);  // 8-byte aligned pairs
        size_t off = pool.size();
        if (off + v.size() + 2 > (size_t)INT32_MAX) throw std::runtime_error("plan pool too large");
        pool.insert(pool.end(), v.begin(), v.end());
        pool.push_back(0);
        pool.push_back(0);
        return (int32_t)off;
    };
    d.rl_nph = r.nph;
    d.rl_nv = r.nv;
    d.o_rl_ini_src = put(r.ini_src, false);
    d.o_rl_ini_cnt = put(r.ini_cnt, false);
    d.o_rl_div_ptr = put(r.div_ptr, false);
    d.o_rl_div = put(r.div, true);
    d.o_rl_piv_ptr = put(r.piv_ptr, false);
    d.o_rl_piv = put(r.piv, false);
    d.o_rl_piv_a = put(r.piv_a, false);
    d.o_rl_task_ptr = put(r.task_ptr, false);
    d.o_rl_task_nlong = put(r.task_nlong, false);
    d.o_rl_task_tgt = put(r.task_tgt, false);
    d.o_rl_task_beg = put(r.task_beg, false);
    d.o_rl_pair = put(r.pair, true);
    if (pool.size() & 3) pool.resize((pool.size() + 3) & ~(size_t)3, 0);  // 16-byte records
    d.o_rl_sgl_ptr = put(r.sgl_ptr, false);
    if (pool.size() & 3) pool.resize((pool.size() + 3) & ~(size_t)3, 0);
    // (target, L, U) in one little-endian 64-bit word, kRlSlotBits each
    if (r.nv >= (1 << kRlSlotBits)) throw std::runtime_error("right-looking schedule: too many value slots");
    std::vector<int32_t> packed(2 * (size_t)r.nsgl);
    for (int32_t k = 0; k < r.nsgl; k++) {
        const uint64_t w = (uint64_t)(uint32_t)r.sgl[4 * (size_t)k] |
                           ((uint64_t)(uint32_t)r.sgl[4 * (size_t)k + 1] << kRlSlotBits) |
                           ((uint64_t)(uint32_t)r.sgl[4 * (size_t)k + 2] << (2 * kRlSlotBits));
        packed[2 * (size_t)k] = (int32_t)(uint32_t)w;
        packed[2 * (size_t)k + 1] = (int32_t)(uint32_t)(w >> 32);
    }
    d.o_rl_sgl = put(packed, false);
}

}  // namespace rsgpu
