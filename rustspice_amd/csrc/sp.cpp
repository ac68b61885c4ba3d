// sp.cpp - construction, scheduling and host emulation of step programs.
#include "sp.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace rsgpu {
namespace {

enum Role : uint8_t { R_A = 1, R_B = 2, R_C = 4, R_WAR = 8, R_BAR = 16 };

struct SeqOp {
    uint32_t kind, d, b, c;
};

struct Pred {
    int32_t op;
    uint8_t role;
};

// Operand slots an op reads (the destination is always read-modify-write).
inline int reads_of(const SeqOp& o, uint32_t* slot, uint8_t* role, uint32_t pseudo_base) {
    int k = 0;
    slot[k] = o.kind == SP_ODSUB ? pseudo_base + o.d : o.d;
    role[k++] = R_A;
    if (o.kind == SP_UPD || o.kind == SP_SUB || o.kind == SP_ODSUB || o.kind == SP_DIV) {
        slot[k] = o.b;
        role[k++] = R_B;
    }
    if (o.kind == SP_UPD || o.kind == SP_SUB || o.kind == SP_ODSUB) {
        slot[k] = o.c;
        role[k++] = R_C;
    }
    return k;
}

// List scheduling of a sequential op list onto 64 lockstep lanes.
// extra[i] lists ops that must precede op i without forwarding (barriers).
void schedule(const std::vector<SeqOp>& ops, const std::vector<std::vector<int32_t>>& extra,
              uint32_t nslots, uint32_t npseudo, SpProgram& out) {
    const int32_t m = (int32_t)ops.size();
    std::vector<std::vector<Pred>> preds(m);
    std::vector<std::vector<int32_t>> succ(m);
    {
        const uint32_t total = nslots + npseudo;
        std::vector<int32_t> last_w(total, -1);
        std::vector<std::vector<int32_t>> readers(total);
        for (int32_t i = 0; i < m; i++) {
            uint32_t sl[3];
            uint8_t ro[3];
            const int k = reads_of(ops[i], sl, ro, nslots);
            auto add = [&](int32_t u, uint8_t r) {
                for (auto& p : preds[i])
                    if (p.op == u) {
                        p.role |= r;
                        return;
                    }
                preds[i].push_back({u, r});
            };
            for (int q = 0; q < k; q++) {
                if (last_w[sl[q]] >= 0) add(last_w[sl[q]], ro[q]);
            }
            const uint32_t dw = sl[0];
            for (int32_t r : readers[dw])
                if (r != i) add(r, R_WAR);
            for (int q = 0; q < k; q++)
                if (sl[q] != dw) readers[sl[q]].push_back(i);
            readers[dw].clear();
            last_w[dw] = i;
            for (int32_t u : extra[i]) add(u, R_BAR);
        }
    }
    for (int32_t i = 0; i < m; i++)
        for (auto& p : preds[i]) succ[p.op].push_back(i);
    // priority: longest path to a sink
    std::vector<int32_t> prio(m, 0);
    for (int32_t i = m - 1; i >= 0; i--)
        for (int32_t s : succ[i]) prio[i] = std::max(prio[i], prio[s] + 1);

    std::vector<int32_t> npend(m), ready, placed, keep;
    for (int32_t i = 0; i < m; i++) {
        npend[i] = (int32_t)preds[i].size();
        if (!npend[i]) ready.push_back(i);
    }
    std::vector<std::vector<int32_t>> rows;  // per step: ops, lane = position
    int32_t done = 0;
    while (done < m) {
        std::sort(ready.begin(), ready.end(), [&](int32_t x, int32_t y) {
            return prio[x] != prio[y] ? prio[x] > prio[y] : x < y;
        });
        const size_t take = std::min<size_t>(ready.size(), kSpLanes);
        placed.assign(ready.begin(), ready.begin() + take);
        keep.assign(ready.begin() + take, ready.end());
        for (int32_t v : placed)
            for (int32_t w : succ[v])
                if (--npend[w] == 0) keep.push_back(w);
        ready.swap(keep);
        done += (int32_t)placed.size();
        rows.push_back(placed);
    }
    out.nsteps = (int32_t)rows.size();
    out.nops = m;
    // whole chunks, an even number of them, plus two chunks of no-ops that the
    // kernel's double-buffered loop reads ahead
    const int32_t chunks = ((out.nsteps + kSpChunk - 1) / kSpChunk + 1) / 2 * 2;
    const int32_t padded = (chunks + 2) * kSpChunk;
    out.nsteps_pad = padded;
    out.words.assign((size_t)padded * kSpLanes, 0ull);
    for (int32_t s = 0; s < out.nsteps; s++)
        for (size_t l = 0; l < rows[s].size(); l++) {
            const SeqOp& o = ops[rows[s][l]];
            out.words[(size_t)s * kSpLanes + l] = sp_word(o.kind, o.d, o.b, o.c);
        }
}

int32_t find_sorted(const std::vector<int32_t>& a, int32_t lo, int32_t hi, int32_t key) {
    auto it = std::lower_bound(a.begin() + lo, a.begin() + hi, key);
    if (it == a.begin() + hi || *it != key) return -1;
    return (int32_t)(it - a.begin());
}

}  // namespace

SpLayout sp_layout(const PlanHost& p) {
    SpLayout L;
    L.nl = p.nl;
    L.nu = p.nu;
    L.n = p.n;
    L.nod = p.nod;
    L.t_base = p.nl + p.nu;
    L.od_base = L.t_base + p.n;
    L.nslots = L.od_base + p.nod;
    return L;
}

bool build_sp_programs(const PlanHost& p, SpProgram& fac, SpProgram& sol, std::string* why) {
    const SpLayout L = sp_layout(p);
    if (L.nslots > 65535 || p.n > 65535) {
        if (why) *why = "pattern has more than 65535 value slots";
        return false;
    }
    const int32_t n = p.n;
    // ---- refactor: columns in order, U entries ascending (native_klu.rs:1276-1291)
    {
        std::vector<SeqOp> ops;
        std::vector<int32_t> pos2slot(n, -1);
        for (int32_t k = 0; k < n; k++) {
            const int32_t us = p.ucp[k], ue = p.ucp[k + 1] - 1, ls = p.lcp[k], le = p.lcp[k + 1];
            const uint32_t dslot = (uint32_t)(L.nu ? p.nl + ue : 0);
            if (p.colflag[k] & 1) {
                ops.push_back({SP_DIAG1, dslot, 0, 0});
                continue;
            }
            for (int32_t u = us; u < ue; u++) pos2slot[p.uri[u]] = p.nl + u;
            pos2slot[k] = (int32_t)dslot;
            for (int32_t l = ls; l < le; l++) pos2slot[p.lri[l]] = l;
            for (int32_t u = us; u < ue; u++) {
                const int32_t j = p.uri[u];
                for (int32_t l = p.lcp[j]; l < p.lcp[j + 1]; l++) {
                    const int32_t t = pos2slot[p.lri[l]];
                    if (t < 0) throw std::runtime_error("step program: update outside the column pattern");
                    ops.push_back({SP_UPD, (uint32_t)t, (uint32_t)l, (uint32_t)(p.nl + u)});
                }
            }
            ops.push_back({SP_DIAG, dslot, 0, 0});
            for (int32_t l = ls; l < le; l++) ops.push_back({SP_DIV, (uint32_t)l, dslot, 0});
            for (int32_t u = us; u < ue; u++) pos2slot[p.uri[u]] = -1;
            pos2slot[k] = -1;
            for (int32_t l = ls; l < le; l++) pos2slot[p.lri[l]] = -1;
        }
        std::vector<std::vector<int32_t>> extra(ops.size());
        schedule(ops, extra, (uint32_t)L.nslots, 0, fac);
    }
    // ---- solve: blocks last to first; off-diagonal rows, L (cols ascending), U (cols descending)
    {
        std::vector<SeqOp> ops;
        std::vector<std::vector<int32_t>> extra;
        for (int32_t b = p.nblk - 1; b >= 0; b--) {
            const int32_t bs = p.blk_ptr[b], be = p.blk_ptr[b + 1];
            std::vector<int32_t> tails;
            if (p.blk_has_od[b]) {
                for (int32_t r = bs; r < be; r++) {
                    int32_t last = -1;
                    for (int32_t e = p.odp[r]; e < p.odp[r + 1]; e++) {
                        last = (int32_t)ops.size();
                        ops.push_back({SP_ODSUB, (uint32_t)r, (uint32_t)(L.od_base + e),
                                       (uint32_t)(L.t_base + p.odsrc[e])});
                        extra.emplace_back();
                    }
                    if (last >= 0) tails.push_back(last);
                }
            }
            for (int32_t r = bs; r < be; r++)
                for (int32_t e = p.lrp[r]; e < p.lrp[r + 1]; e++) {
                    ops.push_back({SP_SUB, (uint32_t)(L.t_base + r), (uint32_t)p.lrv[e],
                                   (uint32_t)(L.t_base + p.lrc[e])});
                    extra.push_back(tails);
                }
            for (int32_t r = be - 1; r >= bs; r--) {
                for (int32_t e = p.urp[r]; e < p.urp[r + 1]; e++) {
                    ops.push_back({SP_SUB, (uint32_t)(L.t_base + r), (uint32_t)(p.nl + p.urv[e]),
                                   (uint32_t)(L.t_base + p.urc[e])});
                    extra.push_back(tails);
                }
                ops.push_back({SP_DIV, (uint32_t)(L.t_base + r), (uint32_t)(p.nl + p.ucp[r + 1] - 1), 0});
                extra.push_back(tails);
            }
        }
        schedule(ops, extra, (uint32_t)L.nslots, (uint32_t)n, sol);
    }
    return true;
}

void sp_scatter_slots(const PlanHost& p, const int32_t* rpi, std::vector<int32_t>& sslot) {
    const int32_t n = p.n;
    sslot.assign(p.sc_ptr[n], -1);
    for (int32_t g = 0; g < n; g++) {
        const int32_t us = p.ucp[g], ue = p.ucp[g + 1] - 1;
        for (int32_t s = p.sc_ptr[g]; s < p.sc_ptr[g + 1]; s++) {
            const int32_t pos = rpi[p.sc_dst[s]];
            int32_t slot = -1;
            if (pos == g) {
                slot = p.nl + ue;
            } else if (pos < g) {
                const int32_t q = find_sorted(p.uri, us, ue, pos);
                if (q >= 0) slot = p.nl + q;
            } else {
                slot = find_sorted(p.lri, p.lcp[g], p.lcp[g + 1], pos);
            }
            sslot[s] = slot;
        }
    }
}

void sp_emulate(const SpProgram& prog, const SpLayout& L, double* v, const int32_t* rpi) {
    uint32_t da[kSpLanes];
    double a[kSpLanes], b[kSpLanes], c[kSpLanes];
    for (int s = 0; s < prog.nsteps; s++) {
        const uint64_t* w = &prog.words[(size_t)s * kSpLanes];
        for (int l = 0; l < kSpLanes; l++) {  // all reads of the step
            const uint32_t kind = (uint32_t)(w[l] >> 48) & 15u;
            const uint32_t d = (uint32_t)w[l] & 0xFFFFu;
            da[l] = kind == SP_ODSUB ? (uint32_t)(L.t_base + rpi[d]) : d;
            a[l] = v[da[l]];
            b[l] = v[(w[l] >> 16) & 0xFFFFu];
            c[l] = v[(w[l] >> 32) & 0xFFFFu];
        }
        for (int l = 0; l < kSpLanes; l++) {  // then all writes
            const uint32_t kind = (uint32_t)(w[l] >> 48) & 15u;
            if (kind == SP_NOP) continue;
            double r = a[l];
            switch (kind) {
                case SP_UPD:
                    if (c[l] != 0.0) r = a[l] - b[l] * c[l];
                    break;
                case SP_SUB:
                case SP_ODSUB:
                    r = a[l] - b[l] * c[l];
                    break;
                case SP_DIV:
                    r = a[l] / b[l];
                    break;
                case SP_DIAG:
                    if (std::fabs(a[l]) < 1e-14) r = a[l] >= 0.0 ? 1e-14 : -1e-14;
                    break;
                case SP_DIAG1:
                    if (std::fabs(a[l]) < 1e-300) r = 1e-14;
                    break;
                default:
                    break;
            }
            v[da[l]] = r;
        }
    }
}

}  // namespace rsgpu
