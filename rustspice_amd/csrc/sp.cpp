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
    if (o.kind == SP_UPD || o.kind == SP_SUB || o.kind == SP_ODSUB) {
        slot[k] = o.b;
        role[k++] = R_B;
    }
    if (o.kind == SP_UPD || o.kind == SP_SUB || o.kind == SP_ODSUB || o.kind == SP_DIV) {
        slot[k] = o.c;
        role[k++] = R_C;
    }
    return k;
}

// Forwarding field for a dependency of the given roles: exactly one operand,
// no barrier (a barrier orders against a write whose slot is not static).
inline uint32_t fwd_of(uint8_t role) {
    if (role & R_BAR) return SP_FWD_NONE;
    switch (role & (R_A | R_B | R_C)) {
        case R_A: return SP_FWD_A;
        case R_C: return SP_FWD_C;
        default: return SP_FWD_NONE;
    }
}

// List scheduling of a sequential op list onto 64 lockstep lanes under the
// read lag and forwarding rules of sp.hpp. extra[i] lists ops that must
// precede op i without forwarding (barriers).
void schedule(const std::vector<SeqOp>& ops, const std::vector<std::vector<int32_t>>& extra,
              uint32_t nslots, uint32_t npseudo, uint32_t idle_kind, SpProgram& out) {
    const int32_t m = (int32_t)ops.size();
    std::vector<std::vector<Pred>> preds(m);
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
    std::vector<std::vector<Pred>> succ(m);
    for (int32_t i = 0; i < m; i++)
        for (auto& p : preds[i]) succ[p.op].push_back({i, p.role});
    // priority: longest path to a sink; a forwardable edge costs one step, an
    // LDS dependency kSpLag, an ordering-only (write-after-read) edge none
    auto weight = [](uint8_t role) -> int32_t {
        if (!(role & (R_A | R_B | R_C | R_BAR))) return 0;
        return fwd_of(role) ? 1 : kSpLag;
    };
    std::vector<int32_t> prio(m, 0);
    for (int32_t i = m - 1; i >= 0; i--)
        for (const Pred& s : succ[i]) prio[i] = std::max(prio[i], prio[s.op] + weight(s.role));

    std::vector<int32_t> npend(m), step(m, -1), lane(m, -1);
    // per released op: earliest step without forwarding, and the one late
    // producer it could forward from (step(p1) + 1, on lane(p1))
    std::vector<int32_t> est(m, 0), p1(m, -1), fwd(m, 0);
    std::vector<uint32_t> fkind(m, 0);
    std::vector<int32_t> released, next;
    auto release = [&](int32_t v) {
        int32_t e = 0, best = -1, bestv = -1, second = 0;
        for (const Pred& p : preds[v]) {
            const int32_t t = (p.role & (R_A | R_B | R_C | R_BAR)) ? step[p.op] + kSpLag : step[p.op];
            e = std::max(e, t);
            if (t > bestv) {
                second = std::max(second, bestv);
                bestv = t;
                best = p.op;
            } else {
                second = std::max(second, t);
            }
        }
        est[v] = e;
        p1[v] = -1;
        if (best >= 0) {
            uint8_t role = 0;
            for (const Pred& p : preds[v])
                if (p.op == best) role = p.role;
            const uint32_t f = fwd_of(role);
            if (f && second <= step[best] + 1) {
                p1[v] = best;
                fkind[v] = f;
            }
        }
        released.push_back(v);
    };
    for (int32_t i = 0; i < m; i++) {
        npend[i] = (int32_t)preds[i].size();
        if (!npend[i]) release(i);
    }
    std::vector<std::vector<int32_t>> rows;  // per step: op on each lane (-1 none)
    std::vector<int32_t> cf, cn;
    int32_t done = 0;
    int64_t nfwd = 0;
    std::vector<uint8_t> slow_chunk;
    auto is_slow = [&](int32_t v) { return ops[v].kind >= SP_DIV; };
    for (int32_t s = 0; done < m; s++) {
        std::vector<int32_t> row(kSpLanes, -1);
        cf.clear();
        cn.clear();
        for (int32_t v : released) {
            if (est[v] <= s) cn.push_back(v);
            else if (p1[v] >= 0 && step[p1[v]] + 1 == s) cf.push_back(v);
        }
        // a chunk is slow when a slow op can start at its first step; fast
        // chunks hold no slow ops
        if (s % kSpChunk == 0) {
            bool any = false;
            for (int32_t v : cn) any = any || is_slow(v);
            for (int32_t v : cf) any = any || is_slow(v);
            slow_chunk.push_back(any ? 1 : 0);
        }
        if (!slow_chunk.back()) {
            cn.erase(std::remove_if(cn.begin(), cn.end(), is_slow), cn.end());
            cf.erase(std::remove_if(cf.begin(), cf.end(), is_slow), cf.end());
        }
        auto by_prio = [&](int32_t x, int32_t y) { return prio[x] != prio[y] ? prio[x] > prio[y] : x < y; };
        std::sort(cf.begin(), cf.end(), by_prio);
        std::sort(cn.begin(), cn.end(), by_prio);
        std::vector<int32_t> placed;
        for (int32_t v : cf) {
            const int32_t l = lane[p1[v]];
            if (row[l] >= 0) continue;
            row[l] = v;
            fwd[v] = (int32_t)fkind[v];
            placed.push_back(v);
        }
        int32_t l = 0;
        for (int32_t v : cn) {
            while (l < kSpLanes && row[l] >= 0) l++;
            if (l == kSpLanes) break;
            row[l] = v;
            fwd[v] = 0;
            placed.push_back(v);
        }
        for (int32_t q = 0; q < kSpLanes; q++)
            if (row[q] >= 0) {
                step[row[q]] = s;
                lane[row[q]] = q;
            }
        next.clear();
        for (int32_t v : released)
            if (step[v] < 0) next.push_back(v);
        released.swap(next);
        for (int32_t v : placed) {
            if (fwd[v]) nfwd++;
            for (const Pred& w : succ[v])
                if (--npend[w.op] == 0) release(w.op);
        }
        done += (int32_t)placed.size();
        rows.push_back(std::move(row));
    }
    out.nsteps = (int32_t)rows.size();
    out.nops = m;
    out.nfwd = nfwd;
    // whole chunks, an even number of them (the kernel's loop runs two
    // buffers), plus two chunks of idle ops that its read-ahead touches
    const int32_t chunks = ((out.nsteps + kSpChunk - 1) / kSpChunk + 1) / 2 * 2;
    const int32_t padded = (chunks + 2) * kSpChunk;
    out.nsteps_pad = padded;
    const uint32_t dummy = nslots;  // idle lanes: a SUB on the dummy slot
    out.words.assign((size_t)padded * kSpLanes, sp_word(idle_kind, dummy, dummy, dummy));
    for (int32_t s = 0; s < out.nsteps; s++)
        for (int32_t q = 0; q < kSpLanes; q++) {
            const int32_t v = rows[s][q];
            if (v < 0) continue;
            const SeqOp& o = ops[v];
            const bool rb = o.kind == SP_UPD || o.kind == SP_SUB || o.kind == SP_ODSUB;
            const bool rc = rb || o.kind == SP_DIV;
            out.words[(size_t)s * kSpLanes + q] =
                sp_word(o.kind, o.d, rb ? o.b : dummy, rc ? o.c : dummy, (uint32_t)fwd[v]);
        }
    out.nslow = 0;
    for (size_t c = 0; c < slow_chunk.size(); c++)
        if (slow_chunk[c]) {
            out.nslow++;
            for (size_t i = c * kSpChunk * kSpLanes; i < (c + 1) * kSpChunk * kSpLanes; i++)
                out.words[i].fl |= kSpSlowChunk;
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
    if (L.nslots + 1 > kSpMaxSlots) {
        if (why) *why = "pattern has more value slots than LDS can hold";
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
            for (int32_t l = ls; l < le; l++) ops.push_back({SP_DIV, (uint32_t)l, 0, dslot});
            for (int32_t u = us; u < ue; u++) pos2slot[p.uri[u]] = -1;
            pos2slot[k] = -1;
            for (int32_t l = ls; l < le; l++) pos2slot[p.lri[l]] = -1;
        }
        std::vector<std::vector<int32_t>> extra(ops.size());
        schedule(ops, extra, (uint32_t)L.nslots, 0, SP_UPD, fac);
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
                ops.push_back({SP_DIV, (uint32_t)(L.t_base + r), 0, (uint32_t)(p.nl + p.ucp[r + 1] - 1)});
                extra.push_back(tails);
            }
        }
        schedule(ops, extra, (uint32_t)L.nslots, (uint32_t)n, SP_SUB, sol);
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
    // operands of the steps s, s+1, s+2 in flight, as the kernel holds them
    struct Fetched {
        uint32_t d;
        double a, b, c;
    };
    std::vector<Fetched> q[3];
    for (auto& x : q) x.assign(kSpLanes, {0, 0.0, 0.0, 0.0});
    double r[kSpLanes] = {};
    auto at = [&](uint32_t b8) -> double& { return v[b8 / 8]; };
    auto fetch = [&](int s, std::vector<Fetched>& f) {
        if (s >= prog.nsteps_pad) return;
        const SpOpWord* w = &prog.words[(size_t)s * kSpLanes];
        for (int l = 0; l < kSpLanes; l++) {
            f[l].d = sp_kind(w[l]) == SP_ODSUB ? (uint32_t)(L.t_base + rpi[w[l].d8 / 8]) * 8u : w[l].d8;
            f[l].a = at(f[l].d);
            f[l].b = at(w[l].b8);
            f[l].c = at(w[l].c8);
        }
    };
    fetch(0, q[0]);
    fetch(1, q[1]);
    for (int s = 0; s < prog.nsteps; s++) {
        fetch(s + 2, q[(s + 2) % 3]);  // issued before this step's writes
        const SpOpWord* w = &prog.words[(size_t)s * kSpLanes];
        const std::vector<Fetched>& f = q[s % 3];
        for (int l = 0; l < kSpLanes; l++) {
            const uint32_t kind = sp_kind(w[l]);
            const uint32_t fw = sp_fwd(w[l]);
            const double a = fw == SP_FWD_A ? r[l] : f[l].a;
            const double b = f[l].b;
            const double c = fw == SP_FWD_C ? r[l] : f[l].c;
            double x = (kind == SP_UPD && c == 0.0) ? a : a - b * c;
            switch (kind) {
                case SP_DIV:
                    x = a / c;
                    break;
                case SP_DIAG:
                    x = std::fabs(a) < 1e-14 ? (a >= 0.0 ? 1e-14 : -1e-14) : a;
                    break;
                case SP_DIAG1:
                    x = std::fabs(a) < 1e-300 ? 1e-14 : a;
                    break;
                default:
                    break;
            }
            r[l] = x;
            at(f[l].d) = x;
        }
    }
}

}  // namespace rsgpu
