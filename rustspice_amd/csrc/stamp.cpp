// stamp.cpp - symbolic replay of the reference stamp order; see stamp.hpp.
#include "stamp.hpp"

#include <algorithm>
#include <stdexcept>

namespace rsgpu {

namespace {
struct Ins {
    i64 col, row;
    int32_t slot;  // -1 = ground pin
};
}  // namespace

i64 stamp_slot_bases(const rs_device* dev, i64 ndev, std::vector<int32_t>& sbase) {
    sbase.resize(ndev);
    i64 next = ndev * kSlotsPerDevice;
    for (i64 d = 0; d < ndev; d++) {
        if (is_bsim(dev[d])) {
            sbase[d] = (int32_t)next;
            next += kBsimSlots;
        } else {
            sbase[d] = (int32_t)(d * kSlotsPerDevice);
        }
        if (next > INT32_MAX) throw std::runtime_error("too many devices");
    }
    return next;
}

void build_stamp_pattern(const rs_device* dev, i64 ndev, i64 node_count, int mode,
                         StampPattern& out, const int32_t* sbase, const uint8_t* gmask) {
    out = StampPattern();
    out.node_count = node_count;
    out.aux_of.assign(ndev, -1);
    std::vector<Ins> ins;
    std::vector<std::pair<i64, int32_t>> rhs;  // (row, slot) in insertion order
    i64 n = node_count;
    auto add = [&](i64 i, i64 j, int32_t slot) {  // StampContext::add -> insert(col=j,row=i)
        if (j >= n || i >= n || i < 0 || j < 0) return;  // SparseBuilder::insert drops
        ins.push_back({j, i, slot});
    };
    auto add_rhs = [&](i64 i, int32_t slot) {
        if (i >= 0 && i < n) rhs.emplace_back(i, slot);
    };
    const bool dc = mode == MODE_DC;
    for (i64 d = 0; d < ndev; d++) {
        const rs_device& v = dev[d];
        if (d * kSlotsPerDevice + kSlotsPerDevice > INT32_MAX) throw std::runtime_error("too many devices");
        const int32_t b = sbase ? sbase[d] : (int32_t)(d * kSlotsPerDevice);
        const i64 a = v.node[0], c = v.node[1];
        switch (v.kind) {
        case RS_DEV_R:  // stamp.rs:242-259
        case RS_DEV_D:  // stamp.rs:351-387
            add(a, a, b + 0); add(c, c, b + 1); add(a, c, b + 2); add(c, a, b + 3);
            if (v.kind == RS_DEV_D) { add_rhs(a, b + 4); add_rhs(c, b + 5); }
            break;
        case RS_DEV_C:  // stamp.rs:759-787 (BE), 871-915 (TRAP); open in DC
            if (dc) break;
            add(a, a, b + 0); add(c, c, b + 1); add(a, c, b + 2); add(c, a, b + 3);
            add_rhs(a, b + 4); add_rhs(c, b + 5);
            break;
        case RS_DEV_I:  // stamp.rs:261-276, 331-349
            if (dc && (v.flags & RS_F_WAVEFORM)) break;
            add_rhs(a, b + 0); add_rhs(c, b + 1);
            break;
        case RS_DEV_V: {  // stamp.rs:278-325; aux allocated before the adds
            if (dc && (v.flags & RS_F_WAVEFORM)) break;
            i64 k = n;
            n++;  // allocate_aux resizes the builder and rhs
            out.aux_of[d] = (int32_t)k;
            add(a, k, b + 0); add(c, k, b + 1); add(k, a, b + 2); add(k, c, b + 3);
            add_rhs(k, b + 4);
            break;
        }
        case RS_DEV_M: {  // stamp.rs:389-553
            const i64 dr = v.node[0], ga = v.node[1], so = v.node[2], bu = v.node[3];
            add(dr, dr, b + 0); add(so, so, b + 1); add(dr, so, b + 2); add(so, dr, b + 3);
            add(dr, ga, b + 4); add(dr, so, b + 5); add(so, ga, b + 6); add(so, so, b + 7);
            if (!is_bsim(v)) {  // level 1: gmbs = 0, no bulk stamps
                add_rhs(dr, b + 8); add_rhs(so, b + 9);
                break;
            }
            if (!sbase) throw std::runtime_error("BSIM device without a slot table");
            const uint8_t gm = gmask ? gmask[d] : 0;
            if (gm & 1) { add(dr, bu, b + 8); add(dr, so, b + 9); add(so, bu, b + 10); add(so, so, b + 11); }
            add_rhs(dr, b + 12); add_rhs(so, b + 13);
            if (gm & 2) {
                add(dr, dr, b + 14); add(bu, bu, b + 15); add(dr, bu, b + 16); add(bu, dr, b + 17);
                add_rhs(dr, b + 18); add_rhs(bu, b + 19);
            }
            if (gm & 4) {
                add(ga, ga, b + 20); add(so, so, b + 21); add(ga, so, b + 22); add(so, ga, b + 23);
                add_rhs(ga, b + 24); add_rhs(so, b + 25);
            }
            if (gm & 8) {
                add(ga, ga, b + 26); add(dr, dr, b + 27); add(ga, dr, b + 28); add(dr, ga, b + 29);
                add_rhs(ga, b + 30); add_rhs(dr, b + 31);
            }
            break;
        }
        default:
            throw std::runtime_error("unsupported device kind");
        }
    }
    add(0, 0, -1);  // ground pin (engine.rs:279-280)
    out.n = n;
    // finalize (mna.rs:320-347): per column, stable sort by row
    std::vector<i64> cnt(n + 1, 0);
    for (auto& e : ins) cnt[e.col + 1]++;
    for (i64 j = 0; j < n; j++) cnt[j + 1] += cnt[j];
    out.ap = cnt;
    std::vector<Ins> sorted(ins.size());
    {
        std::vector<i64> f(cnt.begin(), cnt.end() - 1);
        for (auto& e : ins) sorted[f[e.col]++] = e;
    }
    for (i64 j = 0; j < n; j++)
        std::stable_sort(sorted.begin() + out.ap[j], sorted.begin() + out.ap[j + 1],
                         [](const Ins& x, const Ins& y) { return x.row < y.row; });
    out.ai.resize(sorted.size());
    out.ax_src.resize(sorted.size());
    for (size_t q = 0; q < sorted.size(); q++) {
        out.ai[q] = sorted[q].row;
        out.ax_src[q] = sorted[q].slot;
    }
    out.rhs_ptr.assign(n + 1, 0);
    for (auto& r : rhs) out.rhs_ptr[r.first + 1]++;
    for (i64 i = 0; i < n; i++) out.rhs_ptr[i + 1] += out.rhs_ptr[i];
    out.rhs_src.resize(rhs.size());
    std::vector<int32_t> f(out.rhs_ptr.begin(), out.rhs_ptr.end() - 1);
    for (auto& r : rhs) out.rhs_src[f[r.first]++] = r.second;
}

}  // namespace rsgpu
