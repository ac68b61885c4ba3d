// stamp.hpp - device stamping (the DeviceStamp contract of
// crates/sim-core/src/stamp.rs:94-121 run for a whole circuit at once).
//
// Host side, once per (circuit, mode): replay the reference's stamp order
// symbolically (instances in order, each stamp's add(i, j, v) =
// SparseBuilder::insert(col=j, row=i), aux rows allocated on first use, the
// ground pin last, then a stable per-column sort by row) and record for every
// entry of the resulting unmerged CSC which device stamp produced it. RHS rows
// get the ordered list of their contributions the same way.
// Device side, every Newton iteration: one thread per device evaluates its
// stamp values into a small per-device array; one thread per CSC entry copies
// its value (ax is the reference's unmerged array, bit for bit), one thread
// per row sums its RHS contributions from 0.0 in insertion order.
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/rs_circuit.h"

namespace rsgpu {

using i64 = int64_t;

enum StampMode { MODE_DC = 0, MODE_TRAN_BE = 1, MODE_TRAN_TRAP = 2 };

// Fixed per-device slot block in the stamp-value array: matrix values first,
// then RHS values.
constexpr int kSlotsPerDevice = 10;

struct StampPattern {
    i64 n = 0;          // matrix size: node_count + aux rows of this mode
    i64 node_count = 0;
    std::vector<i64> ap, ai;        // unmerged CSC (as the engine hands it over)
    std::vector<int32_t> ax_src;    // slot index per CSC entry (-1: ground pin 1.0)
    std::vector<int32_t> rhs_ptr;   // n + 1
    std::vector<int32_t> rhs_src;   // slot index per contribution, insertion order
    std::vector<int32_t> aux_of;    // per device: aux row (node_count + k) or -1
};

// Builds the pattern of one mode. DC skips capacitors and waveform sources
// (stamp.rs:137, 261-300); TRAN stamps every device.
void build_stamp_pattern(const rs_device* dev, i64 ndev, i64 node_count, int mode,
                         StampPattern& out);

// Device-side stamp parameters per call. flags & kStampJacobian: MOSFET
// matrix values are the exact partial derivatives of the stamped current
// (used by the adjoint), not the Newton linearisation of stamp.rs.
constexpr int32_t kStampJacobian = 1;
struct StampArgs {
    double t, dt, gmin, source_scale;
    int32_t mode;
    int32_t flags;
};

}  // namespace rsgpu
