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
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/rs_circuit.h"

namespace rsgpu {

using i64 = int64_t;

enum StampMode { MODE_DC = 0, MODE_TRAN_BE = 1, MODE_TRAN_TRAP = 2 };

// Fixed per-device slot block in the stamp-value array: matrix values first,
// then RHS values.
constexpr int kSlotsPerDevice = 10;
// BSIM MOSFETs (level != 1) use a block of their own after all fixed blocks
// (the circuit's slot base table points there), in stamp.rs:389-553 order:
//   0-3 gds, 4-7 gm, 8-11 gmbs, 12-13 RHS (drain, source),
//   BSIM4 only: 14-17 gsub, 18-19 RHS (drain, bulk), 20-23 gigs, 24-25 RHS
//   (gate, source), 26-29 gigd, 30-31 RHS (gate, drain).
// The gmbs and BSIM4 groups are stamped only when their currents exceed the
// reference's thresholds; the evaluation reports which in a group mask (bit 0
// gmbs, 1 substrate, 2 gate-source, 3 gate-drain), and the pattern of a
// Newton iteration is the one built for the mask of every BSIM device.
constexpr int kBsimSlots = 32;
__host__ __device__ inline bool is_bsim(const rs_device& v) { return v.kind == RS_DEV_M && v.p[RS_MP_LEVEL] != 1.0; }

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
// (stamp.rs:137, 261-300); TRAN stamps every device. sbase: slot base of
// every device (NULL: d * kSlotsPerDevice, no BSIM devices); gmask: group
// mask of every device (BSIM devices only; NULL: all groups off).
void build_stamp_pattern(const rs_device* dev, i64 ndev, i64 node_count, int mode,
                         StampPattern& out, const int32_t* sbase = nullptr,
                         const uint8_t* gmask = nullptr);
// Slot bases: fixed blocks first, then one kBsimSlots block per BSIM device
// in device order. Returns the total slot count.
i64 stamp_slot_bases(const rs_device* dev, i64 ndev, std::vector<int32_t>& sbase);

// Device-side stamp parameters per call. flags & kStampJacobian: MOSFET
// matrix values are the exact partial derivatives of the stamped current
// (used by the adjoint), not the Newton linearisation of stamp.rs.
constexpr int32_t kStampJacobian = 1;
// stamp_dev only: the RHS rows assembled by a workgroup run on the circuit's
// side stream; rhs_join() makes the stream wait for them before d_rhs is read
constexpr int32_t kStampDeferLongRhs = 2;
struct StampArgs {
    double t, dt, gmin, source_scale;
    int32_t mode;
    int32_t flags;
    const rs_bsim_model* models = nullptr;  // BSIM model table (device memory) or NULL
    const double* t_dev = nullptr;          // time read from device memory instead of t
    const int32_t* skip = nullptr;          // evaluation skipped while *skip != 0
    // adaptive transient graphs: mode and dt read from device memory; in BE
    // mode the capacitors read the BE history arrays below instead of the
    // kernel's own (the reference keeps separate BE and TRAP states,
    // engine.rs:351-355)
    const int32_t* mode_dev = nullptr;
    const double* dt_dev = nullptr;
    const double* be_cap_v = nullptr;
    const double* be_cap_i = nullptr;
    const uint8_t* be_has_v = nullptr;
    const uint8_t* be_has_i = nullptr;
};

}  // namespace rsgpu
