// ac.hpp - AC (small-signal) sweep: every frequency's complex MNA assembled
// and solved by a dense complex LU in one launch, one workgroup per frequency.
//
// The reference (Engine::run_ac_result, crates/sim-core/src/engine.rs:864-949)
// loops over frequencies on one core; the frequencies are independent, so
// here they are the unit of parallelism. Per frequency the kernel
//   - sums each dense entry's stamp contributions in the order the reference
//     inserts them (stamp.rs:1525-1747 into complex_mna.rs's per-column lists,
//     stable-sorted by row, then `a[row][col] += v` in that order);
//   - factors with ComplexDenseSolver's partial pivoting (complex_solver.rs:
//     50-136): |z| = hypot, first row of the largest modulus, singular below
//     1e-30, rows swapped whole, multipliers stored in place;
//   - back-substitutes each row's terms in ascending column order.
// Complex operations use num-complex 0.4.6's formulas with separately rounded
// products (build flag -ffp-contract=off).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/rs_circuit.h"

namespace rsgpu {

// One contribution: value index << 2 | negate << 1 | imaginary. A real
// contribution adds (+-vals[i], 0); an imaginary one adds (0, +-(omega * vals[i])).
struct AcDev {
    int32_t n = 0;      // complex MNA size (nodes + V-source aux rows)
    int32_t nent = 0;   // dense entries with contributions
    const int32_t* ent_pos = nullptr;  // row * n + col
    const int32_t* ent_ptr = nullptr;  // nent + 1
    const int32_t* con = nullptr;
    const double* vals = nullptr;      // base values (device evaluation, then 1.0)
    const double* rhs = nullptr;       // 2n: the excitation (frequency independent)
    const double* freqs = nullptr;
};

// Small-signal values at the DC point: per device 3 doubles (R, D: g; C: c;
// M: gds floored at 1e-12, gm, gmbs) and a mask (bit 0: gmbs is stamped).
hipError_t launch_ac_eval(const rs_device* dev, int64_t ndev, const rs_bsim_model* models,
                          const double* x_dc, int64_t nx, double* vals, uint8_t* mask,
                          hipStream_t s);
// Frequencies [f0, f0 + nf): a is nf * n * n complex (scratch), x nf * n
// complex (the solutions), status nf (1 solved, 0 singular pivot).
hipError_t launch_ac_solve(const AcDev& p, int64_t f0, int64_t nf, double* a, double* x,
                           int32_t* status, hipStream_t s);

}  // namespace rsgpu
