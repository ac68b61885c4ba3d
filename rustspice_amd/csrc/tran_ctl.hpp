// tran_ctl.hpp - the adaptive transient's step controller on the device, so
// that Engine::run_tran_result_with_params (crates/sim-core/src/engine.rs:
// 463-602) runs as captured graphs with no host read per Newton iteration.
//
// A graph holds kTranSlots Newton-iteration slots. Each slot is
//   k_tctl_pre    (state machine: start an attempt: breakpoint and settling
//                  limits, clamp to tstop, restart from the accepted x)
//   eval, assemble, refactor, solve, k_newton_update   (one Newton iteration
//                  of the BE or TRAP solve the state names; all gated)
//   k_tctl_post   (a finished BE solve starts the TRAP solve; a finished TRAP
//                  solve gets Milne's LTE, the PI controller's verdict and
//                  the next dt; rejections as engine.rs:512-601)
//   copies, the accepted point into a ring of kTranRing points, and the two
//   capacitor histories (stamp.rs:572-668), gated on acceptance.
// The host launches the graph, reads the 200-byte state and the new points,
// and relaunches until the run has finished. The controller's powers
// (analysis.rs:173, 530-536) are evaluated in double-double and rounded once
// (ddpow.hpp); when the exact power lies within 0.02 ulp of a rounding
// midpoint (about 4 % of calls), where libm's pow (error bound 0.52 ulp) is
// not fixed by the exact value, the state machine stops with phase
// kPhaseHostPow and the host supplies the libm value before relaunching.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rsgpu {

constexpr int kTranSlots = 16;
constexpr int kTranRing = 16;  // > kTranSlots / 2 (an accepted step spends two slots at least)

enum TranPhase : int32_t {
    kPhaseBegin = 0,    // the next slot starts an attempt
    kPhaseBe = 1,       // BE Newton iterations
    kPhaseTrap = 2,     // TRAP Newton iterations
    kPhaseFinished = 3, // t reached tstop, or the run failed
    kPhaseHostPow = 4,  // waiting for the host to evaluate pw_in[] with libm
    kPhaseResume = 5    // the host wrote pw_out[]: finish the attempt's verdict
};

struct TranCtl {
    // k_newton_update's state block: done, converged, iterations, pad, damping,
    // |dx| of the previous iteration
    int32_t nw_done, nw_conv, nw_iters, nw_pad;
    double nw_damping, nw_prev;
    double t, dt, t_target, dt_step;  // dt_step: the dt of the attempt being solved
    double prev_err;                  // the PI controller's previous error (has_prev)
    double err;                       // Milne LTE ratio of the attempt (analysis.rs:147-200)
    double pw_in[3], pw_out[3];       // powers for the host: base with exponents k_p, k_i, 0.5
    int32_t phase, mode, status;      // status 0 running or converged, 2 failed
    int32_t accepted, rejected, consec, ctl_consec, has_prev, settle;
    int32_t accept_now;               // 1 in the slot that accepted a step
    int32_t cp_begin, cp_be, cp_acc;  // copy gates of this slot
    int32_t iters_total;
    int32_t lte_ok;                   // LTE verdict (max ratio <= 1)
    int32_t pw_need;                  // bits 0..2: pw_in[] entries the host must evaluate
    int64_t npoints;                  // stored points, the initial one included
};

struct TranParams {
    double tstop, min_dt, max_dt, abs_tol, rel_tol, damping;
    const double* bp;   // breakpoints, ascending (waveform.rs BreakpointManager)
    int32_t nbp;
    int32_t n;          // TRAN matrix size
};

// One slot's controller kernels (one workgroup each) and gated copies.
hipError_t launch_tctl_pre(TranCtl* c, TranParams p, hipStream_t s);
hipError_t launch_tctl_post(TranCtl* c, TranParams p, const double* x_be, const double* x, hipStream_t s);
// dst[i] = src[i] for i < n when *gate (grid-wide)
hipError_t launch_gated_copy(double* dst, const double* src, int64_t n, const int32_t* gate, hipStream_t s);
// on acceptance: xa = x and the ring's slot (npoints - 1) % kTranRing gets (t, x)
hipError_t launch_tstore(const TranCtl* c, const double* x, double* xa, double* ring_t, double* ring_x,
                         int64_t n, hipStream_t s);

}  // namespace rsgpu
