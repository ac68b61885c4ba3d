// profile.hpp - optional per-kernel timing with HIP events recorded on the
// launch stream around each launch (enabled by rs_profile_enable).
#pragma once
#include <hip/hip_runtime.h>

namespace rsgpu {
// PK_SN_GATHER: each k_sn_gather launch inside a supernodal factor (the
// factor as a whole is also counted under PK_REFACTOR). PK_BBD_SCHUR: a
// BbdSolver's dense Schur product (k_bbd_schur, or the MFMA tiles with their
// packing and chunk sum).
enum ProfKernel {
    PK_REFACTOR = 0, PK_SOLVE = 1, PK_EVAL = 2, PK_ASSEMBLE = 3, PK_BATCH = 4, PK_SN_GATHER = 5, PK_BBD_SCHUR = 6, PK_COUNT = 7
};
bool prof_on();
// Returns an event pair slot to record around a launch, or -1 when disabled.
int prof_begin(int kernel, hipStream_t s);
void prof_end(int slot, hipStream_t s);
}  // namespace rsgpu
