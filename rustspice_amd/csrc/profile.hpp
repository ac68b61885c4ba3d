// profile.hpp - optional per-kernel timing with HIP events recorded on the
// launch stream around each launch (enabled by rs_profile_enable).
#pragma once
#include <hip/hip_runtime.h>

namespace rsgpu {
// PK_SN_GATHER: each k_sn_gather launch inside a supernodal factor (the
// factor as a whole is also counted under PK_REFACTOR).
enum ProfKernel {
    PK_REFACTOR = 0, PK_SOLVE = 1, PK_EVAL = 2, PK_ASSEMBLE = 3, PK_BATCH = 4, PK_SN_GATHER = 5, PK_COUNT = 6
};
bool prof_on();
// Returns an event pair slot to record around a launch, or -1 when disabled.
int prof_begin(int kernel, hipStream_t s);
void prof_end(int slot, hipStream_t s);
}  // namespace rsgpu
