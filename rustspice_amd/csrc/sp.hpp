// sp.hpp - step programs: the refactor and the solve of one pattern as a
// host-scheduled list of scalar operations executed in lockstep by the 64
// lanes of one wavefront.
//
// Every operation of NativeKLU's left-looking refactor (native_klu.rs:1262-1347
// with reused pivots) and of its BTF/L/U solve (native_klu.rs:1491-1600) is
// one op on a value slot held in LDS:
//   UPD   v[d] = v[c] != 0 ? v[d] - v[b]*v[c] : v[d]   (work[r] -= l*u_jk, skip u_jk == 0)
//   SUB   v[d] = v[d] - v[b]*v[c]                       (triangular solve update)
//   DIV   v[d] = v[d] / v[b]                            (L column, U solve)
//   DIAG  clamp |v[d]| < 1e-14 to +-1e-14               (native_klu.rs:1328-1331)
//   DIAG1 |v[d]| < 1e-300 -> 1e-14                      (singleton block, :1215-1237)
//   ODSUB as SUB on slot t_base + RPI[d] (BTF off-diagonal row d, pre-pivot position)
// Each slot receives its updates in the reference's order, so the result is
// bit-identical; only independent operations are reordered.
//
// Execution model (what the scheduler guarantees and the kernel relies on):
// step s runs at most one op per lane; all lanes read their operands from LDS,
// then all lanes write. The LDS operations of one wavefront execute in program
// order, so an op may read any slot written at an earlier step, by any lane.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "plan.hpp"

namespace rsgpu {

enum SpKind : uint32_t {
    SP_NOP = 0,
    SP_UPD = 1,
    SP_SUB = 2,
    SP_ODSUB = 3,
    SP_DIV = 4,  // kinds >= SP_DIV need the slow path
    SP_DIAG = 5,
    SP_DIAG1 = 6,
};
constexpr int kSpLanes = 64;
constexpr int kSpChunk = 16;  // steps per prefetched chunk in the kernel
// op word: dst[0:16) b[16:32) c[32:48) kind[48:52)
inline uint64_t sp_word(uint32_t kind, uint32_t d, uint32_t b, uint32_t c) {
    return (uint64_t)d | ((uint64_t)b << 16) | ((uint64_t)c << 32) | ((uint64_t)kind << 48);
}

// Value slots: lx [0,nl) | ux [nl,nl+nu) | T [t_base, t_base+n) | off-diagonal ax [od_base, +nod)
struct SpLayout {
    int32_t nl = 0, nu = 0, n = 0, nod = 0, t_base = 0, od_base = 0, nslots = 0;
};

struct SpProgram {
    int32_t nsteps = 0;     // scheduled steps
    int32_t nsteps_pad = 0; // an even number of whole chunks, plus two chunks of NOPs
    int64_t nops = 0;
    std::vector<uint64_t> words;  // nsteps_pad * 64, step-major
};

SpLayout sp_layout(const PlanHost& p);

// Builds the refactor and solve programs of a plan whose positions are
// pre-pivot (the batched kernel's identity-pivot analysis). Returns false
// with a reason when a slot index does not fit in 16 bits.
bool build_sp_programs(const PlanHost& p, SpProgram& fac, SpProgram& sol, std::string* why);

// Slot of every scatter run for the given pivots (RPI[pre] = position), -1
// when the position lies outside the column's static pattern (dropped, as the
// reference's work vector entry is never read back).
void sp_scatter_slots(const PlanHost& p, const int32_t* rpi, std::vector<int32_t>& sslot);

// Host model of the kernel's execution of one program (tests).
void sp_emulate(const SpProgram& prog, const SpLayout& L, double* v, const int32_t* rpi);

}  // namespace rsgpu
