// sp.hpp - step programs: the refactor and the solve of one pattern as a
// host-scheduled list of scalar operations executed in lockstep by the 64
// lanes of one wavefront.
//
// Every operation of NativeKLU's left-looking refactor (native_klu.rs:1262-1347
// with reused pivots) and of its BTF/L/U solve (native_klu.rs:1491-1600) is
// one op on a value slot held in LDS:
//   UPD   v[d] = v[c] != 0 ? v[d] - v[b]*v[c] : v[d]   (work[r] -= l*u_jk, skip u_jk == 0)
//   SUB   v[d] = v[d] - v[b]*v[c]                       (triangular solve update)
//   DIV   v[d] = v[d] / v[c]                            (L column, U solve)
//   DIAG  clamp |v[d]| < 1e-14 to +-1e-14               (native_klu.rs:1328-1331)
//   DIAG1 |v[d]| < 1e-300 -> 1e-14                      (singleton block, :1215-1237)
//   ODSUB as SUB on slot t_base + RPI[d] (BTF off-diagonal row d, pre-pivot position)
// Each slot receives its updates in the reference's order, so the result is
// bit-identical; only independent operations are reordered. Idle lanes run a
// SUB on a dummy slot (the last one), so every lane writes at every step.
//
// Execution model (what the scheduler guarantees and the kernel relies on):
// step s runs at most one op per lane. A lane issues the LDS reads of its op
// of step s + 2 before it computes and writes its op of step s, so an operand
// read from LDS sees the writes of steps <= s - 3 (kSpLag). One operand of an
// op may instead be the result its lane computed at step s - 1 (forwarded in
// a register, field fwd: 1 a, 2 b, 3 c); dependency chains then advance one
// step per op instead of three. The LDS operations of one wavefront execute in
// program order, which makes the read/write ordering above exact.
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
constexpr int kSpChunk = 12;  // steps per prefetched chunk in the kernel (a multiple of 3)
constexpr int kSpLag = 3;     // steps between an LDS write and a read that sees it
enum SpFwd : uint32_t { SP_FWD_NONE = 0, SP_FWD_A = 1, SP_FWD_C = 2 };
constexpr int32_t kSpMaxSlots = 1 << 15;  // LDS byte addresses of slots fit in 18 bits
// One op of one lane: operand slots as LDS byte addresses (slot * 8), so the
// kernel uses them without decoding (ODSUB keeps its row r as r * 8 in d8),
// and a flag word: kind in bits 0-2, kSpSlowChunk (bit 3) on every word of a
// chunk that may hold kinds >= SP_DIV, SP_FWD_A as bit 31 (fl < 0 as a signed
// integer), SP_FWD_C as bit 16 (fl > 0xFFFF as a signed integer). Only slow
// chunks pay for the division and clamps; the scheduler places those kinds in
// slow chunks only.
struct SpOpWord {
    uint32_t d8, b8, c8, fl;
};
static_assert(sizeof(SpOpWord) == 16, "op word layout");
constexpr uint32_t kSpSlowChunk = 8u;
inline SpOpWord sp_word(uint32_t kind, uint32_t d, uint32_t b, uint32_t c, uint32_t fwd = 0) {
    uint32_t fl = kind;
    if (fwd == SP_FWD_A) fl |= 0x80000000u;
    if (fwd == SP_FWD_C) fl |= 0x10000u;
    return {d * 8u, b * 8u, c * 8u, fl};
}
inline uint32_t sp_kind(const SpOpWord& w) { return w.fl & 7u; }
inline uint32_t sp_fwd(const SpOpWord& w) {
    return (int32_t)w.fl < 0 ? SP_FWD_A : ((int32_t)w.fl > 0xFFFF ? SP_FWD_C : SP_FWD_NONE);
}

// Value slots: lx [0,nl) | ux [nl,nl+nu) | T [t_base, t_base+n) | off-diagonal ax
// [od_base, +nod) | dummy (slot nslots; nslots + 1 slots are allocated)
struct SpLayout {
    int32_t nl = 0, nu = 0, n = 0, nod = 0, t_base = 0, od_base = 0, nslots = 0;
};

struct SpProgram {
    int32_t nsteps = 0;     // scheduled steps
    int32_t nsteps_pad = 0; // an even number of whole chunks, plus two chunks of idle ops
    int64_t nops = 0;
    int64_t nfwd = 0;       // ops that take one operand by forwarding
    int32_t nslow = 0;      // slow chunks
    std::vector<SpOpWord> words;  // nsteps_pad * 64, step-major
};

SpLayout sp_layout(const PlanHost& p);

// Step programs of one solver's own plan (positions after its pivots, so the
// row map RPI is the identity), run by k_sp_refactor / k_sp_solve: offsets of
// the op words in a word buffer and of the scatter slots (nsc) and RPI (n) in
// an int buffer.
struct SpSolverDev {
    int32_t n, nl, nu, nod, nslots, t_base, od_base, nsc;
    int32_t fac_chunks, sol_chunks, o_ss, o_rpi;
    int32_t o_lrun, n_lrun;  // scatter runs longer than kSpShortRun, summed by a wavefront
    int64_t o_fac, o_sol;
};
constexpr int32_t kSpShortRun = 16;

// Builds the refactor and solve programs of a plan whose positions are
// pre-pivot (the batched kernel's identity-pivot analysis). Returns false
// with a reason when the slots exceed kSpMaxSlots.
bool build_sp_programs(const PlanHost& p, SpProgram& fac, SpProgram& sol, std::string* why);

// Slot of every scatter run for the given pivots (RPI[pre] = position), -1
// when the position lies outside the column's static pattern (dropped, as the
// reference's work vector entry is never read back).
void sp_scatter_slots(const PlanHost& p, const int32_t* rpi, std::vector<int32_t>& sslot);

// Host model of the kernel's execution of one program (tests); v holds
// nslots + 1 values.
void sp_emulate(const SpProgram& prog, const SpLayout& L, double* v, const int32_t* rpi);

}  // namespace rsgpu
