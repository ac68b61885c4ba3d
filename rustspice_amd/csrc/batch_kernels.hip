// batch_kernels.hip - persistent batched DC operating point kernel.
//
// One wavefront (a workgroup of 64) owns one instance at a time and runs its
// whole Newton loop (newton.rs:65-185 inside engine.rs:249-311): device
// evaluation and assembly (stamp.rs, mna.rs), the first numeric factor with
// threshold pivoting inside the static pattern and later refactors with the
// reused pivots (native_klu.rs:1194-1382), the BTF/L/U solve
// (native_klu.rs:1442-1551), the convergence test, damping and x update.
// Every floating-point operation is the reference's, in the reference's order.
//
// Latency is what bounds one instance, so each factor column runs from a
// precomputed column program: the (j, row, L index) pairs of its left-looking
// updates are loaded for up to 64 pairs at once, then applied group by group
// in ascending j with only LDS traffic. Work vector, x, x_new and the pivot
// arrays always live in LDS; the L/U values too when they fit (lumax > 0).
// The workgroup is one wavefront, whose LDS and global accesses execute in
// program order, so steps are separated by compiler fences only.
#include <hip/hip_runtime.h>

#include <climits>

#include "batch.hpp"
#include "device_eval.hpp"
#include "sp.hpp"

namespace {

using namespace rsgpu;

constexpr int kW = 64;
constexpr int kPhases = 8;  // eval, gather, rhs, factor, refactor, slots, solve, newton

// LDS exchange between the lanes of this wavefront: LDS instructions of one
// wavefront execute in program order, so a compiler fence is enough.
__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Global-memory exchange between lanes: workgroup-scope fence (waits for the
// wavefront's outstanding vector memory operations).
__device__ __forceinline__ void gsync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// ---- step programs (sp.hpp) --------------------------------------------------
// Operands of one step, read two steps ahead of its execution; d is the
// destination's LDS byte address.
struct SpIn {
    double a, b, c;
    uint32_t d;
};

// Step-program slots are addressed by their LDS byte address: the slots start
// the kernel's dynamic LDS, and the kernel has no static LDS (checked on the
// host before step programs are enabled), so slot byte offset = LDS address.
typedef __attribute__((address_space(3))) double lds_double;
__device__ __forceinline__ lds_double& vat(uint32_t b8) {
    return *reinterpret_cast<lds_double*>(static_cast<size_t>(b8));
}

// Issues the LDS reads of one step (w = {d8, b8, c8, flags}). An off-diagonal
// update's destination is the pivot position of its row (wavefront-uniform
// branch; kOd programs only).
template <bool kOd>
__device__ __forceinline__ void sp_fetch(const uint4& w, const int* RPI, int t_base, SpIn& f) {
    uint32_t d = w.x;
    if (kOd) {
        const uint32_t kind = w.w & 7u;
        if (__ballot(kind == SP_ODSUB)) {
            if (kind == SP_ODSUB) d = (uint32_t)(t_base + RPI[d >> 3]) << 3;
        }
    }
    f.d = d;
    f.a = vat(d);
    f.b = vat(w.y);
    f.c = vat(w.z);
}

// Executes one step. Operand a (flag word < 0) or c (flag word > 0xFFFF) may
// be the lane's previous result r. The plain op is UPD in a refactor program
// (kUpd: skip when c == 0) and SUB in a solve program; kinds >= SP_DIV occur
// in slow chunks only (kSlow), whose steps take a wavefront-uniform branch
// for them. Every lane writes (idle lanes to the dummy slot).
template <bool kUpd, bool kSlow>
__device__ __forceinline__ void sp_exec(const uint4& w, const SpIn& f, double& r) {
    const int32_t fl = (int32_t)w.w;
    const double a = fl < 0 ? r : f.a;
    const double c = fl > 0xFFFF ? r : f.c;
    double x = a - f.b * c;
    if (kUpd) x = c == 0.0 ? a : x;
    if (kSlow) {
        if (__ballot((fl & 7) >= (int32_t)SP_DIV)) {
            const int kind = fl & 7;
            if (kind == SP_DIV) x = a / c;
            if (kind == SP_DIAG) x = fabs(a) < 1e-14 ? (a >= 0.0 ? 1e-14 : -1e-14) : a;
            if (kind == SP_DIAG1) x = fabs(a) < 1e-300 ? 1e-14 : a;
        }
    }
    r = x;
    vat(f.d) = x;
}

// One chunk of kSpChunk steps from cur; steps 10 and 11 read ahead into nxt,
// whose loads (the chunk after cur) go out after the first step. The empty
// asm statements keep the buffer loads where they are written, so at the
// chunk head only the loads that step needs are in flight.
template <bool kUpd, bool kOd, bool kSlow>
__device__ __forceinline__ void sp_chunk(const uint4 (&cur)[kSpChunk], uint4 (&nxt)[kSpChunk],
                                         const uint4*& p, SpIn (&f)[3], double& r,
                                         const int* RPI, int t_base) {
    sp_fetch<kOd>(cur[2], RPI, t_base, f[2]);
    sp_exec<kUpd, kSlow>(cur[0], f[0], r);
    asm volatile("" ::: "memory");
    p += kSpChunk * kW;
#pragma unroll
    for (int k = 0; k < kSpChunk; ++k) nxt[k] = p[k * kW];
    asm volatile("" ::: "memory");
#pragma unroll
    for (int k = 1; k < kSpChunk; ++k) {
        sp_fetch<kOd>(k + 2 < kSpChunk ? cur[k + 2] : nxt[k + 2 - kSpChunk], RPI, t_base,
                      f[(k + 2) % 3]);
        sp_exec<kUpd, kSlow>(cur[k], f[k % 3], r);
    }
    asm volatile("" ::: "memory");
}

// Runs one program on the 64 lanes of this wavefront. Step s reads its
// operands while step s - 2 executes (the scheduler keeps every LDS operand
// kSpLag steps behind its producer, or forwards it), so the LDS latency
// overlaps two steps of work. Op words stream from global memory through two
// register buffers of kSpChunk steps; programs are padded to an even number
// of chunks plus two chunks of idle ops, so the loop needs no bounds checks.
// kSpChunk is a multiple of 3, which keeps the rotation of the three operand
// sets in step with the loop. A chunk's kind (slow or not) is uniform: one
// scalar branch per chunk.
static_assert(kSpChunk % 3 == 0, "operand rotation");
template <bool kUpd, bool kOd>
__device__ __forceinline__ void sp_run(const uint4* __restrict__ prog, int nchunks,
                                       const int* RPI, int t_base, int lane) {
    uint4 qa[kSpChunk], qb[kSpChunk];
    SpIn f[3];
    double r = 0.0;
    const uint4* p = prog + lane;
#pragma unroll
    for (int k = 0; k < kSpChunk; ++k) qa[k] = p[k * kW];
    sp_fetch<kOd>(qa[0], RPI, t_base, f[0]);
    sp_fetch<kOd>(qa[1], RPI, t_base, f[1]);
    for (int ch = 0; ch < nchunks; ch += 2) {
        if (__builtin_amdgcn_readfirstlane(qa[0].w) & kSpSlowChunk)
            sp_chunk<kUpd, kOd, true>(qa, qb, p, f, r, RPI, t_base);
        else
            sp_chunk<kUpd, kOd, false>(qa, qb, p, f, r, RPI, t_base);
        if (__builtin_amdgcn_readfirstlane(qb[0].w) & kSpSlowChunk)
            sp_chunk<kUpd, kOd, true>(qb, qa, p, f, r, RPI, t_base);
        else
            sp_chunk<kUpd, kOd, false>(qb, qa, p, f, r, RPI, t_base);
    }
}

// ---- gathers from the device slot values --------------------------------------
__device__ __forceinline__ double sv_at(const double* sv, int i) { return i < 0 ? 1.0 : sv[i]; }

// Ordered sum 0.0 + v0 + v1 + v2 of a short gather record (count <= 3);
// records of long sums contribute nothing here.
__device__ __forceinline__ double rec_sum(const int4 r, const double* sv) {
    const int c = r.x <= 3 ? r.x : 0;
    const double v0 = c > 0 ? sv_at(sv, r.y) : 0.0;
    const double v1 = c > 1 ? sv_at(sv, r.z) : 0.0;
    const double v2 = c > 2 ? sv_at(sv, r.w) : 0.0;
    double acc = 0.0;
    if (c > 0) acc += v0;
    if (c > 1) acc += v1;
    if (c > 2) acc += v2;
    return acc;
}

constexpr int kStage = 4 * kW;  // doubles of LDS staging for long sums

// Ordered sum 0.0 + sv[idx[0]] + ... + sv[idx[cnt-1]] of a long list: all
// lanes load a chunk into LDS, lane 0 adds it in order. Valid on lane 0.
__device__ double coop_sum(const int32_t* idx, int cnt, const double* sv, double* S, int lane) {
    double acc = 0.0;
    for (int b = 0; b < cnt; b += kStage) {
        const int m = cnt - b < kStage ? cnt - b : kStage;
        int id[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int q = j * kW + lane;
            id[j] = q < m ? idx[b + q] : 0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int q = j * kW + lane;
            if (q < m) S[q] = sv_at(sv, id[j]);
        }
        wsync();
        if (lane == 0) {  // reads batched 16 at a time; the adds stay in order
            int q = 0;
            for (; q + 16 <= m; q += 16) {
                double t[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) t[j] = S[q + j];
#pragma unroll
                for (int j = 0; j < 16; ++j) acc += t[j];
            }
            for (; q < m; ++q) acc += S[q];
        }
        wsync();
    }
    return acc;
}

// Index of key in the ascending a[lo, hi), or -1.
__device__ __forceinline__ int find_sorted(const int32_t* a, int lo, int hi, int key) {
    const int end = hi;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < end && a[lo] == key) ? lo : -1;
}

// Block-wide exchange for kWv wavefronts per instance (one: the wavefront
// helpers above).
template <int kWv>
__device__ __forceinline__ void bsync() {
    if (kWv == 1) wsync();
    else __syncthreads();
}
template <int kWv>
__device__ __forceinline__ void gbsync() {
    if (kWv == 1) gsync();
    else __syncthreads();
}

// kWv wavefronts per instance: the data-parallel phases (evaluation, gathers,
// convergence test, vector updates) use all of them; the factor, the solve and
// the ordered norms run on wavefront 0, whose step programs are latency-bound.
template <int kWv>
__global__ __launch_bounds__(64 * kWv) void k_batch_dcop(const BPat* __restrict__ pats,
                                                   const BInst* __restrict__ insts,
                                                   const int32_t* __restrict__ order, int ninst,
                                                   int* __restrict__ counter,
                                                   const int32_t* __restrict__ ip,
                                                   const SpOpWord* __restrict__ prog,
                                                   const rs_device* __restrict__ devs,
                                                   double* __restrict__ scr,
                                                   int32_t* __restrict__ iscr,
                                                   BResult* __restrict__ res, BOpts o, int nmax,
                                                   int lumax,
                                                   unsigned long long* __restrict__ phase_cyc) {
    extern __shared__ double lds[];
    double* VL = lds;           // L/U/diag values when lumax > 0 (step-program slots)
    double* W = VL + lumax;     // factor work vector; also the solve vector
    double* X = W + nmax;
    double* XN = X + nmax;      // rhs, then x_new
    double* S = XN + nmax;      // staging of long ordered sums, kStage per wavefront
    double* SB = S + kStage * kWv;  // broadcast slots: 4 doubles, then 8 ints
    int* SBi = reinterpret_cast<int*>(SB + 4);
    // (kWv > 1) squares of dx and x_new for the ordered norms
    double* Q = SB + 8;
    int* RPI = reinterpret_cast<int*>(Q + (kWv > 1 ? 2 * nmax : 0));
    int* RP = RPI + nmax;
    constexpr int kT = kW * kWv;
    const int tid = threadIdx.x, lane = tid & (kW - 1), wave = tid / kW;
    double* Sw = S + wave * kStage;
    // optional per-phase cycle accounting (measurement only)
    unsigned long long pc[kPhases] = {};
    unsigned long long t_last = 0;
#define PHASE(k)                                                  \
    if (phase_cyc) {                                              \
        const unsigned long long t_now = __builtin_amdgcn_s_memtime(); \
        pc[k] += t_now - t_last;                                  \
        t_last = t_now;                                           \
    }

    for (;;) {
        int qi;
        if (kWv == 1) {
            int qv = 0;
            if (lane == 0) qv = atomicAdd(counter, 1);
            qi = __shfl(qv, 0);
        } else {
            if (tid == 0) SBi[0] = atomicAdd(counter, 1);
            __syncthreads();
            qi = SBi[0];
            __syncthreads();
        }
        if (qi >= ninst) {
            if (phase_cyc && tid == 0)
                for (int k = 0; k < kPhases; ++k) atomicAdd(phase_cyc + k, pc[k]);
            break;
        }
        const int ii = order[qi];
        const BInst I = insts[ii];
        const BPat P = pats[I.pat];
        const int n = P.n, nc = P.node_count;
        const rs_device* dev = devs + I.o_dev;
        double* sv = scr + I.o_sv;
        double* lx = lumax ? VL : scr + I.o_lx;
        double* ux = lumax ? VL + P.nl : scr + I.o_ux;
        double* ud = lumax ? VL + P.nl + P.nu : scr + I.o_ud;
        const int32_t* ax_src = ip + P.o_ax_src;
        const int32_t* rhs_src = ip + P.o_rhs_src;
        const int32_t* sc_ptr = ip + P.o_sc_ptr;
        const int32_t* sc_pre = ip + P.o_sc_pre;
        const int32_t* lcp = ip + P.o_lcp;
        const int32_t* lri = ip + P.o_lri;
        const int32_t* ucp = ip + P.o_ucp;
        const int32_t* uri = ip + P.o_uri;
        const int32_t* colflag = ip + P.o_colflag;
        const int32_t* up_ptr = ip + P.o_up_ptr;
        const int32_t* up_j = ip + P.o_up_j;
        const int32_t* up_row = ip + P.o_up_row;
        const int32_t* up_l = ip + P.o_up_l;
        const int32_t* pin = ip + P.o_pin;
        const int32_t* pout = ip + P.o_pout;
        const int32_t* lrp = ip + P.o_lrp;
        const int32_t* lrc = ip + P.o_lrc;
        const int32_t* lrv = ip + P.o_lrv;
        const int32_t* urp = ip + P.o_urp;
        const int32_t* urc = ip + P.o_urc;
        const int32_t* urv = ip + P.o_urv;
        const int32_t* blk_ptr = ip + P.o_blk_ptr;
        const int32_t* lsl_blk = ip + P.o_lsl_blk;
        const int32_t* lsl_ptr = ip + P.o_lsl_ptr;
        const int32_t* lsl_rows = ip + P.o_lsl_rows;
        const int32_t* usl_blk = ip + P.o_usl_blk;
        const int32_t* usl_ptr = ip + P.o_usl_ptr;
        const int32_t* usl_rows = ip + P.o_usl_rows;
        const int32_t* odp = ip + P.o_odp;
        const int32_t* odax = ip + P.o_odax;
        const int32_t* odsrc = ip + P.o_odsrc;
        const int32_t* has_od = ip + P.o_blk_has_od;
        const int32_t* sc_col = ip + P.o_sc_col;
        const int4* screc = reinterpret_cast<const int4*>(ip + P.o_screc);
        const int4* rhsrec = reinterpret_cast<const int4*>(ip + P.o_rhsrec);
        const int32_t* lrun = ip + P.o_lrun;
        const int32_t* lrow = ip + P.o_lrow;
        double* rsum = scr + I.o_ax;  // run sums for the column-by-column factor
        int32_t* ss = iscr + I.o_ss;
        const bool sp = lumax > 0 && P.sp_ok;  // step programs (values in LDS)
        const int nsc = sc_ptr[n];

        for (int i = tid; i < n; i += kT) {
            W[i] = 0.0;
            X[i] = 0.0;
        }
        int xl = nc;  // the Newton x has node_count entries until convergence
        bool factored = false;
        BResult fin{0, 0, 1, 0, 0.0};
        bsync<kWv>();
        if (phase_cyc) t_last = __builtin_amdgcn_s_memtime();

        const double gstart = fmax(o.gmin * 1e3, 1e-6);
        for (int gi = 0; gi <= o.gmin_steps && !fin.converged; ++gi) {
            const double gmin = o.gmin_steps == 0
                                    ? o.gmin
                                    : gstart + (o.gmin - gstart) * ((double)gi / (double)o.gmin_steps);
            for (int si = 0; si <= o.source_steps && !fin.converged; ++si) {
                const double scale = o.source_steps == 0 ? 1.0 : (double)si / (double)o.source_steps;
                BResult r{0, 0, 1, fin.full_factors, 0.0};
                double damping = o.damping;
                double prev = 1.7976931348623157e308;
                for (int it = 0; it < o.max_iters; ++it) {
                    r.iterations = it + 1;
                    PHASE(7)
                    // ---- stamp: evaluate, gather ax and the RHS contributions ----
                    StampArgs a{0.0, 0.0, gmin, scale, MODE_DC, 0};
                    {  // device records loaded four at a time: one memory round trip per four
                        int d = tid;
                        for (; d + 3 * kT < I.ndev; d += 4 * kT) {
                            rs_device v[4];
#pragma unroll
                            for (int j = 0; j < 4; ++j) v[j] = dev[d + j * kT];
#pragma unroll
                            for (int j = 0; j < 4; ++j)
                                eval_device(v[j], d + j * kT, X, xl, a, nullptr, nullptr, nullptr,
                                            nullptr, sv + (size_t)(d + j * kT) * kSlotsPerDevice);
                        }
                        for (; d < I.ndev; d += kT) {
                            const rs_device v = dev[d];
                            eval_device(v, d, X, xl, a, nullptr, nullptr, nullptr, nullptr,
                                        sv + (size_t)d * kSlotsPerDevice);
                        }
                    }
                    gbsync<kWv>();
                    PHASE(0)
                    if (!factored)
                        for (int i = tid; i < n; i += kT) {
                            RPI[i] = i;
                            RP[i] = i;
                        }
                    // scatter-run sums go into the L/U slots for a step-program
                    // refactor, else into rsum for the column-by-column factor
                    const bool spr = sp && factored;
                    if (spr) {
                        for (int i = tid; i < P.nl + P.nu; i += kT) VL[i] = 0.0;
                        bsync<kWv>();
                    }
                    // RHS rows and scatter runs from their short records ...
                    for (int base = 0; base < n; base += 4 * kT) {
                        int4 rr[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int i = base + j * kT + tid;
                            rr[j] = i < n ? rhsrec[i] : make_int4(0, 0, 0, 0);
                        }
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int i = base + j * kT + tid;
                            const double v = rec_sum(rr[j], sv);
                            if (i < n && rr[j].x <= 3) XN[i] = v;
                        }
                    }
                    PHASE(1)
                    for (int base = 0; base < nsc; base += 4 * kT) {
                        int4 rr[4];
                        int tt[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int q = base + j * kT + tid;
                            rr[j] = q < nsc ? screc[q] : make_int4(0, 0, 0, 0);
                            tt[j] = (spr && q < nsc) ? ss[q] : -1;
                        }
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const int q = base + j * kT + tid;
                            const double v = rec_sum(rr[j], sv);
                            if (q < nsc && rr[j].x <= 3) {
                                if (!spr) rsum[q] = v;
                                else if (tt[j] >= 0) VL[tt[j]] = v;
                            }
                        }
                    }
                    // ... and the long ordered sums of both, one wavefront each
                    for (int j = wave; j < P.n_lrow + P.n_lrun; j += kWv) {
                        if (j < P.n_lrow) {
                            const int i = lrow[j];
                            const int4 rr = rhsrec[i];
                            const double v = coop_sum(rhs_src + rr.y, rr.x, sv, Sw, lane);
                            if (lane == 0) XN[i] = v;
                        } else {
                            const int q = lrun[j - P.n_lrow];
                            const int4 rr = screc[q];
                            const double v = coop_sum(ax_src + rr.y, rr.x, sv, Sw, lane);
                            if (lane == 0) {
                                if (!spr) rsum[q] = v;
                                else if (ss[q] >= 0) VL[ss[q]] = v;
                            }
                        }
                    }
                    if (spr) bsync<kWv>();
                    else gbsync<kWv>();
                    PHASE(2)
                    // ---- factor (first: with pivot search; later: reuse) ----
                    if (spr) {  // refactor as a step program
                        if (wave == 0)
                            sp_run<true, false>(reinterpret_cast<const uint4*>(prog + P.o_spf),
                                                P.spf_chunks, RPI, P.t_base, lane);
                        bsync<kWv>();
                        PHASE(4)
                    } else if (wave == 0)
                    for (int g = 0; g < n; ++g) {
                        for (int s = sc_ptr[g] + lane; s < sc_ptr[g + 1]; s += kW)
                            W[RPI[sc_pre[s]]] = rsum[s];
                        wsync();
                        const int p1 = up_ptr[g + 1];
                        for (int base = up_ptr[g]; base < p1; base += kW) {
                            const int idx = base + lane;
                            const bool act = idx < p1;
                            const int jj = act ? up_j[idx] : INT_MAX;
                            const int rr = act ? up_row[idx] : 0;
                            const double lv = act ? lx[up_l[idx]] : 0.0;
                            const int cend = (p1 - base) < kW ? (p1 - base) : kW;
                            int cur = 0;
                            while (cur < cend) {  // groups of equal j, ascending
                                const int jcur = __shfl(jj, cur);
                                const double u = W[jcur];
                                const bool mine = act && jj == jcur;
                                const unsigned long long m = __ballot(mine);
                                if (u != 0.0 && mine) W[rr] = W[rr] - lv * u;
                                wsync();
                                cur += __popcll(m);
                            }
                        }
                        const bool single = (colflag[g] & 1) != 0;
                        const int ls = lcp[g], le = lcp[g + 1];
                        const int us = ucp[g], ue = ucp[g + 1] - 1;
                        if (!factored && !single && le > ls) {
                            const double dabs = fabs(W[g]);
                            double bv = -1.0;
                            int bl = INT_MAX;
                            for (int l = ls + lane; l < le; l += kW) {
                                const double v = fabs(W[lri[l]]);
                                if (v > dabs && v > bv) {
                                    bv = v;
                                    bl = l;
                                }
                            }
                            for (int off = kW / 2; off > 0; off >>= 1) {
                                const double ov = __shfl_xor(bv, off);
                                const int ol = __shfl_xor(bl, off);
                                if (ov > bv || (ov == bv && ol < bl)) {
                                    bv = ov;
                                    bl = ol;
                                }
                            }
                            if (bl != INT_MAX && dabs < o.pivot_tol * bv && lane == 0) {
                                const int mrow = lri[bl];
                                const double t = W[g];
                                W[g] = W[mrow];
                                W[mrow] = t;
                                const int okk = RP[g], om = RP[mrow];
                                RP[g] = om;
                                RP[mrow] = okk;
                                RPI[okk] = mrow;
                                RPI[om] = g;
                            }
                            wsync();
                        }
                        for (int u = us + lane; u < ue; u += kW) ux[u] = W[uri[u]];
                        double diag = W[g];
                        if (single) {
                            if (fabs(diag) < 1e-300) diag = 1e-14;
                        } else if (fabs(diag) < 1e-14) {
                            diag = diag >= 0.0 ? 1e-14 : -1e-14;
                        }
                        if (lane == 0) {
                            ux[ue] = diag;
                            ud[g] = diag;
                        }
                        for (int l = ls + lane; l < le; l += kW) lx[l] = W[lri[l]] / diag;
                        wsync();
                        for (int s = sc_ptr[g] + lane; s < sc_ptr[g + 1]; s += kW) W[RPI[sc_pre[s]]] = 0.0;
                        for (int l = ls + lane; l < le; l += kW) W[lri[l]] = 0.0;
                        for (int u = us + lane; u < ue; u += kW) W[uri[u]] = 0.0;
                        if (lane == 0) W[g] = 0.0;
                        if (lumax) wsync();
                        else gsync();  // L/U values of this column are read by later columns
                    }
                    if (!spr) gbsync<kWv>();  // the factor of wavefront 0 for all
                    if (!factored) {
                        PHASE(3)
                        r.full_factors++;
                        if (sp) {  // scatter slots under the pivots just found
                            for (int s = tid; s < nsc; s += kT) {
                                const int g = sc_col[s], pos = RPI[sc_pre[s]];
                                const int ue = ucp[g + 1] - 1;
                                int slot = -1;
                                if (pos == g) {
                                    slot = P.nl + ue;
                                } else if (pos < g) {
                                    const int q = find_sorted(uri, ucp[g], ue, pos);
                                    slot = q < 0 ? -1 : P.nl + q;
                                } else {
                                    slot = find_sorted(lri, lcp[g], lcp[g + 1], pos);
                                }
                                ss[s] = slot;
                            }
                            gbsync<kWv>();
                            PHASE(5)
                        }
                    }
                    factored = true;
                    if (sp) {  // ---- solve as a step program ----
                        for (int e = tid; e < P.nslots - P.od_base; e += kT)
                            VL[P.od_base + e] = sv_at(sv, ax_src[odax[e]]);
                        for (int i = tid; i < n; i += kT) VL[P.t_base + RPI[pin[i]]] = XN[i];
                        bsync<kWv>();
                        if (wave == 0) {
                            const uint4* sprog = reinterpret_cast<const uint4*>(prog + P.o_sps);
                            if (P.nslots > P.od_base)
                                sp_run<false, true>(sprog, P.sps_chunks, RPI, P.t_base, lane);
                            else
                                sp_run<false, false>(sprog, P.sps_chunks, RPI, P.t_base, lane);
                        }
                        bsync<kWv>();
                        for (int i = tid; i < n; i += kT) XN[i] = VL[P.t_base + pout[i]];
                        bsync<kWv>();
                    } else {
                    if (wave == 0) {
                    // ---- solve, in W (all zero now; re-zeroed afterwards) ----
                    double* T = W;
                    for (int i = lane; i < n; i += kW) T[RPI[pin[i]]] = XN[i];
                    wsync();
                    for (int b = P.nblk - 1; b >= 0; --b) {
                        if (has_od[b]) {
                            for (int p = blk_ptr[b] + lane; p < blk_ptr[b + 1]; p += kW) {
                                const int e0 = odp[p], e1 = odp[p + 1];
                                if (e0 == e1) continue;
                                const int pos = RPI[p];
                                double t = T[pos];
                                for (int e = e0; e < e1; ++e) t = t - sv_at(sv, ax_src[odax[e]]) * T[odsrc[e]];
                                T[pos] = t;
                            }
                            wsync();
                        }
                        for (int lev = lsl_blk[b]; lev < lsl_blk[b + 1]; ++lev) {
                            for (int q = lsl_ptr[lev] + lane; q < lsl_ptr[lev + 1]; q += kW) {
                                const int rr = lsl_rows[q];
                                const int e0 = lrp[rr], e1 = lrp[rr + 1];
                                if (e0 == e1) continue;
                                double t = T[rr];
                                for (int e = e0; e < e1; ++e) t = t - lx[lrv[e]] * T[lrc[e]];
                                T[rr] = t;
                            }
                            wsync();
                        }
                        for (int lev = usl_blk[b]; lev < usl_blk[b + 1]; ++lev) {
                            for (int q = usl_ptr[lev] + lane; q < usl_ptr[lev + 1]; q += kW) {
                                const int rr = usl_rows[q];
                                double t = T[rr];
                                for (int e = urp[rr]; e < urp[rr + 1]; ++e) t = t - ux[urv[e]] * T[urc[e]];
                                T[rr] = t / ud[rr];
                            }
                            wsync();
                        }
                    }
                    for (int i = lane; i < n; i += kW) XN[i] = T[pout[i]];
                    wsync();
                    for (int i = lane; i < n; i += kW) W[i] = 0.0;
                    }
                    bsync<kWv>();
                    }
                    PHASE(6)
                    // ---- convergence, damping, update (newton.rs:93-107) ----
                    const int nd = n < xl ? n : xl;
                    int ok = 1;
                    for (int i = tid; i < nd; i += kT) {
                        const double dxi = XN[i] - X[i];
                        if (!(fabs(dxi) <= o.abs_tol + o.rel_tol * fabs(XN[i]))) ok = 0;
                    }
                    bool conv = __ballot(ok == 0) == 0ull;
                    if (kWv > 1) {
                        if (lane == 0) SBi[1 + wave] = conv ? 1 : 0;
                        __syncthreads();
                        for (int w = 0; w < kWv; ++w) conv = conv && SBi[1 + w] != 0;
                    }
                    // |dx| steers the damping of the next iteration; |x_new| is only
                    // reported, so it is summed on the iteration that ends the loop
                    const bool need_d = !conv, need_f = conv || it + 1 >= o.max_iters;
                    double dxn_l = 0.0, fn_l = 0.0;
                    if (kWv > 1) {
                        // the squares in parallel; the sums stay in order on one lane
                        for (int i = tid; i < n; i += kT) {
                            const double dxi = XN[i] - X[i];
                            Q[i] = i < nd ? dxi * dxi : 0.0;
                            Q[nmax + i] = XN[i] * XN[i];
                        }
                        __syncthreads();
                        if (tid == 0) {
                            if (need_d) {
                                double sd = 0.0;
                                int i = 0;
                                for (; i + 16 <= nd; i += 16) {
                                    double qd[16];
#pragma unroll
                                    for (int j = 0; j < 16; ++j) qd[j] = Q[i + j];
#pragma unroll
                                    for (int j = 0; j < 16; ++j) sd += qd[j];
                                }
                                for (; i < nd; ++i) sd += Q[i];
                                dxn_l = sqrt(sd);
                            }
                            if (need_f) {
                                double sf = 0.0;
                                int i = 0;
                                for (; i + 16 <= n; i += 16) {
                                    double qf[16];
#pragma unroll
                                    for (int j = 0; j < 16; ++j) qf[j] = Q[nmax + i + j];
#pragma unroll
                                    for (int j = 0; j < 16; ++j) sf += qf[j];
                                }
                                for (; i < n; ++i) sf += Q[nmax + i];
                                fn_l = sqrt(sf);
                            }
                        }
                    } else if (tid == 0) {
                        // the ordered sums (newton.rs:98-101), reads batched 8 at a time
                        if (need_d) {
                            double sd = 0.0;
                            int i = 0;
                            for (; i + 8 <= nd; i += 8) {
                                double xn[8], xo[8];
#pragma unroll
                                for (int j = 0; j < 8; ++j) {
                                    xn[j] = XN[i + j];
                                    xo[j] = X[i + j];
                                }
#pragma unroll
                                for (int j = 0; j < 8; ++j) {
                                    const double dxi = xn[j] - xo[j];
                                    sd += dxi * dxi;
                                }
                            }
                            for (; i < nd; ++i) {
                                const double dxi = XN[i] - X[i];
                                sd += dxi * dxi;
                            }
                            dxn_l = sqrt(sd);
                        }
                        if (need_f) {
                            double sf = 0.0;
                            int i = 0;
                            for (; i + 8 <= n; i += 8) {
                                double xn[8];
#pragma unroll
                                for (int j = 0; j < 8; ++j) xn[j] = XN[i + j];
#pragma unroll
                                for (int j = 0; j < 8; ++j) sf += xn[j] * xn[j];
                            }
                            for (; i < n; ++i) sf += XN[i] * XN[i];
                            fn_l = sqrt(sf);
                        }
                    }
                    double dxn;
                    if (kWv == 1) {
                        dxn = __shfl(dxn_l, 0);
                        const double fn = __shfl(fn_l, 0);
                        if (need_f) r.final_norm = fn;
                    } else {
                        if (tid == 0) {
                            SB[0] = dxn_l;
                            SB[1] = fn_l;
                        }
                        __syncthreads();
                        dxn = SB[0];
                        if (need_f) r.final_norm = SB[1];
                    }
                    if (conv) {
                        for (int i = tid; i < n; i += kT) X[i] = XN[i];
                        xl = n;
                        r.converged = 1;
                        r.reason = 0;
                        bsync<kWv>();
                        break;
                    }
                    if (dxn > prev && damping > o.damping_min) {
                        const double d2 = damping * 0.5;
                        damping = d2 >= o.damping_min ? d2 : o.damping_min;
                    }
                    prev = dxn;
                    for (int i = tid; i < nd; i += kT) X[i] = X[i] + damping * (XN[i] - X[i]);
                    bsync<kWv>();
                }
                fin = r;
            }
        }
        PHASE(7)
        if (tid == 0) res[ii] = fin;
        // the solution, or the last damped iterate when Newton did not
        // converge (its node_count entries; the aux entries stay zero)
        for (int i = tid; i < n; i += kT) scr[I.o_x + i] = X[i];
        if (lumax) {  // leave the factor in global scratch as well
            for (int l = tid; l < P.nl; l += kT) scr[I.o_lx + l] = lx[l];
            for (int u = tid; u < P.nu; u += kT) scr[I.o_ux + u] = ux[u];
            for (int i = tid; i < n; i += kT) scr[I.o_ud + i] = ux[ucp[i + 1] - 1];
        }
        __syncthreads();
    }
}

size_t lds_bytes(int nmax, int lumax, int waves) {
    return (size_t)nmax * ((waves > 1 ? 5 : 3) * sizeof(double) + 2 * sizeof(int)) +
           (size_t)(lumax + kStage * waves + 8) * sizeof(double);
}

// ---- one matrix by its step programs (single-matrix solver) ------------------
// At most 512 threads: the program's two 12-step op-word buffers and operand
// sets then keep their registers (a 1,024-thread bound leaves 128 VGPRs and
// spills inside the chunk loop, whose reload waits for every op-word load in
// flight). Shared memory: the nslots + 1 value slots from address 0, then 64
// doubles per wavefront for the ordered sums of long scatter runs.
constexpr int kSpMaxThreads = 512;
// VL (dynamic LDS from address 0, as sp_run requires) holds the value slots:
// lx | ux | T | off-diagonal ax | dummy. The refactor scatters A into the
// slots (duplicates summed in CSC order from 0.0, fill entries 0), runs the
// refactor program on wavefront 0 and writes the values back in the solver's
// layout; the solve loads them, runs the solve program and writes x.
__global__ __launch_bounds__(kSpMaxThreads) void k_sp_refactor(const SpSolverDev* __restrict__ sdp,
                                                     const PlanDev* __restrict__ pdp,
                                                     const int32_t* __restrict__ ip,
                                                     const int32_t* __restrict__ sip,
                                                     const SpOpWord* __restrict__ words,
                                                     const double* __restrict__ ax,
                                                     double* __restrict__ vals,
                                                     const int32_t* __restrict__ skip) {
    extern __shared__ double VL[];
    if (skip && *skip) return;
    const SpSolverDev sd = *sdp;
    const PlanDev d = *pdp;
    const int tid = threadIdx.x, bs = blockDim.x, lane = tid & (kW - 1), wave = tid / kW;
    for (int a = tid; a <= sd.nslots; a += bs) VL[a] = 0.0;
    __syncthreads();
    const int32_t* sc_src = ip + d.o_sc_src;
    const int32_t* sc_cnt = ip + d.o_sc_cnt;
    const int32_t* ss = sip + sd.o_ss;
    for (int s = tid; s < sd.nsc; s += bs) {
        const int slot = ss[s];
        const int cnt = sc_cnt[s];
        if (slot < 0 || cnt > kSpShortRun) continue;
        double v = 0.0;
        const int src = sc_src[s];
        for (int q = 0; q < cnt; ++q) v += ax[src + q];
        VL[slot] = v;
    }
    {  // long runs: the wavefront loads 64 terms at a time, lane 0 adds them in order
        double* S = VL + sd.nslots + 1 + wave * kW;
        const int32_t* lrun = sip + sd.o_lrun;
        for (int j = wave; j < sd.n_lrun; j += bs / kW) {
            const int s = lrun[j];
            const int src = sc_src[s], cnt = sc_cnt[s];
            double v = 0.0;
            for (int b = 0; b < cnt; b += kW) {
                const int m = cnt - b < kW ? cnt - b : kW;
                if (lane < m) S[lane] = ax[src + b + lane];
                wsync();
                if (lane == 0)
                    for (int q = 0; q < m; ++q) v += S[q];
                wsync();
            }
            if (lane == 0) VL[ss[s]] = v;
        }
    }
    __syncthreads();
    if (wave == 0)
        sp_run<true, false>(reinterpret_cast<const uint4*>(words + sd.o_fac), sd.fac_chunks,
                            sip + sd.o_rpi, sd.t_base, lane);
    __syncthreads();
    double* lx = vals + d.o_lx;
    double* ux = vals + d.o_ux;
    double* ud = vals + d.o_ud;
    const int32_t* ucp = ip + d.o_ucp;
    for (int a = tid; a < sd.nl; a += bs) lx[a] = VL[a];
    for (int a = tid; a < sd.nu; a += bs) ux[a] = VL[sd.nl + a];
    for (int g = tid; g < sd.n; g += bs) ud[g] = VL[sd.nl + ucp[g + 1] - 1];
    if (d.flags & PLAN_CSR) {
        const int32_t* lrv = ip + d.o_lrv;
        const int32_t* urv = ip + d.o_urv;
        double* lxr = vals + d.o_lxr;
        double* uxr = vals + d.o_uxr;
        for (int e = tid; e < d.nl; e += bs) lxr[e] = VL[lrv[e]];
        for (int e = tid; e < d.nu - d.n; e += bs) uxr[e] = VL[sd.nl + urv[e]];
    }
}

__global__ __launch_bounds__(kSpMaxThreads) void k_sp_solve(const SpSolverDev* __restrict__ sdp,
                                                  const PlanDev* __restrict__ pdp,
                                                  const int32_t* __restrict__ ip,
                                                  const int32_t* __restrict__ sip,
                                                  const SpOpWord* __restrict__ words,
                                                  const double* __restrict__ ax,
                                                  const double* __restrict__ vals,
                                                  const double* __restrict__ rhs,
                                                  double* __restrict__ x,
                                                  const int32_t* __restrict__ skip) {
    extern __shared__ double VL[];
    if (skip && *skip) return;
    const SpSolverDev sd = *sdp;
    const PlanDev d = *pdp;
    const int tid = threadIdx.x, bs = blockDim.x, lane = tid & (kW - 1), wave = tid / kW;
    const double* lx = vals + d.o_lx;
    const double* ux = vals + d.o_ux;
    const int32_t* odax = ip + d.o_odax;
    const int32_t* perm_in = ip + d.o_perm_in;
    const int32_t* perm_out = ip + d.o_perm_out;
    for (int a = tid; a < sd.nl; a += bs) VL[a] = lx[a];
    for (int a = tid; a < sd.nu; a += bs) VL[sd.nl + a] = ux[a];
    for (int e = tid; e < sd.nod; e += bs) VL[sd.od_base + e] = ax[odax[e]];
    for (int i = tid; i < sd.n; i += bs) VL[sd.t_base + perm_in[i]] = rhs[i];
    if (tid == 0) VL[sd.nslots] = 0.0;
    __syncthreads();
    if (wave == 0) {
        const uint4* prog = reinterpret_cast<const uint4*>(words + sd.o_sol);
        if (sd.nod > 0)
            sp_run<false, true>(prog, sd.sol_chunks, sip + sd.o_rpi, sd.t_base, lane);
        else
            sp_run<false, false>(prog, sd.sol_chunks, sip + sd.o_rpi, sd.t_base, lane);
    }
    __syncthreads();
    for (int i = tid; i < sd.n; i += bs) x[i] = VL[sd.t_base + perm_out[i]];
}

}  // namespace

namespace rsgpu {

hipError_t launch_batch_dcop(const BPat* pats, const BInst* insts, const int32_t* order, int ninst,
                             int* counter, const int32_t* ip, const SpOpWord* prog,
                             const rs_device* devs, double* scr, int32_t* iscr, BResult* res,
                             const BOpts& o, int nmax, int lumax, int grid, int waves, hipStream_t s,
                             unsigned long long* phase_cyc) {
    if (waves == 4)
        hipLaunchKernelGGL(k_batch_dcop<4>, dim3(grid), dim3(4 * kW), lds_bytes(nmax, lumax, 4), s,
                           pats, insts, order, ninst, counter, ip, prog, devs, scr, iscr, res, o,
                           nmax, lumax, phase_cyc);
    else
        hipLaunchKernelGGL(k_batch_dcop<1>, dim3(grid), dim3(kW), lds_bytes(nmax, lumax, 1), s,
                           pats, insts, order, ninst, counter, ip, prog, devs, scr, iscr, res, o,
                           nmax, lumax, phase_cyc);
    return hipGetLastError();
}

size_t batch_lds_bytes(int nmax, int lumax, int waves) { return lds_bytes(nmax, lumax, waves); }

hipError_t launch_sp_refactor(const SpSolverDev* sd, const PlanDev* pd, const int32_t* ip,
                              const int32_t* sip, const SpOpWord* words, const double* ax,
                              double* vals, int nslots, int threads, hipStream_t s,
                              const int32_t* skip) {
    hipLaunchKernelGGL(k_sp_refactor, dim3(1), dim3(threads),
                       (size_t)(nslots + 1 + threads) * sizeof(double), s, sd, pd, ip, sip, words, ax,
                       vals, skip);
    return hipGetLastError();
}

hipError_t launch_sp_solve(const SpSolverDev* sd, const PlanDev* pd, const int32_t* ip,
                           const int32_t* sip, const SpOpWord* words, const double* ax,
                           const double* vals, const double* rhs, double* x, int nslots, int threads,
                           hipStream_t s, const int32_t* skip) {
    hipLaunchKernelGGL(k_sp_solve, dim3(1), dim3(threads), (size_t)(nslots + 1) * sizeof(double), s,
                       sd, pd, ip, sip, words, ax, vals, rhs, x, skip);
    return hipGetLastError();
}

// Static LDS of the single-matrix step-program kernels (must be 0, as for
// k_batch_dcop).
size_t sp_solver_static_lds() {
    size_t mx = 0;
    for (const void* k : {reinterpret_cast<const void*>(k_sp_refactor),
                          reinterpret_cast<const void*>(k_sp_solve)}) {
        hipFuncAttributes a{};
        if (hipFuncGetAttributes(&a, k) != hipSuccess) return (size_t)-1;
        mx = a.sharedSizeBytes > mx ? a.sharedSizeBytes : mx;
    }
    return mx;
}

// Static LDS of k_batch_dcop: step programs address their slots by absolute
// LDS address, which is only right when this is 0.
size_t batch_static_lds() {
    size_t mx = 0;
    for (const void* k : {reinterpret_cast<const void*>(k_batch_dcop<1>),
                          reinterpret_cast<const void*>(k_batch_dcop<4>)}) {
        hipFuncAttributes a{};
        if (hipFuncGetAttributes(&a, k) != hipSuccess) return (size_t)-1;
        mx = a.sharedSizeBytes > mx ? a.sharedSizeBytes : mx;
    }
    return mx;
}

int batch_blocks_per_cu(int nmax, int lumax, int waves) {
    int nb = 0;
    hipError_t e = waves == 4
                       ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch_dcop<4>, 4 * kW,
                                                                      lds_bytes(nmax, lumax, 4))
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch_dcop<1>, kW,
                                                                      lds_bytes(nmax, lumax, 1));
    if (e != hipSuccess || nb < 1) nb = 1;
    return nb;
}

}  // namespace rsgpu
