// kernels.hip - CDNA4 (gfx950) kernels for the static-pattern sparse LU.
//
// k_refactor: numeric refactorization on a fixed pattern and pivot sequence
//   (native_klu.rs:1261-1373 with reused pivots; sparse_lu.rs:451-534 is the
//   same loop with identity pivots). One workgroup per matrix; columns run in
//   dependency levels, one wavefront per column. Each wavefront owns a dense
//   work vector (LDS when it fits). Per column the U entries are applied in
//   ascending row order, each lane owning distinct rows of L(:,j), so every
//   work[row] sees the reference's updates in the reference's order: results
//   are bit-identical to the sequential loop (no FMA contraction).
// k_refactor_rl: the same arithmetic scheduled right-looking (rlplan.cpp):
//   pivots of one phase divide their L columns, then every entry updated in
//   that phase applies its updates in the reference's order. The critical
//   path is the number of phases instead of the sum of the column chains;
//   the values live in LDS when they fit.
// k_solve: P, L (unit), U, P^T sweeps of native_klu.rs:1442-1488 and the BTF
//   block back-substitution of :1491-1551, row-oriented so each unknown is a
//   sequential dot product in the reference's update order, rows of one
//   dependency level in parallel. Long rows take a wavefront: the products
//   are formed in parallel and one lane subtracts them in order.
#include <hip/hip_runtime.h>

#include "plan.hpp"

using rsgpu::PlanDev;

namespace {

constexpr int kWave = 64;
constexpr int kSglBatch = 8;  // one-update tasks in flight per thread

__device__ __forceinline__ void wave_sync() {
    // Orders this wave's LDS/global accesses before the following ones
    // (lanes exchange data through memory between steps).
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Row-ordered copies of L and U (without diagonal) for the solve, which then
// reads each row's values contiguously instead of through lrv/urv.
__device__ __forceinline__ void csr_copies(const PlanDev& d, const int32_t* __restrict__ ip,
                                           double* __restrict__ vpool, const double* lx,
                                           const double* ux) {
    const int32_t* lrv = ip + d.o_lrv;
    const int32_t* urv = ip + d.o_urv;
    double* lxr = vpool + d.o_lxr;
    double* uxr = vpool + d.o_uxr;
    const int nlr = d.nl, nur = d.nu - d.n;
    for (int e = threadIdx.x; e < nlr; e += blockDim.x) lxr[e] = lx[lrv[e]];
    for (int e = threadIdx.x; e < nur; e += blockDim.x) uxr[e] = ux[urv[e]];
}

// t - P[0] - P[1] - ... - P[c-1], subtracted in this order; the reads of P
// (LDS) are issued sixteen at a time, so only the subtractions are serial.
__device__ __forceinline__ double ordered_sub(double t, const double* P, int c) {
    int i = 0;
    for (; i + 16 <= c; i += 16) {
        double p[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) p[j] = P[i + j];
#pragma unroll
        for (int j = 0; j < 16; ++j) t = t - p[j];
    }
    if (i + 8 <= c) {
        double p[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) p[j] = P[i + j];
#pragma unroll
        for (int j = 0; j < 8; ++j) t = t - p[j];
        i += 8;
    }
    if (i + 4 <= c) {
        double p[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) p[j] = P[i + j];
#pragma unroll
        for (int j = 0; j < 4; ++j) t = t - p[j];
        i += 4;
    }
    if (i + 2 <= c) {
        const double p0 = P[i], p1 = P[i + 1];
        t = t - p0;
        t = t - p1;
        i += 2;
    }
    if (i < c) t = t - P[i];
    return t;
}

__device__ __forceinline__ double clamp_pivot(double diag, bool singleton) {
    if (singleton) return fabs(diag) < 1e-300 ? 1e-14 : diag;
    if (fabs(diag) < 1e-14) return diag >= 0.0 ? 1e-14 : -1e-14;
    return diag;
}

template <bool kLds>
__global__ __launch_bounds__(1024) void k_refactor(const PlanDev* __restrict__ descs,
                                                  const int32_t* __restrict__ ip,
                                                  const double* __restrict__ axpool,
                                                  double* __restrict__ vpool,
                                                  double* __restrict__ wpool,
                                                  const int32_t* __restrict__ skip) {
    extern __shared__ double smem[];
    if (skip && *skip) return;
    const PlanDev d = descs[blockIdx.x];
    const int n = d.n;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    double* work = kLds ? smem + (size_t)wid * n : wpool + d.o_work + (size_t)wid * n;
    if (kLds) {
        for (int i = threadIdx.x; i < nw * n; i += blockDim.x) smem[i] = 0.0;
        __syncthreads();
    }
    const int32_t* lcp = ip + d.o_lcp;
    const int32_t* lri = ip + d.o_lri;
    const int32_t* ucp = ip + d.o_ucp;
    const int32_t* uri = ip + d.o_uri;
    const int32_t* colflag = ip + d.o_colflag;
    const int32_t* flev_ptr = ip + d.o_flev_ptr;
    const int32_t* flev_cols = ip + d.o_flev_cols;
    const int32_t* sc_ptr = ip + d.o_sc_ptr;
    const int32_t* sc_src = ip + d.o_sc_src;
    const int32_t* sc_cnt = ip + d.o_sc_cnt;
    const int32_t* sc_dst = ip + d.o_sc_dst;
    const double* ax = axpool + d.o_ax;
    double* lx = vpool + d.o_lx;
    double* ux = vpool + d.o_ux;
    double* ud = vpool + d.o_ud;

    for (int lev = 0; lev < d.nflev; ++lev) {
        const int c0 = flev_ptr[lev], c1 = flev_ptr[lev + 1];
        for (int c = c0 + wid; c < c1; c += nw) {
            const int g = flev_cols[c];
            // scatter A(:,g): duplicates summed in CSC order starting from 0.0
            for (int s = sc_ptr[g] + lane; s < sc_ptr[g + 1]; s += kWave) {
                double v = 0.0;
                const int src = sc_src[s], cnt = sc_cnt[s];
                for (int q = 0; q < cnt; ++q) v += ax[src + q];
                work[sc_dst[s]] = v;
            }
            wave_sync();
            const int us = ucp[g], ue = ucp[g + 1] - 1;
            for (int u = us; u < ue; ++u) {
                const int j = uri[u];
                const double ujk = work[j];
                if (ujk != 0.0) {
                    for (int l = lcp[j] + lane; l < lcp[j + 1]; l += kWave) {
                        const int row = lri[l];
                        work[row] = work[row] - lx[l] * ujk;
                    }
                }
                wave_sync();
            }
            for (int u = us + lane; u < ue; u += kWave) ux[u] = work[uri[u]];
            const double diag = clamp_pivot(work[g], colflag[g] & rsgpu::COL_SINGLETON);
            if (lane == 0) {
                ux[ue] = diag;
                ud[g] = diag;
            }
            for (int l = lcp[g] + lane; l < lcp[g + 1]; l += kWave) lx[l] = work[lri[l]] / diag;
            wave_sync();
            // clear touched entries (keeps the all-zero invariant of work)
            for (int s = sc_ptr[g] + lane; s < sc_ptr[g + 1]; s += kWave) work[sc_dst[s]] = 0.0;
            for (int l = lcp[g] + lane; l < lcp[g + 1]; l += kWave) work[lri[l]] = 0.0;
            for (int u = us + lane; u < ue; u += kWave) work[uri[u]] = 0.0;
            if (lane == 0) work[g] = 0.0;
            wave_sync();
        }
        __syncthreads();
    }
    if (d.flags & rsgpu::PLAN_CSR) csr_copies(d, ip, vpool, lx, ux);
}


// Values V: lx | pad | ux | pad (indices of the schedule), in LDS (kLdsV) or
// in the value pool; G is the value pool at lx (ud follows V there).
template <bool kLdsV>
__global__ __launch_bounds__(1024) void k_refactor_rl(const PlanDev* __restrict__ descs,
                                                     const int32_t* __restrict__ ip,
                                                     const double* __restrict__ axpool,
                                                     double* __restrict__ vpool,
                                                     const int32_t* __restrict__ skip) {
    extern __shared__ double smem[];
    if (skip && *skip) return;
    const PlanDev d = descs[blockIdx.x];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    const int tid = threadIdx.x, bs = blockDim.x;
    double* S = smem + wid * kWave;  // ordered-sum staging of this wavefront
    double* G = vpool + d.o_lx;
    double* V = kLdsV ? smem + nw * kWave : G;
    const int nv = d.rl_nv;
    const int ux0 = d.nl + 1, ud0 = d.nl + 1 + d.nu + 1;
    const double* ax = axpool + d.o_ax;
    const int32_t* ini_src = ip + d.o_rl_ini_src;
    const int32_t* ini_cnt = ip + d.o_rl_ini_cnt;
    const int32_t* div_ptr = ip + d.o_rl_div_ptr;
    const int2* div = reinterpret_cast<const int2*>(ip + d.o_rl_div);
    const int32_t* piv_ptr = ip + d.o_rl_piv_ptr;
    const int32_t* piv = ip + d.o_rl_piv;
    const int32_t* piv_a = ip + d.o_rl_piv_a;
    const int32_t* task_ptr = ip + d.o_rl_task_ptr;
    const int32_t* task_nlong = ip + d.o_rl_task_nlong;
    const int32_t* task_tgt = ip + d.o_rl_task_tgt;
    const int32_t* task_beg = ip + d.o_rl_task_beg;
    const int2* pair = reinterpret_cast<const int2*>(ip + d.o_rl_pair);
    const int32_t* sgl_ptr = ip + d.o_rl_sgl_ptr;
    const unsigned long long* sgl = reinterpret_cast<const unsigned long long*>(ip + d.o_rl_sgl);

    // A scattered into every slot: duplicates summed in CSC order from 0.0
    for (int a = tid; a < nv; a += bs) {
        double v = 0.0;
        const int c = ini_cnt[a], s0 = ini_src[a];
        for (int q = 0; q < c; ++q) v += ax[s0 + q];
        V[a] = v;
    }
    __syncthreads();
    for (int p = 0; p < d.rl_nph; ++p) {
        // (a) L columns of this phase's pivots over their clamped diagonals
        for (int k = div_ptr[p] + tid; k < div_ptr[p + 1]; k += bs) {
            const int2 dv = div[k];
            V[dv.x] = V[dv.x] / clamp_pivot(V[dv.y], false);
        }
        __syncthreads();
        // (b) the clamped diagonals, then the ordered updates of the phase
        for (int k = piv_ptr[p] + tid; k < piv_ptr[p + 1]; k += bs) {
            const int j = piv[k], pa = piv_a[k];
            const int a = pa & 0x7FFFFFFF;
            const double diag = clamp_pivot(V[a], pa < 0);
            V[a] = diag;
            G[ud0 + j] = diag;
        }
        const int t0 = task_ptr[p], tl = t0 + task_nlong[p], t1 = task_ptr[p + 1];
        for (int k = t0 + wid; k < tl; k += nw) {
            const int tgt = task_tgt[k], b = task_beg[k], e = task_beg[k + 1];
            double v = V[tgt];
            for (int base = b; base < e; base += kWave) {
                const int q = base + lane;
                double pr = 0.0;
                bool act = false;
                if (q < e) {
                    const int2 pq = pair[q];
                    const double uu = V[pq.y];
                    if (uu != 0.0) {
                        pr = V[pq.x] * uu;
                        act = true;
                    }
                }
                // a skipped update stages +0.0, and v - (+0.0) is v bit for bit
                S[lane] = act ? pr : 0.0;
                wave_sync();
                if (lane == 0) v = ordered_sub(v, S, min(kWave, e - base));
                wave_sync();
            }
            if (lane == 0) V[tgt] = v;
        }
        // one-update tasks, kSglBatch per thread with every load issued first
        const int g0 = sgl_ptr[p], g1 = sgl_ptr[p + 1];
        // (records packed as target | L << 21 | U << 42; slots past the
        // phase's end load nothing)
        constexpr unsigned long long kSlot = (1ull << rsgpu::kRlSlotBits) - 1;
        for (int base = g0 + tid; base < g1; base += kSglBatch * bs) {
            unsigned long long rc[kSglBatch];
#pragma unroll
            for (int i = 0; i < kSglBatch; ++i) {
                const int k = base + i * bs;
                rc[i] = k < g1 ? sgl[k] : 0ull;
            }
            double vt[kSglBatch], vl[kSglBatch], vu[kSglBatch];
#pragma unroll
            for (int i = 0; i < kSglBatch; ++i) {
                vu[i] = 0.0;
                if (base + i * bs < g1) {
                    vt[i] = V[rc[i] & kSlot];
                    vl[i] = V[(rc[i] >> rsgpu::kRlSlotBits) & kSlot];
                    vu[i] = V[rc[i] >> (2 * rsgpu::kRlSlotBits)];
                }
            }
#pragma unroll
            for (int i = 0; i < kSglBatch; ++i)
                if (base + i * bs < g1 && vu[i] != 0.0) V[rc[i] & kSlot] = vt[i] - vl[i] * vu[i];
        }
        // short tasks: pairs and values loaded eight at a time (every load
        // in bounds: a task has at least one pair), updates in order
        for (int k = tl + tid; k < t1; k += bs) {
            const int tgt = task_tgt[k], b = task_beg[k], e = task_beg[k + 1];
            double v = V[tgt];
            for (int q0 = b; q0 < e; q0 += 8) {
                int2 pq[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) pq[i] = pair[min(q0 + i, e - 1)];
                double lv[8], uv[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    uv[i] = V[pq[i].y];
                    lv[i] = V[pq[i].x];
                }
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    if (q0 + i < e && uv[i] != 0.0) v = v - lv[i] * uv[i];
            }
            V[tgt] = v;
        }
        __syncthreads();
    }
    if (kLdsV)
        for (int a = tid; a < nv; a += bs) G[a] = V[a];
    if (d.flags & rsgpu::PLAN_CSR) csr_copies(d, ip, vpool, V, V + ux0);
}

// One level of a row-oriented triangular sweep: T[r] = (T[r] - sum_e v_e * T[c_e]) (/ diag)
// with the subtractions in entry order. Entry values come from the row-ordered
// copy (kCsr) or through the position map. Long rows (the level's first
// nlong) take a wavefront each.
template <bool kCsr, bool kDiv>
__device__ __forceinline__ void sweep_level(double* T, double* S, const int32_t* rows, int q0,
                                            int nlong, int q1, const int32_t* __restrict__ rp,
                                            const int32_t* __restrict__ rc,
                                            const int32_t* __restrict__ rv,
                                            const double* __restrict__ vals,
                                            const double* __restrict__ ud) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    if (q1 - q0 <= 2 * nw) nlong = q1 - q0;  // a narrow level: every row takes a wavefront
    for (int q = q0 + wid; q < q0 + nlong; q += nw) {
        const int r = rows[q];
        const int e0 = rp[r], e1 = rp[r + 1];
        double t = T[r];
        for (int base = e0; base < e1; base += kWave) {
            const int e = base + lane;
            double pr = 0.0;
            if (e < e1) pr = vals[kCsr ? e : rv[e]] * T[rc[e]];
            S[lane] = pr;
            wave_sync();
            if (lane == 0) t = ordered_sub(t, S, min(kWave, e1 - base));
            wave_sync();
        }
        if (lane == 0) T[r] = kDiv ? t / ud[r] : t;
    }
    for (int q = q0 + nlong + threadIdx.x; q < q1; q += blockDim.x) {
        const int r = rows[q];
        const int e0 = rp[r], e1 = rp[r + 1];
        if (!kDiv && e0 == e1) continue;
        double t = T[r];
        for (int e = e0; e < e1; ++e) t = t - vals[kCsr ? e : rv[e]] * T[rc[e]];
        T[r] = kDiv ? t / ud[r] : t;
    }
}

template <bool kCsr>
__device__ __forceinline__ void solve_body(const PlanDev& d, const int32_t* __restrict__ ip,
                                           const double* __restrict__ ax,
                                           const double* __restrict__ vpool, double* T, double* S) {
    const int32_t* lrp = ip + d.o_lrp;
    const int32_t* lrc = ip + d.o_lrc;
    const int32_t* lrv = ip + d.o_lrv;
    const int32_t* urp = ip + d.o_urp;
    const int32_t* urc = ip + d.o_urc;
    const int32_t* urv = ip + d.o_urv;
    const int32_t* lsl_blk = ip + d.o_lsl_blk;
    const int32_t* lsl_ptr = ip + d.o_lsl_ptr;
    const int32_t* lsl_rows = ip + d.o_lsl_rows;
    const int32_t* lsl_nlong = ip + d.o_lsl_nlong;
    const int32_t* usl_blk = ip + d.o_usl_blk;
    const int32_t* usl_ptr = ip + d.o_usl_ptr;
    const int32_t* usl_rows = ip + d.o_usl_rows;
    const int32_t* usl_nlong = ip + d.o_usl_nlong;
    const int32_t* blk_ptr = ip + d.o_blk_ptr;
    const int32_t* odp = ip + d.o_odp;
    const int32_t* odax = ip + d.o_odax;
    const int32_t* odsrc = ip + d.o_odsrc;
    const int32_t* has_od = ip + d.o_blk_has_od;
    const double* lv = kCsr ? vpool + d.o_lxr : vpool + d.o_lx;
    const double* uv = kCsr ? vpool + d.o_uxr : vpool + d.o_ux;
    const double* ud = vpool + d.o_ud;
    for (int b = d.nblk - 1; b >= 0; --b) {
        if (has_od[b]) {
            for (int r = blk_ptr[b] + threadIdx.x; r < blk_ptr[b + 1]; r += blockDim.x) {
                const int e0 = odp[r], e1 = odp[r + 1];
                if (e0 == e1) continue;
                double t = T[r];
                for (int e = e0; e < e1; ++e) t = t - ax[odax[e]] * T[odsrc[e]];
                T[r] = t;
            }
            __syncthreads();
        }
        for (int lev = lsl_blk[b]; lev < lsl_blk[b + 1]; ++lev) {
            sweep_level<kCsr, false>(T, S, lsl_rows, lsl_ptr[lev], lsl_nlong[lev], lsl_ptr[lev + 1],
                                     lrp, lrc, lrv, lv, ud);
            __syncthreads();
        }
        for (int lev = usl_blk[b]; lev < usl_blk[b + 1]; ++lev) {
            sweep_level<kCsr, true>(T, S, usl_rows, usl_ptr[lev], usl_nlong[lev], usl_ptr[lev + 1],
                                    urp, urc, urv, uv, ud);
            __syncthreads();
        }
    }
}

// Shared memory: 64 doubles per wavefront (ordered sums of long rows), then
// the solve vector when kLds.
template <bool kLds>
__global__ __launch_bounds__(1024) void k_solve(const PlanDev* __restrict__ descs,
                                               const int32_t* __restrict__ ip,
                                               const double* __restrict__ axpool,
                                               const double* __restrict__ vpool,
                                               const double* __restrict__ rhspool,
                                               double* __restrict__ xpool,
                                               double* __restrict__ wpool,
                                               const int32_t* __restrict__ skip,
                                               double* __restrict__ tvec) {
    extern __shared__ double smem[];
    if (skip && *skip) return;
    const PlanDev d = descs[blockIdx.x];
    const int n = d.n;
    const int nw = blockDim.x / kWave;
    double* S = smem + (threadIdx.x / kWave) * kWave;
    double* T = kLds ? smem + nw * kWave : (tvec ? tvec : wpool + d.o_tvec);
    const int32_t* perm_in = ip + d.o_perm_in;
    const int32_t* perm_out = ip + d.o_perm_out;
    const double* ax = axpool + d.o_ax;
    const double* rhs = rhspool + d.o_vec;
    double* x = xpool + d.o_vec;

    for (int i = threadIdx.x; i < n; i += blockDim.x) T[perm_in[i]] = rhs[i];
    __syncthreads();
    if (d.flags & rsgpu::PLAN_CSR)
        solve_body<true>(d, ip, ax, vpool, T, S);
    else
        solve_body<false>(d, ip, ax, vpool, T, S);
    for (int i = threadIdx.x; i < n; i += blockDim.x) x[i] = T[perm_out[i]];
}

// ---- flat pipelined level solve (solve mode 3) ----
// Each level in two phases: (a) every entry's product v * T[col] into the
// LDS buffer P, one slot per thread; (b) one thread per row subtracts its
// products from T[row] in entry order (then divides by the U diagonal): the
// same operations in the same order as the row sweeps, so the result is
// bit-identical. The static slot and row data of level l + 2 and the values
// and diagonals of level l + 1 are loaded while level l computes; loads are
// unconditional from in-bounds indices (slot 0, record 0 when there is
// nothing to load) and loaded registers are never rewritten, so a level
// waits for loads issued a level earlier only.
constexpr int kFlatSlots = 2;  // slots per thread loaded ahead

struct FlStatic {
    int col[kFlatSlots], src[kFlatSlots];
    int4 row;
};
struct FlDyn {
    double v[kFlatSlots];
    double dg;
};

// Barrier for data exchanged through LDS only: global loads issued before it
// stay in flight.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

__device__ __forceinline__ FlStatic fl_static(const int4* LV, int l, int nlev, int tid, int bs,
                                              const int32_t* __restrict__ col, const int32_t* __restrict__ src,
                                              const int4* __restrict__ rows) {
    FlStatic f;
    const int lc = l < nlev ? l : nlev - 1;
    const int4 a = LV[lc], b = LV[lc + 1];
#pragma unroll
    for (int i = 0; i < kFlatSlots; ++i) {
        const int sl = a.x + tid + i * bs;
        const int q = l < nlev && sl < b.x ? sl : 0;
        f.col[i] = col[q];
        f.src[i] = src[q];
    }
    const int k = a.y + tid;
    f.row = rows[l < nlev && k < b.y ? k : 0];
    return f;
}

__device__ __forceinline__ FlDyn fl_dyn(const FlStatic& f, const double* __restrict__ vals,
                                        const double* __restrict__ ud) {
    FlDyn g;
#pragma unroll
    for (int i = 0; i < kFlatSlots; ++i) g.v[i] = vals[f.src[i]];
    g.dg = ud[f.row.x];
    return g;
}

__device__ __forceinline__ void fl_level(double* T, double* P, const int4* LV, int l, const FlStatic& f,
                                         const FlDyn& g, int tid, int bs, const int32_t* __restrict__ col,
                                         const int32_t* __restrict__ src, const int4* __restrict__ rows,
                                         const double* __restrict__ vals, const double* __restrict__ ud) {
    const int4 a = LV[l], b = LV[l + 1];
    const int ns = b.x - a.x;
#pragma unroll
    for (int i = 0; i < kFlatSlots; ++i)
        if (tid + i * bs < ns) P[tid + i * bs] = g.v[i] * T[f.col[i]];
    for (int t = tid + kFlatSlots * bs; t < ns; t += bs) P[t] = vals[src[a.x + t]] * T[col[a.x + t]];
    lds_barrier();
    const int div = a.z;
    if (a.y + tid < b.y) {  // the thread's first row: record and diagonal loaded ahead
        double t = T[f.row.x];
        t = ordered_sub(t, P + (f.row.y - a.x), f.row.z - f.row.y);
        T[f.row.x] = div ? t / g.dg : t;
    }
    for (int k = a.y + tid + bs; k < b.y; k += bs) {
        const int4 rr = rows[k];
        double t = T[rr.x];
        t = ordered_sub(t, P + (rr.y - a.x), rr.z - rr.y);
        T[rr.x] = div ? t / ud[rr.x] : t;
    }
    lds_barrier();
}

// Shared memory: T (n doubles, rounded up to an even count), P (sq_maxent
// doubles, rounded up to even), the level table (sq_nlev + 1 int4).
__global__ __launch_bounds__(1024) void k_solve_lp(const PlanDev* __restrict__ descs,
                                                  const int32_t* __restrict__ ip,
                                                  const double* __restrict__ vpool,
                                                  const double* __restrict__ rhspool,
                                                  double* __restrict__ xpool,
                                                  const int32_t* __restrict__ skip) {
    extern __shared__ double smem[];
    if (skip && *skip) return;
    const PlanDev d = descs[blockIdx.x];
    const int n = d.n, nlev = d.sq_nlev;
    const int tid = threadIdx.x, bs = blockDim.x;
    double* T = smem;
    double* P = T + ((n + 1) & ~1);
    int4* LV = reinterpret_cast<int4*>(P + ((d.sq_maxent + 1) & ~1));
    const int4* glev = reinterpret_cast<const int4*>(ip + d.o_sq_lev);
    const int4* rows = reinterpret_cast<const int4*>(ip + d.o_sq_row);
    const int32_t* col = ip + d.o_sq_col;
    const int32_t* src = ip + d.o_sq_src;
    const int32_t* perm_in = ip + d.o_perm_in;
    const int32_t* perm_out = ip + d.o_perm_out;
    const double* vals = vpool + d.o_lx;
    const double* ud = vpool + d.o_ud;
    const double* rhs = rhspool + d.o_vec;
    double* x = xpool + d.o_vec;
    for (int i = tid; i < n; i += bs) T[perm_in[i]] = rhs[i];
    for (int l = tid; l <= nlev; l += bs) LV[l] = glev[l];
    __syncthreads();
    // static data rotates over three register sets, values over two
    FlStatic A0 = fl_static(LV, 0, nlev, tid, bs, col, src, rows), A1 = fl_static(LV, 1, nlev, tid, bs, col, src, rows),
             A2;
    FlDyn G0 = fl_dyn(A0, vals, ud), G1;
#define FL_STEP(L, AC, AN, AF, GC, GN)                                         \
    if ((L) < nlev) {                                                         \
        GN = fl_dyn(AN, vals, ud);                                            \
        AF = fl_static(LV, (L) + 2, nlev, tid, bs, col, src, rows);           \
        fl_level(T, P, LV, (L), AC, GC, tid, bs, col, src, rows, vals, ud);   \
    }
    for (int l = 0; l < nlev; l += 6) {
        FL_STEP(l, A0, A1, A2, G0, G1)
        FL_STEP(l + 1, A1, A2, A0, G1, G0)
        FL_STEP(l + 2, A2, A0, A1, G0, G1)
        FL_STEP(l + 3, A0, A1, A2, G1, G0)
        FL_STEP(l + 4, A1, A2, A0, G0, G1)
        FL_STEP(l + 5, A2, A0, A1, G1, G0)
    }
#undef FL_STEP
    __syncthreads();
    for (int i = tid; i < n; i += bs) x[i] = T[perm_out[i]];
}

// out[0] = max |ax|, out[1] = max |ux| (+inf when some ux is not finite):
// the element growth of a refactor on reused pivots. One workgroup.
__global__ __launch_bounds__(256) void k_growth(const double* __restrict__ ax, int nnz,
                                                const double* __restrict__ ux, int nu,
                                                double* __restrict__ out) {
    __shared__ double sa[256], su[256];
    double ma = 0.0, mu = 0.0;
    for (int i = threadIdx.x; i < nnz; i += blockDim.x) ma = fmax(ma, fabs(ax[i]));
    for (int i = threadIdx.x; i < nu; i += blockDim.x) {
        const double v = fabs(ux[i]);
        mu = (v <= 1.7976931348623157e308) ? fmax(mu, v) : __builtin_huge_val();
    }
    sa[threadIdx.x] = ma;
    su[threadIdx.x] = mu;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sa[threadIdx.x] = fmax(sa[threadIdx.x], sa[threadIdx.x + w]);
            su[threadIdx.x] = fmax(su[threadIdx.x], su[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sa[0];
        out[1] = su[0];
    }
}

}  // namespace

namespace rsgpu {

hipError_t launch_growth(const double* ax, int nnz, const double* ux, int nu, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_growth, dim3(1), dim3(256), 0, s, ax, nnz, ux, nu, out);
    return hipGetLastError();
}

// mode: bit 0 LDS (work vectors, or the values of the right-looking
// schedule), bit 1 right-looking schedule.
hipError_t launch_refactor(const PlanDev* descs, int nmat, const int32_t* ip, const double* ax,
                           double* vals, double* wpool, int max_n, int mode, int threads,
                           hipStream_t stream, int rl_nv, const int32_t* skip) {
    const bool lds = mode & 1;
    if (mode & 2) {
        size_t shm = (size_t)(threads / kWave) * kWave * sizeof(double);
        if (lds) {
            shm += (size_t)rl_nv * sizeof(double);
            hipLaunchKernelGGL(k_refactor_rl<true>, dim3(nmat), dim3(threads), shm, stream, descs, ip,
                               ax, vals, skip);
        } else {
            hipLaunchKernelGGL(k_refactor_rl<false>, dim3(nmat), dim3(threads), shm, stream, descs,
                               ip, ax, vals, skip);
        }
        return hipGetLastError();
    }
    size_t shm = lds ? (size_t)(threads / kWave) * max_n * sizeof(double) : 0;
    if (lds)
        hipLaunchKernelGGL(k_refactor<true>, dim3(nmat), dim3(threads), shm, stream, descs, ip, ax,
                           vals, wpool, skip);
    else
        hipLaunchKernelGGL(k_refactor<false>, dim3(nmat), dim3(threads), 0, stream, descs, ip, ax,
                           vals, wpool, skip);
    return hipGetLastError();
}

// Solve mode 3 (k_solve_lp) of one matrix whose plan has level records.
size_t solve_lp_lds(int n, int maxent, int nlev) {
    return (size_t)((n + 1) & ~1) * sizeof(double) + (size_t)((maxent + 1) & ~1) * sizeof(double) +
           (size_t)(nlev + 1) * 4 * sizeof(int32_t);
}
hipError_t launch_solve_lp(const PlanDev* descs, const int32_t* ip, const double* vals, const double* rhs,
                           double* x, int n, int maxent, int nlev, int threads, hipStream_t stream,
                           const int32_t* skip) {
    hipLaunchKernelGGL(k_solve_lp, dim3(1), dim3(threads), solve_lp_lds(n, maxent, nlev), stream, descs, ip,
                       vals, rhs, x, skip);
    return hipGetLastError();
}

hipError_t launch_solve(const PlanDev* descs, int nmat, const int32_t* ip, const double* ax,
                        const double* vals, const double* rhs, double* x, double* wpool, int max_n,
                        bool lds, int threads, hipStream_t stream, const int32_t* skip, double* tvec) {
    if (tvec && nmat != 1) return hipErrorInvalidValue;
    size_t shm = (size_t)(threads / kWave) * kWave * sizeof(double) +
                 (lds ? (size_t)max_n * sizeof(double) : 0);
    if (lds)
        hipLaunchKernelGGL(k_solve<true>, dim3(nmat), dim3(threads), shm, stream, descs, ip, ax, vals,
                           rhs, x, wpool, skip, tvec);
    else
        hipLaunchKernelGGL(k_solve<false>, dim3(nmat), dim3(threads), shm, stream, descs, ip, ax, vals,
                           rhs, x, wpool, skip, tvec);
    return hipGetLastError();
}

}  // namespace rsgpu
