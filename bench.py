#!/usr/bin/env python3
"""Benchmark: refactor+solve/s of the GPU LinearSolver drop-in on the
500-stage level-1 CMOS inverter chain transient (.tran 5p 10n, max step 5p).

One step = one Engine TRAN run of that circuit through the GPU backend
(crates/sim-core/src/engine.rs:333-632): device stamping, refactor and solve
in every Newton iteration. Under the reference's own semantics the TRAN's
initial DC operating point ends MaxIters after 100 Newton iterations (the
level-1 PMOS frame of bsim/evaluate.rs:298-321 gives the chain no DC
solution), so a step is 100 stamp+refactor+solve iterations on the n=504 DC
matrix followed by the reference's "DC operating point failed" exit.

value = refactor+solve calls per second over the timed steps (inputs resident
in device memory). The DC OP runs device-resident: stamping, factor, solve and
the convergence test of all 100 iterations in one k_batch_dcop launch. The
host-driven loop (one refactor and one solve launch per iteration, the
convergence test on the host) is measured beside it.
roofline = the dominant kernel's algorithmic bytes per launch over its average
launch time from HIP events on its launch stream during the timed region.
cpu_baseline = the C restatement of NativeKluSolver + Engine (oracle/) on one
host core, same circuit, same semantics, checked against the GPU result.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "refactor+solve/s and batched DC-OPs/s at <=1e-9 rel err vs FP64 KLU"
HBM_PEAK_GBS = 8000.0


def _same(a, b):
    """Equal doubles, NaN == NaN (NaN payloads differ between x86 and GPU)."""
    return (a != a and b != b) or a == b


def _algorithmic_bytes(nnz_a, nnz_l, nnz_u, n):
    """Compulsory traffic per launch (FP64 values 8 B, int32 indices 4 B)."""
    refactor = 12 * nnz_a + 12 * (nnz_l + nnz_u) + 8 * n
    solve = 12 * (nnz_l + nnz_u) + 24 * n
    return refactor, solve


def _measured_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload (scaled by a plain copy's counters), or None."""
    path = os.path.join(ROOT, "profiles", "k_batch_dcop_chain500_traffic.md")
    if kernel != "k_batch_dcop" or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(circ, tran_args, budget_s, gpu_res):
    """Oracle (C restatement) timed on one host core: same circuit, same
    reference semantics. Returns the baseline dict and a parity verdict."""
    from oracle import oracle as O

    runs, calls = 0, 0
    t0 = time.perf_counter()
    last = None
    while True:
        r = O.tran(circ, store=False, **tran_args)
        runs += 1
        calls += r["factor_calls"]
        last = r
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    parity = (last["status"] == gpu_res["status"] and last["dc_iters"] == gpu_res["dc_iters"])
    # bit-level check of the Newton trajectory through the DC OP entry
    xo, ro = O.dc_op(circ)
    return dict(value=calls / dt, unit="refactor+solve/s", cores=1, kind="port",
                sample=f"{runs} full TRAN runs of the 500-stage chain ({calls} refactor+solve "
                       f"calls, {dt:.1f} s), oracle/ NativeKLU restatement, 1 thread"), \
        dict(status_match=parity, oracle_dc_final_norm=ro.final_norm, oracle_dc_iters=ro.iterations)


def batched_dc_ops(args, E, NL, S, torch):
    """4,096 Monte-Carlo inverter chains (stages ~ U[50,400], VTO += N(0,20mV),
    u0 *= N(1,0.05)), every DC OP in one persistent launch; compared with
    looping the plain GPU backend (rs_dc_op per circuit) on a sample."""
    import numpy as np

    t0 = time.perf_counter()
    circs = NL.monte_carlo_chains(args.batch)
    b = E.GpuBatch(circs)
    setup_s = time.perf_counter() - t0
    info = b.info()
    b.dc_op()  # warm-up
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.batch_steps):
        b.dc_op()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ev0.elapsed_time(ev1) / args.batch_steps
    out = b.results()
    iters = np.array([r.iterations for _, r in out])
    conv = int(sum(r.converged for _, r in out))
    dcops = args.batch * args.batch_steps / wall
    # looping the plain single-circuit GPU backend on a sample of instances
    k = min(args.loop_sample, len(circs))
    idx = np.linspace(0, len(circs) - 1, k).astype(int)
    gcs = [(E.GpuCircuit(circs[i]), S.GpuSolver(circs[i].node_count)) for i in idx]
    for g, _ in gcs:  # plain backend: refactor/solve launches driven from the host
        g.set_device_newton(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    same = True
    for (g, s), i in zip(gcs, idx):
        _, r = g.dc_op(s)
        same &= (r.iterations == out[i][1].iterations and _same(r.final_norm, out[i][1].final_norm))
    torch.cuda.synchronize()
    loop_per = (time.perf_counter() - t0) / k
    return dict(
        workload=f"{args.batch} Monte-Carlo level-1 inverter chains, stages U[50,400], DC OP each "
                 "(reference semantics: NativeKLU, no gmin/source stepping)",
        dc_ops_per_s=dcops, refactor_solve_per_s=float(iters.sum()) * args.batch_steps / wall,
        ms_per_batch=1e3 * wall / args.batch_steps, kernel_ms_per_batch=kern_ms,
        instances=args.batch, patterns=info["patterns"], grid=info["grid"], nmax=info["nmax"],
        converged=conv, newton_iters_total=int(iters.sum()), setup_s=setup_s,
        loop_plain_backend_dc_ops_per_s=1.0 / loop_per, loop_sample=k,
        speedup_vs_loop_plain=dcops * loop_per, loop_results_identical=bool(same))


def adjoint_cost(E, NL, S, torch, reps=5):
    """DC adjoint (dV/dp for every device) on the 64x64 resistor mesh against
    the forward DC OP through the same backend (SparseLU, GPU)."""
    c = NL.resistor_mesh(64, 64)
    g = E.GpuCircuit(c)
    s = S.GpuSolver(c.node_count, S.SPARSELU)
    x, r = g.dc_op(s)
    probe = c.node_count // 2
    g.dc_adjoint(x, probe)  # first call: host analysis of the Jacobian pattern
    g.dc_adjoint(x, probe)  # second: the first device refactor times the launch shapes
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.dc_op(s)
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        dp, _ = g.dc_adjoint(x, probe)
    torch.cuda.synchronize()
    t_adj = (time.perf_counter() - t0) / reps
    # one forward refactor + solve of the same matrix (GPU refactor path)
    ap, ai, ax, rhs, n = g.stamp(x)
    m = S.GpuSolver(n, S.SPARSELU)
    m.factor(ap, ai, ax)
    m.factor(ap, ai, ax)  # the first refactor times the launch shapes
    b = rhs.copy()
    m.solve(b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        m.factor(ap, ai, ax)
        b = rhs.copy()
        m.solve(b)
    torch.cuda.synchronize()
    t_rs = (time.perf_counter() - t0) / reps
    return dict(workload="64x64 resistor mesh (n=4,098), dV(center)/dp for all 12,162 devices (R, V, gmin R)",
                forward_dc_op_ms=1e3 * t_fwd, forward_refactor_solve_ms=1e3 * t_rs,
                adjoint_ms=1e3 * t_adj, adjoint_over_dc_op=t_adj / t_fwd,
                adjoint_over_refactor_solve=t_adj / t_rs, newton_iters=r.iterations,
                sensitivities=int(len(dp)))


def mesh_refactor_solve(E, NL, S, torch, np, reps=10):
    """Refactor + solve of the 64x64 mesh DC matrix (n=4,098), NativeKLU
    semantics, inputs resident in HBM, with the autotuned launch shapes."""
    c = NL.resistor_mesh(64, 64)
    g = E.GpuCircuit(c)
    ap, ai, ax, rhs, n = g.stamp(np.zeros(c.node_count))
    s = S.GpuSolver(n, S.NATIVEKLU)
    s.factor(ap, ai, ax)
    d_ax = torch.tensor(ax, dtype=torch.float64, device="cuda")
    d_b = torch.tensor(rhs, dtype=torch.float64, device="cuda")
    d_x = torch.empty_like(d_b)
    s.refactor_device(d_ax)  # the first refactor and solve time the launch shapes
    s.solve_device(d_b, d_x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        s.refactor_device(d_ax)
        s.solve_device(d_b, d_x)
    torch.cuda.synchronize()
    t_gpu = (time.perf_counter() - t0) / reps
    res = dict(workload="64x64 resistor mesh DC matrix (n=4,098), refactor + solve, NativeKLU semantics",
               gpu_ms=1e3 * t_gpu, launch_shape=s.launch_shape())
    return res, (ap, ai, ax, rhs, n, d_x.cpu().numpy())


# Per device kind: terminal voltages read, stamp values written, FP64 flops
# of one evaluation (R: g = 1/R; C, TRAP: g = 2C/dt and the history current;
# level-1 MOSFET: region, current and three conductances; diode: exp and g).
_EVAL_COST = {1: (2, 4, 1), 2: (2, 6, 6), 3: (2, 5, 0), 4: (2, 2, 0), 5: (2, 6, 10), 6: (4, 10, 25)}


def kernel_rooflines(E, NL, S, torch, np, reps=10, nsteps=20):
    """Achieved rate of each single-matrix kernel against its roofline, HIP
    events around every launch (rs_profile): k_refactor* and k_solve* on the
    64x64 mesh DC matrix (B_f = 12 nnz(A) + 12 (nnz L + nnz U) + 8 n,
    B_s = 12 (nnz L + nnz U) + 24 n bytes per launch); k_eval and k_assemble
    in a host-driven fixed-step TRAP transient of the 32x32 RC mesh with 32
    inverter taps (eval: device records, terminal voltages and capacitor
    history read, stamp values written, and its FP64 flops; assemble: 20 B per
    matrix and RHS contribution (slot index, value read and written) + 8 n)."""
    from rustspice_amd import _lib

    out = {}
    c = NL.resistor_mesh(64, 64)
    g = E.GpuCircuit(c)
    ap, ai, ax, rhs, n = g.stamp(np.zeros(c.node_count))
    s = S.GpuSolver(n, S.NATIVEKLU)
    s.factor(ap, ai, ax)
    d_ax = torch.tensor(ax, dtype=torch.float64, device="cuda")
    d_b = torch.tensor(rhs, dtype=torch.float64, device="cuda")
    d_x = torch.empty_like(d_b)
    s.refactor_device(d_ax)
    s.solve_device(d_b, d_x)
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    for _ in range(reps):
        s.refactor_device(d_ax)
        s.solve_device(d_b, d_x)
    torch.cuda.synchronize()
    prof = _lib.profile_get()
    _lib.profile_enable(False)
    ps = S.HostFactor(n, ap, ai, ax).plan_stats()
    b_ref, b_sol = _algorithmic_bytes(len(ai), ps["nnz_l"], ps["nnz_u"], n)
    f_sol = 2 * (ps["nnz_l"] + ps["nnz_u"] - n) + n
    sh = s.launch_shape()
    for name, b, mode in (("refactor", b_ref, sh["ref_mode"]), ("solve", b_sol, sh["sol_mode"])):
        p = prof[name]
        gbs = b / (p["avg_us"] * 1e-6) / 1e9
        out[f"{name}_mesh64"] = dict(avg_us=p["avg_us"], count=p["count"], bytes=b, gbs=gbs,
                                     hbm_frac=gbs / HBM_PEAK_GBS, mode=mode, threads=sh[f"{name[:3]}_threads"])
    out["solve_mesh64"]["gflops"] = f_sol / (prof["solve"]["avg_us"] * 1e-6) / 1e9
    s.close()

    c = NL.rc_mesh_inverters(32, 32)
    g = E.GpuCircuit(c)
    s = S.GpuSolver(c.node_count)
    g.tran_fixed(s, NL.C5_TRAN["dt"], 2)
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    g.tran_fixed(s, NL.C5_TRAN["dt"], nsteps)
    torch.cuda.synchronize()
    prof = _lib.profile_get()
    _lib.profile_enable(False)
    kinds = [int(k) for k in c.devices["kind"]]
    reads = sum(c.devices.dtype.itemsize + 8 * _EVAL_COST[k][0] + (16 if k == 2 else 0) for k in kinds)
    writes = sum(8 * _EVAL_COST[k][1] for k in kinds)
    flops = sum(_EVAL_COST[k][2] for k in kinds)
    pe = prof["eval"]
    out["eval_rcmesh32"] = dict(avg_us=pe["avg_us"], count=pe["count"], bytes=reads + writes,
                                gbs=(reads + writes) / (pe["avg_us"] * 1e-6) / 1e9,
                                fp64_gflops=flops / (pe["avg_us"] * 1e-6) / 1e9,
                                fp64_frac=flops / (pe["avg_us"] * 1e-6) / 78.6e12, devices=len(kinds))
    apt, ait = g.pattern(E.MODE_TRAN_TRAP)
    nt = len(apt) - 1
    n_rhs = sum(_EVAL_COST[k][1] for k in kinds) - len(ait)  # stamp values not in the matrix
    b_asm = 20 * (len(ait) + max(n_rhs, 0)) + 8 * nt
    pa = prof["assemble"]
    out["assemble_rcmesh32"] = dict(avg_us=pa["avg_us"], count=pa["count"], bytes=b_asm,
                                    gbs=b_asm / (pa["avg_us"] * 1e-6) / 1e9,
                                    hbm_frac=b_asm / (pa["avg_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS)
    s.close()
    return out


def tran_adjoint_cost(E, NL, S, nx=32, nsteps=500, every=25):
    """Transient adjoint on the RC mesh driven by level-1 inverters (the
    adjoint workload at nx x nx with nx*nx/32 taps): dV(centre at t =
    nsteps*2 ps)/dp for every R, C and MOSFET W, checkpoint every `every`
    steps, against one forward fixed-step TRAP transient through the same
    plain GPU backend (adjoint_over_forward) and against the adjoint's own
    forward pass, which solves every step to its exact root
    (adjoint_over_exact_forward)."""
    import numpy as np

    c = NL.rc_mesh_inverters(nx, nx)
    probe = NL.rc_mesh_probe(nx, nx)
    dt = NL.C5_TRAN["dt"]
    g = E.GpuCircuit(c)
    s = S.GpuSolver(c.node_count)
    st, _, _ = g.tran_fixed(s, dt, 2)  # warm-up: patterns, first factors, launch shapes
    t0 = time.perf_counter()
    st, xf, it = g.tran_fixed(s, dt, nsteps)
    t_fwd = time.perf_counter() - t0
    g.tran_adjoint(s, dt, 2, probe, every)  # warm-up of the adjoint's solvers
    t0 = time.perf_counter()
    dp, xa, info = g.tran_adjoint(s, dt, nsteps, probe, every)
    t_adj = time.perf_counter() - t0
    if info["status"] != 0:
        return dict(workload=f"{nx}x{nx} RC mesh adjoint", adjoint_info=info, forward_status=int(st))
    # every state kept in host memory instead of checkpoints: no recomputation
    t0 = time.perf_counter()
    dp_all, _, info_all = g.tran_adjoint(s, dt, nsteps, probe, 0)
    t_all = time.perf_counter() - t0
    kinds = [d["kind"] for d in c.devices]
    return dict(workload=f"{nx}x{nx} RC mesh, {nx * nx // 32} inverter taps, {nsteps} TRAP steps of 2 ps, "
                         f"checkpoint every {every}; dV(centre)/d(every R, C, W)",
                n=int(len(xa)), forward_status=int(st), forward_ms=1e3 * t_fwd, forward_newton_iters=int(it),
                adjoint_ms=1e3 * t_adj, adjoint_over_forward=t_adj / t_fwd,
                adjoint_over_exact_forward=t_adj / max(info["forward_ms"] * 1e-3, 1e-12), adjoint_info=info,
                keep_all_states=dict(adjoint_ms=1e3 * t_all, adjoint_over_forward=t_all / t_fwd,
                                     states=int(info_all["checkpoints"]),
                                     recomputed_steps=int(info_all["recomputed_steps"]),
                                     same_dp_as_checkpointed=bool(np.array_equal(dp_all, dp))),
                sensitivities=dict(R=kinds.count(NL.RS_DEV_R), C=kinds.count(NL.RS_DEV_C),
                                   W=kinds.count(NL.RS_DEV_M)),
                probe_v=float(xa[probe]),
                max_abs_dp=dict(R=float(np.abs(dp[np.array(kinds) == NL.RS_DEV_R]).max()),
                                C=float(np.abs(dp[np.array(kinds) == NL.RS_DEV_C]).max()),
                                W=float(np.abs(dp[np.array(kinds) == NL.RS_DEV_M]).max())))


def tran_graph_cost(E, NL, S):
    """Fixed-step TRAP transients host-driven against every step after the
    first as a captured hipGraph (bit-identical results)."""
    out = {}
    for name, c, dt, n in (("rc_ladder30", NL.rc_ladder(30), 1e-10, 500),
                           ("rcmesh16_taps8", NL.rc_mesh_inverters(16, 16), 2e-12, 200)):
        row = {}
        for graph in (False, True):
            g = E.GpuCircuit(c)
            g.set_tran_graph(graph)
            s = S.GpuSolver(c.node_count)
            g.tran_fixed(s, dt, 3)
            t0 = time.perf_counter()
            st, x, it = g.tran_fixed(s, dt, n)
            row["graph_ms" if graph else "host_ms"] = 1e3 * (time.perf_counter() - t0)
            row["newton_iters"] = int(it)
            row["graph_steps" if graph else "status"] = g.graph_steps() if graph else int(st)
        row["speedup"] = row["host_ms"] / row["graph_ms"]
        out[name] = row
    return out


def cpu_baseline_mesh(mesh, np, reps=10):
    """The oracle's refactor + solve of the same mesh matrix on one host core,
    and whether the GPU result is bit-identical to it."""
    from oracle import oracle as O

    ap, ai, ax, rhs, n, xg = mesh
    o = O.NativeKlu(n)
    o.factor(ap, ai, ax)
    t0 = time.perf_counter()
    for _ in range(reps):
        o.factor(ap, ai, ax)
        bb = rhs.copy()
        o.solve(bb)
    t_cpu = (time.perf_counter() - t0) / reps
    return dict(cpu_oracle_ms=1e3 * t_cpu, cores=1, kind="port",
                bit_identical=bool(np.array_equal(xg.view(np.int64), bb.view(np.int64))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch", type=int, default=4096, help="Monte-Carlo instances (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--loop-sample", type=int, default=16)
    ap.add_argument("--host-steps", type=int, default=3)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the per-kernel rooflines (mesh refactor/solve, RC mesh eval/assemble)")
    args = ap.parse_args()
    if args.copy_mib:
        import torch
        a = torch.empty(args.copy_mib << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": a.numel() * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the solver runs one matrix per device; see DESIGN.md for sharding)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from rustspice_amd import _lib
    from rustspice_amd import engine as E
    from rustspice_amd import netlists as NL
    from rustspice_amd import solver as S

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.kernels_only:
        print(json.dumps({"kernel_rooflines": kernel_rooflines(E, NL, S, torch, np)}))
        return
    circ = NL.inverter_chain(500)
    tran_args = dict(NL.C2_TRAN)
    g = E.GpuCircuit(circ)
    s = S.GpuSolver(circ.node_count, S.NATIVEKLU)

    for _ in range(args.warmup):
        g.tran(s, store=False, **tran_args)
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    t0 = time.perf_counter()
    res = []
    for _ in range(args.steps):
        res.append(g.tran(s, store=False, **tran_args))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = _lib.profile_get()
    _lib.profile_enable(False)
    elapsed = t1 - t0
    calls = sum(r["factor_calls"] for r in res)
    value = calls / elapsed
    st = s.stats()
    ap_dc, ai_dc = g.pattern(E.MODE_DC)
    n_dc = len(ap_dc) - 1
    # L/U sizes of the DC matrix from the host analysis (the device-resident
    # DC OP leaves no factorization on the solver handle)
    _, _, ax0, _, _ = g.stamp(np.zeros(circ.node_count), mode=E.MODE_DC)
    ps = S.HostFactor(n_dc, ap_dc, ai_dc, ax0).plan_stats()
    st.update(nnz_l=ps["nnz_l"], nnz_u=ps["nnz_u"], n_factor_levels=ps["factor_levels"],
              n_lsolve_levels=ps["lsolve_levels"], n_usolve_levels=ps["usolve_levels"])
    b_ref, b_sol = _algorithmic_bytes(len(ai_dc), st["nnz_l"], st["nnz_u"], n_dc)
    kern = {}
    # k_batch_dcop runs dc_iters refactor+solve iterations per launch
    b_batch = (b_ref + b_sol) * res[-1]["dc_iters"]
    for name, b in (("refactor", b_ref), ("solve", b_sol), ("batch", b_batch)):
        p = prof[name]
        if p["count"]:
            avg_s = p["avg_us"] * 1e-6
            kern[name] = dict(avg_us=p["avg_us"], count=p["count"], bytes=b,
                              gbs=b / avg_s / 1e9, total_ms=p["ms_total"])
    for name in ("eval", "assemble"):
        p = prof[name]
        kern[name] = dict(avg_us=p["avg_us"], count=p["count"], total_ms=p["ms_total"])
    dom = max(("refactor", "solve", "batch"), key=lambda k: kern.get(k, {}).get("total_ms", 0.0))
    kname = {"batch": "k_batch_dcop"}.get(dom, f"k_{dom}")

    # the same TRAN with every Newton iteration driven from the host
    gh = E.GpuCircuit(circ)
    gh.set_device_newton(False)
    sh = S.GpuSolver(circ.node_count, S.NATIVEKLU)
    gh.tran(sh, store=False, **tran_args)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hres = [gh.tran(sh, store=False, **tran_args) for _ in range(args.host_steps)]
    torch.cuda.synchronize()
    h_el = time.perf_counter() - t0
    host_driven = dict(refactor_solve_per_s=sum(r["factor_calls"] for r in hres) / h_el,
                       ms_per_step=1e3 * h_el / args.host_steps, steps=args.host_steps,
                       same_result=bool(hres[-1]["status"] == res[-1]["status"] and
                                        hres[-1]["dc_iters"] == res[-1]["dc_iters"] and
                                        hres[-1]["factor_calls"] == res[-1]["factor_calls"]))
    roof = dict(bound="hbm", achieved=kern[dom]["gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                frac=kern[dom]["gbs"] / HBM_PEAK_GBS, traffic=_measured_traffic(kname), kernel=kname,
                bytes_per_launch=kern[dom]["bytes"], avg_launch_us=kern[dom]["avg_us"])

    line = dict(
        metric=METRIC, value=value, unit="refactor+solve/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=1e3 * elapsed / args.steps, higher_is_better=True,
        scaling="weak", vs_baseline=None, dtype="f64", data="synthetic",
        config=dict(workload="500-stage level-1 CMOS inverter chain, .tran 5p 10n max step 5p "
                             "(configs[1] circuit); reference semantics: initial DC OP ends "
                             "MaxIters at 100 Newton iterations",
                    solver="NativeKLU semantics, device-resident Newton (one launch per DC OP)",
                    n=n_dc,
                    nnz_unmerged=len(ai_dc), nnz_l=st["nnz_l"], nnz_u=st["nnz_u"],
                    factor_levels=st["n_factor_levels"], lsolve_levels=st["n_lsolve_levels"],
                    usolve_levels=st["n_usolve_levels"]),
        tran_status=res[-1]["status"], dc_iters_per_step=res[-1]["dc_iters"],
        refactor_solve_calls=calls, kernels=kern, roofline=roof, host_driven=host_driven,
    )
    if args.batch > 0:
        line["batched"] = batched_dc_ops(args, E, NL, S, torch)
    if not args.no_adjoint:
        line["adjoint"] = adjoint_cost(E, NL, S, torch)
        line["tran_adjoint"] = tran_adjoint_cost(E, NL, S)
        line["tran_graph"] = tran_graph_cost(E, NL, S)
        line["mesh_refactor_solve"], mesh = mesh_refactor_solve(E, NL, S, torch, np)
        line["kernel_rooflines"] = kernel_rooflines(E, NL, S, torch, np)
    if not args.no_cpu:
        cb, par = cpu_baseline(circ, tran_args, args.cpu_seconds, res[-1])
        xg, rg = g.dc_op(s)
        par["gpu_dc_final_norm"] = rg.final_norm
        par["bit_identical_trajectory"] = (_same(rg.final_norm, par["oracle_dc_final_norm"]) and
                                           rg.iterations == par["oracle_dc_iters"])
        line["cpu_baseline"] = cb
        line["parity"] = par
        if not args.no_adjoint:
            line["mesh_refactor_solve"]["cpu_baseline"] = cpu_baseline_mesh(mesh, np)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
