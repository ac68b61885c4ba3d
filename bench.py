#!/usr/bin/env python3
"""Benchmark: refactor+solve/s of the GPU LinearSolver drop-in on the RC
power-grid transient (1024x1024 grid, n = 1,048,833: segment R U[0.05, 0.5]
ohm, node C U[1, 10] fF, 1.0 V pads every 64 nodes, 4,096 pulsed 10 mA loads;
trapezoidal steps of 1 ps through BbdSolver with 64 partitions).

One step = one trapezoidal timestep of that circuit (crates/sim-core/src/
engine.rs:333-632 with a fixed step): device stamping, then in every Newton
iteration a refactor and a solve of the whole matrix. BbdSolver(64) on this
pattern finds a single block (every node reaches ground through its
capacitor), so it factors the whole matrix with its SparseLU fallback
(bbd_solver.rs:414-504, sparse_lu.rs:440-618); here that is the supernodal
GPU factor, bit-identical to SparseLU. The initial DC operating point runs
through SparseLU directly: BbdSolver's DC pattern has a 225k-node border.

value = refactor+solve calls per second over the timed steps (inputs resident
in device memory; the Newton loop is driven from the host).
roofline = the dominant kernel (k_sn_gather, the descendant contributions of
the supernodal factor): its multiply-subtract flops per launch over its
average launch time from HIP events on its launch stream, against the FP64
peak (the factor does far more flops per compulsory byte than the FP64/HBM
balance of about 10: compute-bound).
cpu_baseline = the oracle's restatement of the reference's column-parallel
factor (native_klu.rs:402-576) and solve on the same matrix, on the host's
cores, checked value for value against the GPU factor.

The earlier workloads (500-stage inverter chain, 4,096 batched Monte-Carlo
DC-OPs, the adjoints, the 64x64 mesh) follow as secondary sections.
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


def _same_arr(a, b):
    """Equal arrays of doubles, NaN == NaN."""
    import numpy as np

    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and bool(np.array_equal(a, b, equal_nan=True))


def _algorithmic_bytes(nnz_a, nnz_l, nnz_u, n):
    """Compulsory traffic per launch (FP64 values 8 B, int32 indices 4 B)."""
    refactor = 12 * nnz_a + 12 * (nnz_l + nnz_u) + 8 * n
    solve = 12 * (nnz_l + nnz_u) + 24 * n
    return refactor, solve


FP64_PEAK_TFLOPS = 78.6


def _log(msg):
    """Progress line on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def _cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def _traffic_from_profile(name, key="traffic_bytes_per_launch:"):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key):
            return float(line.split(":")[1])
    return None


def power_grid_tran(args, E, NL, S, torch, np):
    """The headline: DC operating point, then warm-up and timed trapezoidal
    steps of the power grid through BbdSolver(64), with HIP-event kernel
    timing over the timed steps and the KCL residual of one more step."""
    import scipy.sparse as sp
    from rustspice_amd import _lib

    nx = args.grid
    dt = NL.C3_TRAN["dt"]
    t0 = time.perf_counter()
    c = NL.power_grid(nx, nx)
    gc = E.GpuCircuit(c)
    _log(f"power grid {nx}x{nx}: {c.ndev} devices, {c.node_count} nodes ({time.perf_counter() - t0:.1f} s)")
    dcs = S.GpuSolver(c.node_count, S.SPARSELU)
    t1 = time.perf_counter()
    st_dc, x, it_dc = gc.tran_fixed(dcs, dt, 0)  # DC operating point, no steps
    dc_s = time.perf_counter() - t1
    dc_info = dcs.supernodal_info()
    _log(f"DC operating point (SparseLU): status {st_dc}, {it_dc} Newton iterations, {dc_s:.1f} s")
    dcs.close()
    g = S.GpuSolver(c.node_count, S.BBD)
    g.set_bbd_blocks(64)
    t = 0.0
    t1 = time.perf_counter()
    st_w, x, it_w = gc.tran_steps(g, t, dt, args.warmup, x)
    t += args.warmup * dt
    warm_s = time.perf_counter() - t1
    info = g.supernodal_info()
    bbd = g.bbd_info()
    _log(f"{args.warmup} warm-up steps (first: analysis of the TRAN pattern): status {st_w}, "
         f"{it_w} iterations, {warm_s:.1f} s")
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    t1 = time.perf_counter()
    st, x, iters = gc.tran_steps(g, t, dt, args.steps, x)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t1
    prof = _lib.profile_get()
    _lib.profile_enable(False)
    t += args.steps * dt
    _log(f"{args.steps} timed steps: status {st}, {iters} Newton iterations, {elapsed:.2f} s")
    # KCL residual of one more step: its matrix and right-hand side from the
    # device stamp before the step, its solution from the step itself
    ap, ai, ax, rhs, n = gc.stamp(x, mode=E.MODE_TRAN_TRAP, t=t + dt, dt=dt)
    st_k, xk, it_k = gc.tran_steps(g, t, dt, 1, x)
    A = sp.csc_matrix((ax, ai, ap), shape=(n, n))
    r = A @ xk[:n] - rhs
    scale = np.abs(A).max(axis=1).toarray().ravel() * np.abs(xk[:n]).max() + np.abs(rhs)
    rel = np.abs(r) / np.maximum(scale, 1e-300)
    nc = c.node_count
    v = xk[1:nc] - xk[0]
    kcl = dict(step_status=int(st_k), newton_iters=int(it_k),
               max_rel_residual_nonground_rows=float(rel[1:].max()),
               ground_row_rel_residual=float(rel[0]), ground_potential_x0=float(xk[0]),
               node_v_minus_ground_min=float(v.min()), node_v_minus_ground_max=float(v.max()))
    _log(f"KCL residual (rows other than ground): {kcl['max_rel_residual_nonground_rows']:.2e}")
    calls = int(iters)
    value = calls / elapsed
    pat_sizes = np.zeros(3, np.int64)
    g._L.rs_get_factor_pattern(g.h, pat_sizes, None, None, None, None, None)
    nnz_l, nnz_u = int(pat_sizes[1]), int(pat_sizes[2])
    kern = {}
    for name in ("refactor", "solve", "sn_gather"):
        p = prof[name]
        kern[name] = dict(avg_us=p["avg_us"], count=p["count"], total_ms=p["ms_total"])
    nfac = max(kern["refactor"]["count"], 1)
    g_flops = 2.0 * info["gather_updates"]  # per factor
    f_flops = 2.0 * info["updates"] + nnz_l
    launches = kern["sn_gather"]["count"] / nfac
    gather_tflops = g_flops * nfac / (kern["sn_gather"]["total_ms"] * 1e-3) / 1e12
    kern["sn_gather"].update(flops_per_launch=g_flops / max(launches, 1), launches_per_factor=launches,
                             tflops=gather_tflops)
    kern["refactor"].update(flops=f_flops, tflops=f_flops / (kern["refactor"]["avg_us"] * 1e-6) / 1e12,
                            compulsory_bytes=12 * (len(ai) + nnz_l + nnz_u) + 8 * n)
    kern["refactor"]["flops_per_compulsory_byte"] = f_flops / kern["refactor"]["compulsory_bytes"]
    kern["solve"].update(bytes=12 * (nnz_l + nnz_u) + 24 * n)
    kern["solve"]["gbs"] = kern["solve"]["bytes"] / (kern["solve"]["avg_us"] * 1e-6) / 1e9
    roof = dict(bound="mfma", achieved=gather_tflops, peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                frac=gather_tflops / FP64_PEAK_TFLOPS,
                traffic=_traffic_from_profile("k_sn_gather_power_grid1024_traffic.md"),
                kernel="k_sn_gather", flops_per_launch=kern["sn_gather"]["flops_per_launch"],
                avg_launch_us=kern["sn_gather"]["avg_us"],
                note="FP64 vector ALU (multiply and subtract rounded separately, as the reference "
                     "computes them); peak = MI355X FP64 78.6 TFLOP/s")
    res = dict(value=value, elapsed=elapsed, calls=calls, steps_per_s=args.steps / elapsed,
               status=int(st), kernels=kern, roofline=roof, kcl=kcl,
               dc=dict(status=int(st_dc), newton_iters=int(it_dc), seconds=dc_s,
                       solver="SparseLU (supernodal GPU factor)", supernodes=dc_info["supernodes"]),
               warmup=dict(steps=args.warmup, seconds=warm_s, newton_iters=int(it_w)),
               n=int(n), nnz_unmerged=int(len(ai)), nnz_l=nnz_l, nnz_u=nnz_u,
               supernodal=info, bbd=bbd, devices=int(c.ndev))
    return res, (g, gc, ap, ai, ax, rhs, n)


def cpu_baseline_power_grid(state, budget_s, np):
    """The oracle's column-parallel factor (native_klu.rs:402-576, identity
    pivots = SparseLU's arithmetic) and solve on the grid's TRAN matrix and
    the GPU factor's own pattern. 1 thread: a systematic sample of every m-th
    column, each recomputed from the GPU's L values of the columns before it
    (the per-column work is independent given those), scaled by m. All host
    threads (at most 16): the whole factor. Both checked value for value
    against the GPU factor of the same matrix."""
    from oracle import oracle as O

    g, gc, ap, ai, ax, rhs, n = state
    model, ncpu = _cpu_info()
    nthreads = max(1, min(16, ncpu))
    t0 = time.perf_counter()
    g.factor(ap, ai, ax)  # the GPU factor of exactly this matrix
    pat = g.factor_pattern()
    lx_g, ux_g, ud_g = g.factor_values()
    _log(f"CPU baseline: GPU factor and pattern fetched ({time.perf_counter() - t0:.1f} s)")
    cp = O.ColPar(pat)
    cp.lx[: len(lx_g)] = lx_g
    cp.ux[: len(ux_g)] = ux_g
    cp.ud[:n] = ud_g
    # calibrate the stride on a small sample, then time one that fills ~budget/2
    m = 4096
    cols = np.arange(0, n, m, dtype=np.int64)
    t_cal = cp.columns(ap, ai, ax, cols)
    est_full = t_cal * m
    m = int(max(1, min(4096, est_full / max(budget_s * 0.5, 1e-3))))
    cols = np.arange(m // 2, n, m, dtype=np.int64)
    t_s = cp.columns(ap, ai, ax, cols)
    t1_factor = t_s * m
    lx, ux, ud = cp.values()
    sample_ok = bool(np.array_equal(lx.view(np.int64), lx_g.view(np.int64)) and
                     np.array_equal(ux.view(np.int64), ux_g.view(np.int64)) and
                     np.array_equal(ud.view(np.int64), ud_g.view(np.int64)))
    _log(f"CPU baseline 1 thread: {len(cols)} sampled columns (every {m}-th) in {t_s:.1f} s "
         f"-> {t1_factor:.1f} s per factor")
    t0 = time.perf_counter()
    xs = cp.solve(rhs)
    t_solve = time.perf_counter() - t0
    xg = rhs.copy()
    g.solve(xg)
    res = dict(unit="refactor+solve/s", cores=1, kind="port",
               value=1.0 / (t1_factor + t_solve),
               sample=(f"oracle column-parallel factor restatement, 1 thread: {len(cols)} of {n} columns "
                       f"(every {m}-th), each from the GPU's L values of earlier columns, {t_s:.1f} s, "
                       f"scaled x{m} to {t1_factor:.1f} s per factor; full solve {t_solve:.2f} s"),
               factor_s_est=t1_factor, solve_s=t_solve, sample_values_identical=sample_ok,
               solution_identical_to_gpu=bool(np.array_equal(xs.view(np.int64), xg.view(np.int64))),
               cpu_model=model, nproc=ncpu)
    par = None
    if t1_factor / nthreads < 3.0 * budget_s:
        cp2 = O.ColPar(pat)
        t0 = time.perf_counter()
        cp2.factor(ap, ai, ax, nthreads=nthreads)
        t_par = time.perf_counter() - t0
        lx2, ux2, ud2 = cp2.values()
        same = bool(np.array_equal(lx2.view(np.int64), lx_g.view(np.int64)) and
                    np.array_equal(ux2.view(np.int64), ux_g.view(np.int64)) and
                    np.array_equal(ud2.view(np.int64), ud_g.view(np.int64)))
        x2 = cp2.solve(rhs)
        par = dict(value=1.0 / (t_par + t_solve), unit="refactor+solve/s", cores=nthreads,
                   kind="port", factor_s=t_par, solve_s=t_solve, values_identical=same,
                   solution_identical=bool(np.array_equal(x2.view(np.int64), xs.view(np.int64))),
                   sample=f"whole factor on {nthreads} threads ({t_par:.1f} s) + 1-thread solve")
        _log(f"CPU baseline {nthreads} threads: whole factor {t_par:.1f} s, identical={same}")
    else:
        par = dict(skipped=f"estimated {t1_factor / nthreads:.0f} s on {nthreads} threads")
    res["column_parallel"] = par
    return res


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
    # the same sample through the device-resident plain backend (each DC OP
    # one k_batch_dcop launch of its own)
    gds = [(E.GpuCircuit(circs[i]), S.GpuSolver(circs[i].node_count)) for i in idx]
    for g, s in gds:
        g.dc_op(s)  # warm-up: patterns and plans
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    same_dev = True
    for (g, s), i in zip(gds, idx):
        _, r = g.dc_op(s)
        same_dev &= (r.iterations == out[i][1].iterations and _same(r.final_norm, out[i][1].final_norm))
    torch.cuda.synchronize()
    dev_per = (time.perf_counter() - t0) / k
    return dict(
        workload=f"{args.batch} Monte-Carlo level-1 inverter chains, stages U[50,400], DC OP each "
                 "(reference semantics: NativeKLU, no gmin/source stepping)",
        dc_ops_per_s=dcops, refactor_solve_per_s=float(iters.sum()) * args.batch_steps / wall,
        ms_per_batch=1e3 * wall / args.batch_steps, kernel_ms_per_batch=kern_ms,
        instances=args.batch, patterns=info["patterns"], grid=info["grid"], nmax=info["nmax"],
        converged=conv, newton_iters_total=int(iters.sum()), setup_s=setup_s,
        loop_plain_backend_dc_ops_per_s=1.0 / loop_per, loop_sample=k,
        speedup_vs_loop_plain=dcops * loop_per, loop_results_identical=bool(same),
        loop_device_resident_dc_ops_per_s=1.0 / dev_per,
        speedup_vs_loop_device_resident=dcops * dev_per,
        loop_device_resident_results_identical=bool(same_dev))


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


def tran_adjoint_cost(E, NL, S, nx=256, nsteps=50, every=25):
    """Transient adjoint on the RC mesh driven by level-1 inverters (the
    adjoint workload: nx x nx with nx*nx/32 taps, 2,048 at 256x256):
    dV(centre at t = nsteps*2 ps)/dp for every R, C and MOSFET W with a
    checkpoint every `every` steps, against one plain fixed-step TRAP
    transient of the same steps through SparseLuSolver (adjoint_over_forward;
    at n >= 20,000 both factor supernodally, the plain one in the reference's
    arithmetic) and against the adjoint's own forward pass, which solves every
    step to its exact root (adjoint_over_exact_forward). The first call of
    each (the symbolic analyses of the DC and TRAN patterns and their plans)
    is timed apart as setup."""
    import numpy as np

    c = NL.rc_mesh_inverters(nx, nx)
    probe = NL.rc_mesh_probe(nx, nx)
    dt = NL.C5_TRAN["dt"]
    g = E.GpuCircuit(c)
    s = S.GpuSolver(c.node_count, S.SPARSELU)
    t0 = time.perf_counter()
    st, _, _ = g.tran_fixed(s, dt, 1)  # setup: patterns, analyses, plans
    setup_fwd = time.perf_counter() - t0
    _log(f"  adjoint workload {nx}x{nx}: plain forward setup {setup_fwd:.1f} s")
    t0 = time.perf_counter()
    st, xf, it = g.tran_fixed(s, dt, nsteps)
    t_fwd = time.perf_counter() - t0
    _log(f"  plain forward {nsteps} steps {t_fwd:.1f} s, status {st}, {it} Newton iterations")
    t0 = time.perf_counter()
    g.tran_adjoint(s, dt, 1, probe, every)  # setup of the adjoint's solvers and plans
    setup_adj = time.perf_counter() - t0
    t0 = time.perf_counter()
    dp, xa, info = g.tran_adjoint(s, dt, nsteps, probe, every)
    t_adj = time.perf_counter() - t0
    _log(f"  adjoint {nsteps} steps every {every}: {t_adj:.1f} s, status {info['status']}")
    if info["status"] != 0:
        return dict(workload=f"{nx}x{nx} RC mesh adjoint", adjoint_info=info, forward_status=int(st))
    # every state kept in host memory instead of checkpoints: no recomputation
    t0 = time.perf_counter()
    dp_all, _, info_all = g.tran_adjoint(s, dt, nsteps, probe, 0)
    t_all = time.perf_counter() - t0
    kinds = np.array([d["kind"] for d in c.devices])
    sn = s.supernodal_info()
    return dict(workload=f"{nx}x{nx} RC mesh, {nx * nx // 32} inverter taps, {nsteps} TRAP steps of 2 ps, "
                         f"checkpoint every {every}; dV(centre)/d(every R, C, W)",
                n=int(len(xa)), devices=int(c.ndev), solver="SparseLU (supernodal at n >= 20,000)",
                supernodal=dict(active=int(sn["active"]), supernodes=int(sn["supernodes"]),
                                levels=int(sn["levels"])),
                setup_s=dict(plain_forward=setup_fwd, adjoint=setup_adj),
                forward_status=int(st), forward_ms=1e3 * t_fwd, forward_newton_iters=int(it),
                forward_note=None if st == 0 else
                "the reference-semantics TRAP loop through SparseLuSolver stops (MaxIters): the ground "
                "node's row is kept and its pivot clamped to 1e-14, so the node voltages drift with "
                "the ground potential; the adjoint's own forward (exact roots, no clamped pivot) is "
                "the forward it is measured against",
                adjoint_ms=1e3 * t_adj, adjoint_over_forward=t_adj / t_fwd if st == 0 else None,
                adjoint_over_exact_forward=t_adj / max(info["forward_ms"] * 1e-3, 1e-12), adjoint_info=info,
                keep_all_states=dict(adjoint_ms=1e3 * t_all,
                                     adjoint_over_exact_forward=t_all / max(info_all["forward_ms"] * 1e-3, 1e-12),
                                     states=int(info_all["checkpoints"]),
                                     recomputed_steps=int(info_all["recomputed_steps"]),
                                     same_dp_as_checkpointed=bool(np.array_equal(dp_all, dp))),
                sensitivities=dict(R=int((kinds == NL.RS_DEV_R).sum()), C=int((kinds == NL.RS_DEV_C).sum()),
                                   W=int((kinds == NL.RS_DEV_M).sum())),
                probe_v=float(xa[probe]),
                max_abs_dp=dict(R=float(np.abs(dp[kinds == NL.RS_DEV_R]).max()),
                                C=float(np.abs(dp[kinds == NL.RS_DEV_C]).max()),
                                W=float(np.abs(dp[kinds == NL.RS_DEV_M]).max())))


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


def tran_device_cost(E, NL, S):
    """The reference's adaptive TRAN (rs_tran) host-driven against its loop on
    the device controller after the first accepted step (bit-identical
    results), on a PULSE-driven RC ladder."""
    b = NL._Builder()
    b.add(NL.RS_DEV_V, (1, 0), 0.0, NL.pulse_params(0.0, 1.0, 1e-10, 5e-11, 5e-11, 3e-10, 1e-9),
          flags=NL.RS_F_WAVEFORM)
    for i in range(30):
        b.add(NL.RS_DEV_R, (1 + i, 2 + i), 1e3)
        b.add(NL.RS_DEV_C, (2 + i, 0), 1e-13)
    c = b.build(32, "pulse_ladder30")
    args = dict(tstep=1e-11, tstop=3e-9, tmax=5e-11)
    row = dict(workload="30-section RC ladder, PULSE input, .tran 10p 3n max 50p (NativeKLU semantics)")
    res = {}
    for dev in (False, True):
        best = None
        for _ in range(3):
            g = E.GpuCircuit(c)
            g.set_tran_graph(dev)
            s = S.GpuSolver(c.node_count)
            t0 = time.perf_counter()
            r = g.tran(s, **args)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        res[dev] = r
        row["device_ms" if dev else "host_ms"] = 1e3 * best
        if dev:
            row["device_info"] = g.tran_graph_info()
    row["accepted"], row["rejected"] = res[True]["accepted"], res[True]["rejected"]
    row["newton_iters"] = res[True]["newton_iters_total"]
    row["same_result"] = _same_arr(res[True]["sols"], res[False]["sols"]) and \
        _same_arr(res[True]["times"], res[False]["times"])
    row["speedup"] = row["host_ms"] / row["device_ms"]
    return row


def ac_sweep_cost(NL, np, nodes=200, per_decade=100):
    """An AC sweep (engine.rs:829-970) of a seeded 200-node R/C/V/I/diode/
    CMOS network over 1 Hz .. 10 GHz, all frequencies in one launch, against
    the oracle on one host core (a sample of frequencies, scaled)."""
    from oracle import oracle as O
    from rustspice_amd import ac as A

    circ, mag, ph = NL.ac_network(2024, nodes=nodes, with_nonlinear=True)
    x = np.random.RandomState(7).uniform(0.0, 1.0, circ.node_count + 1)
    freqs = A.frequencies("dec", per_decade, 1.0, 1e10)
    g = A.GpuAc(circ, mag, ph)
    g.run(x, freqs)  # warm-up (allocates the scratch of this sweep)
    t0 = time.perf_counter()
    xg, mg, n, k = g.run(x, freqs)
    wall = time.perf_counter() - t0
    t = g.timing()
    sample = freqs[:: max(1, len(freqs) // 16)]
    t1 = time.perf_counter()
    xo, mo, _ = O.ac(circ, mag, ph, x, sample)
    cpu_s = (time.perf_counter() - t1) * len(freqs) / len(sample)
    idx = np.searchsorted(freqs, sample)
    flops = 8.0 * n ** 3 / 3.0 * len(freqs)
    return dict(workload=f"{nodes}-node R/C/V/I/diode/CMOS network, .ac dec {per_decade} 1 10g "
                         f"({len(freqs)} frequencies), complex MNA n={n}",
                frequencies=int(len(freqs)), solved=int(k), n=int(n),
                gpu_ms=1e3 * wall, kernel_ms=t["ms"], launches=t["launches"],
                frequencies_per_s=len(freqs) / wall, kernel_gflops=flops / (t["ms"] * 1e-3) / 1e9,
                cpu_baseline=dict(frequencies_per_s=len(freqs) / cpu_s, cores=1, kind="port",
                                  sample=f"{len(sample)} of the frequencies through oracle/orc_ac.c, scaled"),
                speedup_vs_one_core=cpu_s / wall,
                sample_bit_identical=_same_arr(xg[idx].real, xo.real) and _same_arr(xg[idx].imag, xo.imag)
                and _same_arr(mg[idx], mo))


def bbd_schur_cost(NL, S, np, width=2048, depth=4, rails=32, reps=20):
    """BbdSolver(2)'s dense Schur product S = B_T - sum_k C_k B_kk^-1 R_k
    (bbd_solver.rs:744-810) on two tap banks of `width` chains joined by a
    bus of 2 x `rails` border rows that tap every edge cell (border 64, the
    dense Schur limit; K = 2 x width coupling columns, C and Y half dense):
    the reference's term order (k_bbd_schur, one cell per thread) against the
    FP64 MFMA tiles (packing + v_mfma_f64_16x16x4f64 split-K tiles + ordered
    chunk sum), HIP events around each launch; x of both against the oracle."""
    from oracle import oracle as O
    from rustspice_amd import _lib

    n, ap, ai, ax = NL.tap_bank_matrix(width, depth, rails)
    b = np.random.RandomState(5).randn(n)
    o = O.Bbd(n, 2)
    o.prepare(n)
    o.factor(ap, ai, ax)
    xo = b.copy()
    o.solve(xo)
    out = {}
    for mode, name in ((0, "reference_order"), (2, "mfma_tiles")):
        g = S.GpuSolver(n, S.BBD)
        g.set_bbd_blocks(2)
        g.set_bbd_mfma(mode)
        g.factor(ap, ai, ax)
        _lib.profile_enable(True)
        t0 = time.perf_counter()
        for _ in range(reps):
            g.factor(ap, ai, ax)
        wall = (time.perf_counter() - t0) / reps
        prof = _lib.profile_get()["bbd_schur"]
        _lib.profile_enable(False)
        info, binfo = g.bbd_mfma_info(), g.bbd_info()
        xg = b.copy()
        g.solve(xg)
        bs, k = binfo["border"], info["k"]
        bsp = (bs + 15) // 16 * 16
        tile_flops = 2.0 * bsp * bsp * ((k + 15) // 16 * 16)
        out[name] = dict(schur_us=prof["avg_us"], launches=prof["count"], factor_ms=1e3 * wall,
                         used_tiles=info["used"],
                         rel_err_vs_oracle=float(np.max(np.abs(xg - xo)) / np.max(np.abs(xo))),
                         bit_identical_to_oracle=_same_arr(xg, xo))
        if info["used"]:
            out[name]["tile_tflops"] = tile_flops / (prof["avg_us"] * 1e-6) / 1e12
            out[name]["roofline"] = dict(bound="mfma", achieved=out[name]["tile_tflops"], peak=78.6,
                                         unit="TFLOP/s", frac=out[name]["tile_tflops"] / 78.6)
    useful = 2.0 * 2 * rails * width * rails  # reference terms with nonzero y, one multiply-subtract each
    out.update(workload=f"two tap banks of {width} chains x {depth} cells joined by 2x{rails} rails "
                        f"(n={n}), BbdSolver(2), border {bs}, K={k} coupling columns, "
                        f"packed C {info['c_density']:.2f} dense",
               n=int(n), border=int(bs), k=int(k), tile_flops=tile_flops, useful_flops=useful,
               speedup=out["reference_order"]["schur_us"] / out["mfma_tiles"]["schur_us"])
    return out


def cpu_baseline_mesh(mesh, np, reps=10):
    """The oracle's refactor + solve of the same mesh matrix on one host core
    and the column-parallel refactor (native_klu.rs:402-576) on all host
    threads, whether the GPU result is bit-identical to them, and how far the
    reference-semantics answer lies from an exact FP64 solve of the matrix
    (ground row and column removed, node voltages against ground)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
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
    model, ncpu = _cpu_info()
    out = dict(cpu_oracle_ms=1e3 * t_cpu, cores=1, kind="port", cpu_model=model, nproc=ncpu,
               bit_identical=bool(np.array_equal(xg.view(np.int64), bb.view(np.int64))))
    blocks = [o.block(b) for b in range(o.stats()["nblocks"])]
    if len(blocks) == 1:  # one BTF block: the column-parallel refactor applies
        blk = blocks[0]
        rp = blk["row_perm"]
        rpi = np.empty_like(rp)
        rpi[rp] = np.arange(len(rp))
        for th in (1, max(1, min(16, ncpu))):
            cp = O.ColPar(blk, row_perm_inv=rpi)
            cp.factor(ap, ai, ax, nthreads=th)
            t0 = time.perf_counter()
            for _ in range(reps):
                cp.factor(ap, ai, ax, nthreads=th)
                xc = cp.solve(rhs)
            out[f"column_parallel_{th}t_ms"] = 1e3 * (time.perf_counter() - t0) / reps
            out[f"column_parallel_{th}t_bit_identical"] = bool(np.array_equal(xc.view(np.int64),
                                                                              xg.view(np.int64)))
    A = sp.csc_matrix((ax, ai, ap), shape=(n, n))
    xe = np.zeros(n)
    xe[1:] = spl.spsolve(A[1:, 1:].tocsc(), rhs[1:])
    v_ref, v_ex = xg[1:] - xg[0], xe[1:]
    out["distance_from_exact"] = dict(
        max_rel_node_voltages=float(np.abs(v_ref - v_ex).max() / np.abs(v_ex).max()),
        max_rel_unreferenced=float(np.abs(xg[1:] - v_ex).max() / np.abs(v_ex).max()),
        ground_potential_x0=float(xg[0]),
        note="exact: scipy spsolve with the ground row and column removed")
    return out


def chain_tran(args, E, NL, S, torch, np):
    """The 500-stage level-1 inverter chain, .tran 5p 10n max step 5p, through
    the NativeKLU-semantics GPU backend. Under the reference's semantics the
    TRAN's initial DC operating point ends MaxIters after 100 Newton
    iterations (the level-1 PMOS frame of bsim/evaluate.rs:298-321 gives the
    chain no DC solution), so a run is 100 stamp+refactor+solve iterations
    on the n=504 DC matrix, device-resident in one k_batch_dcop launch."""
    from rustspice_amd import _lib

    circ = NL.inverter_chain(500)
    tran_args = dict(NL.C2_TRAN)
    g = E.GpuCircuit(circ)
    s = S.GpuSolver(circ.node_count, S.NATIVEKLU)
    for _ in range(2):
        g.tran(s, store=False, **tran_args)
    torch.cuda.synchronize()
    _lib.profile_enable(True)
    t0 = time.perf_counter()
    res = [g.tran(s, store=False, **tran_args) for _ in range(args.chain_runs)]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = _lib.profile_get()
    _lib.profile_enable(False)
    calls = sum(r["factor_calls"] for r in res)
    ap_dc, ai_dc = g.pattern(E.MODE_DC)
    n_dc = len(ap_dc) - 1
    _, _, ax0, _, _ = g.stamp(np.zeros(circ.node_count), mode=E.MODE_DC)
    ps = S.HostFactor(n_dc, ap_dc, ai_dc, ax0).plan_stats()
    b_ref, b_sol = _algorithmic_bytes(len(ai_dc), ps["nnz_l"], ps["nnz_u"], n_dc)
    kern = {}
    b_batch = (b_ref + b_sol) * res[-1]["dc_iters"]  # dc_iters refactor+solve per launch
    for name, b in (("refactor", b_ref), ("solve", b_sol), ("batch", b_batch)):
        p = prof[name]
        if p["count"]:
            kern[name] = dict(avg_us=p["avg_us"], count=p["count"], bytes=b,
                              gbs=b / (p["avg_us"] * 1e-6) / 1e9, total_ms=p["ms_total"])
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    kname = {"batch": "k_batch_dcop"}.get(dom, f"k_{dom}")
    gh = E.GpuCircuit(circ)
    gh.set_device_newton(False)
    sh = S.GpuSolver(circ.node_count, S.NATIVEKLU)
    gh.tran(sh, store=False, **tran_args)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hres = [gh.tran(sh, store=False, **tran_args) for _ in range(args.host_steps)]
    torch.cuda.synchronize()
    h_el = time.perf_counter() - t0
    out = dict(
        workload="500-stage level-1 CMOS inverter chain, .tran 5p 10n max step 5p; reference "
                 "semantics: the initial DC OP ends MaxIters at 100 Newton iterations",
        solver="NativeKLU semantics, device-resident Newton (one launch per DC OP)",
        refactor_solve_per_s=calls / elapsed, ms_per_run=1e3 * elapsed / args.chain_runs,
        runs=args.chain_runs, n=n_dc, nnz_unmerged=len(ai_dc), nnz_l=ps["nnz_l"], nnz_u=ps["nnz_u"],
        tran_status=res[-1]["status"], dc_iters_per_run=res[-1]["dc_iters"], kernels=kern,
        roofline=dict(bound="hbm", achieved=kern[dom]["gbs"], peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=kern[dom]["gbs"] / HBM_PEAK_GBS, traffic=_measured_traffic(kname),
                      kernel=kname, bytes_per_launch=kern[dom]["bytes"], avg_launch_us=kern[dom]["avg_us"]),
        host_driven=dict(refactor_solve_per_s=sum(r["factor_calls"] for r in hres) / h_el,
                         ms_per_run=1e3 * h_el / args.host_steps, runs=args.host_steps,
                         same_result=bool(hres[-1]["status"] == res[-1]["status"] and
                                          hres[-1]["dc_iters"] == res[-1]["dc_iters"] and
                                          hres[-1]["factor_calls"] == res[-1]["factor_calls"])))
    if not args.no_cpu:
        cb, par = cpu_baseline(circ, tran_args, min(args.cpu_seconds, 10.0), res[-1])
        xg, rg = g.dc_op(s)
        par["gpu_dc_final_norm"] = rg.final_norm
        par["bit_identical_trajectory"] = (_same(rg.final_norm, par["oracle_dc_final_norm"]) and
                                           rg.iterations == par["oracle_dc_iters"])
        out["cpu_baseline"] = cb
        out["parity"] = par
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40, help="timed trapezoidal steps of the power grid")
    ap.add_argument("--warmup", type=int, default=2, help="untimed steps (the first analyzes the pattern)")
    ap.add_argument("--grid", type=int, default=1024, help="power grid side (1024: the full workload)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--headline-only", action="store_true", help="skip the secondary sections")
    ap.add_argument("--chain-runs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=4096, help="Monte-Carlo instances (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--loop-sample", type=int, default=16)
    ap.add_argument("--host-steps", type=int, default=3)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-grid", type=int, default=256, help="RC mesh side of the transient adjoint")
    ap.add_argument("--adjoint-steps", type=int, default=50,
                    help="TRAP steps of the transient adjoint (500: the workload's 1 ns)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the transient adjoint section (with --adjoint-steps 500: the full workload)")
    ap.add_argument("--bbd-only", action="store_true",
                    help="only the BbdSolver dense Schur product (reference order against MFMA tiles)")
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

    from rustspice_amd import engine as E
    from rustspice_amd import netlists as NL
    from rustspice_amd import solver as S

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.kernels_only:
        print(json.dumps({"kernel_rooflines": kernel_rooflines(E, NL, S, torch, np)}))
        return
    if args.bbd_only:
        print(json.dumps({"bbd_schur": bbd_schur_cost(NL, S, np)}))
        return
    if args.adjoint_only:
        print(json.dumps({"tran_adjoint": tran_adjoint_cost(E, NL, S, nx=args.adjoint_grid,
                                                            nsteps=args.adjoint_steps)}))
        return
    hl, state = power_grid_tran(args, E, NL, S, torch, np)
    line = dict(
        metric=METRIC, value=hl["value"], unit="refactor+solve/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=1e3 * hl["elapsed"] / args.steps, higher_is_better=True,
        scaling="weak", vs_baseline=None, dtype="f64", data="synthetic",
        config=dict(workload=f"RC power-grid transient, {args.grid}x{args.grid} grid (n={hl['n']}), "
                             "segment R U[0.05,0.5] ohm, node C U[1,10] fF, 1.0 V pads every 64 nodes, "
                             f"{args.grid * args.grid // 256} pulsed 10 mA loads; trapezoidal steps of 1 ps "
                             "through BbdSolver(64) (whole-matrix SparseLU fallback); DC OP via SparseLU",
                    solver="BbdSolver(64) -> SparseLU semantics, supernodal GPU factor and solve",
                    n=hl["n"], nnz_unmerged=hl["nnz_unmerged"], nnz_l=hl["nnz_l"], nnz_u=hl["nnz_u"],
                    devices=hl["devices"], supernodes=hl["supernodal"]["supernodes"],
                    levels=hl["supernodal"]["levels"]),
        timesteps_per_s=hl["steps_per_s"], newton_iters=hl["calls"], tran_status=hl["status"],
        kernels=hl["kernels"], roofline=hl["roofline"], kcl=hl["kcl"], dc_op=hl["dc"],
        warmup_detail=hl["warmup"], supernodal=hl["supernodal"], bbd=hl["bbd"])
    if not args.no_cpu:
        cb = cpu_baseline_power_grid(state, args.cpu_seconds, np)
        line["cpu_baseline"] = cb
    del state
    if not args.headline_only:
        _log("secondary: 500-stage inverter chain")
        line["chain500"] = chain_tran(args, E, NL, S, torch, np)
        if args.batch > 0:
            _log("secondary: batched Monte-Carlo DC-OPs")
            line["batched"] = batched_dc_ops(args, E, NL, S, torch)
        if not args.no_adjoint:
            _log("secondary: adjoints, graphs, mesh refactor+solve")
            line["adjoint"] = adjoint_cost(E, NL, S, torch)
            line["tran_adjoint"] = tran_adjoint_cost(E, NL, S, nx=args.adjoint_grid, nsteps=args.adjoint_steps)
            line["tran_graph"] = tran_graph_cost(E, NL, S)
            line["tran_device_controller"] = tran_device_cost(E, NL, S)
            line["ac_sweep"] = ac_sweep_cost(NL, np)
            line["bbd_schur"] = bbd_schur_cost(NL, S, np)
            line["mesh_refactor_solve"], mesh = mesh_refactor_solve(E, NL, S, torch, np)
            line["kernel_rooflines"] = kernel_rooflines(E, NL, S, torch, np)
            if not args.no_cpu:
                line["mesh_refactor_solve"]["cpu_baseline"] = cpu_baseline_mesh(mesh, np)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
