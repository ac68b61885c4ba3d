"""Transient sensitivities by the discrete adjoint (rs_tran_adjoint; no
reference counterpart) against the dense FP64 CPU adjoint and central
finite differences of the exact-root forward (tests/tran_adjoint_ref.py,
itself pinned by its own finite differences on the CPU).

The workload is the RC mesh driven by level-1 inverters at seeded taps; the
checked instances give each tap a load to ground, so every MOSFET stays off
the level-1 kink at vds = 0 where a central-difference Jacobian is not the
model's derivative.
"""
import numpy as np
import pytest

from rustspice_amd import engine as E
from rustspice_amd import netlists as NL
from rustspice_amd import solver as S
from tests import tran_adjoint_ref as R

pytestmark = pytest.mark.gpu

DT = 2e-12


def _gpu(c, nsteps, probe, every, sn=1):
    g = E.GpuCircuit(c)
    s = S.GpuSolver(c.node_count)
    s.set_supernodal(sn)
    dp, x, info = g.tran_adjoint(s, DT, nsteps, probe, every)
    assert info["status"] == 0 and info["inexact_steps"] == 0, info
    return dp, x, info


def test_forward_is_the_exact_root_trajectory(gpu):
    c = NL.rc_mesh_inverters(4, 4, 2, load=1e4)
    probe = NL.rc_mesh_probe(4, 4)
    _, x, info = _gpu(c, 12, probe, 5)
    xs, _ = R.forward(c, DT, 12)
    assert np.abs(x[: c.node_count] - xs[-1][: c.node_count]).max() <= 1e-12
    assert info["checkpoints"] == 3 and info["recomputed_steps"] == 12
    assert info["adjoint_solves"] == 13 and info["polish_iters"] > 0


def test_matches_dense_cpu_adjoint(gpu):
    c = NL.rc_mesh_inverters(4, 4, 2, load=1e4)
    probe = NL.rc_mesh_probe(4, 4)
    N = 12
    dp, _, _ = _gpu(c, N, probe, 5)
    params = [(k, *R.param(c, k)) for k in range(c.ndev)
              if c.devices[k]["kind"] in (NL.RS_DEV_R, NL.RS_DEV_C, NL.RS_DEV_M)]
    ref, _, _ = R.adjoint(c, DT, N, probe, params)
    for kind in (NL.RS_DEV_R, NL.RS_DEV_C, NL.RS_DEV_M):
        idx = [i for i, (k, *_) in enumerate(params) if c.devices[k]["kind"] == kind]
        want = ref[idx]
        got = dp[[params[i][0] for i in idx]]
        scale = np.abs(want).max()
        assert scale > 0.0
        # each sensitivity within 1e-6 of itself or of the kind's largest (the
        # attainable accuracy of lambda is eps * cond(J), about 1e-7 here)
        assert np.allclose(got, want, rtol=1e-6, atol=1e-6 * scale), (kind, got, want)


def test_matches_central_differences(gpu):
    """h = 1e-6 relative on seeded parameters whose sensitivity is resolved
    (well above the 1e-12 / h round-off floor of the difference quotient)."""
    c = NL.rc_mesh_inverters(4, 4, 2, load=1e4)
    probe = NL.rc_mesh_probe(4, 4)
    N = 12
    dp, _, _ = _gpu(c, N, probe, 4)
    rng = np.random.RandomState(5)
    cand = [k for k in range(c.ndev) if c.devices[k]["kind"] in (NL.RS_DEV_R, NL.RS_DEV_C, NL.RS_DEV_M)]
    checked = 0
    for k in rng.permutation(cand)[:24]:
        p0 = R.param(c, k)[0]()
        fd = R.finite_difference(c, DT, N, probe, int(k), h_rel=1e-6)
        floor = 1e-12 / (1e-6 * abs(p0))
        if abs(fd) < 1e3 * floor:
            continue
        assert abs(dp[k] - fd) <= 1e-6 * abs(fd) + floor, (int(k), dp[k], fd)
        checked += 1
    assert checked >= 6


def test_checkpoint_interval_does_not_change_the_gradient(gpu):
    c = NL.rc_mesh_inverters(4, 4, 2, load=1e4)
    probe = NL.rc_mesh_probe(4, 4)
    runs = [_gpu(c, 10, probe, every) for every in (1, 3, 1000, 0)]
    # the recomputed segments reproduce the forward states bit for bit, and
    # so the gradient is bit-identical too
    for dp, x, info in runs[1:]:
        assert np.array_equal(x, runs[0][1])
        assert np.array_equal(dp, runs[0][0])
    # every = 0 keeps all 11 states and recomputes nothing
    assert [r[2]["checkpoints"] for r in runs] == [10, 4, 1, 11]
    assert [r[2]["recomputed_steps"] for r in runs] == [10, 10, 10, 0]


def test_unloaded_instance_matches_differences_of_the_gpu_forward(gpu):
    """The workload's own construction (no tap loads: the DC OP holds the
    PMOS at vds = 0, where the model's one-sided derivative is the one the
    solution follows): central differences (h = 1e-6 relative) of the GPU's
    exact-root forward itself, on an 8x8 instance over 20 steps."""
    c = NL.rc_mesh_inverters(8, 8, 4)
    probe = NL.rc_mesh_probe(8, 8)
    N = 20
    dp, _, _ = _gpu(c, N, probe, 7)
    ks = [k for k in range(c.ndev) if c.devices[k]["kind"] == NL.RS_DEV_M][:4]
    ks += [k for k in range(c.ndev) if c.devices[k]["kind"] == NL.RS_DEV_C][30:33]
    ks += [0]  # Vdd
    checked = 0
    for k in ks:
        get, put = R.param(c, k)
        p0 = get()
        h = 1e-6 * abs(p0)
        put(p0 + h)
        xp = _gpu(c, N, probe, N)[1][probe]
        put(p0 - h)
        xm = _gpu(c, N, probe, N)[1][probe]
        put(p0)
        fd = (xp - xm) / (2.0 * h)
        floor = 1e-12 / h
        if abs(fd) < 1e3 * floor:
            continue
        assert abs(dp[k] - fd) <= 1e-6 * abs(fd) + floor, (k, dp[k], fd)
        checked += 1
    assert checked >= 4


def test_no_value_is_read_before_it_is_written(gpu):
    """With every fresh FP64 device buffer filled with NaN, the forward state
    and the gradient are bit-identical to a run on ordinary buffers. (The
    exact-Newton residual of the first TRAP step once read x from a buffer
    that the TRAN pattern's allocation had just replaced, so its result
    depended on what the allocator's memory held.)"""
    from rustspice_amd import _lib
    c = NL.rc_mesh_inverters(4, 4, 2, load=1e4)
    probe = NL.rc_mesh_probe(4, 4)
    base = [_gpu(c, 10, probe, every) for every in (3, 0)]
    _lib.set_poison(True)
    try:
        pois = [_gpu(c, 10, probe, every) for every in (3, 0)]
    finally:
        _lib.set_poison(False)
    for (dp, x, _), (dpp, xp, _) in zip(base, pois):
        assert np.array_equal(x, xp) and np.array_equal(dp, dpp)


@pytest.mark.parametrize("nx,taps,load,N", [(4, 2, 1e4, 12), (8, 4, None, 20)])
def test_supernodal_factors_agree_with_the_level_kernels(gpu, nx, taps, load, N):
    """With the adjoint's solvers on the supernodal path (forced at this size;
    the default above 20,000 unknowns), the exact-root forward is the same
    SparseLU arithmetic, so its states are bit-identical; the backward factors
    J^T on its own instead of solving with J's factor transposed, so the
    gradient agrees to round-off. The checkpoint interval still leaves the
    gradient bit-identical."""
    c = NL.rc_mesh_inverters(nx, nx, taps, load=load)
    probe = NL.rc_mesh_probe(nx, nx)
    dp0, x0, _ = _gpu(c, N, probe, 5, sn=0)
    runs = [_gpu(c, N, probe, every, sn=2) for every in (5, 3, 0)]
    dp, x, info = runs[0]
    assert np.array_equal(x, x0)
    for kind in (NL.RS_DEV_R, NL.RS_DEV_C, NL.RS_DEV_M):
        idx = np.array([k for k in range(c.ndev) if c.devices[k]["kind"] == kind])
        scale = np.abs(dp0[idx]).max()
        assert scale > 0.0
        assert np.allclose(dp[idx], dp0[idx], rtol=1e-9, atol=1e-9 * scale), kind
    for dpe, xe, _ in runs[1:]:
        assert np.array_equal(xe, x) and np.array_equal(dpe, dp)


def test_supernodal_adjoint_matches_dense_cpu_adjoint(gpu):
    c = NL.rc_mesh_inverters(4, 4, 2, load=1e4)
    probe = NL.rc_mesh_probe(4, 4)
    N = 12
    dp, _, _ = _gpu(c, N, probe, 5, sn=2)
    params = [(k, *R.param(c, k)) for k in range(c.ndev)
              if c.devices[k]["kind"] in (NL.RS_DEV_R, NL.RS_DEV_C, NL.RS_DEV_M)]
    ref, _, _ = R.adjoint(c, DT, N, probe, params)
    for kind in (NL.RS_DEV_R, NL.RS_DEV_C, NL.RS_DEV_M):
        idx = [i for i, (k, *_) in enumerate(params) if c.devices[k]["kind"] == kind]
        want = ref[idx]
        got = dp[[params[i][0] for i in idx]]
        scale = np.abs(want).max()
        assert np.allclose(got, want, rtol=1e-6, atol=1e-6 * scale), (kind, got, want)


def test_full_size_workload_matches_differences_of_the_gpu_forward(gpu):
    """The adjoint workload at its full size (256x256 RC mesh, 2,048 inverter
    taps, no loads, n = 65,541): the adjoint's solvers factor supernodally
    (n >= 20,000) on the symmetric closure of the MOSFET-unsymmetric pattern,
    J^T on a plan of its own. Over 25 steps (50 ps, one checkpoint segment),
    central differences (h = 1e-6 relative) of the GPU's exact-root forward,
    each parameter updated in place (rs_circuit_set_device), on seeded R, C
    and W next to the probe, where the sensitivities stand well above the
    difference quotient's round-off floor; checkpoints every 5 steps leave
    the gradient bit-identical."""
    nx = 256
    c = NL.rc_mesh_inverters(nx, nx)
    probe = NL.rc_mesh_probe(nx, nx)
    N = 25
    g = E.GpuCircuit(c)
    s = S.GpuSolver(c.node_count, S.SPARSELU)  # its DC OP starts the forward (NativeKLU: no plan at this size)
    dp, x, info = g.tran_adjoint(s, DT, N, probe, 0)
    print(f"full-size adjoint: n={len(x)}, forward {info['forward_ms']:.0f} ms", flush=True)
    assert info["status"] == 0 and info["inexact_steps"] == 0, info
    dp1, x1, _ = g.tran_adjoint(s, DT, N, probe, 5)
    assert np.array_equal(x1, x) and np.array_equal(dp1, dp)

    def dist(k):
        nodes = [int(v) - 3 for v in c.devices[k]["node"][:2] if int(v) >= 3]
        pi, pj = divmod(probe - 3, nx)
        return min((abs(v // nx - pi) + abs(v % nx - pj) for v in nodes), default=10 ** 9)

    rng = np.random.RandomState(7)
    picks = []
    for kind, radius, count in ((NL.RS_DEV_R, 1, 2), (NL.RS_DEV_C, 1, 3), (NL.RS_DEV_M, 12, 3)):
        cand = [k for k in range(c.ndev) if c.devices[k]["kind"] == kind and dist(k) <= radius]
        assert len(cand) >= count
        picks += [int(k) for k in rng.choice(cand, count, replace=False)]
    checked = 0
    for k in picks:
        get, put = R.param(c, k)
        p0 = get()
        h = 1e-6 * abs(p0)
        vals = []
        for p in (p0 + h, p0 - h):
            put(p)
            g.set_device(k)
            dpx, xx, inf = g.tran_adjoint(s, DT, N, probe, 0)
            assert inf["status"] == 0
            vals.append(xx[probe])
        put(p0)
        g.set_device(k)
        fd = (vals[0] - vals[1]) / (2.0 * h)
        print(f"device {k}: adjoint {dp[k]:.6e} difference {fd:.6e}", flush=True)
        # the exact-root forward stops once |dx| <= 1e-12 + 1e-10 |x| and its
        # chord steps contract by 10x at least, so two perturbed forwards
        # carry a few 1e-12 V of solver noise at this size: the quotient's floor
        floor = 5e-12 / h
        assert abs(dp[k] - fd) <= 1e-6 * abs(fd) + floor, (k, dp[k], fd)
        checked += abs(fd) >= 3.0 * floor
    assert checked >= 4
    # the restored parameters give the first gradient back
    dp2, x2, _ = g.tran_adjoint(s, DT, N, probe, 0)
    assert np.array_equal(x2, x) and np.array_equal(dp2, dp)


def test_32x32_instance_matches_dense_cpu_adjoint(gpu):
    """The adjoint workload's 32x32 instance (n = 1,031, 32 taps, each with a
    load to ground so that no MOSFET sits on the level-1 kink, where a
    central-difference Jacobian is not the model's derivative) over 50 steps
    with a checkpoint every 25, against the dense FP64 CPU adjoint (colour
    group Jacobians) on 64 seeded parameters: 24 R, 24 C and 16 W."""
    nx, N = 32, 50
    c = NL.rc_mesh_inverters(nx, nx, load=1e4)
    probe = NL.rc_mesh_probe(nx, nx)
    dp, _, _ = _gpu(c, N, probe, 25)
    rng = np.random.RandomState(3)
    ks = []
    for kind, count in ((NL.RS_DEV_R, 24), (NL.RS_DEV_C, 24), (NL.RS_DEV_M, 16)):
        cand = [k for k in range(c.ndev) if c.devices[k]["kind"] == kind]
        ks += [int(k) for k in rng.choice(cand, count, replace=False)]
    ks.sort()
    ref, _, _ = R.adjoint(c, DT, N, probe, [(k, *R.param(c, k)) for k in ks], colored=True)
    kinds = np.array([c.devices[k]["kind"] for k in ks])
    for kind in (NL.RS_DEV_R, NL.RS_DEV_C, NL.RS_DEV_M):
        sel = kinds == kind
        assert sel.sum() >= 1
        want, got = ref[sel], dp[np.array(ks)[sel]]
        # the kind's largest over all devices sets the absolute scale
        scale = np.abs(dp[[k for k in range(c.ndev) if c.devices[k]["kind"] == kind]]).max()
        assert np.allclose(got, want, rtol=1e-6, atol=1e-6 * scale), (kind, got, want)
