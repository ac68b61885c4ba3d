"""Circuit-level front end: device stamping and the DC/TRAN Newton drivers.

Mirrors ``Engine`` (crates/sim-core/src/engine.rs) for the hot path:
``dc_op`` is ``run_dc_result`` (engine.rs:249-311), ``tran`` is
``run_tran_result_with_params`` (engine.rs:333-632), ``stamp`` is one call of
the engine's MNA build closure (engine.rs:262-283). The solver argument is a
:class:`rustspice_amd.solver.GpuSolver` (the LinearSolver drop-in).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .solver import GpuSolver, SolverError

MODE_DC, MODE_TRAN_BE, MODE_TRAN_TRAP = 0, 1, 2


class NewtonOpts(C.Structure):
    _fields_ = [
        ("max_iters", C.c_int32), ("gmin_steps", C.c_int32), ("source_steps", C.c_int32),
        ("pad", C.c_int32), ("abs_tol", C.c_double), ("rel_tol", C.c_double),
        ("gmin", C.c_double), ("damping", C.c_double), ("damping_min", C.c_double),
    ]


def default_opts(**kw) -> NewtonOpts:
    """newton_config_from_options (engine.rs:987-999) over the option defaults
    of options.rs:79-145: itl1=100, vntol=1e-6, reltol=1e-3, gmin=1e-12."""
    o = NewtonOpts(100, 0, 0, 0, 1e-6, 1e-3, 1e-12, 1.0, 0.1)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class TranAdjointInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("checkpoints", C.c_int32), ("max_refine", C.c_int32),
                ("polish_iters", C.c_int32), ("forward_iters", C.c_int64), ("recompute_iters", C.c_int64),
                ("recomputed_steps", C.c_int64), ("adjoint_solves", C.c_int64),
                ("inexact_steps", C.c_int64), ("forward_ms", C.c_double), ("recompute_ms", C.c_double),
                ("backward_ms", C.c_double)]


class NewtonResult(C.Structure):
    _fields_ = [("converged", C.c_int32), ("iterations", C.c_int32), ("reason", C.c_int32),
                ("pad", C.c_int32), ("final_norm", C.c_double)]


class TranResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("accepted", C.c_int32), ("rejected", C.c_int32),
                ("newton_iters_total", C.c_int32), ("factor_calls", C.c_int32),
                ("dc_iters", C.c_int32), ("npoints", C.c_int64)]


class SweepResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("factor_calls", C.c_int32),
                ("pad", C.c_int32), ("npoints", C.c_int64), ("nsolved", C.c_int64),
                ("fail_value", C.c_double)]


def _bind():
    L = _lib.lib()
    if getattr(L, "_engine_bound", False):
        return L
    vp = C.c_void_p
    L.rs_circuit_create.restype = vp
    L.rs_circuit_create.argtypes = [vp, C.c_int64, C.c_int64]
    L.rs_circuit_destroy.argtypes = [vp]
    L.rs_circuit_last_error.restype = C.c_int32
    L.rs_circuit_last_error.argtypes = [vp, C.c_char_p, C.c_int64]
    L.rs_circuit_pattern.restype = C.c_int32
    L.rs_circuit_pattern.argtypes = [vp, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                     vp, vp]
    L.rs_circuit_set_initial_guess.restype = C.c_int32
    L.rs_circuit_set_initial_guess.argtypes = [vp, vp, C.c_int64]
    L.rs_circuit_set_device.restype = C.c_int32
    L.rs_circuit_set_device.argtypes = [vp, C.c_int64, vp]
    L.rs_circuit_set_cap_state.restype = C.c_int32
    L.rs_circuit_set_cap_state.argtypes = [vp] * 5
    L.rs_stamp_device.restype = C.c_int32
    L.rs_stamp_device.argtypes = [vp, C.c_int32, vp, C.c_int64, C.c_double, C.c_double, C.c_double,
                                  C.c_double, vp, C.c_int64, vp, C.c_int64, C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64), vp]
    L.rs_stamp.restype = C.c_int32
    L.rs_stamp.argtypes = [vp, C.c_int32, vp, C.c_int64, C.c_double, C.c_double, C.c_double,
                           C.c_double, vp, C.c_int64, vp, C.c_int64, C.POINTER(C.c_int64),
                           C.POINTER(C.c_int64)]
    L.rs_dc_op.restype = C.c_int32
    L.rs_dc_op.argtypes = [vp, vp, C.POINTER(NewtonOpts), vp, C.c_int64, C.POINTER(C.c_int64),
                           C.POINTER(NewtonResult)]
    L.rs_tran.restype = C.c_int32
    L.rs_tran.argtypes = [vp, vp, C.POINTER(NewtonOpts), C.c_double, C.c_double, C.c_double,
                          C.c_double, C.c_int64, vp, vp, C.c_int64, C.POINTER(TranResult)]
    L.rs_tran_fixed.restype = C.c_int32
    L.rs_tran_fixed.argtypes = [vp, vp, C.POINTER(NewtonOpts), C.c_double, C.c_int64, vp,
                                C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                C.POINTER(C.c_int32)]
    L.rs_tran_steps.restype = C.c_int32
    L.rs_tran_steps.argtypes = [vp, vp, C.POINTER(NewtonOpts), C.c_double, C.c_double, C.c_int64, vp,
                                C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.rs_circuit_set_device_newton.argtypes = [vp, C.c_int32]
    L.rs_circuit_set_bsim_models.restype = C.c_int32
    L.rs_circuit_set_bsim_models.argtypes = [vp, vp, C.c_int64]
    L.rs_mos_eval.restype = C.c_int32
    L.rs_mos_eval.argtypes = [vp, C.c_int64, vp, vp, vp, C.c_int64, vp]
    L.rs_dc_sweep_points.restype = C.c_int64
    L.rs_dc_sweep_points.argtypes = [C.c_double, C.c_double, C.c_double, vp, C.c_int64]
    L.rs_dc_sweep.restype = C.c_int32
    L.rs_dc_sweep.argtypes = [vp, vp, C.POINTER(NewtonOpts), C.c_int64, C.c_double, C.c_double,
                              C.c_double, C.c_int64, vp, vp, C.c_int64, C.POINTER(SweepResult)]
    L.rs_circuit_set_tran_graph.restype = None
    L.rs_circuit_set_tran_graph.argtypes = [vp, C.c_int32]
    L.rs_circuit_graph_steps.restype = C.c_int64
    L.rs_circuit_graph_steps.argtypes = [vp]
    L.rs_circuit_tran_graph_info.restype = C.c_int32
    L.rs_circuit_tran_graph_info.argtypes = [vp, vp]
    L.rs_tran_adjoint.restype = C.c_int32
    L.rs_tran_adjoint.argtypes = [vp, vp, C.POINTER(NewtonOpts), C.c_double, C.c_int64, C.c_int64,
                                  C.c_int64, vp, vp, C.c_int64, C.POINTER(C.c_int64),
                                  C.POINTER(TranAdjointInfo)]
    L.rs_dc_adjoint.restype = C.c_int32
    L.rs_dc_adjoint.argtypes = [vp, vp, C.c_int64, C.c_int64, C.c_double, vp, vp]
    L.rs_batch_create.restype = vp
    L.rs_batch_create.argtypes = [vp, vp, vp, C.c_int64, C.c_int32]
    L.rs_batch_destroy.argtypes = [vp]
    L.rs_batch_last_error.restype = C.c_int32
    L.rs_batch_last_error.argtypes = [vp, C.c_char_p, C.c_int64]
    L.rs_batch_info.restype = C.c_int32
    L.rs_batch_info.argtypes = [vp, vp]
    L.rs_batch_dc_op.restype = C.c_int32
    L.rs_batch_dc_op.argtypes = [vp, C.POINTER(NewtonOpts), vp]
    L.rs_batch_phase_profile.restype = C.c_int32
    L.rs_batch_phase_profile.argtypes = [vp, C.c_int32]
    L.rs_batch_phase_cycles.restype = C.c_int32
    L.rs_batch_phase_cycles.argtypes = [vp, vp]
    L.rs_batch_results.restype = C.c_int32
    L.rs_batch_results.argtypes = [vp, vp, vp, vp]
    L._engine_bound = True
    return L


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class GpuCircuit:
    """A device-resident instance list (include/rs_circuit.h records)."""

    def __init__(self, circuit):
        self._L = _bind()
        self.circuit = circuit
        self._dev = np.ascontiguousarray(circuit.devices)
        self.h = self._L.rs_circuit_create(_ptr(self._dev), circuit.ndev, circuit.node_count)
        self._check(0)
        models = getattr(circuit, "models", None)
        if models is not None and len(models):
            self._models = np.ascontiguousarray(models)
            self._check(self._L.rs_circuit_set_bsim_models(self.h, _ptr(self._models),
                                                           len(self._models)))

    def set_device(self, k: int):
        """rs_circuit_set_device: uploads device k's record from the host
        circuit again after its values changed (same kind, nodes and flags)."""
        rec = np.ascontiguousarray(self.circuit.devices[k: k + 1])
        self._check(self._L.rs_circuit_set_device(self.h, int(k), _ptr(rec)))

    def close(self):
        if getattr(self, "h", None):
            self._L.rs_circuit_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        buf = C.create_string_buffer(512)
        self._L.rs_circuit_last_error(self.h, buf, 512)
        if rc != 0 or (buf.value and rc == 0 and b"hipMalloc" in buf.value):
            raise SolverError(rc or 7, buf.value.decode(errors="replace"))

    def dc_adjoint(self, x, probe: int, gmin: float = 1e-12):
        """dx[probe]/dp for every device at the DC solution x (adjoint method):
        returns (per-device sensitivities, lambda)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        ap, _ = self.pattern(MODE_DC)
        n = len(ap) - 1
        dp = np.zeros(max(self.circuit.ndev, 1))
        lam = np.zeros(n)
        self._check(self._L.rs_dc_adjoint(self.h, _ptr(x), len(x), probe, gmin, _ptr(dp), _ptr(lam)))
        return dp[: self.circuit.ndev], lam

    def set_device_newton(self, on: bool):
        """True: a fresh DC OP runs as one device launch (same results)."""
        self._L.rs_circuit_set_device_newton(self.h, int(on))

    def pattern(self, mode=MODE_DC):
        n, nnz = C.c_int64(), C.c_int64()
        self._check(self._L.rs_circuit_pattern(self.h, mode, C.byref(n), C.byref(nnz), None, None))
        ap = np.zeros(n.value + 1, np.int64)
        ai = np.zeros(max(nnz.value, 1), np.int64)
        self._check(self._L.rs_circuit_pattern(self.h, mode, None, None, _ptr(ap), _ptr(ai)))
        return ap, ai[: nnz.value]

    def set_cap_state(self, cap_v, cap_i, has_v, has_i):
        arrs = [np.ascontiguousarray(cap_v, np.float64), np.ascontiguousarray(cap_i, np.float64),
                np.ascontiguousarray(has_v, np.uint8), np.ascontiguousarray(has_i, np.uint8)]
        for a in arrs:  # the C entry point reads ndev entries of each
            if a.ndim != 1 or len(a) != self.circuit.ndev:
                raise ValueError(f"capacitor state arrays need {self.circuit.ndev} entries, got {a.shape}")
        self._check(self._L.rs_circuit_set_cap_state(self.h, *[_ptr(a) for a in arrs]))

    def set_initial_guess(self, x=None):
        """The next tran()'s starting point for its DC operating point (the
        netlist's .ic values by node); None clears it."""
        if x is None:
            self._check(self._L.rs_circuit_set_initial_guess(self.h, None, 0))
            return
        self._ic = np.ascontiguousarray(x, np.float64)
        self._check(self._L.rs_circuit_set_initial_guess(self.h, _ptr(self._ic), len(self._ic)))

    def stamp(self, x=None, mode=MODE_DC, t=0.0, dt=0.0, gmin=1e-12, source_scale=1.0):
        """One MNA build: returns (ap, ai, ax, rhs, n) like the engine closure.
        (With BSIM devices the pattern depends on x: it is read after the stamp.)"""
        x = np.ascontiguousarray(np.zeros(self.circuit.node_count) if x is None else x, np.float64)
        ap, ai = self.pattern(mode)
        cap_ax, cap_n = len(ai), len(ap) - 1
        nnz, n = C.c_int64(), C.c_int64()
        for _ in range(2):
            ax = np.zeros(max(cap_ax, 1))
            rhs = np.zeros(max(cap_n, 1))
            rc = self._L.rs_stamp(self.h, mode, _ptr(x), len(x), t, dt, gmin, source_scale,
                                  _ptr(ax), cap_ax, _ptr(rhs), cap_n, C.byref(nnz), C.byref(n))
            if rc == 0:
                break
            if nnz.value <= cap_ax and n.value <= cap_n:
                self._check(rc)
            cap_ax, cap_n = nnz.value, n.value
        self._check(rc)
        ap, ai = self.pattern(mode)
        assert len(ai) == nnz.value and len(ap) - 1 == n.value
        return ap, ai, ax[: nnz.value], rhs[: n.value], n.value

    def stamp_device(self, x_dev, ax_dev, rhs_dev, mode=MODE_DC, t=0.0, dt=0.0, gmin=1e-12,
                     source_scale=1.0, stream=None):
        """rs_stamp_device: x, ax and rhs are float64 device tensors (ax and
        rhs at least as long as the pattern's nnz and n). Returns (nnz, n)."""
        nnz, n = C.c_int64(), C.c_int64()
        s = 0 if stream is None else int(getattr(stream, "cuda_stream", stream))
        self._check(self._L.rs_stamp_device(self.h, mode, C.c_void_p(x_dev.data_ptr()), x_dev.numel(), t, dt,
                                            gmin, source_scale, C.c_void_p(ax_dev.data_ptr()), ax_dev.numel(),
                                            C.c_void_p(rhs_dev.data_ptr()), rhs_dev.numel(), C.byref(nnz),
                                            C.byref(n), C.c_void_p(s)))
        return nnz.value, n.value

    def dc_op(self, solver: GpuSolver, opts=None, last=False):
        """Returns (x or None, NewtonResult). With last=True, x is returned
        whether or not Newton converged: the solution, or the iterate the loop
        ended with (node_count entries)."""
        opts = opts or default_opts()
        cap = self.circuit.node_count + self.circuit.ndev + 8
        x = np.zeros(cap)
        xl = C.c_int64()
        r = NewtonResult()
        self._check(self._L.rs_dc_op(self.h, solver.h, C.byref(opts), _ptr(x), cap, C.byref(xl),
                                     C.byref(r)))
        return (x[: xl.value].copy() if (r.converged or last) else None), r

    def tran(self, solver: GpuSolver, tstep, tstop, tstart=0.0, tmax=None, opts=None,
             max_points=100000, store=True):
        opts = opts or default_opts()
        tmax = tstop if tmax is None else tmax
        stride = self.circuit.node_count + self.circuit.ndev + 8
        times = np.zeros(max_points) if store else None
        sols = np.zeros(max_points * stride) if store else None
        r = TranResult()
        self._check(self._L.rs_tran(self.h, solver.h, C.byref(opts), tstep, tstop, tstart, tmax,
                                    max_points, _ptr(times) if store else None,
                                    _ptr(sols) if store else None, stride, C.byref(r)))
        out = {k: getattr(r, k) for k, _ in TranResult._fields_}
        if store:
            k = min(int(r.npoints), max_points)
            out["times"] = times[:k].copy()
            out["sols"] = sols[: k * stride].reshape(k, stride).copy()
        return out

    def dc_sweep(self, solver: GpuSolver, source: int, start, stop, step, opts=None):
        """Engine::run_dc_sweep_result (engine.rs:684-820) over the value of
        device `source`. Returns dict(status, iterations, factor_calls,
        values, sols[nsolved, stride], fail_value); status 0 Converged,
        1 MaxIters, 2 Failed, 3 source not found."""
        opts = opts or default_opts()
        n = self._L.rs_dc_sweep_points(start, stop, step, None, 0)
        if n < 0:
            raise ValueError("DC sweep range is not finite")
        stride = self.circuit.node_count + self.circuit.ndev + 8
        values = np.zeros(max(n, 1))
        sols = np.zeros((max(n, 1), stride))
        res = SweepResult()
        self._check(self._L.rs_dc_sweep(self.h, solver.h, C.byref(opts), source, start, stop, step,
                                        n, _ptr(values), _ptr(sols), stride, C.byref(res)))
        return dict(status=res.status, iterations=res.iterations, factor_calls=res.factor_calls,
                    values=values[: res.npoints].copy(), sols=sols[: res.nsolved].copy(),
                    fail_value=res.fail_value)

    def set_tran_graph(self, on=True):
        """rs_circuit_set_tran_graph: tran_fixed's steps as captured hipGraphs."""
        self._L.rs_circuit_set_tran_graph(self.h, int(bool(on)))

    def graph_steps(self) -> int:
        return int(self._L.rs_circuit_graph_steps(self.h))

    def tran_graph_info(self) -> dict:
        """Adaptive rs_tran runs on the device controller, graph launches,
        controller powers evaluated by the host's libm."""
        out = np.zeros(3, np.int64)
        self._L.rs_circuit_tran_graph_info(self.h, _ptr(out))
        return dict(runs=int(out[0]), launches=int(out[1]), host_pows=int(out[2]))

    def tran_adjoint(self, solver: GpuSolver, dt, nsteps, probe: int, every: int = 25, opts=None):
        """rs_tran_adjoint: d x_N[probe] / dp for every device after the
        fixed-step TRAP transient of tran_fixed, by the discrete adjoint with a
        checkpoint every `every` steps (every <= 0: every state kept, no
        recomputation). Returns (dp, x_N, info dict)."""
        opts = opts or default_opts()
        cap = self.circuit.node_count + self.circuit.ndev + 8
        x = np.zeros(cap)
        dp = np.zeros(max(self.circuit.ndev, 1))
        xl = C.c_int64()
        info = TranAdjointInfo()
        self._check(self._L.rs_tran_adjoint(self.h, solver.h, C.byref(opts), dt, nsteps, probe, every,
                                            _ptr(dp), _ptr(x), cap, C.byref(xl), C.byref(info)))
        out = {k: getattr(info, k) for k, _ in TranAdjointInfo._fields_}
        return dp[: self.circuit.ndev], x[: xl.value].copy(), out

    def tran_fixed(self, solver: GpuSolver, dt, nsteps, opts=None):
        opts = opts or default_opts()
        cap = self.circuit.node_count + self.circuit.ndev + 8
        x = np.zeros(cap)
        xl, it, st = C.c_int64(), C.c_int32(), C.c_int32()
        self._check(self._L.rs_tran_fixed(self.h, solver.h, C.byref(opts), dt, nsteps, _ptr(x), cap,
                                          C.byref(xl), C.byref(it), C.byref(st)))
        return st.value, x[: xl.value].copy(), it.value

    def tran_steps(self, solver: GpuSolver, t0, dt, nsteps, x, opts=None):
        """Continues a fixed-step TRAP transient from time t0 and state x with
        the circuit's current capacitor history (after tran_fixed or a
        previous call). Returns (status, x, newton_iters)."""
        opts = opts or default_opts()
        cap = max(len(x), self.circuit.node_count + self.circuit.ndev + 8)
        xb = np.zeros(cap)
        xb[: len(x)] = x
        xl, it, st = C.c_int64(len(x)), C.c_int32(), C.c_int32()
        self._check(self._L.rs_tran_steps(self.h, solver.h, C.byref(opts), t0, dt, nsteps, _ptr(xb), cap,
                                          C.byref(xl), C.byref(it), C.byref(st)))
        return st.value, xb[: xl.value].copy(), it.value


def mos_eval(models, model_idx, level, bias):
    """rs_mos_eval: BSIM evaluation of n bias points on the GPU.
    bias: (n, 7) = w, l, vd, vg, vs, vb, temp. Returns (n, 13)."""
    L = _bind()
    models = np.ascontiguousarray(models)
    mi = np.ascontiguousarray(model_idx, dtype=np.int32)
    lv = np.ascontiguousarray(level, dtype=np.int32)
    b = np.ascontiguousarray(bias, dtype=np.float64)
    n = len(mi)
    out = np.zeros((n, 13))
    rc = L.rs_mos_eval(_ptr(models), len(models), _ptr(mi), _ptr(lv), _ptr(b), n, _ptr(out))
    if rc:
        raise SolverError(rc, "rs_mos_eval")
    return out


class GpuBatch:
    """Many independent DC operating points solved in one persistent launch
    (one wavefront per instance, work queue in descending size order)."""

    def __init__(self, circuits, kind=0):
        self._L = _bind()
        self.circuits = list(circuits)
        devs = [np.ascontiguousarray(c.devices) for c in self.circuits]
        self._dev = np.ascontiguousarray(np.concatenate(devs))
        self._ptr = np.zeros(len(devs) + 1, np.int64)
        self._ptr[1:] = np.cumsum([len(d) for d in devs])
        self._nc = np.array([c.node_count for c in self.circuits], np.int64)
        self.h = self._L.rs_batch_create(_ptr(self._dev), _ptr(self._ptr), _ptr(self._nc),
                                         len(self.circuits), kind)
        self._check(0)

    def _check(self, rc):
        buf = C.create_string_buffer(512)
        self._L.rs_batch_last_error(self.h, buf, 512)
        if rc != 0 or buf.value:
            raise SolverError(rc or 7, buf.value.decode(errors="replace"))

    def close(self):
        if getattr(self, "h", None):
            self._L.rs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        out = np.zeros(5, np.int64)
        self._L.rs_batch_info(self.h, _ptr(out))
        return dict(instances=int(out[0]), patterns=int(out[1]), nmax=int(out[2]), grid=int(out[3]),
                    lds_values=int(out[4]))

    PHASES = ("eval", "gather", "rhs", "factor", "refactor", "slots", "solve", "newton")

    def phase_profile(self, on=True):
        """Per-phase shader-clock accounting of later dc_op calls (measurement)."""
        self._check(self._L.rs_batch_phase_profile(self.h, int(on)))

    def phase_cycles(self):
        out = np.zeros(8, np.int64)
        self._check(self._L.rs_batch_phase_cycles(self.h, _ptr(out)))
        return dict(zip(self.PHASES, map(int, out)))

    def dc_op(self, opts=None, stream=None):
        """Enqueues the batched solve (asynchronous on `stream`)."""
        opts = opts or default_opts()
        self._opts = opts
        s = 0 if stream is None else int(getattr(stream, "cuda_stream", stream))
        self._check(self._L.rs_batch_dc_op(self.h, C.byref(opts), C.c_void_p(s)))

    def results(self, last=False):
        """List of (x or None, NewtonResult) per instance (x: full MNA vector).
        With last=True, x is returned for every instance: the solution, or the
        iterate its Newton loop ended with (node_count entries)."""
        k = len(self.circuits)
        res = (NewtonResult * k)()
        xp = np.zeros(k + 1, np.int64)
        self._check(self._L.rs_batch_results(self.h, C.cast(res, C.c_void_p), None, _ptr(xp)))
        x = np.zeros(max(int(xp[-1]), 1))
        self._check(self._L.rs_batch_results(self.h, C.cast(res, C.c_void_p), _ptr(x), _ptr(xp)))
        out = []
        for i in range(k):
            xi = x[xp[i]:xp[i + 1]]
            if res[i].converged:
                out.append((xi.copy(), res[i]))
            elif last:
                out.append((xi[: self.circuits[i].node_count].copy(), res[i]))
            else:
                out.append((None, res[i]))
        return out
