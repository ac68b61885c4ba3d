"""Netlist-driven simulation on the GPU: ``Engine::run_with_store`` for SPICE text.

``Simulator`` mirrors crates/sim-core/src/engine.rs:43-247: one circuit, one
solver for every analysis (so a NativeKLU pattern's pivots carry over from one
analysis to the next, as the reference's ``Engine.solver`` does), and the
netlist's analyses run in order into a :class:`~rustspice_amd.results.ResultStore`:

- ``.op``   -> ``rs_dc_op``    (``run_dc_result``, engine.rs:249-311)
- ``.dc``   -> ``rs_dc_sweep`` (``run_dc_sweep_result``, engine.rs:684-820)
- ``.tran`` -> ``rs_tran``     (``run_tran_result_with_params``, engine.rs:333-632),
  starting from the ``.ic`` values
- ``.ac``   -> ``rs_ac``       (``run_ac_result``, engine.rs:829-970)

Every analysis runs through librsgpu.so; there is no CPU path.
"""
from __future__ import annotations

import numpy as np

from . import engine as E
from . import netlist as N
from . import results as R
from .solver import BBD, GpuSolver

_STATUS = {0: "Converged", 1: "MaxIters", 2: "Failed"}


class Simulator:
    """A parsed netlist on the GPU. ``solver_kind`` overrides ``.option solver=``."""

    def __init__(self, circ: N.SpiceCircuit, solver_kind: int | None = None):
        self.circ = circ
        self.low = N.lower(circ)
        self.kind = N.solver_kind(circ) if solver_kind is None else solver_kind
        self.gc = E.GpuCircuit(self.low.circuit)
        self.solver = GpuSolver(circ.node_count, self.kind)
        if self.kind == BBD:
            self.solver.set_bbd_blocks(4)  # SolverType::Bbd is create_bbd_solver(n, 4) (solver.rs:176)
        threads = int(circ.options.values["solver_parallel"])
        if threads:
            self.solver.set_parallel_threads(threads)
        self.opts = E.default_opts(**N.newton_options(circ))

    @classmethod
    def from_text(cls, text: str, **kw) -> "Simulator":
        return cls(N.load(text=text), **kw)

    @classmethod
    def from_file(cls, path: str, **kw) -> "Simulator":
        return cls(N.load(path=path), **kw)

    def close(self):
        self.gc.close()
        self.solver.close()

    @property
    def node_names(self) -> list:
        return list(self.circ.node_names)

    # -- analyses ----------------------------------------------------------
    def op(self) -> R.RunResult:
        x, r = self.gc.dc_op(self.solver, self.opts)
        status = _STATUS.get(int(r.reason), "Failed") if not r.converged else "Converged"
        return R.RunResult("Op", status, int(r.iterations), self.node_names,
                           solution=list(x) if r.converged else [])

    def dc(self, source: str, start: float, stop: float, step: float) -> R.RunResult:
        want = N._lower(source)
        idx = next((i for i, inst in enumerate(self.circ.instances) if N._lower(inst.name) == want), None)
        if idx is None:
            return R.RunResult("Dc", "Failed", 0, self.node_names,
                               message=f"DC sweep source '{source}' not found", sweep_var=source)
        rec = self.low.record_of.get(idx)
        if rec is None:
            raise N.UnsupportedDevice(f"DC sweep source '{source}' has no stamped record")
        d = self.gc.dc_sweep(self.solver, rec, start, stop, step, self.opts)
        status = _STATUS.get(int(d["status"]), "Failed")
        sols = [list(s[: self._dc_n()]) for s in d["sols"]]
        msg = None
        if status == "MaxIters":
            msg = f"Failed to converge at sweep point {N.rust_f64_display(d['fail_value'])}"
        elif status == "Failed":
            msg = f"Solver failure at sweep point {N.rust_f64_display(d['fail_value'])}"
        return R.RunResult("Dc", status, int(d["iterations"]), self.node_names,
                           solution=sols[-1] if sols else [], message=msg, sweep_var=source,
                           sweep_values=list(d["values"]), sweep_solutions=sols)

    def tran(self, tstep: float, tstop: float, tstart: float = 0.0, tmax: float | None = None,
             max_points: int = 1_000_000) -> R.RunResult:
        self.gc.set_initial_guess(N.to_numpy_ic(self.circ))
        try:
            d = self.gc.tran(self.solver, tstep, tstop, tstart, tstop if tmax is None else tmax, self.opts,
                             max_points=max_points)
        finally:
            self.gc.set_initial_guess(None)
        if d["npoints"] == 0:  # the initial DC operating point failed (engine.rs:431-447)
            return R.RunResult("Tran", "Failed", 0, self.node_names,
                               message="DC operating point failed to converge")
        status = "Converged" if d["status"] == 0 else "Failed"
        n_tr = self._tran_n()
        rows = [list(s[: (self._dc_n() if i == 0 else n_tr)]) for i, s in enumerate(d.get("sols", []))]
        # RunResult.iterations of a TRAN is its accepted step count (engine.rs:618)
        return R.RunResult("Tran", status, int(d["accepted"]), self.node_names,
                           solution=rows[-1] if rows else [], tran_times=list(d.get("times", [])),
                           tran_solutions=rows,
                           message=f"Adaptive: {d['accepted']} accepted, {d['rejected']} rejected")

    def ac(self, sweep: str, points: int, fstart: float, fstop: float) -> R.RunResult:
        from . import ac as A

        return A.run_ac(self, sweep, points, fstart, fstop)

    def run(self, analysis: tuple) -> R.RunResult:
        kind = analysis[0]
        if kind == "op":
            return self.op()
        if kind == "dc":
            return self.dc(*analysis[1:])
        if kind == "tran":
            return self.tran(*analysis[1:])
        if kind == "ac":
            return self.ac(*analysis[1:])
        raise ValueError(f"unknown analysis {analysis!r}")

    def run_all(self, store: R.ResultStore | None = None) -> R.ResultStore:
        """Every analysis of the netlist in order (``.op`` when it names none)."""
        store = store if store is not None else R.ResultStore()
        for a in self.circ.analyses:
            store.add_run(self.run(a))
        return store

    def _dc_n(self) -> int:
        ap, _ = self.gc.pattern(E.MODE_DC)
        return len(ap) - 1

    def _tran_n(self) -> int:
        ap, _ = self.gc.pattern(E.MODE_TRAN_TRAP)
        return len(ap) - 1


def simulate(text: str | None = None, path: str | None = None, solver_kind: int | None = None) -> R.ResultStore:
    """Parse a netlist, run its analyses on the GPU and return the results."""
    sim = Simulator(N.load(text=text, path=path), solver_kind=solver_kind)
    try:
        return sim.run_all()
    finally:
        sim.close()
