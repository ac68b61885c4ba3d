"""AC small-signal sweep on the GPU: ``Engine::run_ac_result`` (engine.rs:829-970).

The reference stamps and solves one frequency after another with a dense
complex LU (complex_solver.rs:50-136). ``GpuAc`` hands the whole sweep to one
launch of librsgpu.so's ``k_ac_solve``, one workgroup per frequency, with the
same arithmetic per frequency (see rsgpu.h, ``rs_ac_*``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from . import results as R
from .solver import SolverError

SWEEPS = {"dec": 0, "oct": 1, "lin": 2}


def _bind():
    L = _lib.lib()
    if getattr(L, "_ac_bound", False):
        return L
    vp, i64 = C.c_void_p, C.c_int64
    L.rs_ac_frequencies.restype = i64
    L.rs_ac_frequencies.argtypes = [C.c_int32, i64, C.c_double, C.c_double, vp, i64]
    L.rs_ac_create.restype = vp
    L.rs_ac_create.argtypes = [vp, i64, i64, vp, i64, vp, vp]
    L.rs_ac_run.restype = C.c_int32
    L.rs_ac_run.argtypes = [vp, vp, i64, vp, i64, vp, vp, i64, C.POINTER(i64), C.POINTER(i64)]
    L.rs_ac_timing.restype = C.c_int32
    L.rs_ac_timing.argtypes = [vp, vp]
    L.rs_ac_last_error.restype = C.c_int32
    L.rs_ac_last_error.argtypes = [vp, C.c_char_p, i64]
    L.rs_ac_destroy.argtypes = [vp]
    L._ac_bound = True
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def frequencies(sweep: str, points: int, fstart: float, fstop: float) -> np.ndarray:
    """generate_frequency_sweep (engine.rs:1002-1044)."""
    L = _bind()
    k = SWEEPS[sweep]
    n = L.rs_ac_frequencies(k, points, fstart, fstop, None, 0)
    out = np.zeros(max(n, 1))
    L.rs_ac_frequencies(k, points, fstart, fstop, _ptr(out), n)
    return out[:n].copy()


class GpuAc:
    """The AC sweep of one circuit: ``circuit`` holds the records the AC
    stamps read (a V/I source is stamped whether or not its DC value parses),
    ``ac_mag``/``ac_phase`` each record's excitation (0 when none)."""

    def __init__(self, circuit, ac_mag, ac_phase):
        self._L = _bind()
        self.circuit = circuit
        self._dev = np.ascontiguousarray(circuit.devices)
        models = getattr(circuit, "models", None)
        self._models = np.ascontiguousarray(models) if models is not None and len(models) else None
        self._mag = np.ascontiguousarray(ac_mag, np.float64)
        self._ph = np.ascontiguousarray(ac_phase, np.float64)
        if len(self._mag) != circuit.ndev or len(self._ph) != circuit.ndev:
            raise ValueError("ac_mag and ac_phase need one entry per device")
        self.h = self._L.rs_ac_create(_ptr(self._dev), circuit.ndev, circuit.node_count, _ptr(self._models),
                                      0 if self._models is None else len(self._models), _ptr(self._mag),
                                      _ptr(self._ph))

    def close(self):
        if getattr(self, "h", None):
            self._L.rs_ac_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, x_dc, freqs):
        """Returns (x[nsolved, n] complex, magph[nsolved, node_count, 2] of
        (dB, degrees), n, nsolved)."""
        x_dc = np.ascontiguousarray(x_dc, np.float64)
        freqs = np.ascontiguousarray(freqs, np.float64)
        nv = int(np.count_nonzero(self._dev["kind"] == 3)) if self.circuit.ndev else 0
        cap = self.circuit.node_count + nv
        nf = len(freqs)
        xo = np.zeros(max(nf * cap * 2, 1))
        mp = np.zeros(max(nf * self.circuit.node_count * 2, 1))
        n, ns = C.c_int64(), C.c_int64()
        rc = self._L.rs_ac_run(self.h, _ptr(x_dc), len(x_dc), _ptr(freqs), nf, _ptr(xo), _ptr(mp), cap,
                               C.byref(n), C.byref(ns))
        if rc:
            buf = C.create_string_buffer(512)
            self._L.rs_ac_last_error(self.h, buf, 512)
            raise SolverError(rc, buf.value.decode(errors="replace"))
        k = ns.value
        x = xo[: nf * cap * 2].reshape(nf, cap, 2)[:k, : n.value]
        return (x[..., 0] + 1j * x[..., 1], mp[: nf * self.circuit.node_count * 2]
                .reshape(nf, self.circuit.node_count, 2)[:k], n.value, k)

    def timing(self) -> dict:
        out = np.zeros(5)
        self._L.rs_ac_timing(self.h, _ptr(out))
        return dict(ms=out[0], launches=int(out[1]), n=int(out[2]), entries=int(out[3]), contributions=int(out[4]))


def run_ac(sim, sweep: str, points: int, fstart: float, fstop: float) -> R.RunResult:
    """Engine::run_ac_result (engine.rs:829-970) for a Simulator: its DC
    operating point, then the sweep; Failed at the first frequency whose
    solve fails, with the frequencies solved before it."""
    op = sim.op()
    if op.status != "Converged":
        return R.RunResult("Ac", "Failed", 0, sim.node_names,
                           message="DC operating point failed to converge")
    freqs = frequencies(sweep, points, fstart, fstop)
    low = sim.low
    ac = GpuAc(low.ac_circuit, low.ac_mag, low.ac_phase)
    try:
        _, magph, _, k = ac.run(np.array(op.solution), freqs)
    finally:
        ac.close()
    sols = [[(float(m), float(p)) for m, p in magph[f]] for f in range(k)]
    if k < len(freqs):
        from .netlist import rust_f64_display

        return R.RunResult("Ac", "Failed", k, sim.node_names, solution=op.solution,
                           message=f"AC solve failed at frequency {rust_f64_display(freqs[k])} Hz",
                           ac_frequencies=list(freqs[:k]), ac_solutions=sols)
    return R.RunResult("Ac", "Converged", k, sim.node_names, solution=op.solution,
                       ac_frequencies=list(freqs), ac_solutions=sols)
