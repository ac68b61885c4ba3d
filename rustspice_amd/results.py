"""Run results and the reference's result file formats.

``RunResult`` / ``ResultStore`` mirror crates/sim-core/src/result_store.rs:1-74.
The writers produce the reference's text formats byte for byte, the ``Date:``
line aside (it is the time of writing):

- PSF text: crates/sim-core/src/psf.rs:7-237 (``write_psf_text``,
  ``write_psf_sweep``, ``write_psf_ac``, ``write_psf_tran``);
- ngspice raw: crates/sim-core/src/raw.rs:12-258 (``write_raw_op``,
  ``write_raw_sweep``, ``write_raw_tran``, ``write_raw_ac``);
- CSV: crates/sim-core/src/csv_export.rs:10-146;
- JSON: crates/sim-core/src/json_export.rs:12-218.

Numbers are written as Rust's ``{:.N e}`` writes them (``1.000000e-3``:
no ``+`` sign and no zero padding in the exponent).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

#: CARGO_PKG_VERSION of crates/sim-core (Cargo.toml:3)
VERSION = "0.1.0"

ANALYSIS = ("Op", "Dc", "Tran", "Ac")
STATUS = ("Converged", "MaxIters", "Failed")


@dataclass
class RunResult:
    """result_store.rs:19-44. ``analysis`` and ``status`` hold the variant
    names (``"Op"``, ``"Converged"``)."""
    analysis: str
    status: str
    iterations: int
    node_names: list
    solution: list = field(default_factory=list)
    message: str | None = None
    sweep_var: str | None = None
    sweep_values: list = field(default_factory=list)
    sweep_solutions: list = field(default_factory=list)
    tran_times: list = field(default_factory=list)
    tran_solutions: list = field(default_factory=list)
    ac_frequencies: list = field(default_factory=list)
    ac_solutions: list = field(default_factory=list)  # per frequency: [(mag_db, phase_deg)] per node
    id: int = 0


class ResultStore:
    """result_store.rs:46-70."""

    def __init__(self):
        self.runs: list[RunResult] = []

    def add_run(self, run: RunResult) -> int:
        run.id = len(self.runs)
        self.runs.append(run)
        return run.id

    def write_psf_text(self, run_id: int, path: str, precision: int = 6) -> None:
        if not 0 <= run_id < len(self.runs):
            raise FileNotFoundError("run not found")
        write_psf_text(self.runs[run_id], path, precision)


# ---------------------------------------------------------------------------
# Rust number formatting
# ---------------------------------------------------------------------------
def _exp_text(mant: str, exp: int) -> str:
    return f"{mant}e{exp}"


def fmt_e(v: float, precision: int | None = None) -> str:
    """Rust ``{:.prec$e}`` (or ``{:e}`` with no precision: the shortest digits
    that read back). NaN is ``NaN``, infinities ``inf``/``-inf``."""
    v = float(v)
    if v != v:
        return "NaN"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    if precision is None:
        sign = "-" if math.copysign(1.0, v) < 0 else ""
        if v == 0.0:
            return sign + "0e0"
        r = repr(abs(v))
        mant, _, e = r.partition("e")
        e = int(e) if e else 0
        ip, _, fp = mant.partition(".")
        digits = (ip + fp).lstrip("0")
        lead = len(ip + fp) - len(digits)
        exp = len(ip) + e - lead - 1
        digits = digits.rstrip("0") or "0"
        m = digits[0] + ("." + digits[1:] if len(digits) > 1 else "")
        return sign + _exp_text(m, exp)
    s = "%.*e" % (precision, v)
    mant, _, e = s.partition("e")
    return _exp_text(mant, int(e))


def _pad_left(s: str, width: int) -> str:
    """``{:width$}`` of a string: left-aligned."""
    return s + " " * max(width - len(s), 0)


def _pad_right(s: str, width: int) -> str:
    """``{:>width$}``."""
    return " " * max(width - len(s), 0) + s


def _days_to_ymd(days: int):
    """psf.rs:254-287 / raw.rs:222-251."""
    year = 1970
    while True:
        diy = 366 if _leap(year) else 365
        if days < diy:
            break
        days -= diy
        year += 1
    months = [31, 29 if _leap(year) else 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31]
    month = 1
    for dm in months:
        if days < dm:
            break
        days -= dm
        month += 1
    return year, month, days + 1


def _leap(y: int) -> bool:
    return (y % 4 == 0 and y % 100 != 0) or y % 400 == 0


def now_text(secs: int | None = None) -> str:
    """``chrono_lite_now`` (psf.rs:228-251): UTC ``YYYY-MM-DDTHH:MM:SSZ``."""
    secs = int(time.time()) if secs is None else int(secs)
    y, mo, d = _days_to_ymd(secs // 86400)
    t = secs % 86400
    return f"{y:04}-{mo:02}-{d:02}T{t // 3600:02}:{(t % 3600) // 60:02}:{t % 60:02}Z"


def _write(path: str, text: str) -> None:
    with open(path, "w", encoding="utf-8", newline="") as f:
        f.write(text)


def _get(sol, i: int, default=0.0):
    return sol[i] if i < len(sol) else default


def _is_current(name: str) -> bool:
    return name.startswith("I(") or name.startswith("i(") or "#branch" in name


# ---------------------------------------------------------------------------
# PSF text (psf.rs)
# ---------------------------------------------------------------------------
def _psf_head() -> list[str]:
    return ["PSF_TEXT\n", f"# Generated by RustSpice v{VERSION}\n", f"# Date: {now_text()}\n", "\n"]


def write_psf_text(run: RunResult, path: str, precision: int = 6) -> None:
    """psf.rs:7-43: a single-point result (OP or a final state)."""
    out = _psf_head()
    out += ["[Header]\n", f"analysis = {run.analysis}\n", f"status = {run.status}\n",
            f"iterations = {run.iterations}\n", "\n", "[Signals]\n"]
    out += [f"V({n})\n" for n in run.node_names]
    out += ["\n", "[Values]\n"]
    width = max((len(n) + 3 for n in run.node_names), default=4)
    for i, v in enumerate(run.solution):
        name = run.node_names[i] if i < len(run.node_names) else f"n{i}"
        out.append(f"{_pad_left(f'V({name})', width)}  {fmt_e(v, precision)}\n")
    _write(path, "".join(out))


def write_psf_sweep(source, sweep_values, node_names, sweep_results, path, precision=6) -> None:
    """psf.rs:46-103."""
    out = _psf_head()
    out += ["[Sweep]\n", f"source = {source}\n", f"points = {len(sweep_values)}\n", "\n", "[Signals]\n",
            f"{source}\n"]
    out += [f"V({n})\n" for n in node_names]
    out += ["\n", "[Data]\n"]
    w = precision + 8
    out.append(_pad_right(source, w) + "".join(f"  {_pad_right(f'V({n})', w)}" for n in node_names) + "\n")
    for i, sv in enumerate(sweep_values):
        row = _pad_right(fmt_e(sv, precision), w)
        if i < len(sweep_results):
            sol = sweep_results[i]
            row += "".join(f"  {_pad_right(fmt_e(_get(sol, k), precision), w)}" for k in range(len(node_names)))
        out.append(row + "\n")
    _write(path, "".join(out))


def write_psf_ac(frequencies, node_names, ac_solutions, path, precision=6) -> None:
    """psf.rs:106-165."""
    out = _psf_head()
    out += ["[AC Analysis]\n", "type = frequency_response\n", f"points = {len(frequencies)}\n", "\n",
            "[Signals]\n", "frequency\n"]
    for n in node_names:
        if n != "0":
            out += [f"VM({n})\n", f"VP({n})\n"]
    out += ["\n", "[Data]\n"]
    w = precision + 10
    head = _pad_right("frequency", w)
    for n in node_names:
        if n != "0":
            head += f"  {_pad_right(f'VM({n})', w)}  {_pad_right(f'VP({n})', w)}"
    out.append(head + "\n")
    for i, f in enumerate(frequencies):
        row = _pad_right(fmt_e(f, precision), w)
        if i < len(ac_solutions):
            sol = ac_solutions[i]
            for k, n in enumerate(node_names):
                if n != "0" and k < len(sol):
                    m, p = sol[k]
                    row += f"  {_pad_right(fmt_e(m, precision), w)}  {_pad_right(fmt_e(p, precision), w)}"
        out.append(row + "\n")
    _write(path, "".join(out))


def write_psf_tran(times, node_names, solutions, path, precision=6) -> None:
    """psf.rs:168-225."""
    out = _psf_head()
    out += ["[Transient Analysis]\n", f"points = {len(times)}\n"]
    if len(times):
        out += [f"tstart = {fmt_e(times[0])}\n", f"tstop = {fmt_e(times[-1])}\n"]
    out += ["\n", "[Signals]\n", "time\n"]
    out += [f"V({n})\n" for n in node_names if n != "0"]
    out += ["\n", "[Data]\n"]
    w = precision + 8
    out.append(_pad_right("time", w) + "".join(f"  {_pad_right(f'V({n})', w)}" for n in node_names if n != "0")
               + "\n")
    for i, t in enumerate(times):
        row = _pad_right(fmt_e(t, precision), w)
        if i < len(solutions):
            sol = solutions[i]
            row += "".join(f"  {_pad_right(fmt_e(_get(sol, k), precision), w)}"
                           for k, n in enumerate(node_names) if n != "0")
        out.append(row + "\n")
    _write(path, "".join(out))


# ---------------------------------------------------------------------------
# ngspice raw (raw.rs)
# ---------------------------------------------------------------------------
def _raw_head(plotname: str, flags: str, nvars: int, npoints: int) -> list[str]:
    """raw.rs:197-210."""
    return [f"Title: RustSpice v{VERSION} Simulation\n", f"Date: {now_text()}\n", f"Plotname: {plotname}\n",
            f"Flags: {flags}\n", f"No. Variables: {nvars}\n", f"No. Points: {npoints}\n"]


def write_raw_op(run: RunResult, path: str, precision: int = 6) -> None:
    """raw.rs:13-52 (ground filtered; an empty solution reads as zeros)."""
    names = [n for n in run.node_names if n != "0"]
    vals = [_get(run.solution, i) for i, n in enumerate(run.node_names) if n != "0"]
    if not vals:
        raise IndexError("write_raw_op needs at least one non-ground node")
    out = _raw_head("Operating Point", "real", len(names), 1)
    out.append("Variables:\n")
    for i, n in enumerate(names):
        out.append(f"\t{i}\tv({n})\t{'current' if _is_current(n) else 'voltage'}\n")
    out.append("Values:\n")
    out.append(f" 0\t{fmt_e(vals[0], precision)}\n")
    out += [f"\t{fmt_e(v, precision)}\n" for v in vals[1:]]
    _write(path, "".join(out))


def _raw_points(out, xs, node_names, sols, precision, first):
    for i, x in enumerate(xs):
        out.append(f" {i}\t{first(x)}\n")
        if i < len(sols):
            sol = sols[i]
            out += [f"\t{fmt_e(_get(sol, k), precision)}\n" for k, n in enumerate(node_names) if n != "0"]


def write_raw_sweep(source, sweep_values, node_names, sweep_results, path, precision=6) -> None:
    """raw.rs:55-101."""
    names = [n for n in node_names if n != "0"]
    out = _raw_head("DC transfer characteristic", "real", 1 + len(names), len(sweep_values))
    out += ["Variables:\n", f"\t0\t{source.lower()}\tvoltage\n"]
    out += [f"\t{i + 1}\tv({n})\tvoltage\n" for i, n in enumerate(names)]
    out.append("Values:\n")
    _raw_points(out, sweep_values, node_names, sweep_results, precision, lambda x: fmt_e(x, precision))
    _write(path, "".join(out))


def write_raw_tran(times, node_names, solutions, path, precision=6) -> None:
    """raw.rs:104-148."""
    names = [n for n in node_names if n != "0"]
    out = _raw_head("Transient Analysis", "real", 1 + len(names), len(times))
    out += ["Variables:\n", "\t0\ttime\ttime\n"]
    out += [f"\t{i + 1}\tv({n})\tvoltage\n" for i, n in enumerate(names)]
    out.append("Values:\n")
    _raw_points(out, times, node_names, solutions, precision, lambda x: fmt_e(x, precision))
    _write(path, "".join(out))


def db_phase_to_complex(mag_db: float, phase_deg: float):
    """raw.rs:217-224."""
    mag = math.pow(10.0, mag_db / 20.0)
    ph = phase_deg * math.pi / 180.0
    return mag * math.cos(ph), mag * math.sin(ph)


def write_raw_ac(frequencies, node_names, ac_solutions, path, precision=6) -> None:
    """raw.rs:154-194: values as ``real,imag`` from (dB, degrees)."""
    names = [n for n in node_names if n != "0"]
    out = _raw_head("AC Analysis", "complex", 1 + len(names), len(frequencies))
    out += ["Variables:\n", "\t0\tfrequency\tfrequency\n"]
    out += [f"\t{i + 1}\tv({n})\tvoltage\n" for i, n in enumerate(names)]
    out.append("Values:\n")
    for i, f in enumerate(frequencies):
        out.append(f" {i}\t{fmt_e(f, precision)},{fmt_e(0.0, precision)}\n")
        if i < len(ac_solutions):
            sol = ac_solutions[i]
            for k, n in enumerate(node_names):
                if n != "0" and k < len(sol):
                    re_, im_ = db_phase_to_complex(*sol[k])
                    out.append(f"\t{fmt_e(re_, precision)},{fmt_e(im_, precision)}\n")
    _write(path, "".join(out))


# ---------------------------------------------------------------------------
# CSV (csv_export.rs) and JSON (json_export.rs)
# ---------------------------------------------------------------------------
def write_csv_op(run: RunResult, path: str, precision: int = 6) -> None:
    out = ["node,type,value\n"]
    for i, n in enumerate(run.node_names):
        if n != "0":
            out.append(f"{n},{'current' if _is_current(n) else 'voltage'},{fmt_e(_get(run.solution, i), precision)}\n")
    _write(path, "".join(out))


def _csv_series(head: str, xs, node_names, sols, precision) -> str:
    out = [head + "".join(f",V({n})" for n in node_names if n != "0") + "\n"]
    for i, x in enumerate(xs):
        row = fmt_e(x, precision)
        if i < len(sols):
            row += "".join(f",{fmt_e(_get(sols[i], k), precision)}" for k, n in enumerate(node_names) if n != "0")
        out.append(row + "\n")
    return "".join(out)


def write_csv_sweep(source, sweep_values, node_names, sweep_results, path, precision=6) -> None:
    _write(path, _csv_series(source, sweep_values, node_names, sweep_results, precision))


def write_csv_tran(times, node_names, solutions, path, precision=6) -> None:
    _write(path, _csv_series("time", times, node_names, solutions, precision))


def write_csv_ac(frequencies, node_names, ac_solutions, path, precision=6) -> None:
    out = ["frequency" + "".join(f",|V({n})|_dB,phase({n})" for n in node_names if n != "0") + "\n"]
    for i, f in enumerate(frequencies):
        row = fmt_e(f, precision)
        if i < len(ac_solutions):
            sol = ac_solutions[i]
            for k, n in enumerate(node_names):
                if n != "0":
                    m, p = sol[k] if k < len(sol) else (0.0, 0.0)
                    row += f",{fmt_e(m, precision)},{fmt_e(p, precision)}"
        out.append(row + "\n")
    _write(path, "".join(out))


def _json_head(analysis: str) -> list[str]:
    return ["{\n", '  "format": "rustspice-json",\n', f'  "version": "{VERSION}",\n', f'  "analysis": "{analysis}",\n']


def write_json_op(run: RunResult, path: str, precision: int = 6) -> None:
    out = ["{\n", '  "format": "rustspice-json",\n', f'  "version": "{VERSION}",\n',
           f'  "analysis": "{run.analysis}",\n', f'  "status": "{run.status}",\n',
           f'  "iterations": {run.iterations},\n', '  "variables": [\n']
    items = [(i, n) for i, n in enumerate(run.node_names) if n != "0"]
    for j, (i, n) in enumerate(items):
        comma = "," if j < len(items) - 1 else ""
        out.append(f'    {{"name": "{n}", "type": "{"current" if _is_current(n) else "voltage"}", '
                   f'"value": {fmt_e(_get(run.solution, i), precision)}}}{comma}\n')
    out += ["  ]\n", "}\n"]
    _write(path, "".join(out))


def _json_series(analysis, extra, first_var, xs, node_names, sols, precision) -> str:
    out = _json_head(analysis) + extra + [f'  "points": {len(xs)},\n', '  "variables": [\n', first_var]
    out += [f',\n    {{"name": "{n}", "type": "voltage"}}' for n in node_names if n != "0"]
    out += ["\n  ],\n", '  "data": [\n']
    for i, x in enumerate(xs):
        row = "    [" + fmt_e(x, precision)
        if i < len(sols):
            row += "".join(f", {fmt_e(_get(sols[i], k), precision)}" for k, n in enumerate(node_names) if n != "0")
        out.append(row + "]" + ("," if i < len(xs) - 1 else "") + "\n")
    out += ["  ]\n", "}\n"]
    return "".join(out)


def write_json_sweep(source, sweep_values, node_names, sweep_results, path, precision=6) -> None:
    _write(path, _json_series("Dc", [f'  "sweep_source": "{source}",\n'],
                              f'    {{"name": "{source}", "type": "sweep"}}', sweep_values, node_names,
                              sweep_results, precision))


def write_json_tran(times, node_names, solutions, path, precision=6) -> None:
    _write(path, _json_series("Tran", [], '    {"name": "time", "type": "time"}', times, node_names, solutions,
                              precision))


def write_json_ac(frequencies, node_names, ac_solutions, path, precision=6) -> None:
    out = _json_head("Ac") + [f'  "points": {len(frequencies)},\n', '  "variables": [\n',
                              '    {"name": "frequency", "type": "frequency"}']
    out += [f',\n    {{"name": "{n}", "type": "ac"}}' for n in node_names if n != "0"]
    out += ["\n  ],\n", '  "data": [\n']
    for i, f in enumerate(frequencies):
        row = "    {" + f'"frequency": {fmt_e(f, precision)}'
        if i < len(ac_solutions):
            sol = ac_solutions[i]
            vals = []
            for k, n in enumerate(node_names):
                if n != "0":
                    m, p = sol[k] if k < len(sol) else (0.0, 0.0)
                    vals.append(f'{{"mag_db": {fmt_e(m, precision)}, "phase_deg": {fmt_e(p, precision)}}}')
            row += ', "values": [' + ", ".join(vals) + "]"
        out.append(row + "}" + ("," if i < len(frequencies) - 1 else "") + "\n")
    out += ["  ]\n", "}\n"]
    _write(path, "".join(out))
