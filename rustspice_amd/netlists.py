"""Seeded synthetic circuits for the five benchmark/parity workloads.

Each generator returns a :class:`Circuit`: a numpy record array of
``rs_device`` (``include/rs_circuit.h``) plus the node count.  The records are
what RustSpice's stamp loop would see after parsing a netlist once
(crates/sim-core/src/stamp.rs:94-553): numbers are stored as the doubles the
reference's ``parse_number_with_suffix`` (stamp.rs:721-747) yields for the
netlist text, e.g. ``10p`` is ``10 * 1e-12``.

Randomness uses the reference's 64-bit LCG (tests/bbd_benchmark_tests.rs:290-301:
``state = state * 6364136223846793005 + 1442695040888963407`` wrapping) with
``u = (state >> 11) * 2**-53`` and Box-Muller normals; the seed of workload k
is ``0x52530000 + k``.

Node numbering follows the reference's NodeTable (circuit.rs:31-57): ground is
0 and nodes are numbered by first appearance; every generator lists devices so
that the first appearance order equals the ids documented below.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass

import numpy as np

RS_DEV_R, RS_DEV_C, RS_DEV_V, RS_DEV_I, RS_DEV_D, RS_DEV_M = 1, 2, 3, 4, 5, 6
RS_F_PMOS, RS_F_WAVEFORM = 1, 2

DEVICE_DTYPE = np.dtype(
    [
        ("kind", np.int32),
        ("flags", np.int32),
        ("node", np.int32, (4,)),
        ("value", np.float64),
        ("p", np.float64, (8,)),
    ],
    align=True,
)
assert DEVICE_DTYPE.itemsize == 96

#: EPSILON_OX of crates/sim-devices/src/bsim/params.rs:10
EPSILON_OX = 3.9 * 8.854e-12
TOX_DEFAULT = 1.5e-8  # params.rs:344


def sfx(mantissa: float, mult: float) -> float:
    """Value of a SPICE token ``<mantissa><suffix>`` as the reference parses it."""
    return float(mantissa) * mult


P, N, U, F = 1e-12, 1e-9, 1e-6, 1e-15


class Lcg:
    """tests/bbd_benchmark_tests.rs:295-298 LCG."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state * 6364136223846793005 + 1442695040888963407) & self.MASK
        return self.state

    def uniform(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def normal(self) -> float:
        u1 = 1.0 - self.uniform()  # (0, 1]
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def generate_rhs(n: int, seed: int) -> np.ndarray:
    """bbd_benchmark_tests.rs:290-301 generate_rhs."""
    out = np.empty(n, dtype=np.float64)
    st = seed & Lcg.MASK
    for i in range(n):
        st = (st * 6364136223846793005 + 1442695040888963407) & Lcg.MASK
        out[i] = float(st >> 33) / float(1 << 31) - 1.0
    return out


@dataclass
class Circuit:
    devices: np.ndarray  # DEVICE_DTYPE records
    node_count: int
    name: str = ""
    models: np.ndarray | None = None  # BSIM_MODEL_DTYPE records for MOSFETs of level != 1

    @property
    def ndev(self) -> int:
        return int(self.devices.shape[0])


class _Builder:
    def __init__(self):
        self.rows = []

    def add(self, kind, nodes, value=0.0, p=None, flags=0):
        node = list(nodes) + [0] * (4 - len(nodes))
        pp = list(p or []) + [0.0] * (8 - len(p or []))
        self.rows.append((kind, flags, node, float(value), pp))

    def build(self, node_count, name, models=None):
        arr = np.zeros(len(self.rows), dtype=DEVICE_DTYPE)
        for i, (k, f, nd, v, pp) in enumerate(self.rows):
            arr[i]["kind"] = k
            arr[i]["flags"] = f
            arr[i]["node"] = nd
            arr[i]["value"] = v
            arr[i]["p"] = pp
        if models is not None:
            models = np.ascontiguousarray(np.array(list(models), dtype=BSIM_MODEL_DTYPE))
        return Circuit(arr, node_count, name, models)


def mos_params(vth0: float, u0: float, w: float, l: float, tox: float = TOX_DEFAULT):
    """p[] slots of a level-1 MOSFET (rs_mos_param)."""
    return [1.0, vth0, u0, tox, w, l]


#: rs_bsim_model (include/rs_circuit.h): the BsimParams fields evaluate_bsim_dc
#: and evaluate_bsim4_dc read, plus the instance's sa and sb
BSIM_FIELDS = ("vth0 k1 k2 dvt0 dvt1 dvt2 eta0 dsub drout nlx nfactor "
               "u0 ua ub uc vsat pclm tox lint wint rdsw prt "
               "tnom ute kt1 kt1l kt2 saref sbref ku0 kvth0 tku0 "
               "alpha0 alpha1 beta0 beta1 jtss jtsd vtss vtsd nsti sa sb").split()
BSIM_MODEL_DTYPE = np.dtype([(f, np.float64) for f in BSIM_FIELDS] +
                            [("pmos", np.int32), ("pad", np.int32)], align=True)
assert BSIM_MODEL_DTYPE.itemsize == 43 * 8 + 8

#: BsimParams::nmos_default (crates/sim-devices/src/bsim/params.rs:309-440)
BSIM_NMOS_DEFAULT = dict(
    vth0=0.7, k1=0.5, k2=0.0, dvt0=2.2, dvt1=0.53, dvt2=-0.032, eta0=0.08, dsub=0.56,
    drout=0.56, nlx=1.74e-7, nfactor=1.0, u0=500.0, ua=2.25e-9, ub=5.87e-19, uc=-4.65e-11,
    vsat=1.5e5, pclm=1.3, tox=1.5e-8, lint=0.0, wint=0.0, rdsw=0.0, prt=0.0, tnom=300.15,
    ute=-1.5, kt1=-0.11, kt1l=0.0, kt2=0.022, saref=1e-6, sbref=1e-6, ku0=0.0, kvth0=0.0,
    tku0=0.0, alpha0=0.0, alpha1=0.0, beta0=30.0, beta1=0.0, jtss=0.0, jtsd=0.0, vtss=10.0,
    vtsd=10.0, nsti=1.0, sa=0.0, sb=0.0)
#: BsimParams::pmos_default (params.rs:443-455): the fields it changes
BSIM_PMOS_CHANGES = dict(vth0=-0.7, u0=150.0, ute=-1.0, kt1=-0.08)


def bsim_model(pmos: bool = False, tnom_c: float | None = None, **overrides) -> np.ndarray:
    """One rs_bsim_model record: the reference's defaults for the type, then the
    given instance/model parameters as build_bsim_params (bsim/mod.rs:83-388)
    applies them. `tnom_c` is the netlist's TNOM in Celsius (+273.15)."""
    rec = np.zeros((), dtype=BSIM_MODEL_DTYPE)
    vals = dict(BSIM_NMOS_DEFAULT)
    if pmos:
        vals.update(BSIM_PMOS_CHANGES)
    if tnom_c is not None:
        vals["tnom"] = tnom_c + 273.15
    for k, v in overrides.items():
        if k not in vals:
            raise KeyError(k)
        vals[k] = float(v)
    for k, v in vals.items():
        rec[k] = v
    rec["pmos"] = 1 if pmos else 0
    return rec


_RUST_F64 = re.compile(r"[+-]?((\d+\.?\d*|\.\d+)(e[+-]?\d+)?|inf|infinity|nan)")


def _rust_f64(s: str):
    """str::parse::<f64> on a lower-case string: an optional sign, then
    digits with an optional point and exponent, or inf/infinity/nan; no
    whitespace, underscores or hex (which Python's float() would take)."""
    return float(s) if _RUST_F64.fullmatch(s) else None


def parse_number(s: str):
    """parse_number (crates/sim-devices/src/bsim/mod.rs:390-420): a number
    with an optional SI suffix (meg, mil, f p n u m k g t; lower-cased,
    trimmed); when the part before a one-letter suffix does not parse, the
    whole string is tried as a plain number. None when nothing parses."""
    t = s.lower().strip()
    if t.endswith("meg"):
        num, mult = t[:-3], 1e6
    elif t.endswith("mil"):
        num, mult = t[:-3], 25.4e-6
    else:
        sfx = {"f": 1e-15, "p": 1e-12, "n": 1e-9, "u": 1e-6, "m": 1e-3, "k": 1e3, "g": 1e9, "t": 1e12}
        cut = max(len(t) - 1, 0)
        num, mult = (t[:cut], sfx[t[cut:]]) if t[cut:] in sfx else (t, 1.0)
    v = _rust_f64(num)
    return v * mult if v is not None else _rust_f64(t)


#: build_bsim_params (bsim/mod.rs:83-388): (BsimParams field, netlist keys
#: tried in order); tnom is given in Celsius (+273.15) and pemod is cast to
#: an unsigned integer
BSIM_PARAM_KEYS = (
    [("vth0", ("vth0", "vto", "vth"))] +
    [(k, (k,)) for k in "k1 k2 dvt0 dvt1 dvt2 eta0 dsub nlx nfactor".split()] +
    [("u0", ("u0", "uo"))] +
    [(k, (k,)) for k in ("ua ub uc vsat a0 ags pclm pdiblc1 pdiblc2 pdiblcb drout tox lint wint rdsw "
                         "rsh tnom ute kt1 kt1l kt2 cgso cgdo cgbo dvt0w dvt1w dvt2w voff voffl minv "
                         "k3 k3b w0 lpe0 lpeb vfb ute0 ute1 pemod up lp ud ud1 eu vs vsattemp lambda "
                         "vtl lc alpha0 alpha1 beta0 beta1 saref sbref wlod ku0 kvth0 ku0mult tku0 "
                         "jtss jtsd nsti vtss vtsd pvag fprout pdits delta").split()])


def build_bsim_params(params: dict, level: int, is_pmos: bool):
    """build_bsim_params (bsim/mod.rs:83-388): the type's defaults, then every
    parameter the map gives (keys as stored in the map; a key's value that
    does not parse falls through to the next alias). Returns (record, parsed):
    the rs_bsim_model record the device evaluation reads (level and type
    apart) and every BsimParams field the map set, including the ones the
    evaluation does not read."""
    parsed = {}
    for field, keys in BSIM_PARAM_KEYS:
        for k in keys:
            v = params.get(k.lower())
            v = parse_number(v) if v is not None else None
            if v is not None:
                break
        if v is None:
            continue
        if field == "tnom":
            v = v + 273.15
        elif field == "pemod":  # `v as u32`: truncation, saturating, NaN -> 0
            v = 0 if v != v else int(min(max(v, 0.0), 4294967295.0))
        parsed[field] = v
    rec = bsim_model(pmos=is_pmos, **{k: v for k, v in parsed.items() if k in BSIM_FIELDS})
    parsed["level"] = int(level)
    parsed["pmos"] = bool(is_pmos)
    return rec, parsed


def bsim_params(level: int, model: int, w: float, l: float, temp: float = 300.15):
    """p[] slots of a BSIM MOSFET (rs_mos_param): level 49 BSIM3, 54 BSIM4."""
    return [float(level), 0.0, 0.0, 0.0, w, l, float(model), temp]


def u0_from_kp(kp: float, tox: float = TOX_DEFAULT) -> float:
    """Level 1 in the reference ignores KP (bsim/mod.rs:441-455) and uses
    beta = u0 * 1e-4 * cox; a netlist that wants a given KP sets u0 to this."""
    return kp / (1e-4 * (EPSILON_OX / tox))


def pulse_params(v1, v2, td, tr, tf, pw, per):
    return [v1, v2, td, tr, tf, pw, per]


# --------------------------------------------------------------------------
# C1: 64x64 resistor mesh, DC OP
# --------------------------------------------------------------------------
def resistor_mesh(nx: int = 64, ny: int = 64, seed: int = 0x52530000 + 1) -> Circuit:
    """Mesh of nx*ny nodes; node (i, j) has id 1 + i*nx + j.

    Device order: 1 Gohm from every node to ground (row-major, which fixes the
    node numbering), the 1 V source at (0,0), horizontal then vertical mesh
    resistors with R = 100 * 100**u, and 1 ohm from the far corner to ground.
    """
    rng = Lcg(seed)
    b = _Builder()
    nid = lambda i, j: 1 + i * nx + j  # noqa: E731
    for i in range(ny):
        for j in range(nx):
            b.add(RS_DEV_R, (nid(i, j), 0), sfx(1, 1e9))
    b.add(RS_DEV_V, (nid(0, 0), 0), 1.0)
    for i in range(ny):
        for j in range(nx - 1):
            b.add(RS_DEV_R, (nid(i, j), nid(i, j + 1)), 100.0 * 100.0 ** rng.uniform())
    for i in range(ny - 1):
        for j in range(nx):
            b.add(RS_DEV_R, (nid(i, j), nid(i + 1, j)), 100.0 * 100.0 ** rng.uniform())
    b.add(RS_DEV_R, (nid(ny - 1, nx - 1), 0), 1.0)
    return b.build(1 + nx * ny, f"mesh{nx}x{ny}")


# --------------------------------------------------------------------------
# C2 / C4: CMOS inverter chains with level-1 MOSFETs
# --------------------------------------------------------------------------
NMOS_KP, PMOS_KP = 200e-6, 100e-6


def inverter_chain(
    stages: int = 500,
    vin_pulse: bool = True,
    vin_dc: float = 0.0,
    vto_n=None,
    vto_p=None,
    u0_scale_n=None,
    u0_scale_p=None,
    load_c: float = sfx(5, F),
) -> Circuit:
    """Chain of `stages` inverters. Nodes: 0 ground, 1 vdd, 2 in, 3+i out_i.

    Devices: Vdd (1 V DC), Vin (PULSE(0 1 0 10p 10p 990p 2n) or DC vin_dc),
    then per stage: NMOS (d=out, g=prev, s=0, b=0), PMOS (d=out, g=prev,
    s=vdd, b=vdd, type=pmos), C load to ground. NMOS W/L = 1u/0.1u,
    PMOS W/L = 2u/0.1u, VTO = +/-0.4 V, KP mapped to u0 (u0_from_kp).
    Optional per-stage arrays perturb VTO and scale u0 (Monte-Carlo).
    """
    b = _Builder()
    b.add(RS_DEV_V, (1, 0), 1.0)
    if vin_pulse:
        b.add(RS_DEV_V, (2, 0), 0.0,
              pulse_params(0.0, 1.0, 0.0, sfx(10, P), sfx(10, P), sfx(990, P), sfx(2, N)),
              flags=RS_F_WAVEFORM)
    else:
        b.add(RS_DEV_V, (2, 0), vin_dc)
    u0n, u0p = u0_from_kp(NMOS_KP), u0_from_kp(PMOS_KP)
    wn, wp, l = sfx(1, U), sfx(2, U), sfx(0.1, U)
    for i in range(stages):
        out, prev = 3 + i, (2 if i == 0 else 2 + i)
        vn = 0.4 if vto_n is None else float(vto_n[i])
        vp = -0.4 if vto_p is None else float(vto_p[i])
        sn = 1.0 if u0_scale_n is None else float(u0_scale_n[i])
        sp = 1.0 if u0_scale_p is None else float(u0_scale_p[i])
        b.add(RS_DEV_M, (out, prev, 0, 0), 0.0, mos_params(vn, u0n * sn, wn, l))
        b.add(RS_DEV_M, (out, prev, 1, 1), 0.0, mos_params(vp, u0p * sp, wp, l), flags=RS_F_PMOS)
        b.add(RS_DEV_C, (out, 0), load_c)
    return b.build(3 + stages, f"inv{stages}")


#: .tran 5p 10n 0 5p of the 500-stage chain workload
C2_TRAN = dict(tstep=sfx(5, P), tstop=sfx(10, N), tstart=0.0, tmax=sfx(5, P))


class LcgStream:
    """The same LCG as :class:`Lcg`, drawn in vectorised blocks with a
    precomputed jump table (uint64 arithmetic wraps modulo 2**64)."""

    BLOCK = 1 << 16
    _A = None
    _C = None

    def __init__(self, seed: int):
        if LcgStream._A is None:
            a, c = 6364136223846793005, 1442695040888963407
            A = np.empty(self.BLOCK, np.uint64)
            Cc = np.empty(self.BLOCK, np.uint64)
            pa, pc = 1, 0
            for j in range(self.BLOCK):
                pa = (pa * a) & Lcg.MASK
                pc = (pc * a + c) & Lcg.MASK
                A[j], Cc[j] = pa, pc
            LcgStream._A, LcgStream._C = A, Cc
        self.state = np.uint64(seed & Lcg.MASK)
        self.buf = np.zeros(0)
        self.pos = 0

    def _refill(self):
        with np.errstate(over="ignore"):
            st = LcgStream._A * self.state + LcgStream._C
        self.state = st[-1]
        u = (st >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        self.buf = np.concatenate([self.buf[self.pos:], u])
        self.pos = 0

    def uniforms(self, k: int) -> np.ndarray:
        while len(self.buf) - self.pos < k:
            self._refill()
        out = self.buf[self.pos:self.pos + k]
        self.pos += k
        return out

    def normals(self, k: int) -> np.ndarray:
        """Box-Muller, one normal per (u1, u2) pair taken in stream order."""
        u = self.uniforms(2 * k).reshape(k, 2)
        return np.sqrt(-2.0 * np.log(1.0 - u[:, 0])) * np.cos(2.0 * np.pi * u[:, 1])


def _chain_records(stages, vin_dc, vn, sn, vp, sp, load_c=sfx(5, F), with_caps=True):
    """Vectorised inverter_chain(vin_pulse=False) records (same order)."""
    per = 3 if with_caps else 2
    arr = np.zeros(2 + per * stages, dtype=DEVICE_DTYPE)
    arr[0]["kind"], arr[0]["node"], arr[0]["value"] = RS_DEV_V, (1, 0, 0, 0), 1.0
    arr[1]["kind"], arr[1]["node"], arr[1]["value"] = RS_DEV_V, (2, 0, 0, 0), vin_dc
    i = np.arange(stages)
    out = 3 + i
    prev = np.where(i == 0, 2, 2 + i)
    mn = arr[2::per][:stages]
    mp = arr[3::per][:stages]
    u0n, u0p = u0_from_kp(NMOS_KP), u0_from_kp(PMOS_KP)
    wn, wp, l = sfx(1, U), sfx(2, U), sfx(0.1, U)
    mn["kind"] = RS_DEV_M
    mn["node"] = np.stack([out, prev, np.zeros_like(i), np.zeros_like(i)], 1)
    mn["p"][:, 0], mn["p"][:, 1], mn["p"][:, 2] = 1.0, vn, u0n * sn
    mn["p"][:, 3], mn["p"][:, 4], mn["p"][:, 5] = TOX_DEFAULT, wn, l
    mp["kind"], mp["flags"] = RS_DEV_M, RS_F_PMOS
    mp["node"] = np.stack([out, prev, np.ones_like(i), np.ones_like(i)], 1)
    mp["p"][:, 0], mp["p"][:, 1], mp["p"][:, 2] = 1.0, vp, u0p * sp
    mp["p"][:, 3], mp["p"][:, 4], mp["p"][:, 5] = TOX_DEFAULT, wp, l
    arr[2::per][:stages] = mn
    arr[3::per][:stages] = mp
    if with_caps:
        cc = arr[4::per][:stages]
        cc["kind"] = RS_DEV_C
        cc["node"] = np.stack([out, np.zeros_like(i), np.zeros_like(i), np.zeros_like(i)], 1)
        cc["value"] = load_c
        arr[4::per][:stages] = cc
    return arr


def monte_carlo_chains(count: int = 4096, seed: int = 0x52530000 + 4, lo: int = 50,
                       hi: int = 400):
    """Monte-Carlo inverter chains: stages ~ U[lo, hi], VTO += N(0, 20 mV),
    u0 *= N(1, 0.05) per MOSFET, DC input at 0 V. Draw order per instance:
    one uniform for the stage count, then per stage four normals (NMOS dVTO,
    NMOS u0 scale, PMOS dVTO, PMOS u0 scale), each from two uniforms."""
    rng = LcgStream(seed)
    out = []
    for _ in range(count):
        stages = lo + int(rng.uniforms(1)[0] * (hi - lo + 1))
        z = rng.normals(4 * stages).reshape(stages, 4)
        recs = _chain_records(stages, 0.0, 0.4 + 0.02 * z[:, 0], 1.0 + 0.05 * z[:, 1],
                              -0.4 + 0.02 * z[:, 2], 1.0 + 0.05 * z[:, 3])
        out.append(Circuit(recs, 3 + stages, f"mc_inv{stages}"))
    return out


def rc_mesh_inverters(nx: int = 256, ny: int = 256, taps: int | None = None,
                      seed: int = 0x52530000 + 5, load: float | None = None) -> Circuit:
    """RC mesh driven at seeded taps by level-1 inverters with a shared PULSE
    input (the adjoint workload; 256x256 with 2,048 taps at full size).

    Nodes: 0 ground, 1 vdd, 2 in, mesh node (i, j) = 3 + i*nx + j.
    Devices: Vdd (1 V DC); Vin PULSE(0 1 0 10p 10p 990p 2n); 1 Gohm from in
    to ground, which defines the input in the DC OP, where the PULSE source
    is not stamped (stamp.rs DC rule); horizontal then vertical mesh
    resistors R = 1 + 9u ohm; a capacitor C = (1 + 9u) fF from every mesh node
    to ground (row-major); then per tap, in draw order, an NMOS (d=tap, g=in,
    s=0, b=0) and a PMOS (d=tap, g=in, s=vdd, b=vdd) with the inverter-chain
    sizes. Taps (default nx*ny/32) are distinct mesh nodes drawn by a partial
    Fisher-Yates shuffle with the LCG. load: a resistor from every tap to
    ground after the inverters (none by default).
    """
    rng = Lcg(seed)
    taps = max(1, nx * ny // 32) if taps is None else taps
    b = _Builder()
    nid = lambda i, j: 3 + i * nx + j  # noqa: E731
    b.add(RS_DEV_V, (1, 0), 1.0)
    b.add(RS_DEV_V, (2, 0), 0.0,
          pulse_params(0.0, 1.0, 0.0, sfx(10, P), sfx(10, P), sfx(990, P), sfx(2, N)),
          flags=RS_F_WAVEFORM)
    b.add(RS_DEV_R, (2, 0), sfx(1, 1e9))
    for i in range(ny):
        for j in range(nx - 1):
            b.add(RS_DEV_R, (nid(i, j), nid(i, j + 1)), 1.0 + 9.0 * rng.uniform())
    for i in range(ny - 1):
        for j in range(nx):
            b.add(RS_DEV_R, (nid(i, j), nid(i + 1, j)), 1.0 + 9.0 * rng.uniform())
    for i in range(ny):
        for j in range(nx):
            b.add(RS_DEV_C, (nid(i, j), 0), (1.0 + 9.0 * rng.uniform()) * F)
    pool = list(range(nx * ny))
    u0n, u0p = u0_from_kp(NMOS_KP), u0_from_kp(PMOS_KP)
    wn, wp, l = sfx(1, U), sfx(2, U), sfx(0.1, U)
    for t in range(taps):
        k = t + int(rng.uniform() * (len(pool) - t))
        pool[t], pool[k] = pool[k], pool[t]
        node = 3 + pool[t]
        b.add(RS_DEV_M, (node, 2, 0, 0), 0.0, mos_params(0.4, u0n, wn, l))
        b.add(RS_DEV_M, (node, 2, 1, 1), 0.0, mos_params(-0.4, u0p, wp, l), flags=RS_F_PMOS)
    if load is not None:
        for t in range(taps):
            b.add(RS_DEV_R, (3 + pool[t], 0), load)
    return b.build(3 + nx * ny, f"rcmesh{nx}x{ny}t{taps}")


def power_grid(nx: int = 1024, ny: int = 1024, pitch: int = 64, loads: int | None = None,
               seed: int = 0x52530000 + 3) -> Circuit:
    """RC power grid: nx*ny nodes, segment R ~ U[0.05, 0.5] ohm, node C ~
    U[1, 10] fF to ground, a 1.0 V Vdd pad every `pitch` nodes in both
    directions, and `loads` current sources PULSE(0 10m td 0 0 20p) with td ~
    U[0, 100] ps drawing current from distinct seeded nodes to ground
    (default: one per 256 nodes, 4,096 at 1024x1024).

    Nodes: 0 ground, grid node (i, j) = 1 + i*nx + j. Devices, in this order
    (which makes first appearance row-major): the node capacitors (row-major),
    the horizontal then the vertical segment resistors, the pads (row-major
    over i, j multiples of `pitch`), then the loads in draw order. Draw order:
    the capacitor values, the horizontal then vertical resistor values, then
    per load one uniform for its node (a partial Fisher-Yates shuffle of the
    grid nodes) and one for its delay. At 1024x1024 the MNA has n = 1,048,833
    (ground, the nodes and one aux row per pad)."""
    loads = max(1, nx * ny // 256) if loads is None else loads
    rng = LcgStream(seed)
    nn = nx * ny
    nh, nv = ny * (nx - 1), (ny - 1) * nx
    pads = [(i, j) for i in range(0, ny, pitch) for j in range(0, nx, pitch)]
    arr = np.zeros(nn + nh + nv + len(pads) + loads, dtype=DEVICE_DTYPE)
    node = np.arange(nn, dtype=np.int32).reshape(ny, nx) + 1
    o = 0
    cc = arr[o:o + nn]
    cc["kind"] = RS_DEV_C
    cc["node"][:, 0] = node.ravel()
    cc["value"] = (1.0 + 9.0 * rng.uniforms(nn)) * F
    o += nn
    rh = arr[o:o + nh]
    rh["kind"] = RS_DEV_R
    rh["node"][:, 0] = node[:, :-1].ravel()
    rh["node"][:, 1] = node[:, 1:].ravel()
    rh["value"] = 0.05 + 0.45 * rng.uniforms(nh)
    o += nh
    rv = arr[o:o + nv]
    rv["kind"] = RS_DEV_R
    rv["node"][:, 0] = node[:-1, :].ravel()
    rv["node"][:, 1] = node[1:, :].ravel()
    rv["value"] = 0.05 + 0.45 * rng.uniforms(nv)
    o += nv
    for (i, j) in pads:
        arr[o]["kind"] = RS_DEV_V
        arr[o]["node"][0] = node[i, j]
        arr[o]["value"] = 1.0
        o += 1
    pool = np.arange(nn, dtype=np.int64)
    u = rng.uniforms(2 * loads).reshape(loads, 2)
    for t in range(loads):
        k = t + int(u[t, 0] * (nn - t))
        pool[t], pool[k] = pool[k], pool[t]
        arr[o]["kind"] = RS_DEV_I
        arr[o]["flags"] = RS_F_WAVEFORM
        arr[o]["node"][0] = 1 + pool[t]
        arr[o]["p"][:7] = pulse_params(0.0, sfx(10, 1e-3), sfx(100, P) * u[t, 1], 0.0, 0.0, sfx(20, P), 0.0)
        o += 1
    return Circuit(arr, 1 + nn, f"pgrid{nx}x{ny}")


#: 200 TRAP steps of 1 ps
C3_TRAN = dict(dt=sfx(1, P), nsteps=200)


#: 500 TRAP steps of 2 ps, checkpoint every 25, probe at the mesh centre
C5_TRAN = dict(dt=sfx(2, P), nsteps=500, every=25)


def rc_mesh_probe(nx: int, ny: int) -> int:
    return 3 + (ny // 2) * nx + nx // 2


def rc_ladder(n: int = 20, r: float = 1e3, c: float = 1e-12, vin: float = 1.0) -> Circuit:
    """Small RC ladder (V source, n series R, C from each tap to ground)."""
    b = _Builder()
    b.add(RS_DEV_V, (1, 0), vin)
    for i in range(n):
        b.add(RS_DEV_R, (1 + i, 2 + i), r)
        b.add(RS_DEV_C, (2 + i, 0), c)
    return b.build(2 + n, f"rc_ladder{n}")


def diode_circuit(v: float = 1.0, r: float = 1e3) -> Circuit:
    """V - R - diode to ground (is=1e-14, n=1)."""
    b = _Builder()
    b.add(RS_DEV_V, (1, 0), v)
    b.add(RS_DEV_R, (1, 2), r)
    b.add(RS_DEV_D, (2, 0), 0.0, [1e-14, 1.0])
    return b.build(3, "diode")


# --------------------------------------------------------------------------
# AC sweeps
# --------------------------------------------------------------------------
def ac_network(seed: int, nodes: int = 12, with_nonlinear: bool = True):
    """A seeded small-signal test network: ground, nodes 1..nodes; a 1 V
    source with AC 1 at node 1, a resistor chain plus random R and C, an AC
    current source, and (optionally) a diode and a level-1 CMOS stage.
    Returns (circuit, ac_mag, ac_phase), one magnitude and phase per device."""
    rng = np.random.RandomState(seed)
    b = _Builder()
    mag, ph = [], []

    def add(kind, nodes_, value, p=None, flags=0, m=0.0, phase=0.0):
        b.add(kind, nodes_, value, p, flags)
        mag.append(m)
        ph.append(phase)

    add(RS_DEV_V, (1, 0), 1.0, m=1.0, phase=0.0)
    for i in range(1, nodes):
        add(RS_DEV_R, (i, i + 1), float(10 ** rng.uniform(2, 4)))
    for _ in range(nodes):
        a, c = rng.randint(0, nodes + 1, 2)
        if a != c:
            add(RS_DEV_R, (int(a), int(c)), float(10 ** rng.uniform(2, 5)))
    for i in range(2, nodes + 1):
        add(RS_DEV_C, (i, 0), float(10 ** rng.uniform(-13, -10)))
        add(RS_DEV_R, (i, 0), 1e6)
    for _ in range(3):
        a, c = rng.randint(1, nodes + 1, 2)
        add(RS_DEV_C, (int(a), int(c)), float(10 ** rng.uniform(-13, -11)))
    add(RS_DEV_I, (0, nodes), 1e-4, m=float(rng.uniform(0.1, 1.0)), phase=float(rng.uniform(-90, 90)))
    if with_nonlinear:
        add(RS_DEV_D, (nodes // 2, 0), 0.0, [1e-14, 1.0])
        add(RS_DEV_M, (nodes - 1, 2, 0, 0), 0.0, mos_params(0.4, 500.0, 1e-6, 1e-7))
        add(RS_DEV_M, (nodes - 1, 2, 1, 1), 0.0, mos_params(-0.4, 250.0, 2e-6, 1e-7), flags=RS_F_PMOS)
    return b.build(nodes + 1, f"ac{seed}"), np.array(mag), np.array(ph)


def tap_bank_matrix(width: int = 2048, depth: int = 4, rails: int = 32, seed: int = 0x52530000 + 6):
    """Symmetric conductance matrix of two tap banks joined by a rail bus,
    shaped so that BbdSolver(2) (bbd.rs bfs_bisect) cuts exactly the bus:
    border = 2 * rails rows, each rail tapping every cell of its bank's edge
    row, so the border's coupling blocks C_k and Y_k = B_kk^-1 R_k are half
    dense (the dense Schur case, border <= 64, for rails <= 32).

    Nodes: 0 a lead of degree 1 (the search starts there), 1 the bank's feed,
    then bank A as `width` chains of `depth` cells (layer t is one search
    level), rails H_A (each tapping all of A's last layer), rails H_B (H_A[i]
    to H_B[i]), bank B (H_B taps its first layer), its feed and its lead.
    The half split lands between H_A and H_B because each side holds
    width * depth + 2 nodes. Branch conductances are uniform in [0.5, 2] S and
    every node has 0.01 S to ground. Returns (n, ap, ai, ax) in CSC with
    sorted rows and the diagonal present."""
    rs = np.random.RandomState(seed)
    W, D, w = int(width), int(depth), int(rails)
    a_lead, a_feed = 0, 1
    a0 = 2
    ha = a0 + W * D
    hb = ha + w
    b0 = hb + w
    b_feed = b0 + W * D
    b_lead = b_feed + 1
    n = b_lead + 1
    edges = [(a_lead, a_feed)]
    edges += [(a_feed, a0 + i) for i in range(W)]
    for t in range(D - 1):
        edges += [(a0 + t * W + i, a0 + (t + 1) * W + i) for i in range(W)]
    for r in range(w):
        edges += [(a0 + (D - 1) * W + i, ha + r) for i in range(W)]
        edges.append((ha + r, hb + r))
        edges += [(hb + r, b0 + i) for i in range(W)]
    for t in range(D - 1):
        edges += [(b0 + t * W + i, b0 + (t + 1) * W + i) for i in range(W)]
    edges += [(b0 + (D - 1) * W + i, b_feed) for i in range(W)]
    edges.append((b_feed, b_lead))
    e = np.array(edges, np.int64)
    g = rs.uniform(0.5, 2.0, len(e))
    diag = np.full(n, 0.01)
    np.add.at(diag, e[:, 0], g)
    np.add.at(diag, e[:, 1], g)
    rows = np.concatenate([e[:, 0], e[:, 1], np.arange(n)])
    cols = np.concatenate([e[:, 1], e[:, 0], np.arange(n)])
    vals = np.concatenate([-g, -g, diag])
    order = np.lexsort((rows, cols))
    rows, cols, vals = rows[order], cols[order], vals[order]
    ap = np.zeros(n + 1, np.int64)
    np.add.at(ap, cols + 1, 1)
    return n, np.cumsum(ap), rows, vals
