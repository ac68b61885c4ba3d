"""DC adjoint sensitivities (no reference counterpart) against central finite
differences of the DC operating point and a dense FP64 CPU adjoint.

The adjoint differentiates the exact root of F(x, p) = A(x) x - rhs(x), the
reference's stamped currents (stamp.rs). The operating points for the finite
differences are therefore polished to that root: the oracle's DC OP (the
reference Newton loop with SparseLU) first, then dense Newton steps with a
numerical Jacobian of the oracle's F. The polish matters: with a voltage
source whose aux pivot SparseLU clamps to 1e-14 (sparse_lu.rs:493-499), the
reference loop stops at a point whose node voltages are off by ~1e-3 V (the
inverter case below), so finite differences of the unpolished solver measure
that solver, not the circuit.
"""
import numpy as np
import pytest

from oracle import oracle as O
from rustspice_amd import engine as E
from rustspice_amd import netlists as NL
from tests.helpers import csc_dense

pytestmark = pytest.mark.gpu

U0N, U0P = NL.u0_from_kp(NL.NMOS_KP), NL.u0_from_kp(NL.PMOS_KP)


def _cs_amp():
    b = NL._Builder()
    b.add(NL.RS_DEV_V, (1, 0), 1.8)
    b.add(NL.RS_DEV_V, (2, 0), 1.0)
    b.add(NL.RS_DEV_R, (1, 3), 2e3)
    b.add(NL.RS_DEV_M, (3, 2, 0, 0), 0.0, NL.mos_params(0.4, U0N, 1e-6, 0.1e-6))
    b.add(NL.RS_DEV_R, (3, 0), 1e5)
    b.add(NL.RS_DEV_I, (3, 0), 1e-5)
    return b.build(4, "cs_amp")


def _inv1():
    b = NL._Builder()
    b.add(NL.RS_DEV_V, (1, 0), 1.0)
    b.add(NL.RS_DEV_V, (2, 0), 0.2)  # a smooth bias: finite differences agree to 1e-8 here
    b.add(NL.RS_DEV_M, (3, 2, 0, 0), 0.0, NL.mos_params(0.4, U0N, 1e-6, 0.1e-6))
    b.add(NL.RS_DEV_M, (3, 2, 1, 1), 0.0, NL.mos_params(-0.4, U0P, 2e-6, 0.1e-6), flags=NL.RS_F_PMOS)
    b.add(NL.RS_DEV_R, (3, 0), 1e4)
    return b.build(4, "inv1")


def _pmos_cs():
    b = NL._Builder()
    b.add(NL.RS_DEV_V, (1, 0), 1.8)
    b.add(NL.RS_DEV_V, (2, 0), 0.8)
    b.add(NL.RS_DEV_M, (3, 2, 1, 1), 0.0, NL.mos_params(-0.4, U0P, 2e-6, 0.1e-6), flags=NL.RS_F_PMOS)
    b.add(NL.RS_DEV_R, (3, 0), 2e3)
    return b.build(4, "pmos_cs")


def _ladder(stages=12):
    """V - R - R - ... with a shunt R and C to ground at every node (DC current flows)."""
    b = NL._Builder()
    b.add(NL.RS_DEV_V, (1, 0), 1.0)
    for k in range(stages):
        b.add(NL.RS_DEV_R, (1 + k, 2 + k), 100.0 * (1 + k))
        b.add(NL.RS_DEV_R, (2 + k, 0), 1e4 / (1 + 0.1 * k))
        b.add(NL.RS_DEV_C, (2 + k, 0), 1e-12)
    return b.build(stages + 2, "ladder")


CIRCS = {"ladder": _ladder, "mesh": lambda: NL.resistor_mesh(6, 6),
         "diode": NL.diode_circuit, "cs_amp": _cs_amp, "inv1": _inv1, "pmos_cs": _pmos_cs}


def _dc(c):
    s = O.NativeKlu(c.node_count + 8)
    s.set_sparselu()
    o = O.default_opts()
    o.abs_tol, o.rel_tol, o.max_iters = 1e-15, 1e-13, 400
    x, r = O.dc_op(c, opts=o, solver=s)
    if x is None:  # tolerances beyond round-off: take the last iterate of a looser run
        o.abs_tol, o.rel_tol = 1e-13, 1e-11
        x, r = O.dc_op(c, opts=o, solver=s)
    assert x is not None, c.name
    return x


def _root(c, x=None):
    """Exact root of the oracle's F(x) = A(x) x - rhs(x): dense Newton with a
    central-difference Jacobian, from the reference DC OP."""
    x = _dc(c) if x is None else x.copy()
    nc = c.node_count

    def F(v):
        ap, ai, ax, rhs, n = O.build_mna(c, x=v[:nc], mode=0)
        return csc_dense(ap, ai, ax, n) @ v - rhs

    for _ in range(40):
        n = len(x)
        J = np.zeros((n, n))
        for j in range(n):
            h = 1e-7 * max(1.0, abs(x[j]))
            xp, xm = x.copy(), x.copy()
            xp[j] += h
            xm[j] -= h
            J[:, j] = (F(xp) - F(xm)) / (2.0 * h)
        dx = np.linalg.solve(J, -F(x))
        x = x + dx
        if np.abs(dx).max() <= 1e-15 * max(1.0, np.abs(x).max()):
            break
    return x


def _param(c, k):
    """(getter, setter) of the differentiated parameter of device k."""
    d = c.devices[k]
    if d["kind"] == NL.RS_DEV_M:
        return lambda: float(d["p"][4]), lambda v: d["p"].__setitem__(4, v)
    if d["kind"] == NL.RS_DEV_D:
        return lambda: float(d["p"][0]), lambda v: d["p"].__setitem__(0, v)
    return lambda: float(d["value"]), lambda v: d.__setitem__("value", v)


@pytest.mark.parametrize("name", sorted(CIRCS))
def test_adjoint_matches_central_differences(gpu, name):
    c = CIRCS[name]()
    x = _root(c)
    probe = c.node_count // 2 if name == "mesh" else c.node_count - 1
    g = E.GpuCircuit(c)
    dp, lam = g.dc_adjoint(x, probe)
    h_rel = 1e-6
    checked = 0
    for k in range(c.ndev):
        if c.devices[k]["kind"] == NL.RS_DEV_C:
            assert dp[k] == 0.0
            continue
        get, put = _param(c, k)
        p0 = get()
        h = h_rel * abs(p0) if p0 != 0.0 else h_rel
        put(p0 + h)
        xp = _root(c, x)
        put(p0 - h)
        xm = _root(c, x)
        put(p0)
        fd = (xp[probe] - xm[probe]) / (2.0 * h)
        # resolution of the difference quotient: solution round-off over the step
        floor = 1e-7 * max(1.0, abs(x[probe])) / max(abs(p0), 1e-30)
        assert abs(dp[k] - fd) <= 1e-6 * max(abs(fd), abs(dp[k])) + floor, (name, k, dp[k], fd)
        checked += abs(fd) > 100.0 * floor
    assert checked >= 3


@pytest.mark.parametrize("name", ["ladder", "mesh", "diode", "cs_amp"])
def test_lambda_matches_dense_cpu_adjoint(gpu, name):
    c = CIRCS[name]()
    x = _root(c)
    probe = c.node_count // 2 if name == "mesh" else c.node_count - 1
    ap, ai, ax, rhs, n = O.build_mna(c, x=x[: c.node_count], mode=0)
    A = csc_dense(ap, ai, ax, n)
    e = np.zeros(n)
    e[probe] = 1.0
    lam_cpu = np.linalg.solve(A.T, e)
    _, lam = E.GpuCircuit(c).dc_adjoint(x, probe)
    assert np.allclose(lam, lam_cpu, rtol=1e-9, atol=1e-12 * np.abs(lam_cpu).max())
