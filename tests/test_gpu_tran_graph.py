"""Fixed-step TRAP transients with every step after the first as a captured
hipGraph (rs_circuit_set_tran_graph; no reference counterpart): the device
Newton update must reproduce the host-driven loop (newton.rs:65-117 via
engine.cpp's run_newton) bit for bit, including its damping, and fall back
to it where the graph path does not apply."""
import numpy as np
import pytest

from oracle import oracle as O
from rustspice_amd import engine as E
from rustspice_amd import netlists as NL
from rustspice_amd import solver as S
from tests.helpers import same_bits

pytestmark = pytest.mark.gpu

CASES = {
    "rcmesh8": (lambda: NL.rc_mesh_inverters(8, 8), 2e-12, 60),
    "rcmesh8_loaded": (lambda: NL.rc_mesh_inverters(8, 8, 4, load=1e4), 2e-12, 60),
    "ladder": (lambda: NL.rc_ladder(30), 1e-10, 50),
    "diode": (NL.diode_circuit, 1e-9, 20),
}


def _run(c, dt, n, graph, kind=S.NATIVEKLU, opts=None):
    g = E.GpuCircuit(c)
    g.set_tran_graph(graph)
    s = S.GpuSolver(c.node_count, kind)
    st, x, it = g.tran_fixed(s, dt, n, opts)
    return st, x, it, g.graph_steps()


@pytest.mark.parametrize("kind", [S.NATIVEKLU, S.SPARSELU])
@pytest.mark.parametrize("name", sorted(CASES))
def test_graph_steps_equal_host_driven_steps(gpu, name, kind):
    make, dt, n = CASES[name]
    c = make()
    st0, x0, it0, gs0 = _run(c, dt, n, False, kind)
    st1, x1, it1, gs1 = _run(c, dt, n, True, kind)
    assert gs0 == 0
    assert st1 == st0 and it1 == it0
    assert same_bits(x1, x0)
    if st0 == 0:
        assert gs1 == n - 1


def test_graph_steps_match_the_oracle(gpu):
    """The graph path against the CPU restatement of the same fixed-step
    transient (reference semantics, NativeKLU)."""
    c = NL.rc_mesh_inverters(8, 8)
    st, x, it, gs = _run(c, 2e-12, 40, True)
    rc, xo, ito = O.tran_fixed(c, 2e-12, 40)
    assert st == 0 and rc == 0 and gs == 39
    assert it == ito
    assert same_bits(x, xo[: len(x)])


def test_damped_and_failing_steps(gpu):
    """Tight tolerances with few iterations: steps that run out of Newton
    iterations stop the transient exactly where the host loop stops."""
    c = NL.rc_mesh_inverters(8, 8)
    o = E.default_opts()
    o.max_iters, o.abs_tol, o.rel_tol = 3, 1e-15, 1e-15
    r0 = _run(c, 2e-12, 30, False, opts=o)
    r1 = _run(c, 2e-12, 30, True, opts=o)
    assert r1[0] == r0[0] and r1[2] == r0[2]
    assert same_bits(r1[1], r0[1])
