"""rs_tran's adaptive loop on the device controller (rs_circuit_set_tran_graph)
against the host-driven loop and the oracle: accepted and rejected steps,
Newton iterations, every stored time and solution bit for bit, on PULSE-driven
RC circuits under NativeKLU and SparseLU semantics (these runs accept 13-16
steps, reject some, and end the way the reference's loop ends: ten
consecutive rejections)."""
import numpy as np
import pytest

from oracle import oracle as O
from rustspice_amd import engine as E
from rustspice_amd import netlists as NL
from rustspice_amd import solver as S
from tests.helpers import same_bits

pytestmark = pytest.mark.gpu


def pulse_rc(c=1e-13):
    b = NL._Builder()
    b.add(NL.RS_DEV_V, (1, 0), 0.0, NL.pulse_params(0.0, 1.0, 1e-10, 5e-11, 5e-11, 3e-10, 1e-9),
          flags=NL.RS_F_WAVEFORM)
    b.add(NL.RS_DEV_R, (1, 2), 1e3)
    b.add(NL.RS_DEV_C, (2, 0), c)
    b.add(NL.RS_DEV_R, (2, 0), 1e4)
    return b.build(3, "pulse_rc")


def pulse_ladder(n=10):
    b = NL._Builder()
    b.add(NL.RS_DEV_V, (1, 0), 0.0, NL.pulse_params(0.0, 1.0, 1e-10, 5e-11, 5e-11, 3e-10, 1e-9),
          flags=NL.RS_F_WAVEFORM)
    for i in range(n):
        b.add(NL.RS_DEV_R, (1 + i, 2 + i), 1e3)
        b.add(NL.RS_DEV_C, (2 + i, 0), 1e-13)
    return b.build(2 + n, f"pladder{n}")


CASES = {
    "pulse_rc": (pulse_rc, dict(tstep=2e-11, tstop=1e-9, tmax=2e-11)),
    "pulse_rc_long": (pulse_rc, dict(tstep=2e-11, tstop=3e-9, tmax=1e-10)),
    "ladder10": (lambda: pulse_ladder(10), dict(tstep=2e-11, tstop=3e-9, tmax=1e-10)),
    "ladder30": (lambda: pulse_ladder(30), dict(tstep=1e-11, tstop=3e-9, tmax=5e-11)),
}


@pytest.mark.parametrize("kind", [S.NATIVEKLU, S.SPARSELU])
@pytest.mark.parametrize("name", list(CASES))
def test_device_controller_matches_host_loop_and_oracle(gpu, name, kind):
    print(f"\n[tran-device] {name} kind {kind}", flush=True)
    make, args = CASES[name]
    c = make()
    host = E.GpuCircuit(c).tran(S.GpuSolver(c.node_count, kind), **args)
    gc = E.GpuCircuit(c)
    gc.set_tran_graph(True)
    dev = gc.tran(S.GpuSolver(c.node_count, kind), **args)
    s = O.NativeKlu(c.node_count)
    if kind == S.SPARSELU:
        s.set_sparselu()
    ref = O.tran(c, solver=s, **args)
    info = gc.tran_graph_info()
    assert info["runs"] == 1 and info["launches"] > 0, info
    assert dev["accepted"] > 1
    for k in ("status", "accepted", "rejected", "newton_iters_total", "factor_calls", "dc_iters", "npoints"):
        assert dev[k] == host[k] == ref[k], (k, dev[k], host[k], ref[k])
    assert same_bits(dev["times"], host["times"]) and same_bits(dev["times"], ref["times"])
    assert same_bits(dev["sols"], host["sols"]) and same_bits(dev["sols"], ref["sols"])
