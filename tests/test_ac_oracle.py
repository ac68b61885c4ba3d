"""The AC oracle (oracle/orc_ac.c) pinned without the reference, on the CPU.

- generate_frequency_sweep against its closed forms (10^(i/10) per decade,
  doubling per octave, equal steps), and librsgpu.so's host copy of it
  (rs_ac_frequencies, no GPU needed) equal to the oracle's bit for bit;
- the RC lowpass of the reference's rc_filter.cir fixture against
  H = 1 / (1 + j w R C) at every frequency;
- random R/C/V/I/diode/level-1 MOSFET networks against numpy's complex solve
  of the admittance matrix built independently from the closed-form small-
  signal conductances (1e-9 relative);
- a floating subnetwork is singular at every frequency (nothing solved).
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from rustspice_amd import ac as A
from rustspice_amd import netlist as N
from rustspice_amd import netlists as NL

NETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "netlists")


@pytest.mark.parametrize("sweep,points,f0,f1", [("dec", 10, 1.0, 1e6), ("dec", 7, 3.0, 2e5), ("oct", 4, 1e3, 1e5),
                                                 ("lin", 11, 0.0, 1e3), ("lin", 1, 5.0, 9.0), ("dec", 10, 1e3, 1e3)])
def test_frequency_sweep(sweep, points, f0, f1):
    f = O.ac_frequencies(sweep, points, f0, f1)
    assert np.array_equal(f, A.frequencies(sweep, points, f0, f1))
    if sweep == "dec":
        dec = math.log10(f1) - math.log10(f0)
        assert len(f) == math.ceil(dec * points) + 1
        ref = f0 * 10.0 ** (np.arange(len(f)) * dec / max(len(f) - 1, 1))
        assert np.allclose(f, ref, rtol=1e-13, atol=0)
    elif sweep == "oct":
        octs = math.log2(f1 / f0)
        assert len(f) == math.ceil(octs * points) + 1
        assert np.allclose(f, f0 * 2.0 ** (np.arange(len(f)) * octs / (len(f) - 1)), rtol=1e-13)
    else:
        assert len(f) == points
        assert np.allclose(f, f0 + np.arange(points) * (f1 - f0) / max(points - 1, 1), rtol=1e-15)


def test_rc_lowpass_closed_form():
    low = N.lower(N.load(path=os.path.join(NETS, "rc_filter.cir")))
    x, r = O.dc_op(low.circuit)
    assert r.converged
    f = O.ac_frequencies("dec", 10, 1.0, 1e6)
    xa, mp, n = O.ac(low.ac_circuit, low.ac_mag, low.ac_phase, x, f)
    assert n == 4 and len(xa) == len(f)
    h = 1.0 / (1.0 + 1j * 2 * np.pi * f * 1e3 * 1e-6)
    assert np.abs(xa[:, 2] - h).max() < 1e-14 and np.abs(xa[:, 1] - 1.0).max() < 1e-15
    assert np.allclose(mp[:, 2, 0], 20 * np.log10(np.abs(h)), rtol=1e-12, atol=1e-12)
    assert np.allclose(mp[:, 2, 1], np.degrees(np.angle(h)), rtol=1e-12, atol=1e-12)
    # ground: 0, up to round-off once |w C| > 1 makes the 'out' row the first pivot
    assert np.abs(xa[:, 0]).max() < 1e-15 and np.all(mp[f < 1e5, 0, 0] == -600.0)
    k = int(np.argmin(np.abs(f - 1.0 / (2 * np.pi * 1e-3))))
    assert abs(mp[k, 2, 1] + 45.0) < 3.0  # phase near -45 degrees around fc = 159 Hz


def random_ac_circuit(seed, nodes=12, with_nonlinear=True):
    return NL.ac_network(seed, nodes, with_nonlinear)


def independent_solve(circ, mag, ph, x, freq):
    """Y(jw) V = I from the closed-form small-signal values, numpy's solve."""
    d = circ.devices
    nv = int(np.count_nonzero(d["kind"] == NL.RS_DEV_V))
    n = circ.node_count + nv
    Y = np.zeros((n, n), complex)
    I = np.zeros(n, complex)
    w = 2 * np.pi * freq
    k = circ.node_count
    xv = lambda i: x[i] if i < len(x) else 0.0  # noqa: E731

    def four(a, b, g):
        Y[a, a] += g; Y[b, b] += g; Y[a, b] -= g; Y[b, a] -= g

    for rec, m, p in zip(d, mag, ph):
        a, c = int(rec["node"][0]), int(rec["node"][1])
        kind = rec["kind"]
        if kind == NL.RS_DEV_R:
            four(a, c, 1.0 / rec["value"])
        elif kind == NL.RS_DEV_C:
            four(a, c, 1j * w * rec["value"])
        elif kind == NL.RS_DEV_V:
            Y[a, k] += 1; Y[c, k] -= 1; Y[k, a] += 1; Y[k, c] -= 1
            I[k] += m * np.exp(1j * np.radians(p))
            k += 1
        elif kind == NL.RS_DEV_I:
            v = m * np.exp(1j * np.radians(p))
            I[a] -= v; I[c] += v
        elif kind == NL.RS_DEV_D:
            vt = 0.02585 * rec["p"][1]
            four(a, c, max(rec["p"][0] / vt * math.exp((xv(a) - xv(c)) / vt), 1e-12))
        elif kind == NL.RS_DEV_M:
            dr, g, s = (int(q) for q in rec["node"][:3])
            out = O.level1(rec, xv(dr), xv(g), xv(s))
            gm, gds = out["gm"], max(out["gds"], 1e-12)
            four(dr, s, gds)
            Y[dr, g] += gm; Y[dr, s] -= gm; Y[s, g] -= gm; Y[s, s] += gm
    Y[0, :] = 0
    Y[0, 0] = 1
    I[0] = 0
    return np.linalg.solve(Y, I)


@pytest.mark.parametrize("seed", range(6))
def test_random_networks_against_numpy(seed):
    circ, mag, ph = random_ac_circuit(seed, nodes=8 + 4 * seed, with_nonlinear=seed % 2 == 0)
    # the small-signal values are taken at a given operating point: any x
    x = np.random.RandomState(100 + seed).uniform(0.0, 1.0, circ.node_count + 1)
    f = O.ac_frequencies("dec", 5, 10.0, 1e10)
    xa, mp, n = O.ac(circ, mag, ph, x, f)
    assert len(xa) == len(f)
    for i, fr in enumerate(f):
        ref = independent_solve(circ, mag, ph, x, fr)
        scale = np.abs(ref).max()
        assert np.abs(xa[i] - ref).max() <= 1e-9 * scale, (seed, fr)
    db = 20 * np.log10(np.maximum(np.abs(xa[:, : circ.node_count]), 1e-300))
    ok = np.abs(xa[:, : circ.node_count]) > 1e-30
    assert np.allclose(mp[..., 0][ok], db[ok], rtol=1e-12, atol=1e-9)


def test_floating_node_is_singular_at_every_frequency():
    b = NL._Builder()
    b.add(NL.RS_DEV_V, (1, 0), 1.0)
    b.add(NL.RS_DEV_R, (1, 0), 1e3)
    b.add(NL.RS_DEV_C, (2, 3), 1e-12)  # nodes 2, 3: no path to ground
    circ = b.build(4, "floating")
    xa, mp, n = O.ac(circ, [1.0, 0, 0], [0.0, 0, 0], np.zeros(5), O.ac_frequencies("dec", 2, 1.0, 100.0))
    assert len(xa) == 0 and n == 5
