"""CPU checks of librsgpu.so's host side: it loads, exports every symbol the
public headers declare, and its integer analysis (AMD, BTF, L/U pattern) and
first pivoting factor equal the oracle's bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from rustspice_amd import _lib
from rustspice_amd import netlists as NL
from rustspice_amd import solver as S
from tests.helpers import random_csc, same_bits

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for h in ("rsgpu.h", "rsgpu_host.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(rs_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _declared_symbols()
    assert declared, "no declarations parsed"
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert set(_lib.EXPORTS) == declared


def _cmp_factor(ap, ai, ax, n, kind):
    o = O.NativeKlu(n)
    if kind == S.SPARSELU:
        o.set_sparselu()
    assert o.factor(ap, ai, ax) == 0
    h = S.HostFactor(n, ap, ai, ax, kind=kind)
    assert h.nblocks() == o.stats()["nblocks"]
    for b in range(h.nblocks()):
        a, c = o.block(b), h.block(b)
        assert a["n"] == c["n"]
        for k in ("col_perm", "l_col_ptr", "l_row_idx", "u_col_ptr", "u_row_idx"):
            assert np.array_equal(a[k], c[k]), (b, k)
        if a["n"] > 1:
            assert np.array_equal(a["row_perm"], c["row_perm"])
            for k in ("l_values", "u_values", "u_diag"):
                assert same_bits(a[k], c[k]), (b, k)
        else:
            assert same_bits(a["u_diag"], c["u_diag"])
    return h


def test_random_matrices_match_oracle():
    rs = np.random.RandomState(1)
    for _ in range(300):
        n = int(rs.randint(1, 60))
        ap, ai = random_csc(rs, n, rs.rand() * 0.3)
        pa, _ = O.amd_order(n, ap, ai)
        pb, _ = S.host_amd_order(n, ap, ai)
        assert np.array_equal(pa, pb)
        ba, bb = O.btf_decompose(n, ap, ai), S.host_btf(n, ap, ai)
        for k in ("row_perm", "col_perm", "block_ptr"):
            assert np.array_equal(ba[k], bb[k]), k
        assert ba["structural_rank"] == bb["structural_rank"]
        ax = rs.randn(len(ai))
        for kind in (S.NATIVEKLU, S.SPARSELU):
            _cmp_factor(ap, ai, ax, n, kind)


@pytest.mark.parametrize("kind", [S.NATIVEKLU, S.SPARSELU])
def test_mesh_matrix_matches_oracle(kind):
    c = NL.resistor_mesh()
    ap, ai, ax, rhs, n = O.build_mna(c)
    h = _cmp_factor(ap, ai, ax, n, kind)
    ps = h.plan_stats()
    assert ps["nnz_l"] == 70589 and ps["nnz_u"] == 74687


def test_chain_matrices_match_oracle():
    c = NL.inverter_chain(500)
    rs = np.random.RandomState(0)
    ap, ai, ax, rhs, n = O.build_mna(c, x=rs.rand(c.node_count))
    h = _cmp_factor(ap, ai, ax, n, S.NATIVEKLU)
    assert h.nblocks() == 2  # trailing structurally singular singleton (Q14)
    for inst in NL.monte_carlo_chains(6):
        ap, ai, ax, rhs, n = O.build_mna(inst, x=rs.rand(inst.node_count))
        _cmp_factor(ap, ai, ax, n, S.NATIVEKLU)


def test_amd_large_ground_node_fast_path():
    """The ground node touches every node; the product's degree shortcut for it
    must still give the reference order (64x64 mesh and a 128x128 mesh)."""
    for nx in (64, 128):
        c = NL.resistor_mesh(nx, nx)
        ap, ai, ax, rhs, n = O.build_mna(c)
        pa, _ = O.amd_order(n, ap, ai)
        pb, _ = S.host_amd_order(n, ap, ai)
        assert np.array_equal(pa, pb)


def test_gpu_entry_points_fail_loudly_without_device():
    """No CPU fallback: without a GPU the solver reports a device error."""
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except Exception:
        pass
    g = S.GpuSolver(2)
    with pytest.raises(S.SolverError) as e:
        g.factor([0, 2, 3], [0, 1, 1], [2.0, 1.0, 3.0])
    assert e.value.variant == "DeviceError"


def test_stamp_pattern_matches_oracle_mna():
    """The product's symbolic replay of the stamp loop gives the engine's
    unmerged CSC pattern (mna.rs:320-347) in DC and TRAN modes."""
    rs = np.random.RandomState(2)
    circs = [NL.resistor_mesh(), NL.inverter_chain(500), NL.rc_ladder(20), NL.diode_circuit()]
    circs += NL.monte_carlo_chains(3)
    for c in circs:
        for mode in (0, 1):
            ts = None
            if mode:
                ts = (rs.rand(c.ndev), rs.rand(c.ndev), np.ones(c.ndev), np.ones(c.ndev), 1)
            ap, ai, _, _, n = O.build_mna(c, mode=mode, t=1e-12, dt=1e-12, tstate=ts)
            pa, pi = S.host_stamp_pattern(c, mode)
            assert np.array_equal(ap, pa) and np.array_equal(ai, pi), (c.name, mode)


@pytest.mark.parametrize("nx", [32, 80])
def test_amd_power_grid_matches_oracle(nx):
    """The RC power grid's TRAN pattern (every node tied to ground through its
    capacitor, pads with aux rows): the product's element shortcuts (an
    element inside the new one skipped, a single outside element counted by
    its size) give the literal restatement's order."""
    c = NL.power_grid(nx, nx, pitch=16)
    ap, ai = S.host_stamp_pattern(c, mode=1)
    n = len(ap) - 1
    pa, _ = O.amd_order(n, ap, ai)
    pb, _ = S.host_amd_order(n, ap, ai)
    assert np.array_equal(pa, pb)


def test_amd_random_grids_with_cliques_match_oracle():
    """Random sparse graphs plus a few dense rows and cliques, where elements
    overlap L_p partly, wholly or not at all."""
    rs = np.random.RandomState(11)
    for _ in range(60):
        n = int(rs.randint(20, 200))
        M = rs.rand(n, n) < rs.rand() * 0.05
        for _ in range(int(rs.randint(0, 4))):  # cliques
            k = rs.choice(n, size=int(rs.randint(3, 12)), replace=False)
            M[np.ix_(k, k)] = True
        if rs.rand() < 0.5:  # a dense row/column, like ground
            g = int(rs.randint(n))
            M[g, :] = M[:, g] = True
        M = M | M.T
        np.fill_diagonal(M, True)
        cols = [np.nonzero(M[:, j])[0] for j in range(n)]
        ap = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int64)
        ai = np.concatenate(cols).astype(np.int64)
        pa, _ = O.amd_order(n, ap, ai)
        pb, _ = S.host_amd_order(n, ap, ai)
        assert np.array_equal(pa, pb)
