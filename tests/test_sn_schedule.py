"""Supernodal factor and solve schedule (the order of the GPU kernels) run on
the host against the oracle's SparseLU arithmetic (sparse_lu.rs, restated
in oracle/orc_klu.c): factor values, pivots and the solution bit for bit."""
import numpy as np
import pytest

from oracle import oracle as O
from rustspice_amd import netlists as NL
from rustspice_amd import solver as S
from tests.helpers import same_bits


def _check(ap, ai, ax, rhs, n, hot_min=0):
    o = O.NativeKlu(n)
    o.set_sparselu()
    assert o.factor(ap, ai, ax) == 0
    xo = rhs.copy()
    assert o.solve(xo) == 0
    want = o.block(0)
    x, lx, ux, ud, st = S.host_sn_selftest(n, ap, ai, ax, rhs, hot_min)
    assert same_bits(lx, want["l_values"])
    assert same_bits(ux, want["u_values"])
    assert same_bits(ud, want["u_diag"])
    assert same_bits(x, xo)
    return st


def _tran_mna(c, seed):
    rs = np.random.RandomState(seed)
    cv, ci = rs.rand(c.ndev), rs.randn(c.ndev) * 1e-4
    hv, hi = np.ones(c.ndev, np.uint8), np.ones(c.ndev, np.uint8)
    return O.build_mna(c, x=rs.rand(c.node_count), mode=1, t=3e-12, dt=1e-12,
                       tstate=(cv, ci, hv, hi, 1))


@pytest.mark.parametrize("nx,pitch", [(8, 4), (16, 8), (24, 8), (40, 16), (72, 32)])
@pytest.mark.parametrize("hot_min", [0, 8, 40])
def test_power_grid_tran_matrix(nx, pitch, hot_min):
    """hot_min sets how many supernodes must hold a position for its diagonal
    to be summed on its own: small values exercise that path (the ground
    node's diagonal at every size, separator nodes at the smallest)."""
    c = NL.power_grid(nx, nx, pitch=pitch)
    ap, ai, ax, rhs, n = _tran_mna(c, nx)
    st = _check(ap, ai, ax, rhs, n, hot_min)
    assert st["levels"] < st["supernodes"]
    if hot_min and nx >= 40:
        assert st["hot_entries"] >= 1


def test_resistor_mesh_dc_matrix():
    c = NL.resistor_mesh(20, 20)
    ap, ai, ax, rhs, n = O.build_mna(c)
    _check(ap, ai, ax, rhs, n, hot_min=6)


def test_random_symmetric_patterns_with_zeros_and_clamps():
    """Random structurally symmetric matrices, some values exactly zero
    (the zero-multiplier skip) and some pivots below the 1e-14 clamp."""
    rs = np.random.RandomState(5)
    for _ in range(80):
        n = int(rs.randint(2, 70))
        M = rs.rand(n, n) < rs.rand() * 0.2
        M = M | M.T
        np.fill_diagonal(M, rs.rand(n) < 0.9)
        cols = [np.nonzero(M[:, j])[0] for j in range(n)]
        ap = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int64)
        ai = np.concatenate(cols + [np.zeros(0, np.int64)]).astype(np.int64)
        ax = rs.randn(len(ai))
        ax[rs.rand(len(ai)) < 0.15] = 0.0
        ax[rs.rand(len(ai)) < 0.05] = 1e-16
        rhs = rs.randn(n)
        _check(ap, ai, ax, rhs, n)


def _csc(M, rs):
    cols = [np.nonzero(M[:, j])[0] for j in range(M.shape[0])]
    ap = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int64)
    ai = np.concatenate(cols + [np.zeros(0, np.int64)]).astype(np.int64)
    ax = rs.randn(len(ai))
    ax[rs.rand(len(ai)) < 0.15] = 0.0
    ax[rs.rand(len(ai)) < 0.05] = 1e-16
    return ap, ai, ax


def _transpose(ap, ai, ax, n):
    """CSC of A^T with duplicates kept in A's CSC order."""
    cols = np.repeat(np.arange(n), np.diff(ap))
    order = np.lexsort((np.arange(len(ai)), ai))  # by row of A, stable
    bp = np.concatenate([[0], np.cumsum(np.bincount(ai, minlength=n))]).astype(np.int64)
    return bp, cols[order].astype(np.int64), ax[order]


def test_random_unsymmetric_patterns_on_the_symmetric_closure():
    """Random structurally unsymmetric matrices: the plan holds the symmetric
    closure of the fill (explicit zeros) and still reproduces SparseLU's
    factor and solve bit for bit, zero values and clamped pivots included."""
    rs = np.random.RandomState(11)
    seen = 0
    for _ in range(120):
        n = int(rs.randint(2, 70))
        M = rs.rand(n, n) < rs.rand() * 0.2
        np.fill_diagonal(M, rs.rand(n) < 0.9)
        ap, ai, ax = _csc(M, rs)
        rhs = rs.randn(n)
        st = _check(ap, ai, ax, rhs, n)
        seen += st["symmetric"] == 0
    assert seen > 60


def test_inverter_tapped_mesh_tran_matrix():
    """MOSFET stamps (gm at (drain, gate) without (gate, drain)) make the
    inverter-tapped RC mesh structurally unsymmetric."""
    for nx, taps, seed in [(12, 9, 3), (20, 12, 4)]:
        c = NL.rc_mesh_inverters(nx, nx, taps)
        ap, ai, ax, rhs, n = _tran_mna(c, seed)
        st = _check(ap, ai, ax, rhs, n, hot_min=8)
        assert st["symmetric"] == 0
    ap = np.array([0, 2, 3], np.int64)
    ai = np.array([0, 1, 1], np.int64)
    _check(ap, ai, np.array([2.0, -1.0, 3.0]), np.array([1.0, 2.0]), 2)


def test_transposed_plan_solves_the_transposed_system():
    """The plan of A^T on A's symbolic factor (the adjoint's factor): where
    A^T's own SparseLU symbolic picks the same order, its solution equals
    SparseLU on the explicit transpose bit for bit; always within round-off
    of a dense solve."""
    rs = np.random.RandomState(12)
    same_order = 0
    cases = []
    for _ in range(60):
        n = int(rs.randint(2, 60))
        M = rs.rand(n, n) < rs.rand() * 0.2
        np.fill_diagonal(M, True)
        ap, ai, ax = _csc(M, rs)
        ax[np.repeat(np.arange(n), np.diff(ap)) == ai] += 4.0  # a dominant diagonal: a tight dense check
        cases.append((ap, ai, ax, n, True))
    c = NL.rc_mesh_inverters(12, 12, 9)  # ground kept: singular but for the clamped pivot
    ap, ai, ax, _, n = _tran_mna(c, 5)
    cases.append((ap, ai, ax, n, False))
    for ap, ai, ax, n, dense in cases:
        rhs = rs.randn(n)
        x, _, _, _, _ = S.host_sn_selftest(n, ap, ai, ax, rhs, transpose=True)
        if dense:
            A = np.zeros((n, n))
            np.add.at(A, (ai, np.repeat(np.arange(n), np.diff(ap))), ax)
            xd = np.linalg.solve(A.T, rhs)
            assert np.allclose(x, xd, rtol=1e-8, atol=1e-8 * np.abs(xd).max())
        bp, bi, bx = _transpose(ap, ai, ax, n)
        o = O.NativeKlu(n)
        o.set_sparselu()
        assert o.factor(bp, bi, bx) == 0
        oa = O.NativeKlu(n)
        oa.set_sparselu()
        assert oa.factor(ap, ai, ax) == 0
        if np.array_equal(o.block(0)["col_perm"], oa.block(0)["col_perm"]):
            same_order += 1
            xo = rhs.copy()
            assert o.solve(xo) == 0
            assert same_bits(x, xo)
        else:
            assert dense  # the mesh's order is checked bit for bit
    assert same_order > 30
