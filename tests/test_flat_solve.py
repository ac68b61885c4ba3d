"""Flat pipelined solve (solve mode 3): its host model over the plan's level
records must reproduce NativeKluSolver's / SparseLuSolver's solve bit for bit
(native_klu.rs:1442-1488; sparse_lu.rs solve_internal), whatever order the
products and rows of a level run in. A record with a wrong slot range, a
wrong column or value offset, or a row placed in a level before one of its
dependencies changes the bits or is rejected as out of range.
"""
import numpy as np
import pytest

from oracle import oracle as O
from rustspice_amd import netlists as NL
from rustspice_amd import solver as S
from tests.helpers import random_csc, same_bits


def _oracle(ap, ai, ax, rhs, n, kind):
    o = O.NativeKlu(n)
    if kind == S.SPARSELU:
        o.set_sparselu()
    o.factor(ap, ai, ax)
    r = rhs.copy()
    o.solve(r)
    return r


CIRCUITS = {
    "mesh16": lambda: NL.resistor_mesh(16, 16),
    "mesh64": lambda: NL.resistor_mesh(64, 64),
    "ladder": lambda: NL.rc_ladder(30),
    "rcmesh8_taps": lambda: NL.rc_mesh_inverters(8, 8, 4, load=1e4),
}


@pytest.mark.parametrize("kind", [S.NATIVEKLU, S.SPARSELU])
@pytest.mark.parametrize("name", sorted(CIRCUITS))
def test_circuit_solve_bit_exact(name, kind):
    c = CIRCUITS[name]()
    ap, ai, ax, rhs, n = O.build_mna(c, x=np.zeros(c.node_count), mode=0)
    x, st = S.host_lp_selftest(n, ap, ai, ax, rhs, kind)
    assert same_bits(x, _oracle(ap, ai, ax, rhs, n, kind))
    # empty L levels are dropped, so never more levels than the row sweeps
    assert 1 <= st["levels"] <= st["sweep_levels"]
    assert st["max_level_entries"] <= st["entries"]


def test_mesh64_entries_are_the_strict_factor_entries():
    c = NL.resistor_mesh(64, 64)
    ap, ai, ax, rhs, n = O.build_mna(c, x=np.zeros(c.node_count), mode=0)
    _, st = S.host_lp_selftest(n, ap, ai, ax, rhs)
    ps = S.HostFactor(n, ap, ai, ax).plan_stats()
    assert ps["offdiag"] == 0
    # every strict L entry and every off-diagonal U entry is one slot (nnz_u
    # counts the diagonal); every U row and every non-empty L row one record
    assert st["entries"] == ps["nnz_l"] + ps["nnz_u"] - n
    assert n < st["rows"] < 2 * n
    assert st["max_level_entries"] == 3996  # the level of the densest L row


def test_random_matrices_with_pivoting():
    """Random patterns and values that make the first factor pivot. Plans
    with BTF off-diagonal entries have no level records and are refused."""
    rs = np.random.RandomState(23)
    checked = refused = 0
    for trial in range(80):
        n = int(rs.randint(2, 70))
        ap, ai = random_csc(rs, n, rs.uniform(0.05, 0.3))
        if len(ai) == 0:
            continue
        ax = rs.randn(len(ai)) * 10.0 ** rs.randint(-3, 3, len(ai))
        rhs = rs.randn(n)
        for kind in (S.NATIVEKLU, S.SPARSELU):
            want = _oracle(ap, ai, ax, rhs, n, kind)
            try:
                x, _ = S.host_lp_selftest(n, ap, ai, ax, rhs, kind)
            except S.SolverError:
                refused += 1
                continue
            assert same_bits(x, want), (trial, kind)
            checked += 1
    assert checked > 60


def test_transposed_records_solve_the_transposed_system():
    """The adjoint's transposed records (U^T on the refactor's column levels,
    then L^T) solve J^T y = e, SparseLU as the adjoint factors it (ordering 0
    and 4: unknowns without a diagonal entry last), against a dense solve.
    The matrices are diagonally dominant, so the factor is J to round-off (on
    a circuit matrix SparseLU clamps the ground pivot and F is not J)."""
    rs = np.random.RandomState(5)
    checked = 0
    for trial in range(40):
        n = int(rs.randint(5, 80))
        ap, ai = random_csc(rs, n, rs.uniform(0.05, 0.3))
        A = np.zeros((n, n))
        for j in range(n):
            for q in range(ap[j], ap[j + 1]):
                A[ai[q], j] += rs.randn()
        A += 10.0 * np.eye(n)
        cols = [np.nonzero(A[:, j])[0] for j in range(n)]
        ap2 = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int64)
        ai2 = np.concatenate(cols).astype(np.int64)
        ax2 = np.concatenate([A[c, j] for j, c in enumerate(cols)])
        e = rs.randn(n)
        want = np.linalg.solve(A.T, e)
        for ordering in (0, 4):
            y, st = S.host_lp_selftest(n, ap2, ai2, ax2, e, S.SPARSELU, ordering, transposed=True)
            assert np.abs(y - want).max() <= 1e-12 * np.abs(want).max(), (trial, ordering)
            checked += 1
    assert checked == 80
