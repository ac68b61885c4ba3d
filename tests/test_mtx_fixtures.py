"""The reference's real circuit matrices (tests/fixtures/matrices, used by
bbd_benchmark_tests.rs:315-360): the oracle's BbdSolver (4 partitions, as
create_solver(SolverType::Bbd) builds it, solver.rs:174) against its
SparseLU within the reference's own bands (1e-3, 1e-3, 0.1 relative, and a
residual no more than 100x SparseLU's), the oracle's factors against what
the mathematics fixes, and the loader against the reference's sizes."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.golden import mtx


def _solve(solver, n, ap, ai, ax, rhs):
    assert solver.analyze(ap, ai) == 0
    assert solver.factor(ap, ai, ax) == 0
    x = rhs.copy()
    assert solver.solve(x) == 0
    return x


@pytest.mark.parametrize("name,n,tol", mtx.FIXTURES)
def test_bbd_agrees_with_sparselu(name, n, tol):
    nn, ap, ai, ax = mtx.parse_matrix_market(name)
    assert nn == n
    rhs = mtx.generate_rhs(n, 42)
    s = O.NativeKlu(n)
    s.set_sparselu()
    s.prepare(n)
    x_ref = _solve(s, n, ap, ai, ax, rhs)
    b = O.Bbd(n, 4)
    b.prepare(n)
    x_bbd = _solve(b, n, ap, ai, ax, rhs)
    res_ref = mtx.relative_residual(n, ap, ai, ax, x_ref, rhs)
    res_bbd = mtx.relative_residual(n, ap, ai, ax, x_bbd, rhs)
    xn = np.sqrt((x_ref * x_ref).sum())
    rel = np.abs(x_bbd - x_ref).max() / (xn / np.sqrt(n)) if xn > 1e-30 else np.abs(x_bbd - x_ref).max()
    assert rel < tol, (name, rel)
    if res_ref > 0.0:
        assert res_bbd / res_ref < 100.0, (name, res_ref, res_bbd)
    info = b.info()
    assert info["blocks"] >= 1


def test_loader_sums_duplicates_and_fills_patterns():
    """rajat01 is a pattern matrix: 10 on the diagonal, -0.1 elsewhere (a
    duplicate position sums); the real matrices keep their values."""
    n, ap, ai, ax = mtx.parse_matrix_market("rajat01")
    assert len(ap) == n + 1 and ap[-1] == len(ai) == len(ax)
    for j in range(0, n, 97):
        rows = ai[ap[j]:ap[j + 1]]
        assert np.all(np.diff(rows) > 0)  # sorted, merged
        vals = ax[ap[j]:ap[j + 1]]
        assert np.all(vals[rows == j] % 10.0 == 0.0)
        off = vals[rows != j]
        assert np.allclose(off / -0.1, np.round(off / -0.1))
    r = mtx.generate_rhs(4, 42)
    assert np.all(np.abs(r) <= 1.0) and len(set(r)) == 4


@pytest.mark.parametrize("name", ["adder_dcop_01", "circuit_1"])
def test_sparselu_factor_is_an_lu_of_the_permuted_matrix(name):
    """SparseLU (no pivots): L U reproduces the symmetrically permuted matrix
    to rounding wherever no pivot was clamped (sparse_lu.rs:440-537)."""
    n, ap, ai, ax = mtx.parse_matrix_market(name)
    s = O.NativeKlu(n)
    s.set_sparselu()
    assert s.factor(ap, ai, ax) == 0
    blk = s.block(0)
    cp = blk["col_perm"]
    import scipy.sparse as sp
    A = sp.csc_matrix((ax, ai, ap), shape=(n, n))
    P = sp.csc_matrix((np.ones(n), (cp, np.arange(n))), shape=(n, n))
    B = (P @ A @ P.T).tocsc()
    L = sp.csc_matrix((blk["l_values"], blk["l_row_idx"], blk["l_col_ptr"]), shape=(n, n)) + sp.identity(n)
    U = sp.csc_matrix((blk["u_values"], blk["u_row_idx"], blk["u_col_ptr"]), shape=(n, n))
    clamped = np.nonzero(np.abs(blk["u_diag"]) == 1e-14)[0]
    E = (L @ U - B).tocoo()
    bound = (abs(L) @ abs(U)).tocsr()  # rounding is relative to |L| |U| entrywise
    err = np.abs(E.data)
    lim = 1e-13 * np.asarray(bound[E.row, E.col]).ravel() + 1e-300
    ok = (err <= lim) | ((E.row == E.col) & np.isin(E.row, clamped))  # a replaced pivot
    assert ok.all(), (name, int((~ok).sum()), float((err / lim)[~ok].max()))
    # and a dropped term shows: the same check fails once one L value is zeroed
    lv = blk["l_values"].copy()
    lv[np.argmax(np.abs(lv))] = 0.0
    L2 = sp.csc_matrix((lv, blk["l_row_idx"], blk["l_col_ptr"]), shape=(n, n)) + sp.identity(n)
    E2 = (L2 @ U - B).tocoo()
    lim2 = 1e-13 * np.asarray(bound[E2.row, E2.col]).ravel() + 1e-300
    assert (np.abs(E2.data) > lim2).any()
