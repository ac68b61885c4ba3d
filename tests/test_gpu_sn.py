"""GPU parity of the supernodal factor and solve (SparseLU semantics,
sparse_lu.rs; BbdSolver's SparseLU fallback, bbd_solver.rs:414-504) on the RC
power grid: factor values, pivots, solutions and whole fixed-step transients
bit for bit against the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from rustspice_amd import _lib
from rustspice_amd import engine as E
from rustspice_amd import netlists as NL
from rustspice_amd import solver as S
from tests.helpers import same_bits

pytestmark = pytest.mark.gpu


def _tran_mna(c, seed):
    rs = np.random.RandomState(seed)
    cv, ci = rs.rand(c.ndev), rs.randn(c.ndev) * 1e-4
    hv, hi = np.ones(c.ndev, np.uint8), np.ones(c.ndev, np.uint8)
    return O.build_mna(c, x=rs.rand(c.node_count), mode=1, t=3e-12, dt=1e-12,
                       tstate=(cv, ci, hv, hi, 1))


def _oracle(ap, ai, ax, rhs, n):
    o = O.NativeKlu(n)
    o.set_sparselu()
    assert o.factor(ap, ai, ax) == 0
    x = rhs.copy()
    assert o.solve(x) == 0
    return o.block(0), x


@pytest.mark.parametrize("nx,pitch", [(24, 8), (48, 16), (96, 32)])
def test_factor_and_solve_bit_exact(gpu, nx, pitch):
    """First factor, refactor on new values and solves, every value equal to
    SparseLU's."""
    c = NL.power_grid(nx, nx, pitch=pitch)
    ap, ai, ax, rhs, n = _tran_mna(c, nx)
    ax2 = ax * (1.0 + 0.25 * np.random.RandomState(nx + 1).rand(len(ax)))
    g = S.GpuSolver(n, S.SPARSELU)
    g.set_supernodal(2)
    for k, vals in enumerate((ax, ax2)):
        want, xo = _oracle(ap, ai, vals, rhs, n)
        g.factor(ap, ai, vals)
        lx, ux, ud = g.factor_values()
        assert same_bits(lx, want["l_values"]) and same_bits(ux, want["u_values"])
        assert same_bits(ud, want["u_diag"])
        xg = rhs.copy()
        g.solve(xg)
        assert same_bits(xg, xo), (k, np.abs(xg - xo).max())
    info = g.supernodal_info()
    st = g.stats()
    assert info["active"] == 1 and info["levels"] < info["supernodes"]
    assert 0 < info["gather_updates"] < info["updates"]
    pat = g.factor_pattern()
    for k in ("col_perm", "l_col_ptr", "l_row_idx", "u_col_ptr", "u_row_idx"):
        assert np.array_equal(pat[k], want[k]), k
    assert (st["factor_count"], st["refactor_count"]) == (1, 1)
    assert st["rcond"] == O_rcond(want)


def O_rcond(blk):
    d = np.abs(blk["u_diag"])
    return d.min() / d.max()


def test_auto_mode_large_grid(gpu):
    """n >= 20,000 switches the SparseLU solver to the supernodal path by
    itself (160x160 grid, n = 25,626)."""
    c = NL.power_grid(160, 160, pitch=32)
    ap, ai, ax, rhs, n = _tran_mna(c, 5)
    g = S.GpuSolver(n, S.SPARSELU)
    g.factor(ap, ai, ax)
    assert g.supernodal_info()["active"] == 1
    _, xo = _oracle(ap, ai, ax, rhs, n)
    for b in (rhs, np.random.RandomState(9).randn(n)):
        xg = b.copy()
        g.solve(xg)
        xo = b.copy()
        o = O.NativeKlu(n)
        o.set_sparselu()
        o.factor(ap, ai, ax)
        o.solve(xo)
        assert same_bits(xg, xo)


def test_no_value_read_before_written(gpu):
    c = NL.power_grid(48, 48, pitch=16)
    ap, ai, ax, rhs, n = _tran_mna(c, 2)
    _lib.set_poison(True)
    try:
        g = S.GpuSolver(n, S.SPARSELU)
        g.set_supernodal(2)
        g.factor(ap, ai, ax)
        xg = rhs.copy()
        g.solve(xg)
    finally:
        _lib.set_poison(False)
    _, xo = _oracle(ap, ai, ax, rhs, n)
    assert same_bits(xg, xo)


@pytest.mark.parametrize("nx,pitch", [(32, 16), (64, 32)])
def test_tran_fixed_bbd_fallback_bit_exact(gpu, nx, pitch):
    """The power-grid transient through BbdSolver(64). The DC OP's pattern
    (capacitors open) splits into real blocks and a border; every TRAP step's
    pattern collapses (ground joins every node through its capacitor), so the
    whole matrix goes to SparseLU, here the supernodal path. DC OP and 20
    trapezoidal steps of 1 ps equal the oracle BbdSolver's run bit for bit,
    with the same Newton iterations."""
    c = NL.power_grid(nx, nx, pitch=pitch)
    o = O.Bbd(c.node_count, 64)
    rco, xo, ito = O.tran_fixed(c, 1e-12, 20, solver=o)
    g = S.GpuSolver(c.node_count, S.BBD)
    g.set_bbd_blocks(64)
    g.set_supernodal(2)
    st, xg, itg = E.GpuCircuit(c).tran_fixed(g, 1e-12, 20)
    assert rco == st == 0 and ito == itg
    assert same_bits(xg, xo)
    assert g.bbd_info()["blocks"] == 1 and g.supernodal_info()["active"] == 1
    # the fallback's factor through the BbdSolver handle: pattern and values
    pat = g.factor_pattern()
    assert pat["n"] == len(xg) and pat["u_col_ptr"][-1] == len(pat["u_row_idx"])
    lx, ux, ud = g.factor_values()
    assert len(lx) == len(pat["l_row_idx"]) and len(ux) == len(pat["u_row_idx"])
    assert same_bits(ud, ux[pat["u_col_ptr"][1:] - 1])


def test_large_grid_matches_host_schedule(gpu):
    """256x256 grid (n = 65,553): the oracle's literal AMD is quadratic here,
    so the GPU is held to the host execution of the same schedule, which the
    CPU tests pin to the oracle bit for bit on smaller grids. Wide supernodes
    (the blocked panel path) and large gather records (staged through LDS)
    both occur at this size."""
    c = NL.power_grid(256, 256)
    ap, ai, ax, rhs, n = _tran_mna(c, 4)
    xh, lxh, uxh, udh, st = S.host_sn_selftest(n, ap, ai, ax, rhs)
    assert st["max_w"] >= 256
    g = S.GpuSolver(n, S.SPARSELU)
    g.factor(ap, ai, ax)
    assert g.supernodal_info()["active"] == 1
    lx, ux, ud = g.factor_values()
    assert same_bits(ud, udh) and same_bits(lx, lxh) and same_bits(ux, uxh)
    xg = rhs.copy()
    g.solve(xg)
    assert same_bits(xg, xh)
