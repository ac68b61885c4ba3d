"""BbdSolver and DenseSolver back ends of librsgpu.so against the oracle.

CPU: the library's host decomposition (bbd.rs) equals the oracle's on the
reference's test matrices, random patterns, grids and MNA matrices.
GPU (marked): factor + solve through the C-ABI equals the oracle's BbdSolver
(bbd_solver.rs) / DenseSolver (solver.rs:595-724) bit for bit, on the dense
and the sparse Schur paths and on the ground-node fallback (Q7).
"""
import numpy as np
import pytest

from oracle import oracle as O
from rustspice_amd import netlists as NL
from rustspice_amd import solver as S
from tests.helpers import random_csc, same_bits
from tests.test_bbd_oracle import _grid

REF_MATRICES = [  # bbd_solver.rs / bbd.rs test matrices
    (3, [0, 1, 2, 3], [0, 1, 2], [2.0, 3.0, 4.0]),
    (2, [0, 2, 4], [0, 1, 0, 1], [3.0, 1.0, 1.0, 2.0]),
    (4, [0, 2, 4, 6, 8], [0, 1, 0, 1, 2, 3, 2, 3], [2.0, 1.0, 1.0, 3.0, 4.0, 1.0, 1.0, 2.0]),
    (3, [0, 2, 5, 7], [0, 1, 0, 1, 2, 1, 2], [4.0, 1.0, 1.0, 5.0, 2.0, 2.0, 6.0]),
    (5, [0, 2, 5, 7, 10, 13], [0, 1, 0, 1, 4, 2, 3, 2, 3, 4, 1, 3, 4],
     [4.0, 1.0, 1.0, 4.0, 1.0, 4.0, 1.0, 1.0, 4.0, 1.0, 1.0, 1.0, 3.0]),
    (8, [0, 2, 5, 8, 10, 13, 16, 19, 22],
     [0, 1, 0, 1, 2, 1, 2, 6, 3, 4, 3, 4, 5, 4, 5, 7, 2, 6, 7, 5, 6, 7],
     [2.0, -1.0, -1.0, 2.0, -1.0, -1.0, 2.0, 1.0, 2.0, -1.0, -1.0, 2.0, -1.0, -1.0, 2.0, 1.0,
      1.0, 3.0, 0.5, 1.0, 0.5, 3.0]),
]


def _cases():
    rs = np.random.RandomState(17)
    out = [(n, np.array(ap), np.array(ai), np.array(ax, float)) for n, ap, ai, ax in REF_MATRICES]
    for k in (6, 12, 40):
        out.append(_grid(k, rs))
    out.append(_grid(10, rs, ground=True))
    for _ in range(12):
        n = int(rs.randint(4, 80))
        ap, ai = random_csc(rs, n, rs.uniform(0.02, 0.15), sym=True, diag=True)
        out.append((n, ap, ai, rs.randn(len(ai)) + 5.0 * (ai == np.repeat(np.arange(n), np.diff(ap)))))
    return out


@pytest.mark.parametrize("nb", [2, 4, 8])
def test_host_decomposition_equals_oracle(nb):
    for n, ap, ai, _ in _cases():
        d = S.host_bbd_decompose(n, ap, ai, nb)
        o = O.bbd_decompose(n, ap, ai, nb)
        assert d["num_blocks"] == o["num_blocks"]
        assert np.array_equal(d["perm"], o["perm"]) and np.array_equal(d["block_ptr"], o["block_ptr"])
        assert d["border_start"] == o["border_start"]
    c = NL.resistor_mesh(16, 16)
    ap, ai, ax, rhs, n = O.build_mna(c)
    assert S.host_bbd_decompose(n, ap, ai, 64)["num_blocks"] == 1  # Q7 collapse


@pytest.mark.gpu
@pytest.mark.parametrize("nb", [2, 8])
def test_gpu_bbd_bit_exact(gpu, nb):
    rs = np.random.RandomState(23)
    seen = set()
    for n, ap, ai, ax in _cases():
        o = O.Bbd(n, nb)
        o.prepare(n)
        g = S.GpuSolver(n, S.BBD)
        g.set_bbd_blocks(nb)
        assert o.factor(ap, ai, ax) == 0
        g.factor(ap, ai, ax)
        info = g.bbd_info()
        oi = o.info()
        assert (info["blocks"], info["border"], info["schur"]) == (oi["blocks"], oi["border"], oi["schur"])
        seen.add(info["schur"])
        for _ in range(2):
            b = rs.randn(n)
            xo, xg = b.copy(), b.copy()
            assert o.solve(xo) == 0
            g.solve(xg)
            assert same_bits(xg, xo), (n, info)
        # a refactor with new values on the same pattern
        ax2 = ax * (1.0 + 0.01 * rs.randn(len(ax)))
        o.factor(ap, ai, ax2)
        g.factor(ap, ai, ax2)
        b = rs.randn(n)
        xo, xg = b.copy(), b.copy()
        o.solve(xo)
        g.solve(xg)
        assert same_bits(xg, xo)
    assert {-1, 1} <= seen and (nb != 2 or 0 in seen)  # fallback, dense and sparse Schur


@pytest.mark.gpu
def test_gpu_dense_solver_bit_exact(gpu):
    rs = np.random.RandomState(29)
    for n in (1, 3, 16, 50, 70):
        A = rs.randn(n, n) + 2.0 * np.eye(n)
        ap = np.arange(0, n * n + 1, n)
        ai = np.tile(np.arange(n), n)
        ax = A.T.ravel().copy()
        b = rs.randn(n)
        rc, xo = O.dense_factor_solve(n, ap, ai, ax, b)
        g = S.GpuSolver(n, S.DENSE)
        g.factor(ap, ai, ax)
        xg = b.copy()
        g.solve(xg)
        assert rc == 0 and same_bits(xg, xo), n
    g = S.GpuSolver(2, S.DENSE)  # a zero column fails like the reference (FactorFailed)
    with pytest.raises(S.SolverError):
        g.factor(np.array([0, 2, 2]), np.array([0, 1]), np.array([1.0, 2.0]))


# ---- dense Schur product on FP64 MFMA tiles ---------------------------------
TAP_BANKS = [(256, 3, 20), (64, 2, 8), (512, 2, 32)]  # border 40 (padded to 48), 16, 64


@pytest.mark.parametrize("shape", TAP_BANKS)
def test_tap_bank_border_is_the_rail_bus(shape):
    """The generator's claim the MFMA tests rest on: BbdSolver(2)'s bisection
    (the oracle's and the library's alike) cuts exactly the two rail rows."""
    W, D, w = shape
    n, ap, ai, ax = NL.tap_bank_matrix(W, D, w)
    d = S.host_bbd_decompose(n, ap, ai, 2)
    o = O.bbd_decompose(n, ap, ai, 2)
    assert np.array_equal(d["perm"], o["perm"]) and d["border_start"] == o["border_start"]
    assert d["num_blocks"] == 2 and n - d["border_start"] == 2 * w
    hub = set(range(2 + W * D, 2 + W * D + 2 * w))
    assert set(int(v) for v in d["perm"][d["border_start"]:]) == hub


def _rel_inf(x, ref):
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


MFMA_TOL = 1e-9  # relative (inf-norm) to the reference's term order


@pytest.mark.gpu
@pytest.mark.parametrize("shape", TAP_BANKS)
def test_gpu_bbd_mfma_schur_matches_oracle(gpu, shape):
    """Tiles on (mode 2, and mode 1's density rule, which takes them here):
    x within 1e-9 of the oracle's BbdSolver, through a refactor with new
    values; the default (mode 0) stays bit-exact on the same solver."""
    W, D, w = shape
    n, ap, ai, ax = NL.tap_bank_matrix(W, D, w)
    rs = np.random.RandomState(31)
    o = O.Bbd(n, 2)
    o.prepare(n)
    for mode in (2, 1, 0):
        g = S.GpuSolver(n, S.BBD)
        g.set_bbd_blocks(2)
        g.set_bbd_mfma(mode)
        for vals in (ax, ax * (1.0 + 0.01 * rs.randn(len(ax)))):
            assert o.factor(ap, ai, vals) == 0
            g.factor(ap, ai, vals)
            info = g.bbd_mfma_info()
            assert info["k"] == 2 * W and abs(info["c_density"] - 0.5) < 1e-3
            big = W >= 64  # mode 1 needs K >= 64
            assert info["used"] == (mode == 2 or (mode == 1 and big)), info
            b = rs.randn(n)
            xo, xg = b.copy(), b.copy()
            assert o.solve(xo) == 0
            g.solve(xg)
            if info["used"]:
                assert _rel_inf(xg, xo) <= MFMA_TOL, (mode, _rel_inf(xg, xo))
            else:
                assert same_bits(xg, xo)


@pytest.mark.gpu
def test_gpu_bbd_mfma_forced_on_every_dense_schur_case(gpu):
    """Mode 2 on every dense-Schur case of the bit-exact suite (tiny K, one
    padded tile, random patterns): within 1e-9 of the oracle; mode 1 takes
    the tiles only where its density rule says so, else stays bit-exact."""
    rs = np.random.RandomState(37)
    forced = 0
    for nb in (2, 8):
        for n, ap, ai, ax in _cases():
            o = O.Bbd(n, nb)
            o.prepare(n)
            assert o.factor(ap, ai, ax) == 0
            if o.info()["schur"] != 1:
                continue
            for mode in (2, 1):
                g = S.GpuSolver(n, S.BBD)
                g.set_bbd_blocks(nb)
                g.set_bbd_mfma(mode)
                g.factor(ap, ai, ax)
                info = g.bbd_mfma_info()
                b = rs.randn(n)
                xo, xg = b.copy(), b.copy()
                assert o.solve(xo) == 0
                g.solve(xg)
                if mode == 1 and not info["used"]:
                    assert same_bits(xg, xo), (n, nb, info)
                elif info["k"] > 0:
                    assert info["used"]
                    forced += 1
                    assert _rel_inf(xg, xo) <= MFMA_TOL, (n, nb, _rel_inf(xg, xo))
    assert forced >= 4
