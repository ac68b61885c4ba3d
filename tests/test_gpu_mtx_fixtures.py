"""The reference's real circuit matrices (tests/fixtures/matrices) through
the GPU solvers: SparseLU (plain and with the supernodal path asked for),
NativeKLU (AMD, BTF, threshold pivots: the first factor on the host, the
refactor on the device) and BbdSolver(4), each against the oracle's same
solver bit for bit, on the first factor and on a refactor with new values."""
import numpy as np
import pytest

from oracle import oracle as O
from rustspice_amd import solver as S
from tests.golden import mtx
from tests.helpers import same_bits

pytestmark = pytest.mark.gpu

KINDS = ["sparselu", "sparselu_sn", "nativeklu", "bbd4"]


def _oracle(kind, n):
    if kind == "bbd4":
        return O.Bbd(n, 4)
    o = O.NativeKlu(n)
    if kind.startswith("sparselu"):
        o.set_sparselu()
    return o


def _gpu(kind, n):
    g = S.GpuSolver(n, {"sparselu": S.SPARSELU, "sparselu_sn": S.SPARSELU, "nativeklu": S.NATIVEKLU,
                        "bbd4": S.BBD}[kind])
    if kind == "sparselu_sn":
        g.set_supernodal(2)  # used where the pattern is structurally symmetric
    if kind == "bbd4":
        g.set_bbd_blocks(4)
    return g


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("name,n,tol", mtx.FIXTURES)
def test_fixture_matches_oracle(gpu, name, n, tol, kind):
    nn, ap, ai, ax = mtx.parse_matrix_market(name)
    rhs = mtx.generate_rhs(n, 42)
    ax2 = ax * (1.0 + 0.1 * np.random.RandomState(n).rand(len(ax)))
    o, g = _oracle(kind, n), _gpu(kind, n)
    for vals in (ax, ax2):
        assert o.factor(ap, ai, vals) == 0
        g.factor(ap, ai, vals)
        xo, xg = rhs.copy(), rhs.copy()
        assert o.solve(xo) == 0
        g.solve(xg)
        assert same_bits(xg, xo), (name, kind, np.nanmax(np.abs(xg - xo)))
