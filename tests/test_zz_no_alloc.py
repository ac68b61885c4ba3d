"""rs_solve_device_noalloc (no reference counterpart: the device solve for
callers that cannot allocate) against rs_solve_device on the same handle,
bit for bit, twice in a row with the second call on a smaller matrix; on the
supernodal path, on the level-set path with its solve vector in global
memory (the scratch array) and with it in LDS."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rustspice_amd import _lib
from rustspice_amd import netlists as NL
from rustspice_amd import solver as S
from tests.helpers import same_bits

pytestmark = pytest.mark.gpu


def _grid_tran(nx, seed):
    c = NL.power_grid(nx, nx, pitch=max(4, nx // 3))
    rs = np.random.RandomState(seed)
    cv, ci = rs.rand(c.ndev), rs.randn(c.ndev) * 1e-4
    hv, hi = np.ones(c.ndev, np.uint8), np.ones(c.ndev, np.uint8)
    return O.build_mna(c, x=rs.rand(c.node_count), mode=1, t=3e-12, dt=1e-12,
                       tstate=(cv, ci, hv, hi, 1))


def _mesh_dc(nx):
    return O.build_mna(NL.resistor_mesh(nx, nx))


def _both(g, rhs):
    """x from rs_solve_device and from rs_solve_device_noalloc."""
    d_b = torch.tensor(rhs, dtype=torch.float64, device="cuda")
    x1 = torch.full_like(d_b, float("nan"))
    x2 = torch.full_like(d_b, float("nan"))
    g.solve_device(d_b, x1)
    g.solve_device_noalloc(d_b, x2)
    torch.cuda.synchronize()
    return x1.cpu().numpy(), x2.cpu().numpy()


def _solver(mats, kind, setup):
    ap, ai, ax, rhs, n = mats
    g = S.GpuSolver(n, kind)
    if kind == S.SPARSELU:
        g.set_supernodal(2)
    g.factor(ap, ai, ax)
    if setup is not None:
        g.set_launch_shape(256, 0, 256, setup)
    g.factor(ap, ai, ax)  # a device refactor: the factor lives on the GPU
    return g, rhs, n


# (larger matrix, smaller matrix, solver kind, forced solve mode or None)
CASES = {
    "supernodal": (lambda: _grid_tran(40, 1), lambda: _grid_tran(20, 2), S.SPARSELU, None),
    "levels_global": (lambda: _mesh_dc(24), lambda: _mesh_dc(12), S.NATIVEKLU, 0),
    "levels_lds": (lambda: _mesh_dc(24), lambda: _mesh_dc(12), S.NATIVEKLU, 1),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_noalloc_solve_equals_solve_device(gpu, case):
    big, small, kind, mode = CASES[case]
    g1, rhs1, n1 = _solver(big(), kind, mode)
    g2, rhs2, n2 = _solver(small(), kind, mode)
    assert n2 < n1 <= _lib.lib().rs_noalloc_max_n()
    for g in (g1, g2):
        if mode is None:
            assert g.supernodal_info()["active"] == 1
        else:
            assert g.launch_shape()["sol_mode"] == mode
    for g, rhs in ((g1, rhs1), (g2, rhs2)):
        for b in (rhs, np.random.RandomState(len(rhs)).randn(len(rhs))):
            x1, x2 = _both(g, b)
            assert np.isfinite(x1).any()
            assert same_bits(x1, x2)
    # the first handle again after the smaller call: the shared array holds
    # the last call's contents, which the solve overwrites before reading
    x1, x2 = _both(g1, rhs1)
    assert same_bits(x1, x2)
    g1.close()
    g2.close()


def test_noalloc_rejects_bbd_handles(gpu):
    ap, ai, ax, rhs, n = _mesh_dc(8)
    g = S.GpuSolver(n, S.BBD)
    g.factor(ap, ai, ax)
    d_b = torch.tensor(rhs, dtype=torch.float64, device="cuda")
    x = torch.empty_like(d_b)
    with pytest.raises(S.SolverError):
        g.solve_device_noalloc(d_b, x)
