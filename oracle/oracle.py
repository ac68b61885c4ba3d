"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, never by the product package.  Build with
``make -C oracle`` (``__graft_entry__.build()`` does it).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class NewtonOpts(C.Structure):
    _fields_ = [
        ("max_iters", C.c_int32), ("gmin_steps", C.c_int32), ("source_steps", C.c_int32),
        ("pad", C.c_int32), ("abs_tol", C.c_double), ("rel_tol", C.c_double),
        ("gmin", C.c_double), ("damping", C.c_double), ("damping_min", C.c_double),
    ]


def default_opts(**kw) -> NewtonOpts:
    """engine.rs:987-999 with options.rs:79-145 defaults."""
    o = NewtonOpts(100, 0, 0, 0, 1e-6, 1e-3, 1e-12, 1.0, 0.1)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class NewtonResult(C.Structure):
    _fields_ = [("converged", C.c_int32), ("iterations", C.c_int32), ("reason", C.c_int32),
                ("final_norm", C.c_double)]


class TranResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("accepted", C.c_int32), ("rejected", C.c_int32),
                ("newton_iters_total", C.c_int32), ("factor_calls", C.c_int32),
                ("dc_iters", C.c_int32), ("npoints", C.c_int64)]


class SweepResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("npoints", C.c_int64),
                ("nsolved", C.c_int64)]


class Mna(C.Structure):
    _fields_ = [("n", C.c_int64), ("nnz", C.c_int64), ("ap", C.POINTER(C.c_int64)),
                ("ai", C.POINTER(C.c_int64)), ("ax", C.POINTER(C.c_double)),
                ("rhs", C.POINTER(C.c_double))]


class TState(C.Structure):
    _fields_ = [("cap_v", C.POINTER(C.c_double)), ("cap_i", C.POINTER(C.c_double)),
                ("has_v", C.POINTER(C.c_uint8)), ("has_i", C.POINTER(C.c_uint8)),
                ("method", C.c_int)]


class BbdDecomp(C.Structure):
    _fields_ = [("n", C.c_int64), ("num_blocks", C.c_int64), ("border_start", C.c_int64),
                ("border_size", C.c_int64), ("perm", C.POINTER(C.c_int64)),
                ("inv_perm", C.POINTER(C.c_int64)), ("block_ptr", C.POINTER(C.c_int64))]


class Btf(C.Structure):
    _fields_ = [("n", C.c_int64), ("row_perm", C.POINTER(C.c_int64)),
                ("col_perm", C.POINTER(C.c_int64)), ("row_perm_inv", C.POINTER(C.c_int64)),
                ("col_perm_inv", C.POINTER(C.c_int64)), ("block_ptr", C.POINTER(C.c_int64)),
                ("num_blocks", C.c_int64), ("structural_rank", C.c_int64),
                ("num_singletons", C.c_int64), ("max_block_size", C.c_int64)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.orc_amd_order.argtypes = [C.c_int64, i64p, i64p, i64p, i64p]
        L.orc_btf_decompose.argtypes = [C.c_int64, i64p, i64p, C.POINTER(Btf)]
        L.orc_btf_free.argtypes = [C.POINTER(Btf)]
        L.orc_klu_new.restype = vp
        L.orc_klu_new.argtypes = [C.c_int64]
        L.orc_klu_free.argtypes = [vp]
        L.orc_klu_prepare.argtypes = [vp, C.c_int64]
        L.orc_klu_analyze.argtypes = [vp, C.c_int64, i64p, C.c_int64, i64p]
        L.orc_klu_factor.argtypes = [vp, C.c_int64, i64p, C.c_int64, i64p, f64p]
        L.orc_klu_solve.argtypes = [vp, C.c_int64, f64p]
        L.orc_klu_reset_pattern.argtypes = [vp]
        L.orc_klu_set_config.argtypes = [vp, C.c_double, C.c_int, C.c_int]
        L.orc_klu_set_sparselu.argtypes = [vp]
        L.orc_klu_stats.argtypes = [vp, i64p, C.POINTER(C.c_double)]
        L.orc_klu_block_n.restype = C.c_int64
        L.orc_klu_block_n.argtypes = [vp, C.c_int64]
        L.orc_klu_block_sizes.argtypes = [vp, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_klu_block_get.argtypes = [vp, C.c_int64] + [vp] * 9
        L.orc_klu_btf_get.argtypes = [vp, i64p, i64p, i64p]
        L.orc_build_mna.argtypes = [vp, C.c_int64, C.c_int64, vp, C.c_int64, C.c_int,
                                    C.c_double, C.c_double, vp, C.c_double, C.c_double,
                                    C.POINTER(Mna)]
        L.orc_mna_free.argtypes = [C.POINTER(Mna)]
        L.orc_level1.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, f64p]
        L.orc_pulse_eval.restype = C.c_double
        L.orc_pulse_eval.argtypes = [f64p, C.c_double]
        L.orc_dc_op.argtypes = [vp, C.c_int64, C.c_int64, C.POINTER(NewtonOpts), vp, f64p,
                                C.c_int64, C.POINTER(C.c_int64), C.POINTER(NewtonResult)]
        L.orc_set_initial_guess.argtypes = [vp, C.c_int64]
        L.orc_ac_frequencies.restype = C.c_int64
        L.orc_ac_frequencies.argtypes = [C.c_int, C.c_int64, C.c_double, C.c_double, vp, C.c_int64]
        L.orc_ac.restype = C.c_int64
        L.orc_ac.argtypes = [vp, C.c_int64, C.c_int64, vp, vp, vp, C.c_int64, vp, C.c_int64, vp, vp,
                             C.c_int64, C.POINTER(C.c_int64)]
        L.orc_set_initial_guess.restype = None
        L.orc_tran.argtypes = [vp, C.c_int64, C.c_int64, C.POINTER(NewtonOpts), C.c_double,
                               C.c_double, C.c_double, C.c_double, vp, C.c_int64, vp, vp,
                               C.c_int64, C.POINTER(TranResult)]
        L.orc_tran_fixed.argtypes = [vp, C.c_int64, C.c_int64, C.POINTER(NewtonOpts), C.c_double,
                                     C.c_int64, vp, f64p, C.c_int64, C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int32)]
        L.orc_lte_milne.argtypes = [f64p, f64p, C.c_int64, C.c_double, C.c_double, f64p]
        L.orc_set_bsim_models.argtypes = [vp, C.c_int64]
        L.orc_set_mutation.argtypes = [C.c_int]
        L.orc_bsim_eval.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_double, f64p]
        L.orc_dc_sweep_points.restype = C.c_int64
        L.orc_dc_sweep_points.argtypes = [C.c_double, C.c_double, C.c_double, vp, C.c_int64]
        L.orc_dc_sweep.argtypes = [vp, C.c_int64, C.c_int64, C.POINTER(NewtonOpts), C.c_int64,
                                   C.c_double, C.c_double, C.c_double, vp, C.c_int64, vp,
                                   C.c_int64, C.POINTER(SweepResult)]
        L.orc_bbd_partition.restype = C.c_int64
        L.orc_bbd_partition.argtypes = [C.c_int64, i64p, i64p, C.c_int64, i64p]
        L.orc_bbd_decompose.argtypes = [C.c_int64, i64p, i64p, C.c_int64, C.POINTER(BbdDecomp)]
        L.orc_bbd_decomp_free.argtypes = [C.POINTER(BbdDecomp)]
        L.orc_bbd_new.restype = vp
        L.orc_bbd_new.argtypes = [C.c_int64, C.c_int64]
        L.orc_bbd_free.argtypes = [vp]
        L.orc_bbd_prepare.argtypes = [vp, C.c_int64]
        L.orc_bbd_reset_pattern.argtypes = [vp]
        L.orc_bbd_analyze.argtypes = [vp, C.c_int64, i64p, C.c_int64, i64p]
        L.orc_bbd_factor.argtypes = [vp, C.c_int64, i64p, C.c_int64, i64p, f64p]
        L.orc_bbd_solve.argtypes = [vp, C.c_int64, f64p]
        L.orc_bbd_info.argtypes = [vp, i64p]
        L.orc_dense_build.argtypes = [C.c_int64, C.c_int64, i64p, i64p, f64p, f64p]
        L.orc_dense_factor.argtypes = [C.c_int64, f64p, i64p, C.c_int]
        L.orc_dense_solve.argtypes = [C.c_int64, f64p, i64p, f64p, C.c_int]
        L.orc_colpar_factor.restype = C.c_int
        L.orc_colpar_factor.argtypes = [C.c_int64, i64p, i64p, vp, i64p, i64p, i64p, i64p, i64p, i64p,
                                        f64p, f64p, f64p, f64p, C.c_int]
        L.orc_colpar_columns.restype = C.c_double
        L.orc_colpar_columns.argtypes = [C.c_int64, i64p, i64p, vp, i64p, i64p, i64p, i64p, i64p, i64p,
                                         f64p, f64p, f64p, f64p, i64p, C.c_int64]
        L.orc_lu_solve.argtypes = [C.c_int64, i64p, vp, i64p, i64p, f64p, i64p, i64p, f64p, f64p]
        _LIB = L
    return _LIB


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def amd_order(n, ap, ai):
    perm = np.zeros(max(n, 1), np.int64)
    inv = np.zeros(max(n, 1), np.int64)
    lib().orc_amd_order(n, _i64(ap), _i64(ai), perm, inv)
    return perm[:n], inv[:n]


def btf_decompose(n, ap, ai):
    b = Btf()
    lib().orc_btf_decompose(n, _i64(ap), _i64(ai), C.byref(b))
    nb = b.num_blocks
    arr = lambda p, k: np.ctypeslib.as_array(p, shape=(k,)).copy() if k else np.zeros(0, np.int64)  # noqa: E731
    out = dict(row_perm=arr(b.row_perm, n), col_perm=arr(b.col_perm, n),
               row_perm_inv=arr(b.row_perm_inv, n), col_perm_inv=arr(b.col_perm_inv, n),
               block_ptr=arr(b.block_ptr, nb + 1), num_blocks=nb,
               structural_rank=b.structural_rank, num_singletons=b.num_singletons,
               max_block_size=b.max_block_size)
    lib().orc_btf_free(C.byref(b))
    return out


class NativeKlu:
    """Oracle NativeKluSolver (native_klu.rs), LinearSolver-shaped."""

    def __init__(self, n: int):
        self.h = lib().orc_klu_new(n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_klu_free(self.h)
            self.h = None

    def set_config(self, pivot_tol=0.001, ordering=0, use_btf=True):
        lib().orc_klu_set_config(self.h, pivot_tol, ordering, int(use_btf))

    def set_sparselu(self):
        """SparseLuSolver semantics (sparse_lu.rs): no pivoting, no BTF."""
        lib().orc_klu_set_sparselu(self.h)

    def prepare(self, n):
        lib().orc_klu_prepare(self.h, n)

    def analyze(self, ap, ai):
        ap, ai = _i64(ap), _i64(ai)
        return lib().orc_klu_analyze(self.h, len(ap), ap, len(ai), ai)

    def factor(self, ap, ai, ax):
        ap, ai = _i64(ap), _i64(ai)
        return lib().orc_klu_factor(self.h, len(ap), ap, len(ai), ai, _f64(ax))

    def solve(self, rhs):
        """Solves in place on a float64 array; returns the status code."""
        assert rhs.dtype == np.float64 and rhs.flags.c_contiguous
        return lib().orc_klu_solve(self.h, len(rhs), rhs)

    def reset_pattern(self):
        lib().orc_klu_reset_pattern(self.h)

    def stats(self):
        out = np.zeros(6, np.int64)
        rc = C.c_double()
        lib().orc_klu_stats(self.h, out, C.byref(rc))
        keys = ["factor_count", "refactor_count", "nnz_l", "nnz_u", "nblocks", "noffdiag"]
        d = {k: int(v) for k, v in zip(keys, out)}
        d["rcond"] = rc.value
        return d

    def block(self, b=0):
        L = lib()
        n = L.orc_klu_block_n(self.h, b)
        nl, nu = C.c_int64(), C.c_int64()
        L.orc_klu_block_sizes(self.h, b, C.byref(nl), C.byref(nu))
        nl, nu = nl.value, nu.value
        d = dict(col_perm=np.zeros(n, np.int64), row_perm=np.zeros(n, np.int64),
                 l_col_ptr=np.zeros(n + 1, np.int64), l_row_idx=np.zeros(max(nl, 1), np.int64),
                 u_col_ptr=np.zeros(n + 1, np.int64), u_row_idx=np.zeros(max(nu, 1), np.int64),
                 l_values=np.zeros(max(nl, 1)), u_values=np.zeros(max(nu, 1)), u_diag=np.zeros(n))
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        L.orc_klu_block_get(self.h, b, *[ptr(d[k]) for k in
                            ["col_perm", "row_perm", "l_col_ptr", "l_row_idx", "u_col_ptr",
                             "u_row_idx", "l_values", "u_values", "u_diag"]])
        d["l_row_idx"] = d["l_row_idx"][:nl]
        d["l_values"] = d["l_values"][:nl]
        d["u_row_idx"] = d["u_row_idx"][:nu]
        d["u_values"] = d["u_values"][:nu]
        d["n"] = n
        return d

    def btf(self, n):
        rp, cp, bp = np.zeros(n, np.int64), np.zeros(n, np.int64), np.zeros(n + 1, np.int64)
        nb = lib().orc_klu_btf_get(self.h, rp, cp, bp)
        if nb == 0:
            return None
        return dict(row_perm=rp, col_perm=cp, block_ptr=bp[: nb + 1], num_blocks=nb)


def bbd_partition(n, ap, ai, num_blocks):
    """GreedyBisectionPartitioner::partition (bbd.rs:103-171): (assignment
    with -1 for border nodes, block count)."""
    a = np.zeros(max(n, 1), np.int64)
    nb = lib().orc_bbd_partition(n, _i64(ap), _i64(ai), num_blocks, a)
    return a[:n], int(nb)


def bbd_decompose(n, ap, ai, num_blocks):
    """bbd_decompose (bbd.rs:355-485) as a dict."""
    d = BbdDecomp()
    lib().orc_bbd_decompose(n, _i64(ap), _i64(ai), num_blocks, C.byref(d))
    nb = d.num_blocks
    out = dict(num_blocks=int(nb), border_start=int(d.border_start), border_size=int(d.border_size),
               perm=np.array(d.perm[:n], np.int64), inv_perm=np.array(d.inv_perm[:n], np.int64),
               block_ptr=np.array(d.block_ptr[: nb + 1], np.int64))
    lib().orc_bbd_decomp_free(C.byref(d))
    return out


class Bbd:
    """Oracle BbdSolver (bbd_solver.rs) with GreedyBisectionPartitioner."""

    def __init__(self, n, num_blocks=2):
        self.h = lib().orc_bbd_new(n, num_blocks)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_bbd_free(self.h)
            self.h = None

    def prepare(self, n):
        lib().orc_bbd_prepare(self.h, n)

    def analyze(self, ap, ai):
        ap, ai = _i64(ap), _i64(ai)
        return lib().orc_bbd_analyze(self.h, len(ap), ap, len(ai), ai)

    def factor(self, ap, ai, ax):
        ap, ai = _i64(ap), _i64(ai)
        return lib().orc_bbd_factor(self.h, len(ap), ap, len(ai), ai, _f64(ax))

    def solve(self, rhs):
        assert rhs.dtype == np.float64 and rhs.flags.c_contiguous
        return lib().orc_bbd_solve(self.h, len(rhs), rhs)

    def reset_pattern(self):
        lib().orc_bbd_reset_pattern(self.h)

    def info(self):
        out = np.zeros(4, np.int64)
        lib().orc_bbd_info(self.h, out)
        return dict(blocks=int(out[0]), border=int(out[1]), schur=int(out[2]),
                    factor_count=int(out[3]))


def dense_factor_solve(n, ap, ai, ax, rhs, schur=False):
    """DenseSolver (solver.rs:595-724) factor + solve, or DenseSchurSolver's
    checks with schur=True. Returns (status, x)."""
    lu = np.zeros(max(n * n, 1))
    piv = np.zeros(max(n, 1), np.int64)
    L = lib()
    ap = _i64(ap)
    rc = L.orc_dense_build(n, len(ap), ap, _i64(ai), _f64(ax), lu)
    if rc == 0:
        rc = L.orc_dense_factor(n, lu, piv, int(schur))
    x = np.array(rhs, np.float64)
    if rc == 0:
        rc = L.orc_dense_solve(n, lu, piv, x, int(schur))
    return rc, x


_MODELS_REF = None


def set_bsim_models(models):
    """The BSIM model table orc_build_mna uses (kept alive here)."""
    global _MODELS_REF
    if models is None or len(models) == 0:
        _MODELS_REF = None
        lib().orc_set_bsim_models(None, 0)
        return
    _MODELS_REF = np.ascontiguousarray(models)
    lib().orc_set_bsim_models(_MODELS_REF.ctypes.data_as(C.c_void_p), len(_MODELS_REF))


def _devptr(circ):
    set_bsim_models(getattr(circ, "models", None))
    return circ.devices.ctypes.data_as(C.c_void_p)


def bsim_eval(model, level, w, l, vd, vg, vs, vb, temp=300.15):
    """evaluate_bsim_dc / evaluate_bsim4_dc of one bias point: dict."""
    m = np.ascontiguousarray(np.array(model))
    out = np.zeros(13)
    lib().orc_bsim_eval(m.ctypes.data_as(C.c_void_p), int(level), w, l, vd, vg, vs, vb, temp, out)
    keys = ["ids", "gm", "gds", "gmbs", "ieq", "region", "vth", "isub", "gsub", "igs", "igd",
            "gigs", "gigd"]
    r = dict(zip(keys, map(float, out)))
    r["region"] = int(r["region"])
    return r


def build_mna(circ, x=None, mode=0, t=0.0, dt=0.0, tstate=None, gmin=1e-12,
              source_scale=1.0):
    """MNA of one Newton iteration: returns (ap, ai, ax, rhs, n) like the engine
    closure (engine.rs:262-283). tstate = (cap_v, cap_i, has_v, has_i, method)."""
    L = lib()
    x = _f64(np.zeros(circ.node_count) if x is None else x)
    m = Mna()
    ts = None
    keep = None
    if tstate is not None:
        cv, ci, hv, hi, meth = tstate
        keep = [_f64(cv), _f64(ci), np.ascontiguousarray(hv, np.uint8),
                np.ascontiguousarray(hi, np.uint8)]
        ts = TState(keep[0].ctypes.data_as(C.POINTER(C.c_double)),
                    keep[1].ctypes.data_as(C.POINTER(C.c_double)),
                    keep[2].ctypes.data_as(C.POINTER(C.c_uint8)),
                    keep[3].ctypes.data_as(C.POINTER(C.c_uint8)), meth)
    L.orc_build_mna(_devptr(circ), circ.ndev, circ.node_count, x.ctypes.data_as(C.c_void_p),
                    len(x), mode, t, dt, C.cast(C.byref(ts), C.c_void_p) if ts else None,
                    gmin, source_scale, C.byref(m))
    n, nnz = m.n, m.nnz
    ap = np.ctypeslib.as_array(m.ap, shape=(n + 1,)).copy()
    ai = np.ctypeslib.as_array(m.ai, shape=(max(nnz, 1),))[:nnz].copy()
    ax = np.ctypeslib.as_array(m.ax, shape=(max(nnz, 1),))[:nnz].copy()
    rhs = np.ctypeslib.as_array(m.rhs, shape=(max(n, 1),))[:n].copy()
    L.orc_mna_free(C.byref(m))
    return ap, ai, ax, rhs, n


def level1(dev, vd, vg, vs, vb=0.0):
    out = np.zeros(5)
    rec = np.ascontiguousarray(dev)
    lib().orc_level1(rec.ctypes.data_as(C.c_void_p), vd, vg, vs, vb, out)
    return dict(ids=out[0], gm=out[1], gds=out[2], gmbs=out[3], ieq=out[4])


def set_mutation(m: int):
    """Plants a mistake in the oracle (0 restores the reference behaviour;
    1 flips the PMOS current sign of the level-1 model). Tests use it to show
    that a parity check fails on a plausible bug."""
    lib().orc_set_mutation(int(m))


def pulse_eval(p7, t):
    return lib().orc_pulse_eval(_f64(p7), t)


def dc_op(circ, opts=None, solver=None, last=False):
    """Returns (x or None, NewtonResult). With last=True, x is returned
    whether or not Newton converged: the solution, or the x the loop ended
    with (node_count entries, as newton.rs:93-107 leaves it)."""
    opts = opts or default_opts()
    own = solver is None
    s = solver or NativeKlu(circ.node_count)
    cap = circ.node_count + circ.ndev + 8
    x = np.zeros(cap)
    xl = C.c_int64()
    res = NewtonResult()
    lib().orc_dc_op(_devptr(circ), circ.ndev, circ.node_count, C.byref(opts), s.h, x, cap,
                    C.byref(xl), C.byref(res))
    if own:
        del s
    return (x[: xl.value].copy() if (res.converged or last) else None), res


def tran(circ, tstep, tstop, tstart=0.0, tmax=None, opts=None, solver=None, max_points=100000,
         store=True, x0=None):
    opts = opts or default_opts()
    ic = np.ascontiguousarray(np.zeros(0) if x0 is None else x0, np.float64)
    lib().orc_set_initial_guess(ic.ctypes.data_as(C.c_void_p), len(ic))
    s = solver or NativeKlu(circ.node_count)
    tmax = tstop if tmax is None else tmax
    stride = circ.node_count + circ.ndev + 8
    times = np.zeros(max_points) if store else None
    sols = np.zeros(max_points * stride) if store else None
    res = TranResult()
    lib().orc_tran(_devptr(circ), circ.ndev, circ.node_count, C.byref(opts), tstep, tstop, tstart,
                   tmax, s.h, max_points,
                   times.ctypes.data_as(C.c_void_p) if store else None,
                   sols.ctypes.data_as(C.c_void_p) if store else None, stride, C.byref(res))
    out = dict(status=res.status, accepted=res.accepted, rejected=res.rejected,
               newton_iters_total=res.newton_iters_total, factor_calls=res.factor_calls,
               dc_iters=res.dc_iters, npoints=int(res.npoints))
    lib().orc_set_initial_guess(None, 0)
    if store:
        k = min(int(res.npoints), max_points)
        out["times"] = times[:k].copy()
        out["sols"] = sols[: k * stride].reshape(k, stride).copy()
    return out


def tran_fixed(circ, dt, nsteps, opts=None, solver=None):
    opts = opts or default_opts()
    s = solver or NativeKlu(circ.node_count)
    cap = circ.node_count + circ.ndev + 8
    x = np.zeros(cap)
    xl = C.c_int64()
    it = C.c_int32()
    rc = lib().orc_tran_fixed(_devptr(circ), circ.ndev, circ.node_count, C.byref(opts), dt, nsteps,
                              s.h, x, cap, C.byref(xl), C.byref(it))
    return rc, x[: xl.value].copy(), it.value


def dc_sweep_points(start, stop, step):
    n = lib().orc_dc_sweep_points(start, stop, step, None, 0)
    v = np.zeros(max(n, 1))
    lib().orc_dc_sweep_points(start, stop, step, v.ctypes.data_as(C.c_void_p), n)
    return v[:n].copy()


def dc_sweep(circ, source, start, stop, step, opts=None, solver=None):
    """engine.rs:684-820 on device index `source`: dict(status, iterations,
    values, sols[nsolved, stride])."""
    opts = opts or default_opts()
    s = solver or NativeKlu(circ.node_count)
    values = dc_sweep_points(start, stop, step)
    stride = circ.node_count + circ.ndev + 8
    sols = np.zeros((max(len(values), 1), stride))
    res = SweepResult()
    lib().orc_dc_sweep(_devptr(circ), circ.ndev, circ.node_count, C.byref(opts), source, start,
                       stop, step, s.h, len(values), sols.ctypes.data_as(C.c_void_p), stride,
                       C.byref(res))
    return dict(status=res.status, iterations=res.iterations, values=values,
                nsolved=int(res.nsolved), sols=sols[: res.nsolved].copy())


class ColPar:
    """The reference's column-parallel factor (native_klu.rs:402-576) and the
    matching triangular solve on a given one-block pattern: col_perm[original]
    = position, L/U column pointers and row indices in position space (U
    diagonal last). row_perm_inv None: identity pivots (SparseLU's factor)."""

    def __init__(self, pat, row_perm_inv=None):
        self.n = int(pat["n"])
        self.cp = _i64(pat["col_perm"])
        self.cpi = np.empty(self.n, np.int64)
        self.cpi[self.cp] = np.arange(self.n, dtype=np.int64)
        self.lcp, self.lri = _i64(pat["l_col_ptr"]), _i64(pat["l_row_idx"])
        self.ucp, self.uri = _i64(pat["u_col_ptr"]), _i64(pat["u_row_idx"])
        self.rpi = None if row_perm_inv is None else _i64(row_perm_inv)
        self.lx = np.zeros(max(len(self.lri), 1))
        self.ux = np.zeros(max(len(self.uri), 1))
        self.ud = np.zeros(max(self.n, 1))

    def _rpi(self):
        return None if self.rpi is None else self.rpi.ctypes.data_as(C.c_void_p)

    def factor(self, ap, ai, ax, nthreads=1):
        lib().orc_colpar_factor(self.n, self.cp, self.cpi, self._rpi(), self.lcp, self.lri, self.ucp,
                                self.uri, _i64(ap), _i64(ai), _f64(ax), self.lx, self.ux, self.ud,
                                int(nthreads))

    def columns(self, ap, ai, ax, cols):
        """Recompute the listed columns on one thread from the L values
        already held (a full factor first); returns the seconds taken."""
        return lib().orc_colpar_columns(self.n, self.cp, self.cpi, self._rpi(), self.lcp, self.lri,
                                        self.ucp, self.uri, _i64(ap), _i64(ai), _f64(ax), self.lx,
                                        self.ux, self.ud, _i64(cols), len(cols))

    def solve(self, rhs):
        x = _f64(rhs).copy()
        lib().orc_lu_solve(self.n, self.cp, self._rpi(), self.lcp, self.lri, self.lx, self.ucp,
                           self.uri, self.ux, x)
        return x

    def values(self):
        return self.lx[: len(self.lri)], self.ux[: len(self.uri)], self.ud[: self.n]


def lte_milne(xb, xt, abs_tol=1e-6, rel_tol=1e-3):
    out = np.zeros(5)
    xb, xt = _f64(xb), _f64(xt)
    lib().orc_lte_milne(xb, xt, min(len(xb), len(xt)), abs_tol, rel_tol, out)
    return dict(error_norm=out[0], max_node=int(out[1]), accept=bool(out[2]),
                suggested_factor=out[3], raw_lte=out[4])


AC_SWEEPS = {"dec": 0, "oct": 1, "lin": 2}


def ac_frequencies(sweep, points, fstart, fstop):
    """generate_frequency_sweep (engine.rs:1002-1044)."""
    k = AC_SWEEPS[sweep]
    n = lib().orc_ac_frequencies(k, points, fstart, fstop, None, 0)
    out = np.zeros(max(n, 1))
    lib().orc_ac_frequencies(k, points, fstart, fstop, out.ctypes.data_as(C.c_void_p), n)
    return out[:n].copy()


def ac(circ, ac_mag, ac_phase, x_dc, freqs):
    """The AC sweep of engine.rs:864-949 at the DC solution x_dc. Returns
    (x[nsolved, n] complex, magph[nsolved, node_count, 2] (dB, degrees), n)."""
    if circ.models is not None and len(circ.models):
        set_bsim_models(circ.models)
    freqs = np.ascontiguousarray(freqs, np.float64)
    ndev = circ.ndev
    nv = int(np.count_nonzero(circ.devices["kind"] == 3)) if ndev else 0
    cap_n = circ.node_count + nv
    nf = len(freqs)
    xo = np.zeros(max(nf * cap_n * 2, 1))
    mp = np.zeros(max(nf * circ.node_count * 2, 1))
    nout = C.c_int64()
    am = np.ascontiguousarray(ac_mag, np.float64)
    ph = np.ascontiguousarray(ac_phase, np.float64)
    xd = np.ascontiguousarray(x_dc, np.float64)
    k = lib().orc_ac(_devptr(circ), ndev, circ.node_count, am.ctypes.data_as(C.c_void_p),
                     ph.ctypes.data_as(C.c_void_p), xd.ctypes.data_as(C.c_void_p), len(xd),
                     freqs.ctypes.data_as(C.c_void_p), nf, xo.ctypes.data_as(C.c_void_p),
                     mp.ctypes.data_as(C.c_void_p), cap_n, C.byref(nout))
    x = xo[: nf * cap_n * 2].reshape(nf, cap_n, 2)[:k]
    return x[..., 0] + 1j * x[..., 1], mp[: nf * circ.node_count * 2].reshape(nf, circ.node_count, 2)[:k], nout.value
