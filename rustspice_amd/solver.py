"""Python mirror of RustSpice's ``LinearSolver`` trait over librsgpu.so.

``GpuSolver`` has the trait's methods (crates/sim-core/src/solver.rs:141-166)
with the same argument meaning: ``prepare(n)``, ``analyze(ap, ai)``,
``factor(ap, ai, ax)``, ``solve(rhs)`` (in place), ``reset_pattern()``,
``name()``, ``set_parallel_threads(t)``. Errors raise :class:`SolverError`
carrying the ``SolverError`` variant (solver.rs:76-92).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

NATIVEKLU, SPARSELU, BBD, DENSE = 0, 1, 2, 3

_ERR_NAMES = {
    1: "AnalyzeFailed", 2: "FactorFailed", 3: "SolveFailed", 4: "SingularMatrix",
    5: "IllConditioned", 6: "InvalidMatrix", 7: "DeviceError",
}


class SolverError(RuntimeError):
    """A SolverError variant (solver.rs:76-92) with its payload: `pivot` for
    SingularMatrix, `rcond` for IllConditioned."""

    def __init__(self, code: int, message: str, pivot=None, rcond=None):
        self.code = code
        self.variant = _ERR_NAMES.get(code, f"Unknown({code})")
        self.pivot, self.rcond = pivot, rcond
        super().__init__(f"{self.variant}: {message}")


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class GpuSolver:
    """``create_solver(SolverType::NativeKlu | SparseLu | Bbd | Dense, n)`` on
    the MI355X (solver.rs:169-200)."""

    def __init__(self, n: int, kind: int = NATIVEKLU):
        self._L = _lib.lib()
        self.h = self._L.rs_solver_create(kind, n)
        if not self.h:
            raise ValueError(f"unknown solver kind {kind}")
        self.kind = kind
        self.n = n

    def close(self):
        if getattr(self, "h", None):
            self._L.rs_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != 0:
            buf = C.create_string_buffer(512)
            self._L.rs_last_error(self.h, buf, 512)
            pivot = int(self._L.rs_error_pivot(self.h)) if rc == 4 else None
            rcond = self.stats()["rcond"] if rc == 5 else None
            raise SolverError(rc, buf.value.decode(errors="replace"), pivot=pivot, rcond=rcond)

    # --- LinearSolver ---------------------------------------------------------
    def prepare(self, n: int):
        self._L.rs_prepare(self.h, n)
        self.n = n

    def analyze(self, ap, ai):
        ap, ai = _i64(ap), _i64(ai)
        self._check(self._L.rs_analyze(self.h, ap, len(ap), ai, len(ai)))

    def factor(self, ap, ai, ax):
        ap, ai, ax = _i64(ap), _i64(ai), _f64(ax)
        self._check(self._L.rs_factor(self.h, ap, len(ap), ai, len(ai), ax, len(ax)))

    def solve(self, rhs: np.ndarray):
        """Overwrites ``rhs`` (float64, contiguous) with the solution."""
        if rhs.dtype != np.float64 or not rhs.flags.c_contiguous:
            raise TypeError("rhs must be a contiguous float64 array")
        self._check(self._L.rs_solve(self.h, rhs, len(rhs)))

    def tsolve(self, rhs: np.ndarray):
        """rs_tsolve: overwrites ``rhs`` with x of A^T x = rhs (the stored L/U)."""
        if rhs.dtype != np.float64 or not rhs.flags.c_contiguous:
            raise TypeError("rhs must be a contiguous float64 array")
        self._check(self._L.rs_tsolve(self.h, rhs, len(rhs)))

    def tsolve_device(self, rhs_dev, x_dev, stream=None):
        self._check(self._L.rs_tsolve_device(self.h, C.c_void_p(rhs_dev.data_ptr()),
                                             C.c_void_p(x_dev.data_ptr()),
                                             C.c_void_p(_stream_ptr(stream))))

    def reset_pattern(self):
        self._L.rs_reset_pattern(self.h)

    def set_bbd_blocks(self, num_blocks: int):
        """BbdSolver::new(.., num_blocks): target diagonal blocks (BBD kind)."""
        self._L.rs_set_bbd_blocks(self.h, num_blocks)

    def bbd_info(self):
        out = np.zeros(4, np.int64)
        self._check(self._L.rs_bbd_info(self.h, out))
        return dict(blocks=int(out[0]), border=int(out[1]), schur=int(out[2]),
                    factor_count=int(out[3]))

    def set_bbd_mfma(self, mode: int):
        """Dense Schur product on FP64 MFMA tiles (BBD kind): 0 off (the
        reference's term order, bit-exact), 1 where the packed operands are
        dense enough, 2 always. Agrees with the reference to rounding only."""
        self._L.rs_set_bbd_mfma(self.h, mode)

    def bbd_mfma_info(self):
        out = np.zeros(4, np.int64)
        self._check(self._L.rs_bbd_mfma_info(self.h, out))
        return dict(mode=int(out[0]), used=bool(out[1]), k=int(out[2]),
                    c_density=out[3] / 1000.0)

    def name(self) -> str:
        return self._L.rs_solver_name(self.h).decode()

    def set_parallel_threads(self, threads: int):
        self._L.rs_set_parallel_threads(self.h, threads)

    # --- NativeKluSolver extras ------------------------------------------------
    def set_config(self, pivot_tol=0.001, ordering=0, use_btf=True):
        self._L.rs_set_config(self.h, pivot_tol, ordering, int(use_btf))

    def set_autotune(self, on=True):
        """Time the kernel launch shapes on the first refactor/solve of each
        pattern and keep the fastest (default on)."""
        self._L.rs_set_autotune(self.h, int(bool(on)))

    def set_supernodal(self, mode: int = 1):
        """0 off, 1 SparseLU patterns of n >= 20,000 (default), 2 every
        SparseLU pattern: factor and solve by supernodal panels on the
        device (an unsymmetric pattern on its symmetric closure). The
        transient adjoint's own solvers follow this setting."""
        self._L.rs_set_supernodal(self.h, int(mode))

    def supernodal_info(self) -> dict:
        out = np.zeros(9, np.int64)
        self._L.rs_supernodal_info(self.h, out)
        keys = ["active", "supernodes", "levels", "contributions", "panel_doubles", "device_bytes",
                "updates", "max_w", "gather_updates"]
        return dict(zip(keys, map(int, out)))

    def set_growth_check(self, limit: float):
        """Re-factor with new pivots when a device refactor's growth
        max|U|/max|A| exceeds `limit` (0 disables: the reference's behaviour)."""
        self._L.rs_set_growth_check(self.h, float(limit))

    def growth(self) -> dict:
        return dict(fallbacks=int(self._L.rs_growth_fallbacks(self.h)),
                    last=float(self._L.rs_last_growth(self.h)))

    def set_launch_shape(self, ref_threads, ref_mode, sol_threads, sol_mode):
        """ref_mode: 0/1 column kernel (1: LDS work vectors), 2/3 right-looking
        schedule (3: values in LDS), 4 step program; sol_mode: 0 solve vector
        in global memory, 1 in LDS, 2 step program, 3 pipelined level sweeps."""
        self._check(self._L.rs_set_launch_shape(self.h, ref_threads, int(ref_mode), sol_threads,
                                                int(sol_mode)))

    def launch_shape(self) -> dict:
        out = (C.c_int64 * 6)()
        flags = self._L.rs_launch_shape(self.h, out)
        return dict(ref_threads=out[0], ref_mode=int(out[1]), sol_threads=out[2],
                    sol_mode=int(out[3]), ref_us=out[4] / 1000.0, sol_us=out[5] / 1000.0,
                    tuned_ref=bool(flags & 1), tuned_sol=bool(flags & 2), forced=bool(flags & 4))

    def stats(self) -> dict:
        s = _lib.SolverStats()
        self._L.rs_get_stats(self.h, C.byref(s))
        return s.as_dict()

    def rcond(self) -> float:
        return self.stats()["rcond"]

    # --- device-resident hot path (torch tensors on cuda) ---------------------
    def refactor_device(self, ax_dev, stream=None):
        self._check(self._L.rs_refactor_device(self.h, C.c_void_p(ax_dev.data_ptr()),
                                               C.c_void_p(_stream_ptr(stream))))

    def solve_device(self, rhs_dev, x_dev, stream=None):
        self._check(self._L.rs_solve_device(self.h, C.c_void_p(rhs_dev.data_ptr()),
                                            C.c_void_p(x_dev.data_ptr()),
                                            C.c_void_p(_stream_ptr(stream))))

    def solve_device_noalloc(self, rhs_dev, x_dev, stream=None):
        """rs_solve_device_noalloc: the device solve with no allocation."""
        self._check(self._L.rs_solve_device_noalloc(self.h, C.c_void_p(rhs_dev.data_ptr()),
                                                    C.c_void_p(x_dev.data_ptr()),
                                                    C.c_void_p(_stream_ptr(stream))))

    def factor_pattern(self):
        """One-block factor pattern: col_perm[original] = position, L/U
        column pointers and row indices (position space)."""
        sz = np.zeros(3, np.int64)
        self._check(self._L.rs_get_factor_pattern(self.h, sz, None, None, None, None, None))
        n, nl, nu = map(int, sz)
        d = dict(col_perm=np.zeros(n, np.int64), l_col_ptr=np.zeros(n + 1, np.int64),
                 l_row_idx=np.zeros(max(nl, 1), np.int64), u_col_ptr=np.zeros(n + 1, np.int64),
                 u_row_idx=np.zeros(max(nu, 1), np.int64))
        self._check(self._L.rs_get_factor_pattern(self.h, sz, d["col_perm"], d["l_col_ptr"], d["l_row_idx"],
                                                  d["u_col_ptr"], d["u_row_idx"]))
        d["l_row_idx"], d["u_row_idx"] = d["l_row_idx"][:nl], d["u_row_idx"][:nu]
        d["n"] = n
        return d

    def factor_values(self):
        sz = np.full(3, -1, np.int64)
        self._L.rs_get_factor_pattern(self.h, sz, None, None, None, None, None)
        if sz[0] < 0:  # a BbdSolver with several blocks: no factor values here
            self._check(self._L.rs_get_factor_values(self.h, None, None, None))
        dim, n = int(sz[0]), (int(sz[1]), int(sz[2]))
        lx = np.zeros(max(n[0], 1))
        ux = np.zeros(max(n[1], 1))
        ud = np.zeros(max(dim, 1))
        self._check(self._L.rs_get_factor_values(self.h, lx, ux, ud))
        return lx[: n[0]], ux[: n[1]], ud[:dim]


def _stream_ptr(stream) -> int:
    if stream is None:
        return 0
    return int(getattr(stream, "cuda_stream", stream))


# --- host-side analysis (no device memory) -------------------------------------
def host_amd_order(n, ap, ai):
    perm, inv = np.zeros(max(n, 1), np.int64), np.zeros(max(n, 1), np.int64)
    _lib.lib().rs_host_amd_order(n, _i64(ap), _i64(ai), perm, inv)
    return perm[:n], inv[:n]


def host_btf(n, ap, ai):
    rp, cp, bp = np.zeros(max(n, 1), np.int64), np.zeros(max(n, 1), np.int64), np.zeros(n + 1, np.int64)
    rank = C.c_int64()
    nb = _lib.lib().rs_host_btf(n, _i64(ap), _i64(ai), rp, cp, bp, C.byref(rank))
    return dict(row_perm=rp[:n], col_perm=cp[:n], block_ptr=bp[: nb + 1], num_blocks=int(nb),
                structural_rank=rank.value)


class GpuSolverBatch:
    """rs_solver_batch_*: many CSC matrices (patterns may differ), each with
    its first factor (pivots) on the host, then refactored in one launch and
    solved in another from device vectors concatenated by matrix."""

    def __init__(self, mats, kind=NATIVEKLU):
        """mats: list of (ap, ai, ax) CSC triples."""
        self._L = _lib.lib()
        self.h = None
        if not mats:
            raise SolverError(6, "Invalid matrix: empty batch")
        n = np.array([len(ap) - 1 for ap, _, _ in mats], np.int64)
        ap_cat = np.concatenate([_i64(ap) for ap, _, _ in mats])
        ai_cat = np.concatenate([_i64(ai) for _, ai, _ in mats])
        ax_cat = np.concatenate([_f64(ax) for _, _, ax in mats])
        self.h = self._L.rs_solver_batch_create(kind, len(mats), n, ap_cat, ai_cat, ax_cat)
        err = self.last_error()
        if err:
            raise SolverError(6, err)
        sz = np.zeros(3, np.int64)
        self.ax_off = np.zeros(len(mats), np.int64)
        self.vec_off = np.zeros(len(mats), np.int64)
        self._L.rs_solver_batch_layout(self.h, sz, self.ax_off, self.vec_off)
        self.count, self.nnz_total, self.n_total = map(int, sz)
        self.n = n

    def last_error(self):
        buf = C.create_string_buffer(512)
        self._L.rs_solver_batch_last_error(self.h, buf, 512)
        return buf.value.decode()

    def _check(self, rc):
        if rc:
            raise SolverError(rc, self.last_error())

    def close(self):
        if getattr(self, "h", None):
            self._L.rs_solver_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def refactor_device(self, ax_dev, stream=None):
        self._check(self._L.rs_solver_batch_refactor(self.h, C.c_void_p(ax_dev.data_ptr()),
                                                     C.c_void_p(_stream_ptr(stream))))

    def solve_device(self, rhs_dev, x_dev, stream=None):
        self._check(self._L.rs_solver_batch_solve(self.h, C.c_void_p(rhs_dev.data_ptr()),
                                                  C.c_void_p(x_dev.data_ptr()),
                                                  C.c_void_p(_stream_ptr(stream))))

    def factor_values(self, m, nnz_l, nnz_u):
        lx, ux, ud = np.zeros(max(nnz_l, 1)), np.zeros(max(nnz_u, 1)), np.zeros(int(self.n[m]))
        self._check(self._L.rs_solver_batch_factor_values(self.h, m, lx, ux, ud))
        return lx[:nnz_l], ux[:nnz_u], ud


class HostFactor:
    """First (pivoting) factor as computed on the host before the GPU takes over."""

    def __init__(self, n, ap, ai, ax, kind=NATIVEKLU, pivot_tol=0.001, ordering=0, use_btf=True):
        self._L = _lib.lib()
        self.h = self._L.rs_host_factor_create(kind, n, _i64(ap), _i64(ai), _f64(ax), pivot_tol,
                                               ordering, int(use_btf))

    def __del__(self):
        if getattr(self, "h", None):
            self._L.rs_host_factor_destroy(self.h)
            self.h = None

    def nblocks(self):
        return int(self._L.rs_host_factor_nblocks(self.h))

    def block(self, b=0):
        n, nl, nu = C.c_int64(), C.c_int64(), C.c_int64()
        self._L.rs_host_factor_block_sizes(self.h, b, C.byref(n), C.byref(nl), C.byref(nu))
        n, nl, nu = n.value, nl.value, nu.value
        d = dict(col_perm=np.zeros(n, np.int64), row_perm=np.zeros(n, np.int64),
                 l_col_ptr=np.zeros(n + 1, np.int64), l_row_idx=np.zeros(max(nl, 1), np.int64),
                 u_col_ptr=np.zeros(n + 1, np.int64), u_row_idx=np.zeros(max(nu, 1), np.int64),
                 l_values=np.zeros(max(nl, 1)), u_values=np.zeros(max(nu, 1)), u_diag=np.zeros(n))
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
        self._L.rs_host_factor_block_get(self.h, b, *[ptr(d[k]) for k in
                                         ["col_perm", "row_perm", "l_col_ptr", "l_row_idx",
                                          "u_col_ptr", "u_row_idx", "l_values", "u_values",
                                          "u_diag"]])
        for k, m in (("l_row_idx", nl), ("l_values", nl), ("u_row_idx", nu), ("u_values", nu)):
            d[k] = d[k][:m]
        d["n"] = n
        return d

    def plan_stats(self):
        out = np.zeros(8, np.int64)
        self._L.rs_host_factor_plan_stats(self.h, out)
        keys = ["factor_levels", "max_level_width", "lsolve_levels", "usolve_levels",
                "scatter_runs", "offdiag", "nnz_l", "nnz_u"]
        return dict(zip(keys, map(int, out)))


def host_stamp_pattern(circuit, mode=0):
    """Unmerged CSC pattern the stamp loop produces (host replay, no device)."""
    L = _lib.lib()
    dev = np.ascontiguousarray(circuit.devices)
    n, nnz = C.c_int64(), C.c_int64()
    rc = L.rs_host_stamp_pattern(dev.ctypes.data_as(C.c_void_p), circuit.ndev, circuit.node_count,
                                 mode, C.byref(n), C.byref(nnz), None, None)
    if rc:
        raise SolverError(rc, "unsupported device")
    ap = np.zeros(n.value + 1, np.int64)
    ai = np.zeros(max(nnz.value, 1), np.int64)
    L.rs_host_stamp_pattern(dev.ctypes.data_as(C.c_void_p), circuit.ndev, circuit.node_count, mode,
                            None, None, ap.ctypes.data_as(C.c_void_p), ai.ctypes.data_as(C.c_void_p))
    return ap, ai[: nnz.value]


def host_sp_selftest(n, ap, ai, ax0, ax1, rhs, kind=NATIVEKLU, own_plan=False):
    """Host model of the GPU step programs: first factor on ax0 (pivots),
    refactor on ax1, solve rhs. own_plan: the programs of the single-matrix
    solver's plan (positions after the pivots) instead of the batched
    kernel's identity-pivot plan. Returns (x, stats)."""
    L = _lib.lib()
    x = np.zeros(n)
    st = np.zeros(8, np.int64)
    k = kind | (0x100 if own_plan else 0)
    rc = L.rs_host_sp_selftest(k, n, _i64(ap), _i64(ai), _f64(ax0), _f64(ax1), _f64(rhs), x, st)
    if rc:
        raise SolverError(rc, "step program")
    keys = ["refactor_steps", "refactor_ops", "solve_steps", "solve_ops", "refactor_slow_chunks",
            "solve_slow_chunks", "refactor_forwarded", "solve_forwarded"]
    return x, dict(zip(keys, map(int, st)))


def host_lp_selftest(n, ap, ai, ax, rhs, kind=NATIVEKLU, ordering=0, transposed=False):
    """Host model of the flat pipelined solve (solve mode 3): first factor on
    ax (with the given ordering), then the level records run level by level
    (transposed: the records of J^T x = rhs used by the adjoint). Returns (x,
    stats), or raises SolverError when the plan has no level records."""
    L = _lib.lib()
    x = np.zeros(n)
    st = np.zeros(5, np.int64)
    k = kind | (ordering << 8) | (0x10000 if transposed else 0)
    rc = L.rs_host_lp_selftest(k, n, _i64(ap), _i64(ai), _f64(ax), _f64(rhs), x, st)
    if rc:
        raise SolverError(rc, "level records")
    keys = ["levels", "max_level_entries", "entries", "rows", "sweep_levels"]
    return x, dict(zip(keys, map(int, st)))


def host_rl_selftest(n, ap, ai, ax0, ax1, kind=NATIVEKLU):
    """Host model of the right-looking refactor schedule: first factor on ax0
    (pivots), refactor on ax1 phase by phase. Returns (lx, ux, ud, stats) in
    the solver's position order."""
    L = _lib.lib()
    hf = HostFactor(n, ap, ai, ax0, kind)
    st0 = hf.plan_stats()
    lx, ux, ud = np.zeros(max(st0["nnz_l"], 1)), np.zeros(max(st0["nnz_u"], 1)), np.zeros(max(n, 1))
    st = np.zeros(8, np.int64)
    rc = L.rs_host_rl_selftest(kind, n, _i64(ap), _i64(ai), _f64(ax0), _f64(ax1), lx, ux, ud, st)
    if rc:
        raise SolverError(rc, "no right-looking schedule")
    keys = ["phases", "tasks", "updates", "long_tasks", "column_levels", "nnz_l", "nnz_u", "slots"]
    stats = dict(zip(keys, map(int, st)))
    return lx[: stats["nnz_l"]], ux[: stats["nnz_u"]], ud[:n], stats


def host_bbd_decompose(n, ap, ai, num_blocks):
    """The library's bbd_decompose (host integer work) as a dict."""
    L = _lib.lib()
    perm = np.zeros(max(n, 1), np.int64)
    bp = np.zeros(num_blocks + 2, np.int64)
    b0 = C.c_int64()
    nb = L.rs_host_bbd_decompose(n, _i64(ap), _i64(ai), num_blocks, perm, bp, C.byref(b0))
    return dict(num_blocks=int(nb), perm=perm[:n], block_ptr=bp[: nb + 1], border_start=b0.value,
                border_size=n - b0.value if nb > 1 else 0)


def host_sn_selftest(n, ap, ai, ax=None, rhs=None, hot_min=0, transpose=False):
    """Host execution of the supernodal schedule (the order of the GPU
    kernels): SparseLU symbolic factor, supernodal plan and, with ax, the
    factor of ax and the solve of rhs (transpose: the plan of A^T on A's
    symbolic factor, x solving A^T x = rhs; lx and ux are then not
    computed). Returns (x, lx, ux, ud, stats); the arrays are None when ax
    is None."""
    L = _lib.lib()
    st = np.zeros(13, np.int64)
    ap, ai = _i64(ap), _i64(ai)
    fl = 1 if transpose else 0
    rc = L.rs_host_sn_selftest(n, ap, ai, None, None, None, None, None, None, hot_min, fl, st)
    if rc:
        raise SolverError(rc, "no supernodal plan")
    keys = ["supernodes", "levels", "contributions", "panel_doubles", "max_w", "max_h", "updates",
            "gather_tiles", "nnz_l", "nnz_u", "hot_entries", "tile_records", "symmetric"]
    stats = dict(zip(keys, map(int, st)))
    if ax is None:
        return None, None, None, None, stats
    ax = _f64(ax)
    rhs = _f64(rhs if rhs is not None else np.zeros(n))
    x = np.zeros(max(n, 1))
    lx, ux = np.zeros(max(stats["nnz_l"], 1)), np.zeros(max(stats["nnz_u"], 1))
    ud = np.zeros(max(n, 1))
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    rc = L.rs_host_sn_selftest(n, ap, ai, p(ax), p(rhs), p(x), p(lx), p(ux), p(ud), hot_min, fl, st)
    if rc:
        raise SolverError(rc, "no supernodal plan")
    return x[:n], lx[: stats["nnz_l"]], ux[: stats["nnz_u"]], ud[:n], stats
