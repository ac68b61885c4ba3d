"""Dense FP64 CPU reference of the fixed-step trapezoidal transient and its
discrete adjoint (the checker for rs_tran_adjoint; no reference counterpart).

Its Jacobians are central differences, so the circuits it checks must keep
every MOSFET away from the level-1 model's kink at vds = 0 (the reference
swaps drain and source there without reversing the current, evaluate.rs):
rc_mesh_inverters(..., load=R) gives each tap a DC path to ground.

Forward: the DC OP, then TRAP steps of dt, each solved to the exact root of
the oracle's stamped currents F_k(x) = A(x) x - rhs(x) (oracle.build_mna with
the capacitor history), with the accepted-step history update of
stamp.rs:615-668 between steps (cap_v = v_k, cap_i = C (v_k - v_{k-1}) / dt).

Adjoint: J_k^T lambda_k = e_probe [k = N] - (dF_{k+1}/dx_k)^T lambda_{k+1}
- (dF_{k+2}/dx_k)^T lambda_{k+2}, with every partial derivative taken by
central differences of the oracle's F (J_k in x_k; the cross-step terms
through the capacitor history the oracle stamps), and
dx_N[probe]/dp = -sum_k lambda_k^T dF_k/dp with dF_k/dp by central
differences too. Nothing here restates the GPU's closed forms.
"""
import numpy as np
import scipy.sparse as sp

from oracle import oracle as O
from rustspice_amd import netlists as NL

DC, TRAN = 0, 1


def _dense(ap, ai, ax, n):
    """The CSC matrix as a dense array, duplicates summed in CSC order."""
    A = np.zeros((n, n))
    np.add.at(A, (ai[: ap[n]], np.repeat(np.arange(n), np.diff(ap[: n + 1]))), ax[: ap[n]])
    return A


def _F(c, x, mode, t, dt, ts):
    """F(x) = A(x) x - rhs(x) with the history ts = (cv, ci, hv, hi)."""
    nc = c.node_count
    tstate = None if mode == DC else (ts[0], ts[1], ts[2], ts[3], 1)
    ap, ai, ax, rhs, n = O.build_mna(c, x=x[:nc], mode=mode, t=t, dt=dt, tstate=tstate)
    xx = np.zeros(n)
    m = min(n, len(x))
    xx[:m] = x[:m]
    return sp.csc_matrix((ax[: ap[n]], ai[: ap[n]], ap[: n + 1]), shape=(n, n)) @ xx - rhs, n


def _jac(f, x, h_rel=1e-7):
    """Central-difference Jacobian, one column at a time."""
    f0 = f(x)
    J = np.zeros((len(f0), len(x)))
    for j in range(len(x)):
        h = h_rel * max(1.0, abs(x[j]))
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        J[:, j] = (f(xp) - f(xm)) / (2.0 * h)
    return J


def _coloring(rows_of, ncol):
    """Greedy column colouring: two columns that share a row never share a
    colour (Curtis, Powell and Reid), so one pair of evaluations per colour
    gives every column of the group's difference quotients."""
    cols_of_row = {}
    for j in range(ncol):
        for i in rows_of[j]:
            cols_of_row.setdefault(int(i), []).append(j)
    color = np.full(ncol, -1, np.int64)
    for j in range(ncol):
        used = {color[q] for i in rows_of[j] for q in cols_of_row[int(i)] if color[q] >= 0}
        k = 0
        while k in used:
            k += 1
        color[j] = k
    return color


def _jac_colored(f, x, rows_of, color, nodes, defect, h_rel=1e-7):
    """Central-difference Jacobian by colour groups; rows_of[j] must hold
    every row that column j can change except row 0 (columns beyond it:
    none). Row 0, the ground node's current balance, couples every column:
    every device current that enters a node leaves another, so the node
    rows sum to what the stamps add beyond such pairs (defect: the row sum
    of the stamped matrix's node rows, a tie of the ground node to itself),
    and row 0 is defect minus the sum of the other node rows."""
    f0 = f(x)
    J = np.zeros((len(f0), len(x)))
    ncol = min(len(x), len(rows_of))
    for k in range(int(color[:ncol].max()) + 1 if ncol else 0):
        cols = [j for j in range(ncol) if color[j] == k and len(rows_of[j])]
        if not cols:
            continue
        h = np.array([h_rel * max(1.0, abs(x[j])) for j in cols])
        xp, xm = x.copy(), x.copy()
        xp[cols] += h
        xm[cols] -= h
        d = f(xp) - f(xm)
        for j, hj in zip(cols, h):
            r = rows_of[j]
            r = r[r < len(f0)]
            J[r, j] = d[r] / (2.0 * hj)
    J[0, :] = defect[: len(x)] - J[1:nodes, :].sum(axis=0)
    return J


def _pattern(c, mode, states, n_x, nodes_only=False):
    """rows_of[j]: the rows of column j in the union of the mode's stamped
    patterns at the given states (the Jacobian's entries and the capacitor
    history's all lie in it) less the ground row, a colouring of the
    columns, and the node count. nodes_only: only node columns (what a
    later step's history reads of an earlier x)."""
    hist = (np.zeros(c.ndev), np.zeros(c.ndev), np.ones(c.ndev, np.uint8), np.ones(c.ndev, np.uint8), 1)
    rows = [set() for _ in range(n_x)]
    defect = np.zeros(n_x)
    for q, x in enumerate(states):
        ap, ai, ax, _, n = O.build_mna(c, x=x[: c.node_count], mode=mode, t=0.0,
                                       dt=1e-12 if mode == TRAN else 0.0, tstate=None if mode == DC else hist)
        for j in range(min(n, n_x)):
            rows[j].update(int(i) for i in ai[ap[j]: ap[j + 1]])
        if q == 0 and not nodes_only:  # the capacitor history's currents come in pairs: no defect
            m = min(n, n_x)
            defect[:m] = _dense(ap, ai, ax, n)[: c.node_count, :m].sum(axis=0)
    rows_of = [np.array(sorted(r - {0}), np.int64) if (not nodes_only or j < c.node_count)
               else np.zeros(0, np.int64) for j, r in enumerate(rows)]
    return rows_of, _coloring(rows_of, n_x), c.node_count, defect


def _newton(c, x0, mode, t, dt, ts, n):
    """Root of F by the reference's Newton matrix (its fixed point is the root
    of F whether or not the matrix is F's exact Jacobian), to round-off."""
    x = np.zeros(n)
    m = min(n, len(x0))
    x[:m] = x0[:m]
    nc = c.node_count
    tstate = None if mode == DC else (ts[0], ts[1], ts[2], ts[3], 1)
    prev = np.inf
    for _ in range(200):
        ap, ai, ax, rhs, nn = O.build_mna(c, x=x[:nc], mode=mode, t=t, dt=dt, tstate=tstate)
        dx = np.linalg.solve(_dense(ap, ai, ax, nn), rhs) - x
        x = x + dx
        mdx, mx = np.abs(dx).max(), max(1.0, np.abs(x).max())
        if mdx <= 1e-15 * mx or (mdx <= 1e-12 * mx and mdx >= prev):  # converged, or at round-off
            break
        prev = mdx
    return x


def _caps(c):
    return [k for k in range(c.ndev) if c.devices[k]["kind"] == NL.RS_DEV_C]


def _vc(c, k, x):
    a, b = c.devices[k]["node"][0], c.devices[k]["node"][1]
    return x[a] - x[b]


def forward(c, dt, nsteps):
    """x_0 .. x_N and the histories ts_k (in force during step k)."""
    nd = c.ndev
    _, n_dc = _F(c, np.zeros(c.node_count), DC, 0.0, 0.0, None)
    xs = [_newton(c, O.dc_op(c)[0], DC, 0.0, 0.0, None, n_dc)]
    cv, ci = np.zeros(nd), np.zeros(nd)
    hv, hi = np.zeros(nd, np.uint8), np.zeros(nd, np.uint8)
    for k in _caps(c):
        cv[k], hv[k] = _vc(c, k, xs[0]), 1
    hist, t = [None], 0.0
    n_tr = None
    for step in range(1, nsteps + 1):
        t = t + dt
        ts = (cv.copy(), ci.copy(), hv.copy(), hi.copy())
        if n_tr is None:
            _, n_tr = _F(c, xs[-1], TRAN, t, dt, ts)
        xs.append(_newton(c, xs[-1], TRAN, t, dt, ts, n_tr))
        hist.append(ts)
        for k in _caps(c):
            v = _vc(c, k, xs[-1])
            ci[k] = c.devices[k]["value"] * (v - (cv[k] if hv[k] else v)) / dt
            cv[k], hi[k] = v, 1
    return xs, hist


def _hist_after(c, xs, k, dt):
    """History in force during step k + 1, from x_0..x_k (for differentiation;
    the update of forward(), over all capacitors at once)."""
    nd = c.ndev
    cv, ci = np.zeros(nd), np.zeros(nd)
    hv, hi = np.zeros(nd, np.uint8), np.zeros(nd, np.uint8)
    q = np.nonzero(c.devices["kind"] == NL.RS_DEV_C)[0]
    if len(q):
        a, b = c.devices["node"][q, 0], c.devices["node"][q, 1]
        cap = c.devices["value"][q]
        cv[q], hv[q] = xs[0][a] - xs[0][b], 1
        for j in range(1, k + 1):
            v = xs[j][a] - xs[j][b]
            ci[q] = cap * (v - cv[q]) / dt
            cv[q], hi[q] = v, 1
    return cv, ci, hv, hi


def adjoint(c, dt, nsteps, probe, params, colored=False):
    """dx_N[probe]/dp for params = [(device, getter, setter)], by the dense
    discrete adjoint. colored: Jacobians by column colour groups of the
    stamped pattern (one pair of evaluations per colour instead of per
    column; the instances of a thousand unknowns). Returns (dp, lambdas, xs)."""
    xs, hist = forward(c, dt, nsteps)
    N = nsteps
    times = [0.0]
    for _ in range(N):
        times.append(times[-1] + dt)

    def Fk(k, xk, xprev):
        """F_k with x_{k-1}, x_{k-2} replaced by xprev = (x_{k-1}, x_{k-2})."""
        if k == 0:
            return _F(c, xk, DC, 0.0, 0.0, None)[0]
        seq = list(xs[: k - 2]) + [xprev[1], xprev[0]] if k >= 2 else [xprev[0]]
        ts = _hist_after(c, seq, k - 1, dt)
        return _F(c, xk, TRAN, times[k], dt, ts)[0]

    pats = {}

    def jac(f, k, cross):
        if not colored:
            return _jac(f, xs[k])
        mode = TRAN if (cross or k > 0) else DC
        key = (mode, len(xs[k]), cross)
        if key not in pats:
            pats[key] = _pattern(c, mode, xs[1:] if mode == TRAN else xs[:1], len(xs[k]), nodes_only=cross)
        return _jac_colored(f, xs[k], *pats[key])

    lam = [None] * (N + 1)
    for k in range(N, -1, -1):
        n = len(xs[k])
        rhs = np.zeros(n)
        if k == N:
            rhs[probe] = 1.0
        for j in (k + 1, k + 2):
            if j > N:
                continue
            xprev = (xs[j - 1], xs[j - 2] if j >= 2 else None)
            slot = j - k - 1  # x_k is x_{j-1} (slot 0) or x_{j-2} (slot 1)

            def fj(v, j=j, xprev=xprev, slot=slot):
                xp = list(xprev)
                w = np.zeros(len(xp[slot]))
                w[: len(v)] = v[: len(w)]
                xp[slot] = w
                return Fk(j, xs[j], xp)

            D = jac(fj, k, True)
            rhs -= D.T @ lam[j]
        Jk = jac(lambda v, k=k: Fk(k, v, (xs[k - 1] if k >= 1 else None,
                                          xs[k - 2] if k >= 2 else None)), k, False)
        lam[k] = np.linalg.solve(Jk.T, rhs)
    dp = []
    for dev, get, put in params:
        # F is linear in C, W, source values and 1/R: a large step keeps the
        # quotient's round-off (|F| eps / h) far below the sensitivity
        p0 = get()
        h = 1e-3 * abs(p0) if p0 != 0.0 else 1e-6
        tot = 0.0
        for k in range(N + 1):
            xprev = (xs[k - 1] if k >= 1 else None, xs[k - 2] if k >= 2 else None)
            put(p0 + h)
            fp = Fk(k, xs[k], xprev)
            put(p0 - h)
            fm = Fk(k, xs[k], xprev)
            put(p0)
            tot -= lam[k] @ ((fp - fm) / (2.0 * h))
        dp.append(tot)
    return np.array(dp), lam, xs


def param(c, k):
    """(getter, setter) of device k's differentiated parameter."""
    d = c.devices[k]
    if d["kind"] == NL.RS_DEV_M:
        return lambda: float(d["p"][4]), lambda v: d["p"].__setitem__(4, v)
    if d["kind"] == NL.RS_DEV_D:
        return lambda: float(d["p"][0]), lambda v: d["p"].__setitem__(0, v)
    return lambda: float(d["value"]), lambda v: d.__setitem__("value", v)


def finite_difference(c, dt, nsteps, probe, k, h_rel=1e-6):
    """Central difference of the polished forward's x_N[probe] in device k's parameter."""
    get, put = param(c, k)
    p0 = get()
    h = h_rel * abs(p0) if p0 != 0.0 else h_rel
    put(p0 + h)
    xp = forward(c, dt, nsteps)[0][-1][probe]
    put(p0 - h)
    xm = forward(c, dt, nsteps)[0][-1][probe]
    put(p0)
    return (xp - xm) / (2.0 * h)
