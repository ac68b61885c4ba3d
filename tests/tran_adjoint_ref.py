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

from oracle import oracle as O
from rustspice_amd import netlists as NL
from tests.helpers import csc_dense

DC, TRAN = 0, 1


def _F(c, x, mode, t, dt, ts):
    """F(x) = A(x) x - rhs(x) with the history ts = (cv, ci, hv, hi)."""
    nc = c.node_count
    tstate = None if mode == DC else (ts[0], ts[1], ts[2], ts[3], 1)
    ap, ai, ax, rhs, n = O.build_mna(c, x=x[:nc], mode=mode, t=t, dt=dt, tstate=tstate)
    xx = np.zeros(n)
    m = min(n, len(x))
    xx[:m] = x[:m]
    return csc_dense(ap, ai, ax, n) @ xx - rhs, n


def _jac(f, x, h_rel=1e-7):
    f0 = f(x)
    J = np.zeros((len(f0), len(x)))
    for j in range(len(x)):
        h = h_rel * max(1.0, abs(x[j]))
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        J[:, j] = (f(xp) - f(xm)) / (2.0 * h)
    return J


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
        dx = np.linalg.solve(csc_dense(ap, ai, ax, nn), rhs) - x
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
    """History in force during step k + 1, from x_0..x_k (for differentiation)."""
    nd = c.ndev
    cv, ci = np.zeros(nd), np.zeros(nd)
    hv, hi = np.zeros(nd, np.uint8), np.zeros(nd, np.uint8)
    for q in _caps(c):
        cv[q], hv[q] = _vc(c, q, xs[0]), 1
    for j in range(1, k + 1):
        for q in _caps(c):
            v = _vc(c, q, xs[j])
            ci[q] = c.devices[q]["value"] * (v - cv[q]) / dt
            cv[q], hi[q] = v, 1
    return cv, ci, hv, hi


def adjoint(c, dt, nsteps, probe, params):
    """dx_N[probe]/dp for params = [(device, getter, setter)], by the dense
    discrete adjoint. Returns (dp, lambdas, xs)."""
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

            D = _jac(fj, xs[k])
            rhs -= D.T @ lam[j]
        Jk = _jac(lambda v, k=k: Fk(k, v, (xs[k - 1] if k >= 1 else None,
                                           xs[k - 2] if k >= 2 else None)), xs[k])
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
