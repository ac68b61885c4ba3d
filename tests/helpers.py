"""Shared helpers for the parity tests."""
import numpy as np


def same_bits(a, b):
    """Bit-identical float64 arrays. NaNs count as equal whatever their sign
    and payload: x86 and the GPU produce different default NaN encodings, and
    IEEE 754 leaves the payload of an invalid operation unspecified."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return np.array_equal(a[~na].view(np.int64), b[~nb].view(np.int64))


def random_csc(rs, n, dens, sym=None, diag=None, dup=None):
    """Random CSC pattern (sorted rows, optional duplicate entries)."""
    M = rs.rand(n, n) < dens
    if sym if sym is not None else rs.rand() < 0.5:
        M = M | M.T
    if diag if diag is not None else rs.rand() < 0.7:
        np.fill_diagonal(M, True)
    cols = [np.nonzero(M[:, j])[0] for j in range(n)]
    if dup if dup is not None else rs.rand() < 0.3:
        cols = [np.sort(np.concatenate([c, c[:1]])) for c in cols]
    ap = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int64)
    ai = np.concatenate(cols + [np.zeros(0, np.int64)]).astype(np.int64)
    return ap, ai


def csc_dense(ap, ai, ax, n):
    A = np.zeros((n, n))
    for j in range(n):
        for q in range(ap[j], ap[j + 1]):
            A[ai[q], j] += ax[q]
    return A
