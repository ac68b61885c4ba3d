"""Loads the reference's Matrix Market fixtures (tests/fixtures/matrices,
copied as data into tests/golden/matrices) the way its BBD benchmark tests
do (crates/sim-core/tests/bbd_benchmark_tests.rs:32-140): entries sorted by
(column, row) with a stable sort, duplicates summed in file order, a
`pattern` matrix given 10.0 on the diagonal and -0.1 elsewhere, a
`symmetric` one mirrored. Also that file's right-hand side generator
(:253-264) and its relative residual (:147-171). Test data only."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MATRICES = os.path.join(HERE, "matrices")


def parse_matrix_market(name):
    with open(os.path.join(MATRICES, name + ".mtx")) as f:
        lines = f.read().splitlines()
    header = lines[0].lower()
    assert lines[0].startswith("%%MatrixMarket")
    pattern, symmetric = "pattern" in header, "symmetric" in header
    i = 1
    while lines[i].startswith("%"):
        i += 1
    nr, nc, _ = (int(t) for t in lines[i].split()[:3])
    assert nr == nc
    ents = []
    for line in lines[i + 1:]:
        t = line.strip()
        if not t or t.startswith("%"):
            continue
        p = t.split()
        if len(p) < 2:
            continue
        r, c = int(p[0]) - 1, int(p[1]) - 1
        v = (10.0 if r == c else -0.1) if pattern else float(p[2])
        ents.append((r, c, v))
        if symmetric and r != c:
            ents.append((c, r, v))
    ents.sort(key=lambda e: (e[1], e[0]))  # stable, like sort_by
    ded = []
    for r, c, v in ents:
        if ded and ded[-1][0] == r and ded[-1][1] == c:
            ded[-1][2] += v
        else:
            ded.append([r, c, v])
    ap = np.zeros(nr + 1, np.int64)
    for _, c, _ in ded:
        ap[c + 1] += 1
    ap = np.cumsum(ap).astype(np.int64)
    ai = np.array([e[0] for e in ded], np.int64)
    ax = np.array([e[2] for e in ded], np.float64)
    return nr, ap, ai, ax


def generate_rhs(n, seed=42):
    """The 64-bit LCG of generate_rhs: v = (state >> 33) / 2^31 - 1."""
    out = np.empty(n)
    s = seed
    for i in range(n):
        s = (s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        out[i] = float(s >> 33) / float(1 << 31) - 1.0
    return out


def relative_residual(n, ap, ai, ax, x, b):
    r = np.array(b, np.float64).copy()
    for col in range(n):
        for q in range(ap[col], ap[col + 1]):
            r[ai[q]] -= ax[q] * x[col]
    rn, bn = np.sqrt((r * r).sum()), np.sqrt((np.asarray(b) ** 2).sum())
    return rn if bn < 1e-30 else rn / bn


#: the reference's fixtures: (name, n, BBD-vs-SparseLU tolerance)
FIXTURES = [("adder_dcop_01", 1813, 1e-3), ("circuit_1", 2624, 1e-3), ("rajat01", 6833, 0.1)]
