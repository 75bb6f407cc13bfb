"""GF(2) helpers and independent formulations used to pin the oracle."""
from __future__ import annotations

import itertools

import numpy as np


def gf2_rank(m: np.ndarray) -> int:
    a = (np.asarray(m) & 1).astype(np.uint8).copy()
    rows, cols = a.shape
    r = 0
    for c in range(cols):
        piv = np.nonzero(a[r:, c])[0]
        if piv.size == 0:
            continue
        p = r + piv[0]
        if p != r:
            a[[r, p]] = a[[p, r]]
        nz = np.nonzero(a[:, c])[0]
        for i in nz:
            if i != r:
                a[i] ^= a[r]
        r += 1
        if r == rows:
            break
    return r


def gf2_inv(m: np.ndarray) -> np.ndarray:
    k = m.shape[0]
    a = np.concatenate([(m & 1).astype(np.uint8), np.eye(k, dtype=np.uint8)], axis=1)
    for c in range(k):
        piv = np.nonzero(a[c:, c])[0]
        assert piv.size, "singular"
        p = c + piv[0]
        if p != c:
            a[[c, p]] = a[[p, c]]
        for i in np.nonzero(a[:, c])[0]:
            if i != c:
                a[i] ^= a[c]
    return a[:, k:]


def nullspace_generator(h: np.ndarray) -> np.ndarray:
    """A generator of the null space of H by plain elimination (any basis)."""
    m, n = h.shape
    a = (h & 1).astype(np.uint8).copy()
    piv_cols = []
    r = 0
    for c in range(n):
        nz = np.nonzero(a[r:, c])[0]
        if nz.size == 0:
            continue
        p = r + nz[0]
        a[[r, p]] = a[[p, r]]
        for i in np.nonzero(a[:, c])[0]:
            if i != r:
                a[i] ^= a[r]
        piv_cols.append(c)
        r += 1
        if r == m:
            break
    free = [c for c in range(n) if c not in piv_cols]
    g = np.zeros((len(free), n), np.uint8)
    for i, f in enumerate(free):
        g[i, f] = 1
        for row, pc in enumerate(piv_cols):
            if a[row, f]:
                g[i, pc] = 1
    return g


def osd_expected(gen: np.ndarray, llr: np.ndarray, order: int, max_flip: int, max_tests: int):
    """Order-p OSD stated as linear algebra: the most reliable basis (greedy over
    the stable |L|-descending order), candidates u = msg on the basis, metric =
    sum |L| over disagreements, first minimum wins.  Returns (codeword, metric)."""
    k, n = gen.shape
    rel = np.abs(llr.astype(np.float64))
    hard = (llr < 0).astype(np.uint8)
    perm = np.argsort(-rel, kind="stable")
    basis = []
    for c in perm:
        if gf2_rank(gen[:, basis + [c]]) > len(basis):
            basis.append(int(c))
            if len(basis) == k:
                break
    ginv = gf2_inv(gen[:, basis])
    base = hard[basis]
    weak = np.argsort(rel[basis], kind="stable")
    pool = list(weak[: min(k, max_flip)])
    best, best_m, tests = None, None, 0
    for o in range(order + 1):
        if len(pool) < o:
            break
        for comb in itertools.combinations(range(len(pool)), o):
            if tests >= max_tests:
                break
            msg = base.copy()
            for j in comb:
                msg[pool[j]] ^= 1
            u = (msg.astype(np.int64) @ ginv.astype(np.int64)) % 2
            cw = ((u @ gen.astype(np.int64)) % 2).astype(np.uint8)
            met = float(rel[cw != hard].sum())
            tests += 1
            if best is None or met < best_m:
                best, best_m = cw, met
    return best, best_m


def maxlog_app(codebook: np.ndarray, llr: np.ndarray) -> np.ndarray:
    """Exact max-log APP by enumeration: L_i = min_{c_i=1} sum c.llr - min_{c_i=0} sum c.llr."""
    cost = codebook.astype(np.float64) @ llr.astype(np.float64)
    out = np.zeros(codebook.shape[1])
    for i in range(codebook.shape[1]):
        out[i] = cost[codebook[:, i] == 1].min() - cost[codebook[:, i] == 0].min()
    return out


def codebook(gen: np.ndarray) -> np.ndarray:
    k = gen.shape[0]
    msgs = ((np.arange(2 ** k)[:, None] >> np.arange(k)[None, :]) & 1).astype(np.int64)
    return ((msgs @ gen.astype(np.int64)) % 2).astype(np.uint8)


def random_code(n: int, m: int, rng: np.random.Generator, density: float = 0.35):
    """Random H (m x n) whose last m columns are invertible, as CSR."""
    while True:
        h = (rng.random((m, n)) < density).astype(np.uint8)
        if gf2_rank(h[:, n - m:]) == m and h.sum(axis=1).min() >= 2:
            break
    rp = np.zeros(m + 1, np.int32)
    ci = []
    for r in range(m):
        cols = np.nonzero(h[r])[0]
        ci.extend(cols.tolist())
        rp[r + 1] = len(ci)
    return h, rp, np.array(ci, np.int32)
