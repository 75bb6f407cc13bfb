"""Pins the CPU oracle's layered normalized min-sum (CPU only).

The reference selects a "colored layered" GPU BP schedule for n >= 1000
(source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp:3658) inside a decoder
library that is absent from the reference tree, so the schedule's exact
definition is chosen here (greedy first-fit row colouring in row order, rows
of a colour processed as independent layers).  It is pinned here by:
  - the colouring: every colour an independent set, greedy-minimal, order
    sorted by (colour, row), recomputed independently in numpy;
  - closed forms: a single check (one layered pass = one flooding pass, which
    is pinned against the reference), two checks sharing a variable (the
    second row sees the first row's update), damping on exact binary values;
  - order invariance inside a colour (the layers really are independent);
  - cycle-free code: alpha = 1 min-sum reaches the exact max-log APP;
  - DVB-S2 codes: layered converges in markedly fewer iterations than flooding.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from ldpc_helpers import codebook, maxlog_app, nullspace_generator

DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sicc-baseband-proto_amd", "data")


def _csr(rows):
    rp = np.zeros(len(rows) + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    return rp, np.array(sum(rows, []), np.int32)


def _greedy_numpy(n, m, rp, ci):
    rows = [set(ci[rp[r]:rp[r + 1]].tolist()) for r in range(m)]
    color = []
    for r in range(m):
        used = {color[r2] for r2 in range(r) if rows[r] & rows[r2]}
        c = 0
        while c in used:
            c += 1
        color.append(c)
    return np.array(color)


@pytest.mark.parametrize("which", ["ccsds128", "random"])
def test_layer_order_is_a_valid_greedy_colouring(which):
    if which == "ccsds128":
        n, m, rp, ci = oracle.alist_parse(os.path.join(DATA, "CCSDS_ldpc_n128_k64.alist"))
    else:
        rng = np.random.default_rng(3)
        n, m = 60, 40
        rows = [sorted(rng.choice(n, rng.integers(2, 7), replace=False).tolist()) for _ in range(m)]
        rp, ci = _csr(rows)
    order, color = oracle.layer_order(n, m, rp, ci)
    assert np.array_equal(color, _greedy_numpy(n, m, rp, ci))
    assert np.array_equal(np.sort(order), np.arange(m))
    assert np.array_equal(order, np.lexsort((np.arange(m), color)))
    for c in range(color.max() + 1):
        cols = np.concatenate([ci[rp[r]:rp[r + 1]] for r in np.flatnonzero(color == c)])
        assert len(np.unique(cols)) == len(cols)  # independent set


def test_dvbs2_colouring_independent():
    n, m, rp, ci = oracle.alist_parse(os.path.join(DATA, "DVB_S2_short_N16200_rate_1_2.alist"))
    order, color = oracle.layer_order(n, m, rp, ci)
    assert color.max() + 1 == 13
    for c in range(13):
        cols = np.concatenate([ci[rp[r]:rp[r + 1]] for r in np.flatnonzero(color == c)])
        assert len(np.unique(cols)) == len(cols)


def test_single_check_equals_flooding_and_closed_form():
    rp, ci = _csr([[0, 1, 2, 3]])
    llr = np.array([1.5, -0.25, 3.0, -2.0], np.float32)
    lay = oracle.nms_layered(4, 1, rp, ci, llr, max_iter=1, alpha=0.75, early_stop=False)
    code = oracle.Code(4, 1, rp, ci, front_info=False)
    flo = code.nms(llr, max_iter=1, alpha=0.75, early_stop=False)
    assert np.array_equal(lay["app"], flo["app"])
    a = np.abs(llr)
    sgn = np.prod(np.sign(llr))
    exp = [llr[i] + np.float32(0.75) * sgn * np.sign(llr[i]) * np.delete(a, i).min() for i in range(4)]
    assert np.allclose(lay["app"], exp, rtol=0, atol=1e-6)


def test_two_checks_second_row_sees_first_update():
    # rows {0,1,2} and {2,3}: the layered pass updates APP_2 with row 0 before row 1 reads it
    rp, ci = _csr([[0, 1, 2], [2, 3]])
    llr = np.array([2.0, 3.0, -0.5, 1.0], np.float32)
    got = oracle.nms_layered(4, 2, rp, ci, llr, order=np.array([0, 1], np.int32), max_iter=1, alpha=1.0,
                             early_stop=False)["app"]
    # row 0: signs (+,+,-): r_0 = -0.5, r_1 = -0.5, r_2 = +2 -> APP = [1.5, 2.5, 1.5, 1.0]
    # row 1 on q = APP: q_2 = 1.5, q_3 = 1.0 -> r_2 = 1.0, r_3 = 1.5 -> APP_2 = 2.5, APP_3 = 2.5
    assert np.array_equal(got, np.array([1.5, 2.5, 2.5, 2.5], np.float32))
    flo = oracle.Code(4, 2, rp, ci, front_info=False).nms(llr, max_iter=1, alpha=1.0, early_stop=False)["app"]
    assert not np.array_equal(got, flo)  # flooding: row 1 reads the channel value of bit 2
    # the other row order is the other sequential schedule
    rev = oracle.nms_layered(4, 2, rp, ci, llr, order=np.array([1, 0], np.int32), max_iter=1, alpha=1.0,
                             early_stop=False)["app"]
    assert not np.array_equal(got, rev)


def test_damping_and_offset_closed_form():
    rp, ci = _csr([[0, 1, 2]])
    llr = np.array([1.0, -2.0, 4.0], np.float32)
    d, alpha = 0.25, 0.5
    r1 = np.array([-1.0, 0.5, -0.5], np.float32) * np.float32(alpha) * 2  # alpha * sign * min_others
    for iters in (1, 2):
        got = oracle.nms_layered(3, 1, rp, ci, llr, max_iter=iters, alpha=alpha, damping=d, early_stop=False)
        r = np.zeros(3, np.float32)
        for _ in range(iters):
            r = np.float32(1 - d) * r1 + np.float32(d) * r
        assert np.array_equal(got["app"], llr + r)
    off = oracle.nms_layered(3, 1, rp, ci, llr, max_iter=1, alpha=1.0, offset=0.75, early_stop=False)
    # mags (others' min) 2, 1, 1 -> minus 0.75 -> 1.25, 0.25, 0.25; sign product -1
    assert np.array_equal(off["app"], llr + np.array([-1.25, 0.25, -0.25], np.float32))


def test_order_inside_a_colour_does_not_matter():
    n, m, rp, ci = oracle.alist_parse(os.path.join(DATA, "CCSDS_ldpc_n128_k64.alist"))
    code = oracle.Code(n, m, rp, ci, True)
    order, color = code.layer_order()
    rng = np.random.default_rng(4)
    info = rng.integers(0, 2, (20, code.k), dtype=np.uint8)
    cw = code.encode(info).astype(np.float64)
    llr = (2.0 / 0.3 * ((1 - 2 * cw) + np.sqrt(0.3) * rng.standard_normal(cw.shape))).astype(np.float32)
    shuffled = order.copy()
    for c in range(color.max() + 1):
        idx = np.flatnonzero(color[order] == c)
        shuffled[idx] = rng.permutation(order[idx])
    assert not np.array_equal(shuffled, order)
    for f in range(20):
        a = code.nms_layered(llr[f], order, max_iter=30, alpha=0.75, damping=0.1)
        b = code.nms_layered(llr[f], shuffled, max_iter=30, alpha=0.75, damping=0.1)
        assert np.array_equal(a["app"], b["app"]) and a["iterations"] == b["iterations"]


def test_tree_code_reaches_maxlog_map():
    rp, ci = _csr([[0, 1, 2], [2, 3, 4], [4, 5, 6]])
    code = oracle.Code(7, 3, rp, ci, front_info=False)
    book = codebook(nullspace_generator(code.dense_h()))
    rng = np.random.default_rng(12)
    for _ in range(50):
        llr = rng.normal(0.5, 2.0, 7).astype(np.float32)
        r = code.nms_layered(llr, max_iter=12, alpha=1.0, early_stop=False)
        assert np.allclose(r["app"], maxlog_app(book, llr), rtol=1e-5, atol=1e-5)


def test_dvbs2_layered_converges_faster_than_flooding():
    n, m, rp, ci = oracle.alist_parse(os.path.join(DATA, "DVB_S2_short_N16200_rate_1_2.alist"))
    code = oracle.Code(n, m, rp, ci, True)
    assert code.acc and code.k == 7200
    rng = np.random.default_rng(5)
    info = rng.integers(0, 2, (6, code.k), dtype=np.uint8)
    cw = code.encode(info)
    s2 = 1.0 / (2 * 0.5 * 10 ** (2.2 / 10))
    llr = (2 / s2 * ((1 - 2 * cw.astype(np.float64)) + np.sqrt(s2) * rng.standard_normal(cw.shape))).astype(np.float32)
    li, fi = [], []
    for f in range(6):
        a = code.nms_layered(llr[f], max_iter=60, alpha=0.75)
        b = code.nms(llr[f], max_iter=60, alpha=0.75)
        assert a["ok"] and b["ok"] and np.array_equal(a["hard"], cw[f]) and np.array_equal(b["hard"], cw[f])
        li.append(a["iterations"])
        fi.append(b["iterations"])
    assert np.mean(li) < 0.7 * np.mean(fi), (li, fi)
