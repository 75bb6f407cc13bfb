"""Pins the CPU oracle's LDPC code model, min-sum BP and OSD (CPU only).

Each check is independent of the oracle's own formulas: closed forms, brute
force on tiny codes, linear-algebra restatements, and -- where the reference's
C++ CPU path builds from its sources (oracle/_ref) -- the reference itself.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import refbind
from ldpc_helpers import codebook, gf2_inv, gf2_rank, maxlog_app, nullspace_generator, osd_expected, random_code


@pytest.fixture(scope="module")
def ccsds(ccsds_path):
    n, m, rp, ci = oracle.alist_parse(ccsds_path)
    return oracle.Code(n, m, rp, ci, front_info=True)


def test_alist_ccsds_shape(ccsds_path):
    n, m, rp, ci = oracle.alist_parse(ccsds_path)
    assert (n, m, rp[-1]) == (128, 64, 512)
    assert set(np.diff(rp).tolist()) == {8}
    deg = np.bincount(ci, minlength=n)
    assert deg[:64].tolist() == [5] * 64 and deg[64:].tolist() == [3] * 64


def test_rref_generator_is_systematic_and_in_code(ccsds):
    g = ccsds.generator()
    h = ccsds.dense_h()
    assert ccsds.k == 64
    assert np.array_equal(ccsds.info_cols, np.arange(64))
    assert not ((h.astype(np.int64) @ g.T.astype(np.int64)) % 2).any()
    assert gf2_rank(g) == 64
    assert np.array_equal(g[:, :64], np.eye(64, dtype=np.uint8))


def test_encode_valid_and_linear(ccsds):
    rng = np.random.default_rng(3)
    a = rng.integers(0, 2, (50, 64), dtype=np.uint8)
    b = rng.integers(0, 2, (50, 64), dtype=np.uint8)
    ca, cb, cab = ccsds.encode(a), ccsds.encode(b), ccsds.encode(a ^ b)
    h = ccsds.dense_h().astype(np.int64)
    assert not ((ca.astype(np.int64) @ h.T) % 2).any()
    assert np.array_equal(ca ^ cb, cab)
    assert np.array_equal(ca[:, :64], a)


@pytest.mark.skipif(not refbind.available(), reason="reference CPU path not built (oracle/_ref)")
def test_code_model_matches_reference(ccsds, ccsds_path):
    ref = refbind.RefCode(ccsds_path, True)
    assert (ref.n, ref.m, ref.k) == (ccsds.n, ccsds.m, ccsds.k)
    assert np.array_equal(ref.info_cols, ccsds.info_cols)
    assert np.array_equal(ref.pivot_cols, ccsds.pivot_cols)
    info = np.random.default_rng(5).integers(0, 2, (200, 64), dtype=np.uint8)
    assert np.array_equal(ref.encode(info), ccsds.encode(info))


def test_nms_single_check_closed_form():
    # one check over three bits: r_e = alpha * (product of other signs) * min of other magnitudes
    c = oracle.Code(3, 1, np.array([0, 3]), np.array([0, 1, 2]), front_info=False)
    llr = np.array([2.0, -0.5, 1.5], np.float32)
    r = c.nms(llr, max_iter=1, alpha=0.75)
    expect = np.array([2.0 - 0.75 * 0.5, -0.5 + 0.75 * 1.5, 1.5 - 0.75 * 0.5], np.float32)
    assert np.array_equal(r["app"], expect)
    assert r["ok"] and r["iterations"] == 1 and not r["hard"].any()
    # equal minima: the first minimum edge receives min2 == min1
    llr = np.array([1.0, -1.0, 3.0], np.float32)
    r = c.nms(llr, max_iter=1, alpha=1.0, early_stop=False)
    assert np.array_equal(r["app"], np.array([1.0 - 1.0, -1.0 + 1.0, 3.0 - 1.0], np.float32))
    # offset min-sum: magnitude reduced by the offset, floored at zero
    r = c.nms(np.array([2.0, -0.5, 1.5], np.float32), max_iter=1, alpha=1.0, offset=1.0, early_stop=False)
    assert np.array_equal(r["app"], np.array([2.0, -0.5 + 0.5, 1.5], np.float32))


def test_nms_tree_code_converges_to_maxlog_map():
    # chain of three checks sharing bits 2 and 4: a cycle-free factor graph,
    # on which min-sum (alpha = 1) reaches the exact max-log APP.
    rows = [[0, 1, 2], [2, 3, 4], [4, 5, 6]]
    rp = np.array([0, 3, 6, 9])
    ci = np.array(sum(rows, []))
    c = oracle.Code(7, 3, rp, ci, front_info=False)
    h = c.dense_h()
    book = codebook(nullspace_generator(h))
    assert book.shape[0] == 16
    rng = np.random.default_rng(11)
    for _ in range(50):
        llr = rng.normal(0.5, 2.0, 7).astype(np.float32)
        r = c.nms(llr, max_iter=12, alpha=1.0, early_stop=False)
        exp = llr.astype(np.float64) + 0.0
        exp = maxlog_app(book, llr)
        assert np.allclose(r["app"], exp, rtol=1e-5, atol=1e-5)


def test_nms_clean_codeword_decodes_in_one_iteration(ccsds):
    rng = np.random.default_rng(2)
    info = rng.integers(0, 2, (20, 64), dtype=np.uint8)
    cws = ccsds.encode(info)
    for cw in cws:
        llr = (10.0 * (1.0 - 2.0 * cw)).astype(np.float32)
        r = ccsds.nms(llr, max_iter=50, alpha=0.75)
        assert r["ok"] and r["iterations"] == 1
        assert np.array_equal(r["hard"], cw)


@pytest.mark.skipif(not refbind.available(), reason="reference CPU path not built (oracle/_ref)")
@pytest.mark.parametrize("ebn0,alpha,iters", [(1.5, 0.75, 50), (3.0, 0.75, 50), (2.0, 0.8, 8)])
def test_nms_bit_exact_vs_reference_min_sum(ccsds, ccsds_path, ebn0, alpha, iters):
    ref = refbind.RefCode(ccsds_path, True)
    rng = np.random.default_rng(int(ebn0 * 10) + iters)
    info = rng.integers(0, 2, (1500, 64), dtype=np.uint8)
    llr = oracle.awgn_llr(ccsds.encode(info), ebn0, 0.5, rng)
    rcode, rinfo, rok, rit = ref.min_sum(llr, iters, alpha)
    for i in range(llr.shape[0]):
        o = ccsds.nms(llr[i], max_iter=iters, alpha=alpha)
        assert o["ok"] == bool(rok[i])
        assert o["iterations"] == rit[i]
        assert np.array_equal(o["hard"], rcode[i])


def test_best_syndrome_tracking(ccsds):
    # The hand-off LLR is the APP of the first iteration with the lowest
    # syndrome weight, or the channel LLR when no iteration improves on it.
    rng = np.random.default_rng(9)
    h = ccsds.dense_h().astype(np.int64)
    llr_all = oracle.awgn_llr(ccsds.encode(rng.integers(0, 2, (200, 64), dtype=np.uint8)), 0.5, 0.5, rng)
    checked = 0
    for llr in llr_all:
        full = ccsds.nms(llr, max_iter=10, alpha=0.75)
        if full["ok"]:
            assert np.array_equal(full["best_app"], full["app"])
            continue
        best, best_w = llr, int(((h @ (llr < 0)) % 2).sum())
        for t in range(1, 11):
            app = ccsds.nms(llr, max_iter=t, alpha=0.75, early_stop=False)["app"]
            w = int(((h @ (app < 0)) % 2).sum())
            if w < best_w:
                best, best_w = app, w
        assert np.array_equal(full["best_app"], best)
        assert full["best_synw"] == best_w
        checked += 1
    assert checked > 20


def test_osd_output_is_codeword_and_argmin_over_teps(ccsds):
    rng = np.random.default_rng(21)
    g = ccsds.generator()
    g_indep = nullspace_generator(ccsds.dense_h())  # a different basis of the same code
    h = ccsds.dense_h().astype(np.int64)
    for order, flips, tests in [(0, 64, 4096), (1, 64, 4096), (2, 32, 4096), (2, 64, 300)]:
        for _ in range(6):
            llr = rng.normal(1.0, 2.0, 128).astype(np.float32)
            cw, met, nt = oracle.osd(g, llr, order, flips, tests)
            assert not ((h @ cw) % 2).any()
            exp_cw, exp_m = osd_expected(g_indep, llr, order, flips, tests)
            assert np.array_equal(cw, exp_cw)
            assert met == pytest.approx(exp_m, rel=1e-12)


def test_osd_full_order_is_maximum_likelihood():
    rng = np.random.default_rng(4)
    h, rp, ci = random_code(16, 8, rng)
    c = oracle.Code(16, 8, rp, ci, front_info=True)
    g = c.generator()
    book = codebook(g)
    for _ in range(30):
        llr = rng.normal(0.3, 1.5, 16).astype(np.float32)
        cw, met, nt = oracle.osd(g, llr, order=8, max_flip=8, max_tests=1 << 8)
        rel = np.abs(llr.astype(np.float64))
        hard = llr < 0
        metrics = np.array([rel[b.astype(bool) != hard].sum() for b in book])
        assert nt == 256
        assert met == pytest.approx(metrics.min(), rel=1e-12)
        assert np.array_equal(cw, book[np.argmin(metrics)])


def test_osd_independent_of_generator_basis(ccsds):
    rng = np.random.default_rng(8)
    g = ccsds.generator()
    while True:
        a = rng.integers(0, 2, (64, 64), dtype=np.uint8)
        if gf2_rank(a) == 64:
            break
    g2 = ((a.astype(np.int64) @ g.astype(np.int64)) % 2).astype(np.uint8)
    assert gf2_inv(a) is not None
    for _ in range(10):
        llr = rng.normal(0.8, 2.0, 128).astype(np.float32)
        c1, m1, _ = oracle.osd(g, llr, 2, 64, 4096)
        c2, m2, _ = oracle.osd(g2, llr, 2, 64, 4096)
        assert np.array_equal(c1, c2) and m1 == m2


def test_osd_ties_stable_sort_and_first_best(ccsds):
    # all |L| equal: the reliability order is the index order and many TEPs tie;
    # integer metrics make the expected first-minimum exact.
    rng = np.random.default_rng(13)
    g = ccsds.generator()
    for _ in range(5):
        llr = np.where(rng.random(128) < 0.15, -1.0, 1.0).astype(np.float32)
        cw, met, _ = oracle.osd(g, llr, 2, 16, 4096)
        exp_cw, exp_m = osd_expected(g, llr, 2, 16, 4096)
        assert met == exp_m
        assert np.array_equal(cw, exp_cw)
