"""The cuda_osd_* calls the reference's own wrappers make, on the GPU, through the
raw C ABI (ctypes), bit-exact against the CPU oracle or the synchronous decode.

Reference call sites (source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp):
  cuda_osd_create_stream_decoder_with_matrices   :3251-3256 with the dense H and
      systematic parity of build_dense_h_for_cuda_osd / build_systematic_parity_for_cuda_osd (:3120-3160)
  cuda_osd_submit_device_codewords / flush / poll :3356-3432 (async gpu-sim loop)
  cuda_osd_decode_bp_csr_device_batch            :3753
  cuda_osd_get_decoded_codewords / get_total_decode_ms (matlab_decoder_c_api.h:120-121)
plus the stream decoder's latency-budget launch (cuda_osd_stream_config_t.max_latency_us,
matlab_decoder_c_api.h:17-32): a partly filled batch is launched by a poll once its oldest
codeword has waited max_latency_us, with no flush.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import pytest

import oracle
import sicc_amd as sa

pytestmark = pytest.mark.gpu

u64p = C.POINTER(C.c_ulonglong)
u32p = C.POINTER(C.c_uint)
u8p = C.POINTER(C.c_ubyte)


def _code():
    n, m, rp, ci = oracle.alist_parse(sa.CCSDS_128_64)
    return oracle.Code(n, m, rp, ci, True), (n, m, rp, ci)


def _poll_all(lib, h, k, want, flushed, deadline_s=5.0):
    """Poll until `want` codewords came back; returns {(frame_id, block): (info, used)}."""
    got = {}
    cap = 4096
    info = np.zeros((cap, k), np.uint8)
    ts = np.zeros(cap, np.uint64)
    ids = np.zeros(cap, np.uint64)
    blk = np.zeros(cap, np.uint32)
    used = np.zeros(cap, np.uint8)
    n = C.c_int(0)
    t_end = time.time() + deadline_s
    while len(got) < want and time.time() < t_end:
        sa.check(lib.cuda_osd_poll_ready(C.c_ulonglong(h), cap, info.ctypes.data_as(u8p), ts.ctypes.data_as(u64p),
                                         ids.ctypes.data_as(u64p), blk.ctypes.data_as(u32p),
                                         used.ctypes.data_as(u8p), C.byref(n)))
        if flushed:
            assert n.value > 0 or len(got) == want  # LINK:3514: zero after flush is an error
        for i in range(n.value):
            got[(int(ids[i]), int(blk[i]))] = (info[i].copy(), int(used[i]))
        if n.value == 0:
            time.sleep(0.0005)
    return got


def test_device_submit_flush_poll_matches_oracle():
    """cuda_osd_submit_device_codewords of device LLRs (the gpu-sim loop's call,
    LINK:3356-3386) in uneven pieces, then flush and poll until drained; every
    (frame_id, block) result bit-exact against the oracle BP + OSD with the
    stream decoder's defaults (order 2, 32 flip positions, 4 096 tests)."""
    import torch

    print("[case] abi device submit/flush/poll", flush=True)
    lib = sa.load()
    code, _ = _code()
    rng = np.random.default_rng(61)
    N = 3000
    info = rng.integers(0, 2, (N, code.k), dtype=np.uint8)
    llr = oracle.awgn_llr(code.encode(info), 2.5, 0.5, rng)
    d_llr = torch.from_numpy(llr).to("cuda:0")
    torch.cuda.synchronize()
    cfg = sa.stream_config(max_batch=1024, min_batch=1, max_iter=50, alpha=0.75, latency_us=10 ** 9)
    h = C.c_ulonglong(0)
    sa.check(lib.cuda_osd_create_stream_decoder(C.byref(cfg), C.byref(h)))
    try:
        ids = (np.arange(N) // 30).astype(np.uint64)
        blk = (np.arange(N) % 30).astype(np.uint32)
        ts = np.arange(N, dtype=np.uint64)
        s0 = 0
        for piece in (1, 700, 1023, 276, 1000):
            sa.check(lib.cuda_osd_submit_device_codewords(
                h, C.c_void_p(d_llr.data_ptr() + s0 * code.n * 4), piece, ts[s0:].ctypes.data_as(u64p),
                ids[s0:].ctypes.data_as(u64p), blk[s0:].ctypes.data_as(u32p)))
            s0 += piece
        assert s0 == N
        sa.check(lib.cuda_osd_flush(h))
        got = _poll_all(lib, h.value, code.k, N, flushed=True)
        assert len(got) == N
        ref = code.bposd(llr, max_iter=50, alpha=0.75, order=2, max_flip=32, max_tests=4096)
        for i in range(N):
            gi, gu = got[(int(ids[i]), int(blk[i]))]
            assert np.array_equal(gi, ref["info"][i]), i
            assert gu == ref["used_osd"][i], i
        st = sa.StreamStats()
        sa.check(lib.cuda_osd_get_stats(h, C.byref(st)))
        assert st.submitted_codewords == N and st.decoded_codewords == N
        assert st.osd_fallback_codewords == int(ref["used_osd"].sum())
        assert st.bp_success_codewords == int(ref["bp_ok"].sum())
        assert st.decoded_batches == 3  # 1024 + 1024 + 952 (the flushed remainder)
        assert st.current_pending_codewords == 0
        assert lib.cuda_osd_get_decoded_codewords(h.value) == N
        assert lib.cuda_osd_get_total_decode_ms(h.value) > 0.0
        assert lib.cuda_osd_get_total_decode_ms(h.value) == pytest.approx(st.total_decode_ms)
    finally:
        lib.cuda_osd_destroy_stream_decoder(h)


def test_latency_budget_launches_without_flush():
    """A batch below max_batch_size is launched once its oldest codeword has
    waited max_latency_us (here 3 ms), by a poll, without flush; results equal
    the synchronous decode of the same LLRs.  With an unbounded budget the same
    partial batch stays queued."""
    print("[case] abi latency-budget launch", flush=True)
    lib = sa.load()
    code, _ = _code()
    rng = np.random.default_rng(62)
    N = 100
    info = rng.integers(0, 2, (N, code.k), dtype=np.uint8)
    llr = oracle.awgn_llr(code.encode(info), 3.0, 0.5, rng)
    ids = np.arange(N, dtype=np.uint64)
    blk = np.zeros(N, np.uint32)
    for budget_us, expect in ((3000, N), (10 ** 9, 0)):
        with sa.Decoder.stream(sa.stream_config(max_batch=4096, min_batch=1, latency_us=budget_us)) as dec:
            sync = dec.decode_host(llr)["info"]
            dec.submit(llr, frame_ids=ids, block_indices=blk)
            got = _poll_all(lib, dec.handle, code.k, N, flushed=False, deadline_s=0.25)
            assert len(got) == expect, (budget_us, len(got))
            if expect:
                for i in range(N):
                    assert np.array_equal(got[(i, 0)][0], sync[i]), i
                st = dec.stats()
                assert st.decoded_batches == 1 and st.decoded_codewords == N
            else:
                dec.flush()
                assert len(_poll_all(lib, dec.handle, code.k, N, flushed=True)) == N


def test_stream_decoder_from_reference_matrices():
    """cuda_osd_create_stream_decoder_with_matrices with the dense H and the
    systematic parity rows as the reference builds them (LINK:3120-3160: row i
    = parity bits of encode(e_i), bit p at word p / 64, bit p % 64): decodes
    bit-exact against the oracle; a parity table that disagrees with H is
    rejected."""
    print("[case] abi create_with_matrices", flush=True)
    lib = sa.load()
    code, (n, m, rp, ci) = _code()
    k = code.k
    H = code.dense_h()
    W = (n - k + 63) // 64
    gen = code.generator()
    par = np.zeros((k, W), np.uint64)
    for i in range(k):
        for p in range(n - k):
            if gen[i, k + p]:
                par[i, p // 64] |= np.uint64(1) << np.uint64(p % 64)
    cfg = sa.stream_config(max_batch=4096, max_iter=50, alpha=0.75)
    lib.cuda_osd_create_stream_decoder_with_matrices.argtypes = [C.POINTER(sa.StreamConfig), u8p, u64p, C.c_int,
                                                                 u64p]
    h = C.c_ulonglong(0)
    Hc = np.ascontiguousarray(H, np.uint8)
    sa.check(lib.cuda_osd_create_stream_decoder_with_matrices(C.byref(cfg), Hc.ctypes.data_as(u8p),
                                                              par.ctypes.data_as(u64p), W, C.byref(h)))
    dec = sa.Decoder(h.value, "stream")
    with dec:
        dec.set_osd(1, 64, 4096)
        rng = np.random.default_rng(63)
        info = rng.integers(0, 2, (2000, k), dtype=np.uint8)
        llr = oracle.awgn_llr(code.encode(info), 2.5, 0.5, rng)
        out = dec.decode_host(llr)
        ref = code.bposd(llr, max_iter=50, alpha=0.75, order=1, max_flip=64)
        assert np.array_equal(out["info"], ref["info"])
        assert np.array_equal(out["used_osd"], ref["used_osd"])
        assert np.array_equal(dec.generator(), gen)
    bad = par.copy()
    bad[3, 0] ^= np.uint64(1)
    h2 = C.c_ulonglong(0)
    st = lib.cuda_osd_create_stream_decoder_with_matrices(C.byref(cfg), Hc.ctypes.data_as(u8p),
                                                          bad.ctypes.data_as(u64p), W, C.byref(h2))
    assert st == sa.STATUS_INVALID_ARGUMENT and h2.value == 0
    assert "does not match" in lib.cuda_osd_get_last_error().decode()


def test_bp_csr_device_batch_matches_oracle():
    """cuda_osd_decode_bp_csr_device_batch (device LLRs, host outputs, LINK:3753)
    and its host twin cuda_osd_decode_bp_csr_batch: code bits and success flags
    bit-exact against the oracle's flooding min-sum, decode time reported."""
    import torch

    print("[case] abi bp_csr device batch", flush=True)
    lib = sa.load()
    code, (n, m, rp, ci) = _code()
    rng = np.random.default_rng(64)
    N = 5000
    info = rng.integers(0, 2, (N, code.k), dtype=np.uint8)
    llr = oracle.awgn_llr(code.encode(info), 2.0, 0.5, rng)
    d_llr = torch.from_numpy(llr).to("cuda:0")
    torch.cuda.synchronize()
    with sa.Decoder.bp_csr(n, m, rp, ci, max_iter=40, alpha=0.8, schedule=0) as dec:
        bits = np.zeros((N, n), np.uint8)
        ok = np.zeros(N, np.uint8)
        ms = C.c_double(0)
        sa.check(lib.cuda_osd_decode_bp_csr_device_batch(dec.handle, C.c_void_p(d_llr.data_ptr()), N,
                                                         bits.ctypes.data_as(u8p), ok.ctypes.data_as(u8p),
                                                         C.byref(ms)))
        assert ms.value > 0.0
        hb, hok, _ = dec.bp_csr_batch(llr)
    for i in range(N):
        r = code.nms(llr[i], max_iter=40, alpha=0.8)
        assert np.array_equal(bits[i], r["hard"]), i
        assert ok[i] == r["ok"], i
    assert np.array_equal(hb, bits) and np.array_equal(hok, ok)
    assert 0 < ok.sum() < N
