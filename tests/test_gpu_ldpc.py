"""GPU parity of the HIP BP / OSD decoders against the CPU oracle, through the C ABI.

Integer outputs (bits, flags, iteration counts) must be identical; float
outputs (APP LLRs, OSD distances) are bit-identical by construction and are
compared exactly (the stated tolerance, 1e-5 relative, is not needed).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import sicc_amd as sa

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ccsds(ccsds_path):
    n, m, rp, ci = oracle.alist_parse(ccsds_path)
    return oracle.Code(n, m, rp, ci, front_info=True)


def _frames(code, count, ebn0, seed):
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 2, (count, code.k), dtype=np.uint8)
    cw = code.encode(info)
    return info, cw, oracle.awgn_llr(cw, ebn0, code.k / code.n, rng)


@pytest.mark.parametrize("ebn0", [1.0, 2.5, 4.0])
def test_bp_bit_exact(ccsds, ebn0):
    print(f"[case] BP flooding CCSDS(128,64) Eb/N0={ebn0}", flush=True)
    _, _, llr = _frames(ccsds, 3000, ebn0, 100 + int(ebn0 * 10))
    with sa.Decoder.stream(sa.stream_config(max_iter=50, alpha=0.75)) as dec:
        dec.set_osd(-1)
        g = dec.decode_host(llr, want_app=True)
    for i in range(llr.shape[0]):
        o = ccsds.nms(llr[i], max_iter=50, alpha=0.75)
        assert g["bp_ok"][i] == o["ok"], i
        assert g["iterations"][i] == o["iterations"], i
        assert np.array_equal(g["code"][i], o["hard"]), i
        assert np.array_equal(g["app"][i], o["app"]), i
    assert not g["used_osd"].any()


@pytest.mark.parametrize("offset", [0.0, 0.15])
@pytest.mark.parametrize("osd_order", [-1, 1])
def test_hard_output_launch_matches_oracle(ccsds, osd_order, offset):
    # Device decode asking only for hard outputs (information bits, flags,
    # iteration counts) runs the lean single-wave BP kernel, whose zero-offset
    # build drops the offset clamp; both builds must agree with the oracle
    # exactly like the full-output launches above.
    import torch

    print(f"[case] hard-output device decode, OSD order {osd_order}, offset {offset}", flush=True)
    _, _, llr = _frames(ccsds, 4000, 2.0, 515 + osd_order)
    dev = torch.device("cuda:0")
    d_llr = torch.from_numpy(llr).to(dev)
    cw = llr.shape[0]
    d_info = torch.zeros((cw, ccsds.k), dtype=torch.uint8, device=dev)
    d_ok = torch.zeros(cw, dtype=torch.uint8, device=dev)
    d_used = torch.zeros(cw, dtype=torch.uint8, device=dev)
    d_it = torch.zeros(cw, dtype=torch.int32, device=dev)
    with sa.Decoder.stream(sa.stream_config(max_iter=20, alpha=0.75, offset=offset)) as dec:
        dec.set_osd(osd_order, 64, 4096)
        dec.decode_device(d_llr.data_ptr(), cw, d_info=d_info.data_ptr(), d_ok=d_ok.data_ptr(),
                          d_used=d_used.data_ptr(), d_iters=d_it.data_ptr(),
                          stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
    info, ok, used, its = (t.cpu().numpy() for t in (d_info, d_ok, d_used, d_it))
    if osd_order < 0:
        for i in range(0, cw, 7):
            o = ccsds.nms(llr[i], max_iter=20, alpha=0.75, offset=offset)
            assert ok[i] == o["ok"], i
            assert its[i] == o["iterations"], i
            assert np.array_equal(info[i], o["hard"][ccsds.info_cols]), i
    else:
        o = ccsds.bposd(llr, max_iter=20, alpha=0.75, offset=offset, order=1, max_flip=64, max_tests=4096,
                        threads=8)
        assert o["used_osd"].sum() > 10
        assert np.array_equal(used, o["used_osd"])
        assert np.array_equal(ok, o["bp_ok"])
        assert np.array_equal(info, o["info"])


@pytest.mark.parametrize("steal", [0, 4])
@pytest.mark.parametrize("seed", [5, 8])
def test_regular_weight3_hard_output_launch(seed, steal):
    # A regular column-weight-3, rate-1/2 QC code of 128 variables whose rows
    # hold at most 8 edges, with hard outputs only, runs the single-wave
    # hard-output build (static striding and work stealing) on a second code
    # shape; it must agree with the oracle exactly.
    import torch

    print(f"[case] regular weight-3 QC n=128 seed {seed}, steal chunk {steal}", flush=True)
    n, mb = 128, 8
    m, rp, ci = sa.qc_code(n, mb, 3, seed=seed)
    assert np.all(np.bincount(ci, minlength=n) == 3) and np.diff(rp).max() <= 8
    code = oracle.Code(n, m, rp, ci, front_info=True)
    rng = np.random.default_rng(77 + seed)
    cw = 3000
    info = rng.integers(0, 2, (cw, code.k), dtype=np.uint8)
    llr = oracle.awgn_llr(code.encode(info), 2.5, code.k / n, rng)
    dev = torch.device("cuda:0")
    d_llr = torch.from_numpy(llr).to(dev)
    d_info = torch.zeros((cw, code.k), dtype=torch.uint8, device=dev)
    d_ok = torch.zeros(cw, dtype=torch.uint8, device=dev)
    d_it = torch.zeros(cw, dtype=torch.int32, device=dev)
    with sa.Decoder.bp_csr(n, m, rp, ci, max_iter=30, alpha=0.75, schedule=0) as dec:
        dec.set_tune(1, 4, 0, steal)
        dec.decode_device(d_llr.data_ptr(), cw, d_info=d_info.data_ptr(), d_ok=d_ok.data_ptr(),
                          d_iters=d_it.data_ptr(), stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
    got_info, ok, its = (t.cpu().numpy() for t in (d_info, d_ok, d_it))
    assert 0 < ok.sum() < cw
    for i in range(cw):
        o = code.nms(llr[i], max_iter=30, alpha=0.75)
        assert ok[i] == o["ok"], i
        assert its[i] == o["iterations"], i
        assert np.array_equal(got_info[i], o["hard"][code.info_cols]), i


def test_full_headline_batch_properties(ccsds):
    # The headline workload at full size (1 048 576 codewords, Eb/N0 4.0 dB, BP 50
    # iterations + order-1 OSD): every output is a codeword (H c = 0 for BP
    # successes and for OSD results alike), the hard-output launch agrees with the
    # full-output launch on every codeword, FER <= 1e-3, and a seeded sample is
    # bit-exact against the oracle.
    import torch

    print("[case] full (128,64) batch, 1048576 codewords", flush=True)
    cw = 1 << 20
    rng = np.random.default_rng(4242)
    info = rng.integers(0, 2, (cw, ccsds.k), dtype=np.uint8)
    llr = oracle.awgn_llr(ccsds.encode(info), 4.0, 0.5, rng)
    dev = torch.device("cuda:0")
    d_llr = torch.from_numpy(llr).to(dev)
    d_code = torch.zeros((cw, ccsds.n), dtype=torch.uint8, device=dev)
    d_info = torch.zeros((cw, ccsds.k), dtype=torch.uint8, device=dev)
    d_info2 = torch.zeros((cw, ccsds.k), dtype=torch.uint8, device=dev)
    d_ok = torch.zeros(cw, dtype=torch.uint8, device=dev)
    d_used = torch.zeros(cw, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    with sa.Decoder.stream(sa.stream_config(max_iter=50, alpha=0.75)) as dec:
        dec.set_osd(1, 64, 4096)
        dec.decode_device(d_llr.data_ptr(), cw, d_info=d_info.data_ptr(), d_code=d_code.data_ptr(),
                          d_ok=d_ok.data_ptr(), d_used=d_used.data_ptr(), stream=stream)
        dec.decode_device(d_llr.data_ptr(), cw, d_info=d_info2.data_ptr(), stream=stream)
        torch.cuda.synchronize()
    h = torch.from_numpy(ccsds.dense_h().astype(np.float32)).to(dev)
    synd = torch.remainder(d_code.float() @ h.t(), 2.0)
    assert int(synd.abs().sum().item()) == 0
    assert torch.equal(d_info, d_info2)
    errors = int((d_info.cpu() != torch.from_numpy(info)).any(dim=1).sum().item())
    assert errors / cw <= 1e-3, errors
    assert int(d_used.sum().item()) > 1000
    sample = 20000
    o = ccsds.bposd(llr[:sample], max_iter=50, alpha=0.75, order=1, max_flip=64, max_tests=4096, threads=8)
    assert np.array_equal(d_code[:sample].cpu().numpy(), o["code"])
    assert np.array_equal(d_used[:sample].cpu().numpy(), o["used_osd"])


def test_full_order2_osd_batch_properties(ccsds):
    # The OSD-heavy workload at full size: 262 144 frames at Eb/N0 1.0 dB, 20 BP
    # iterations, every failure to order-2 OSD (2 081 test patterns): all outputs
    # are codewords, about three quarters of the frames go through OSD, and a seeded
    # sample is bit-exact against the oracle.
    import torch

    print("[case] full order-2 OSD batch, 262144 frames", flush=True)
    cw = 1 << 18
    rng = np.random.default_rng(777)
    info = rng.integers(0, 2, (cw, ccsds.k), dtype=np.uint8)
    llr = oracle.awgn_llr(ccsds.encode(info), 1.0, 0.5, rng)
    dev = torch.device("cuda:0")
    d_llr = torch.from_numpy(llr).to(dev)
    d_code = torch.zeros((cw, ccsds.n), dtype=torch.uint8, device=dev)
    d_used = torch.zeros(cw, dtype=torch.uint8, device=dev)
    with sa.Decoder.stream(sa.stream_config(max_iter=20, alpha=0.75)) as dec:
        dec.set_osd(2, 64, 4096)
        dec.decode_device(d_llr.data_ptr(), cw, d_code=d_code.data_ptr(), d_used=d_used.data_ptr(),
                          stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
    h = torch.from_numpy(ccsds.dense_h().astype(np.float32)).to(dev)
    assert int(torch.remainder(d_code.float() @ h.t(), 2.0).abs().sum().item()) == 0
    used = d_used.cpu().numpy()
    assert 0.7 < used.mean() < 0.85, used.mean()  # BP fails on ~3/4 of the frames at 1.0 dB
    sample = 3000
    o = ccsds.bposd(llr[:sample], max_iter=20, alpha=0.75, order=2, max_flip=64, max_tests=4096, threads=8)
    assert np.array_equal(d_code[:sample].cpu().numpy(), o["code"])
    assert np.array_equal(used[:sample], o["used_osd"])


@pytest.mark.parametrize("order,flips", [(1, 64), (2, 32), (2, 64)])
def test_bposd_bit_exact(ccsds, order, flips):
    print(f"[case] BP-OSD order={order} flips={flips}", flush=True)
    _, cw, llr = _frames(ccsds, 2000, 1.5, 7 + order * 3 + flips)
    with sa.Decoder.stream(sa.stream_config(max_iter=20, alpha=0.75)) as dec:
        dec.set_osd(order, flips, 4096)
        g = dec.decode_host(llr)
    o = ccsds.bposd(llr, max_iter=20, alpha=0.75, order=order, max_flip=flips, max_tests=4096, threads=8)
    assert o["used_osd"].sum() > 100
    assert np.array_equal(g["used_osd"], o["used_osd"])
    assert np.array_equal(g["bp_ok"], o["bp_ok"])
    assert np.array_equal(g["code"], o["code"])
    assert np.array_equal(g["info"], o["info"])


@pytest.mark.parametrize("order,flips", [(1, 128), (2, 48)])
def test_bposd_k128_bit_exact(order, flips):
    # CCSDS (256,128): OSD over a 128-row generator (two column words per lane)
    import os

    print(f"[case] BP-OSD (256,128) order={order} flips={flips}", flush=True)
    path = os.path.join(sa.DATA_DIR, "CCSDS_ldpc_n256_k128.alist")
    n, m, rp, ci = oracle.alist_parse(path)
    code = oracle.Code(n, m, rp, ci, front_info=True)
    _, _, llr = _frames(code, 600, 1.5, 91 + order)
    cfg = sa.stream_config(n=256, k=128, max_iter=20, alpha=0.75)
    with sa.Decoder.stream(cfg, alist=path) as dec:
        dec.set_osd(order, flips, 4096)
        g = dec.decode_host(llr)
    o = code.bposd(llr, max_iter=20, alpha=0.75, order=order, max_flip=flips, max_tests=4096, threads=8)
    assert o["used_osd"].sum() > 50
    assert np.array_equal(g["used_osd"], o["used_osd"])
    assert np.array_equal(g["bp_ok"], o["bp_ok"])
    assert np.array_equal(g["code"], o["code"])
    assert np.array_equal(g["info"], o["info"])


def test_bposd_damping_offset_bit_exact(ccsds):
    print("[case] BP-OSD damping=0.15 offset=0.15", flush=True)
    _, _, llr = _frames(ccsds, 1500, 2.0, 55)
    cfg = sa.stream_config(max_iter=8, alpha=0.8, offset=0.15, damping=0.15)
    with sa.Decoder.stream(cfg) as dec:
        dec.set_osd(1, 64, 4096)
        g = dec.decode_host(llr)
    o = ccsds.bposd(llr, max_iter=8, alpha=0.8, offset=0.15, damping=0.15, order=1, max_flip=64, threads=8)
    assert np.array_equal(g["bp_ok"], o["bp_ok"])
    assert np.array_equal(g["code"], o["code"])


def test_osd_only_bit_exact(ccsds):
    print("[case] OSD-only order 2 / 2081 TEPs", flush=True)
    rng = np.random.default_rng(31)
    _, _, llr = _frames(ccsds, 400, 1.0, 31)
    gen = ccsds.generator()
    with sa.Decoder.osd_only() as dec:
        info, dist = dec.osd_batch(llr)
    for i in range(llr.shape[0]):
        cw, met, tests = oracle.osd(gen, llr[i], 2, 64, 4096)
        assert tests == 2081
        assert np.array_equal(info[i], cw[:64]), i
        assert dist[i] == np.float32(met), i
    del rng


def test_generator_and_encoder_match_oracle(ccsds):
    with sa.Decoder.stream() as dec:
        assert (dec.n, dec.k, dec.m, dec.nnz) == (128, 64, 64, 512)
        rp, ci = dec.csr()
        assert np.array_equal(rp, ccsds.rp) and np.array_equal(ci, ccsds.ci)
        assert np.array_equal(dec.generator(), ccsds.generator())
        info = np.random.default_rng(1).integers(0, 2, (100, 64), dtype=np.uint8)
        assert np.array_equal(dec.encode(info), ccsds.encode(info))


def test_builtin_stream_decoder_and_poll(ccsds):
    print("[case] stream submit/flush/poll", flush=True)
    _, _, llr = _frames(ccsds, 700, 2.0, 77)
    cfg = sa.stream_config(max_batch=256, min_batch=64, max_iter=20)
    with sa.Decoder.stream(cfg, builtin=True) as dec:
        dec.set_osd(1, 64, 4096)
        ref = dec.decode_host(llr)
        ids = np.arange(700) // 10
        blk = np.arange(700) % 10
        for s in range(0, 700, 90):
            dec.submit(llr[s:s + 90], frame_ids=ids[s:s + 90], block_indices=blk[s:s + 90])
        dec.flush()
        got = {}
        while len(got) < 700:
            info, fid, b, used = dec.poll(128)
            assert len(info) > 0
            for j in range(len(info)):
                got[int(fid[j]) * 10 + int(b[j])] = (info[j].copy(), int(used[j]))
        st = dec.stats()
        assert st.submitted_codewords == 700 and st.decoded_codewords == 700
        assert st.osd_fallback_codewords == int(ref["used_osd"].sum())
    for i in range(700):
        assert np.array_equal(got[i][0], ref["info"][i])
        assert got[i][1] == ref["used_osd"][i]


def test_bp_csr_decoder_other_codes():
    print("[case] BP CSR: CCSDS(256,128) and QC(1024, rate 1/2)", flush=True)
    import os

    path = os.path.join(sa.DATA_DIR, "CCSDS_ldpc_n256_k128.alist")
    n, m, rp, ci = oracle.alist_parse(path)
    m_qc, rp_qc, ci_qc = sa.qc_code(1024, 8, 3, seed=5)
    for (n, m, rp, ci), ebn0 in [((n, m, rp, ci), 2.0), ((1024, m_qc, rp_qc, ci_qc), 1.8)]:
        code = oracle.Code(n, m, rp, ci, front_info=True)
        _, _, llr = _frames(code, 300, ebn0, n)
        with sa.Decoder.bp_csr(n, m, rp, ci, max_iter=30, alpha=0.75, schedule=0) as dec:
            bits, ok, _ = dec.bp_csr_batch(llr)
        for i in range(llr.shape[0]):
            o = code.nms(llr[i], max_iter=30, alpha=0.75)
            assert ok[i] == o["ok"]
            assert np.array_equal(bits[i], o["hard"])


@pytest.mark.parametrize("which", ["ccsds128", "ccsds256", "qc1024", "qc2048_r34", "qc512_r23"])
def test_every_launch_shape_bit_exact(which):
    # every (team, codewords-per-group) shape the autotuner may pick decodes
    # bit-identically to the oracle, including the multi-wave team kernels
    import os

    import torch

    if which.startswith("ccsds"):
        n = int(which[5:])
        path = os.path.join(sa.DATA_DIR, f"CCSDS_ldpc_n{n}_k{n // 2}.alist")
        n, m, rp, ci = oracle.alist_parse(path)
    else:
        n = int(which[2:6].rstrip("_r"))
        mb = {"": 8, "r34": 4, "r23": 5}[which.split("_")[1] if "_" in which else ""]
        m, rp, ci = sa.qc_code(n, mb, 3, seed=n + mb)
    print(f"[case] launch shapes {which} n={n} m={m}", flush=True)
    code = oracle.Code(n, m, rp, ci, front_info=True)
    _, _, llr = _frames(code, 96, 1.5 if n <= 256 else 2.0, 3 * n + m)
    ref = [code.nms(llr[i], max_iter=25, alpha=0.75) for i in range(llr.shape[0])]
    d_llr = torch.from_numpy(llr).cuda()
    with sa.Decoder.bp_csr(n, m, rp, ci, max_iter=25, alpha=0.75, schedule=0) as dec:
        for team in (1, 2, 4, 8, 16):
            for tpb in ((1, 2, 4) if team == 1 else (1,)):
                try:
                    dec.set_tune(team, tpb, 0)
                except sa.SiccError:
                    continue
                bits, ok, _ = dec.bp_csr_batch(llr)
                for i in range(llr.shape[0]):
                    assert ok[i] == ref[i]["ok"], (team, tpb, i)
                    assert np.array_equal(bits[i], ref[i]["hard"]), (team, tpb, i)
        best = dec.autotune(d_llr.data_ptr(), llr.shape[0])
        assert best["team"] in (1, 2, 4, 8, 16) and best["bp_ms"] > 0
        bits, ok, _ = dec.bp_csr_batch(llr)
        assert all(np.array_equal(bits[i], ref[i]["hard"]) for i in range(llr.shape[0]))


def test_error_paths():
    lib = sa.load()
    import ctypes as C

    code, _ = None, None
    assert lib.cuda_osd_flush(C.c_ulonglong(123456)) == sa.STATUS_INVALID_ARGUMENT
    assert b"unknown handle" in lib.cuda_osd_get_last_error()
    with sa.Decoder.stream() as dec:
        with pytest.raises(sa.SiccError) as e:
            dec.decode_host(np.zeros((0, 128), np.float32))
        assert e.value.status == sa.STATUS_INVALID_ARGUMENT
        with pytest.raises(sa.SiccError):
            dec.set_osd(4)
    bad = sa.stream_config(n=100, k=50)
    with pytest.raises(sa.SiccError):
        sa.Decoder.stream(bad, builtin=True)
    del code
