"""GPU parity of the frame-level decode (sicc_frame_* C ABI) -- the reference's
decode_frame / finish_decode_from_llr (uhd_ldpc_ofdm_link.cpp:5331-5603,
6615-6705) returning FrameDecodeResult (:273-286) -- against the reference's own
results (tests/golden/frame_small.npz, written by its CPU decode_frame) and the
CPU oracle chain.

Bars: detection, parity, frame_ok, frame id, bit and pre-FEC error counts,
iterations and decoded information bits identical; SNR and EVM (double sums of
float terms whose FP64 FFT inputs come from a different FFT algorithm) within
1e-6 relative; CFO as the sync tests (1e-6 relative).
"""
from __future__ import annotations

import os
from types import SimpleNamespace

import numpy as np
import pytest

import oracle
import sicc_amd as sa
import waveform as wf

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KEYS = ("detected", "parity_ok", "frame_ok", "frame_id", "bit_errors", "pre_fec_bit_errors", "iterations")


def _fixture_phy(z):
    nfft, cp, S, A, half, B, flen, blocks = (int(v) for v in z["dims"])
    return SimpleNamespace(nfft=nfft, cp=cp, num_symbols=S, active_sc=A, pre_half_len=half, bits_per_symbol=B,
                           frame_len=flen, blocks=blocks, fs=float(z["fs"][0]), preamble=z["preamble"],
                           pre_len=len(z["preamble"]), used_shifted=z["used_shifted"],
                           pilot_symbols=z["pilot_symbols"], data_symbols=z["data_symbols"],
                           pilot_grid=z["pilot_grid"], constellation=z["constellation"])


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(G, "frame_small.npz"))


def _ccsds_bp(iters, alpha):
    n, m, rp, ci = oracle.alist_parse(sa.CCSDS_128_64)
    return sa.Decoder.bp_csr(n, m, rp, ci, max_iter=iters, alpha=alpha, schedule=0)


def _check_results(res, info, ri, rd, ref_info_bits, check_snr=True):
    F = len(res)
    for f in range(F):
        for j, key in enumerate(KEYS):
            assert int(res[key][f]) == int(ri[f][j]), (f, key, res[key][f], ri[f][j])
        if check_snr:
            assert res["snr_db"][f] == pytest.approx(rd[f][2], rel=1e-6), f
            assert res["evm_rms"][f] == pytest.approx(rd[f][3], rel=1e-6), f
        if res["snr_gated"][f]:
            assert not info[f].any()
        else:
            assert np.array_equal(info[f], ref_info_bits[f]), f


@pytest.mark.parametrize("gated", [False, True])
def test_demod_decode_matches_reference_decode_frame(golden, gated):
    """The caller's sync (the reference's find_frame start, peak and CFO) as
    decode_frame takes them; results equal the reference decode_frame."""
    import torch

    z = golden
    print(f"[case] frame golden demod+decode gated={gated}", flush=True)
    phy = _fixture_phy(z)
    iters, alpha = z["bp"]
    nv = float(z["noise_var"][0])
    F, span = z["bufs"].shape
    dev = torch.device("cuda:0")
    sync = np.zeros(F, sa.SYNC_DTYPE)
    sync["start"] = z["sync"][:, 0].astype(np.int32)
    sync["metric"] = z["sync"][:, 1]
    sync["cfo_hz"] = z["sync"][:, 2]
    sync["cfo_hz_d"] = z["sync"][:, 2]
    gate = float(z["gate_db"][0]) if gated else -1e30
    d_rx = torch.from_numpy(np.ascontiguousarray(z["bufs"]).view(np.float32)).to(dev)
    d_sync = torch.from_numpy(sync.view(np.uint8)).to(dev)
    d_ref = torch.from_numpy(np.ascontiguousarray(z["ref_info"])).to(dev)
    d_info = torch.empty((F, phy.blocks * 64), dtype=torch.uint8, device=dev)
    d_res = torch.empty(F * sa.FRAME_RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    with sa.OfdmReceiver(phy, 128, phy.blocks, noise_var=nv) as rxr, _ccsds_bp(int(iters), float(alpha)) as dec, \
            sa.FrameDecoder(rxr, dec, snr_gate_db=gate) as fd:
        fd.demod_decode_device(d_rx.data_ptr(), F, span, d_sync.data_ptr(), d_info.data_ptr(), d_res.data_ptr(),
                               d_ref=d_ref.data_ptr())
        torch.cuda.synchronize()
    res = d_res.cpu().numpy().view(sa.FRAME_RESULT_DTYPE)
    info = d_info.cpu().numpy()
    ri, rd = (z["gate_int"], z["gate_dbl"]) if gated else (z["res_int"], z["res_dbl"])
    _check_results(res, info, ri, rd, z["info"])
    assert np.array_equal(res["peak"], z["sync"][:, 1].astype(np.float32).astype(np.float64))
    assert np.array_equal(res["cfo_hz"], z["sync"][:, 2])
    assert (res["blocks"] == phy.blocks).all()
    assert np.array_equal(res["snr_gated"].astype(bool), gated & (rd[:, 2] < gate))


def test_demod_decode_reproduces_ofdm_small_decode_frame():
    """The single frame of tests/golden/ofdm_small.npz at its own size: the
    reference decode_frame's information bits (df_info) and success (df_ok),
    from the frame at start 0 with its CFO and noise variance (df_args)."""
    import torch

    print("[case] frame golden ofdm_small demod+decode", flush=True)
    z = np.load(os.path.join(G, "ofdm_small.npz"))
    phy = _fixture_phy(z)
    cfo, nv, iters, alpha = (float(v) for v in z["df_args"])
    frame = np.ascontiguousarray(z["df_frame"].astype(np.complex64))
    dev = torch.device("cuda:0")
    sync = np.zeros(1, sa.SYNC_DTYPE)
    sync["metric"] = 1.0
    sync["cfo_hz"] = cfo
    sync["cfo_hz_d"] = cfo
    d_rx = torch.from_numpy(frame.view(np.float32)).to(dev)
    d_sync = torch.from_numpy(sync.view(np.uint8)).to(dev)
    d_info = torch.empty((1, phy.blocks * 64), dtype=torch.uint8, device=dev)
    d_res = torch.empty(sa.FRAME_RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    with sa.OfdmReceiver(phy, 128, phy.blocks, noise_var=nv) as rxr, _ccsds_bp(int(iters), alpha) as dec, \
            sa.FrameDecoder(rxr, dec) as fd:
        fd.demod_decode_device(d_rx.data_ptr(), 1, frame.shape[0], d_sync.data_ptr(), d_info.data_ptr(),
                               d_res.data_ptr())
        torch.cuda.synchronize()
    res = d_res.cpu().numpy().view(sa.FRAME_RESULT_DTYPE)
    assert np.array_equal(d_info.cpu().numpy()[0], z["df_info"])
    assert bool(res["parity_ok"][0]) == bool(z["df_ok"][0])
    assert res["detected"][0] == 1 and res["blocks"][0] == phy.blocks


def test_receive_decode_host_matches_reference(golden):
    """The whole chain from the buffers (this library's sync, then demod, BP and
    aggregation) through the host entry point: sync starts equal the
    reference's find_frame, every result field equals decode_frame's."""
    z = golden
    print("[case] frame golden receive+decode host", flush=True)
    phy = _fixture_phy(z)
    iters, alpha = z["bp"]
    with sa.OfdmReceiver(phy, 128, phy.blocks, noise_var=float(z["noise_var"][0]), threshold=0.45) as rxr, \
            _ccsds_bp(int(iters), float(alpha)) as dec, sa.FrameDecoder(rxr, dec) as fd:
        info, res = fd.receive_decode_host(z["bufs"], ref_info=z["ref_info"])
        info0, res0 = fd.receive_decode_host(z["bufs"])  # no reference payload
    assert np.array_equal(res["cfo_hz"].astype(np.float64), res["cfo_hz"])
    for f in range(len(res)):
        assert res["cfo_hz"][f] == pytest.approx(z["sync"][f, 2], rel=1e-6), f
        assert res["peak"][f] == pytest.approx(z["sync"][f, 1], rel=1e-5), f
    _check_results(res, info, z["res_int"], z["res_dbl"], z["info"])
    # without a reference test: frame_id 0, counts -1, frame_ok 0 (LINK:5584-5588), same bits and parity
    assert np.array_equal(info0, info)
    assert (res0["frame_id"] == 0).all() and (res0["bit_errors"] == -1).all()
    assert (res0["pre_fec_bit_errors"] == -1).all() and (res0["frame_ok"] == 0).all()
    assert np.array_equal(res0["parity_ok"], res["parity_ok"])
    assert np.array_equal(res0["iterations"], res["iterations"])


@pytest.mark.parametrize("kind", ["bp", "bposd"])
def test_llr_path_matches_oracle(kind):
    """finish_decode_from_llr: LLR frames of 5 (128,64) blocks; every field
    against the oracle's BP(-OSD) and aggregation, with reference payloads
    that match, and ones with flipped bits (bit errors on decoded frames)."""
    print(f"[case] frame LLR path {kind}", flush=True)
    n, m, rp, ci = oracle.alist_parse(sa.CCSDS_128_64)
    code = oracle.Code(n, m, rp, ci, True)
    rng = np.random.default_rng(404 + (kind == "bposd"))
    F, B, k = 48, 5, code.k
    info = rng.integers(0, 2, (F * B, k), dtype=np.uint8)
    llr = oracle.awgn_llr(code.encode(info), 2.5, 0.5, rng)
    ref = info.reshape(F, B * k).copy()
    ref[::7, 3] ^= 1  # reference payloads that differ from what was sent
    if kind == "bp":
        dec = sa.Decoder.bp_csr(n, m, rp, ci, max_iter=50, alpha=0.75, schedule=0)
        o = code.bposd(llr, max_iter=50, alpha=0.75, order=-1)
    else:
        dec = sa.Decoder.stream(sa.stream_config(max_iter=50, alpha=0.75))
        dec.set_osd(1, 64, 4096)
        o = code.bposd(llr, max_iter=50, alpha=0.75, order=1, max_flip=64)
    with dec, sa.FrameDecoder(None, dec, blocks_per_frame=B) as fd:
        got, res = fd.decode_llr_host(llr.reshape(F, B * n), ref_info=ref)
        got0, res0 = fd.decode_llr_host(llr.reshape(F, B * n))
    assert np.array_equal(got, o["info"].reshape(F, B * k))
    assert np.array_equal(got0, got)
    par = (o["bp_ok"] | o["used_osd"]).reshape(F, B) if kind == "bposd" else o["bp_ok"].reshape(F, B)
    its = o["iterations"].reshape(F, B)
    for f in range(F):
        rc = code.encode(ref[f].reshape(B, k)).reshape(-1)
        e = oracle.frame_finish(B, n, k, par[f], its[f], got[f], llr.reshape(F, B * n)[f], ref[f], rc)
        for key in ("parity_ok", "frame_ok", "frame_id", "bit_errors", "pre_fec_bit_errors", "iterations"):
            assert int(res[key][f]) == e[key], (f, key)
        assert res["detected"][f] == 1 and res["blocks"][f] == B
        assert res["bp_ok_blocks"][f] == int(o["bp_ok"].reshape(F, B)[f].sum())
        assert res["osd_blocks"][f] == (int(o["used_osd"].reshape(F, B)[f].sum()) if kind == "bposd" else 0)
        assert np.isnan(res["evm_rms"][f]) and res["snr_db"][f] == 0.0
        e0 = oracle.frame_finish(B, n, k, par[f], its[f], got[f], llr.reshape(F, B * n)[f])
        assert int(res0["parity_ok"][f]) == e0["parity_ok"] and int(res0["frame_ok"][f]) == 0
    assert 0 < res["frame_ok"].sum() < F
    assert (res["bit_errors"][::7] > 0).all()


def test_cfg2_receive_decode_vs_oracle_chain():
    """The 1024-point, 600-subcarrier, 16-QAM receive workload with the rate-1/2
    n = 1024 QC code (30 iterations) at full size, 4 096 frames through the
    6-tap fading channel, in one frame-decode call: results consistent with the
    separately launched receiver and decoder for every frame (bits, sync),
    error counts equal to numpy's on every frame, and a spread sample of frames
    equal to the oracle chain (find_frame, demod with SNR / EVM, min-sum per
    block, aggregation)."""
    import torch

    print("[case] frame cfg2 4096 frames", flush=True)
    dev = torch.device("cuda:0")
    phy = wf.OfdmPhy(bits_per_symbol=4, active_sc=600, num_symbols=14, pilot_symbols=[0])
    rng = np.random.default_rng(4242)
    F, n = 4096, 1024
    blocks = phy.coded_bits // n
    m, rp, ci = sa.qc_code(n, 8, 3, seed=64)
    dec = sa.Decoder.bp_csr(n, m, rp, ci, max_iter=30, alpha=0.75, schedule=0)
    k = dec.k
    info = rng.integers(0, 2, (F * blocks, k), dtype=np.uint8)
    cw = dec.encode(info)
    coded = np.concatenate([cw.reshape(F, blocks * n),
                            rng.integers(0, 2, (F, phy.coded_bits - blocks * n), dtype=np.uint8)], axis=1)
    span = phy.frame_len + 256
    rx = np.empty((F, span), np.complex64)
    nv = 0.0
    for s0 in range(0, F, 512):
        ch = wf.channel(phy, wf.tx_frames(phy, coded[s0:s0 + 512]), rng, span=span, snr_db=14.0)
        rx[s0:s0 + 512] = ch.rx
        nv = ch.noise_var_freq
    ref = info.reshape(F, blocks * k)
    d_rx = torch.from_numpy(rx.view(np.float32)).to(dev)
    d_ref = torch.from_numpy(ref).to(dev)
    d_info = torch.empty((F, blocks * k), dtype=torch.uint8, device=dev)
    d_res = torch.empty(F * sa.FRAME_RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    d_sync = torch.empty(F * 24, dtype=torch.uint8, device=dev)
    d_llr = torch.empty((F, blocks * n), dtype=torch.float32, device=dev)
    d_blk = torch.empty(F * blocks, dtype=torch.uint8, device=dev)
    d_it = torch.empty(F * blocks, dtype=torch.int32, device=dev)
    with dec, sa.OfdmReceiver(phy, n, blocks, max_search=256, noise_var=nv, threshold=0.2) as rxr, \
            sa.FrameDecoder(rxr, dec) as fd:
        fd.receive_decode_device(d_rx.data_ptr(), F, span, d_info.data_ptr(), d_res.data_ptr(),
                                 d_ref=d_ref.data_ptr(), d_sync=d_sync.data_ptr(), d_llr=d_llr.data_ptr(),
                                 d_block_ok=d_blk.data_ptr(), d_block_it=d_it.data_ptr())
        torch.cuda.synchronize()
        res = d_res.cpu().numpy().view(sa.FRAME_RESULT_DTYPE)
        got = d_info.cpu().numpy()
        sync = d_sync.cpu().numpy().view(sa.SYNC_DTYPE)
        llr = d_llr.cpu().numpy()
        blk = d_blk.cpu().numpy().reshape(F, blocks)
        its = d_it.cpu().numpy().reshape(F, blocks)
        # the same batch through the separate receiver and decoder launches
        d_llr2 = torch.empty_like(d_llr)
        d_info2 = torch.empty_like(d_info)
        d_sync2 = torch.empty_like(d_sync)
        rxr.receive_device(d_rx.data_ptr(), F, span, d_sync2.data_ptr(), d_llr2.data_ptr())
        torch.cuda.synchronize()
        dec.decode_device(d_llr2.data_ptr(), F * blocks, d_info=d_info2.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(d_llr, d_llr2)
        assert torch.equal(d_sync, d_sync2)
    det = res["detected"].astype(bool)
    assert np.array_equal(det, sync["start"] >= 0)
    assert det.mean() >= 0.99
    sep = d_info2.cpu().numpy()
    assert np.array_equal(got[det], sep[det])
    assert not got[~det].any()
    # every frame: counts and flags against numpy
    be = (got != ref).sum(axis=1)
    assert np.array_equal(res["bit_errors"][det], be[det])
    assert np.array_equal(res["parity_ok"][det].astype(bool), blk[det].all(axis=1))
    assert np.array_equal(res["frame_ok"].astype(bool), det & blk.all(axis=1) & (be == 0))
    assert np.array_equal(res["iterations"][det], np.floor((2 * its[det].sum(1) + blocks) / (2 * blocks)))
    hard_err = ((llr < 0) != coded[:, : blocks * n]).sum(axis=1)
    assert np.array_equal(res["pre_fec_bit_errors"][det], hard_err[det])
    assert 0 < res["frame_ok"].sum() < F  # the fading channel leaves both outcomes
    # sample against the oracle chain
    code = oracle.Code(n, m, rp, ci, True)
    for f in range(0, F, 170):
        ok, s, peak, cfo, _ = oracle.find_frame(rx[f], phy.pre_half_len, phy.preamble, min(16, phy.cp), 0.2,
                                                phy.fs, max_search=256)
        assert ok == bool(det[f]) and (not ok or s == sync["start"][f]), f
        if not ok:
            continue
        ol, snr, evm = oracle.demod_frame(rx[f][s:s + phy.frame_len], phy, sync["cfo_hz_d"][f], nv, stats=True)
        ol = ol[: blocks * n]
        par, it, oi = [], [], []
        for b in range(blocks):
            r = code.nms(llr[f, b * n:(b + 1) * n], max_iter=30, alpha=0.75)
            par.append(r["ok"])
            it.append(r["iterations"])
            oi.append(r["hard"][code.info_cols])
        oi = np.concatenate(oi)
        assert np.array_equal(oi, got[f]), f
        e = oracle.frame_finish(blocks, n, k, par, it, oi, llr[f], ref[f], coded[f, : blocks * n])
        for key in ("parity_ok", "frame_ok", "frame_id", "bit_errors", "pre_fec_bit_errors", "iterations"):
            assert int(res[key][f]) == e[key], (f, key)
        assert res["snr_db"][f] == pytest.approx(snr, rel=1e-6), f
        assert res["evm_rms"][f] == pytest.approx(evm, rel=1e-6), f
        # the demodulator's LLRs against the oracle's, as the receive tests bound them
        scale = np.sqrt(np.mean(ol.astype(np.float64) ** 2))
        assert np.all(np.abs(llr[f] - ol) <= 1e-5 * np.abs(ol) + 1e-5 * scale), f


def test_frame_decoder_argument_errors():
    n, m, rp, ci = oracle.alist_parse(sa.CCSDS_128_64)
    with sa.Decoder.bp_csr(n, m, rp, ci) as dec:
        with pytest.raises(sa.SiccError):
            sa.FrameDecoder(None, dec, blocks_per_frame=0)
        phy = wf.OfdmPhy(bits_per_symbol=2)
        with sa.OfdmReceiver(phy, 256, 4) as rxr:
            with pytest.raises(sa.SiccError, match="code_n"):
                sa.FrameDecoder(rxr, dec)
        with sa.FrameDecoder(None, dec, blocks_per_frame=2) as fd:
            with pytest.raises(sa.SiccError, match="no OFDM receiver"):
                fd.receive_decode_host(np.zeros((1, 4096), np.complex64))
    with sa.Decoder.osd_only() as od:
        with pytest.raises(sa.SiccError, match="BP"):
            sa.FrameDecoder(None, od, blocks_per_frame=1)
