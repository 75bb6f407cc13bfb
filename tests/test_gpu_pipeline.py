"""GPU parity of usrp_gpu_pipeline_bridge::Pipeline (through its C layer).

TX: the HIP transmitter against the oracle's restatement of the reference GPU
transmitter (frame within 1e-5 of its peak: the device inverse FFT is FP32,
like the reference GPU transmitter's single-precision cuFFT, the oracle's FP64,
and the AWGN goes through logf / cosf, whose last-ulp rounding differs between
the device and host libraries).  RX: sync
and LLRs against the oracle's find_frame / demod_frame, with the bound of the
batched receiver tests.  Loopback: noise-free TX -> RX gives sign(LLR) = coded
bit (the reference pipeline bench's known-answer check).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import sicc_amd as sa
import waveform as wf

from test_oracle_ofdm import dvbs2_rows

pytestmark = pytest.mark.gpu

N, K = 128, 64


def _phy(bits=2, active_sc=512, num_symbols=6, pilots=(0, 3), nfft=1024):
    return wf.OfdmPhy(bits_per_symbol=bits, active_sc=active_sc, num_symbols=num_symbols,
                      pilot_symbols=np.array(pilots, np.int32), nfft=nfft, cp=max(72 * nfft // 1024, 8),
                      pre_half_len=nfft // 2)


def _tables(phy, gen=None, dvb=None, precomp=None, actual=None):
    pi, di = oracle.symbol_index_tables(phy.num_symbols, phy.pilot_symbols, phy.data_symbols)
    p0, p1, al = phy.interp_tables()
    t = dict(used_indices=phy.used_shifted, pilot_symbol_index=pi, data_symbol_index=di,
             pilot_symbols=phy.pilot_symbols, data_symbols=phy.data_symbols, interp_p0=p0, interp_p1=p1,
             interp_alpha=al.astype(np.float32), pilot_grid=phy.pilot_grid, preamble=phy.preamble,
             precomp_h=precomp, actual_h=actual)
    if gen is not None:
        t["generator_kn"] = gen
    if dvb is not None:
        t["dvb_s2_accumulator"] = True
        t["row_info_offsets"], t["row_info_cols"] = dvb
    return t


def _code():
    n_, m_, rp, ci = oracle.alist_parse(sa.CCSDS_128_64)
    return oracle.Code(n_, m_, rp, ci, True)


def _pipeline(phy, amp=0.2, noise_var=1e-3, max_rx=0, **tables):
    blocks = phy.coded_bits // N
    p = sa.Pipeline(phy, N, K, blocks, slots=3, max_rx_search_samples=max_rx, tx_amplitude=amp, noise_var=noise_var)
    p.upload_static_tables(**_tables(phy, **tables))
    return p, blocks


def _close(got, ref):
    tol = 1e-5 * np.abs(ref).max()
    err = np.abs(got.astype(np.complex128) - ref).max()
    assert err <= tol, (err, tol)


def test_tx_matches_oracle_generator_channel_and_noise():
    print("[case] pipeline TX generator path", flush=True)
    phy = _phy()
    gen = _code().generator()
    rng = np.random.default_rng(21)
    h = (rng.standard_normal((phy.num_symbols, phy.active_sc)) +
         1j * rng.standard_normal((phy.num_symbols, phy.active_sc))).astype(np.complex64)
    g = (0.5 + 0.5 * rng.random((phy.num_symbols, phy.active_sc))).astype(np.complex64)
    for precomp, actual in ((None, None), (g, None), (None, h), (g, h)):
        p, blocks = _pipeline(phy, amp=0.25, gen=gen, precomp=precomp, actual=actual)
        info = rng.integers(0, 2, blocks * K, dtype=np.uint8)
        _, ref = oracle.tx_frame(info, N, K, phy, 0.25, gen_kn=gen, precomp_h=precomp, actual_h=actual)
        _close(p.submit_tx_frame(1, info), ref)
        d = p.submit_tx_frame_device(2, info)
        p.synchronize_tx(2)
        _close(p.read_frame(d), ref)
        # frequency-domain AWGN only when asked for
        _, refn = oracle.tx_frame(info, N, K, phy, 0.25, gen_kn=gen, precomp_h=precomp, actual_h=actual,
                                  noise_var=0.02, seed=(5 << 32) + 9)
        d = p.submit_tx_frame_device(0, info, noise_var=0.02, seed=(5 << 32) + 9, add_noise=True)
        p.synchronize_tx(0)
        _close(p.read_frame(d), refn)
        d = p.submit_tx_frame_device(0, info, noise_var=0.02, seed=3, add_noise=False)
        p.synchronize_tx(0)
        _close(p.read_frame(d), ref)
        p.close()


@pytest.mark.parametrize("nfft,active_sc", [(128, 64), (512, 256), (2048, 1024)])
def test_tx_fft_sizes_match_oracle(nfft, active_sc):
    """The transmitter's inverse FFT at sizes other than 1024 (ending with a
    radix-2 stage), against the oracle TX at the frame tolerance."""
    print(f"[case] pipeline TX nfft={nfft}", flush=True)
    phy = _phy(active_sc=active_sc, nfft=nfft)
    gen = _code().generator()
    p, blocks = _pipeline(phy, amp=0.25, gen=gen)
    rng = np.random.default_rng(nfft)
    info = rng.integers(0, 2, blocks * K, dtype=np.uint8)
    _, ref = oracle.tx_frame(info, N, K, phy, 0.25, gen_kn=gen)
    _close(p.submit_tx_frame(1, info), ref)
    p.close()


def test_tx_dvbs2_encoder_matches_oracle():
    print("[case] pipeline TX DVB-S2 accumulator path", flush=True)
    phy = _phy(bits=4, active_sc=512, num_symbols=5, pilots=(0,))
    dvb = dvbs2_rows(K, N - K, 5, 22)
    p, blocks = _pipeline(phy, dvb=dvb)
    rng = np.random.default_rng(23)
    for _ in range(3):
        info = rng.integers(0, 2, blocks * K, dtype=np.uint8)
        _, ref = oracle.tx_frame(info, N, K, phy, 0.2, dvb_rows=dvb)
        _close(p.submit_tx_frame(0, info), ref)
    p.close()


def test_loopback_sync_and_demod_paths():
    print("[case] pipeline loopback", flush=True)
    phy = _phy(bits=4, active_sc=512, num_symbols=6, pilots=(0, 3))
    code = _code()
    gen = code.generator()
    lead = 300
    p, blocks = _pipeline(phy, amp=0.2, noise_var=1e-4, max_rx=phy.frame_len + lead, gen=gen)
    rng = np.random.default_rng(24)
    for trial in range(3):
        info = rng.integers(0, 2, blocks * K, dtype=np.uint8)
        coded = code.encode(info.reshape(blocks, K)).reshape(-1)
        frame = p.submit_tx_frame(trial % 3, info)
        # find_frame on a buffer with `lead` samples of weak noise before the frame
        buf = np.concatenate([(1e-4 * rng.standard_normal(lead)).astype(np.complex64), frame])
        start, metric, cfo = p.find_frame(trial % 3, buf)
        ok, ostart, opeak, ocfo, _ = oracle.find_frame(buf, phy.pre_half_len, phy.preamble, min(16, phy.cp),
                                                       0.45, phy.fs)
        assert ok and start == ostart == lead
        assert metric == pytest.approx(opeak, rel=1e-5) and cfo == pytest.approx(ocfo, abs=1e-3)
        # demod from host, device and device-to-device entry points agree, and
        # match the oracle demodulator; noise-free: sign(LLR) = coded bit
        llr_h, sync = p.demod_frame_to_host_llr(0, frame)
        assert sync[0] == 0
        ref = oracle.demod_frame(frame, phy, float(np.float32(sync[2])), 1e-4)
        scale = np.sqrt(np.mean(ref.astype(np.float64) ** 2))
        assert np.all(np.abs(llr_h - ref) <= 1e-5 * np.abs(ref) + 1e-5 * scale)
        assert np.array_equal((llr_h < 0).astype(np.uint8), coded)
        d_llr, sync2 = p.demod_frame_to_device_llr(1, frame)
        assert sync2 == sync
        assert np.array_equal(p.read_llr(d_llr), llr_h)
        d_frame = p.submit_tx_frame_device(2, info)
        p.synchronize_tx(2)
        assert p.find_frame_device(2, d_frame, phy.frame_len)[0] == 0
        # the device-frame entry takes the float CFO of the SyncResult
        d_llr2 = p.demod_frame_device_to_device_llr(2, d_frame, sync)
        llr_d = p.read_llr(d_llr2)
        assert np.all(np.abs(llr_d - ref) <= 1e-5 * np.abs(ref) + 1e-5 * scale)
        assert np.array_equal((llr_d < 0).astype(np.uint8), coded)
    p.close()


def test_batched_tx_equals_single_frames_and_decodes():
    print("[case] pipeline batched TX", flush=True)
    import torch

    phy = _phy(bits=2, active_sc=512, num_symbols=6, pilots=(0, 3))
    gen = _code().generator()
    p, blocks = _pipeline(phy, amp=0.2, gen=gen)
    frames = 40
    rng = np.random.default_rng(25)
    info = rng.integers(0, 2, (frames, blocks * K), dtype=np.uint8)
    d_info = torch.from_numpy(info).cuda()
    d_out = torch.empty((frames, phy.frame_len, 2), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()  # the library's stream does not order after torch's
    p.tx_batch_device(d_info.data_ptr(), frames, d_out.data_ptr(), noise_var=0.01, seed=1000)
    torch.cuda.synchronize()
    got = d_out.cpu().numpy().view(np.complex64).reshape(frames, phy.frame_len)
    for f in (0, 1, 17, frames - 1):
        d = p.submit_tx_frame_device(0, info[f], noise_var=0.01, seed=1000 + f, add_noise=True)
        p.synchronize_tx(0)
        assert np.array_equal(p.read_frame(d), got[f])
        _, ref = oracle.tx_frame(info[f], N, K, phy, 0.2, gen_kn=gen, noise_var=0.01, seed=1000 + f)
        _close(got[f], ref)
    # the batched receiver recovers the bits from the device-resident frames
    with sa.OfdmReceiver(phy, N, blocks, max_search=64, noise_var=0.01 * 0.2 ** 2) as rxr:
        sync = torch.empty((frames, 24), dtype=torch.uint8, device="cuda")
        llr = torch.empty((frames, blocks * N), dtype=torch.float32, device="cuda")
        rxr.receive_device(d_out.data_ptr(), frames, phy.frame_len, sync.data_ptr(), llr.data_ptr())
        torch.cuda.synchronize()
    code = _code()
    coded = code.encode(info.reshape(-1, K)).reshape(frames, -1)
    ber = np.mean((llr.cpu().numpy() < 0) != coded)
    assert ber < 0.02, ber
    p.close()


def test_error_paths():
    phy = _phy()
    blocks = phy.coded_bits // N
    with pytest.raises(sa.SiccError, match="coded bits per frame must match"):
        sa.Pipeline(phy, N, K, blocks - 1)
    with pytest.raises(sa.SiccError, match="blocks_per_frame must be set"):
        sa.Pipeline(phy, N, K, 0)
    p = sa.Pipeline(phy, N, K, blocks, slots=2)
    info = np.zeros(blocks * K, np.uint8)
    with pytest.raises(sa.SiccError, match="before uploading"):
        p.submit_tx_frame(0, info)
    bad = _tables(phy, gen=_code().generator())
    bad["interp_alpha"] = bad["interp_alpha"][:-1]
    with pytest.raises(sa.SiccError, match="interp_alpha has wrong size"):
        p.upload_static_tables(**bad)
    p.upload_static_tables(**_tables(phy))  # no generator: RX works, TX refuses
    with pytest.raises(sa.SiccError, match="generator matrix"):
        p.submit_tx_frame(0, info)
    with pytest.raises(sa.SiccError, match="invalid GPU pipeline slot"):
        p.find_frame(2, np.zeros(phy.frame_len, np.complex64))
    with pytest.raises(sa.SiccError, match="exceeds slot allocation"):
        p.find_frame(0, np.zeros(phy.frame_len + 1, np.complex64))
    with pytest.raises(sa.SiccError, match="shorter than the preamble"):
        p.find_frame(0, np.zeros(phy.pre_len - 1, np.complex64))
    start, metric, _ = p.find_frame(0, np.zeros(phy.frame_len, np.complex64))
    assert start == -1 and metric < 0.45
    p.close()
