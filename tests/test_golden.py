"""The oracle against golden fixtures written by the reference's own CPU path
(tests/golden/make_golden.py, from oracle/_ref built out of the reference
sources).  These run everywhere, with or without the reference checkout.

Bars: bit-exact for encoder, min-sum decisions / flags / iteration counts,
decoded frames and FFT outputs; find_frame start exact, peak and CFO to 1e-9
relative (the same double sums in a different association); max-log LLRs
exact; the transmitted frame equal up to the reference's peak normalization.
"""
from __future__ import annotations

import os
from types import SimpleNamespace

import numpy as np
import pytest

import oracle

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ALIST = os.path.join(os.path.dirname(G), "..", "sicc-baseband-proto_amd", "data", "CCSDS_ldpc_n128_k64.alist")


@pytest.fixture(scope="module")
def ldpc():
    return np.load(os.path.join(G, "ldpc_ccsds128.npz"))


@pytest.fixture(scope="module")
def ofdm():
    return np.load(os.path.join(G, "ofdm_small.npz"))


@pytest.fixture(scope="module")
def code():
    n, m, rp, ci = oracle.alist_parse(ALIST)
    return oracle.Code(n, m, rp, ci, True)


def test_encoder_and_code_model(ldpc, code):
    assert np.array_equal(code.info_cols, ldpc["info_cols"])
    assert np.array_equal(code.encode(ldpc["enc_info"]), ldpc["enc_cw"])


@pytest.mark.parametrize("tag", ["a", "b"])
def test_min_sum_bit_exact(ldpc, code, tag):
    iters, norm = ldpc[f"ms_cfg_{tag}"]
    llr = ldpc["ms_llr"]
    for i in range(llr.shape[0]):
        o = code.nms(llr[i], max_iter=int(iters), alpha=float(norm))
        assert np.array_equal(o["hard"], ldpc[f"ms_code_{tag}"][i]), i
        assert np.array_equal(o["hard"][code.info_cols], ldpc[f"ms_info_{tag}"][i]), i
        assert o["ok"] == bool(ldpc[f"ms_ok_{tag}"][i]), i
        assert o["iterations"] == ldpc[f"ms_it_{tag}"][i], i
    assert 0 < ldpc[f"ms_ok_{tag}"].sum() < llr.shape[0]  # both outcomes present


def _phy(z):
    nfft, cp, S, A, half, B, flen, blocks = (int(v) for v in z["dims"])
    return SimpleNamespace(nfft=nfft, cp=cp, num_symbols=S, active_sc=A, pre_half_len=half, bits_per_symbol=B,
                           frame_len=flen, blocks=blocks, fs=float(z["fs"][0]), preamble=z["preamble"],
                           pre_len=len(z["preamble"]), used_shifted=z["used_shifted"],
                           pilot_symbols=z["pilot_symbols"], data_symbols=z["data_symbols"],
                           pilot_grid=z["pilot_grid"], constellation=z["constellation"])


def test_transmitter_matches_up_to_scale(ofdm, code):
    phy = _phy(ofdm)
    _, ours = oracle.tx_frame(ofdm["tx_info"], 128, 64, phy, 0.2, gen_kn=code.generator())
    ref = ofdm["tx_frame"].astype(np.complex128)
    s = np.vdot(ref, ours) / np.vdot(ref, ref)
    assert abs(s.imag) < 1e-6 * abs(s.real)
    assert np.abs(ours - s.real * ref).max() <= 2e-6 * np.abs(ours).max()


def test_find_frame(ofdm):
    phy = _phy(ofdm)
    ok, start, peak, cfo = ofdm["ff_out"]
    o_ok, o_s, o_peak, o_cfo, _ = oracle.find_frame(ofdm["ff_buf"], phy.pre_half_len, phy.preamble,
                                                     min(16, phy.cp), 0.45, phy.fs)
    assert bool(ok) and o_ok and o_s == int(start) == 57
    assert o_peak == pytest.approx(peak, rel=1e-9) and o_cfo == pytest.approx(cfo, rel=1e-9)


def test_fft_to_shifted_grid_bit_exact(ofdm):
    n = ofdm["fft_in"].shape[1]
    for x, ref in zip(ofdm["fft_in"], ofdm["fft_out"]):
        nat = oracle.fft(x.astype(np.complex128))
        assert np.array_equal(nat[(np.arange(n) + n // 2) % n].astype(np.complex64), ref)


def test_maxlog_llr_formula(ofdm):
    # the demodulator's last step as ofdm_oracle.c and ofdm_demod_kernel compute it (float distances,
    # Gray index MSB first, (min_{b=1} - min_{b=0}) / sigma2_eq in double): equal to the reference's output
    phy = _phy(ofdm)
    B = phy.bits_per_symbol
    eq, s2 = ofdm["llr_eq"], ofdm["llr_s2"]
    dre = (eq.real[:, None] - phy.constellation.real[None, :]).astype(np.float32)
    dim = (eq.imag[:, None] - phy.constellation.imag[None, :]).astype(np.float32)
    d2 = (dre * dre + dim * dim).astype(np.float32)
    idx = np.arange(1 << B)
    out = np.zeros((len(eq), B), np.float32)
    for b in range(B):
        one = ((idx >> (B - 1 - b)) & 1) == 1
        out[:, b] = ((d2[:, one].min(1) - d2[:, ~one].min(1)).astype(np.float64) / s2).astype(np.float32)
    assert np.array_equal(out.reshape(-1), ofdm["llr_out"])


def test_decode_frame_end_to_end(ofdm, code):
    phy = _phy(ofdm)
    cfo, nv, iters, norm = ofdm["df_args"]
    llr = oracle.demod_frame(ofdm["df_frame"], phy, float(cfo), float(nv))
    info, ok = [], True
    for b in range(phy.blocks):
        r = code.nms(llr[b * 128:(b + 1) * 128], max_iter=int(iters), alpha=float(norm))
        info.append(r["hard"][code.info_cols])
        ok &= r["ok"]
    assert np.array_equal(np.concatenate(info), ofdm["df_info"])
    assert ok == bool(ofdm["df_ok"][0])


@pytest.fixture(scope="module")
def frames():
    return np.load(os.path.join(G, "frame_small.npz"))


FRAME_KEYS = ("detected", "parity_ok", "frame_ok", "frame_id", "bit_errors", "pre_fec_bit_errors", "iterations")


def oracle_frame(z, code, f, gate_db=None):
    """The oracle chain on frame f of the fixture: demod (with SNR / EVM), NMS per
    block, per-frame aggregation."""
    phy = _phy(z)
    start, peak, cfo = z["sync"][f]
    rx = z["bufs"][f][int(start):int(start) + phy.frame_len]
    iters, norm = z["bp"]
    llr, snr, evm = oracle.demod_frame(rx, phy, float(cfo), float(z["noise_var"][0]), stats=True)
    n = code.n
    par, its, info = [], [], []
    for b in range(phy.blocks):
        r = code.nms(llr[b * n:(b + 1) * n], max_iter=int(iters), alpha=float(norm))
        par.append(r["ok"])
        its.append(r["iterations"])
        info.append(r["hard"][code.info_cols])
    info = np.concatenate(info)
    ref_coded = code.encode(z["ref_info"][f].reshape(phy.blocks, code.k)).reshape(-1)
    gated = gate_db is not None and snr < gate_db
    out = oracle.frame_finish(phy.blocks, n, code.k, par, its, info, llr[:phy.blocks * n], z["ref_info"][f],
                              ref_coded, gated=gated)
    out.update(detected=1, snr_db=snr, evm_rms=evm, info=info)
    return out


@pytest.mark.parametrize("gated", [False, True])
def test_frame_result_matches_reference(frames, code, gated):
    """Every FrameDecodeResult field of the reference decode_frame (reference-test
    mode), with and without the SNR gate: integers and decoded bits exact, SNR and
    EVM to 1e-9 relative (double sums of the same floats)."""
    z = frames
    ri, rd = (z["gate_int"], z["gate_dbl"]) if gated else (z["res_int"], z["res_dbl"])
    gate = float(z["gate_db"][0]) if gated else None
    for f in range(len(z["bufs"])):
        o = oracle_frame(z, code, f, gate)
        for j, key in enumerate(FRAME_KEYS):
            assert o[key] == ri[f][j], (f, key)
        assert o["snr_db"] == pytest.approx(rd[f][2], rel=1e-9, abs=1e-12), f
        assert o["evm_rms"] == pytest.approx(rd[f][3], rel=1e-9), f
        if not (gated and o["snr_db"] < gate):
            assert np.array_equal(o["info"], z["info"][f]), f
    # both outcomes present: decoded and failed frames, gated and ungated ones
    assert 0 < ri[:, 2].sum() < len(ri)
    if gated:
        assert (ri[:, 1] == 0).sum() > (z["res_int"][:, 1] == 0).sum()
