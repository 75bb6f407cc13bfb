"""GPU parity of the long-code BP kernel (layered and flooding schedules).

Bit-exact against the oracle: hard bits, success flags, iteration counts and
APP floats, for DVB-S2 N16200 (rates 1/2 and 5/6, APP in LDS) and N64800
(APP in global scratch), the reference's default offset / damping, a short
code forced onto the long kernel, and the automatic schedule choice
(layered for n >= 1000, as source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp:3658).
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
import sicc_amd as sa

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ccsds():
    return _code("CCSDS_ldpc_n128_k64.alist")


def _code(name):
    n, m, rp, ci = oracle.alist_parse(os.path.join(sa.DATA_DIR, name))
    return oracle.Code(n, m, rp, ci, True)


def _llr(code, count, ebn0, seed):
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 2, (count, code.k), dtype=np.uint8)
    cw = code.encode(info)
    return cw, oracle.awgn_llr(cw, ebn0, code.k / code.n, rng)


def _check(code, llr, schedule, max_iter, alpha=0.75, offset=0.0, damping=0.0):
    with sa.Decoder.bp_csr(code.n, code.m, code.rp, code.ci, max_iter=max_iter, alpha=alpha, offset=offset,
                           damping=damping, schedule=schedule) as dec:
        g = dec.decode_host(llr, want_app=True)
    layered = schedule >= 2 or (schedule < 0 and code.n >= 1000)
    ok_count = 0
    for i in range(llr.shape[0]):
        if layered:
            o = code.nms_layered(llr[i], max_iter=max_iter, alpha=alpha, offset=offset, damping=damping)
        else:
            o = code.nms(llr[i], max_iter=max_iter, alpha=alpha, offset=offset, damping=damping)
        assert g["bp_ok"][i] == o["ok"], i
        assert g["iterations"][i] == o["iterations"], (i, g["iterations"][i], o["iterations"])
        assert np.array_equal(g["code"][i], o["hard"]), i
        assert np.array_equal(g["app"][i].view(np.uint32), o["app"].view(np.uint32)), i
        assert np.array_equal(g["info"][i], o["hard"][code.info_cols]), i
        ok_count += int(o["ok"])
    return ok_count


@pytest.mark.parametrize("schedule", [2, 0])
def test_dvbs2_short_rate_half(schedule):
    print(f"[case] long BP DVB-S2 N16200 r1/2 schedule={schedule}", flush=True)
    code = _code("DVB_S2_short_N16200_rate_1_2.alist")
    _, llr = _llr(code, 24, 1.7, 31)
    ok = _check(code, llr, schedule, max_iter=30)
    assert 0 < ok < 24  # both outcomes exercised


def test_dvbs2_reference_defaults_offset_damping():
    # the reference's CSR decoder defaults: normalization 0.80, offset 0.15, damping 0.15
    print("[case] long BP offset/damping", flush=True)
    code = _code("DVB_S2_short_N16200_rate_1_2.alist")
    _, llr = _llr(code, 12, 1.9, 32)
    _check(code, llr, 2, max_iter=25, alpha=0.8, offset=0.15, damping=0.15)
    _check(code, llr, 0, max_iter=25, alpha=0.8, offset=0.15, damping=0.15)


def test_dvbs2_rate_five_sixths_wide_rows():
    print("[case] long BP DVB-S2 N16200 r5/6 (row degree 15-19)", flush=True)
    code = _code("DVB_S2_short_N16200_rate_5_6.alist")
    _, llr = _llr(code, 16, 4.2, 33)
    _check(code, llr, 2, max_iter=30)
    _check(code, llr, 0, max_iter=30)


def test_dvbs2_normal_frame_app_in_global_memory():
    print("[case] long BP DVB-S2 N64800 r1/2", flush=True)
    code = _code("DVB_S2_N64800_R12.alist")
    _, llr = _llr(code, 6, 1.7, 34)
    _check(code, llr, 2, max_iter=20)
    _check(code, llr, 0, max_iter=20)


def test_short_code_on_long_kernel_and_auto_schedule(ccsds):
    print("[case] long BP short codes / auto schedule", flush=True)
    _, llr = _llr(ccsds, 200, 2.0, 35)
    _check(ccsds, llr, 2, max_iter=30)       # layered through the long kernel
    _check(ccsds, llr, -1, max_iter=30)      # auto: flooding for n < 1000 (short kernel)
    m, rp, ci = sa.qc_code(1024, 8, 3, seed=9)
    qc = oracle.Code(1024, m, rp, ci, True)
    _, llr = _llr(qc, 64, 1.8, 36)
    _check(qc, llr, -1, max_iter=30)         # auto: layered for n >= 1000


def test_long_kernel_rejects_osd_and_tuning():
    code = _code("DVB_S2_short_N16200_rate_1_2.alist")
    with sa.Decoder.bp_csr(code.n, code.m, code.rp, code.ci, schedule=2) as dec:
        with pytest.raises(sa.SiccError, match="short-code"):
            dec.set_tune(1, 1, 0)
