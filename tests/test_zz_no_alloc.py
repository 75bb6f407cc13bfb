"""The allocation-free entry point sicc_decode_device_noalloc, through the C ABI,
against sicc_decode_device on the same handle: every output bit-exact on two
calls in a row, the second with a smaller batch (the shared scratch array keeps
the first call's contents).  Limits and argument errors are checked too.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle
import sicc_amd as sa

pytestmark = pytest.mark.gpu

OUT_KEYS = ("info", "code", "app", "ok", "used", "iters", "dist")


def _outputs(torch, cw, n, k, dev):
    return {"info": torch.full((cw, k), 7, dtype=torch.uint8, device=dev),
            "code": torch.full((cw, n), 7, dtype=torch.uint8, device=dev),
            "app": torch.full((cw, n), 7.0, dtype=torch.float32, device=dev),
            "ok": torch.full((cw,), 7, dtype=torch.uint8, device=dev),
            "used": torch.full((cw,), 7, dtype=torch.uint8, device=dev),
            "iters": torch.full((cw,), -7, dtype=torch.int32, device=dev),
            "dist": torch.full((cw,), 7.0, dtype=torch.float32, device=dev)}


def _call(lib, fn, dec, d_llr, cw, o, stream):
    args = [C.c_void_p(o[key].data_ptr()) for key in OUT_KEYS]
    return getattr(lib, fn)(C.c_ulonglong(dec.handle), C.c_void_p(d_llr.data_ptr()), C.c_int(cw), *args,
                            C.c_void_p(stream))


def _compare(lib, torch, dec, llr, sizes):
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    for cw in sizes:
        print(f"[case] noalloc n={dec.n} codewords={cw}", flush=True)
        d_llr = torch.from_numpy(np.ascontiguousarray(llr[:cw])).to(dev)
        a = _outputs(torch, cw, dec.n, dec.k, dev)
        b = _outputs(torch, cw, dec.n, dec.k, dev)
        sa.check(_call(lib, "sicc_decode_device", dec, d_llr, cw, a, stream))
        sa.check(_call(lib, "sicc_decode_device_noalloc", dec, d_llr, cw, b, stream))
        torch.cuda.synchronize()
        for key in OUT_KEYS:
            assert torch.equal(a[key], b[key]), (cw, key)


def _lib():
    lib = sa.load()
    sig = [C.c_ulonglong, C.c_void_p, C.c_int] + [C.c_void_p] * 8
    lib.sicc_decode_device.argtypes = sig
    lib.sicc_decode_device_noalloc.argtypes = sig
    lib.sicc_decode_noalloc_capacity.argtypes = [C.c_ulonglong, C.POINTER(C.c_int), C.POINTER(C.c_ulonglong)]
    return lib


def test_noalloc_bposd_matches_allocating_entry():
    """CCSDS (128,64) BP + order-1 OSD at 2.5 dB (a third of the frames reach OSD):
    12 000 then 3 000 codewords, plus a batch at the exact capacity."""
    import torch

    lib = _lib()
    n, m, rp, ci = oracle.alist_parse(sa.CCSDS_128_64)
    code = oracle.Code(n, m, rp, ci, True)
    rng = np.random.default_rng(77)
    with sa.Decoder.stream(sa.stream_config(max_iter=50, alpha=0.75)) as dec:
        dec.set_osd(1, 64, 4096)
        cap, nbytes = C.c_int(), C.c_ulonglong()
        sa.check(lib.sicc_decode_noalloc_capacity(dec.handle, C.byref(cap), C.byref(nbytes)))
        assert nbytes.value == 8 << 20 and cap.value == (nbytes.value - 512) // (4 * n + 4) == 16256
        info = rng.integers(0, 2, (cap.value, code.k), dtype=np.uint8)
        llr = oracle.awgn_llr(code.encode(info), 2.5, 0.5, rng)
        _compare(lib, torch, dec, llr, [12000, 3000, cap.value])
        # one codeword over the limit: a clear error, nothing launched
        dev = torch.device("cuda:0")
        d_llr = torch.zeros((cap.value + 1, n), dtype=torch.float32, device=dev)
        o = _outputs(torch, cap.value + 1, n, code.k, dev)
        st = _call(lib, "sicc_decode_device_noalloc", dec, d_llr, cap.value + 1, o, 0)
        assert st == sa.STATUS_INVALID_ARGUMENT
        assert "exceed the scratch limit" in lib.cuda_osd_get_last_error().decode()


def test_noalloc_bp_work_stealing_matches_allocating_entry():
    """A BP-only QC n = 256 decoder with work stealing (the cursor lives in the
    array): 20 000 then 500 codewords."""
    import torch

    lib = _lib()
    m, rp, ci = sa.qc_code(256, 8, 3, seed=1264)
    rng = np.random.default_rng(78)
    with sa.Decoder.bp_csr(256, m, rp, ci, max_iter=50, alpha=0.75, schedule=0) as dec:
        dec.set_tune(1, 4, 0, 32)
        info = rng.integers(0, 2, (20000, dec.k), dtype=np.uint8)
        llr = oracle.awgn_llr(dec.encode(info), 2.0, 0.5, rng)
        _compare(lib, torch, dec, llr, [20000, 500])


def test_noalloc_rejects_the_long_code_kernel():
    import torch

    lib = _lib()
    m, rp, ci = sa.qc_code(1024, 8, 3, seed=64)
    dev = torch.device("cuda:0")
    with sa.Decoder.bp_csr(1024, m, rp, ci, max_iter=10, alpha=0.75, schedule=2) as dec:
        d_llr = torch.ones((4, 1024), dtype=torch.float32, device=dev)
        o = _outputs(torch, 4, 1024, dec.k, dev)
        st = _call(lib, "sicc_decode_device_noalloc", dec, d_llr, 4, o, 0)
        assert st == sa.STATUS_INVALID_ARGUMENT
        assert "short-code flooding kernel" in lib.cuda_osd_get_last_error().decode()
