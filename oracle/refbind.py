"""ctypes binding of oracle/_ref/libref.so, the reference's own CPU path built
from its sources by oracle/build_ref.py.  TEST INFRASTRUCTURE ONLY."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libref.so")

u8p = C.POINTER(C.c_uint8)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int)

_lib = None


def available() -> bool:
    return os.path.exists(LIB_PATH)


def load():
    global _lib
    if _lib is None:
        _lib = C.CDLL(LIB_PATH)
        _lib.ref_last_error.restype = C.c_char_p
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


class RefCode:
    """The reference LdpcCode (read_alist + build_ldpc) and its CPU min-sum."""

    def __init__(self, alist: str, front_info: bool = True):
        lib = load()
        n, m, k = C.c_int(), C.c_int(), C.c_int()
        if lib.ref_code_load(alist.encode(), int(front_info), C.byref(n), C.byref(m), C.byref(k)) != 0:
            raise RuntimeError(lib.ref_last_error().decode())
        self.n, self.m, self.k = n.value, m.value, k.value
        self.info_cols = np.zeros(self.k, np.int32)
        self.pivot_cols = np.zeros(self.m, np.int32)
        lib.ref_code_info_cols(_p(self.info_cols, i32p), _p(self.pivot_cols, i32p))

    def encode(self, info):
        info = np.ascontiguousarray(info, np.uint8).reshape(-1, self.k)
        out = np.zeros((info.shape[0], self.n), np.uint8)
        load().ref_code_encode(_p(info, u8p), info.shape[0], _p(out, u8p))
        return out

    def min_sum(self, llr, max_iter=50, norm=0.75):
        llr = np.ascontiguousarray(llr, np.float32).reshape(-1, self.n)
        c = llr.shape[0]
        code = np.zeros((c, self.n), np.uint8)
        info = np.zeros((c, self.k), np.uint8)
        ok = np.zeros(c, np.uint8)
        it = np.zeros(c, np.int32)
        load().ref_min_sum(_p(llr, f32p), c, max_iter, C.c_double(norm), _p(code, u8p), _p(info, u8p), _p(ok, u8p),
                           _p(it, i32p))
        return code, info, ok, it


class RefPhy:
    """The reference PhyConfig (build_phy_config, LINK:3971) and its CPU receive functions."""

    def __init__(self, nfft=1024, cp=72, num_symbols=14, active_sc=600, pre_half_len=512, modulation="qpsk",
                 pilot_period=14, rate=15.36e6, sync_threshold=0.45, amplitude=0.7, test_seed=20260418):
        lib = load()
        if lib.ref_phy_setup(nfft, cp, num_symbols, active_sc, pre_half_len, modulation.encode(), pilot_period,
                             C.c_double(rate), C.c_double(sync_threshold), C.c_double(amplitude), test_seed) != 0:
            raise RuntimeError(lib.ref_last_error().decode())
        fl, cb, npil, nd, bps, pl = (C.c_int() for _ in range(6))
        lib.ref_phy_dims(C.byref(fl), C.byref(cb), C.byref(npil), C.byref(nd), C.byref(bps), C.byref(pl))
        self.frame_len, self.coded_bits, self.bits_per_symbol = fl.value, cb.value, bps.value
        self.nfft, self.cp, self.num_symbols, self.active_sc = nfft, cp, num_symbols, active_sc
        self.pre_half_len, self.fs, self.sync_threshold = pre_half_len, rate, sync_threshold
        pre = np.zeros(2 * pl.value, np.float32)
        used = np.zeros(active_sc, np.int32)
        ps = np.zeros(npil.value, np.int32)
        ds = np.zeros(nd.value, np.int32)
        pg = np.zeros(2 * npil.value * active_sc, np.float32)
        const = np.zeros(2 << bps.value, np.float32)
        lib.ref_phy_tables(_p(pre, f32p), _p(used, i32p), _p(ps, i32p), _p(ds, i32p), _p(pg, f32p),
                           _p(const, f32p))
        self.preamble = pre.view(np.complex64)
        self.pre_len = len(self.preamble)
        self.used_shifted = used
        self.pilot_symbols = ps
        self.data_symbols = ds
        self.pilot_grid = pg.view(np.complex64).reshape(npil.value, active_sc)
        self.constellation = const.view(np.complex64)

    def build_tx_frame(self, info_bits: np.ndarray) -> np.ndarray:
        info = np.ascontiguousarray(info_bits, np.uint8)
        out = np.zeros(2 * self.frame_len, np.float32)
        if load().ref_build_tx_frame(_p(info, u8p), _p(out, f32p)) < 0:
            raise RuntimeError(load().ref_last_error().decode())
        return out.view(np.complex64)

    def find_frame(self, buf: np.ndarray):
        iq = np.ascontiguousarray(np.asarray(buf, np.complex64).view(np.float32))
        s, p, c = C.c_int(), C.c_double(), C.c_double()
        ok = load().ref_find_frame(_p(iq, f32p), len(buf), C.byref(s), C.byref(p), C.byref(c))
        return bool(ok), s.value, p.value, c.value

    def fft_shifted(self, sym: np.ndarray) -> np.ndarray:
        iq = np.ascontiguousarray(np.asarray(sym, np.complex64).view(np.float32))
        out = np.zeros(2 * len(sym), np.float32)
        load().ref_fft_shifted(_p(iq, f32p), len(sym), _p(out, f32p))
        return out.view(np.complex64)

    def symbol_llr(self, eq: np.ndarray, sigma2: np.ndarray) -> np.ndarray:
        iq = np.ascontiguousarray(np.asarray(eq, np.complex64).view(np.float32))
        s2 = np.ascontiguousarray(sigma2, np.float64)
        out = np.zeros(len(eq) * self.bits_per_symbol, np.float32)
        load().ref_symbol_llr(_p(iq, f32p), len(eq), _p(s2, f64p), _p(out, f32p))
        return out

    def decode_frame(self, frame: np.ndarray, cfo_hz: float, noise_var: float, max_iter: int, norm: float,
                     info_len: int):
        iq = np.ascontiguousarray(np.asarray(frame, np.complex64).view(np.float32))
        info = np.zeros(info_len, np.uint8)
        ok, snr = C.c_int(), C.c_double()
        r = load().ref_decode_frame(_p(iq, f32p), C.c_double(cfo_hz), C.c_double(noise_var), max_iter,
                                    C.c_double(norm), _p(info, u8p), C.byref(ok), C.byref(snr))
        if r < 0:
            raise RuntimeError(load().ref_last_error().decode())
        return info[:r], bool(ok.value)

    def build_info_bits(self, frame_id: int, info_len: int) -> np.ndarray:
        """build_info_bits (LINK:4048): the reference payload of a frame id."""
        out = np.zeros(info_len, np.uint8)
        r = load().ref_build_info_bits(C.c_uint(frame_id), _p(out, u8p))
        if r < 0:
            raise RuntimeError(load().ref_last_error().decode())
        return out[:r]

    def decode_frame_full(self, frame: np.ndarray, cfo_hz: float, peak: float, noise_var: float, max_iter: int,
                          norm: float, info_len: int, snr_gate_db: float = -120.0, reference_test: bool = True,
                          expected_id: int = -1) -> dict:
        """decode_frame (LINK:5331-5603) with every FrameDecodeResult field."""
        iq = np.ascontiguousarray(np.asarray(frame, np.complex64).view(np.float32))
        info = np.zeros(info_len, np.uint8)
        ires = np.zeros(7, np.int64)
        dres = np.zeros(4, np.float64)
        r = load().ref_decode_frame_full(_p(iq, f32p), C.c_double(cfo_hz), C.c_double(peak),
                                         C.c_double(noise_var), max_iter, C.c_double(norm), C.c_double(snr_gate_db),
                                         int(reference_test), C.c_longlong(expected_id), _p(info, u8p),
                                         ires.ctypes.data_as(C.POINTER(C.c_longlong)), _p(dres, f64p))
        if r < 0:
            raise RuntimeError(load().ref_last_error().decode())
        keys = ("detected", "parity_ok", "frame_ok", "frame_id", "bit_errors", "pre_fec_bit_errors", "iterations")
        out = {k: int(v) for k, v in zip(keys, ires)}
        out.update({"peak": dres[0], "cfo_hz": dres[1], "snr_db": dres[2], "evm_rms": dres[3], "info": info[:r]})
        return out
