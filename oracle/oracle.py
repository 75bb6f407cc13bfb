"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke test and bench.py's
cpu_baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

u8p = C.POINTER(C.c_uint8)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int)
u64p = C.POINTER(C.c_uint64)

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run oracle/build.py")
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def alist_parse(path: str):
    lib = load()
    n, m, e = C.c_int(), C.c_int(), C.c_int()
    r = lib.oracle_alist_parse(path.encode(), C.byref(n), C.byref(m), C.byref(e), None, None, 0)
    if r != 0:
        raise ValueError(f"alist parse failed ({r})")
    rp = np.zeros(m.value + 1, np.int32)
    ci = np.zeros(e.value, np.int32)
    r = lib.oracle_alist_parse(path.encode(), C.byref(n), C.byref(m), C.byref(e), _p(rp, i32p), _p(ci, i32p),
                               e.value)
    if r != 0:
        raise ValueError(f"alist parse failed ({r})")
    return n.value, m.value, rp, ci


class Code:
    """Oracle code model: RREF encoder tables (LINK:2811-2997)."""

    def __init__(self, n, m, row_ptr, col_ind, front_info=True):
        lib = load()
        self.n, self.m = int(n), int(m)
        self.rp = np.ascontiguousarray(row_ptr, np.int32)
        self.ci = np.ascontiguousarray(col_ind, np.int32)
        self.words = (self.n + 63) // 64
        self.info_cols = np.zeros(self.n, np.int32)
        self.pivot_cols = np.zeros(self.m, np.int32)
        self.rref = np.zeros(self.m * self.words, np.uint64)
        k, acc = C.c_int(), C.c_int()
        r = lib.oracle_build_ldpc(self.n, self.m, _p(self.rp, i32p), _p(self.ci, i32p), int(front_info), C.byref(k),
                                  _p(self.info_cols, i32p), _p(self.pivot_cols, i32p), _p(self.rref, u64p),
                                  C.byref(acc))
        if r != 0:
            raise ValueError(f"build_ldpc failed ({r})")
        self.k = k.value
        self.acc = acc.value
        self.info_cols = self.info_cols[: self.k].copy()

    def encode(self, info: np.ndarray) -> np.ndarray:
        info = np.ascontiguousarray(info, np.uint8).reshape(-1, self.k)
        out = np.zeros((info.shape[0], self.n), np.uint8)
        load().oracle_encode_batch(info.shape[0], self.n, self.m, self.k, self.acc, _p(self.rp, i32p),
                                   _p(self.ci, i32p), _p(self.info_cols, i32p), _p(self.pivot_cols, i32p),
                                   _p(self.rref, u64p), _p(info, u8p), _p(out, u8p))
        return out

    def generator(self) -> np.ndarray:
        return self.encode(np.eye(self.k, dtype=np.uint8))

    def dense_h(self) -> np.ndarray:
        h = np.zeros((self.m, self.n), np.uint8)
        for r in range(self.m):
            h[r, self.ci[self.rp[r]:self.rp[r + 1]]] = 1
        return h

    def nms(self, llr, max_iter=50, alpha=0.75, offset=0.0, damping=0.0, early_stop=True):
        llr = np.ascontiguousarray(llr, np.float32)
        hard = np.zeros(self.n, np.uint8)
        app = np.zeros(self.n, np.float32)
        best = np.zeros(self.n, np.float32)
        it, bw = C.c_int(), C.c_int()
        ok = load().oracle_nms_decode(self.n, self.m, _p(self.rp, i32p), _p(self.ci, i32p), _p(llr, f32p), max_iter,
                                      C.c_float(alpha), C.c_float(offset), C.c_float(damping), int(early_stop),
                                      _p(hard, u8p), _p(app, f32p), _p(best, f32p), C.byref(it), C.byref(bw))
        return {"ok": bool(ok), "hard": hard, "app": app, "best_app": best, "iterations": it.value,
                "best_synw": bw.value}

    def layer_order(self):
        return layer_order(self.n, self.m, self.rp, self.ci)

    def nms_layered(self, llr, order=None, max_iter=50, alpha=0.75, offset=0.0, damping=0.0, early_stop=True):
        return nms_layered(self.n, self.m, self.rp, self.ci, llr, order, max_iter, alpha, offset, damping, early_stop)

    def bposd(self, llr, gen=None, max_iter=50, alpha=0.75, offset=0.0, damping=0.0, order=1, max_flip=64,
              max_tests=4096, threads=1):
        llr = np.ascontiguousarray(llr, np.float32).reshape(-1, self.n)
        cnt = llr.shape[0]
        gen = np.ascontiguousarray(self.generator() if gen is None else gen, np.uint8)
        code = np.zeros((cnt, self.n), np.uint8)
        info = np.zeros((cnt, self.k), np.uint8)
        ok = np.zeros(cnt, np.uint8)
        used = np.zeros(cnt, np.uint8)
        iters = np.zeros(cnt, np.int32)
        r = load().oracle_bposd_batch(cnt, self.n, self.m, self.k, _p(self.rp, i32p), _p(self.ci, i32p),
                                      _p(gen, u8p), _p(self.info_cols, i32p), _p(llr, f32p), max_iter,
                                      C.c_float(alpha), C.c_float(offset), C.c_float(damping), order, max_flip,
                                      max_tests, threads, _p(code, u8p), _p(info, u8p), _p(ok, u8p), _p(used, u8p),
                                      _p(iters, i32p))
        if r != 0:
            raise RuntimeError("oracle OSD failed (rank deficient generator)")
        return {"code": code, "info": info, "bp_ok": ok, "used_osd": used, "iterations": iters}


def osd(gen: np.ndarray, llr: np.ndarray, order: int, max_flip: int = 64, max_tests: int = 4096):
    gen = np.ascontiguousarray(gen, np.uint8)
    k, n = gen.shape
    llr = np.ascontiguousarray(llr, np.float32)
    code = np.zeros(n, np.uint8)
    met = C.c_double()
    tests = C.c_int()
    r = load().oracle_osd_decode(n, k, _p(gen, u8p), _p(llr, f32p), order, max_flip, max_tests, _p(code, u8p),
                                 C.byref(met), C.byref(tests))
    if r != 0:
        raise RuntimeError("rank deficient generator")
    return code, met.value, tests.value


def awgn_llr(codewords: np.ndarray, ebn0_db: float, rate: float, rng: np.random.Generator) -> np.ndarray:
    """BPSK (0 -> +1) over AWGN, LLR = 2y/sigma^2 with sigma^2 = 1/(2 R Eb/N0)
    (source/cpp_cuda/uhd_cpp/tests/gpu_bp_osd_awgn_bench.cpp:159-172)."""
    sigma2 = 1.0 / (2.0 * rate * 10.0 ** (ebn0_db / 10.0))
    x = 1.0 - 2.0 * codewords.astype(np.float64)
    y = x + np.sqrt(sigma2) * rng.standard_normal(codewords.shape)
    return (2.0 / sigma2 * y).astype(np.float32)


def layer_order(n, m, rp, ci):
    """Greedy first-fit row colouring -> (rows sorted by (colour, row), colour of each row)."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    order = np.zeros(m, np.int32)
    color = np.zeros(m, np.int32)
    load().oracle_layer_order(n, m, _p(rp, i32p), _p(ci, i32p), _p(order, i32p), _p(color, i32p))
    return order, color


def nms_layered(n, m, rp, ci, llr, order=None, max_iter=50, alpha=0.75, offset=0.0, damping=0.0, early_stop=True):
    """Layered normalized min-sum over rows in `order` (default: the greedy colouring order)."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    if order is None:
        order = layer_order(n, m, rp, ci)[0]
    order = np.ascontiguousarray(order, np.int32)
    llr = np.ascontiguousarray(llr, np.float32)
    hard = np.zeros(n, np.uint8)
    app = np.zeros(n, np.float32)
    it = C.c_int()
    ok = load().oracle_nms_layered_decode(n, m, _p(rp, i32p), _p(ci, i32p), _p(order, i32p), _p(llr, f32p),
                                          max_iter, C.c_float(alpha), C.c_float(offset), C.c_float(damping),
                                          int(early_stop), _p(hard, u8p), _p(app, f32p), C.byref(it))
    return {"ok": bool(ok), "hard": hard, "app": app, "iterations": it.value}


# ---------------------------------------------------------------- OFDM ----
def fft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """LINK:3788 radix-2 FFT on a complex vector (FP64)."""
    re = np.ascontiguousarray(x.real, np.float64)
    im = np.ascontiguousarray(x.imag, np.float64)
    load().oracle_fft(_p(re, f64p), _p(im, f64p), len(x), int(inverse))
    return re + 1j * im


def _iq(x: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, np.complex64).view(np.float32))


def find_frame(buf: np.ndarray, half_len: int, preamble: np.ndarray, refine_radius: int, threshold: float,
               fs: float, max_search: int = 0):
    """LINK:5123 find_frame on a complex64 buffer -> (found, start, peak, cfo_hz, coarse_metric)."""
    iq = _iq(buf)
    pre = _iq(preamble)
    start, peak, cfo, coarse = C.c_int(), C.c_double(), C.c_double(), C.c_double()
    ok = load().oracle_find_frame(_p(iq, f32p), len(buf), half_len, _p(pre, f32p), len(preamble), refine_radius,
                                  max_search, C.c_double(threshold), C.c_double(fs), C.byref(start), C.byref(peak),
                                  C.byref(cfo), C.byref(coarse))
    return bool(ok), start.value, peak.value, cfo.value, coarse.value


def demod_frame(frame: np.ndarray, phy, cfo_hz: float, noise_var: float, stats: bool = False):
    """LINK:5331-5521 demodulation of one frame (starting at the preamble) to LLRs;
    with stats, also the frame's SNR (dB) and EVM estimates (LINK:5431-5467):
    returns (llr, snr_db, evm_rms)."""
    iq = _iq(frame)
    used = np.ascontiguousarray(phy.used_shifted, np.int32)
    ps = np.ascontiguousarray(phy.pilot_symbols, np.int32)
    ds = np.ascontiguousarray(phy.data_symbols, np.int32)
    pg = _iq(phy.pilot_grid.reshape(-1))
    const = _iq(phy.constellation)
    out = np.zeros(len(ds) * phy.active_sc * phy.bits_per_symbol, np.float32)
    snr = C.c_double(0.0)
    evm = C.c_double(0.0)
    load().oracle_demod_frame_ex(_p(iq, f32p), phy.nfft, phy.cp, phy.pre_len, phy.num_symbols, phy.active_sc,
                                 _p(used, i32p), len(ps), _p(ps, i32p), _p(pg, f32p), len(ds), _p(ds, i32p),
                                 phy.bits_per_symbol, _p(const, f32p), C.c_double(cfo_hz), C.c_double(phy.fs),
                                 C.c_double(noise_var), _p(out, f32p), C.byref(snr) if stats else None,
                                 C.byref(evm) if stats else None)
    if stats:
        return out, snr.value, evm.value
    return out


def frame_finish(blocks: int, n: int, k: int, block_parity, block_iters, info, llr, ref_info=None,
                 ref_coded=None, gated: bool = False) -> dict:
    """LINK:5482-5591 per-frame aggregation of one frame's blocks (FrameDecodeResult
    fields other than detection, peak, CFO, SNR and EVM)."""
    bp = np.ascontiguousarray(block_parity, np.uint8)
    bi = np.ascontiguousarray(block_iters, np.int32)
    inf = np.ascontiguousarray(info, np.uint8).reshape(-1)
    ll = np.ascontiguousarray(llr, np.float32).reshape(-1)
    ri = None if ref_info is None else np.ascontiguousarray(ref_info, np.uint8).reshape(-1)
    rc = None if ref_coded is None else np.ascontiguousarray(ref_coded, np.uint8).reshape(-1)
    po, it, be, pe, fo = (C.c_int(0) for _ in range(5))
    fid = C.c_uint32(0)
    load().oracle_frame_finish(blocks, n, k, _p(bp, u8p), _p(bi, i32p), _p(inf, u8p), _p(ri, u8p), _p(rc, u8p),
                               _p(ll, f32p), int(gated), C.byref(po), C.byref(it), C.byref(fid), C.byref(be),
                               C.byref(pe), C.byref(fo))
    return {"parity_ok": po.value, "iterations": it.value, "frame_id": fid.value, "bit_errors": be.value,
            "pre_fec_bit_errors": pe.value, "frame_ok": fo.value}


def symbol_index_tables(num_symbols: int, pilot_symbols, data_symbols):
    """Per-symbol pilot / data ordinals (-1 where the symbol is of the other kind)."""
    pi = np.full(num_symbols, -1, np.int32)
    di = np.full(num_symbols, -1, np.int32)
    for i, s in enumerate(pilot_symbols):
        pi[s] = i
    for i, s in enumerate(data_symbols):
        di[s] = i
    return pi, di


def tx_frame(info: np.ndarray, n: int, k: int, phy, amplitude: float, gen_kn: np.ndarray | None = None,
             dvb_rows: tuple | None = None, coded_bits: int | None = None, noise_var: float = 0.0, seed: int = 0,
             precomp_h: np.ndarray | None = None, actual_h: np.ndarray | None = None):
    """PIPE:129-342 transmitter on one frame -> (coded bits, complex64 frame).

    info: [blocks * k] bytes; gen_kn [k][n] bytes, or dvb_rows = (row_info_offsets,
    row_info_cols) for the DVB-S2 accumulator encoder."""
    info = np.ascontiguousarray(info, np.uint8).reshape(-1)
    blocks = info.size // k
    if coded_bits is None:
        coded_bits = len(phy.data_symbols) * phy.active_sc * phy.bits_per_symbol
    pi, di = symbol_index_tables(phy.num_symbols, phy.pilot_symbols, phy.data_symbols)
    used = np.ascontiguousarray(phy.used_shifted, np.int32)
    pg = _iq(phy.pilot_grid.reshape(-1))
    pre = _iq(phy.preamble)
    const = _iq(phy.constellation)
    gen = None if gen_kn is None else np.ascontiguousarray(gen_kn, np.uint8)
    ro = rc = None
    if dvb_rows is not None:
        ro = np.ascontiguousarray(dvb_rows[0], np.int32)
        rc = np.ascontiguousarray(dvb_rows[1], np.int32)
        if rc.size == 0:
            rc = np.zeros(1, np.int32)
    pc = None if precomp_h is None else _iq(np.asarray(precomp_h).reshape(-1))
    ac = None if actual_h is None else _iq(np.asarray(actual_h).reshape(-1))
    frame_len = phy.pre_len + phy.num_symbols * (phy.nfft + phy.cp)
    coded = np.zeros(coded_bits, np.uint8)
    out = np.zeros(2 * frame_len, np.float32)
    r = load().oracle_tx_frame(
        _p(info, u8p), n, k, blocks, _p(gen, u8p), int(dvb_rows is not None), _p(ro, i32p), _p(rc, i32p),
        coded_bits, phy.nfft, phy.cp, phy.pre_len, phy.num_symbols, phy.active_sc, phy.bits_per_symbol,
        _p(pi, i32p), _p(di, i32p), _p(used, i32p), _p(pg, f32p), _p(pre, f32p), _p(const, f32p), _p(pc, f32p),
        _p(ac, f32p), C.c_float(amplitude), C.c_float(noise_var), C.c_uint64(seed), _p(coded, u8p), _p(out, f32p))
    if r != 0:
        raise ValueError("oracle_tx_frame: bad shape")
    return coded, out.view(np.complex64)
