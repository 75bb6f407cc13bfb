"""Synthetic OFDM waveforms for the receive path (numpy, seeded).

Frame layout follows the reference transmitter (build_tx_frame_from_info,
source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp:4590-4629): a two-half
identical Schmidl-Cox preamble, then `num_symbols` OFDM symbols with a cyclic
prefix, pilot symbols at `pilot_symbols`, data symbols carrying Gray-mapped
QAM on `active_sc` subcarriers around DC (LINK:4011-4018), each frame scaled
so its peak magnitude equals `amplitude` (LINK:4620-4627).  The channel adds
exponential-profile Rayleigh multipath, a carrier frequency offset, a timing
offset and AWGN.  These are inputs for tests and the benchmark; the receive
path itself is the HIP library.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def used_shifted_indices(nfft: int, active_sc: int) -> np.ndarray:
    """Shifted-grid indices of the active subcarriers, DC excluded (LINK:4011-4018)."""
    n2, a2 = nfft // 2, active_sc // 2
    return np.array(list(range(n2 - a2, n2)) + list(range(n2 + 1, n2 + a2 + 1)), np.int32)


def constellation(bits_per_symbol: int) -> np.ndarray:
    """Gray QAM table indexed by the symbol's bits read MSB first, unit average
    power (configure_modulation, LINK:2362-2414)."""
    if bits_per_symbol == 1:
        return np.array([1.0, -1.0], np.complex64)
    if bits_per_symbol == 2:
        s = np.float32(1.0 / np.sqrt(2.0))
        return np.array([complex(s, s), complex(s, -s), complex(-s, s), complex(-s, -s)], np.complex64)
    axis = bits_per_symbol // 2
    levels = 1 << axis
    m = 1 << bits_per_symbol
    pts = []
    for idx in range(m):
        gi = idx >> axis
        gq = idx & (levels - 1)
        bi, bq = gi, gq
        t = gi >> 1
        while t:
            bi ^= t
            t >>= 1
        t = gq >> 1
        while t:
            bq ^= t
            t >>= 1
        pts.append(complex((levels - 1) - 2 * bi, (levels - 1) - 2 * bq))
    pts = np.array(pts, np.complex128)
    scale = 1.0 / np.sqrt(np.mean(np.abs(pts) ** 2))
    return (pts * np.float32(scale)).astype(np.complex64)


@dataclass
class OfdmPhy:
    nfft: int = 1024
    cp: int = 72
    num_symbols: int = 14
    active_sc: int = 600
    pre_half_len: int = 512
    bits_per_symbol: int = 4
    pilot_symbols: np.ndarray = field(default_factory=lambda: np.array([0], np.int32))
    fs: float = 15.36e6
    seed: int = 11
    amplitude: float = 0.7

    def __post_init__(self):
        rng = np.random.default_rng(self.seed)
        self.pilot_symbols = np.asarray(self.pilot_symbols, np.int32)
        self.data_symbols = np.array([s for s in range(self.num_symbols) if s not in set(self.pilot_symbols)],
                                     np.int32)
        self.used_shifted = used_shifted_indices(self.nfft, self.active_sc)
        self.constellation = constellation(self.bits_per_symbol)
        q = constellation(2)
        half = q[rng.integers(0, 4, self.pre_half_len)]
        pre = np.concatenate([half, half]).astype(np.complex128)
        pre /= np.sqrt(np.mean(np.abs(pre) ** 2))
        self.preamble = pre.astype(np.complex64)
        self.pre_len = 2 * self.pre_half_len
        self.pilot_grid = q[rng.integers(0, 4, (len(self.pilot_symbols), self.active_sc))].astype(np.complex64)

    @property
    def sym_len(self) -> int:
        return self.nfft + self.cp

    @property
    def frame_len(self) -> int:
        return self.pre_len + self.num_symbols * self.sym_len

    @property
    def coded_bits(self) -> int:
        return len(self.data_symbols) * self.active_sc * self.bits_per_symbol

    def interp_tables(self):
        """Pilot indices and weights per data symbol (LINK:5402-5420)."""
        ps = list(self.pilot_symbols)
        p0s, p1s, al = [], [], []
        for sym in self.data_symbols:
            hi = 0
            while hi + 1 < len(ps) and sym > ps[hi + 1]:
                hi += 1
            p0, p1 = hi, min(hi + 1, len(ps) - 1)
            if sym <= ps[0]:
                p0 = p1 = 0
            elif sym >= ps[-1]:
                p0 = p1 = len(ps) - 1
            a = 0.0 if p0 == p1 else (sym - ps[p0]) / (ps[p1] - ps[p0])
            p0s.append(p0)
            p1s.append(p1)
            al.append(a)
        return np.array(p0s, np.int32), np.array(p1s, np.int32), np.array(al, np.float64)


def map_bits(phy: OfdmPhy, bits: np.ndarray) -> np.ndarray:
    """[F][coded_bits] bits -> [F][coded_bits / B] symbols (MSB-first index)."""
    B = phy.bits_per_symbol
    b = bits.reshape(bits.shape[0], -1, B).astype(np.int64)
    idx = np.zeros(b.shape[:2], np.int64)
    for i in range(B):
        idx = (idx << 1) | b[:, :, i]
    return phy.constellation[idx]


def tx_frames(phy: OfdmPhy, coded_bits: np.ndarray) -> np.ndarray:
    """[F][coded_bits] -> [F][frame_len] complex64 baseband frames."""
    F = coded_bits.shape[0]
    syms = map_bits(phy, coded_bits).reshape(F, len(phy.data_symbols), phy.active_sc)
    grid = np.zeros((F, phy.num_symbols, phy.nfft), np.complex128)
    nat = (phy.used_shifted + phy.nfft // 2) % phy.nfft
    for p, s in enumerate(phy.pilot_symbols):
        grid[:, s, nat] = phy.pilot_grid[p]
    for d, s in enumerate(phy.data_symbols):
        grid[:, s, nat] = syms[:, d]
    td = np.fft.ifft(grid, axis=2)
    body = np.concatenate([td[:, :, -phy.cp:], td], axis=2).reshape(F, -1)
    frames = np.concatenate([np.broadcast_to(phy.preamble, (F, phy.pre_len)), body], axis=1)
    peak = np.maximum(np.abs(frames).max(axis=1, keepdims=True), 1e-12)
    return (frames * (phy.amplitude / peak)).astype(np.complex64)


@dataclass
class ChannelOut:
    rx: np.ndarray          # [F][span] complex64
    offsets: np.ndarray     # [F] true frame start
    cfo_hz: float
    noise_var_time: float   # per-sample AWGN variance
    noise_var_freq: float   # per-subcarrier noise variance after the unnormalized FFT


def channel(phy: OfdmPhy, frames: np.ndarray, rng: np.random.Generator, span: int, taps: int = 6,
            decay_db: float = 3.0, cfo_sc: float = 0.3, max_offset: int = 200, snr_db: float = 14.0,
            fading: bool = True) -> ChannelOut:
    """Rayleigh multipath (exponential power profile, unit total power), CFO,
    a uniform timing offset in [0, max_offset] and AWGN at `snr_db`."""
    F, N = frames.shape
    if span < N + max_offset + taps:
        raise ValueError("span too short for the frame, offset and channel")
    offs = rng.integers(0, max_offset + 1, F)
    rx = np.zeros((F, span), np.complex128)
    if fading:
        p = 10.0 ** (-decay_db * np.arange(taps) / 10.0)
        p /= p.sum()
        h = (rng.standard_normal((F, taps)) + 1j * rng.standard_normal((F, taps))) * np.sqrt(p / 2.0)
    else:
        h = np.zeros((F, taps), np.complex128)
        h[:, 0] = 1.0
    for f in range(F):
        y = np.convolve(frames[f].astype(np.complex128), h[f])
        rx[f, offs[f]:offs[f] + len(y)] = y
    cfo_hz = cfo_sc * phy.fs / phy.nfft
    rx *= np.exp(2j * np.pi * cfo_hz * np.arange(span) / phy.fs)[None, :]
    # SNR is the per-subcarrier SNR of the data carriers: mean |Y|^2 of the first
    # data symbol's active bins (unit-power channel) over the per-bin noise
    # variance nfft * sigma_t^2 of the receiver's unnormalized FFT.
    s0 = phy.pre_len + int(phy.data_symbols[0]) * phy.sym_len + phy.cp
    nat = (phy.used_shifted + phy.nfft // 2) % phy.nfft
    es = float(np.mean(np.abs(np.fft.fft(frames[:, s0:s0 + phy.nfft].astype(np.complex128), axis=1)[:, nat]) ** 2))
    nv_f = es / (10.0 ** (snr_db / 10.0))
    nv = nv_f / phy.nfft
    rx += np.sqrt(nv / 2.0) * (rng.standard_normal(rx.shape) + 1j * rng.standard_normal(rx.shape))
    return ChannelOut(rx.astype(np.complex64), offs.astype(np.int32), cfo_hz, nv, nv_f)
