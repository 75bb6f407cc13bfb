"""Statistical parity with the reference's published decoder-only result.

The reference measured its GPU OSD-only decoder on the CCSDS (128,64) code at
Eb/N0 3 dB over 1 000 codewords: 1 frame error, FER 1.0e-3
(performance_results/tables/decoder_awgn_summary.csv:2), with the setup of
source/cpp_cuda/uhd_cpp/tests/gpu_bp_osd_awgn_bench.cpp:139-172: the all-zero
codeword, y = 1 + sigma n, LLR = (2 / sigma^2) y, sigma^2 = 1 / (2 R Eb/N0),
decoded by cuda_osd_decode_osd_batch in batches of 4 096.  The same call here on
200 000 codewords must give a frame error rate whose Clopper-Pearson 95 %
interval overlaps the published point's (about [2.5e-5, 5.6e-3]).  The noise
comes from numpy, not the reference's mt19937 Box-Muller: only the statistics
can agree, not the frames.
"""
from __future__ import annotations

import numpy as np
import pytest

import sicc_amd as sa

pytestmark = pytest.mark.gpu


def clopper_pearson(k, n, conf=0.95):
    from scipy.stats import beta

    a = 1.0 - conf
    lo = 0.0 if k == 0 else float(beta.ppf(a / 2, k, n - k + 1))
    hi = 1.0 if k == n else float(beta.ppf(1 - a / 2, k + 1, n - k))
    return lo, hi


def test_osd_only_fer_at_3db_inside_published_interval():
    print("[case] OSD-only FER at 3 dB, 200 000 codewords", flush=True)
    N, B, ebn0 = 200_000, 4096, 3.0
    sigma2 = 1.0 / (2.0 * 0.5 * 10.0 ** (ebn0 / 10.0))
    rng = np.random.default_rng(31415)
    fe = 0
    bit_errors = 0
    with sa.Decoder.osd_only(max_batch=B) as dec:
        for s0 in range(0, N, B):
            c = min(B, N - s0)
            y = 1.0 + np.sqrt(sigma2) * rng.standard_normal((c, 128))
            llr = ((2.0 / sigma2) * y).astype(np.float32)
            info, dist = dec.osd_batch(llr)
            fe += int(info.any(axis=1).sum())
            bit_errors += int(info.sum())
    pub = clopper_pearson(1, 1000)
    ours = clopper_pearson(fe, N)
    print(f"FER {fe / N:.3e} ({fe}/{N}), 95% CI [{ours[0]:.2e}, {ours[1]:.2e}]; published CI "
          f"[{pub[0]:.2e}, {pub[1]:.2e}]; BER {bit_errors / (N * 64):.3e}", flush=True)
    assert ours[0] <= pub[1] and pub[0] <= ours[1], (ours, pub)
    assert fe > 0  # enough frames to see errors at this point
