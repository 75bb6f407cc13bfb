"""The C-ABI library loads and exports every function the public headers declare (CPU only)."""
from __future__ import annotations

import glob
import os
import re
import subprocess

import sicc_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_c_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*?[\s\*]+([a-z][a-z0-9_]+)\s*\(", text, re.M):
            if m.group(1) not in ("if", "return", "sizeof"):
                names.add(m.group(1))
    return names


def test_library_loads():
    assert sicc_amd.load() is not None


def test_every_declared_function_is_exported():
    names = declared_c_functions()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", sicc_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = sorted(n for n in names if n not in exported)
    assert not missing, missing


def test_reference_abi_names_present():
    # every entry point of the reference's cuda_osd C ABI and the three extra ones the
    # canonical tree calls (uhd_ldpc_ofdm_link.cpp:3251, 3753)
    ref = ["cuda_osd_create_stream_decoder", "cuda_osd_create_stream_decoder_with_matrices",
           "cuda_osd_destroy_stream_decoder", "cuda_osd_create_osd_decoder", "cuda_osd_destroy_osd_decoder",
           "cuda_osd_create_bp_csr_decoder", "cuda_osd_destroy_bp_csr_decoder", "cuda_osd_decode_bp_csr_batch",
           "cuda_osd_decode_bp_csr_device_batch", "cuda_osd_decode_osd_batch", "cuda_osd_submit_codewords",
           "cuda_osd_submit_device_codewords", "cuda_osd_poll_ready", "cuda_osd_flush", "cuda_osd_get_stats",
           "cuda_osd_get_decoded_codewords", "cuda_osd_get_total_decode_ms", "cuda_osd_find_frame",
           "cuda_osd_get_last_error"]
    lib = sicc_amd.load()
    for n in ref:
        assert hasattr(lib, n), n


def test_qc_code_builder_without_gpu():
    # the code constructor is host-only: circulant structure, column weight 3, full-rank parity part
    import numpy as np

    from ldpc_helpers import gf2_rank

    m, rp, ci = sicc_amd.qc_code(256, 8, 3, seed=9)
    n, Z = 256, 16
    assert m == 128 and rp[-1] == 3 * n
    h = np.zeros((m, n), np.uint8)
    for r in range(m):
        h[r, ci[rp[r]:rp[r + 1]]] = 1
    assert (h.sum(axis=0) == 3).all()
    assert gf2_rank(h[:, n - m:]) == m
    for bi in range(m // Z):
        for bj in range(n // Z):
            blk = h[bi * Z:(bi + 1) * Z, bj * Z:(bj + 1) * Z]
            if blk.any():  # a circulant permutation: row r is row 0 shifted by r
                assert all(np.array_equal(np.roll(blk[0], r), blk[r]) for r in range(Z))
    # no 4-cycles: any two columns share at most one row
    overlap = h.T.astype(np.int32) @ h.astype(np.int32)
    np.fill_diagonal(overlap, 0)
    assert overlap.max() <= 1
