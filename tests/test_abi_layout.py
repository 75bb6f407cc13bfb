"""Struct layouts of the decoder C ABI equal the reference header's (CPU only).

The same probe program is compiled twice with gcc, once against the reference's
own cuda_osd header
(source/cpp_cuda/portable/usrp_video_link_portable_20260513/third_party/cuda_osd/include/matlab_decoder_c_api.h,
read where it lies in the reference checkout) and once against
include/matlab_decoder_c_api.h; sizeof and offsetof of every field of the
three config structs and the stats struct must agree.  The one difference is
by design: cuda_osd_bp_csr_config_t carries the trailing `schedule` int the
canonical tree sets (source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp:3658),
so it is one int longer with every shared field at the same offset.  Skipped
when the reference checkout is absent (the GPU hosts).
"""
from __future__ import annotations

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_HDR_DIR = os.path.join(os.environ.get("SICC_REFERENCE", "/root/reference"), "source", "cpp_cuda", "portable",
                           "usrp_video_link_portable_20260513", "third_party", "cuda_osd", "include")

FIELDS = {
    "cuda_osd_stream_config_t": ["n", "k", "m", "max_batch_size", "min_batch_size", "max_latency_us",
                                 "arrival_ewma_alpha", "default_interarrival_us", "bp_max_iterations",
                                 "bp_normalization", "bp_offset", "bp_damping", "bp_min_abs_llr_accept",
                                 "osd_threads"],
    "cuda_osd_osd_config_t": ["n", "k", "max_batch_size", "osd_threads"],
    "cuda_osd_bp_csr_config_t": ["n", "m", "max_batch_size", "max_iterations", "normalization", "offset",
                                 "damping", "min_abs_llr_accept", "nnz"],
    "cuda_osd_stream_stats_t": ["submitted_codewords", "decoded_codewords", "bp_success_codewords",
                                "osd_fallback_codewords", "decoded_batches", "current_pending_codewords",
                                "max_pending_codewords", "last_submit_timestamp_us", "ewma_interarrival_us",
                                "average_batch_size", "total_decode_ms"],
}


def _probe_source():
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "matlab_decoder_c_api.h"', "int main(void) {"]
    for s, fields in FIELDS.items():
        lines.append(f'printf("{s} sizeof %zu\\n", sizeof({s}));')
        for f in fields:
            lines.append(f'printf("{s} {f} %zu %zu\\n", offsetof({s}, {f}), sizeof((({s}*)0)->{f}));')
    lines.append('printf("status %d %d %d\\n", CUDA_OSD_STATUS_OK, CUDA_OSD_STATUS_INVALID_ARGUMENT, '
                 'CUDA_OSD_STATUS_INTERNAL_ERROR);')
    lines.append("return 0; }")
    return "\n".join(lines) + "\n"


def _layout(include_dir, tmp_path, tag):
    src = tmp_path / f"probe_{tag}.c"
    exe = tmp_path / f"probe_{tag}"
    src.write_text(_probe_source())
    subprocess.run(["gcc", "-std=c11", "-I", include_dir, "-o", str(exe), str(src)], check=True,
                   capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for line in out.splitlines():
        p = line.split()
        if p[0] == "status":
            got["status"] = p[1:]
        elif p[1] == "sizeof":
            got[f"{p[0]} sizeof"] = p[2]
        else:
            got[f"{p[0]} {p[1]}"] = p[2:]  # offset, size
    return got


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_HDR_DIR, "matlab_decoder_c_api.h")),
                    reason="reference checkout absent")
def test_struct_layouts_match_reference_header(tmp_path):
    ref = _layout(REF_HDR_DIR, tmp_path, "ref")
    ours = _layout(os.path.join(ROOT, "include"), tmp_path, "ours")
    for key, val in ref.items():
        if key == "cuda_osd_bp_csr_config_t sizeof":
            continue
        assert ours[key] == val, (key, ours[key], val)
    # the documented trailing extension: one int after nnz
    assert int(ours["cuda_osd_bp_csr_config_t sizeof"]) == int(ref["cuda_osd_bp_csr_config_t sizeof"]) + 4
    assert int(ref["cuda_osd_bp_csr_config_t nnz"][0]) + 4 == int(ref["cuda_osd_bp_csr_config_t sizeof"])
