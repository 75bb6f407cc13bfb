"""Build oracle/_ref/libref.so from the reference's own C++ CPU sources.

TEST INFRASTRUCTURE ONLY.  Compiles, with g++ and nothing else, the reference
translation units of the CPU receive/decode path where they lie in the
reference checkout (default /root/reference, override with SICC_REFERENCE):
    source/cpp_cuda/uhd_cpp/src/uhd_ldpc_ofdm_link.cpp   (via ref_harness.cpp)
    source/cpp_cuda/uhd_cpp/src/matrix_loader.cpp
    source/cpp_cuda/uhd_cpp/src/predictive_amc.cpp
    source/cpp_cuda/uhd_cpp/src/rx_decoder_router.cpp
    source/cpp_cuda/uhd_cpp/src/fec_config.cpp
No reference file is copied; the output goes to oracle/_ref/ only.  When the
checkout is absent (the GPU host) the prebuilt library, if any, is used as is.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_ref")
LIB = os.path.join(OUT_DIR, "libref.so")
REF = os.environ.get("SICC_REFERENCE", "/root/reference")
SRC = os.path.join(REF, "source", "cpp_cuda", "uhd_cpp", "src")
UNITS = ["matrix_loader.cpp", "predictive_amc.cpp", "rx_decoder_router.cpp", "fec_config.cpp"]


def available() -> bool:
    return os.path.exists(os.path.join(SRC, "uhd_ldpc_ofdm_link.cpp"))


def build() -> str | None:
    if not available():
        return LIB if os.path.exists(LIB) else None
    os.makedirs(OUT_DIR, exist_ok=True)
    flags = ["-std=c++17", "-O2", "-fPIC", "-w", f"-I{SRC}"]
    objs = []
    link_cpp = os.path.join(SRC, "uhd_ldpc_ofdm_link.cpp")
    jobs = [(os.path.join(HERE, "ref_harness.cpp"), [f'-DSICC_REF_LINK_CPP="{link_cpp}"'])]
    jobs += [(os.path.join(SRC, u), []) for u in UNITS]
    procs = []
    for src, extra in jobs:
        obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
        objs.append(obj)
        procs.append((src, subprocess.Popen(["g++", *flags, *extra, "-c", src, "-o", obj],
                                            stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    for src, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"reference build failed on {src}:\n{err}")
    tmp = LIB + ".tmp"
    r = subprocess.run(["g++", "-shared", "-o", tmp, *objs, "-lpthread"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"reference link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build())
    sys.exit(0)
