"""Shared test setup: import paths, the `gpu` marker, and the built libraries."""
from __future__ import annotations

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sicc-baseband-proto_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _load_module(name, path):
    import importlib.util

    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _ensure_built():
    ob = _load_module("build_oracle", os.path.join(ROOT, "oracle", "build_oracle.py"))
    if not os.path.exists(ob.LIB):
        ob.build()
    if not os.path.exists(os.path.join(PKG, "lib", "libsicc_amd.so")):
        _load_module("sicc_build", os.path.join(PKG, "build.py")).build()


_ensure_built()


@pytest.fixture(scope="session")
def ccsds_path():
    return os.path.join(PKG, "data", "CCSDS_ldpc_n128_k64.alist")


def gpu_available() -> bool:
    import torch

    return torch.cuda.is_available()
