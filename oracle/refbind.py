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
