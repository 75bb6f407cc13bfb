"""ctypes binding of libsicc_amd.so -- the MI355X receive-and-decode library.

This is the binding a Python caller of the reference's cuda_osd C ABI would
write; tests and bench.py use it.  It adds no compute of
its own: every decode goes through the HIP library, and loading fails loudly
when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsicc_amd.so")
DATA_DIR = os.path.join(HERE, "data")
CCSDS_128_64 = os.path.join(DATA_DIR, "CCSDS_ldpc_n128_k64.alist")

STATUS_OK = 0
STATUS_INVALID_ARGUMENT = 1
STATUS_INTERNAL_ERROR = 2

u8p = C.POINTER(C.c_uint8)
f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int)
u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
f64p = C.POINTER(C.c_double)


class StreamConfig(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("k", C.c_int), ("m", C.c_int),
        ("max_batch_size", C.c_int), ("min_batch_size", C.c_int),
        ("max_latency_us", C.c_ulonglong), ("arrival_ewma_alpha", C.c_double),
        ("default_interarrival_us", C.c_ulonglong), ("bp_max_iterations", C.c_int),
        ("bp_normalization", C.c_float), ("bp_offset", C.c_float), ("bp_damping", C.c_float),
        ("bp_min_abs_llr_accept", C.c_float), ("osd_threads", C.c_int),
    ]


class OsdConfig(C.Structure):
    _fields_ = [("n", C.c_int), ("k", C.c_int), ("max_batch_size", C.c_int), ("osd_threads", C.c_int)]


class BpCsrConfig(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("m", C.c_int), ("max_batch_size", C.c_int), ("max_iterations", C.c_int),
        ("normalization", C.c_float), ("offset", C.c_float), ("damping", C.c_float),
        ("min_abs_llr_accept", C.c_float), ("nnz", C.c_int), ("schedule", C.c_int),
    ]


class StreamStats(C.Structure):
    _fields_ = [
        ("submitted_codewords", C.c_ulonglong), ("decoded_codewords", C.c_ulonglong),
        ("bp_success_codewords", C.c_ulonglong), ("osd_fallback_codewords", C.c_ulonglong),
        ("decoded_batches", C.c_ulonglong), ("current_pending_codewords", C.c_ulonglong),
        ("max_pending_codewords", C.c_ulonglong), ("last_submit_timestamp_us", C.c_ulonglong),
        ("ewma_interarrival_us", C.c_double), ("average_batch_size", C.c_double),
        ("total_decode_ms", C.c_double),
    ]


_lib = None


def load() -> C.CDLL:
    """Load the HIP library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    # torch ships its own HIP and HSA runtimes.  Mapping them first lets the
    # library's libamdhip64.so.7 / libhsa-runtime64.so.1 dependencies resolve to
    # those same copies; loaded the other way round the process would hold two
    # HSA runtimes and torch would find no device.
    import torch  # noqa: F401

    lib = C.CDLL(LIB_PATH)
    lib.cuda_osd_get_last_error.restype = C.c_char_p
    lib.cuda_osd_get_decoded_codewords.restype = C.c_ulonglong
    lib.cuda_osd_get_decoded_codewords.argtypes = [C.c_ulonglong]
    lib.cuda_osd_get_total_decode_ms.restype = C.c_double
    lib.cuda_osd_get_total_decode_ms.argtypes = [C.c_ulonglong]
    for name in ("cuda_osd_destroy_stream_decoder", "cuda_osd_destroy_osd_decoder",
                 "cuda_osd_destroy_bp_csr_decoder", "cuda_osd_flush"):
        getattr(lib, name).argtypes = [C.c_ulonglong]
    lib.sicc_decode_device.argtypes = [C.c_ulonglong, C.c_void_p, C.c_int] + [C.c_void_p] * 8
    lib.sicc_decode_host.argtypes = [C.c_ulonglong, f32p, C.c_int, u8p, u8p, f32p, u8p, u8p, i32p, f32p]
    lib.sicc_set_kernel_timing.argtypes = [C.c_ulonglong, C.c_int]
    lib.sicc_get_kernel_timing.argtypes = [C.c_ulonglong, f64p, f64p, i32p, i32p]
    lib.sicc_set_osd_params.argtypes = [C.c_ulonglong, C.c_int, C.c_int, C.c_int]
    lib.sicc_get_code_info.argtypes = [C.c_ulonglong, i32p, i32p, i32p, i32p]
    lib.sicc_get_code_csr.argtypes = [C.c_ulonglong, i32p, i32p]
    lib.sicc_get_generator.argtypes = [C.c_ulonglong, u8p]
    lib.sicc_encode.argtypes = [C.c_ulonglong, u8p, C.c_int, u8p]
    lib.sicc_qc_code_csr.argtypes = [C.c_int, C.c_int, C.c_int, C.c_ulonglong, i32p, i32p, i32p, i32p]
    lib.cuda_osd_submit_codewords.argtypes = [C.c_ulonglong, f32p, C.c_int, u64p, u64p, u32p]
    lib.cuda_osd_submit_device_codewords.argtypes = [C.c_ulonglong, C.c_void_p, C.c_int, u64p, u64p, u32p]
    lib.cuda_osd_poll_ready.argtypes = [C.c_ulonglong, C.c_int, u8p, u64p, u64p, u32p, u8p, i32p]
    lib.cuda_osd_get_stats.argtypes = [C.c_ulonglong, C.POINTER(StreamStats)]
    lib.cuda_osd_decode_bp_csr_batch.argtypes = [C.c_ulonglong, f32p, C.c_int, u8p, u8p, f64p]
    lib.cuda_osd_decode_bp_csr_device_batch.argtypes = [C.c_ulonglong, C.c_void_p, C.c_int, u8p, u8p, f64p]
    lib.cuda_osd_decode_osd_batch.argtypes = [C.c_ulonglong, f32p, C.c_int, u8p, f32p, f64p, f64p]
    _lib = lib
    return lib


def ptr(a: np.ndarray | None, t):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(t)


class SiccError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"status {status}: {msg}")
        self.status = status


def check(status: int):
    if status != STATUS_OK:
        raise SiccError(status, load().cuda_osd_get_last_error().decode(errors="replace"))


def stream_config(n=128, k=64, max_batch=4096, min_batch=1, max_iter=50, alpha=0.75, offset=0.0,
                  damping=0.0, min_abs_accept=0.0, latency_us=1000) -> StreamConfig:
    return StreamConfig(n=n, k=k, m=n - k, max_batch_size=max_batch, min_batch_size=min_batch,
                        max_latency_us=latency_us, arrival_ewma_alpha=0.15, default_interarrival_us=150,
                        bp_max_iterations=max_iter, bp_normalization=alpha, bp_offset=offset,
                        bp_damping=damping, bp_min_abs_llr_accept=min_abs_accept, osd_threads=256)


class Decoder:
    """Owns one decoder handle of the C ABI."""

    def __init__(self, handle: int, kind: str):
        self.lib = load()
        self.handle = handle
        self.kind = kind
        n, k, m, e = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(self.lib.sicc_get_code_info(handle, C.byref(n), C.byref(k), C.byref(m), C.byref(e)))
        self.n, self.k, self.m, self.nnz = n.value, k.value, m.value, e.value

    # -- construction --
    @classmethod
    def stream(cls, cfg: StreamConfig | None = None, alist: str | None = None, csr=None, builtin=False):
        lib = load()
        cfg = cfg or stream_config()
        h = C.c_ulonglong(0)
        if builtin:
            check(lib.cuda_osd_create_stream_decoder(C.byref(cfg), C.byref(h)))
        elif csr is not None:
            n, m, rp, ci = csr
            rp = np.ascontiguousarray(rp, dtype=np.int32)
            ci = np.ascontiguousarray(ci, dtype=np.int32)
            check(lib.sicc_create_stream_decoder_csr(C.byref(cfg), int(n), int(m), ptr(rp, i32p), ptr(ci, i32p),
                                                     C.byref(h)))
        else:
            check(lib.sicc_create_stream_decoder_alist(C.byref(cfg), (alist or CCSDS_128_64).encode(), C.byref(h)))
        return cls(h.value, "stream")

    @classmethod
    def osd_only(cls, n=128, k=64, max_batch=4096):
        lib = load()
        cfg = OsdConfig(n=n, k=k, max_batch_size=max_batch, osd_threads=256)
        h = C.c_ulonglong(0)
        check(lib.cuda_osd_create_osd_decoder(C.byref(cfg), C.byref(h)))
        return cls(h.value, "osd")

    @classmethod
    def bp_csr(cls, n, m, row_ptr, col_ind, max_iter=50, alpha=0.75, offset=0.0, damping=0.0, max_batch=4096,
               schedule=-1):
        lib = load()
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(col_ind, dtype=np.int32)
        cfg = BpCsrConfig(n=n, m=m, max_batch_size=max_batch, max_iterations=max_iter, normalization=alpha,
                          offset=offset, damping=damping, min_abs_llr_accept=0.0, nnz=int(rp[-1]),
                          schedule=schedule)
        h = C.c_ulonglong(0)
        check(lib.cuda_osd_create_bp_csr_decoder(C.byref(cfg), ptr(rp, i32p), ptr(ci, i32p), C.byref(h)))
        return cls(h.value, "bp")

    def close(self):
        if self.handle:
            fn = {"stream": self.lib.cuda_osd_destroy_stream_decoder, "osd": self.lib.cuda_osd_destroy_osd_decoder,
                  "bp": self.lib.cuda_osd_destroy_bp_csr_decoder}[self.kind]
            check(fn(self.handle))
            self.handle = 0

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- code tables --
    def csr(self):
        rp = np.zeros(self.m + 1, np.int32)
        ci = np.zeros(self.nnz, np.int32)
        check(self.lib.sicc_get_code_csr(self.handle, ptr(rp, i32p), ptr(ci, i32p)))
        return rp, ci

    def generator(self):
        g = np.zeros((self.k, self.n), np.uint8)
        check(self.lib.sicc_get_generator(self.handle, ptr(g, u8p)))
        return g

    def encode(self, info: np.ndarray) -> np.ndarray:
        info = np.ascontiguousarray(info, dtype=np.uint8).reshape(-1, self.k)
        out = np.zeros((info.shape[0], self.n), np.uint8)
        check(self.lib.sicc_encode(self.handle, ptr(info, u8p), info.shape[0], ptr(out, u8p)))
        return out

    def set_osd(self, order: int, max_flip: int = 64, max_tests: int = 4096):
        check(self.lib.sicc_set_osd_params(self.handle, order, max_flip, max_tests))

    # -- decode --
    def decode_host(self, llr: np.ndarray, want_app=False):
        llr = np.ascontiguousarray(llr, dtype=np.float32).reshape(-1, self.n)
        cw = llr.shape[0]
        out = {
            "info": np.zeros((cw, self.k), np.uint8),
            "code": np.zeros((cw, self.n), np.uint8),
            "app": np.zeros((cw, self.n), np.float32) if want_app else None,
            "bp_ok": np.zeros(cw, np.uint8),
            "used_osd": np.zeros(cw, np.uint8),
            "iterations": np.zeros(cw, np.int32),
            "distance": np.zeros(cw, np.float32),
        }
        check(self.lib.sicc_decode_host(self.handle, ptr(llr, f32p), cw, ptr(out["info"], u8p),
                                        ptr(out["code"], u8p), ptr(out["app"], f32p), ptr(out["bp_ok"], u8p),
                                        ptr(out["used_osd"], u8p), ptr(out["iterations"], i32p),
                                        ptr(out["distance"], f32p)))
        return out

    def decode_device(self, d_llr: int, cw: int, d_info=None, d_code=None, d_app=None, d_ok=None, d_used=None,
                      d_iters=None, d_dist=None, stream=None):
        check(self.lib.sicc_decode_device(self.handle, C.c_void_p(d_llr), cw, d_info, d_code, d_app, d_ok, d_used,
                                          d_iters, d_dist, stream))

    def set_tune(self, team: int = 0, teams_per_block: int = 4, max_waves: int = 0, steal_chunk: int = 0):
        self.lib.sicc_set_bp_tune.argtypes = [C.c_ulonglong, C.c_int, C.c_int, C.c_int, C.c_int]
        check(self.lib.sicc_set_bp_tune(self.handle, team, teams_per_block, max_waves, steal_chunk))

    def get_tune(self):
        t, b, w, s = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self.lib.sicc_get_bp_tune.argtypes = [C.c_ulonglong, i32p, i32p, i32p, i32p]
        check(self.lib.sicc_get_bp_tune(self.handle, C.byref(t), C.byref(b), C.byref(w), C.byref(s)))
        return {"team": t.value, "teams_per_block": b.value, "max_waves": w.value, "steal_chunk": s.value}

    def autotune(self, d_llr: int, cw: int):
        t, b, w, s, ms = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_double()
        self.lib.sicc_autotune.argtypes = [C.c_ulonglong, C.c_void_p, C.c_int, i32p, i32p, i32p, i32p, f64p]
        check(self.lib.sicc_autotune(self.handle, C.c_void_p(d_llr), cw, C.byref(t), C.byref(b), C.byref(w),
                                     C.byref(s), C.byref(ms)))
        return {"team": t.value, "teams_per_block": b.value, "max_waves": w.value, "steal_chunk": s.value,
                "bp_ms": ms.value}

    def set_timing(self, on: bool):
        check(self.lib.sicc_set_kernel_timing(self.handle, int(on)))

    def timing(self):
        bp, osd = C.c_double(), C.c_double()
        nb, no = C.c_int(), C.c_int()
        check(self.lib.sicc_get_kernel_timing(self.handle, C.byref(bp), C.byref(osd), C.byref(nb), C.byref(no)))
        return {"bp_ms": bp.value, "osd_ms": osd.value, "bp_launches": nb.value, "osd_launches": no.value}

    def bp_csr_batch(self, llr: np.ndarray):
        llr = np.ascontiguousarray(llr, dtype=np.float32).reshape(-1, self.n)
        cw = llr.shape[0]
        code = np.zeros((cw, self.n), np.uint8)
        ok = np.zeros(cw, np.uint8)
        ms = C.c_double(0)
        check(self.lib.cuda_osd_decode_bp_csr_batch(self.handle, ptr(llr, f32p), cw, ptr(code, u8p), ptr(ok, u8p),
                                                    C.byref(ms)))
        return code, ok, ms.value

    def osd_batch(self, llr: np.ndarray):
        llr = np.ascontiguousarray(llr, dtype=np.float32).reshape(-1, self.n)
        cw = llr.shape[0]
        info = np.zeros((cw, self.k), np.uint8)
        dist = np.zeros(cw, np.float32)
        ms, mbps = C.c_double(0), C.c_double(0)
        check(self.lib.cuda_osd_decode_osd_batch(self.handle, ptr(llr, f32p), cw, ptr(info, u8p), ptr(dist, f32p),
                                                 C.byref(ms), C.byref(mbps)))
        return info, dist

    def submit(self, llr: np.ndarray, frame_ids=None, block_indices=None, timestamps=None):
        llr = np.ascontiguousarray(llr, dtype=np.float32).reshape(-1, self.n)
        cw = llr.shape[0]
        ts = np.ascontiguousarray(timestamps if timestamps is not None else np.zeros(cw), dtype=np.uint64)
        ids = np.ascontiguousarray(frame_ids if frame_ids is not None else np.zeros(cw), dtype=np.uint64)
        blk = np.ascontiguousarray(block_indices if block_indices is not None else np.arange(cw), dtype=np.uint32)
        check(self.lib.cuda_osd_submit_codewords(self.handle, ptr(llr, f32p), cw, ptr(ts, u64p), ptr(ids, u64p),
                                                 ptr(blk, u32p)))

    def flush(self):
        check(self.lib.cuda_osd_flush(self.handle))

    def poll(self, max_cw: int):
        info = np.zeros((max_cw, self.k), np.uint8)
        ts = np.zeros(max_cw, np.uint64)
        ids = np.zeros(max_cw, np.uint64)
        blk = np.zeros(max_cw, np.uint32)
        used = np.zeros(max_cw, np.uint8)
        got = C.c_int(0)
        check(self.lib.cuda_osd_poll_ready(self.handle, max_cw, ptr(info, u8p), ptr(ts, u64p), ptr(ids, u64p),
                                           ptr(blk, u32p), ptr(used, u8p), C.byref(got)))
        g = got.value
        return info[:g], ids[:g], blk[:g], used[:g]

    def stats(self) -> StreamStats:
        s = StreamStats()
        check(self.lib.cuda_osd_get_stats(self.handle, C.byref(s)))
        return s


class Group(C.Structure):
    _fields_ = [("member", C.c_int), ("count", C.c_int), ("d_llr", C.c_void_p), ("d_info", C.c_void_p),
                ("d_bp_ok", C.c_void_p), ("d_iterations", C.c_void_p), ("d_used_osd", C.c_void_p)]


class MultiDecoder:
    """Mixed-code batch decoder over existing Decoder objects (sicc_multi_* ABI)."""

    def __init__(self, decoders):
        self.lib = load()
        self.decoders = list(decoders)
        hs = (C.c_ulonglong * len(self.decoders))(*[d.handle for d in self.decoders])
        h = C.c_ulonglong(0)
        self.lib.sicc_multi_create.argtypes = [C.POINTER(C.c_ulonglong), C.c_int, C.POINTER(C.c_ulonglong)]
        self.lib.sicc_multi_destroy.argtypes = [C.c_ulonglong]
        self.lib.sicc_multi_decode_device.argtypes = [C.c_ulonglong, C.POINTER(Group), C.c_int, C.c_int,
                                                      C.c_void_p]
        check(self.lib.sicc_multi_create(hs, len(self.decoders), C.byref(h)))
        self.handle = h.value

    def decode(self, groups, use_graph=False, stream=None):
        arr = (Group * len(groups))(*groups)
        check(self.lib.sicc_multi_decode_device(self.handle, arr, len(groups), int(use_graph), stream))

    def close(self):
        if self.handle:
            check(self.lib.sicc_multi_destroy(self.handle))
            self.handle = 0

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class ServerStats(C.Structure):
    _fields_ = [("jobs_submitted", C.c_ulonglong), ("jobs_completed", C.c_ulonglong),
                ("codewords_completed", C.c_ulonglong), ("launches", C.c_ulonglong),
                ("max_inflight_launches", C.c_ulonglong), ("pool_bytes", C.c_ulonglong),
                ("pool_high_water", C.c_ulonglong), ("mean_jobs_per_launch", C.c_double)]


class Server:
    """Continuous-batching decode server (sicc_server_* ABI)."""

    def __init__(self, decoders, budget_bytes=16 << 30, slots=4, max_jobs_per_launch=0, max_launch_codewords=65536):
        self.lib = load()
        lib = self.lib
        lib.sicc_server_create.argtypes = [C.POINTER(C.c_ulonglong), C.c_int, C.c_ulonglong, C.c_int, C.c_int,
                                           C.c_int, C.POINTER(C.c_ulonglong)]
        lib.sicc_server_destroy.argtypes = [C.c_ulonglong]
        lib.sicc_server_submit.argtypes = [C.c_ulonglong, C.c_ulonglong, C.c_int, C.c_void_p, C.c_int, C.c_int]
        lib.sicc_server_poll.argtypes = [C.c_ulonglong, C.c_int, C.POINTER(C.c_ulonglong), i32p]
        lib.sicc_server_fetch.argtypes = [C.c_ulonglong, C.c_ulonglong, u8p, u8p, u8p, f64p]
        lib.sicc_server_drain.argtypes = [C.c_ulonglong]
        lib.sicc_server_drive.argtypes = [C.c_ulonglong, C.c_int, i32p, i32p, C.POINTER(C.c_ulonglong), f64p, f64p,
                                          f64p]
        lib.sicc_server_stats.argtypes = [C.c_ulonglong, C.POINTER(ServerStats)]
        self.decoders = list(decoders)
        hs = (C.c_ulonglong * len(self.decoders))(*[d.handle for d in self.decoders])
        h = C.c_ulonglong(0)
        check(lib.sicc_server_create(hs, len(self.decoders), budget_bytes, slots, max_jobs_per_launch,
                                     max_launch_codewords, C.byref(h)))
        self.handle = h.value

    def submit(self, job_id, member, llr, device_ptr=None, count=None):
        if device_ptr is not None:
            check(self.lib.sicc_server_submit(self.handle, job_id, member, C.c_void_p(device_ptr), count, 1))
        else:
            llr = np.ascontiguousarray(llr, np.float32)
            check(self.lib.sicc_server_submit(self.handle, job_id, member, llr.ctypes.data_as(C.c_void_p),
                                              llr.shape[0], 0))

    def poll(self, max_jobs=1024):
        ids = (C.c_ulonglong * max_jobs)()
        n = C.c_int(0)
        check(self.lib.sicc_server_poll(self.handle, max_jobs, ids, C.byref(n)))
        return [ids[i] for i in range(n.value)]

    def fetch(self, job_id, count, k):
        info = np.zeros((count, k), np.uint8)
        ok = np.zeros(count, np.uint8)
        used = np.zeros(count, np.uint8)
        lat = C.c_double(0)
        check(self.lib.sicc_server_fetch(self.handle, job_id, ptr(info, u8p), ptr(ok, u8p), ptr(used, u8p),
                                         C.byref(lat)))
        return info, ok, used, lat.value

    def drain(self):
        check(self.lib.sicc_server_drain(self.handle))

    def set_batching(self, min_launch_codewords: int, max_wait_us: float):
        self.lib.sicc_server_set_batching.argtypes = [C.c_ulonglong, C.c_int, C.c_double]
        check(self.lib.sicc_server_set_batching(self.handle, min_launch_codewords, max_wait_us))

    def hold(self, on: bool):
        self.lib.sicc_server_hold.argtypes = [C.c_ulonglong, C.c_int]
        check(self.lib.sicc_server_hold(self.handle, int(on)))

    def drive(self, members, counts, d_ptrs, arrivals_us):
        n = len(members)
        mem = np.ascontiguousarray(members, np.int32)
        cnt = np.ascontiguousarray(counts, np.int32)
        ptrs = (C.c_ulonglong * n)(*[int(p) for p in d_ptrs])
        arr = np.ascontiguousarray(arrivals_us, np.float64)
        lat = np.zeros(n, np.float64)
        wall = C.c_double(0)
        check(self.lib.sicc_server_drive(self.handle, n, ptr(mem, i32p), ptr(cnt, i32p), ptrs, ptr(arr, f64p),
                                         ptr(lat, f64p), C.byref(wall)))
        return lat, wall.value

    def stats(self):
        s = ServerStats()
        check(self.lib.sicc_server_stats(self.handle, C.byref(s)))
        return s

    def launch_log(self, max_records=65536):
        """Recent launches (sicc_server_launch_log): structured array of steady-clock
        microsecond times (first_submit, taken, issued, done), jobs, codewords, slot."""
        rec = np.zeros(max_records, LAUNCH_RECORD_DTYPE)
        got = C.c_int(0)
        self.lib.sicc_server_launch_log.argtypes = [C.c_ulonglong, C.c_void_p, C.c_int, i32p]
        check(self.lib.sicc_server_launch_log(self.handle, rec.ctypes.data, max_records, C.byref(got)))
        return rec[:got.value]

    def close(self):
        if self.handle:
            check(self.lib.sicc_server_destroy(self.handle))
            self.handle = 0

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


LAUNCH_RECORD_DTYPE = np.dtype([("t_first_submit_us", np.uint64), ("t_taken_us", np.uint64),
                                ("t_issued_us", np.uint64), ("t_done_us", np.uint64), ("jobs", np.int32),
                                ("codewords", np.int32), ("slot", np.int32), ("lanes", np.int32)])


class OfdmConfig(C.Structure):
    _fields_ = [
        ("nfft", C.c_int), ("cp", C.c_int), ("preamble_len", C.c_int), ("pre_half_len", C.c_int),
        ("num_symbols", C.c_int), ("active_sc", C.c_int), ("bits_per_symbol", C.c_int), ("code_n", C.c_int),
        ("blocks_per_frame", C.c_int), ("max_rx_search_samples", C.c_int), ("refine_radius", C.c_int),
        ("use_pilot_channel_est", C.c_int), ("sample_rate", C.c_float), ("sync_threshold", C.c_float),
        ("demod_noise_var", C.c_float),
    ]


SYNC_DTYPE = np.dtype([("start", np.int32), ("metric", np.float32), ("cfo_hz", np.float32),
                       ("coarse_metric", np.float32), ("cfo_hz_d", np.float64)])


class OfdmReceiver:
    """Owns one OFDM receiver handle (sicc_ofdm_* C ABI)."""

    def __init__(self, phy, code_n: int, blocks_per_frame: int, max_search: int = 0, noise_var: float = 1e-3,
                 threshold: float = 0.45, refine_radius: int = -1, use_pilots: bool = True):
        self.lib = load()
        lib = self.lib
        lib.sicc_ofdm_create.argtypes = [C.POINTER(OfdmConfig), f32p, i32p, C.c_int, i32p, f32p,
                                         C.POINTER(C.c_ulonglong)]
        lib.sicc_ofdm_destroy.argtypes = [C.c_ulonglong]
        lib.sicc_ofdm_llr_per_frame.argtypes = [C.c_ulonglong]
        lib.sicc_ofdm_receive_host.argtypes = [C.c_ulonglong, f32p, C.c_int, C.c_int, C.c_void_p, f32p]
        lib.sicc_ofdm_receive_device.argtypes = [C.c_ulonglong, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                                 C.c_void_p, C.c_void_p]
        lib.sicc_ofdm_sync_device.argtypes = [C.c_ulonglong, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        lib.sicc_ofdm_demod_device.argtypes = [C.c_ulonglong, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                               C.c_void_p]
        self.cfg = OfdmConfig(nfft=phy.nfft, cp=phy.cp, preamble_len=phy.pre_len, pre_half_len=phy.pre_half_len,
                              num_symbols=phy.num_symbols, active_sc=phy.active_sc,
                              bits_per_symbol=phy.bits_per_symbol, code_n=code_n, blocks_per_frame=blocks_per_frame,
                              max_rx_search_samples=max_search, refine_radius=refine_radius,
                              use_pilot_channel_est=int(use_pilots), sample_rate=phy.fs, sync_threshold=threshold,
                              demod_noise_var=noise_var)
        pre = np.ascontiguousarray(np.asarray(phy.preamble, np.complex64).view(np.float32))
        used = np.ascontiguousarray(phy.used_shifted, np.int32)
        ps = np.ascontiguousarray(phy.pilot_symbols, np.int32)
        pg = np.ascontiguousarray(np.asarray(phy.pilot_grid, np.complex64).reshape(-1).view(np.float32))
        h = C.c_ulonglong(0)
        check(lib.sicc_ofdm_create(C.byref(self.cfg), ptr(pre, f32p), ptr(used, i32p), len(ps), ptr(ps, i32p),
                                   ptr(pg, f32p), C.byref(h)))
        self.handle = h.value
        self.llr_per_frame = lib.sicc_ofdm_llr_per_frame(self.handle)

    def close(self):
        if self.handle:
            check(self.lib.sicc_ofdm_destroy(self.handle))
            self.handle = 0

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def receive_host(self, rx: np.ndarray):
        rx = np.ascontiguousarray(np.asarray(rx, np.complex64))
        frames, span = rx.shape
        sync = np.zeros(frames, SYNC_DTYPE)
        llr = np.zeros((frames, self.llr_per_frame), np.float32)
        check(self.lib.sicc_ofdm_receive_host(self.handle, rx.view(np.float32).ctypes.data_as(f32p), frames, span,
                                              sync.ctypes.data_as(C.c_void_p), ptr(llr, f32p)))
        return sync, llr

    def receive_device(self, d_rx: int, frames: int, span: int, d_sync: int, d_llr: int, stream=None):
        check(self.lib.sicc_ofdm_receive_device(self.handle, C.c_void_p(d_rx), frames, span, C.c_void_p(d_sync),
                                                C.c_void_p(d_llr), stream))


class FrameConfig(C.Structure):
    _fields_ = [("blocks_per_frame", C.c_int), ("snr_gate_db", C.c_double)]


FRAME_RESULT_DTYPE = np.dtype([
    ("detected", np.int32), ("parity_ok", np.int32), ("frame_ok", np.int32), ("frame_id", np.uint32),
    ("bit_errors", np.int32), ("pre_fec_bit_errors", np.int32), ("iterations", np.int32), ("blocks", np.int32),
    ("bp_ok_blocks", np.int32), ("osd_blocks", np.int32), ("snr_gated", np.int32), ("reserved", np.int32),
    ("peak", np.float64), ("cfo_hz", np.float64), ("snr_db", np.float64), ("evm_rms", np.float64)])
assert FRAME_RESULT_DTYPE.itemsize == 80


class FrameDecoder:
    """Owns one frame decoder (sicc_frame_* C ABI): an OFDM receiver (or none,
    for LLR input) bound to an LDPC decoder, returning FrameDecodeResult records."""

    def __init__(self, receiver: "OfdmReceiver | None", decoder: "Decoder", blocks_per_frame: int = 0,
                 snr_gate_db: float = -1e30):
        self.lib = load()
        lib = self.lib
        vp = C.c_void_p
        lib.sicc_frame_decoder_create.argtypes = [C.c_ulonglong, C.c_ulonglong, C.POINTER(FrameConfig),
                                                  C.POINTER(C.c_ulonglong)]
        lib.sicc_frame_decoder_destroy.argtypes = [C.c_ulonglong]
        lib.sicc_frame_decoder_reserve.argtypes = [C.c_ulonglong, C.c_int]
        lib.sicc_frame_decoder_info.argtypes = [C.c_ulonglong, i32p, i32p, i32p, i32p]
        lib.sicc_frame_receive_decode_device.argtypes = [C.c_ulonglong, vp, C.c_int, C.c_int] + [vp] * 8
        lib.sicc_frame_demod_decode_device.argtypes = [C.c_ulonglong, vp, C.c_int, C.c_int] + [vp] * 8
        lib.sicc_frame_decode_llr_device.argtypes = [C.c_ulonglong, vp, C.c_int] + [vp] * 6
        lib.sicc_frame_receive_decode_host.argtypes = [C.c_ulonglong, vp, C.c_int, C.c_int, vp, vp, vp]
        lib.sicc_frame_decode_llr_host.argtypes = [C.c_ulonglong, vp, C.c_int, vp, vp, vp]
        cfg = FrameConfig(blocks_per_frame=blocks_per_frame, snr_gate_db=snr_gate_db)
        h = C.c_ulonglong(0)
        check(lib.sicc_frame_decoder_create(receiver.handle if receiver else 0, decoder.handle, C.byref(cfg),
                                            C.byref(h)))
        self.handle = h.value
        b, n, k, ib = (C.c_int() for _ in range(4))
        check(lib.sicc_frame_decoder_info(self.handle, C.byref(b), C.byref(n), C.byref(k), C.byref(ib)))
        self.blocks, self.n, self.k, self.info_bits = b.value, n.value, k.value, ib.value

    def close(self):
        if self.handle:
            check(self.lib.sicc_frame_decoder_destroy(self.handle))
            self.handle = 0

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def reserve(self, frames: int):
        check(self.lib.sicc_frame_decoder_reserve(self.handle, frames))

    def receive_decode_device(self, d_rx, frames, span, d_info, d_results, d_ref=None, d_sync=None, d_llr=None,
                              d_block_ok=None, d_block_it=None, stream=None):
        check(self.lib.sicc_frame_receive_decode_device(self.handle, d_rx, frames, span, d_ref, d_info, d_results,
                                                        d_sync, d_llr, d_block_ok, d_block_it, stream))

    def demod_decode_device(self, d_rx, frames, span, d_sync, d_info, d_results, d_ref=None, d_llr=None,
                            d_block_ok=None, d_block_it=None, stream=None):
        check(self.lib.sicc_frame_demod_decode_device(self.handle, d_rx, frames, span, d_sync, d_ref, d_info,
                                                      d_results, d_llr, d_block_ok, d_block_it, stream))

    def decode_llr_device(self, d_llr, frames, d_info, d_results, d_ref=None, d_block_ok=None, d_block_it=None,
                          stream=None):
        check(self.lib.sicc_frame_decode_llr_device(self.handle, d_llr, frames, d_ref, d_info, d_results,
                                                    d_block_ok, d_block_it, stream))

    def receive_decode_host(self, rx: np.ndarray, ref_info: np.ndarray | None = None):
        rx = np.ascontiguousarray(np.asarray(rx, np.complex64))
        frames, span = rx.shape
        info = np.zeros((frames, self.info_bits), np.uint8)
        res = np.zeros(frames, FRAME_RESULT_DTYPE)
        ref = None if ref_info is None else np.ascontiguousarray(ref_info, np.uint8)
        check(self.lib.sicc_frame_receive_decode_host(self.handle, rx.ctypes.data, frames, span,
                                                      None if ref is None else ref.ctypes.data, info.ctypes.data,
                                                      res.ctypes.data))
        return info, res

    def decode_llr_host(self, llr: np.ndarray, ref_info: np.ndarray | None = None):
        llr = np.ascontiguousarray(llr, np.float32)
        frames = llr.shape[0]
        info = np.zeros((frames, self.info_bits), np.uint8)
        res = np.zeros(frames, FRAME_RESULT_DTYPE)
        ref = None if ref_info is None else np.ascontiguousarray(ref_info, np.uint8)
        check(self.lib.sicc_frame_decode_llr_host(self.handle, llr.ctypes.data, frames,
                                                  None if ref is None else ref.ctypes.data, info.ctypes.data,
                                                  res.ctypes.data))
        return info, res


def find_frame(signal: np.ndarray, preamble: np.ndarray, half_len: int, threshold: float, fs: float):
    """cuda_osd_find_frame on the GPU -> (start_1based, peak, cfo_hz)."""
    lib = load()
    sr = np.ascontiguousarray(signal.real, np.float64)
    si = np.ascontiguousarray(signal.imag, np.float64)
    pr = np.ascontiguousarray(preamble.real, np.float64)
    pi = np.ascontiguousarray(preamble.imag, np.float64)
    s, pk, cf = C.c_int(), C.c_double(), C.c_double()
    check(lib.cuda_osd_find_frame(ptr(sr, f64p), ptr(si, f64p), len(sr), ptr(pr, f64p), ptr(pi, f64p), len(pr),
                                  half_len, C.c_double(threshold), C.c_double(fs), C.byref(s), C.byref(pk),
                                  C.byref(cf)))
    return s.value, pk.value, cf.value


def qc_code(n: int, mb: int, col_weight: int = 3, seed: int = 1):
    """CSR of the seeded QC-LDPC code built by the library."""
    lib = load()
    m, e = C.c_int(), C.c_int()
    check(lib.sicc_qc_code_csr(n, mb, col_weight, seed, C.byref(m), C.byref(e), None, None))
    rp = np.zeros(m.value + 1, np.int32)
    ci = np.zeros(e.value, np.int32)
    check(lib.sicc_qc_code_csr(n, mb, col_weight, seed, None, None, ptr(rp, i32p), ptr(ci, i32p)))
    return m.value, rp, ci


# ------------------------------------------------ C++ pipeline front ----
class PipelineConfig(C.Structure):
    _fields_ = [
        ("nfft", C.c_int), ("cp", C.c_int), ("preamble_len", C.c_int), ("pre_half_len", C.c_int),
        ("num_symbols", C.c_int), ("active_sc", C.c_int), ("bits_per_symbol", C.c_int), ("code_n", C.c_int),
        ("code_k", C.c_int), ("blocks_per_frame", C.c_int), ("max_rx_search_samples", C.c_int),
        ("sample_rate", C.c_float), ("sync_threshold", C.c_float), ("tx_amplitude", C.c_float),
        ("demod_noise_var", C.c_float), ("use_pilot_channel_est", C.c_int),
    ]


_TABLE_FIELDS = [
    ("generator_kn", C.c_void_p), ("dvb_s2_accumulator", None), ("row_info_offsets", C.c_void_p),
    ("row_info_cols", C.c_void_p), ("used_indices", C.c_void_p), ("pilot_symbol_index", C.c_void_p),
    ("data_symbol_index", C.c_void_p), ("pilot_symbols", C.c_void_p), ("data_symbols", C.c_void_p),
    ("interp_p0", C.c_void_p), ("interp_p1", C.c_void_p), ("interp_alpha", C.c_void_p),
    ("pilot_grid_iq", C.c_void_p), ("preamble_iq", C.c_void_p), ("precomp_h_iq", C.c_void_p),
    ("actual_h_iq", C.c_void_p),
]


def _table_struct_fields():
    out = []
    for name, t in _TABLE_FIELDS:
        if t is None:
            out.append((name, C.c_int))
        else:
            out.append((name, t))
            out.append((name.replace("_iq", "") + "_count", C.c_int))
    return out


class PipelineTables(C.Structure):
    _fields_ = _table_struct_fields()


class PipelineSync(C.Structure):
    _fields_ = [("start", C.c_int), ("metric", C.c_float), ("cfo_hz", C.c_float)]

    def astuple(self):
        return self.start, self.metric, self.cfo_hz


def device_to_host(ptr: int, nbytes: int) -> bytes:
    """Synchronous copy out of device memory (the HIP runtime the library links)."""
    lib = load()
    buf = C.create_string_buffer(nbytes)
    lib.hipDeviceSynchronize()
    rc = lib.hipMemcpy(buf, C.c_void_p(ptr), C.c_size_t(nbytes), 2)  # hipMemcpyDeviceToHost
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed: {rc}")
    return buf.raw


class Pipeline:
    """usrp_gpu_pipeline_bridge::Pipeline through its C layer (sicc_pipeline_*)."""

    def __init__(self, phy, code_n: int, code_k: int, blocks_per_frame: int, slots: int = 3,
                 max_rx_search_samples: int = 0, tx_amplitude: float = 0.2, noise_var: float = 1e-3,
                 threshold: float = 0.45, use_pilots: bool = True):
        self.lib = lib = load()
        lib.sicc_pipeline_create.argtypes = [C.POINTER(PipelineConfig), C.c_int, C.POINTER(C.c_ulonglong)]
        lib.sicc_pipeline_destroy.argtypes = [C.c_ulonglong]
        lib.sicc_pipeline_upload_static_tables.argtypes = [C.c_ulonglong, C.POINTER(PipelineTables)]
        lib.sicc_pipeline_submit_tx_frame.argtypes = [C.c_ulonglong, C.c_int, C.c_void_p, C.c_void_p]
        lib.sicc_pipeline_submit_tx_frame_device.argtypes = [C.c_ulonglong, C.c_int, C.c_void_p, C.c_float,
                                                             C.c_ulonglong, C.c_int, C.POINTER(C.c_void_p)]
        lib.sicc_pipeline_synchronize_tx.argtypes = [C.c_ulonglong, C.c_int]
        lib.sicc_pipeline_find_frame.argtypes = [C.c_ulonglong, C.c_int, C.c_void_p, C.c_int,
                                                 C.POINTER(PipelineSync)]
        lib.sicc_pipeline_find_frame_device.argtypes = [C.c_ulonglong, C.c_int, C.c_void_p, C.c_int,
                                                        C.POINTER(PipelineSync)]
        lib.sicc_pipeline_demod_frame_to_device_llr.argtypes = [C.c_ulonglong, C.c_int, C.c_void_p,
                                                                C.POINTER(PipelineSync), C.POINTER(C.c_void_p)]
        lib.sicc_pipeline_demod_frame_device_to_device_llr.argtypes = [C.c_ulonglong, C.c_int, C.c_void_p,
                                                                       C.POINTER(PipelineSync),
                                                                       C.POINTER(C.c_void_p)]
        lib.sicc_pipeline_demod_frame_to_host_llr.argtypes = [C.c_ulonglong, C.c_int, C.c_void_p, C.c_void_p,
                                                              C.POINTER(PipelineSync)]
        lib.sicc_pipeline_tx_batch_device.argtypes = [C.c_ulonglong, C.c_void_p, C.c_int, C.c_float,
                                                      C.c_ulonglong, C.c_void_p, C.c_void_p]
        lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.phy = phy
        self.cfg = PipelineConfig(
            nfft=phy.nfft, cp=phy.cp, preamble_len=phy.pre_len, pre_half_len=phy.pre_half_len,
            num_symbols=phy.num_symbols, active_sc=phy.active_sc, bits_per_symbol=phy.bits_per_symbol,
            code_n=code_n, code_k=code_k, blocks_per_frame=blocks_per_frame,
            max_rx_search_samples=max_rx_search_samples, sample_rate=phy.fs, sync_threshold=threshold,
            tx_amplitude=tx_amplitude, demod_noise_var=noise_var, use_pilot_channel_est=int(use_pilots))
        self.frame_samples = phy.pre_len + phy.num_symbols * (phy.nfft + phy.cp)
        self.coded_bits = blocks_per_frame * code_n
        self.info_bits = blocks_per_frame * code_k
        h = C.c_ulonglong(0)
        check(lib.sicc_pipeline_create(C.byref(self.cfg), slots, C.byref(h)))
        self.handle = h.value

    def close(self):
        if getattr(self, "handle", 0):
            self.lib.sicc_pipeline_destroy(self.handle)
            self.handle = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_static_tables(self, **t):
        """Keyword arguments named as StaticTables' fields (numpy arrays or None)."""
        conv = {
            "generator_kn": np.uint8, "row_info_offsets": np.int32, "row_info_cols": np.int32,
            "used_indices": np.int32, "pilot_symbol_index": np.int32, "data_symbol_index": np.int32,
            "pilot_symbols": np.int32, "data_symbols": np.int32, "interp_p0": np.int32, "interp_p1": np.int32,
            "interp_alpha": np.float32, "pilot_grid": np.complex64, "preamble": np.complex64,
            "precomp_h": np.complex64, "actual_h": np.complex64,
        }
        st = PipelineTables()
        self._keep = []
        for name, dt in conv.items():
            v = t.get(name)
            field = name + "_iq" if dt is np.complex64 else name
            if v is None:
                continue
            a = np.ascontiguousarray(np.asarray(v, dt).reshape(-1))
            self._keep.append(a)
            setattr(st, field, a.ctypes.data)
            setattr(st, name + "_count", a.size)
        st.dvb_s2_accumulator = int(bool(t.get("dvb_s2_accumulator", False)))
        check(self.lib.sicc_pipeline_upload_static_tables(self.handle, C.byref(st)))

    def submit_tx_frame(self, slot: int, info: np.ndarray) -> np.ndarray:
        info = np.ascontiguousarray(info, np.uint8).reshape(-1)
        out = np.zeros(self.frame_samples, np.complex64)
        check(self.lib.sicc_pipeline_submit_tx_frame(self.handle, slot, info.ctypes.data,
                                                                out.ctypes.data))
        check(self.lib.sicc_pipeline_synchronize_tx(self.handle, slot))
        return out

    def submit_tx_frame_device(self, slot: int, info: np.ndarray, noise_var: float = 0.0, seed: int = 0,
                               add_noise: bool = False) -> int:
        info = np.ascontiguousarray(info, np.uint8).reshape(-1)
        p = C.c_void_p()
        check(self.lib.sicc_pipeline_submit_tx_frame_device(self.handle, slot, info.ctypes.data,
                                                                       noise_var, seed, int(add_noise),
                                                                       C.byref(p)))
        return p.value

    def synchronize_tx(self, slot: int):
        check(self.lib.sicc_pipeline_synchronize_tx(self.handle, slot))

    def read_frame(self, d_ptr: int) -> np.ndarray:
        return np.frombuffer(device_to_host(d_ptr, 8 * self.frame_samples), np.complex64).copy()

    def read_llr(self, d_ptr: int) -> np.ndarray:
        return np.frombuffer(device_to_host(d_ptr, 4 * self.coded_bits), np.float32).copy()

    def find_frame(self, slot: int, samples: np.ndarray):
        s = np.ascontiguousarray(samples, np.complex64)
        r = PipelineSync()
        check(self.lib.sicc_pipeline_find_frame(self.handle, slot, s.ctypes.data, s.size, C.byref(r)))
        return r.astuple()

    def find_frame_device(self, slot: int, d_ptr: int, count: int):
        r = PipelineSync()
        check(self.lib.sicc_pipeline_find_frame_device(self.handle, slot, d_ptr, count, C.byref(r)))
        return r.astuple()

    def demod_frame_to_device_llr(self, slot: int, frame: np.ndarray):
        f = np.ascontiguousarray(frame, np.complex64)
        r = PipelineSync()
        p = C.c_void_p()
        check(self.lib.sicc_pipeline_demod_frame_to_device_llr(self.handle, slot, f.ctypes.data,
                                                                          C.byref(r), C.byref(p)))
        return p.value, r.astuple()

    def demod_frame_device_to_device_llr(self, slot: int, d_frame: int, sync) -> int:
        r = PipelineSync(*sync)
        p = C.c_void_p()
        check(self.lib.sicc_pipeline_demod_frame_device_to_device_llr(self.handle, slot, d_frame,
                                                                                 C.byref(r), C.byref(p)))
        return p.value

    def demod_frame_to_host_llr(self, slot: int, frame: np.ndarray):
        f = np.ascontiguousarray(frame, np.complex64)
        llr = np.zeros(self.coded_bits, np.float32)
        r = PipelineSync()
        check(self.lib.sicc_pipeline_demod_frame_to_host_llr(self.handle, slot, f.ctypes.data,
                                                                        llr.ctypes.data, C.byref(r)))
        return llr, r.astuple()

    def tx_batch_device(self, d_info: int, frames: int, d_out: int, noise_var: float = 0.0, seed: int = 0,
                        stream: int = 0):
        check(self.lib.sicc_pipeline_tx_batch_device(self.handle, d_info, frames, noise_var, seed,
                                                                d_out, stream or None))
