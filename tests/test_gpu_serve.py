"""Continuous-batching server: results alone == results in the mix (bit-exact),
admission bounded by the memory budget, plain and batched modes, native driver."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import sicc_amd as sa
from ldpc_helpers import gf2_rank

pytestmark = pytest.mark.gpu


def _members():
    decs, codes = [], []
    st = sa.Decoder.stream(sa.stream_config(max_iter=30, alpha=0.75))
    st.set_osd(1, 64, 4096)
    n_, m_, rp, ci = oracle.alist_parse(sa.CCSDS_128_64)
    decs.append(st)
    codes.append(oracle.Code(n_, m_, rp, ci, True))
    for n, mb in [(256, 8), (1024, 4), (2048, 8)]:
        m, rp, ci = sa.qc_code(n, mb, 3, seed=n + mb)
        decs.append(sa.Decoder.bp_csr(n, m, rp, ci, max_iter=30, alpha=0.75, schedule=0))
        codes.append(oracle.Code(n, m, rp, ci, True))
    return decs, codes


def _jobs(codes, n_jobs, seed):
    rng = np.random.default_rng(seed)
    jobs = []
    for j in range(n_jobs):
        mem = int(rng.integers(0, len(codes)))
        cnt = int(rng.integers(1, 40))
        c = codes[mem]
        info = rng.integers(0, 2, (cnt, c.k), dtype=np.uint8)
        jobs.append((mem, oracle.awgn_llr(c.encode(info), 2.0, c.k / c.n, rng)))
    return jobs


def _collect(srv, jobs, ids, codes):
    out = {}
    for i, (mem, llr) in zip(ids, jobs):
        out[i] = srv.fetch(i, llr.shape[0], codes[mem].k)
    return out


def test_alone_equals_in_mix():
    print("[case] server: alone vs mixed, 60 jobs over 4 codes", flush=True)
    decs, codes = _members()
    jobs = _jobs(codes, 60, 3)
    with sa.Server(decs, budget_bytes=256 << 20, slots=4) as srv:
        alone = {}
        for j, (mem, llr) in enumerate(jobs):
            srv.submit(1000 + j, mem, llr)
            srv.drain()
            alone[j] = srv.fetch(1000 + j, llr.shape[0], codes[mem].k)
        srv.hold(True)  # queue the burst, then release it: packed, overlapped launches
        for j, (mem, llr) in enumerate(jobs):
            srv.submit(5000 + j, mem, llr)
        with pytest.raises(sa.SiccError, match="held"):
            srv.drain()
        srv.hold(False)
        srv.drain()
        done = srv.poll(4096)
        assert set(range(5000, 5060)) <= set(done)
        for j, (mem, llr) in enumerate(jobs):
            info, ok, used, lat = srv.fetch(5000 + j, llr.shape[0], codes[mem].k)
            a = alone[j]
            assert np.array_equal(info, a[0]) and np.array_equal(ok, a[1]) and np.array_equal(used, a[2]), j
            assert lat > 0
        st = srv.stats()
        assert st.jobs_completed == 120 and st.mean_jobs_per_launch > 1.0
    # and the oracle agrees on BP-only members (the stream member is covered by the BP-OSD tests)
    for j, (mem, llr) in enumerate(jobs[:20]):
        if mem == 0:
            continue
        for i in range(llr.shape[0]):
            o = codes[mem].nms(llr[i], max_iter=30, alpha=0.75)
            assert np.array_equal(alone[j][0][i], o["hard"][codes[mem].info_cols])
    for d in decs:
        d.close()


def test_full_spec_mix_equals_each_decoder_alone():
    """The served workload's shape: the 15-code QC mix (n 128..2048, three
    rates each), jobs of 1..512 codewords, a 16 GB budget, four slots and
    65 536-codeword launches (fused code groups, launches spread over idle
    slots).  250 jobs released as one burst: every job's bits, success and
    OSD flags equal the member decoder's own decode of that job, and a sample
    equals the oracle."""
    print("[case] server: full-spec 15-code mix, 250 jobs of 1..512 codewords", flush=True)
    rng = np.random.default_rng(515)
    shapes = [(n, mb) for n in (128, 256, 512, 1024, 2048) for mb in (8, 5, 4)]
    decs, codes = [], []
    for n, mb in shapes:
        m, rp, ci = sa.qc_code(n, mb, 3, seed=1000 + n + mb)
        decs.append(sa.Decoder.bp_csr(n, m, rp, ci, max_iter=50, alpha=0.75, schedule=0))
        codes.append(oracle.Code(n, m, rp, ci, True))
    jobs = []
    for j in range(250):
        mem = int(rng.integers(0, len(shapes)))
        cnt = int(rng.integers(1, 513))
        c = codes[mem]
        info = rng.integers(0, 2, (cnt, c.k), dtype=np.uint8)
        jobs.append((mem, oracle.awgn_llr(c.encode(info), 2.6, c.k / c.n, rng)))
    with sa.Server(decs, budget_bytes=16 << 30, slots=4, max_launch_codewords=1 << 16) as srv:
        srv.hold(True)
        for j, (mem, llr) in enumerate(jobs):
            srv.submit(j, mem, llr)
        srv.hold(False)
        srv.drain()
        st = srv.stats()
        assert st.jobs_completed == len(jobs) and st.launches < len(jobs)
        for j, (mem, llr) in enumerate(jobs):
            info, ok, used, _ = srv.fetch(j, llr.shape[0], codes[mem].k)
            ref = decs[mem].decode_host(llr)
            assert np.array_equal(info, ref["info"]), j
            assert np.array_equal(ok.astype(bool), ref["bp_ok"].astype(bool)), j
            assert not used.any(), j
            if j < 12:
                for i in range(0, llr.shape[0], 37):
                    o = codes[mem].nms(llr[i], max_iter=50, alpha=0.75)
                    assert np.array_equal(info[i], o["hard"][codes[mem].info_cols]), (j, i)
                    assert bool(ok[i]) == o["ok"], (j, i)
    for d in decs:
        d.close()


def test_budget_admission_and_plain_mode():
    decs, codes = _members()
    jobs = _jobs(codes, 30, 4)
    # a tight budget forces in-order admission behind earlier jobs; plain mode = 1 job per launch
    with sa.Server(decs, budget_bytes=2 << 20, slots=1, max_jobs_per_launch=1) as srv:
        for j, (mem, llr) in enumerate(jobs):
            srv.submit(j, mem, llr)
        srv.drain()
        st = srv.stats()
        assert st.launches == 30 and st.max_inflight_launches == 1
        assert st.pool_high_water <= 2 << 20
        for j, (mem, llr) in enumerate(jobs):
            srv.fetch(j, llr.shape[0], codes[mem].k)
        big = np.zeros((512, 2048), np.float32)
        with pytest.raises(sa.SiccError) as e:
            srv.submit(999, 3, big)
        assert e.value.status == sa.STATUS_INVALID_ARGUMENT
    for d in decs:
        d.close()


def test_native_driver_latencies():
    import torch

    decs, codes = _members()
    jobs = _jobs(codes, 80, 5)
    dev = [torch.from_numpy(llr).cuda() for _, llr in jobs]
    arrivals = np.cumsum(np.random.default_rng(6).exponential(50.0, len(jobs)))
    with sa.Server(decs, budget_bytes=256 << 20, slots=4) as srv:
        lat, wall = srv.drive([m for m, _ in jobs], [l.shape[0] for _, l in jobs], [t.data_ptr() for t in dev],
                              arrivals)
        assert (lat > 0).all() and wall > 0
        assert srv.stats().jobs_completed == 80
        # batching policy: launches wait for 4096 resident codewords or 300 us while one is in flight
        srv.set_batching(4096, 300.0)
        lat2, _ = srv.drive([m for m, _ in jobs], [l.shape[0] for _, l in jobs], [t.data_ptr() for t in dev],
                            np.zeros(len(jobs)))
        st = srv.stats()
        assert st.jobs_completed == 160 and (lat2 > 0).all()
        with pytest.raises(sa.SiccError):
            srv.set_batching(-1, 0.0)
    for d in decs:
        d.close()


def test_duplicate_job_id_rejected():
    """A job id still known to the server is rejected; the first job keeps its
    results and the id is free again once they are fetched."""
    decs, codes = _members()
    jobs = _jobs(codes, 2, 9)
    with sa.Server(decs, budget_bytes=256 << 20, slots=2) as srv:
        mem, llr = jobs[0]
        srv.submit(7, mem, llr)
        with pytest.raises(sa.SiccError, match="already in use"):
            srv.submit(7, jobs[1][0], jobs[1][1])
        srv.drain()
        first = srv.fetch(7, llr.shape[0], codes[mem].k)
        srv.submit(7, mem, llr)
        srv.drain()
        again = srv.fetch(7, llr.shape[0], codes[mem].k)
        for a, b in zip(first, again):
            if isinstance(a, np.ndarray):
                np.testing.assert_array_equal(a, b)
        assert srv.stats().jobs_completed == 2
    for d in decs:
        d.close()


def test_ragged_info_width_round_trip():
    """k = 73 (neither a multiple of 8 nor of 64): results travel bit-packed from
    the GPU and are unpacked by fetch; every codeword equals the oracle."""
    print("[case] server: k=73 code, packed results", flush=True)
    rng = np.random.default_rng(11)
    n, m = 140, 67
    while True:  # column weight 3, row degree <= 16, last m columns invertible
        h = np.zeros((m, n), np.uint8)
        for c in range(n):
            h[rng.choice(m, 3, replace=False), c] = 1
        if h.sum(axis=1).max() <= 16 and h.sum(axis=1).min() >= 2 and gf2_rank(h[:, n - m:]) == m:
            break
    rp = np.concatenate([[0], np.cumsum(h.sum(axis=1))]).astype(np.int32)
    ci = np.concatenate([np.nonzero(r)[0] for r in h]).astype(np.int32)
    code = oracle.Code(n, m, rp, ci, True)
    assert code.k == 73
    dec = sa.Decoder.bp_csr(n, m, rp, ci, max_iter=30, alpha=0.75, schedule=0)
    jobs = []
    for _ in range(12):
        info = rng.integers(0, 2, (int(rng.integers(1, 40)), code.k), dtype=np.uint8)
        jobs.append(oracle.awgn_llr(code.encode(info), 2.5, code.k / n, rng))
    with sa.Server([dec], budget_bytes=64 << 20, slots=2) as srv:
        srv.hold(True)  # one packed launch of all twelve jobs
        for j, llr in enumerate(jobs):
            srv.submit(j, 0, llr)
        srv.hold(False)
        srv.drain()
        for j, llr in enumerate(jobs):
            info, ok, used, _ = srv.fetch(j, llr.shape[0], code.k)
            for i in range(llr.shape[0]):
                o = code.nms(llr[i], max_iter=30, alpha=0.75)
                assert np.array_equal(info[i], o["hard"][code.info_cols]), (j, i)
                assert bool(ok[i]) == o["ok"] and not used[i], (j, i)
    dec.close()


def test_exit_with_servers_open():
    """A process that exits with a server still open (jobs queued, launches in
    flight, handle never destroyed) exits cleanly: the library stops every open
    server's scheduler thread before the HIP runtime tears down."""
    import os
    import subprocess
    import sys

    print("[case] server: process exit with open servers", flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = (
        "import sys; sys.path[:0] = [%r, %r]\n"
        "import numpy as np, oracle, sicc_amd as sa\n"
        "n, m, rp, ci = oracle.alist_parse(sa.CCSDS_128_64)\n"
        "dec = sa.Decoder.bp_csr(n, m, rp, ci, max_iter=30, alpha=0.75, schedule=0)\n"
        "srv = sa.Server([dec], budget_bytes=64 << 20, slots=2, max_launch_codewords=4096)\n"
        "srv2 = sa.Server([dec], budget_bytes=64 << 20, slots=1, max_launch_codewords=4096)\n"
        "rng = np.random.default_rng(5)\n"
        "for j in range(200):\n"
        "    srv.submit(j, 0, rng.standard_normal((64, 128)).astype(np.float32) + 2.0)\n"
        "srv2.submit(0, 0, np.ones((8, 128), np.float32))\n"
        "print('submitted', flush=True)\n"
    ) % (os.path.join(root, "sicc-baseband-proto_amd"), os.path.join(root, "oracle"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "submitted" in r.stdout


def test_failed_create_is_clean():
    """A budget the device cannot hold fails server creation with an error status
    (everything allocated before the failure is released), the failure does not
    stay pending in the HIP runtime (torch's next launch check passes), and the
    next server works."""
    import torch

    print("[case] server: create with an unallocatable budget", flush=True)
    decs, codes = _members()
    with pytest.raises(sa.SiccError):
        sa.Server(decs[:1], budget_bytes=1 << 52, slots=2)
    x = torch.full((1024,), 3, dtype=torch.int32, device="cuda:0")
    assert int(x.sum().item()) == 3072
    jobs = _jobs(codes[:1], 3, 8)
    with sa.Server(decs[:1], budget_bytes=64 << 20, slots=2) as srv:
        for j, (mem, llr) in enumerate(jobs):
            srv.submit(j, mem, llr)
        srv.drain()
        for j, (mem, llr) in enumerate(jobs):
            info, ok, used, _ = srv.fetch(j, llr.shape[0], codes[mem].k)
            assert info.shape == (llr.shape[0], codes[mem].k)
    for d in decs:
        d.close()


def test_many_code_groups_in_one_launch():
    """70 member codes, one job each, released together: one launch carries 70
    code groups (more than the packing kernel stages in LDS); every codeword
    equals the oracle."""
    print("[case] server: 70 code groups in one launch", flush=True)
    rng = np.random.default_rng(21)
    decs, codes, jobs = [], [], []
    for i in range(70):
        m, rp, ci = sa.qc_code(128, 4, 3, seed=500 + i)
        decs.append(sa.Decoder.bp_csr(128, m, rp, ci, max_iter=20, alpha=0.75, schedule=0))
        code = oracle.Code(128, m, rp, ci, True)
        codes.append(code)
        info = rng.integers(0, 2, (3, code.k), dtype=np.uint8)
        jobs.append(oracle.awgn_llr(code.encode(info), 3.0, code.k / 128, rng))
    with sa.Server(decs, budget_bytes=64 << 20, slots=1, max_launch_codewords=4096) as srv:
        srv.hold(True)
        for j, llr in enumerate(jobs):
            srv.submit(j, j, llr)
        srv.hold(False)
        srv.drain()
        assert srv.stats().launches == 1
        for j, llr in enumerate(jobs):
            info, ok, used, _ = srv.fetch(j, llr.shape[0], codes[j].k)
            for i in range(llr.shape[0]):
                o = codes[j].nms(llr[i], max_iter=20, alpha=0.75)
                assert np.array_equal(info[i], o["hard"][codes[j].info_cols]), (j, i)
                assert bool(ok[i]) == o["ok"], (j, i)
    for d in decs:
        d.close()
