"""Mixed-code batched decode (concurrent per-code persistent kernels, hipGraph
replay, work stealing) against the CPU oracle and against eager execution."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import sicc_amd as sa

pytestmark = pytest.mark.gpu

CODES = [(128, 8), (256, 8), (512, 4), (1024, 5), (2048, 8)]  # (n, block rows): rates 1/2, 1/2, 3/4, 11/16, 1/2


def _setup(counts, seed):
    import torch

    rng = np.random.default_rng(seed)
    decs, codes, data = [], [], []
    for (n, mb), cnt in zip(CODES, counts):
        m, rp, ci = sa.qc_code(n, mb, 3, seed=n * 7 + mb)
        code = oracle.Code(n, m, rp, ci, front_info=True)
        info = rng.integers(0, 2, (cnt, code.k), dtype=np.uint8)
        llr = oracle.awgn_llr(code.encode(info), 2.2, code.k / n, rng)
        decs.append(sa.Decoder.bp_csr(n, m, rp, ci, max_iter=30, alpha=0.75, schedule=0))
        codes.append(code)
        d = {"llr": llr, "d_llr": torch.from_numpy(llr).cuda(),
             "d_info": torch.zeros((cnt, code.k), dtype=torch.uint8, device="cuda"),
             "d_ok": torch.zeros(cnt, dtype=torch.uint8, device="cuda"),
             "d_it": torch.zeros(cnt, dtype=torch.int32, device="cuda")}
        data.append(d)
    return decs, codes, data


def _groups(data):
    return [sa.Group(i, d["llr"].shape[0], d["d_llr"].data_ptr(), d["d_info"].data_ptr(), d["d_ok"].data_ptr(),
                     d["d_it"].data_ptr(), None) for i, d in enumerate(data)]


def test_mixed_batch_matches_oracle_and_graph_replays():
    import torch

    print("[case] mixed batch, 5 codes, eager then 3 graph replays", flush=True)
    decs, codes, data = _setup([160, 96, 64, 40, 24], 11)
    # work stealing on two members, static striding on the others
    decs[0].set_tune(1, 4, 0, 4)
    decs[4].set_tune(0, 4, 0, 2)
    with sa.MultiDecoder(decs) as md:
        md.decode(_groups(data), use_graph=False)
        torch.cuda.synchronize()
        eager = [(d["d_info"].cpu().numpy().copy(), d["d_ok"].cpu().numpy().copy(), d["d_it"].cpu().numpy().copy())
                 for d in data]
        for code, d, (info, ok, it) in zip(codes, data, eager):
            for i in range(0, d["llr"].shape[0], 3):
                o = code.nms(d["llr"][i], max_iter=30, alpha=0.75)
                assert ok[i] == o["ok"] and it[i] == o["iterations"]
                assert np.array_equal(info[i], o["hard"][code.info_cols])
        for rep in range(3):
            for d in data:  # eager work between replays: clear the outputs
                d["d_info"].zero_()
                d["d_ok"].fill_(7)
                d["d_it"].fill_(-1)
            torch.cuda.synchronize()
            md.decode(_groups(data), use_graph=True)
            torch.cuda.synchronize()
            for d, (info, ok, it) in zip(data, eager):
                assert np.array_equal(d["d_info"].cpu().numpy(), info), rep
                assert np.array_equal(d["d_ok"].cpu().numpy(), ok), rep
                assert np.array_equal(d["d_it"].cpu().numpy(), it), rep
    for dd in decs:
        dd.close()


def test_graph_recaptured_when_groups_change():
    import torch

    decs, codes, data = _setup([64, 32, 16, 8, 8], 12)
    with sa.MultiDecoder(decs) as md:
        g = _groups(data)
        md.decode(g, use_graph=True)
        md.decode(g[:2], use_graph=True)  # different group set -> new graph
        torch.cuda.synchronize()
        for code, d in zip(codes[:2], data[:2]):
            info = d["d_info"].cpu().numpy()
            for i in range(d["llr"].shape[0]):
                o = code.nms(d["llr"][i], max_iter=30, alpha=0.75)
                assert np.array_equal(info[i], o["hard"][code.info_cols])
    for dd in decs:
        dd.close()


def test_full_mixed_batch():
    # The mixed workload at full size: 65 536 frames drawn uniformly over 15 QC
    # codes (n 128..2048, 8 / 5 / 4 block rows), decoded by one graph launch:
    # replay equal to eager on every frame, and a spread sample of every code
    # bit-exact against the oracle.
    import torch

    print("[case] full mixed batch, 65536 frames over 15 codes", flush=True)
    rng = np.random.default_rng(65536)
    shapes = [(n, mb) for n in (128, 256, 512, 1024, 2048) for mb in (8, 5, 4)]
    member = rng.integers(0, len(shapes), 65536)
    decs, codes, data = [], [], []
    for j, (n, mb) in enumerate(shapes):
        cnt = int((member == j).sum())
        m, rp, ci = sa.qc_code(n, mb, 3, seed=1000 + n + mb)
        code = oracle.Code(n, m, rp, ci, front_info=True)
        info = rng.integers(0, 2, (cnt, code.k), dtype=np.uint8)
        llr = oracle.awgn_llr(code.encode(info), 3.0, code.k / n, rng)
        decs.append(sa.Decoder.bp_csr(n, m, rp, ci, max_iter=50, alpha=0.75, schedule=0))
        codes.append(code)
        data.append({"llr": llr, "d_llr": torch.from_numpy(llr).cuda(),
                     "d_info": torch.zeros((cnt, code.k), dtype=torch.uint8, device="cuda"),
                     "d_ok": torch.zeros(cnt, dtype=torch.uint8, device="cuda"),
                     "d_it": torch.zeros(cnt, dtype=torch.int32, device="cuda")})
    assert sum(d["llr"].shape[0] for d in data) == 65536
    with sa.MultiDecoder(decs) as md:
        md.decode(_groups(data), use_graph=False)
        torch.cuda.synchronize()
        eager = [tuple(d[k].cpu().numpy().copy() for k in ("d_info", "d_ok", "d_it")) for d in data]
        for d in data:
            d["d_info"].zero_()
            d["d_ok"].fill_(7)
            d["d_it"].fill_(-1)
        torch.cuda.synchronize()
        md.decode(_groups(data), use_graph=True)
        torch.cuda.synchronize()
        for d, (info, ok, it) in zip(data, eager):
            assert np.array_equal(d["d_info"].cpu().numpy(), info)
            assert np.array_equal(d["d_ok"].cpu().numpy(), ok)
            assert np.array_equal(d["d_it"].cpu().numpy(), it)
    for code, d, (info, ok, it) in zip(codes, data, eager):
        for i in range(0, d["llr"].shape[0], 97):
            o = code.nms(d["llr"][i], max_iter=50, alpha=0.75)
            assert ok[i] == o["ok"] and it[i] == o["iterations"], (code.n, i)
            assert np.array_equal(info[i], o["hard"][code.info_cols]), (code.n, i)
    for dd in decs:
        dd.close()


def test_two_groups_of_one_member_and_parameter_changes():
    """Several groups of the same BP-OSD member in one call (eager and graph)
    decode concurrently with separate scratch (OSD hand-off, counters, stealing
    cursor): each equals the oracle.  A member's OSD order changed between
    replays is picked up by the next replay (the graph is recaptured), which
    then equals the eager decode with the new order."""
    import torch

    print("[case] mixed: two groups of one member, OSD order change between replays", flush=True)
    n, m, rp, ci = oracle.alist_parse(sa.CCSDS_128_64)
    code = oracle.Code(n, m, rp, ci, True)
    rng = np.random.default_rng(13)
    counts = [3000, 1700, 2500]
    llrs = [oracle.awgn_llr(code.encode(rng.integers(0, 2, (c, code.k), dtype=np.uint8)), 2.0, 0.5, rng)
            for c in counts]
    data = [{"llr": l, "d_llr": torch.from_numpy(l).cuda(),
             "d_info": torch.zeros((l.shape[0], code.k), dtype=torch.uint8, device="cuda"),
             "d_ok": torch.zeros(l.shape[0], dtype=torch.uint8, device="cuda"),
             "d_it": torch.zeros(l.shape[0], dtype=torch.int32, device="cuda"),
             "d_used": torch.zeros(l.shape[0], dtype=torch.uint8, device="cuda")} for l in llrs]

    def groups():
        return [sa.Group(0, d["llr"].shape[0], d["d_llr"].data_ptr(), d["d_info"].data_ptr(), d["d_ok"].data_ptr(),
                         d["d_it"].data_ptr(), d["d_used"].data_ptr()) for d in data]

    def check(order):
        for d in data:
            o = code.bposd(d["llr"], max_iter=50, alpha=0.75, order=order, max_flip=64)
            assert np.array_equal(d["d_info"].cpu().numpy(), o["info"]), order
            assert np.array_equal(d["d_used"].cpu().numpy(), o["used_osd"]), order
            assert np.array_equal(d["d_it"].cpu().numpy(), o["iterations"]), order

    def clear():
        for d in data:
            d["d_info"].fill_(9)
            d["d_used"].fill_(9)
            d["d_it"].fill_(-1)
        torch.cuda.synchronize()

    with sa.Decoder.stream(sa.stream_config(max_iter=50, alpha=0.75)) as dec:
        dec.set_osd(1, 64, 4096)
        dec.set_tune(1, 4, 0, 16)  # work stealing: the cursor is per-group scratch too
        with sa.MultiDecoder([dec]) as md:
            for use_graph in (False, True, True):
                clear()
                md.decode(groups(), use_graph=use_graph)
                torch.cuda.synchronize()
                check(1)
            dec.set_osd(2, 64, 4096)
            for use_graph in (True, False):
                clear()
                md.decode(groups(), use_graph=use_graph)
                torch.cuda.synchronize()
                check(2)


def test_fused_launch_per_code_parameters():
    """Groups of codes that share one BP kernel instance decode in one fused
    launch with their own code, iteration cap, normalization and offset; a pair
    with damping shares another instance.  Batches far larger than the resident
    teams (each team strides over its own code's batch) equal every decoder's
    own decode exactly, eager and in a graph, and the oracle on a sample."""
    import torch

    print("[case] fused BP launch, per-code parameters", flush=True)
    rng = np.random.default_rng(77)
    params = [dict(max_iter=30, alpha=0.75), dict(max_iter=12, alpha=0.8), dict(max_iter=40, alpha=0.7, offset=0.1),
              dict(max_iter=25, alpha=0.75, damping=0.25), dict(max_iter=20, alpha=0.8, damping=0.5)]
    counts = [30000, 20000, 9000, 12000, 7000]
    decs, codes, data = [], [], []
    for i, (pr, cnt) in enumerate(zip(params, counts)):
        m, rp, ci = sa.qc_code(256, 4, 3, seed=900 + i)
        code = oracle.Code(256, m, rp, ci, front_info=True)
        info = rng.integers(0, 2, (cnt, code.k), dtype=np.uint8)
        llr = oracle.awgn_llr(code.encode(info), 2.0, code.k / 256, rng)
        decs.append(sa.Decoder.bp_csr(256, m, rp, ci, schedule=0, **pr))
        codes.append(code)
        data.append({"llr": llr, "d_llr": torch.from_numpy(llr).cuda(),
                     "d_info": torch.zeros((cnt, code.k), dtype=torch.uint8, device="cuda"),
                     "d_ok": torch.zeros(cnt, dtype=torch.uint8, device="cuda"),
                     "d_it": torch.zeros(cnt, dtype=torch.int32, device="cuda")})
    alone = []
    for dec, d in zip(decs, data):
        cnt = d["llr"].shape[0]
        a_info = torch.zeros((cnt, dec.k), dtype=torch.uint8, device="cuda")
        a_ok = torch.zeros(cnt, dtype=torch.uint8, device="cuda")
        a_it = torch.zeros(cnt, dtype=torch.int32, device="cuda")
        dec.decode_device(d["d_llr"].data_ptr(), cnt, d_info=C_ptr(a_info), d_ok=C_ptr(a_ok), d_iters=C_ptr(a_it))
        torch.cuda.synchronize()
        alone.append((a_info.cpu().numpy(), a_ok.cpu().numpy(), a_it.cpu().numpy()))
    with sa.MultiDecoder(decs) as md:
        for use_graph in (False, True, True):
            for d in data:
                d["d_info"].zero_()
                d["d_ok"].fill_(7)
                d["d_it"].fill_(-1)
            torch.cuda.synchronize()
            md.decode(_groups(data), use_graph=use_graph)
            torch.cuda.synchronize()
            for j, (d, (info, ok, it)) in enumerate(zip(data, alone)):
                assert np.array_equal(d["d_info"].cpu().numpy(), info), (j, use_graph)
                assert np.array_equal(d["d_ok"].cpu().numpy(), ok), (j, use_graph)
                assert np.array_equal(d["d_it"].cpu().numpy(), it), (j, use_graph)
    for code, pr, d, (info, ok, it) in zip(codes, params, data, alone):
        assert 0 < ok.sum() < ok.size
        for i in range(0, d["llr"].shape[0], 211):
            o = code.nms(d["llr"][i], **pr)
            assert ok[i] == o["ok"] and it[i] == o["iterations"], i
            assert np.array_equal(info[i], o["hard"][code.info_cols]), i
    for dd in decs:
        dd.close()


def C_ptr(t):
    import ctypes

    return ctypes.c_void_p(t.data_ptr())
