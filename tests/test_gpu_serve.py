"""The serving loop at its workload: the 16-stream mix (6 x 640x360, 6 x
1280x720, 4 x 1920x1080) for 8 distinct frames each.  Frames batched across
all streams give exactly the results and predictions each stream gets alone;
sampled frames of every resolution equal the oracle's predictor derivation
(AMVP from the stream's MV field), serving search, refinement and MC on the same
pictures (reference = the stream's previous frame, edge-padded), and one
stream is followed by the oracle through its whole sequence.  Device-resident
and page-locked submissions give the pageable path's results.  Admission by
HBM budget, malformed block rejection and a stream closed mid-sequence are
exercised too."""
import hashlib

import numpy as np
import pytest

from tests.serve_common import PAD, block_array, next_field, oracle_frame, predictors
from xin26x_amd import capi
from xin26x_amd.serve import MIX_16, DeviceFrames, Server, StreamSource, mix_sources, server_cfg

pytestmark = pytest.mark.gpu
FRAMES = 8
SAMPLED = {0: (2, 6), 6: (0, 4), 12: (3, 7)}     # stream -> frames checked against the oracle


@pytest.fixture(scope="module")
def mix():
    return mix_sources(seed=9)


@pytest.fixture(scope="module")
def small():
    return [StreamSource(w, h, 900 + i) for i, (w, h) in enumerate([(256, 128), (192, 128), (128, 64)])]


def _bufs(src):
    res = np.zeros(len(src.blocks) * 4, np.int32)
    pred = [np.zeros((src.h, src.w), np.uint16), np.zeros((src.h // 2, src.w // 2), np.uint16),
            np.zeros((src.h // 2, src.w // 2), np.uint16)]
    return res, pred


def _digest(pred):
    h = hashlib.sha256()
    for p in pred:
        h.update(p.tobytes())
    return h.hexdigest()


def _drain(srv, n):
    got = []
    for _ in range(1200):
        got += srv.poll(256, 50)
        if len(got) >= n:
            break
    assert len(got) == n, f"{len(got)} of {n} frames reported"
    return got


def _run_alone(srv, src, frames=FRAMES, keep_pred=False):
    rc, sid = srv.open(src.w, src.h)
    assert rc == 0
    out = []
    for t in range(frames):
        res, pred = _bufs(src)
        assert srv.submit(sid, src.frame(t), src.blocks, res, pred, tag=t) == 0
        r = _drain(srv, 1)[0]
        assert r.status == 0 and r.tag == t and r.batchFrames == 1
        out.append((res.reshape(-1, 4).copy(), [p.copy() for p in pred] if keep_pred else None, _digest(pred)))
    assert srv.close_stream(sid) == 0
    return out


@pytest.fixture(scope="module")
def alone(mix):
    """Each stream of the mix served alone, one frame per batch."""
    srv = Server(server_cfg(sum(len(s.blocks) for s in mix)))
    try:
        out = [_run_alone(srv, s, keep_pred=k in SAMPLED) for k, s in enumerate(mix)]
    finally:
        srv.destroy()
    print("alone runs done", flush=True)
    return out


def _serve_mix(srv, mix):
    """Every frame of every stream offered at once; returns the reports and buffers."""
    sids = []
    for s in mix:
        rc, sid = srv.open(s.w, s.h)
        assert rc == 0
        sids.append(sid)
    bufs = {}
    for t in range(FRAMES):
        for k, s in enumerate(mix):
            res, pred = _bufs(s)
            bufs[(k, t)] = (res, pred)
            assert srv.submit(sids[k], s.frame(t), s.blocks, res, pred, tag=t) == 0
    got = _drain(srv, FRAMES * len(mix))
    assert all(r.status == 0 for r in got)
    return got, bufs


def _equal_alone(mix, bufs, alone):
    for k, s in enumerate(mix):
        for t in range(FRAMES):
            res, pred = bufs[(k, t)]
            assert np.array_equal(res.reshape(-1, 4), alone[k][t][0]), (k, t)
            assert _digest(pred) == alone[k][t][2], (k, t)


def test_16_stream_mix_alone_equals_mix_equals_oracle(oracle, mix, alone):
    assert [(s.w, s.h) for s in mix] == MIX_16
    n_all = sum(len(s.blocks) for s in mix)
    print(f"serving mix: {len(mix)} streams, {n_all} blocks per tick, {FRAMES} frames", flush=True)
    # every batch waits for a frame of every stream: deterministic batching
    srv = Server(server_cfg(n_all, min_batch_frames=len(mix)))
    try:
        got, bufs = _serve_mix(srv, mix)
        assert all(r.batchFrames == len(mix) for r in got), sorted({r.batchFrames for r in got})
        _equal_alone(mix, bufs, alone)
        st = srv.stats()
        assert st.framesDone == FRAMES * len(mix) and st.framesDropped == 0
        assert st.batches == FRAMES
    finally:
        srv.destroy()
    print("mix == alone", flush=True)
    # sampled frames of every resolution: the oracle from the GPU's previous frame (its MV field)
    for k, frames in SAMPLED.items():
        s = mix[k]
        blocks = block_array(s)
        for t in frames:
            field = None
            if t > 0:
                field = None
                for u in range(t):       # the field after frame t-1 is written by every earlier frame in order
                    field = next_field(blocks, alone[k][u][0][:, :2], field, s.w, s.h)
            ref = s.frame(t - 1) if t > 0 else s.frame(t)
            eres, epred, _ = oracle_frame(oracle, s, s.frame(t), ref, field)
            assert np.array_equal(alone[k][t][0], eres), (k, t)
            for p in range(3):
                assert np.array_equal(alone[k][t][1][p], epred[p]), (k, t, p)
            print(f"oracle: stream {k} ({s.w}x{s.h}) frame {t} bit-exact", flush=True)
    # the 360p stream followed by the oracle alone through its whole sequence
    s = mix[0]
    field = None
    for t in range(FRAMES):
        eres, _, field = oracle_frame(oracle, s, s.frame(t), s.frame(t - 1) if t else s.frame(t), field)
        assert np.array_equal(alone[0][t][0], eres), t
    # predictors really come from the previous frames' field: frame 3's are not all zero (frame 0 refers
    # to itself, so its sub-quarter MVs round to zero at quarter-pel AMVR; later frames carry the pan)
    f3 = None
    for u in range(3):
        f3 = next_field(block_array(s), alone[0][u][0][:, :2], f3, s.w, s.h)
    refp = np.ascontiguousarray(np.pad(s.frame(2)[0], PAD, mode="edge"))
    assert np.any(predictors(oracle, block_array(s), f3, s.frame(3)[0], refp, s.w, s.h) != 0)


def test_streams_split_over_both_sessions_run_side_by_side(mix, alone):
    """Batches capped so the 360p and 720p streams fill one batch and the
    1080p streams the other: the two sessions' batches share no stream, so
    neither waits for the other, and every frame still equals the stream
    served alone."""
    cap = max(sum(len(s.blocks) for s in mix[:12]), sum(len(s.blocks) for s in mix[12:]))
    srv = Server(server_cfg(cap))
    try:
        got, bufs = _serve_mix(srv, mix)
        assert max(r.batchFrames for r in got) < len(mix)
        _equal_alone(mix, bufs, alone)
        st = srv.stats()
        assert st.framesDone == FRAMES * len(mix) and st.framesDropped == 0
    finally:
        srv.destroy()
    print("split batches == alone", flush=True)


def test_device_and_pinned_submissions_match_pageable(small):
    import torch
    src = small[0]
    srv = Server(server_cfg(len(src.blocks)))
    try:
        want = _run_alone(srv, src, frames=4, keep_pred=True)
        dev = DeviceFrames(src, 4)
        for kind in ("device", "pinned"):
            rc, sid = srv.open(src.w, src.h)
            assert rc == 0
            for t in range(4):
                if kind == "device":
                    res = torch.zeros(len(src.blocks) * 4, dtype=torch.int32, device="cuda")
                    pred = [torch.zeros((src.h >> (p > 0), src.w >> (p > 0)), dtype=torch.int16, device="cuda")
                            for p in range(3)]
                    rc = srv.submit_frame(sid, dev.planes(t), dev.strides(), capi.XG_MEM_DEVICE, res.data_ptr(),
                                          capi.XG_MEM_DEVICE, src.blocks if t == 0 else None,
                                          [p.data_ptr() for p in pred], tag=t)
                else:
                    planes = [torch.from_numpy(p.view(np.int16)).pin_memory() for p in src.frame(t)]
                    res = torch.zeros(len(src.blocks) * 4, dtype=torch.int32, pin_memory=True)
                    pred = [torch.zeros((src.h >> (p > 0), src.w >> (p > 0)), dtype=torch.int16, pin_memory=True)
                            for p in range(3)]
                    rc = srv.submit_frame(sid, [p.data_ptr() for p in planes], [src.w, src.w // 2, src.w // 2],
                                          capi.XG_MEM_HOST_PINNED, res.data_ptr(), capi.XG_MEM_HOST_PINNED,
                                          src.blocks if t == 0 else None, [p.data_ptr() for p in pred], tag=t)
                assert rc == 0
                r = _drain(srv, 1)[0]
                assert r.status == 0 and r.tag == t
                assert np.array_equal(res.cpu().numpy().reshape(-1, 4), want[t][0]), (kind, t)
                for p in range(3):
                    assert np.array_equal(pred[p].cpu().numpy().view(np.uint16), want[t][1][p]), (kind, t, p)
            assert srv.close_stream(sid) == 0
    finally:
        srv.destroy()


def test_reused_partition_needs_a_first_one(small):
    lib = capi.load()
    src = small[2]
    srv = Server(server_cfg(4096))
    try:
        rc, sid = srv.open(src.w, src.h)
        assert rc == 0
        res, _ = _bufs(src)
        f = src.frame(0)
        rc = srv.submit_frame(sid, [p.ctypes.data for p in f], [src.w, src.w // 2, src.w // 2], capi.XG_MEM_HOST,
                              res.ctypes.data, capi.XG_MEM_HOST, None)
        assert rc == capi.XG_ERR_STATE
        assert srv.stats().framesDone == 0
    finally:
        lib.xg_clear_error()
        srv.destroy()


def test_admission_by_budget(small):
    lib = capi.load()
    s0 = small[0]
    probe = Server(server_cfg(4096))
    base = probe.stats().bytesInUse
    rc, sid = probe.open(s0.w, s0.h)
    assert rc == 0
    per_stream = probe.stats().bytesInUse - base
    probe.destroy()
    srv = Server(server_cfg(4096, budget=base + per_stream + per_stream // 2))
    try:
        rc, a = srv.open(s0.w, s0.h)
        assert rc == 0
        rc, b = srv.open(s0.w, s0.h)
        assert rc == capi.XG_ERR_BUDGET
        assert srv.close_stream(a) == 0
        rc, c = srv.open(s0.w, s0.h)
        assert rc == 0
        res, pred = _bufs(s0)
        assert srv.submit(c, s0.frame(0), s0.blocks, res, pred) == 0
        assert _drain(srv, 1)[0].status == 0
    finally:
        lib.xg_clear_error()
        srv.destroy()


def test_malformed_blocks_rejected(small):
    lib = capi.load()
    s0 = small[0]
    srv = Server(server_cfg(4096))
    try:
        rc, sid = srv.open(s0.w, s0.h)
        assert rc == 0
        res, _ = _bufs(s0)
        bad = s0.blocks.copy()
        bad["w"][0] = 12
        assert srv.submit(sid, s0.frame(0), bad, res) == capi.XG_ERR_ARG
        bad = s0.blocks.copy()
        bad["x"][0] = s0.w - 4
        bad["w"][0] = 8
        assert srv.submit(sid, s0.frame(0), bad, res) == capi.XG_ERR_ARG
        assert srv.submit(sid + 5, s0.frame(0), s0.blocks, res) == capi.XG_ERR_STATE
        assert srv.stats().framesDone == 0
    finally:
        lib.xg_clear_error()
        srv.destroy()


def test_close_mid_sequence_leaves_others_intact(small):
    lib = capi.load()
    a_src, b_src = small[0], small[1]
    srv = Server(server_cfg(sum(len(s.blocks) for s in small)))
    try:
        b_alone = _run_alone(srv, b_src, frames=4)
        rc, a = srv.open(a_src.w, a_src.h)
        rc2, b = srv.open(b_src.w, b_src.h)
        assert rc == 0 and rc2 == 0
        keep = []
        for t in range(4):
            ra, _ = _bufs(a_src)
            rb, pb = _bufs(b_src)
            keep.append((ra, rb, pb))
            assert srv.submit(a, a_src.frame(t), a_src.blocks, ra, tag=100 + t) == 0
            assert srv.submit(b, b_src.frame(t), b_src.blocks, rb, pb, tag=t) == 0
        assert srv.close_stream(a) == 0                # drops what has not started
        got = _drain(srv, 8)
        for r in got:
            if r.stream == b:
                assert r.status == 0
                assert np.array_equal(keep[r.tag][1].reshape(-1, 4), b_alone[r.tag][0])
                assert _digest(keep[r.tag][2]) == b_alone[r.tag][2]
            else:
                assert r.status in (0, capi.XG_ERR_STATE)
        assert srv.submit(a, a_src.frame(0), a_src.blocks, keep[0][0]) == capi.XG_ERR_STATE
    finally:
        lib.xg_clear_error()
        srv.destroy()
