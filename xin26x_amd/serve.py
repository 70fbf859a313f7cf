"""Serving driver: synthetic multi-stream sequences pushed through the
library's server (xg_server_*) with timed arrivals.

The 16-stream mix: 6 x 640x360, 6 x 1280x720, 4 x 1920x1080, 10-bit 4:2:0,
30 fps per stream with N(0, 3 ms) arrival jitter.  Each stream's content is a
smooth field (the generator of the 416x240 / 1080p workloads) panned across
frames plus noise; every frame carries the stream's seeded QTMT partition.
"""
import ctypes
import threading
import time

import numpy as np

from . import capi, workloads as wl
from .device import blocks_to_struct, refine_cfg

MIX_16 = [(640, 360)] * 6 + [(1280, 720)] * 6 + [(1920, 1080)] * 4


class StreamSource:
    """Frames of one synthetic stream: `nvar` distinct pictures cycled."""

    def __init__(self, width, height, seed, nvar=6, noise=2.0):
        self.w, self.h = width, height
        m = 48
        fields = [wl.smooth_field(height + 2 * m, width + 2 * m, seed, 0),
                  wl.smooth_field(height // 2 + m, width // 2 + m, seed, 1),
                  wl.smooth_field(height // 2 + m, width // 2 + m, seed, 2)]
        self.frames = []
        for k in range(nvar):
            dx, dy = 2 * (k % 4) + k // 4, k % 3          # integer pan, luma samples
            dx, dy = dx - dx % 2, dy - dy % 2
            y = fields[0][m + dy:m + dy + height, m + dx:m + dx + width]
            c = [f[m // 2 + dy // 2:m // 2 + dy // 2 + height // 2, m // 2 + dx // 2:m // 2 + dx // 2 + width // 2]
                 for f in fields[1:]]
            pl = [y] + c
            pl = [np.clip(np.rint(a + wl.gaussian(seed * 7 + k, a.size, noise, stream=20 + i).reshape(a.shape)),
                          0, 1023).astype(np.uint16) for i, a in enumerate(pl)]
            self.frames.append([np.ascontiguousarray(a) for a in pl])
        parts = []
        for cy in range(0, height, 64):
            for cx in range(0, width, 64):
                parts += wl.qtmt_partition(cx, cy, width, height, seed * 1000 + cy * 131 + cx)
        b = np.zeros((len(parts), 8), np.int64)
        b[:, :4] = np.array(parts, np.int64)
        self.blocks = blocks_to_struct(b)

    def frame(self, t):
        return self.frames[t % len(self.frames)]


def server_cfg(max_blocks, budget=64 << 30, max_streams=16, search_range=16, search_lambda=1024, max_delay_us=0):
    c = capi.XgServerCfg()
    c.hbmBudget = budget
    c.maxStreams = max_streams
    c.maxBatchBlocks = max_blocks
    c.searchRange = search_range
    c.searchLambda = search_lambda
    c.maxDelayUs = max_delay_us
    c.refine = refine_cfg([tuple(int(v) for v in o) for o in wl.CAND32], wl.LAMBDA_REFINE)
    return c


class Server:
    def __init__(self, cfg):
        self.lib = capi.load()
        self.h = ctypes.c_void_p()
        capi.check(self.lib.xg_server_create(ctypes.byref(self.h), ctypes.byref(cfg)), "server_create")

    def open(self, w, h):
        sid = ctypes.c_int32()
        rc = self.lib.xg_server_open_stream(self.h, w, h, ctypes.byref(sid))
        return rc, sid.value

    def close_stream(self, sid):
        return self.lib.xg_server_close_stream(self.h, sid)

    def submit(self, sid, planes, blocks, results, pred=None, tag=0):
        pp = (ctypes.c_void_p * 3)(*[p.ctypes.data for p in planes])
        st = (ctypes.c_ssize_t * 3)(*[p.shape[1] for p in planes])
        if pred is not None:
            pd = (ctypes.c_void_p * 3)(*[p.ctypes.data for p in pred])
            ps = (ctypes.c_ssize_t * 3)(*[p.shape[1] for p in pred])
            pdp, psp = ctypes.cast(pd, ctypes.c_void_p), ctypes.cast(ps, ctypes.c_void_p)
        else:
            pdp = psp = None
        return self.lib.xg_server_submit(self.h, sid, ctypes.cast(pp, ctypes.c_void_p), ctypes.cast(st, ctypes.c_void_p),
                                         blocks.ctypes.data, len(blocks), results.ctypes.data, pdp, psp, tag)

    def poll(self, cap=64, timeout_ms=100):
        arr = (capi.XgFrameResult * cap)()
        n = self.lib.xg_server_poll(self.h, arr, cap, timeout_ms)
        if n < 0:
            raise RuntimeError(f"poll failed {n}")
        return [arr[i] for i in range(n)]

    def stats(self):
        s = capi.XgServerStats()
        self.lib.xg_server_stats_get(self.h, ctypes.byref(s))
        return s

    def destroy(self):
        self.lib.xg_server_destroy(self.h)


def mix_sources(mix=MIX_16, seed=5):
    return [StreamSource(w, h, seed * 100 + i) for i, (w, h) in enumerate(mix)]


def run_mix(frames=300, fps=30.0, jitter_ms=3.0, srcs=None, seed=5, max_delay_us=0):
    """Arrivals at `fps` per stream with N(0, jitter) timing (fps <= 0: every
    frame offered at once, bounded by a 4-frame result ring per stream);
    returns the latency/throughput report."""
    srcs = srcs or mix_sources(seed=seed)
    max_blocks = sum(len(s.blocks) for s in srcs)
    srv = Server(server_cfg(max_blocks, max_delay_us=max_delay_us))
    try:
        sids = []
        for s in srcs:
            rc, sid = srv.open(s.w, s.h)
            capi.check(rc, "open_stream")
            sids.append(sid)
        ring = 4
        results = [[np.zeros(len(s.blocks) * 4, np.int32) for _ in range(ring)] for s in srcs]
        rng = np.random.default_rng(seed)
        events = []
        for i in range(len(srcs)):
            if fps <= 0:
                events += [(0.0, i, t) for t in range(frames)]
                continue
            phase = rng.uniform(0, 1.0 / fps)
            jit = rng.normal(0, jitter_ms * 1e-3, size=frames)
            for t in range(frames):
                events.append((max(0.0, phase + t / fps + jit[t]), i, t))
        events.sort(key=lambda e: (e[0], e[2], e[1]))
        done = [0]
        lat = []
        batch_frames = []
        inflight = [[0] * len(srcs)]
        lock = threading.Condition()

        def poller():
            total = len(events)
            while done[0] < total:
                for r in srv.poll(64, 50):
                    with lock:
                        lat.append((r.doneNs - r.submitNs) / 1e6)
                        batch_frames.append(r.batchFrames)
                        done[0] += 1
                        inflight[0][r.stream] -= 1
                        lock.notify_all()

        th = threading.Thread(target=poller, daemon=True)
        t0 = time.perf_counter()
        th.start()
        for (at, i, t) in events:
            now = time.perf_counter() - t0
            if at > now:
                time.sleep(at - now)
            with lock:   # a stream never has more frames outstanding than its result ring
                while inflight[0][sids[i]] >= ring:
                    lock.wait(0.05)
                inflight[0][sids[i]] += 1
            capi.check(srv.submit(sids[i], srcs[i].frame(t), srcs[i].blocks, results[i][t % ring], tag=t), "submit")
        th.join()
        wall = time.perf_counter() - t0
        st = srv.stats()
        px = sum(s.w * s.h for s in srcs) * frames
        lat = np.array(lat)
        return {"streams": len(srcs), "frames_per_stream": frames, "wall_s": round(wall, 3),
                "mpixel_per_s": round(px / wall / 1e6, 2),
                "offered_mpixel_per_s": round(px / (frames / fps) / 1e6, 2) if fps > 0 else None,
                "frame_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                                     "p99": round(float(np.percentile(lat, 99)), 3),
                                     "max": round(float(lat.max()), 3)},
                "mean_frames_per_batch": round(float(np.mean(batch_frames)), 2),
                "batches": int(st.batches), "dropped": int(st.framesDropped),
                "hbm_bytes_in_use": int(st.bytesInUse)}
    finally:
        srv.destroy()
