"""Serving driver: synthetic multi-stream sequences pushed through the
library's server (xg_server_*) with timed arrivals.

The 16-stream mix: 6 x 640x360, 6 x 1280x720, 4 x 1920x1080, 10-bit 4:2:0,
30 fps per stream with N(0, 3 ms) arrival jitter.  Each stream's frames are a
window panning over a smooth field (the generator of the 416x240 / 1080p
workloads) along a seeded integer trajectory, plus Gaussian noise taken from a
bank at a per-frame offset, so every frame of a sequence is distinct; every
frame carries the stream's seeded QTMT partition.
"""
import concurrent.futures
import ctypes
import threading
import time

import numpy as np

from . import capi, workloads as wl
from .device import blocks_to_struct, refine_cfg

MIX_16 = [(640, 360)] * 6 + [(1280, 720)] * 6 + [(1920, 1080)] * 4


class StreamSource:
    """Frame t of one synthetic stream: the field window at the trajectory's
    (dx, dy) (even, so luma and chroma move together) plus bank noise."""

    def __init__(self, width, height, seed, margin=48, noise=2.0, nbank=3, nvar=None):
        self.w, self.h, self.m = width, height, margin
        m, nm = margin, 16
        self.fields = [wl.smooth_field(height + 2 * m, width + 2 * m, seed, 0).astype(np.int16),
                       wl.smooth_field(height // 2 + m, width // 2 + m, seed, 1).astype(np.int16),
                       wl.smooth_field(height // 2 + m, width // 2 + m, seed, 2).astype(np.int16)]
        self.bank = []
        nrng = np.random.default_rng(seed * 7 + 1)      # seeded; only the content's texture
        for k in range(nbank):
            shp = [(height + 2 * nm, width + 2 * nm), (height // 2 + nm, width // 2 + nm),
                   (height // 2 + nm, width // 2 + nm)]
            self.bank.append([np.rint(nrng.normal(0.0, noise, a * b)).astype(np.int16).reshape(a, b)
                              for (a, b) in shp])
        rng = np.random.default_rng(seed)
        self.phase = rng.uniform(0, 2 * np.pi, 2)
        self.period = (int(rng.integers(60, 120)), int(rng.integers(40, 90)))
        self.nvar = nvar                # when set, frames repeat every nvar (short test sequences)
        parts = []
        for cy in range(0, height, 64):
            for cx in range(0, width, 64):
                parts += wl.qtmt_partition(cx, cy, width, height, seed * 1000 + cy * 131 + cx)
        b = np.zeros((len(parts), 8), np.int64)
        b[:, :4] = np.array(parts, np.int64)
        self.blocks = blocks_to_struct(b)

    def offset(self, t):
        a = (self.m - 8) // 2
        dx = int(np.rint(a * np.sin(2 * np.pi * t / self.period[0] + self.phase[0])))
        dy = int(np.rint(a / 2 * np.sin(2 * np.pi * t / self.period[1] + self.phase[1])))
        return 2 * dx, 2 * dy

    def frame(self, t):
        if self.nvar:
            t %= self.nvar
        dx, dy = self.offset(t)
        nb = self.bank[t % len(self.bank)]
        nx, ny = (t * 7) % 8 * 2, (t * 5) % 8 * 2
        out = []
        for p in range(3):
            s = 1 if p == 0 else 2
            w, h, m = self.w // s, self.h // s, self.m // s
            x0, y0 = m + dx // s, m + dy // s
            a = self.fields[p][y0:y0 + h, x0:x0 + w] + nb[p][ny // s:ny // s + h, nx // s:nx // s + w]
            out.append(np.ascontiguousarray(np.clip(a, 0, 1023).astype(np.uint16)))
        return out


def two_chain_cap(srcs):
    """The smallest batch cap (blocks) that splits one frame of every stream,
    taken in order, into two batches.  With every frame offered at once the
    two sessions then serve disjoint halves of the streams; a batch waits only
    for the other session's batch when they share a stream, so the two halves
    run side by side instead of one after the other."""
    n = [len(s.blocks) for s in srcs]
    return min(max(sum(n[:k]), sum(n[k:])) for k in range(1, len(n))) if len(n) > 1 else sum(n)


def server_cfg(max_blocks, budget=64 << 30, max_streams=16, search_range=16, search_lambda=1024, max_delay_us=0,
               min_batch_frames=0):
    c = capi.XgServerCfg()
    c.hbmBudget = budget
    c.maxStreams = max_streams
    c.maxBatchBlocks = max_blocks
    c.searchRange = search_range
    c.searchLambda = search_lambda
    c.maxDelayUs = max_delay_us
    c.minBatchFrames = min_batch_frames
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
        """Pageable numpy planes and outputs (xg_server_submit)."""
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

    def submit_frame(self, sid, planes, strides, planes_kind, results, outputs_kind, blocks=None, pred=None,
                     pred_strides=None, tag=0):
        """xg_server_submit_frame: planes / results / pred are addresses
        (ints) in the stated memory kinds; blocks a struct array or None."""
        f = capi.XgServerFrame()
        f.planesKind = planes_kind
        for p in range(3):
            f.planes[p] = planes[p]
            f.strides[p] = strides[p]
        if blocks is not None:
            f.blocks = blocks.ctypes.data
            f.nblocks = len(blocks)
        f.outputsKind = outputs_kind
        f.results = results
        if pred is not None:
            for p in range(3):
                f.pred[p] = pred[p]
                f.predStride[p] = pred_strides[p] if pred_strides else 0
        f.tag = tag
        return self.lib.xg_server_submit_frame(self.h, sid, ctypes.byref(f))

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


def mix_sources(mix=MIX_16, seed=5, threads=8):
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda a: StreamSource(a[1][0], a[1][1], seed * 100 + a[0]), enumerate(mix)))


class DeviceFrames:
    """Frames 0..n-1 of a source uploaded once into HBM (one tensor per plane
    per frame set), as frames that arrive from a GPU decoder would sit there."""

    def __init__(self, src, n, threads=8):
        import torch
        self.src, self.n = src, n
        w, h = src.w, src.h
        self.t = [torch.empty((n, h, w), dtype=torch.int16, device="cuda"),
                  torch.empty((n, h // 2, w // 2), dtype=torch.int16, device="cuda"),
                  torch.empty((n, h // 2, w // 2), dtype=torch.int16, device="cuda")]
        chunk = 16
        with concurrent.futures.ThreadPoolExecutor(threads) as ex:
            for c0 in range(0, n, chunk):
                frames = list(ex.map(src.frame, range(c0, min(n, c0 + chunk))))
                for p in range(3):
                    host = np.stack([f[p] for f in frames]).view(np.int16)
                    self.t[p][c0:c0 + len(frames)].copy_(torch.from_numpy(host))
        torch.cuda.synchronize()

    def planes(self, t):
        return [self.t[p][t].data_ptr() for p in range(3)]

    def strides(self):
        return [self.src.w, self.src.w // 2, self.src.w // 2]


def run_mix(frames=300, fps=30.0, jitter_ms=3.0, srcs=None, seed=5, max_delay_us=0, mode="device", dev_frames=None,
            pinned_ring=8, max_blocks=None):
    """Arrivals at `fps` per stream with N(0, jitter) timing (fps <= 0: every
    frame offered at once, bounded by a 4-frame result ring per stream);
    returns the latency/throughput report.
      mode "device": planes already in HBM (dev_frames, one DeviceFrames per
        source), results to page-locked host buffers: no staging anywhere;
      mode "pinned": planes in page-locked host memory (a ring of
        `pinned_ring` distinct frames per stream), read by DMA (PCIe-inclusive);
      mode "host": pageable numpy planes, copied on submit (xg_server_submit).
    max_blocks caps a batch's blocks (default: every stream's frame fits one
    batch); a smaller cap splits the streams over the two sessions' batches."""
    import torch
    srcs = srcs or mix_sources(seed=seed)
    max_blocks = max_blocks or sum(len(s.blocks) for s in srcs)
    srv = Server(server_cfg(max_blocks, max_delay_us=max_delay_us))
    try:
        sids = []
        for s in srcs:
            rc, sid = srv.open(s.w, s.h)
            capi.check(rc, "open_stream")
            sids.append(sid)
        ring = 4
        if mode == "host":
            results = [[np.zeros(len(s.blocks) * 4, np.int32) for _ in range(ring)] for s in srcs]
        else:
            results = [[torch.empty(len(s.blocks) * 4, dtype=torch.int32, pin_memory=True) for _ in range(ring)]
                       for s in srcs]
        pinned = None
        if mode == "pinned":
            pinned = []
            for s in srcs:
                frs = []
                for t in range(pinned_ring):
                    f = s.frame(t)
                    frs.append([torch.from_numpy(p.view(np.int16)).pin_memory() for p in f])
                pinned.append(frs)
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
        first_part = [True] * len(srcs)
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
            s = srcs[i]
            res = results[i][t % ring]
            blocks = s.blocks if first_part[i] else None
            first_part[i] = False
            if mode == "host":
                rc = srv.submit(sids[i], s.frame(t), s.blocks, res, tag=t)
            elif mode == "pinned":
                f = pinned[i][t % pinned_ring]
                rc = srv.submit_frame(sids[i], [p.data_ptr() for p in f], [s.w, s.w // 2, s.w // 2],
                                      capi.XG_MEM_HOST_PINNED, res.data_ptr(), capi.XG_MEM_HOST_PINNED, blocks, tag=t)
            else:
                d = dev_frames[i]
                rc = srv.submit_frame(sids[i], d.planes(t % d.n), d.strides(), capi.XG_MEM_DEVICE, res.data_ptr(),
                                      capi.XG_MEM_HOST_PINNED, blocks, tag=t)
            capi.check(rc, "submit")
        th.join()
        wall = time.perf_counter() - t0
        st = srv.stats()
        px = sum(s.w * s.h for s in srcs) * frames
        lat = np.array(lat)
        return {"mode": mode, "streams": len(srcs), "frames_per_stream": frames, "wall_s": round(wall, 3),
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


def main(argv=None):
    """Runs the 16-stream mix once and prints its report as one JSON line:
    python -m xin26x_amd.serve [--frames N] [--fps F] [--mode device|pinned|host]."""
    import argparse
    import json

    import torch
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--frames", type=int, default=40, help="distinct frames per stream")
    ap.add_argument("--fps", type=float, default=0.0, help="arrival rate per stream (0: every frame at once)")
    ap.add_argument("--mode", choices=("device", "pinned", "host"), default="device")
    ap.add_argument("--batch-blocks", type=int, default=0,
                    help="blocks per batch (0: two_chain_cap, half the streams per batch)")
    ap.add_argument("--one-batch", action="store_true", help="every stream's frame in one batch")
    args = ap.parse_args(argv)
    torch.cuda.set_device(0)
    srcs = mix_sources()
    dev = [DeviceFrames(s, args.frames) for s in srcs] if args.mode == "device" else None
    print(f"serving: {len(srcs)} streams, {args.frames} frames each, mode {args.mode}", flush=True)
    cap = None if args.one_batch else (args.batch_blocks or two_chain_cap(srcs))
    rep = run_mix(frames=args.frames, fps=args.fps, srcs=srcs, mode=args.mode, dev_frames=dev, max_blocks=cap)
    rep["max_batch_blocks"] = cap or sum(len(s.blocks) for s in srcs)
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
