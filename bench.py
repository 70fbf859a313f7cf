#!/usr/bin/env python3
"""ME+MC throughput of the batched fractional-refinement path at 1080p.

Workload (one step = one frame): 1920x1080 10-bit 4:2:0, seeded smooth-field
pictures, per-64x64-CTU motion uniform in +/-32 px at 1/16 pel, seeded QTMT
partition of every CTU (4x4..64x64 incl. 1:2/1:4), 32 fractional candidates
per block: 8-tap luma + 4-tap chroma interpolation, SATD (8x8 / 4x4 / 2x2
Hadamard), (SATD<<8) + bits*lambda; then motion compensation of each block's
best candidate into the prediction picture (luma + chroma).

Reported: batched form (persistent kernels replayed from a HIP graph) as the
headline value, the plain form (one launch per block shape, one workgroup
per block) beside it and, for reference, the plain form with the tuned
candidates-per-workgroup table, the
per-frame latency p50/p99, the refinement kernel against the HBM roofline
(kernel time from HIP events on the session stream), and the CPU oracle
timed on a bounded sample on this host.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5]
"""
import argparse
import concurrent.futures
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

TUNED = os.path.join(ROOT, "xin26x_amd", "tuned", "refine_1080p_cand32.txt")
PMC_REFINE = os.path.join(ROOT, "profiles", "k_refine_1080p.md")
PMC_FDCT = os.path.join(ROOT, "profiles", "k_fdct_mfma_200k_blocks.md")


def valu_roofline(path, kernel_ms):
    """The bound that binds an integer-ALU kernel: its wave-level VALU
    instructions per launch (SQ_INSTS_VALU of the committed profile of the same
    workload) over the live kernel time, against the VALU issue peak."""
    insts = profiled_counter(path, "SQ_INSTS_VALU")
    if not insts or kernel_ms <= 0:
        return None
    achieved = insts / (kernel_ms * 1e-3) / 1e9
    return {"bound": "valu", "achieved": round(achieved, 1), "peak": VALU_PEAK_G, "unit": "G wave-instr/s",
            "frac": round(achieved / VALU_PEAK_G, 4), "insts_per_launch": insts,
            "insts_source": os.path.relpath(path, ROOT)}


def profiled_counter(path, name):
    """Per-launch value of an SQ counter from a committed rocprofv3 summary's
    counter table (`| NAME | 1,234 |`), or None."""
    try:
        with open(path) as f:
            for ln in f:
                cells = [c.strip() for c in ln.strip().strip("|").split("|")]
                if len(cells) == 2 and cells[0] == name:
                    return int(cells[1].replace(",", ""))
    except (OSError, ValueError):
        pass
    return None


VALU_PEAK_G = 256 * 2.4        # G wave-level VALU instructions/s: 256 CUs x 2.4 GHz x one per CU per cycle


def profiled_traffic(path):
    """HBM bytes per launch from a committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (calibrated against a copy kernel), or (None, None)."""
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch:"):
                    return int(ln.split(":")[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12     # lane-ops/s: 256 CUs x 64 lanes x 2.4 GHz
METRIC = "ME+MC Mpixel/s at 1080p (batched vs plain, bit-exact); p50/p99 frame latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads():
    """Host threads a CPU leg may use: the process's CPU affinity, at most 16
    (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def one_and_all_cores(run, seconds):
    """A CPU leg at one host thread (a third of its time budget) and at every
    thread it may use (the rest): `value` is the all-thread figure,
    `value_1core` the one-thread one; bit-exactness flags hold for both runs."""
    n = host_threads()
    one = run(seconds / 3, 1)
    out = dict(run(seconds * 2 / 3, n)) if n > 1 else dict(one)
    out["cores"] = n
    out["value_1core"] = one["value"]
    out["sample_1core"] = one["sample"]
    out["cpu_model"] = cpu_model()
    for k in ("bit_exact_vs_gpu_on_sample", "bit_exact_vs_gpu", "bit_exact_vs_oracle_on_sample",
              "bit_exact_simd_vs_c_on_sample"):
        if k in one:
            out[k] = bool(out.get(k, True)) and bool(one[k])
    return out


def cpu_baseline(fp, seconds, offs, lam, threads=None):
    """The oracle on a bounded sample of the same workload, host threads."""
    from tests.oracle_lib import OrRefinePrm, load_oracle
    oracle = load_oracle()
    pad = 80
    W, H = fp.width, fp.height
    refp = [np.ascontiguousarray(np.pad(p, pad if i == 0 else pad // 2, mode="edge")) for i, p in enumerate(fp.ref)]
    blk = np.ascontiguousarray(fp.blocks.astype(np.int32))
    nc = len(offs)
    offx = np.ascontiguousarray(np.array([o[0] for o in offs], np.int16))
    offy = np.ascontiguousarray(np.array([o[1] for o in offs], np.int16))
    prm = OrRefinePrm(nc, offx.ctypes.data, offy.ctypes.data, lam, 64, 1)
    rsy, rsc = W + 2 * pad, W // 2 + pad
    P = lambda a, off=0: ctypes.c_void_p(a.ctypes.data + off * a.itemsize)
    threads = threads or host_threads()
    chunk = 256
    n = len(blk)

    def work(start):
        cnt = min(chunk, n - start)
        cost = np.zeros(cnt * nc, np.uint32)
        a = [np.zeros(cnt, t) for t in (np.int32, np.uint32, np.int32, np.int32)]
        pred = [np.zeros((H, W), np.uint16), np.zeros((H // 2, W // 2), np.uint16), np.zeros((H // 2, W // 2), np.uint16)]
        oracle.or_RefineFracBlocks(P(fp.cur[0]), P(fp.cur[1]), P(fp.cur[2]), W, W // 2,
                                   P(refp[0], pad * rsy + pad), P(refp[1], (pad // 2) * rsc + pad // 2),
                                   P(refp[2], (pad // 2) * rsc + pad // 2), rsy, rsc, W, H,
                                   P(blk, start * 8), cnt, ctypes.byref(prm), P(cost), P(a[0]), P(a[1]),
                                   P(a[2]), P(a[3]), P(pred[0]), P(pred[1]), P(pred[2]), W, W // 2)
        return int((blk[start:start + cnt, 2] * blk[start:start + cnt, 3]).sum())

    px = 0
    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        pos = 0
        while time.perf_counter() - t0 < seconds:
            starts = [((pos + k) * chunk) % n for k in range(threads * 4)]
            pos += threads * 4
            px += sum(ex.map(work, starts))
    dt = time.perf_counter() - t0
    return {"value": round(px / dt / 1e6, 3), "unit": "Mpixel/s", "cores": threads, "kind": "port",
            "sample": f"oracle (scalar C, -O3) refinement+MC of the workload's blocks in chunks of {chunk}, "
                      f"{threads} threads, {dt:.1f} s, {px / 1e6:.1f} Mpixel of blocks"}


MFMA_I8_PEAK_TOPS = 5000.0    # dense int8 MFMA, MI355X


def session_time(lib, s, stage, fn, steps, warmup):
    """Average device time (ms) of `stage`'s kernel over `steps` calls of fn,
    from the session's HIP events on its stream."""
    from xin26x_amd import capi
    for _ in range(warmup):
        fn()
    capi.check(lib.xg_session_sync(s), "sync")
    capi.check(lib.xg_session_set_timing(s, 1), "timing")
    capi.check(lib.xg_session_reset_timers(s), "timers")
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    capi.check(lib.xg_session_sync(s), "sync")
    wall = (time.perf_counter() - t0) / steps * 1e3
    ms = ctypes.c_float()
    cnt = ctypes.c_int32()
    capi.check(lib.xg_session_kernel_time(s, stage, ctypes.byref(ms), ctypes.byref(cnt)), "kernel time")
    capi.check(lib.xg_session_set_timing(s, 0), "timing")
    return ms.value, wall


_RESIDUALS = {}


def residual_workload():
    """The 200,000-block residual workload, generated once per process."""
    from xin26x_amd import workloads as wl
    if not _RESIDUALS:
        _RESIDUALS["blocks"] = wl.residual_blocks(n=200000)
    return _RESIDUALS["blocks"]


REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libxin_ref.so")
REF8_LIB = os.path.join(ROOT, "oracle", "_ref", "libxin_ref8.so")


def ref_interp_cpu(fp, offs, seconds, threads=None, simd8=False):
    """The reference's own code on the host: the interpolation share of the
    same refinement (every block's candidates through the reference's
    interpolation, built from its sources into oracle/_ref by
    oracle/Makefile, driven by oracle/ref_interp.c) in chunks of blocks on
    host threads.

    simd8=False: the 10-bit build (h266_luma_mc.c / h266_chroma_mc.c, C only);
    the last luma prediction of each chunk is compared with the oracle's.
    simd8=True: the reference's default 8-bit build through its encoder
    tables with the SSSE3 and AVX2 entries, on the same frames shifted to 8
    bits; one chunk is run again through the C entries of the same build and
    compared.  None when that library was not built."""
    path = REF8_LIB if simd8 else REF_LIB
    if not os.path.exists(path):
        return None
    from tests.oracle_lib import load_oracle
    ref = ctypes.CDLL(path)
    ref.ref_RefineInterpShare.restype = ctypes.c_int64
    level = None
    if simd8:
        ref.ref_SetDispatch.restype = ctypes.c_int32
        level = ref.ref_SetDispatch(2)
    oracle = load_oracle()
    threads = threads or host_threads()
    pad = 80
    W, H = fp.width, fp.height
    planes = [(p >> 2).astype(np.uint8) for p in fp.ref] if simd8 else fp.ref
    pt = np.uint8 if simd8 else np.uint16
    refp = [np.ascontiguousarray(np.pad(p, pad if i == 0 else pad // 2, mode="edge")) for i, p in enumerate(planes)]
    rsy, rsc = W + 2 * pad, W // 2 + pad
    blk = np.ascontiguousarray(fp.blocks.astype(np.int32))
    offx = np.ascontiguousarray(np.array([o[0] for o in offs], np.int16))
    offy = np.ascontiguousarray(np.array([o[1] for o in offs], np.int16))
    P = lambda a, off=0: ctypes.c_void_p(a.ctypes.data + off * a.itemsize)
    chunk, n, nc = 256, len(blk), len(offs)

    def interp(start):
        cnt = min(chunk, n - start)
        oy, ou, ov = np.zeros(64 * 64, pt), np.zeros(32 * 32, pt), np.zeros(32 * 32, pt)
        px = ref.ref_RefineInterpShare(P(refp[0], pad * rsy + pad), P(refp[1], (pad // 2) * rsc + pad // 2),
                                       P(refp[2], (pad // 2) * rsc + pad // 2), ctypes.c_ssize_t(rsy),
                                       ctypes.c_ssize_t(rsc), W, H, P(blk, start * 8), cnt, nc, P(offx), P(offy), 64,
                                       P(oy), P(ou), P(ov))
        return px, cnt, (oy, ou, ov)

    def work(start):
        px, cnt, (oy, _, _) = interp(start)
        if simd8:
            return px, True
        x, y, w, h, mx, my = (int(v) for v in blk[start + cnt - 1, :6])
        mnx, mny = np.int16(-x - 72), np.int16(-y - 72)
        mxx, mxy = np.int16(W + (64 - w + 8) - x), np.int16(H + (64 - h + 8) - y)
        mvx = min(max(mx + int(offx[-1]), 16 * int(mnx)), 16 * int(mxx))
        mvy = min(max(my + int(offy[-1]), 16 * int(mny)), 16 * int(mxy))
        rx, ry = 16 * x + mvx, 16 * y + mvy
        e = np.zeros(64 * 64, np.uint16)
        oracle.or_LumaInterpolation(P(refp[0], (pad + (ry >> 4)) * rsy + pad + (rx >> 4)), rsy, P(e), 64,
                                    ((ry & 15) << 4) + (rx & 15), w, h, 0)
        same = np.array_equal(oy.reshape(64, 64)[:h, :w], e.reshape(64, 64)[:h, :w])
        return px, same

    px, exact, pos = 0, True, 0
    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        while time.perf_counter() - t0 < seconds:
            starts = [((pos + k) * chunk) % n for k in range(threads * 2)]
            pos += threads * 2
            for p_, ok in ex.map(work, starts):
                px += p_
                exact &= bool(ok)
    dt = time.perf_counter() - t0
    if simd8:
        # Every thread has finished: the tables may be rebuilt.  The last
        # block of one chunk, through the SIMD and then the C entries.
        simd = interp(0)[2]
        ref.ref_SetDispatch(0)
        plain = interp(0)[2]
        exact = all(np.array_equal(a, b) for a, b in zip(simd, plain))
        names = {2: "SSSE3 + AVX2", 1: "SSSE3", 0: "C"}
        return {"value": round(px / dt / 1e6, 3), "unit": "Mpixel/s", "cores": threads, "kind": "reference",
                "sample": f"the reference's default 8-bit build (gcc -O2, built from its sources) through its "
                          f"encoder's interpolation tables with the {names[level]} entries, on the workload's "
                          f"frames shifted to 8 bits: the interpolation share only (luma 8-tap + Cb/Cr 4-tap "
                          f"prediction of all {nc} candidates) in chunks of {chunk}, {threads} threads, "
                          f"{dt:.1f} s, {px / 1e6:.1f} Mpixel of blocks. 8-bit samples, not the workload's 10",
                "bit_exact_simd_vs_c_on_sample": bool(exact)}
    return {"value": round(px / dt / 1e6, 3), "unit": "Mpixel/s", "cores": threads, "kind": "reference",
            "sample": f"the reference's 10-bit h266_luma_mc.c + h266_chroma_mc.c (gcc -O2, built from its sources) "
                      f"on the interpolation share only (luma 8-tap + Cb/Cr 4-tap prediction of all {nc} "
                      f"candidates, no SATD, cost or MC) of the workload's blocks in chunks of {chunk}, {threads} "
                      f"threads, {dt:.1f} s, {px / 1e6:.1f} Mpixel of blocks",
            "bit_exact_vs_oracle_on_sample": bool(exact)}


def transforms_leg(lib, steps, warmup, cpu_seconds):
    """200,000 residual blocks, shapes 4..64 x 4..64, DCT-II or DST-VII/DCT-VIII:
    the batched int8-MFMA forward transform, against the HBM and MFMA rooflines."""
    import torch
    from xin26x_amd import capi
    desc, offs, resid = residual_workload()
    n = len(desc)
    w = 1 << (desc & 15).astype(np.int64)
    h = 1 << ((desc >> 4) & 15).astype(np.int64)
    mts = (desc >> 8) & 15
    keep = lambda side: np.where(mts > 0, np.minimum(side, 16 * (side == 32) + side * (side < 32)), np.minimum(side, 32))
    wk, hk = keep(w), keep(h)
    macs = int((h * w * wk + wk * hk * h).sum())          # stage 1 rows, then the kept columns
    samples = int((w * h).sum())
    alg_bytes = samples * (2 + 4)                          # int16 residual in, int32 coefficients out
    s = ctypes.c_void_p()
    capi.check(lib.xg_session_create(ctypes.byref(s), 16), "session")
    try:
        dres = torch.from_numpy(resid).cuda()
        dco = torch.empty(resid.size, dtype=torch.int32, device="cuda")
        capi.check(lib.xg_fdct_plan(s, offs.ctypes.data, desc.ctypes.data, n), "fdct plan")
        run = lambda: capi.check(lib.xg_fdct_run(s, dres.data_ptr(), dco.data_ptr()), "fdct run")
        k_ms, wall_ms = session_time(lib, s, 2, run, steps, warmup)
        coeff = dco.cpu().numpy()
    finally:
        lib.xg_session_destroy(s)
    k_s = k_ms / 1e3
    out = {"workload": "200,000 residual blocks, sides 4..64 (VVC-like weights), Laplacian(12) int16 residuals, "
                       "DCT-II or DST-VII/DCT-VIII (sides <= 32), 10-bit shifts, zero-out",
           "value": round(n / (wall_ms / 1e3), 1), "unit": "blocks/s",
           "coefficients_per_s": round(samples / (wall_ms / 1e3), 1), "ms_per_batch": round(wall_ms, 4),
           "kernel": "k_fdct_mfma<32> + k_fdct_mfma<64> (blocks up to 32 high, then 64 high)",
           "kernel_ms": round(k_ms, 4),
           "roofline": {"bound": "hbm", "achieved": round(alg_bytes / k_s / 1e9, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(alg_bytes / k_s / 1e9 / HBM_PEAK_GBS, 5),
                        "traffic": profiled_traffic(PMC_FDCT)[0], "traffic_source": profiled_traffic(PMC_FDCT)[1],
                        "alg_bytes_per_launch": alg_bytes},
           "mfma": {"achieved": round(2 * macs / k_s / 1e12, 3), "peak": MFMA_I8_PEAK_TOPS, "unit": "TOP/s",
                    "frac": round(2 * macs / k_s / 1e12 / MFMA_I8_PEAK_TOPS, 6),
                    "useful_macs_per_launch": macs}}
    if cpu_seconds > 0:
        out["cpu_baseline"] = one_and_all_cores(
            lambda sec, t: transforms_cpu(desc, offs, resid, coeff, sec, threads=t), cpu_seconds)
    return out


def transforms_cpu(desc, offs, resid, gpu_coeff, seconds, threads=None):
    """The oracle's integer transform on the host, chunks of the same blocks;
    the chunks it computes are also compared with the GPU's coefficients."""
    from tests.oracle_lib import load_oracle
    oracle = load_oracle()
    threads = threads or host_threads()
    out = np.zeros(resid.size, np.int32)
    chunk, n = 2000, len(desc)
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    done = []

    def work(i0):
        i1 = min(n, i0 + chunk)
        oracle.or_FDctBatch(P(resid), P(offs), P(desc), i0, i1, P(out))
        return i0, i1

    blocks = 0
    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        pos = 0
        while time.perf_counter() - t0 < seconds and pos < n:
            starts = [pos + k * chunk for k in range(threads * 2) if pos + k * chunk < n]
            pos += threads * 2 * chunk
            for i0, i1 in ex.map(work, starts):
                done.append((i0, i1))
                blocks += i1 - i0
    dt = time.perf_counter() - t0
    exact = all(np.array_equal(out[offs[i0]:offs[i1 - 1] + (1 << (desc[i1 - 1] & 15)) * (1 << ((desc[i1 - 1] >> 4) & 15))],
                               gpu_coeff[offs[i0]:offs[i1 - 1] + (1 << (desc[i1 - 1] & 15)) * (1 << ((desc[i1 - 1] >> 4) & 15))])
                for i0, i1 in done)
    return {"value": round(blocks / dt, 1), "unit": "blocks/s", "cores": threads, "kind": "port",
            "sample": f"oracle (scalar C, int64 matrix products, -O3) on the first {blocks} blocks in chunks of "
                      f"{chunk}, {threads} host threads, {dt:.1f} s",
            "bit_exact_vs_gpu_on_sample": bool(exact)}


PMC_RECON = os.path.join(ROOT, "profiles", "k_recon_mfma_200k_blocks.md")


def recon_leg(lib, steps, warmup, cpu_seconds):
    """Reconstruction of the same 200,000 blocks: their forward coefficients
    (computed on the GPU, int16) through the batched inverse transform plus
    the add of a 10-bit prediction, against the HBM and MFMA rooflines."""
    import torch
    from xin26x_amd import capi, workloads as wl
    desc, offs, resid = residual_workload()
    n = len(desc)
    w = 1 << (desc & 15).astype(np.int64)
    h = 1 << ((desc >> 4) & 15).astype(np.int64)
    mts = (desc >> 8) & 15
    nz = lambda side: np.where(mts > 0, np.where(side == 32, 16, side), np.minimum(side, 32))
    nzw, nzh = nz(w), nz(h)
    samples = int((w * h).sum())
    macs = int((nzw * h * nzh + h * w * nzw).sum())        # stage 1 kept columns, then every output sample
    alg_bytes = int((2 * nzh * w + 4 * w * h).sum())        # int16 coefficient rows below the zero-out; pred in, recon out
    pred = np.minimum((wl.uniform01(6, resid.size, stream=4) * 1024).astype(np.int64), 1023).astype(np.uint16)
    s = ctypes.c_void_p()
    capi.check(lib.xg_session_create(ctypes.byref(s), 16), "session")
    try:
        dres = torch.from_numpy(resid).cuda()
        dco = torch.empty(resid.size, dtype=torch.int32, device="cuda")
        capi.check(lib.xg_fdct_plan(s, offs.ctypes.data, desc.ctypes.data, n), "fdct plan")
        capi.check(lib.xg_fdct_run(s, dres.data_ptr(), dco.data_ptr()), "fdct run")
        capi.check(lib.xg_session_sync(s), "sync")
        coeff = dco.clamp(-32768, 32767).to(torch.int16)
        del dres, dco
        dpred = torch.from_numpy(pred.view(np.int16)).cuda()
        drec = torch.empty(resid.size, dtype=torch.int16, device="cuda")
        capi.check(lib.xg_recon_plan(s, offs.ctypes.data, desc.ctypes.data, n), "recon plan")
        run = lambda: capi.check(lib.xg_recon_run(s, coeff.data_ptr(), dpred.data_ptr(), drec.data_ptr()),
                                 "recon run")
        k_ms, wall_ms = session_time(lib, s, 5, run, steps, warmup)
        coeff_h = coeff.cpu().numpy()
        recon = drec.cpu().numpy().view(np.uint16)
    finally:
        lib.xg_session_destroy(s)
    k_s = k_ms / 1e3
    traffic, src = profiled_traffic(PMC_RECON)
    out = {"workload": "200,000 TUs of the transform workload: their forward coefficients (int16) through "
                       "the inverse DCT-II/DST-VII/DCT-VIII (shifts 7 and 10, zero-out) plus a uniform 10-bit "
                       "prediction, clipped",
           "value": round(n / (wall_ms / 1e3), 1), "unit": "blocks/s",
           "samples_per_s": round(samples / (wall_ms / 1e3), 1), "ms_per_batch": round(wall_ms, 4),
           "kernel": "k_recon_mfma", "kernel_ms": round(k_ms, 4),
           "roofline": {"bound": "hbm", "achieved": round(alg_bytes / k_s / 1e9, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(alg_bytes / k_s / 1e9 / HBM_PEAK_GBS, 5),
                        "traffic": traffic, "traffic_source": src, "alg_bytes_per_launch": alg_bytes},
           "mfma": {"achieved": round(2 * macs / k_s / 1e12, 3), "peak": MFMA_I8_PEAK_TOPS, "unit": "TOP/s",
                    "frac": round(2 * macs / k_s / 1e12 / MFMA_I8_PEAK_TOPS, 6),
                    "useful_macs_per_launch": macs}}
    if cpu_seconds > 0:
        out["cpu_baseline"] = one_and_all_cores(
            lambda sec, t: recon_cpu(desc, offs, coeff_h, pred, recon, sec, threads=t), cpu_seconds)
    return out


def recon_cpu(desc, offs, coeff, pred, gpu_recon, seconds, threads=None):
    """The oracle's reconstruction on the host, chunks of the same blocks;
    the chunks it computes are also compared with the GPU's samples."""
    from tests.oracle_lib import load_oracle
    oracle = load_oracle()
    threads = threads or host_threads()
    out = np.zeros(pred.size, np.uint16)
    chunk, n = 2000, len(desc)
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    done = []

    def work(i0):
        i1 = min(n, i0 + chunk)
        oracle.or_ReconBatch(P(coeff), P(pred), P(offs), P(desc), i0, i1, P(out))
        return i0, i1

    blocks = 0
    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        pos = 0
        while time.perf_counter() - t0 < seconds and pos < n:
            starts = [pos + k * chunk for k in range(threads * 2) if pos + k * chunk < n]
            pos += threads * 2 * chunk
            for i0, i1 in ex.map(work, starts):
                done.append((i0, i1))
                blocks += i1 - i0
    dt = time.perf_counter() - t0
    end = lambda i: offs[i] + (1 << (desc[i] & 15)) * (1 << ((desc[i] >> 4) & 15))
    exact = all(np.array_equal(out[offs[i0]:end(i1 - 1)], gpu_recon[offs[i0]:end(i1 - 1)]) for i0, i1 in done)
    return {"value": round(blocks / dt, 1), "unit": "blocks/s", "cores": threads, "kind": "port",
            "sample": f"oracle (scalar C, int64 matrix products, -O3) on the first {blocks} blocks in chunks of "
                      f"{chunk}, {threads} host threads, {dt:.1f} s",
            "bit_exact_vs_gpu_on_sample": bool(exact)}


PMC_QUANT = os.path.join(ROOT, "profiles", "k_quant_200k_blocks.md")
RESIDUAL_QP = 22


def quant_descs(desc):
    """Quantiser descriptors of the transform workload's TUs: 4x4 groups, QP 22
    (table index QP + 12; 96% of the TUs keep non-zero levels, 7% of the
    coefficients), a P frame's dead-zone rounding."""
    return ((desc & 0xFFF) | 2 << 12 | 2 << 15 | (RESIDUAL_QP + 12) << 18).astype(np.uint32)


def residual_loop_leg(lib, steps, warmup, cpu_seconds):
    """The residual loop of the 200,000 TUs on the GPU, per batch: forward
    transform, quantisation + dequantisation (QP 32), reconstruction with a
    prediction; the quantiser against the HBM roofline."""
    import torch
    from xin26x_amd import capi, workloads as wl
    desc, offs, resid = residual_workload()
    n = len(desc)
    dq_desc = quant_descs(desc)
    w = 1 << (desc & 15).astype(np.int64)
    h = 1 << ((desc >> 4) & 15).astype(np.int64)
    region = np.minimum(w, 32) * np.minimum(h, 32)
    samples = int((w * h).sum())
    # coefficients read over the region; levels and dequantised values (int16) written over the TU; the maps
    alg_bytes = int((4 * region + 4 * w * h).sum()) + 8 * n
    pred = np.minimum((wl.uniform01(6, resid.size, stream=4) * 1024).astype(np.int64), 1023).astype(np.uint16)
    s = ctypes.c_void_p()
    capi.check(lib.xg_session_create(ctypes.byref(s), 16), "session")
    try:
        dres = torch.from_numpy(resid).cuda()
        dco = torch.empty(resid.size, dtype=torch.int32, device="cuda")
        dq = torch.empty(resid.size, dtype=torch.int16, device="cuda")
        dr = torch.empty(resid.size, dtype=torch.int16, device="cuda")
        dnz = torch.empty(n, dtype=torch.int64, device="cuda")
        dpred = torch.from_numpy(pred.view(np.int16)).cuda()
        drec = torch.empty(resid.size, dtype=torch.int16, device="cuda")
        capi.check(lib.xg_fdct_plan(s, offs.ctypes.data, desc.ctypes.data, n), "fdct plan")
        capi.check(lib.xg_quant_plan(s, offs.ctypes.data, dq_desc.ctypes.data, n), "quant plan")
        capi.check(lib.xg_recon_plan(s, offs.ctypes.data, desc.ctypes.data, n), "recon plan")

        def chain():
            capi.check(lib.xg_fdct_run(s, dres.data_ptr(), dco.data_ptr()), "fdct run")
            capi.check(lib.xg_quant_run(s, dco.data_ptr(), dq.data_ptr(), dr.data_ptr(), dnz.data_ptr()), "quant")
            capi.check(lib.xg_recon_run(s, dr.data_ptr(), dpred.data_ptr(), drec.data_ptr()), "recon run")

        q_ms, wall_ms = session_time(lib, s, 8, chain, steps, warmup)
        coeff = dco.cpu().numpy()
        got = (dq.cpu().numpy(), dr.cpu().numpy(), dnz.cpu().numpy().view(np.uint64))
        rec = drec.cpu().numpy()
        nz_share = float((got[2] != 0).mean())
        # the same loop with the quantiser fused into the forward transform's last stage
        capi.check(lib.xg_fdct_quant_plan(s, offs.ctypes.data, desc.ctypes.data, dq_desc.ctypes.data, n), "fq plan")

        def fused():
            capi.check(lib.xg_fdct_quant_run(s, dres.data_ptr(), dq.data_ptr(), dr.data_ptr(), dnz.data_ptr()), "fq")
            capi.check(lib.xg_recon_run(s, dr.data_ptr(), dpred.data_ptr(), drec.data_ptr()), "recon run")

        dq.zero_()
        dr.zero_()
        dnz.zero_()
        fq_ms, fwall_ms = session_time(lib, s, 2, fused, steps, warmup)
        fused_same = bool(np.array_equal(dq.cpu().numpy(), got[0]) and np.array_equal(dr.cpu().numpy(), got[1])
                          and np.array_equal(dnz.cpu().numpy().view(np.uint64), got[2])
                          and np.array_equal(drec.cpu().numpy(), rec))
    finally:
        lib.xg_session_destroy(s)
    k_s = q_ms / 1e3
    traffic, src = profiled_traffic(PMC_QUANT)
    out = {"workload": "200,000 TUs of the transform workload per batch: forward transform, quantisation + "
                       f"dequantisation (QP {RESIDUAL_QP}, 4x4 groups, P-frame dead zone), reconstruction with a "
                       "uniform 10-bit prediction",
           "value": round(n / (wall_ms / 1e3), 1), "unit": "blocks/s",
           "samples_per_s": round(samples / (wall_ms / 1e3), 1), "ms_per_batch": round(wall_ms, 4),
           "kernels": "k_fdct_mfma + k_quant + k_recon_mfma", "quant_kernel_ms": round(q_ms, 4),
           "tus_with_nonzero_levels": round(nz_share, 4),
           "fused": {"kernels": "k_fdct_mfma<QUANT> (forward transform + quantiser) + k_recon_mfma",
                     "value": round(n / (fwall_ms / 1e3), 1), "unit": "blocks/s", "ms_per_batch": round(fwall_ms, 4),
                     "fdct_quant_kernel_ms": round(fq_ms, 4), "same_outputs_as_three_kernels": fused_same},
           "quant_roofline": {"bound": "hbm", "achieved": round(alg_bytes / k_s / 1e9, 2), "peak": HBM_PEAK_GBS,
                              "unit": "GB/s", "frac": round(alg_bytes / k_s / 1e9 / HBM_PEAK_GBS, 5),
                              "traffic": traffic, "traffic_source": src, "alg_bytes_per_launch": alg_bytes}}
    if cpu_seconds > 0:
        out["cpu_baseline"] = one_and_all_cores(
            lambda sec, t: quant_cpu(desc, dq_desc, offs, coeff, got, sec, threads=t), cpu_seconds)
    return out


def quant_cpu(desc, dq_desc, offs, coeff, got, seconds, threads=None):
    """The oracle's quantiser on the host over chunks of the same TUs (the GPU's
    forward coefficients as input); the chunks it computes are compared with the
    GPU's levels, dequantised values and maps."""
    from tests.oracle_lib import load_oracle
    oracle = load_oracle()
    threads = threads or host_threads()
    q = np.zeros(coeff.size, np.int16)
    r = np.zeros(coeff.size, np.int16)
    nz = np.zeros(len(desc), np.uint64)
    chunk, n = 4000, len(desc)
    P = lambda a: ctypes.c_void_p(a.ctypes.data)
    done = []

    def work(i0):
        i1 = min(n, i0 + chunk)
        oracle.or_QuantBatch(P(coeff), P(offs), P(dq_desc), i0, i1, P(q), P(r), P(nz))
        return i0, i1

    blocks = passes = 0
    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        while time.perf_counter() - t0 < seconds:      # whole passes over the batch until the time is up
            for i0, i1 in ex.map(work, range(0, n, chunk)):
                if passes == 0:
                    done.append((i0, i1))
                blocks += i1 - i0
            passes += 1
    dt = time.perf_counter() - t0
    end = lambda i: offs[i] + (1 << (desc[i] & 15)) * (1 << ((desc[i] >> 4) & 15))
    exact = all(np.array_equal(q[offs[i0]:end(i1 - 1)], got[0][offs[i0]:end(i1 - 1)]) and
                np.array_equal(r[offs[i0]:end(i1 - 1)], got[1][offs[i0]:end(i1 - 1)]) and
                np.array_equal(nz[i0:i1], got[2][i0:i1]) for i0, i1 in done)
    return {"value": round(blocks / dt, 1), "unit": "blocks/s (quantisation only)", "cores": threads, "kind": "port",
            "sample": f"oracle (scalar C, -O3) quantiser, {passes} passes over the {n} TUs in chunks of {chunk}, "
                      f"{threads} host threads, {dt:.1f} s",
            "bit_exact_vs_gpu_on_sample": bool(exact)}


def mctf_work(W, H):
    """Squared differences one xg_mctf_me frame costs at most (every candidate
    and position inside its padding): 4x level +/-8 on 16x16 blocks, 2x and
    full resolution 10 candidates + +/-5 on 16x16, 8x8 blocks with 10
    candidates, the centre and 2 x 49 fractional positions."""
    w2, h2, w4, h4 = W // 2, H // 2, W // 4, H // 4
    return ((w4 // 16) * (h4 // 16) * 289 * 256 + (w2 // 16) * (h2 // 16) * 131 * 256 +
            (W // 16) * (H // 16) * 131 * 256 + (W // 8) * (H // 8) * 109 * 64)


def mctf_leg(lib, fp, cur, ref, steps, warmup, cpu_seconds):
    """MCTF motion estimation of the 1080p pair: four levels (4x, 2x, full on
    16x16 blocks, 8x8 blocks with fractional rounds)."""
    import torch
    from xin26x_amd import capi
    W, H = fp.width, fp.height
    nb = (W // 8) * (H // 8)
    s = ctypes.c_void_p()
    capi.check(lib.xg_session_create(ctypes.byref(s), 16), "session")
    try:
        mv = torch.empty(2 * nb, dtype=torch.int32, device="cuda")
        err = torch.empty(nb, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        p0, p1 = cur.pic.plane[0], ref.pic.plane[0]
        run = lambda: capi.check(lib.xg_mctf_me(s, ctypes.byref(p0), ctypes.byref(p1), mv.data_ptr(),
                                                err.data_ptr()), "mctf")
        k_ms, wall_ms = session_time(lib, s, 10, run, steps, warmup)
        frame_ms, _ = session_time(lib, s, 9, run, steps, 0)
        got = mv.cpu().numpy().reshape(H // 8, W // 8, 2)
    finally:
        lib.xg_session_destroy(s)
    work = mctf_work(W, H)
    out = {"workload": f"1080p pair, MCTF motion estimation: 4x level (16x16, +/-8), 2x and full resolution (16x16, "
                       f"+/-5 around 3x3 lower candidates), {nb} 8x8 blocks with quarter- and 1/16-pel rounds "
                       f"(8-tap), SSD",
           "value": round(W * H / (wall_ms / 1e3) / 1e6, 1), "unit": "Mpixel/s", "ms_per_frame": round(wall_ms, 4),
           "device_ms_per_frame": round(frame_ms, 4), "kernel": "k_mctf_level (finest level)",
           "kernel_ms": round(k_ms, 4), "sq_diffs_per_frame_max": work,
           "sq_diffs_per_s": round(work / (frame_ms / 1e3), 1)}
    if cpu_seconds > 0:
        out["cpu_baseline"] = one_and_all_cores(lambda sec, t: mctf_cpu(fp, got, sec, threads=t), cpu_seconds)
    return out


def mctf_cpu(fp, got, seconds, threads=None):
    """The oracle's MCTF hierarchy on whole frames of the same pair, every
    level's block rows split over host threads, frames until about `seconds`
    have passed (at least one); the finest MVs are compared with the GPU's."""
    from tests.oracle_lib import load_oracle
    oracle = load_oracle()
    threads = threads or host_threads()
    cur, ref = np.ascontiguousarray(fp.cur[0]), np.ascontiguousarray(fp.ref[0])
    H, W = cur.shape
    P = lambda a, o=0: ctypes.c_void_p(a.ctypes.data + o * a.itemsize)

    def sub(src, pad):
        h, w = src.shape
        o = np.zeros((h // 2 + 2 * pad, w // 2 + 2 * pad), np.uint16)
        oracle.or_MctfSubsample(P(src), w, w, h, P(o, pad * o.shape[1] + pad), o.shape[1], pad)
        return o, np.ascontiguousarray(o[pad:pad + h // 2, pad:pad + w // 2])

    def level(ex, c, rp, pad, bs, rng, lower, factor, frac):
        h, w = c.shape
        nbx, nby = w // bs, h // bs
        mv, err = np.zeros(2 * nbx * nby, np.int32), np.zeros(nbx * nby, np.uint32)
        cuts = [nby * k // threads for k in range(threads + 1)]
        lb, lh = (0, 0) if lower is None else (lower.shape[1], lower.shape[0])
        list(ex.map(lambda k: oracle.or_MctfMeLevel(P(c), w, P(rp, pad * rp.shape[1] + pad), rp.shape[1], w, h, pad,
                                                    bs, rng, None if lower is None else P(lower), lb, lh, factor,
                                                    frac, cuts[k], cuts[k + 1], P(mv), P(err)), range(threads)))
        return mv.reshape(nby, nbx, 2)

    r1 = np.ascontiguousarray(np.pad(ref, 80, mode="edge"))
    frames, exact = 0, True
    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        while frames == 0 or time.perf_counter() - t0 < seconds:
            r2, r2i = sub(ref, 32)
            r4, _ = sub(r2i, 16)
            _, c2 = sub(cur, 0)
            _, c4 = sub(c2, 0)
            m4 = np.ascontiguousarray(level(ex, c4, r4, 16, 16, 8, None, 1, 0))
            m2 = np.ascontiguousarray(level(ex, c2, r2, 32, 16, 5, m4, 2, 0))
            m1 = np.ascontiguousarray(level(ex, cur, r1, 80, 16, 5, m2, 2, 0))
            m0 = level(ex, cur, r1, 80, 8, 0, m1, 1, 1)
            exact &= bool(np.array_equal(m0, got))
            frames += 1
    dt = time.perf_counter() - t0
    return {"value": round(W * H * frames / dt / 1e6, 3), "unit": "Mpixel/s", "cores": threads, "kind": "port",
            "sample": f"oracle (scalar C, -O3) MCTF hierarchy on {frames} whole frame(s), each level's block rows "
                      f"split over {threads} threads, {dt:.1f} s", "bit_exact_vs_gpu": exact}


def lookahead_leg(lib, fp, cur, ref, steps, warmup, cpu_seconds, unit=16, rng=32, lam=1024):
    """Lookahead motion estimation of the 1080p pair: both luma planes downscaled
    2x, every 16x16 unit of the 960x540 picture searched +/-32 (+/-64 at full
    resolution)."""
    import torch
    from xin26x_amd import capi
    hw, hh = fp.width // 2, fp.height // 2
    nb = (hw // unit) * (hh // unit)
    s = ctypes.c_void_p()
    capi.check(lib.xg_session_create(ctypes.byref(s), 16), "session")
    try:
        mv = torch.empty(2 * nb, dtype=torch.int16, device="cuda")
        cost = torch.empty(nb, dtype=torch.int32, device="cuda")
        sad = torch.empty(nb, dtype=torch.int32, device="cuda")
        p0, p1 = cur.pic.plane[0], ref.pic.plane[0]
        run = lambda: capi.check(lib.xg_lookahead_me(s, ctypes.byref(p0), ctypes.byref(p1), unit, rng, lam,
                                                     mv.data_ptr(), cost.data_ptr(), sad.data_ptr()), "lookahead")
        k_ms, wall_ms = session_time(lib, s, 3, run, steps, warmup)
        got = (mv.cpu().numpy().reshape(nb, 2), cost.cpu().numpy().view(np.uint32), sad.cpu().numpy().view(np.uint32))
    finally:
        lib.xg_session_destroy(s)
    evals = nb * (2 * rng + 1) ** 2 * unit * unit
    out = {"workload": f"1080p pair downscaled 2x (960x540), {nb} units of {unit}x{unit}, integer full search "
                       f"+/-{rng} (+/-{2 * rng} at full resolution), SAD + lambda * MV bits",
           "value": round(fp.width * fp.height / (wall_ms / 1e3) / 1e6, 1), "unit": "Mpixel/s (full-resolution)",
           "ms_per_frame": round(wall_ms, 4), "kernel": "k_down2x2 x 2 + k_full_search_frame_p<16>",
           "search_kernel_ms": round(k_ms, 4), "abs_diffs_per_s": round(evals / (k_ms / 1e3), 1)}
    if cpu_seconds > 0:
        out["cpu_baseline"] = one_and_all_cores(
            lambda sec, t: lookahead_cpu(fp, unit, rng, lam, got, sec, threads=t), cpu_seconds)
    return out


def lookahead_cpu(fp, unit, rng, lam, got, seconds, threads=None):
    """The oracle's downscale and full search on the host over rows of units;
    the rows it computes are compared with the GPU's MVs, costs and SADs."""
    from tests.oracle_lib import load_oracle
    oracle = load_oracle()
    P = lambda a, o=0: ctypes.c_void_p(a.ctypes.data + 2 * o)
    W, H, pad = fp.width, fp.height, 80
    hw, hh, hp = W // 2, H // 2, pad // 2
    t0 = time.perf_counter()
    cur = np.ascontiguousarray(fp.cur[0])
    refp = np.ascontiguousarray(np.pad(fp.ref[0], pad, mode="edge"))
    hcur = np.zeros((hh, hw), np.uint16)
    href = np.zeros((hh + 2 * hp, hw + 2 * hp), np.uint16)
    oracle.or_Downscale2x2(P(cur), W, P(hcur), hw, hw, hh)
    oracle.or_Downscale2x2(P(refp), W + 2 * pad, P(href), hw + 2 * hp, hw + 2 * hp, hh + 2 * hp)
    nbx, nb = hw // unit, (hw // unit) * (hh // unit)
    mx, my = np.zeros(nb, np.int16), np.zeros(nb, np.int16)
    c, sd = np.zeros(nb, np.uint32), np.zeros(nb, np.uint32)
    threads = threads or host_threads()

    def work(r):
        b0, b1 = r * nbx, (r + 1) * nbx
        # or_FullSearchFrame writes blocks [b0, b1) at their own indices of the output arrays
        oracle.or_FullSearchFrame(P(hcur), hw, P(href, hp * (hw + 2 * hp) + hp), hw + 2 * hp, hw, hh, unit, unit, rng,
                                  lam, rng, b0, b1, P(mx), P(my), P(c), P(sd))
        return r

    rows = []
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        r = 0
        while time.perf_counter() - t0 < seconds and r < hh // unit:
            batch = list(range(r, min(hh // unit, r + threads)))
            rows += list(ex.map(work, batch))
            r += len(batch)
    dt = time.perf_counter() - t0
    sel = np.concatenate([np.arange(q * nbx, (q + 1) * nbx) for q in rows]) if rows else np.zeros(0, np.int64)
    exact = bool(np.array_equal(got[0][sel, 0], mx[sel]) and np.array_equal(got[0][sel, 1], my[sel])
                 and np.array_equal(got[1][sel], c[sel]) and np.array_equal(got[2][sel], sd[sel]))
    px = len(rows) * unit * 2 * W
    return {"value": round(px / dt / 1e6, 2), "unit": "Mpixel/s (full-resolution)", "cores": threads, "kind": "port",
            "sample": f"oracle downscale + full search on {len(rows)} of {hh // unit} unit rows, {threads} host "
                      f"threads, {dt:.1f} s", "bit_exact_vs_gpu_on_sample": exact}


PMC_BIPRED = os.path.join(ROOT, "profiles", "k_bipred_tiles_1080p.md")


def percall_leg(lib, calls, cpu_calls):
    """Latency of the drop-in table slots as the encoder calls them: host
    buffers in, result back in host memory, one block per call (16x16 SATD,
    16x16 luma 8-tap HorVet interpolation).  The oracle's scalar C of the same
    calls is timed beside it (ctypes call overhead included in both)."""
    rng = np.random.default_rng(11)
    a = np.ascontiguousarray(rng.integers(0, 1024, (64, 64)).astype(np.uint16))
    b = np.ascontiguousarray(rng.integers(0, 1024, (64, 64)).astype(np.uint16))
    dst = np.zeros((16, 16), np.uint16)
    out = np.zeros(1, np.uint32)
    P = lambda arr, off=0: ctypes.c_void_p(arr.ctypes.data + 2 * off)

    def per_call_us(fn, n):
        for _ in range(min(50, n)):
            fn()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        return (time.perf_counter() - t0) / n * 1e6

    satd = lambda: lib.Xg266ComputeSatd(P(a, 8 * 64 + 8), 64, P(b, 8 * 64 + 8), 64, 16, 16, P(out))
    interp = lambda: lib.Xg266LumaInterpHorVet(P(a, 8 * 64 + 8), 64, P(dst), 16, 0x55, 16, 16, 0)
    res = {"satd_16x16_us": round(per_call_us(satd, calls), 2), "luma_horvet_16x16_us": round(per_call_us(interp, calls), 2)}
    gpu_satd = int(out[0])
    if cpu_calls > 0:
        from tests.oracle_lib import load_oracle
        oracle = load_oracle()
        osatd = lambda: oracle.or_ComputeSatd(P(a, 8 * 64 + 8), 64, P(b, 8 * 64 + 8), 64, 16, 16)
        res["cpu_oracle_satd_16x16_us"] = round(per_call_us(osatd, cpu_calls), 2)
        res["bit_exact_satd"] = bool(osatd() == gpu_satd)
    res["note"] = ("each call stages its block into page-locked memory, copies it to HBM, runs one kernel and copies "
                   "the result back, synchronously: the slots exist for drop-in parity, and the batched forms are "
                   "what use the GPU")
    return res


def bipred_leg(lib, fp, cur, ref, pred, steps, warmup, cpu_seconds):
    """B-frame bi-prediction SATD of every 1080p block: the input against the
    rounded average of two predictions (the reference picture and the
    refinement's MC output)."""
    import torch
    from xin26x_amd import capi
    from xin26x_amd.device import blocks_to_struct
    barr = blocks_to_struct(fp.blocks)
    n = len(barr)
    s = ctypes.c_void_p()
    capi.check(lib.xg_session_create(ctypes.byref(s), 16), "session")
    try:
        capi.check(lib.xg_bipred_plan(s, barr.ctypes.data, n), "bipred plan")
        dist = torch.empty(n, dtype=torch.int32, device="cuda")
        pi, p0, p1 = cur.pic.plane[0], ref.pic.plane[0], pred.pic.plane[0]
        run = lambda: capi.check(lib.xg_bipred_run(s, ctypes.byref(pi), ctypes.byref(p0), ctypes.byref(p1), 1,
                                                   dist.data_ptr()), "bipred run")
        k_ms, wall_ms = session_time(lib, s, 7, run, steps, warmup)
        got = dist.cpu().numpy().view(np.uint32)
        pred_h = pred.tensors[0].cpu().numpy().view(np.uint16).reshape(fp.height, fp.width)
    finally:
        lib.xg_session_destroy(s)
    px = int((fp.blocks[:, 2] * fp.blocks[:, 3]).sum())
    alg_bytes = 3 * 2 * px
    k_s = k_ms / 1e3
    out = {"workload": f"1080p, {n} QTMT blocks: SATD of the input against (reference + MC prediction + 1) >> 1 "
                       "(8x8 Hadamard tiles, 4x4 where a side is not a multiple of 8)",
           "value": round(fp.width * fp.height / (wall_ms / 1e3) / 1e6, 1), "unit": "Mpixel/s",
           "blocks_per_s": round(n / (wall_ms / 1e3), 1), "ms_per_frame": round(wall_ms, 4),
           "kernel": "k_bipred_tiles (one launch per frame)", "kernel_ms": round(k_ms, 4),
           "roofline": {"bound": "hbm", "achieved": round(alg_bytes / k_s / 1e9, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(alg_bytes / k_s / 1e9 / HBM_PEAK_GBS, 5),
                        "traffic": profiled_traffic(PMC_BIPRED)[0], "traffic_source": profiled_traffic(PMC_BIPRED)[1],
                        "alg_bytes_per_launch": alg_bytes}}
    if cpu_seconds > 0:
        out["cpu_baseline"] = one_and_all_cores(lambda sec, t: bipred_cpu(fp, pred_h, got, sec, threads=t),
                                                cpu_seconds)
    return out


def bipred_cpu(fp, pred_h, got, seconds, threads=None):
    """or_ComputeAvgSatd over chunks of the same blocks on host threads; every
    block it costs is compared with the GPU's SATD."""
    from tests.oracle_lib import load_oracle
    oracle = load_oracle()
    threads = threads or host_threads()
    inp, r0 = np.ascontiguousarray(fp.cur[0]), np.ascontiguousarray(fp.ref[0])
    p1h = np.ascontiguousarray(pred_h)
    W, blocks = fp.width, fp.blocks[:, :4].tolist()
    P = lambda a, o: ctypes.c_void_p(a.ctypes.data + 2 * o)
    chunk, n = 512, len(blocks)

    def work(i0):
        ok = True
        for i in range(i0, min(n, i0 + chunk)):
            x, y, w, h = blocks[i]
            o = y * W + x
            ok &= oracle.or_ComputeAvgSatd(P(inp, o), W, P(r0, o), P(p1h, o), W, w, h) == int(got[i])
        return min(n, i0 + chunk) - i0, ok

    done, exact, pos = 0, True, 0
    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        while time.perf_counter() - t0 < seconds:
            starts = [((pos + k) * chunk) % n for k in range(threads)]
            pos += threads
            for cnt, ok in ex.map(work, starts):
                done += cnt
                exact &= ok
    dt = time.perf_counter() - t0
    return {"value": round(done / dt, 1), "unit": "blocks/s", "cores": threads, "kind": "port",
            "sample": f"oracle (scalar C) on {done} blocks in chunks of {chunk}, {threads} threads, {dt:.1f} s",
            "bit_exact_vs_gpu_on_sample": bool(exact)}


PMC_DMVR = os.path.join(ROOT, "profiles", "k_dmvr_1080p.md")


def dmvr_blocks(fp, seed=12):
    """The bi-predicted DMVR candidates of the 1080p workload: its QTMT blocks
    with both sides >= 8 and area >= 128; list-0 MV = the block's base MV plus
    a seeded 1/16-pel offset, list-1 MV mirrored with its own offset; the MV
    window of a CU in a 64-sample CTU, kept inside 80 samples of padding."""
    from xin26x_amd import capi, workloads as wl
    b = fp.blocks
    sel = (b[:, 2] >= 8) & (b[:, 3] >= 8) & (b[:, 2] * b[:, 3] >= 128)
    b = b[sel]
    off = wl.randint(seed, 4 * len(b), -12, 12, stream=7).reshape(len(b), 4)
    arr = (capi.XgDmvrBlock * len(b))()
    for i, (x, y, w, h, mx, my) in enumerate(b[:, :6].tolist()):
        arr[i] = capi.XgDmvrBlock(x, y, w, h, mx + int(off[i, 0]), my + int(off[i, 1]),
                                  -mx + int(off[i, 2]), -my + int(off[i, 3]),
                                  -x - 72, -y - 72, fp.width - x - w + 72, fp.height - y - h + 72)
    return arr, b


def dmvr_leg(lib, fp, cur, ref, steps, warmup, cpu_seconds):
    """Decoder-side MV refinement of the 1080p workload's bi-predicted blocks
    (list 0 = the reference picture, list 1 = the current picture as a second
    reference): one workgroup per 16x16 subblock."""
    import torch
    from xin26x_amd import capi
    arr, b = dmvr_blocks(fp)
    n = len(b)
    sub = np.zeros(n + 1, np.int32)
    s = ctypes.c_void_p()
    capi.check(lib.xg_session_create(ctypes.byref(s), 16), "session")
    try:
        capi.check(lib.xg_dmvr_plan(s, arr, n, sub.ctypes.data), "dmvr plan")
        nsub = int(sub[-1])
        mvd = torch.empty(2 * nsub, dtype=torch.int16, device="cuda")
        p0, p1 = ref.pic.plane[0], cur.pic.plane[0]
        run = lambda: capi.check(lib.xg_dmvr_run(s, ctypes.byref(p0), ctypes.byref(p1), mvd.data_ptr()), "dmvr run")
        k_ms, wall_ms = session_time(lib, s, 6, run, steps, warmup)
        got = mvd.cpu().numpy()
    finally:
        lib.xg_session_destroy(s)
    sw, sh = np.minimum(b[:, 2], 16), np.minimum(b[:, 3], 16)
    per = (b[:, 2] // sw) * (b[:, 3] // sh)
    alg_bytes = int((per * 2 * (sw + 5) * (sh + 5) * 2).sum()) + nsub * 4    # both lists' regions, one delta out
    k_s = k_ms / 1e3
    traffic, src = profiled_traffic(PMC_DMVR)
    out = {"workload": f"1080p, {n} bi-predicted QTMT blocks (sides >= 8, area >= 128), {nsub} subblocks of up "
                       "to 16x16; list 0 = base MV + seeded 1/16-pel offset into the reference picture, list 1 = "
                       "mirrored MV into the current picture; bilinear prediction, 25 mirrored SADs, parametric "
                       "sub-pel step",
           "value": round(nsub / (wall_ms / 1e3), 1), "unit": "subblocks/s",
           "blocks_per_s": round(n / (wall_ms / 1e3), 1), "ms_per_frame": round(wall_ms, 4),
           "kernel": "k_dmvr", "kernel_ms": round(k_ms, 4),
           "roofline": {"bound": "hbm", "achieved": round(alg_bytes / k_s / 1e9, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(alg_bytes / k_s / 1e9 / HBM_PEAK_GBS, 5),
                        "traffic": traffic, "traffic_source": src, "alg_bytes_per_launch": alg_bytes}}
    if cpu_seconds > 0:
        out["cpu_baseline"] = one_and_all_cores(
            lambda sec, t: dmvr_cpu(fp, arr, b, sub, got, sec, threads=t), cpu_seconds)
    return out


def dmvr_cpu(fp, arr, b, sub, gpu_mvd, seconds, threads=None):
    """The oracle's refinement on the host, chunks of the same blocks; the
    blocks it computes are also compared with the GPU's deltas."""
    from tests.oracle_lib import OrDmvr, load_oracle
    oracle = load_oracle()
    pad = 80
    r0 = np.ascontiguousarray(np.pad(fp.ref[0], pad, mode="edge"))
    r1 = np.ascontiguousarray(np.pad(fp.cur[0], pad, mode="edge"))
    st = r0.shape[1]
    base0, base1 = r0.ctypes.data + 2 * (pad * st + pad), r1.ctypes.data + 2 * (pad * st + pad)
    out = np.zeros(2 * int(sub[-1]), np.int16)
    threads = threads or host_threads()
    chunk, n = 500, len(b)

    def work(i0):
        for i in range(i0, min(n, i0 + chunk)):
            a = arr[i]
            d = OrDmvr(ref0=base0, stride0=st, ref1=base1, stride1=st, cuPelX=a.x, cuPelY=a.y, width=a.width,
                       height=a.height, mv0x=a.mv0x, mv0y=a.mv0y, mv1x=a.mv1x, mv1y=a.mv1y, minMvX=a.minMvX,
                       minMvY=a.minMvY, maxMvX=a.maxMvX, maxMvY=a.maxMvY)
            oracle.or_DeriveMvFromDmvr(ctypes.byref(d), ctypes.c_void_p(out.ctypes.data + 4 * int(sub[i])))
        return i0, min(n, i0 + chunk)

    done, blocks = [], 0
    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        pos = 0
        while time.perf_counter() - t0 < seconds and pos < n:
            starts = [pos + k * chunk for k in range(threads * 2) if pos + k * chunk < n]
            pos += threads * 2 * chunk
            for i0, i1 in ex.map(work, starts):
                done.append((i0, i1))
                blocks += i1 - i0
    dt = time.perf_counter() - t0
    subs = sum(int(sub[i1] - sub[i0]) for i0, i1 in done)
    exact = all(np.array_equal(out[2 * sub[i0]:2 * sub[i1]], gpu_mvd[2 * sub[i0]:2 * sub[i1]]) for i0, i1 in done)
    return {"value": round(subs / dt, 1), "unit": "subblocks/s", "cores": threads, "kind": "port",
            "sample": f"oracle (scalar C) on the first {blocks} blocks ({subs} subblocks) in chunks of {chunk}, "
                      f"{threads} host threads, {dt:.1f} s",
            "bit_exact_vs_gpu_on_sample": bool(exact)}


def search_leg(lib, steps, warmup, cpu_seconds):
    """The plain 416x240 integer full search: 390 16x16 blocks, +/-16, lambda 4."""
    import torch
    from xin26x_amd import capi, workloads as wl
    from xin26x_amd.device import DevicePicture
    cur, ref = wl.frame_pair_416x240()
    W, H = 416, 240
    nb = (W // 16) * (H // 16)
    dc = DevicePicture([cur, cur[::2, ::2].copy(), cur[::2, ::2].copy()], pad=0)
    dr = DevicePicture([ref, ref[::2, ::2].copy(), ref[::2, ::2].copy()], pad=80)
    s = ctypes.c_void_p()
    capi.check(lib.xg_session_create(ctypes.byref(s), 16), "session")
    try:
        mv = torch.zeros(2 * nb, dtype=torch.int16, device="cuda")
        cost = torch.zeros(nb, dtype=torch.int32, device="cuda")
        sad = torch.zeros(nb, dtype=torch.int32, device="cuda")
        run = lambda: capi.check(lib.xg_full_search_frame(s, ctypes.byref(dc.pic.plane[0]),
                                                          ctypes.byref(dr.pic.plane[0]), 16, 16, 16, 1024, 64,
                                                          mv.data_ptr(), cost.data_ptr(), sad.data_ptr()), "search")
        k_ms, wall_ms = session_time(lib, s, 3, run, steps, warmup)
        gmv = mv.cpu().numpy()
    finally:
        lib.xg_session_destroy(s)
    evals = nb * 33 * 33
    out = {"workload": "416x240 10-bit, 390 16x16 blocks, integer full search +/-16, (SAD<<8)+bits*1024",
           "value": round(W * H / (wall_ms / 1e3) / 1e6, 3), "unit": "Mpixel/s", "ms_per_frame": round(wall_ms, 4),
           "sad_evals_per_s": round(evals / (wall_ms / 1e3), 1),
           "kernel": "k_full_search_frame_p<16> (packed v_sad_u16, four adjacent candidates per lane)",
           "kernel_ms": round(k_ms, 4),
           "kernel_abs_diff_ops_per_s": round(evals * 256 / (k_ms / 1e3), 1),
           "sad_roofline": {"achieved": round(evals * 256 / (k_ms / 1e3) / 1e12, 3), "peak": SAD_PAIR_PEAK,
                            "unit": "T abs-diff/s", "frac": round(evals * 256 / (k_ms / 1e3) / 1e12 / SAD_PAIR_PEAK, 4),
                            "note": "390 workgroups on 256 CUs: one wave of workgroups, so the frame is "
                                    "latency-bound (window staging, then 297 items of 16 rows each)"}}
    if cpu_seconds > 0:
        out["cpu_baseline"] = one_and_all_cores(lambda sec, t: search_cpu(cur, ref, gmv, sec, threads=t),
                                                cpu_seconds)
    return out


def search_cpu(cur, ref, gmv, seconds, threads=None):
    """or_FullSearchFrame on whole frames, the frame's blocks split into one
    range per host thread; every frame is compared with the GPU's MVs."""
    from tests.oracle_lib import load_oracle
    oracle = load_oracle()
    threads = threads or host_threads()
    H, W = cur.shape
    nb = (W // 16) * (H // 16)
    refp = np.ascontiguousarray(np.pad(ref, 80, mode="edge"))
    rs = W + 160
    emx, emy = np.zeros(nb, np.int16), np.zeros(nb, np.int16)
    ec, es = np.zeros(nb, np.uint32), np.zeros(nb, np.uint32)
    P = lambda a, off=0: ctypes.c_void_p(a.ctypes.data + off * a.itemsize)
    cuts = [nb * k // threads for k in range(threads + 1)]

    def work(k):
        oracle.or_FullSearchFrame(P(cur), W, P(refp, 80 * rs + 80), rs, W, H, 16, 16, 16, 1024, 64, cuts[k],
                                  cuts[k + 1], P(emx), P(emy), P(ec), P(es))

    frames, exact = 0, True
    t0 = time.perf_counter()
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        while time.perf_counter() - t0 < seconds:
            list(ex.map(work, range(threads)))
            frames += 1
            exact &= bool(np.array_equal(gmv.reshape(nb, 2)[:, 0], emx) and
                          np.array_equal(gmv.reshape(nb, 2)[:, 1], emy))
    dt = time.perf_counter() - t0
    return {"value": round(W * H * frames / dt / 1e6, 3), "unit": "Mpixel/s", "cores": threads, "kind": "port",
            "sample": f"oracle (scalar C, -O3) full frames, the blocks split over {threads} threads, {frames} "
                      f"frames in {dt:.1f} s", "bit_exact_vs_gpu": exact}


SAD_PAIR_PEAK = 256 * 64 * 2.4e9 * 2 / 1e12   # abs-diffs/s: one v_sad_u16 (two samples) per lane per cycle


def pipeline_leg(lib, steps, warmup, cpu_seconds):
    """3840x2160, two references: full search +/-64, predictor + candidates +
    diamond, half/quarter-pel rings, reference choice, 1/16-pel refinement +
    MC, replayed per frame from one captured HIP graph."""
    from xin26x_amd import capi, workloads as wl
    from xin26x_amd.device import DevicePicture, refine_cfg
    REFINE_OFFS = wl.CAND17_SIXTEENTH
    W, H, R, LAM = 3840, 2160, 64, 1024
    log("pipeline: generating the 4K layered scene")
    fr = wl.layered_frames(W, H)
    cur, ref0, ref1 = fr[2], fr[1], fr[0]
    nbx, nby = W // 16, H // 16
    nb = nbx * nby
    evals = 0                                   # SAD evaluations of the full search (windows clipped at the frame)
    for by in range(nby):
        for bx in range(nbx):
            x, y = bx * 16, by * 16
            nx = min(R, W + 56 - x) - max(-R, -x - 72) + 1
            ny = min(R, H + 56 - y) - max(-R, -y - 72) + 1
            evals += 2 * nx * ny
    s = ctypes.c_void_p()
    capi.check(lib.xg_session_create(ctypes.byref(s), nb), "session")
    try:
        dc, d0, d1 = DevicePicture(cur, 80), DevicePicture(ref0, 80), DevicePicture(ref1, 80)
        pred = DevicePicture.empty(W, H)
        cfg = capi.XgPipelineCfg()
        cfg.range, cfg.sadLambda, cfg.ctuSize = R, LAM, 64
        cfg.refine = refine_cfg(REFINE_OFFS, wl.LAMBDA_REFINE)
        run = lambda mode: capi.check(lib.xg_me_pipeline(s, ctypes.byref(cfg), dc.ref(), d0.ref(), d1.ref(),
                                                         pred.ref(), mode), "pipeline")
        log("pipeline: eager frames (per-kernel timing)")
        k_ms, eager_ms = session_time(lib, s, 3, lambda: run(capi.MODE_BATCHED), steps, warmup)
        frame_ms = ctypes.c_float()
        cnt = ctypes.c_int32()
        capi.check(lib.xg_session_kernel_time(s, 4, ctypes.byref(frame_ms), ctypes.byref(cnt)), "frame time")
        log("pipeline: graph replays")
        for _ in range(warmup):
            run(capi.MODE_GRAPH)
        capi.check(lib.xg_session_sync(s), "sync")
        t0 = time.perf_counter()
        for _ in range(steps):
            run(capi.MODE_GRAPH)
        capi.check(lib.xg_session_sync(s), "sync")
        graph_ms = (time.perf_counter() - t0) / steps * 1e3
        q = np.zeros(4 * nb, np.int16)
        capi.check(lib.xg_me_pipeline_read(s, None, None, None, None, q.ctypes.data, None), "read")
        best = np.zeros(nb * 4, np.int32)
        cand = np.zeros(nb * len(REFINE_OFFS), np.uint32)
        capi.check(lib.xg_session_read(s, cand.ctypes.data, best.ctypes.data), "read")
        gpu = {"q": q, "best": best.reshape(nb, 4), "cand": cand, "pred": pred.planes()}
    finally:
        lib.xg_session_destroy(s)
    k_s = k_ms / 1e3
    absdiff = evals * 256
    out = {"workload": "3840x2160 10-bit 4:2:0, layered scene (pan (+5.5,+2.25) px + 12 moving rectangles, "
                       "MVs up to +/-64 px, occlusion), 2 references, 32,400 16x16 blocks: full search +/-64, "
                       "median predictor + candidates + diamond, half/quarter-pel rings, reference choice, "
                       f"{len(REFINE_OFFS)}-candidate 1/16-pel refinement + MC; one HIP graph per frame",
           "value": round(W * H / (graph_ms / 1e3) / 1e6, 2), "unit": "Mpixel/s", "ms_per_frame": round(graph_ms, 3),
           "eager_ms_per_frame": round(eager_ms, 3), "eager_device_ms_per_frame": round(frame_ms.value, 3),
           "sad_evals_per_frame": evals, "kernel": "k_pipe_full", "kernel_ms": round(k_ms, 3),
           "roofline": {"bound": "valu", "achieved": round(absdiff / k_s / 1e12, 2), "peak": round(SAD_PAIR_PEAK, 1),
                        "unit": "T abs-diff/s", "frac": round(absdiff / k_s / 1e12 / SAD_PAIR_PEAK, 4),
                        "note": "v_sad_u16 sums two packed absolute differences per lane and cycle"},
           "median_quarter_mv_ref0": [int(np.median(q.reshape(-1, 2, 2)[:, 0, 0])),
                                      int(np.median(q.reshape(-1, 2, 2)[:, 0, 1]))]}
    if cpu_seconds > 0:
        out["cpu_baseline"] = one_and_all_cores(
            lambda sec, t: pipeline_cpu(cur, ref0, ref1, gpu, sec, threads=t), cpu_seconds)
        out["bit_exact_on_cpu_sample"] = out["cpu_baseline"]["bit_exact_vs_gpu_on_sample"]
    return out


def pipeline_cpu(cur, ref0, ref1, gpu, seconds, threads=None):
    """The oracle's pipeline on bands of whole 16x16 block rows of the same 4K
    frame (or_MePipelineSample: each band computes its rows' stages plus the
    stage-1 searches of the row above), one band per host thread, rounds until
    about `seconds` have passed; every block it finished is compared with the
    GPU's quarter-pel MVs, 1/16 candidate costs, winners and prediction."""
    from tests.oracle_lib import load_oracle
    from tests.pipeline_common import oracle_pipeline_sample
    oracle = load_oracle()
    threads = threads or host_threads()
    H, W = cur[0].shape
    nbx, nby = W // 16, H // 16
    k = 2                                       # block rows per band (and per thread)
    starts = list(range(0, nby - k + 1, k))
    starts = starts[::3] + starts[1::3] + starts[2::3]       # spread each round over the frame
    done, exact, pos = [], True, 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and pos < len(starts):
        rows = [r for r0 in starts[pos:pos + threads] for r in range(r0, r0 + k)]
        nthr = len(starts[pos:pos + threads])
        pos += threads
        sample = np.array([r * nbx + c for r in rows for c in range(nbx)], np.int32)
        e = oracle_pipeline_sample(oracle, cur, ref0, ref1, sample, threads=nthr)
        n = len(sample)
        exact &= bool(np.array_equal(gpu["q"].reshape(-1, 4)[sample], e["q"].reshape(n, 4)))
        exact &= bool(np.array_equal(gpu["cand"].reshape(-1, len(e["cand"]) // n)[sample], e["cand"].reshape(n, -1)))
        b = gpu["best"][sample]
        exact &= bool(np.array_equal(b[:, 0], e["bx"]) and np.array_equal(b[:, 1], e["by"]) and
                      np.array_equal(b[:, 3], e["bi"]))
        for r in rows:
            for p in range(3):
                sh = 1 if p == 0 else 2
                exact &= bool(np.array_equal(gpu["pred"][p][r * 16 // sh:(r + 1) * 16 // sh],
                                             e["pred_pic"][p][r * 16 // sh:(r + 1) * 16 // sh]))
        done += rows
    dt = time.perf_counter() - t0
    px = len(done) * nbx * 256
    return {"value": round(px / dt / 1e6, 4), "unit": "Mpixel/s", "cores": threads, "kind": "port",
            "bit_exact_vs_gpu_on_sample": exact,
            "sample": f"oracle pipeline (scalar C, -O3) on {len(done)} of {nby} block rows ({len(done) * nbx} blocks) "
                      f"of the same 4K frame, bands of {k} rows, one band per thread (each also runs the stage-1 "
                      f"searches of the row above it), {dt:.1f} s; the quarter-pel MVs, 1/16 candidate costs, "
                      f"winners and prediction of those blocks compared with the GPU's"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="default tiles and schedule")
    ap.add_argument("--autotune", action="store_true", help="tune live even when a stored table exists")
    ap.add_argument("--tiles", default=TUNED, help="stored schedule + tile table (written by --save-tiles)")
    ap.add_argument("--save-tiles", default=None, help="write the live-tuned table here")
    ap.add_argument("--quick", action="store_true",
                    help="profiling run: batched form only (no plain form, latency loop or CPU leg)")
    ap.add_argument("--no-extra", action="store_true", help="skip the transform and 416x240 search legs")
    ap.add_argument("--serve-frames", type=int, default=300,
                    help="frames per stream of the 16-stream serving runs (0 = skip serving)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: there is no multi-GPU path (asked for --gpus {args.gpus}). Across GPUs "
              "the work would shard by streams, or by CTU-row bands of one stream.", file=sys.stderr)
        sys.exit(2)

    import torch
    from xin26x_amd import capi, workloads as wl
    from xin26x_amd.device import DevicePicture, blocks_to_struct, refine_cfg

    torch.cuda.set_device(0)
    lib = capi.load()
    if args.calib_copy_mib:
        nbytes = args.calib_copy_mib << 20
        a = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"calibration copy: {nbytes} bytes read, {nbytes} bytes written")
        del a, b
    log("generating the 1080p workload")
    fp = wl.frame_pair_1080p()
    W, H = fp.width, fp.height
    n = len(fp.blocks)
    offs = [tuple(int(v) for v in o) for o in wl.CAND32]
    nc = len(offs)
    cfg = refine_cfg(offs, wl.LAMBDA_REFINE)

    s = ctypes.c_void_p()
    capi.check(lib.xg_session_create(ctypes.byref(s), n), "session")
    cur, ref = DevicePicture(fp.cur, 80), DevicePicture(fp.ref, 80)
    pred = DevicePicture.empty(W, H)
    capi.check(lib.xg_session_set_picture(s, 0, cur.ref()), "picture")
    capi.check(lib.xg_session_set_picture(s, 1, ref.ref()), "picture")
    barr = blocks_to_struct(fp.blocks)
    capi.check(lib.xg_session_set_blocks(s, barr.ctypes.data, n), "blocks")
    tuning = "default"
    if args.no_autotune:
        pass
    elif args.autotune or args.save_tiles or not os.path.exists(args.tiles):
        log("autotuning the schedule and candidates-per-task per block shape")
        path = args.save_tiles.encode() if args.save_tiles else None
        capi.check(lib.xg_session_autotune(s, ctypes.byref(cfg), 2, path), "autotune")
        tuning = "live"
    else:
        capi.check(lib.xg_session_load_tiles(s, args.tiles.encode()), "load tiles")
        tuning = os.path.relpath(args.tiles, ROOT)

    def run(mode, steps):
        for _ in range(steps):
            capi.check(lib.xg_refine_frac(s, ctypes.byref(cfg), pred.ref(), mode), "refine")

    def timed(mode):
        run(mode, args.warmup)
        capi.check(lib.xg_session_sync(s), "sync")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(mode, args.steps)
        capi.check(lib.xg_session_sync(s), "sync")
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps

    def outputs():
        c = np.empty(n * nc, np.uint32)
        b = np.empty(n * 4, np.int32)
        capi.check(lib.xg_session_read(s, c.ctypes.data, b.ctypes.data), "read")
        return c, b, pred.planes()

    if args.quick:
        t_graph = timed(capi.MODE_GRAPH)
        t_direct = timed(capi.MODE_BATCHED)
        t_plain = t_plain_tuned = float("nan")
        exact, lat = None, np.array([float("nan")])
    else:
        # plain form: one launch per block shape, one workgroup per block (all of its candidates)
        shapes = [(lw, lh) for lw in range(2, 7) for lh in range(2, 7)]
        tuned = {sh: lib.xg_session_get_tile(s, *sh) for sh in shapes}
        for sh in shapes:
            capi.check(lib.xg_session_set_tile(s, sh[0], sh[1], nc), "tile")
        log("plain form (one workgroup per block)")
        t_plain = timed(capi.MODE_PLAIN)
        out_plain = outputs()
        for sh in shapes:
            capi.check(lib.xg_session_set_tile(s, sh[0], sh[1], tuned[sh]), "tile")
        log("plain form with the tuned candidates-per-workgroup table")
        t_plain_tuned = timed(capi.MODE_PLAIN)
        log("batched form (graph replay)")
        t_graph = timed(capi.MODE_GRAPH)
        out_graph = outputs()
        log("batched form (direct launch)")
        t_direct = timed(capi.MODE_BATCHED)
        out_direct = outputs()
        exact = all(all(np.array_equal(a, b) for a, b in zip(out_plain[:2], o[:2])) and
                    all(np.array_equal(a, b) for a, b in zip(out_plain[2], o[2])) for o in (out_graph, out_direct))

        # per-frame latency: submit one frame, wait for it (the faster form)
        lat = []
        for _ in range(max(args.steps, 20)):
            t0 = time.perf_counter()
            run(capi.MODE_BATCHED if t_direct < t_graph else capi.MODE_GRAPH, 1)
            capi.check(lib.xg_session_sync(s), "sync")
            lat.append((time.perf_counter() - t0) * 1e3)
        lat = np.array(lat)

    # dominant kernel (fractional refinement) timed with HIP events on the session stream
    capi.check(lib.xg_session_set_timing(s, 1), "timing")
    capi.check(lib.xg_session_reset_timers(s), "timers")
    run(capi.MODE_BATCHED, args.steps)
    ms = ctypes.c_float()
    cnt = ctypes.c_int32()
    capi.check(lib.xg_session_kernel_time(s, 0, ctypes.byref(ms), ctypes.byref(cnt)), "kernel time")
    capi.check(lib.xg_session_set_timing(s, 0), "timing")
    lib.xg_session_destroy(s)

    frame_px = W * H
    pic_bytes = 2 * (W * H + 2 * (W // 2) * (H // 2))
    # current + reference read once; candidate costs, best records (16 B) and the prediction written once
    alg_bytes = 2 * pic_bytes + n * nc * 4 + n * 16 + pic_bytes
    k_s = ms.value / 1e3
    achieved = alg_bytes / k_s / 1e9
    cand_px = int((fp.blocks[:, 2] * fp.blocks[:, 3]).sum()) * nc
    # the same persistent kernel either way: replayed from a captured graph or
    # launched directly; the faster is the value, both are reported
    t_best = min(t_graph, t_direct)
    form = "direct launch" if t_direct < t_graph else "graph replay"
    value = frame_px / t_best / 1e6
    plain = frame_px / t_plain / 1e6
    plain_tuned = frame_px / t_plain_tuned / 1e6

    transforms = reconstruction = residual_loop = search = pipeline = percall = None
    if not args.quick and not args.no_extra:
        log("full motion-search pipeline, 4K, two references")
        pipeline = pipeline_leg(lib, max(5, args.steps // 4), 2, 0 if args.no_cpu else 10.0)
        log("transforms: 200,000 residual blocks")
        transforms = transforms_leg(lib, args.steps, args.warmup, 0 if args.no_cpu else 5.0)
        log("reconstruction: the same 200,000 blocks")
        reconstruction = recon_leg(lib, args.steps, args.warmup, 0 if args.no_cpu else 5.0)
        log("residual loop: forward transform, quantisation, reconstruction of the same blocks")
        residual_loop = residual_loop_leg(lib, args.steps, args.warmup, 0 if args.no_cpu else 3.0)
        log("plain 416x240 full search")
        search = search_leg(lib, args.steps, args.warmup, 0 if args.no_cpu else 3.0)
        log("per-call table slots: latency")
        percall = percall_leg(lib, 2000, 0 if args.no_cpu else 20000)

    serving = None
    if args.serve_frames > 0 and not args.quick:
        from xin26x_amd import serve
        log("serving: generating the 16-stream mix")
        srcs = serve.mix_sources()
        nfr = args.serve_frames
        log(f"serving: {nfr} distinct frames per stream into HBM (as from a GPU decoder)")
        dev = [serve.DeviceFrames(s_, nfr) for s_ in srcs]
        cap = serve.two_chain_cap(srcs)
        log(f"serving: 16 streams at 30 fps, {nfr} frames each")
        rt = serve.run_mix(frames=nfr, fps=30.0, srcs=srcs, mode="device", dev_frames=dev, max_blocks=cap)
        log("serving: saturated (every frame offered at once)")
        sat = serve.run_mix(frames=nfr, fps=0, srcs=srcs, mode="device", dev_frames=dev, max_blocks=cap)
        log("serving: saturated, every stream's frame in one batch")
        sat1 = serve.run_mix(frames=nfr, fps=0, srcs=srcs, mode="device", dev_frames=dev)
        del dev
        torch.cuda.empty_cache()
        # uploads from host memory bind this run, and two chains would only
        # contend for them: every stream's frame in one batch
        log("serving: saturated from page-locked host frames (PCIe-inclusive)")
        pin = serve.run_mix(frames=min(60, nfr), fps=0, srcs=srcs, mode="pinned")
        serving = {"workload": "16 streams (6x 640x360, 6x 1280x720, 4x 1920x1080) 10-bit 4:2:0, "
                               f"{nfr} distinct frames each; per frame: AMVP predictor choice (spatial candidates "
                               "from the stream's resident MV field, its previous frame's final MVs; prediction "
                               "SAD at quarter pel), integer search +/-16 around the predictor, 32-candidate "
                               "fractional refinement + MC; reference = previous frame; the blocks of many "
                               f"streams' frames in shared launches, batches capped at {cap} blocks so the two "
                               "sessions serve disjoint halves of the streams side by side",
                   "max_batch_blocks": cap,
                   "realtime_30fps": rt, "saturated": sat, "saturated_pinned_host": pin,
                   "saturated_one_batch": sat1,
                   "served_vs_plain": round(sat["mpixel_per_s"] / plain, 3) if plain else None,
                   "note": "device mode: frames resident in HBM before the timed region, results to page-locked "
                           "host memory; pinned mode adds the host-to-device upload of every frame (every "
                           "stream's frame in one batch)"}

    dmvr = lookahead = bipred = mctf = None
    if not args.quick and not args.no_extra:
        log("DMVR: the 1080p bi-predicted blocks")
        dmvr = dmvr_leg(lib, fp, cur, ref, args.steps, args.warmup, 0 if args.no_cpu else 3.0)
        log("lookahead motion estimation at half resolution")
        lookahead = lookahead_leg(lib, fp, cur, ref, args.steps, args.warmup, 0 if args.no_cpu else 5.0)
        log("MCTF motion estimation of the 1080p pair")
        mctf = mctf_leg(lib, fp, cur, ref, args.steps, args.warmup, 0 if args.no_cpu else 6.0)
        log("bi-prediction SATD of the 1080p blocks")
        bipred = bipred_leg(lib, fp, cur, ref, pred, args.steps, args.warmup, 0 if args.no_cpu else 3.0)

    traffic, traffic_src = profiled_traffic(PMC_REFINE)

    cpu = None
    if not args.no_cpu and not args.quick:
        log("CPU oracle baseline")
        cpu = one_and_all_cores(lambda sec, t: cpu_baseline(fp, sec, offs, wl.LAMBDA_REFINE, threads=t),
                                args.cpu_seconds)
        if os.path.exists(REF_LIB):
            log("cpu baseline: the reference's own interpolation")
            cpu["reference_interp"] = one_and_all_cores(lambda sec, t: ref_interp_cpu(fp, offs, sec, threads=t),
                                                        args.cpu_seconds / 2)
        if os.path.exists(REF8_LIB):
            log("cpu baseline: the reference's 8-bit SIMD interpolation")
            cpu["reference_interp_simd8"] = one_and_all_cores(
                lambda sec, t: ref_interp_cpu(fp, offs, sec, threads=t, simd8=True), args.cpu_seconds / 2)

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_best * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded smooth-field 10-bit frames, per-CTU motion, QTMT blocks)",
        "config": {"workload": "1080p fractional refinement + MC: 1920x1080 10-bit 4:2:0, "
                               f"{n} QTMT blocks, {nc} fractional candidates/block, 8-tap luma + 4-tap "
                               "chroma interpolation + SATD, MC of the best candidate",
                   "frame": f"{W}x{H}", "blocks": n, "candidates": nc, "global_batch": 1, "tiles": tuning,
                   "parallelism": "single GPU", "form": form},
        "direct_launch_ms_per_step": round(t_direct * 1e3, 4),
        "graph_replay_ms_per_step": round(t_graph * 1e3, 4),
        "plain_value": round(plain, 2),
        "plain_tuned_value": round(plain_tuned, 2),
        "batched_vs_plain": round(value / plain, 3),
        "bit_exact_batched_vs_plain": exact,
        "frame_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                             "p99": round(float(np.percentile(lat, 99)), 4)},
        "block_cand_evals_per_s": round(n * nc / t_best, 1),
        "luma_cand_px_per_s": round(cand_px / t_best, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": traffic_src,
                     "binding": "VALU issue (integer filters + Hadamard): see valu_roofline; HBM traffic is "
                                + (f"{traffic / alg_bytes:.2f}x the algorithmic bytes, " if traffic else "")
                                + "far from its ceiling",
                     "kernel": "k_refine_persistent", "kernel_ms": round(ms.value, 4), "launches": cnt.value,
                     "alg_bytes_per_launch": alg_bytes},
        "valu_roofline": valu_roofline(PMC_REFINE, ms.value),
        "cpu_baseline": cpu,
        "transforms": transforms,
        "reconstruction": reconstruction,
        "residual_loop": residual_loop,
        "dmvr_1080p": dmvr,
        "lookahead_1080p": lookahead,
        "mctf_1080p": mctf,
        "bipred_1080p": bipred,
        "plain_search_416x240": search,
        "per_call_latency": percall,
        "pipeline_4k": pipeline,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
