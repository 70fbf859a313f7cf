"""Shared set-up of the motion-search pipeline tests: a small layered scene
and the oracle's run of the pipeline on it."""
import ctypes

import numpy as np

from tests.helpers import ptr
from tests.oracle_lib import OrRefinePrm
from xin26x_amd import workloads as wl

PAD = 80
RANGE = 64
SAD_LAMBDA = 1024
REFINE_OFFS = wl.CAND17_SIXTEENTH


def scene(w=256, h=128, seed=7, nobj=3, max_mv=12):
    fr = wl.layered_frames(w, h, seed=seed, nobj=nobj, max_mv=max_mv, sides=(32, 96))
    return fr[2], fr[1], fr[0]          # current, ref0 (previous frame), ref1 (two back)


def oracle_pipeline(oracle, cur, ref0, ref1, rng=RANGE, lam=SAD_LAMBDA, offs=REFINE_OFFS, lam_ref=wl.LAMBDA_REFINE):
    H, W = cur[0].shape
    nb = (W // 16) * (H // 16)
    pads = [[np.ascontiguousarray(np.pad(p, PAD if i == 0 else PAD // 2, mode="edge")) for i, p in enumerate(r)]
            for r in (ref0, ref1)]
    rsy, rsc = W + 2 * PAD, W // 2 + PAD
    at = lambda a, i: ptr(a, (PAD if i == 0 else PAD // 2) * (rsy if i == 0 else rsc) + (PAD if i == 0 else PAD // 2))
    offx = np.ascontiguousarray(np.array([o[0] for o in offs], np.int16))
    offy = np.ascontiguousarray(np.array([o[1] for o in offs], np.int16))
    prm = OrRefinePrm(len(offs), offx.ctypes.data, offy.ctypes.data, lam_ref, 64, 1)
    out = {k: np.zeros(4 * nb, np.int16) for k in ("full", "pred", "int", "q")}
    out.update({k: np.zeros(2 * nb, np.uint32) for k in ("intCost", "qCost")})
    out["ref"] = np.zeros(nb, np.int32)
    out["cand"] = np.zeros(nb * len(offs), np.uint32)
    out["bi"], out["bc"] = np.zeros(nb, np.int32), np.zeros(nb, np.uint32)
    out["bx"], out["by"] = np.zeros(nb, np.int32), np.zeros(nb, np.int32)
    out["pred_pic"] = [np.zeros((H, W), np.uint16), np.zeros((H // 2, W // 2), np.uint16),
                       np.zeros((H // 2, W // 2), np.uint16)]
    P = out["pred_pic"]
    oracle.or_MePipeline(ptr(cur[0]), ptr(cur[1]), ptr(cur[2]), W, W // 2,
                         at(pads[0][0], 0), at(pads[0][1], 1), at(pads[0][2], 2),
                         at(pads[1][0], 0), at(pads[1][1], 1), at(pads[1][2], 2), rsy, rsc,
                         W, H, 16, 16, rng, lam, 64, ctypes.byref(prm),
                         ptr(out["full"]), ptr(out["pred"]), ptr(out["int"]), ptr(out["intCost"]), ptr(out["q"]),
                         ptr(out["qCost"]), ptr(out["ref"]), ptr(out["cand"]), ptr(out["bi"]), ptr(out["bc"]),
                         ptr(out["bx"]), ptr(out["by"]), ptr(P[0]), ptr(P[1]), ptr(P[2]), W, W // 2)
    return out


def oracle_pipeline_sample(oracle, cur, ref0, ref1, sample, rng=RANGE, lam=SAD_LAMBDA, offs=REFINE_OFFS,
                           lam_ref=wl.LAMBDA_REFINE, threads=8):
    """The oracle's pipeline for the 16x16 blocks `sample` only
    (or_MePipelineSample), split over `threads` host threads; outputs indexed
    by sample position in the layout of oracle_pipeline's, and the prediction
    planes filled at the sampled blocks only."""
    from concurrent.futures import ThreadPoolExecutor
    H, W = cur[0].shape
    sample = np.ascontiguousarray(np.asarray(sample, np.int32))
    n = len(sample)
    pads = [[np.ascontiguousarray(np.pad(p, PAD if i == 0 else PAD // 2, mode="edge")) for i, p in enumerate(r)]
            for r in (ref0, ref1)]
    rsy, rsc = W + 2 * PAD, W // 2 + PAD
    at = lambda a, i: ptr(a, (PAD if i == 0 else PAD // 2) * (rsy if i == 0 else rsc) + (PAD if i == 0 else PAD // 2))
    offx = np.ascontiguousarray(np.array([o[0] for o in offs], np.int16))
    offy = np.ascontiguousarray(np.array([o[1] for o in offs], np.int16))
    prm = OrRefinePrm(len(offs), offx.ctypes.data, offy.ctypes.data, lam_ref, 64, 1)
    out = {k: np.zeros(4 * n, np.int16) for k in ("full", "pred", "int", "q")}
    out.update({k: np.zeros(2 * n, np.uint32) for k in ("intCost", "qCost")})
    out["ref"] = np.zeros(n, np.int32)
    out["cand"] = np.zeros(n * len(offs), np.uint32)
    out["bi"], out["bc"] = np.zeros(n, np.int32), np.zeros(n, np.uint32)
    out["bx"], out["by"] = np.zeros(n, np.int32), np.zeros(n, np.int32)
    out["pred_pic"] = [np.zeros((H, W), np.uint16), np.zeros((H // 2, W // 2), np.uint16),
                       np.zeros((H // 2, W // 2), np.uint16)]
    P = out["pred_pic"]
    nc = len(offs)

    def part(i0, i1):
        oracle.or_MePipelineSample(
            ptr(cur[0]), ptr(cur[1]), ptr(cur[2]), W, W // 2,
            at(pads[0][0], 0), at(pads[0][1], 1), at(pads[0][2], 2),
            at(pads[1][0], 0), at(pads[1][1], 1), at(pads[1][2], 2), rsy, rsc,
            W, H, 16, 16, rng, lam, 64, ctypes.byref(prm), ptr(sample, i0), i1 - i0,
            ptr(out["full"], 4 * i0), ptr(out["pred"], 4 * i0), ptr(out["int"], 4 * i0), ptr(out["intCost"], 2 * i0),
            ptr(out["q"], 4 * i0), ptr(out["qCost"], 2 * i0), ptr(out["ref"], i0), ptr(out["cand"], nc * i0),
            ptr(out["bi"], i0), ptr(out["bc"], i0), ptr(out["bx"], i0), ptr(out["by"], i0),
            ptr(P[0]), ptr(P[1]), ptr(P[2]), W, W // 2)

    step = max(1, (n + threads - 1) // threads)
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda i0: part(i0, min(n, i0 + step)), range(0, n, step)))
    return out
