"""The full motion-search pipeline (full search +/-64, median predictor +
candidates + diamond, half/quarter-pel rings, reference choice, 1/16-pel
refinement + MC) against the oracle's composition of the same stages:
every stage's MVs and costs, the final costs, winners and prediction are
bit-exact, eagerly and replayed from the captured graph."""
import ctypes

import numpy as np
import pytest

from tests.pipeline_common import RANGE, REFINE_OFFS, SAD_LAMBDA, oracle_pipeline, scene
from xin26x_amd import capi, workloads as wl
from xin26x_amd.device import DevicePicture, refine_cfg

pytestmark = pytest.mark.gpu


def gpu_pipeline(cur, ref0, ref1, mode, reps=1):
    lib = capi.load()
    H, W = cur[0].shape
    nb = (W // 16) * (H // 16)
    s = ctypes.c_void_p()
    assert lib.xg_session_create(ctypes.byref(s), nb) == 0
    try:
        dc, d0, d1 = DevicePicture(cur, 80), DevicePicture(ref0, 80), DevicePicture(ref1, 80)
        pred = DevicePicture.empty(W, H)
        cfg = capi.XgPipelineCfg()
        cfg.range, cfg.sadLambda, cfg.ctuSize = RANGE, SAD_LAMBDA, 64
        cfg.refine = refine_cfg(REFINE_OFFS, wl.LAMBDA_REFINE)
        outs = []
        for _ in range(reps):
            pred.tensors[0].zero_()
            rc = lib.xg_me_pipeline(s, ctypes.byref(cfg), dc.ref(), d0.ref(), d1.ref(), pred.ref(), mode)
            assert rc == 0, rc
            o = {k: np.zeros(4 * nb, np.int16) for k in ("full", "pred", "int", "q")}
            o.update({k: np.zeros(2 * nb, np.uint32) for k in ("intCost", "qCost")})
            assert lib.xg_me_pipeline_read(s, o["full"].ctypes.data, o["pred"].ctypes.data, o["int"].ctypes.data,
                                           o["intCost"].ctypes.data, o["q"].ctypes.data, o["qCost"].ctypes.data) == 0
            cand = np.zeros(nb * len(REFINE_OFFS), np.uint32)
            best = np.zeros(nb * 4, np.int32)
            assert lib.xg_session_read(s, cand.ctypes.data, best.ctypes.data) == 0
            o["cand"], o["best"], o["pred_pic"] = cand, best.reshape(nb, 4), pred.planes()
            outs.append(o)
        return outs
    finally:
        lib.xg_session_destroy(s)


def check(g, e):
    for k in ("full", "pred", "int", "intCost", "q", "qCost", "cand"):
        assert np.array_equal(g[k], e[k]), k
    assert np.array_equal(g["best"][:, 0], e["bx"]) and np.array_equal(g["best"][:, 1], e["by"])
    assert np.array_equal(g["best"][:, 2].view(np.uint32), e["bc"]) and np.array_equal(g["best"][:, 3], e["bi"])
    for p in range(3):
        assert np.array_equal(g["pred_pic"][p], e["pred_pic"][p]), p


@pytest.mark.parametrize("shape", [(256, 128, 3, 12), (320, 192, 5, 40)])
def test_pipeline_matches_oracle(oracle, shape):
    w, h, nobj, mv = shape
    cur, ref0, ref1 = scene(w, h, seed=7 + nobj, nobj=nobj, max_mv=mv)
    exp = oracle_pipeline(oracle, cur, ref0, ref1)
    got = gpu_pipeline(cur, ref0, ref1, capi.MODE_BATCHED)[0]
    check(got, exp)


def test_pipeline_graph_replays_match(oracle):
    cur, ref0, ref1 = scene()
    exp = oracle_pipeline(oracle, cur, ref0, ref1)
    for got in gpu_pipeline(cur, ref0, ref1, capi.MODE_GRAPH, reps=3):
        check(got, exp)
