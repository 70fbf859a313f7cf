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


def test_pipeline_graph_follows_tile_schedule_and_block_changes(oracle):
    """A captured pipeline graph bakes in the refinement's grid and task lists:
    after a tile change, a schedule change and a foreign block list on the same
    session, the next GRAPH-mode run re-captures and still matches the oracle."""
    cur, ref0, ref1 = scene()
    exp = oracle_pipeline(oracle, cur, ref0, ref1)
    lib = capi.load()
    H, W = cur[0].shape
    nb = (W // 16) * (H // 16)
    s = ctypes.c_void_p()
    assert lib.xg_session_create(ctypes.byref(s), nb + 64) == 0
    try:
        dc, d0, d1 = DevicePicture(cur, 80), DevicePicture(ref0, 80), DevicePicture(ref1, 80)
        pred = DevicePicture.empty(W, H)
        cfg = capi.XgPipelineCfg()
        cfg.range, cfg.sadLambda, cfg.ctuSize = RANGE, SAD_LAMBDA, 64
        cfg.refine = refine_cfg(REFINE_OFFS, wl.LAMBDA_REFINE)
        from xin26x_amd.device import blocks_to_struct
        other = blocks_to_struct(wl.block_grid(W, H, 32, 32))

        def run():
            pred.tensors[0].zero_()
            assert lib.xg_me_pipeline(s, ctypes.byref(cfg), dc.ref(), d0.ref(), d1.ref(), pred.ref(),
                                      capi.MODE_GRAPH) == 0
            o = {k: np.zeros(4 * nb, np.int16) for k in ("full", "pred", "int", "q")}
            o.update({k: np.zeros(2 * nb, np.uint32) for k in ("intCost", "qCost")})
            assert lib.xg_me_pipeline_read(s, o["full"].ctypes.data, o["pred"].ctypes.data, o["int"].ctypes.data,
                                           o["intCost"].ctypes.data, o["q"].ctypes.data, o["qCost"].ctypes.data) == 0
            cand = np.zeros(nb * len(REFINE_OFFS), np.uint32)
            best = np.zeros(nb * 4, np.int32)
            assert lib.xg_session_read(s, cand.ctypes.data, best.ctypes.data) == 0
            o["cand"], o["best"], o["pred_pic"] = cand, best.reshape(nb, 4), pred.planes()
            check(o, exp)

        run()
        assert lib.xg_session_set_tile(s, 4, 4, 3) == 0          # 16x16 blocks: 3 candidates per task
        run()
        assert lib.xg_session_set_sched(s, 2, 5) == 0            # other bands and workgroups per CU
        run()
        assert lib.xg_session_set_blocks(s, other.ctypes.data, len(other)) == 0
        run()                                                     # the 16x16 grid is installed again
        run()
    finally:
        lib.xg_session_destroy(s)


def sample_blocks(full, nbx, nby, seed=5, n_edge=700, n_inner=700):
    """Border blocks, blocks on motion edges (a stage-1 MV differs from a
    4-neighbour's) and interior blocks, seeded."""
    rng = np.random.default_rng(seed)
    f = full.reshape(nby, nbx, 4)
    border = [b for b in range(nbx * nby) if b % nbx in (0, nbx - 1) or b // nbx in (0, nby - 1)]
    diff = np.zeros((nby, nbx), bool)
    diff[:, 1:] |= np.any(f[:, 1:] != f[:, :-1], axis=2)
    diff[:, :-1] |= np.any(f[:, :-1] != f[:, 1:], axis=2)
    diff[1:] |= np.any(f[1:] != f[:-1], axis=2)
    diff[:-1] |= np.any(f[:-1] != f[1:], axis=2)
    bset = set(border)
    edge = [b for b in np.flatnonzero(diff.ravel()).tolist() if b not in bset]
    edge = rng.choice(edge, min(n_edge, len(edge)), replace=False).tolist() if edge else []
    rest = sorted(set(range(nbx * nby)) - bset - set(edge))
    inner = rng.choice(rest, min(n_inner, len(rest)), replace=False).tolist()
    return np.array(sorted(border + edge + inner), np.int32), len(border), len(edge), len(inner)


def test_pipeline_4k_two_references_sampled(oracle):
    """The config-4 workload itself: 3840x2160, two references, 12 objects with
    motion up to +/-64 px and occlusion, full search +/-64.  The GPU runs the
    whole frame; every stage of >= 2,000 sampled blocks (all border blocks,
    blocks on motion edges, interior blocks) is compared with the oracle's
    pipeline for those blocks, which reads only each block's own and three
    neighbours' stage-1 searches."""
    from tests.pipeline_common import oracle_pipeline_sample
    fr = wl.layered_frames()
    cur, ref0, ref1 = fr[2], fr[1], fr[0]
    H, W = cur[0].shape
    nbx, nby = W // 16, H // 16
    print(f"4K pipeline: {W}x{H}, {nbx * nby} blocks, range {RANGE}", flush=True)
    got = gpu_pipeline(cur, ref0, ref1, capi.MODE_GRAPH)[0]
    sample, nbord, nedge, ninner = sample_blocks(got["full"], nbx, nby)
    print(f"sample: {len(sample)} blocks ({nbord} border, {nedge} motion edge, {ninner} interior)", flush=True)
    assert len(sample) >= 2000 and nedge >= 500
    exp = oracle_pipeline_sample(oracle, cur, ref0, ref1, sample, threads=16)
    n = len(sample)
    for k, width in (("full", 4), ("pred", 4), ("int", 4), ("q", 4), ("intCost", 2), ("qCost", 2)):
        assert np.array_equal(got[k].reshape(-1, width)[sample], exp[k].reshape(n, width)), k
    nc = len(REFINE_OFFS)
    assert np.array_equal(got["cand"].reshape(-1, nc)[sample], exp["cand"].reshape(n, nc))
    b = got["best"][sample]
    assert np.array_equal(b[:, 0], exp["bx"]) and np.array_equal(b[:, 1], exp["by"])
    assert np.array_equal(b[:, 2].view(np.uint32), exp["bc"]) and np.array_equal(b[:, 3], exp["bi"])
    for p in range(3):
        s = 1 if p == 0 else 2
        for blk in sample.tolist():
            x, y = (blk % nbx) * 16 // s, (blk // nbx) * 16 // s
            assert np.array_equal(got["pred_pic"][p][y:y + 16 // s, x:x + 16 // s],
                                  exp["pred_pic"][p][y:y + 16 // s, x:x + 16 // s]), (p, blk)
