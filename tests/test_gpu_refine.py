"""Batched fractional refinement + motion compensation (the 1080p workload)
against the CPU oracle: every candidate cost, every block's best candidate,
MV and cost, and the luma/chroma prediction picture, bit-exact; plain,
batched and graph forms agree; results do not depend on the per-shape tile
table."""
import ctypes

import numpy as np
import pytest

from tests.helpers import ptr
from tests.oracle_lib import OrRefinePrm
from xin26x_amd import capi, workloads as wl
from xin26x_amd.device import DevicePicture, blocks_to_struct, refine_cfg

pytestmark = pytest.mark.gpu
PAD = 80


def oracle_refine(oracle, fp, blocks, offsets, lam, use_chroma=1):
    W, H = fp.width, fp.height
    refp = [np.ascontiguousarray(np.pad(p, PAD if i == 0 else PAD // 2, mode="edge")) for i, p in enumerate(fp.ref)]
    rsy, rsc = W + 2 * PAD, W // 2 + PAD
    blk = np.ascontiguousarray(np.asarray(blocks, dtype=np.int32))
    n, nc = len(blk), len(offsets)
    offx = np.ascontiguousarray(np.array([o[0] for o in offsets], np.int16))
    offy = np.ascontiguousarray(np.array([o[1] for o in offsets], np.int16))
    prm = OrRefinePrm(nc, offx.ctypes.data, offy.ctypes.data, lam, 64, use_chroma)
    cost = np.zeros(n * nc, np.uint32)
    bi, bc = np.zeros(n, np.int32), np.zeros(n, np.uint32)
    bx, by = np.zeros(n, np.int32), np.zeros(n, np.int32)
    pred = [np.zeros((H, W), np.uint16), np.zeros((H // 2, W // 2), np.uint16), np.zeros((H // 2, W // 2), np.uint16)]
    oracle.or_RefineFracBlocks(ptr(fp.cur[0]), ptr(fp.cur[1]), ptr(fp.cur[2]), W, W // 2,
                               ptr(refp[0], PAD * rsy + PAD), ptr(refp[1], (PAD // 2) * rsc + PAD // 2),
                               ptr(refp[2], (PAD // 2) * rsc + PAD // 2), rsy, rsc, W, H,
                               ptr(blk), n, ctypes.byref(prm), ptr(cost), ptr(bi), ptr(bc), ptr(bx), ptr(by),
                               ptr(pred[0]), ptr(pred[1]), ptr(pred[2]), W, W // 2)
    return cost.reshape(n, nc), bi, bc, bx, by, pred


class Gpu:
    def __init__(self, fp, blocks, max_blocks=None):
        self.lib = capi.load()
        self.s = ctypes.c_void_p()
        assert self.lib.xg_session_create(ctypes.byref(self.s), max_blocks or len(blocks)) == 0
        self.cur = DevicePicture(fp.cur, pad=PAD)
        self.ref = DevicePicture(fp.ref, pad=PAD)
        self.pred = DevicePicture.empty(fp.width, fp.height)
        assert self.lib.xg_session_set_picture(self.s, 0, self.cur.ref()) == 0
        assert self.lib.xg_session_set_picture(self.s, 1, self.ref.ref()) == 0
        self.set_blocks(blocks)

    def set_blocks(self, blocks):
        self.blocks = blocks_to_struct(blocks)
        self.n = len(blocks)
        assert self.lib.xg_session_set_blocks(self.s, self.blocks.ctypes.data, self.n) == 0

    def run(self, cfg, mode):
        assert self.lib.xg_refine_frac(self.s, ctypes.byref(cfg), self.pred.ref(), mode) == 0
        cost = np.empty(self.n * cfg.ncand, np.uint32)
        best = np.empty(self.n * 4, np.int32)
        assert self.lib.xg_session_read(self.s, ptr(cost), ptr(best)) == 0
        return cost.reshape(self.n, cfg.ncand), best.reshape(self.n, 4), self.pred.planes()

    def close(self):
        self.lib.xg_session_destroy(self.s)


@pytest.fixture(scope="module")
def frame():
    return wl.frame_pair_1080p()


def _check(got, exp, blocks):
    cost, best, pred = got
    ecost, ebi, ebc, ebx, eby, epred = exp
    bad = np.nonzero((cost != ecost).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} blocks with cost mismatches; first {blocks[bad[0]].tolist()}"
    assert np.array_equal(best[:, 3], ebi)
    assert np.array_equal(best[:, 2].view(np.uint32), ebc)
    assert np.array_equal(best[:, 0], ebx) and np.array_equal(best[:, 1], eby)
    for p in range(3):
        assert np.array_equal(pred[p], epred[p]), f"prediction plane {p}"


def test_refine_1080p_matches_oracle_all_modes(oracle, frame):
    offs = [tuple(int(v) for v in o) for o in wl.CAND32]
    exp = oracle_refine(oracle, frame, frame.blocks, offs, wl.LAMBDA_REFINE)
    g = Gpu(frame, frame.blocks)
    try:
        cfg = refine_cfg(offs, wl.LAMBDA_REFINE)
        for mode in (capi.MODE_BATCHED, capi.MODE_PLAIN, capi.MODE_GRAPH, capi.MODE_GRAPH):
            g.pred.tensors[0].zero_()
            _check(g.run(cfg, mode), exp, frame.blocks)
    finally:
        g.close()


def test_refine_tiles_do_not_change_results(oracle, frame):
    blocks = frame.blocks[:3000]
    offs = [tuple(int(v) for v in o) for o in wl.CAND32]
    exp = oracle_refine(oracle, frame, blocks, offs, wl.LAMBDA_REFINE)
    g = Gpu(frame, blocks)
    try:
        cfg = refine_cfg(offs, wl.LAMBDA_REFINE)
        for cpt in (1, 3, 32, 64):
            for lw in range(2, 7):
                for lh in range(2, 7):
                    assert g.lib.xg_session_set_tile(g.s, lw, lh, cpt) == 0
            _check(g.run(cfg, capi.MODE_BATCHED), exp, blocks)
        assert g.lib.xg_session_autotune(g.s, ctypes.byref(cfg), 1, None) == 0
        _check(g.run(cfg, capi.MODE_GRAPH), exp, blocks)
    finally:
        g.close()


def test_refine_edges_clipping_and_luma_only(oracle, frame):
    """Blocks at every frame edge with MVs far outside the picture (clipped to
    the CU MV window), an irregular candidate set, luma-only cost."""
    W, H = frame.width, frame.height
    rng = np.random.default_rng(5)
    blocks = []
    for (x, y) in ((0, 0), (W - 64, 0), (0, H - 8), (W - 16, H - 8), (960, 0), (0, 512)):
        for (w, h) in ((64, 64), (16, 8), (8, 8), (4, 4), (32, 8)):
            xx, yy = min(x, W - w), min(y, H - h)
            for mv in ((-4000, -4000), (4000, 4000), (-4000, 3000), (0, 0), (37, -93)):
                blocks.append((xx, yy, w, h, mv[0], mv[1], int(rng.integers(-99, 99)), int(rng.integers(-99, 99))))
    blocks = np.array(blocks, np.int64)
    offs = [(int(a), int(b)) for a, b in rng.integers(-64, 65, size=(45, 2))] + [(64, 64), (-64, -64), (0, 0)]
    exp = oracle_refine(oracle, frame, blocks, offs, 12345, use_chroma=0)
    g = Gpu(frame, blocks)
    try:
        cfg = refine_cfg(offs, 12345, use_chroma=0)
        cost, best, _ = g.run(cfg, capi.MODE_BATCHED)
        assert np.array_equal(cost, exp[0])
        assert np.array_equal(best[:, 3], exp[1])
        assert np.array_equal(best[:, 0], exp[3]) and np.array_equal(best[:, 1], exp[4])
        # Blocks at one position overlap, so the prediction picture is checked
        # per group of non-overlapping blocks (one size and MV per position).
        idx = np.arange(len(blocks)).reshape(6, 25)
        for gi in range(25):
            sub = blocks[idx[:, gi]]
            e = oracle_refine(oracle, frame, sub, offs, 12345, use_chroma=0)
            g.set_blocks(sub)
            g.pred.tensors[0].zero_()
            _, _, pred = g.run(cfg, capi.MODE_BATCHED)
            for b in sub:
                x, y, w, h = (int(v) for v in b[:4])
                assert np.array_equal(pred[0][y:y + h, x:x + w], e[5][0][y:y + h, x:x + w])
                for p in (1, 2):
                    assert np.array_equal(pred[p][y // 2:(y + h) // 2, x // 2:(x + w) // 2],
                                          e[5][p][y // 2:(y + h) // 2, x // 2:(x + w) // 2])
    finally:
        g.close()


def test_refine_rejects_bad_inputs(frame):
    lib = capi.load()
    g = Gpu(frame, frame.blocks[:10])
    try:
        cfg = refine_cfg([(65, 0)], 100)
        assert lib.xg_refine_frac(g.s, ctypes.byref(cfg), g.pred.ref(), capi.MODE_BATCHED) == capi.XG_ERR_ARG
        bad = frame.blocks[:1].copy()
        bad[0, 0] = frame.width - 4
        bad[0, 2] = 8
        g.set_blocks(bad)
        cfg = refine_cfg([(1, 1)], 100)
        assert lib.xg_refine_frac(g.s, ctypes.byref(cfg), g.pred.ref(), capi.MODE_BATCHED) == capi.XG_ERR_ARG
        b2 = blocks_to_struct(frame.blocks[:1])
        b2["w"] = 12
        assert lib.xg_session_set_blocks(g.s, b2.ctypes.data, 1) == capi.XG_ERR_ARG
    finally:
        lib.xg_clear_error()
        g.close()


def test_refine_revalidates_after_state_changes(frame):
    """Validation is cached per session state: a later bad block list is still
    rejected, a good one accepted again, and a new prediction picture receives
    the motion compensation (the cache key covers the prediction argument)."""
    lib = capi.load()
    blocks = frame.blocks[:40]
    g = Gpu(frame, blocks)
    try:
        cfg = refine_cfg([tuple(int(v) for v in o) for o in wl.CAND32[:8]], wl.LAMBDA_REFINE)
        _, best1, pred1 = g.run(cfg, capi.MODE_GRAPH)
        _, best1b, _ = g.run(cfg, capi.MODE_GRAPH)          # unchanged state: the cached path
        assert np.array_equal(best1, best1b)
        bad = blocks[:1].copy()
        bad[0, 0] = frame.width - 4
        bad[0, 2] = 8
        g.set_blocks(bad)
        assert lib.xg_refine_frac(g.s, ctypes.byref(cfg), g.pred.ref(), capi.MODE_BATCHED) == capi.XG_ERR_ARG
        lib.xg_clear_error()
        g.set_blocks(blocks)
        other = DevicePicture.empty(frame.width, frame.height)
        assert lib.xg_refine_frac(g.s, ctypes.byref(cfg), other.ref(), capi.MODE_GRAPH) == 0
        best2 = np.empty(g.n * 4, np.int32)
        assert lib.xg_session_read(g.s, None, ptr(best2)) == 0
        assert np.array_equal(best2.reshape(g.n, 4), best1)
        for p, (a, b) in enumerate(zip(other.planes(), pred1)):
            sc = 1 if p == 0 else 2
            for (x, y, w, h) in blocks[:, :4].tolist():
                sub = (slice(y // sc, (y + h) // sc), slice(x // sc, (x + w) // sc))
                assert np.array_equal(a[sub], b[sub]), (p, x, y)
    finally:
        lib.xg_clear_error()
        g.close()
