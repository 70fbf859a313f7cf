"""Bi-prediction distortion of a frame's blocks (SATD and SAD of the input
against the rounded average of two predictions) against the oracle's
per-block averaged distortions; every block size class, bit-exact."""
import ctypes

import numpy as np
import pytest

from tests.helpers import ptr, rand_pixels
from xin26x_amd import capi
from xin26x_amd.device import BLOCK_DTYPE

pytestmark = pytest.mark.gpu

W, H = 272, 200


def _plane(torch, a):
    t = torch.from_numpy(np.ascontiguousarray(a).view(np.int16)).cuda()
    return t, capi.XgPlane(data=t.data_ptr(), stride=a.shape[1], width=a.shape[1], height=a.shape[0], pad=0)


def _blocks(rng, n):
    sides = (4, 8, 12, 16, 24, 32, 64, 128)
    arr = np.zeros(n, BLOCK_DTYPE)
    for i in range(n):
        w, h = sides[int(rng.integers(0, 8))], sides[int(rng.integers(0, 8))]
        arr[i]["x"] = int(rng.integers(0, W - w + 1))
        arr[i]["y"] = int(rng.integers(0, H - h + 1))
        arr[i]["w"], arr[i]["h"] = w, h
    return arr


def test_bipred_distortion_matches_oracle(oracle):
    import torch
    lib = capi.load()
    rng = np.random.default_rng(71)
    inp, p0, p1 = (rand_pixels(rng, (H, W)) for _ in range(3))
    blocks = _blocks(rng, 500)
    s = ctypes.c_void_p()
    assert lib.xg_session_create(ctypes.byref(s), 16) == 0
    try:
        assert lib.xg_bipred_plan(s, blocks.ctypes.data, len(blocks)) == 0
        (ti, pi), (t0, q0), (t1, q1) = _plane(torch, inp), _plane(torch, p0), _plane(torch, p1)
        for satd in (1, 0):
            dist = torch.zeros(len(blocks), dtype=torch.int32, device="cuda")
            assert lib.xg_bipred_run(s, ctypes.byref(pi), ctypes.byref(q0), ctypes.byref(q1), satd,
                                     dist.data_ptr()) == 0
            assert lib.xg_session_sync(s) == 0
            got = dist.cpu().numpy().view(np.uint32)
            fn = oracle.or_ComputeAvgSatd if satd else oracle.or_ComputeAvgSad
            for i, b in enumerate(blocks):
                o = int(b["y"]) * W + int(b["x"])
                exp = fn(ptr(inp, o), W, ptr(p0, o), ptr(p1, o), W, int(b["w"]), int(b["h"]))
                assert got[i] == exp, (satd, i, b)
        # a block outside a plane is refused before the launch; a bad size at plan time
        small = capi.XgPlane(data=pi.data, stride=W, width=64, height=64, pad=0)
        assert lib.xg_bipred_run(s, ctypes.byref(small), ctypes.byref(q0), ctypes.byref(q1), 1,
                                 dist.data_ptr()) == capi.XG_ERR_ARG
        lib.xg_clear_error()
        bad = _blocks(rng, 1)
        bad[0]["w"] = 6
        assert lib.xg_bipred_plan(s, bad.ctypes.data, 1) == capi.XG_ERR_ARG
        assert lib.xg_bipred_run(s, ctypes.byref(pi), ctypes.byref(q0), ctypes.byref(q1), 1,
                                 dist.data_ptr()) == capi.XG_ERR_STATE
        lib.xg_clear_error()
    finally:
        lib.xg_session_destroy(s)
