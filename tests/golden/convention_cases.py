"""Frozen outputs of the build-defined conventions: the functions whose
reference implementation is a closed binary (SATD normalisation, Exp-Golomb
MV bits, the averaged SATD's rounding, the forward and inverse transforms'
rounding and zero-out, the full search's raster order and ties, the S16 SAD's
row subsampling).  These are not reference vectors: they freeze what the
oracle defines (each definition pinned by tests/test_oracle_properties.py), so
that a later change to a convention shows up as a fixture mismatch instead of
passing silently through both the oracle and the kernels.
"""
import ctypes

import numpy as np


def inputs(seed=97):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 1024, size=(64, 80)).astype(np.uint16)
    b = rng.integers(0, 1024, size=(64, 80)).astype(np.uint16)
    c = rng.integers(0, 1024, size=(64, 80)).astype(np.uint16)
    r = rng.integers(-1023, 1024, size=(64, 64)).astype(np.int32)
    q = rng.integers(-3000, 3001, size=(64, 64)).astype(np.int32)
    s16a = rng.integers(0, 1024, size=(16, 16)).astype(np.int16)
    s16b = rng.integers(0, 1024, size=(16, 16)).astype(np.int16)
    return a, b, c, r, q, s16a, s16b


def compute(oracle):
    from tests.helpers import ptr
    from tests.oracle_lib import OrMe
    a, b, c, r, q, s16a, s16b = inputs()
    out = {}
    sizes = (2, 4, 8, 16, 32, 64)
    out["satd"] = np.array([oracle.or_ComputeSatd(ptr(a), 80, ptr(b), 80, w, h) for w in sizes for h in sizes],
                           np.uint32)
    out["avg_satd"] = np.array([oracle.or_ComputeAvgSatd(ptr(a), 80, ptr(b), ptr(c), 80, w, h)
                                for w in sizes[1:] for h in sizes[1:]], np.uint32)
    out["eg_bits"] = np.array([oracle.or_ExpGolombBits(v) for v in range(-300, 301)], np.uint32)
    out["sad_s16"] = np.array([oracle.or_ComputeSadS16(ptr(s16a), 16, ptr(s16b), 16, 16, h, sh)
                               for h in (8, 16) for sh in (0, 1, 2)], np.uint32)
    fw = []
    for mts in (0, 2, 3, 4, 5):
        for (w, h) in ((4, 4), (8, 16), (32, 32), (64, 64), (64, 8), (16, 32)):
            if mts and (w > 32 or h > 32):
                continue
            o = np.zeros((h, w), np.int32)
            oracle.or_FDctWxH(ptr(np.ascontiguousarray(r[:h, :w])), w, ptr(o), w, w, h, mts, 0, 0, 10)
            fw.append(o.reshape(-1))
            i = np.zeros((h, w), np.int32)
            oracle.or_IDctWxH(ptr(np.ascontiguousarray(q[:h, :w])), w, ptr(i), w, w, h, mts, 1, 1, 10)
            fw.append(i.reshape(-1))
    out["transforms"] = np.concatenate(fw)
    pad = np.pad(b, 24, mode="edge")
    me = []
    for (w, h, sw, lam) in ((16, 16, 8, 1024), (8, 8, 12, 0), (32, 16, 6, 4096)):
        m = OrMe()
        m.width, m.height, m.searchWidth, m.searchHeight, m.sadLambda = w, h, sw, sw, lam
        m.minMvX, m.minMvY, m.maxMvX, m.maxMvY = -20, -20, 20, 20
        m.bestCost = 0xFFFFFFFF
        m.input, m.inputStride = a.ctypes.data + 2 * (20 * 80 + 20), 80
        m.ref, m.refStride = pad.ctypes.data + 2 * ((24 + 20) * pad.shape[1] + 24 + 20), pad.shape[1]
        oracle.or_MotionEstFull(ctypes.byref(m))
        me += [m.bestMvX, m.bestMvY, m.bestCost, m.bestSad]
    out["full_search"] = np.array(me, np.int64)
    return out
