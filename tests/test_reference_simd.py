"""The reference's 8-bit SIMD interpolation, as the CPU baseline drives it.

oracle/_ref/libxin_ref8.so is the reference's default build (8-bit, _X86_OPT_)
of h266_core/source/h266_luma_mc.c, h266_chroma_mc.c and the SSSE3 / AVX2
files beside them, compiled from their sources by oracle/Makefile, with
oracle/ref_interp.c dispatching through the encoder's tables
(h266_encoder/source/h266_func_init.c:424-458 and :550-584 for C, :824-849
and :891-913 for SSSE3, :1182-1197 for AVX2).  Timing that path is only a fair
baseline if the tables are filled right; this checks that the SIMD tables
give the same samples as the C entries of the same build for every width
class, block height and fractional phase.  Skipped where the reference
sources were not present to build the library.
"""
import ctypes
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_ref", "libxin_ref8.so")

pytestmark = pytest.mark.skipif(not os.path.isfile(LIB), reason="oracle/_ref/libxin_ref8.so not built")

W, H, PAD = 256, 192, 80


def _planes(seed):
    rng = np.random.default_rng(seed)
    y = rng.integers(0, 256, (H + 2 * PAD, W + 2 * PAD), dtype=np.uint8)
    u = rng.integers(0, 256, (H // 2 + PAD, W // 2 + PAD), dtype=np.uint8)
    v = rng.integers(0, 256, (H // 2 + PAD, W // 2 + PAD), dtype=np.uint8)
    return y, u, v


def _predict(lib, planes, level, x, y, w, h, mvx, mvy):
    """One candidate of one block through the tables at `level`; returns the
    luma and both chroma predictions."""
    assert lib.ref_SetDispatch(level) >= 0
    P = lambda a, off=0: ctypes.c_void_p(a.ctypes.data + off * a.itemsize)
    py, pu, pv = planes
    blk = np.array([x, y, w, h, mvx, mvy, 0, 0], np.int32)
    zero = np.zeros(1, np.int16)
    oy, ou, ov = np.zeros(64 * 64, np.uint8), np.zeros(32 * 32, np.uint8), np.zeros(32 * 32, np.uint8)
    rsy, rsc = W + 2 * PAD, W // 2 + PAD
    px = lib.ref_RefineInterpShare(P(py, PAD * rsy + PAD), P(pu, (PAD // 2) * rsc + PAD // 2),
                                   P(pv, (PAD // 2) * rsc + PAD // 2), ctypes.c_ssize_t(rsy), ctypes.c_ssize_t(rsc),
                                   W, H, P(blk), 1, 1, P(zero), P(zero), 64, P(oy), P(ou), P(ov))
    assert px == w * h
    return (oy.reshape(64, 64)[:h, :w].copy(), ou.reshape(32, 32)[:h // 2, :w // 2].copy(),
            ov.reshape(32, 32)[:h // 2, :w // 2].copy())


@pytest.fixture(scope="module")
def lib():
    lib = ctypes.CDLL(LIB)
    lib.ref_RefineInterpShare.restype = ctypes.c_int64
    lib.ref_SetDispatch.restype = ctypes.c_int32
    yield lib
    lib.ref_SetDispatch(0)


@pytest.mark.parametrize("w", [4, 8, 16, 32, 64])
def test_simd_tables_match_c_entries(lib, w):
    planes = _planes(w)
    level = lib.ref_SetDispatch(2)
    assert level >= 1, "host CPU without SSSE3: the SIMD tables are not exercised"
    for h in (4, 8, 16, 64):
        # Luma phases 0..15 in each direction; the chroma phase (1/32) runs
        # over 16..31 and 0..15 as the block moves by half a luma sample.
        for fx in range(16):
            for fy in range(16):
                mvx, mvy = 16 * 3 + fx + (16 if fy & 1 else 0), -16 * 2 + fy
                simd = _predict(lib, planes, level, 64, 32, w, h, mvx, mvy)
                plain = _predict(lib, planes, 0, 64, 32, w, h, mvx, mvy)
                for name, a, b in zip(("Y", "Cb", "Cr"), simd, plain):
                    assert np.array_equal(a, b), f"{name} {w}x{h} mv ({mvx}, {mvy}) differs from the C entries"


def test_dispatch_level_is_capped_by_the_host(lib):
    assert lib.ref_SetDispatch(0) == 0
    assert lib.ref_SetDispatch(2) in (0, 1, 2)
