"""DMVR bilinear golden cases: the case table and a replay driver.

Every call has the XinBiliInterp shape (h266_encoder/include/h266_func_struct.h:150-157):
    f(src, srcStride, dst, dstStride, frac, width, height)
writing width + 4 columns of int16 per row (the DMVR padding).  The table
slot is picked as the encoder does (h266_inter_prediction.c:1061-1069):
[!!fracX][!!fracY] -> Copy, Hor, Vet, HorVet (h266_func_init.c:496-530).  The
same table is replayed by make_bili_golden.py through the reference C
(oracle/_ref), by tests/test_oracle_golden.py through the oracle and by
tests/test_gpu_dmvr.py through the HIP library.
"""
import ctypes

import numpy as np

NAMES = ("BiliInterpCopy", "BiliInterpHor", "BiliInterpVet", "BiliInterpHorVet")
PLANE_H, PLANE_W = 176, 208


def make_plane(seed=31):
    """10-bit samples with flat 0 and 1023 patches and a checkerboard patch."""
    rng = np.random.default_rng(seed)
    p = rng.integers(0, 1024, size=(PLANE_H, PLANE_W)).astype(np.uint16)
    p[0:40, 140:200] = 1023
    p[40:80, 140:200] = 0
    yy, xx = np.mgrid[80:120, 140:200]
    p[80:120, 140:200] = np.where((yy + xx) % 2, 1023, 0)
    return p


def kind(fx, fy):
    return (1 if fx else 0) + (2 if fy else 0)


def build_cases():
    """Rows (kind, frac, width, height, x0, y0): every phase pair on 8 x (8+4);
    every width 1..128 at a few phase pairs; the flat and checkerboard patches."""
    cases = []
    for fy in range(16):
        for fx in range(16):
            cases.append((kind(fx, fy), (fy << 4) | fx, 8, 12, 3, 5))
    for w in (1, 2, 4, 8, 16, 32, 64, 128):
        for (fx, fy) in ((0, 0), (7, 0), (0, 9), (5, 11), (15, 15)):
            for h in (4, 20):
                cases.append((kind(fx, fy), (fy << 4) | fx, w, h, 2 if w <= 64 else 4, 7))
    for (x0, y0) in ((150, 4), (150, 44), (150, 84)):
        for (fx, fy) in ((0, 0), (8, 0), (0, 8), (3, 13)):
            cases.append((kind(fx, fy), (fy << 4) | fx, 32, 20, x0, y0))
    return np.array(cases, np.int32)


def replay(fns, cases, plane):
    """Run fns[kind] over every case; returns the outputs back to back and their offsets."""
    sizes = [int(h) * (int(w) + 4) for (_, _, w, h, _, _) in cases.tolist()]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    out = np.zeros(int(offs[-1]), np.int16)
    base = plane.ctypes.data
    for i, (k, frac, w, h, x0, y0) in enumerate(cases.tolist()):
        src = ctypes.c_void_p(base + 2 * (y0 * PLANE_W + x0))
        dst = ctypes.c_void_p(out.ctypes.data + 2 * int(offs[i]))
        fns[k](src, PLANE_W, dst, w + 4, frac, w, h)
    return out, offs


def resolve(lib, prefix):
    fns = []
    for n in NAMES:
        f = getattr(lib, prefix + n)
        f.restype = None
        f.argtypes = [ctypes.c_void_p, ctypes.c_ssize_t, ctypes.c_void_p, ctypes.c_ssize_t, ctypes.c_int32,
                      ctypes.c_uint32, ctypes.c_uint32]
        fns.append(f)
    return fns
