"""Interpolation golden cases: the case table and a replay driver.

The same table is replayed three ways:
  * by make_golden.py through the xin26x reference C built from its own sources
    (oracle/_ref/libxin_ref.so) to write interp_golden.npz;
  * by tests/test_oracle_golden.py through the CPU oracle;
  * by tests/test_gpu_interp.py through the HIP library's C-ABI.

Every function replayed has the XinMotionInterp / XinMotionInterpS16 shape
(h266_encoder/include/h266_func_struct.h:130-148):
    f(src, srcStride, dst, dstStride, frac, width, height, filterIndex)
or, for the bi-prediction combine, XinMotionInterpAvg/Weight (:150-173).
"""
import ctypes

import numpy as np

# key -> (reference symbol, oracle symbol, C-ABI symbol, input kind, output kind, chroma)
FUNCS = {
    "LumaCopy":          ("Xin266InterpCopy",            "or_InterpCopy",            "Xg266InterpCopy",            "u16", "u16", False),
    "LumaHor":           ("Xin266LumaInterpHor",         "or_LumaInterpHor",         "Xg266LumaInterpHor",         "u16", "u16", False),
    "LumaVet":           ("Xin266LumaInterpVet",         "or_LumaInterpVet",         "Xg266LumaInterpVet",         "u16", "u16", False),
    "LumaHorVet":        ("Xin266LumaInterpHorVet",      "or_LumaInterpHorVet",      "Xg266LumaInterpHorVet",      "u16", "u16", False),
    "LumaCopyS16":       ("Xin266InterpCopyU8S16",       "or_InterpCopyU8S16",       "Xg266InterpCopyU8S16",       "u16", "s16", False),
    "LumaHorS16":        ("Xin266LumaInterpHorU8S16",    "or_LumaInterpHorU8S16",    "Xg266LumaInterpHorU8S16",    "u16", "s16", False),
    "LumaVetS16":        ("Xin266LumaInterpVetU8S16",    "or_LumaInterpVetU8S16",    "Xg266LumaInterpVetU8S16",    "u16", "s16", False),
    "LumaHorVetS16":     ("Xin266LumaInterpHorVetU8S16", "or_LumaInterpHorVetU8S16", "Xg266LumaInterpHorVetU8S16", "u16", "s16", False),
    "LumaVetS16U8":      ("Xin266LumaInterpVetS16U8",    "or_LumaInterpVetS16U8",    "Xg266LumaInterpVetS16U8",    "s16", "u16", False),
    "LumaVetS16S16":     ("Xin266LumaInterpVetS16S16",   "or_LumaInterpVetS16S16",   "Xg266LumaInterpVetS16S16",   "s16", "s16", False),
    "ChromaHor":         ("Xin266ChromaInterpHor",       "or_ChromaInterpHor",       "Xg266ChromaInterpHor",       "u16", "u16", True),
    "ChromaVet":         ("Xin266ChromaInterpVet",       "or_ChromaInterpVet",       "Xg266ChromaInterpVet",       "u16", "u16", True),
    "ChromaHorVet":      ("Xin266ChromaInterpHorVet",    "or_ChromaInterpHorVet",    "Xg266ChromaInterpHorVet",    "u16", "u16", True),
    "ChromaHorS16":      ("Xin266ChromaInterpHorU8S16",  "or_ChromaInterpHorU8S16",  "Xg266ChromaInterpHorU8S16",  "u16", "s16", True),
    "ChromaVetS16":      ("Xin266ChromaInterpVetU8S16",  "or_ChromaInterpVetU8S16",  "Xg266ChromaInterpVetU8S16",  "u16", "s16", True),
    "ChromaHorVetS16":   ("Xin266ChromaInterpHorVetU8S16", "or_ChromaInterpHorVetU8S16", "Xg266ChromaInterpHorVetU8S16", "u16", "s16", True),
    "ChromaVetS16U8":    ("Xin266ChromaInterpVetS16U8",  "or_ChromaInterpVetS16U8",  "Xg266ChromaInterpVetS16U8",  "s16", "u16", True),
    "ChromaVetS16S16":   ("Xin266ChromaInterpVetS16S16", "or_ChromaInterpVetS16S16", "Xg266ChromaInterpVetS16S16", "s16", "s16", True),
    "InterpAvg":         ("Xin266InterpAvgS16U8",        "or_InterpAvgS16U8",        "Xg266InterpAvgS16U8",        "s16x2", "u16", False),
    "InterpWeight":      ("Xin266InterpWeightS16U8",     "or_InterpWeightS16U8",     "Xg266InterpWeightS16U8",     "s16x2", "u16", False),
}
KEYS = list(FUNCS)

PLANE_W, PLANE_H = 160, 96     # source planes; block origin inside (ox, oy)
OUT_STRIDE = 136               # dst stride used for every case (> 128)
# BCW weight pairs (8 - w, w) for bcwWeights = {-2, 3, 4, 5, 10}
# (h266_core/source/h266_common_data.c:36-39, h266_inter_prediction.c:422-431)
BCW_PAIRS = [(10, -2), (5, 3), (4, 4), (3, 5), (-2, 10)]


def make_planes(seed=20261202):
    rng = np.random.default_rng(seed)
    u = rng.integers(0, 1024, size=(PLANE_H, PLANE_W), dtype=np.int64)
    u[10:30, 20:40] = 1023               # flat extremes: clipping at both ends
    u[40:60, 60:80] = 0
    yy, xx = np.mgrid[0:PLANE_H, 0:PLANE_W]
    chk = ((yy + xx) & 1) * 1023          # checkerboard: maximal filter overshoot
    u[60:90, 100:150] = chk[60:90, 100:150]
    s = rng.integers(-14330, 14315, size=(PLANE_H, PLANE_W), dtype=np.int64)
    s[5:25, 5:25] = 14314
    s[30:50, 120:140] = -14330
    s2 = rng.integers(-14330, 14315, size=(PLANE_H, PLANE_W), dtype=np.int64)
    return u.astype(np.uint16), s.astype(np.int16), s2.astype(np.int16)


def build_cases():
    """Rows of (key, frac, w, h, fi, ox, oy, wa, wb)."""
    rows = []
    K = KEYS.index
    o = [0]

    def origin(w, h, halo):
        # walk origins over the plane so different cases see different data
        o[0] += 1
        max_x = PLANE_W - w - halo - 4
        max_y = PLANE_H - h - halo - 4
        ox = 4 + (o[0] * 37) % max(1, max_x - 4)
        oy = 4 + (o[0] * 23) % max(1, max_y - 4)
        return ox, oy

    # every phase pair, every filterIndex, small blocks
    for fi in (0, 1, 2):
        for fy in range(16):
            for fx in range(16):
                frac = (fy << 4) | fx
                for key in ("LumaHorVet", "LumaHorVetS16"):
                    rows.append((K(key), frac, 8, 4, fi) + origin(8, 4, 8) + (0, 0))
        for ph in range(16):
            for key in ("LumaHor", "LumaHorS16"):
                rows.append((K(key), ph, 8, 4, fi) + origin(8, 4, 8) + (0, 0))
            for key in ("LumaVet", "LumaVetS16", "LumaVetS16U8", "LumaVetS16S16"):
                rows.append((K(key), ph << 4, 8, 4, fi) + origin(8, 4, 8) + (0, 0))
    for key in ("LumaCopy", "LumaCopyS16"):
        rows.append((K(key), 0, 8, 4, 0) + origin(8, 4, 8) + (0, 0))
    for fy in range(32):
        for fx in range(32):
            frac = (fy << 5) | fx
            for key in ("ChromaHorVet", "ChromaHorVetS16"):
                rows.append((K(key), frac, 4, 2, 0) + origin(4, 2, 4) + (0, 0))
    for ph in range(32):
        for key in ("ChromaHor", "ChromaHorS16"):
            rows.append((K(key), ph, 4, 2, 0) + origin(4, 2, 4) + (0, 0))
        for key in ("ChromaVet", "ChromaVetS16", "ChromaVetS16U8", "ChromaVetS16S16"):
            rows.append((K(key), ph << 5, 4, 2, 0) + origin(4, 2, 4) + (0, 0))
    # shape sweep (every table width class 1..128) at a few fractional positions
    luma_keys = ("LumaCopy", "LumaHor", "LumaVet", "LumaHorVet", "LumaCopyS16", "LumaHorS16",
                 "LumaVetS16", "LumaHorVetS16", "LumaVetS16U8", "LumaVetS16S16")
    chroma_keys = ("ChromaHor", "ChromaVet", "ChromaHorVet", "ChromaHorS16", "ChromaVetS16",
                   "ChromaHorVetS16", "ChromaVetS16U8", "ChromaVetS16S16")
    shapes = [(w, h) for w in (1, 2, 4, 8, 16, 32, 64, 128) for h in (1, 4, 64)]
    shapes += [(w, h) for w in (4, 16) for h in (2, 8, 16, 32)]
    for (w, h) in shapes:
        for frac, fi in (((5 << 4) | 11, 0), ((8 << 4) | 8, 1), ((3 << 4) | 13, 2)):
            for key in luma_keys:
                rows.append((K(key), frac, w, h, fi) + origin(w, h, 8) + (0, 0))
        if w <= 64 and h <= 64:
            for frac in ((7 << 5) | 19, (16 << 5) | 16):
                for key in chroma_keys:
                    rows.append((K(key), frac, w, h, 0) + origin(w, h, 4) + (0, 0))
    for (w, h) in [(4, 4), (8, 8), (16, 4), (64, 64), (128, 8), (1, 16), (2, 2)]:
        rows.append((K("InterpAvg"), 0, w, h, 0) + origin(w, h, 0) + (0, 0))
        for (wa, wb) in BCW_PAIRS:
            rows.append((K("InterpWeight"), 0, w, h, 0) + origin(w, h, 0) + (wa, wb))
    return np.array(rows, dtype=np.int64)


def _ptr(a, off=0):
    return ctypes.c_void_p(a.ctypes.data + off * a.itemsize)


def replay(resolve, cases, planes):
    """Run every case through resolve(key) -> ctypes function; return
    (out_u16, out_s16, offsets) with each case's w*h outputs packed row-major."""
    u, s, s2 = planes
    total = int(sum(int(c[2]) * int(c[3]) for c in cases))
    out_u = np.zeros(total, np.uint16)
    out_s = np.zeros(total, np.int16)
    offs = np.zeros(len(cases) + 1, np.int64)
    pos = 0
    for i, c in enumerate(cases):
        key = KEYS[int(c[0])]
        frac, w, h, fi, ox, oy, wa, wb = (int(v) for v in c[1:])
        _, _, _, kin, kout, _ = FUNCS[key]
        fn = resolve(key)
        dst = np.zeros((h, OUT_STRIDE), np.uint16 if kout == "u16" else np.int16)
        if kin == "s16x2":
            a0 = _ptr(s, oy * PLANE_W + ox)
            a1 = _ptr(s2, oy * PLANE_W + ox)
            if key == "InterpAvg":
                fn(a0, PLANE_W, a1, PLANE_W, _ptr(dst), OUT_STRIDE, w, h)
            else:
                fn(a0, PLANE_W, a1, PLANE_W, _ptr(dst), OUT_STRIDE, w, h, wa, wb)
        else:
            src = u if kin == "u16" else s
            fn(_ptr(src, oy * PLANE_W + ox), PLANE_W, _ptr(dst), OUT_STRIDE, frac, w, h, fi)
        flat = dst[:, :w].reshape(-1)
        if kout == "u16":
            out_u[pos:pos + w * h] = flat
        else:
            out_s[pos:pos + w * h] = flat
        pos += w * h
        offs[i + 1] = pos
    return out_u, out_s, offs


def resolver(lib, column):
    """column 0: reference symbols, 1: oracle symbols, 2: C-ABI symbols."""
    cache = {}

    def get(key):
        if key not in cache:
            spec = FUNCS[key]
            fn = getattr(lib, spec[column])
            fn.restype = None
            if spec[3] == "s16x2":
                args = [ctypes.c_void_p, ctypes.c_ssize_t, ctypes.c_void_p, ctypes.c_ssize_t,
                        ctypes.c_void_p, ctypes.c_ssize_t, ctypes.c_uint32, ctypes.c_uint32]
                if key == "InterpWeight":
                    args += [ctypes.c_int32, ctypes.c_int32]
                fn.argtypes = args
            else:
                fn.argtypes = [ctypes.c_void_p, ctypes.c_ssize_t, ctypes.c_void_p, ctypes.c_ssize_t,
                               ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
            cache[key] = fn
        return cache[key]
    return get
