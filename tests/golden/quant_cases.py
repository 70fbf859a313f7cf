"""Quantisation golden cases: the case table and a replay driver.

Every call has the XinQuantInvQuant shape (h266_encoder/include/h266_func_struct.h:181-196):
    f(qCoeff, tCoeff, rCoeff, coeffStride, width, height, cgWidth, cgHeight,
      qMult, qAdd, qShift, iqMult, iqAdd, iqShift, &nzCGBitMap)
The parameters follow the caller's derivation (h266_encoder/source/h266_trans_quant.c:100-120)
from the scale tables qMult / iqMult (h266_encoder/source/h266_encoder_create.c:94-105) and
the shifts of Xin266ConstructQuantParam (:1677-1706) at 10 bits (XIN_QP_SHIFT 12,
h266_encoder/include/h266_definition.h:127); the width and height passed are at most 32
(h266_trans_quant.c:212-213).  The same table is replayed by make_quant_golden.py through
the reference C (oracle/_ref), by tests/test_oracle_golden.py through the oracle and by
tests/test_gpu_primitives.py through the HIP library.
"""
import ctypes

import numpy as np

Q_MULT = ((26214, 23302, 20560, 18396, 16384, 14564), (18396, 16384, 14564, 13107, 11651, 10280))
IQ_MULT = ((40, 45, 51, 57, 64, 72), (57, 64, 72, 80, 90, 102))

# (width, height, cgWidth, cgHeight) of the coefficient area passed in
SHAPES = [(4, 4, 4, 4), (8, 8, 4, 4), (16, 16, 4, 4), (32, 32, 4, 4), (4, 8, 4, 4), (8, 4, 4, 4),
          (16, 4, 4, 4), (32, 8, 4, 4), (8, 32, 4, 4), (32, 16, 4, 4), (2, 8, 2, 8), (8, 2, 8, 2),
          (2, 16, 2, 8), (1, 16, 1, 16), (16, 1, 16, 1), (4, 32, 4, 4)]
QPS = (0, 17, 22, 27, 32, 37, 51)


def params(qp, lg_w, lg_h, rdoq, intra):
    """(qMult, qAdd, qShift, iqMult, iqAdd, iqShift) for a transform block of
    2^lg_w x 2^lg_h at QP qp (8-bit scale), non-skip MTS."""
    q = qp + 12
    lg_tu = lg_w + lg_h
    adj, lg_size = lg_tu & 1, lg_tu >> 1
    q_shift = 14 + (15 - 10) + q // 6 - lg_size - adj
    iq_shift = 6 - (15 - 10) + lg_size + adj
    q_add = (1 << (q_shift - 1)) if rdoq else ((171 if intra else 85) << (q_shift - 9))
    return Q_MULT[adj][q % 6], q_add, q_shift, IQ_MULT[adj][q % 6] << (q // 6), 1 << (iq_shift - 1), iq_shift


def build_cases(seed=23):
    """Case table rows: w, h, cgw, cgh, stride, offset, qMult, qAdd, qShift, iqMult, iqAdd, iqShift."""
    rng = np.random.default_rng(seed)
    rows, off = [], 0
    for (w, h, cgw, cgh) in SHAPES:
        lw, lh = max(0, int(np.log2(w))), max(0, int(np.log2(h)))
        for qp in QPS:
            for rdoq, intra in ((0, 0), (0, 1), (1, 0)):
                stride = w + int(rng.integers(0, 5))
                rows.append((w, h, cgw, cgh, stride, off) + params(qp, lw, lh, rdoq, intra))
                off += stride * h
    return np.array(rows, np.int64), off


def make_inputs(total, seed=24):
    """Transform coefficients: mostly Laplacian-sized, some at the int16 limits, some zeros."""
    rng = np.random.default_rng(seed)
    t = np.rint(rng.laplace(0, 300, total)).astype(np.int64)
    big = rng.random(total) < 0.05
    t[big] = rng.choice([-32768, 32767, -20000, 25000], int(big.sum()))
    t[rng.random(total) < 0.2] = 0
    return np.clip(t, -32768, 32767).astype(np.int32)


def replay(fn, cases, t):
    """Run fn over every case; returns q, r (same layout as t) and the per-case CG maps."""
    q = np.zeros_like(t)
    r = np.zeros_like(t)
    maps = np.zeros(len(cases), np.uint64)
    P = lambda a, o: ctypes.c_void_p(a.ctypes.data + 4 * int(o))
    for i, (w, h, cgw, cgh, stride, off, qm, qa, qs, iqm, iqa, iqs) in enumerate(cases.tolist()):
        m = ctypes.c_uint64(0)
        fn(P(q, off), P(t, off), P(r, off), stride, w, h, cgw, cgh, qm, qa, qs, iqm, iqa, iqs, ctypes.byref(m))
        maps[i] = m.value
    return q, r, maps


def resolve(lib, name):
    f = getattr(lib, name)
    f.restype = None
    f.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_ssize_t] + [ctypes.c_uint32] * 4 + [ctypes.c_int32] * 6 + \
                 [ctypes.c_void_p]
    return f
