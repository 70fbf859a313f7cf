"""Batched reconstruction (inverse transform + prediction add) on int8 MFMA
against the CPU oracle: every shape 2..64 x 2..64, DCT-II, the four
DST-VII/DCT-VIII pairs, implicit intra MTS and transform skip, int16
coefficient extremes, garbage in the zero-out region (must not be read),
large mixed batches; bit-exact 10-bit samples."""
import ctypes

import numpy as np
import pytest

from tests.helpers import ptr, rand_pixels
from xin26x_amd import capi, workloads as wl

pytestmark = pytest.mark.gpu


def _nz(typ, n):
    return min(n, 16 if typ else 32)


_MTS = {0: (0, 0), 2: (2, 2), 3: (1, 2), 4: (2, 1), 5: (1, 1)}


def _types(w, h, m, implicit):
    if m == 1:
        return None
    hor, ver = _MTS[m]
    if implicit and m == 0:
        hor = 2 if 4 <= w <= 16 else hor
        ver = 2 if 4 <= h <= 16 else ver
    return hor, ver


def _layout(shapes, tight=False):
    desc = np.array([lw | (lh << 4) | (m << 8) | (imp << 12) for (lw, lh, m, imp) in shapes], np.uint32)
    offs, o = [], 0
    for (lw, lh, _, _) in shapes:
        a = min(8, 1 << lw) if tight else 8
        o = (o + a - 1) // a * a
        offs.append(o)
        o += (1 << lw) * (1 << lh)
    return desc, np.array(offs, np.int64), o


def _coeffs(rng, desc, offs, total, garbage=True):
    """Sparse coefficients in the dequantiser's int16 range, a share at the
    limits; the zero-out region holds garbage (the kernel must not read it)."""
    c = rng.integers(-2000, 2001, total).astype(np.int64)
    c[rng.random(total) < 0.4] = 0
    c[rng.random(total) < 0.01] = 32767
    c[rng.random(total) < 0.01] = -32768
    c = c.astype(np.int16)
    if not garbage:
        for d, o in zip(desc.tolist(), offs.tolist()):
            w, h, m, imp = 1 << (d & 15), 1 << ((d >> 4) & 15), (d >> 8) & 15, (d >> 12) & 1
            t = _types(w, h, m, imp)
            if t is None:
                continue
            blk = c[o:o + w * h].reshape(h, w)
            blk[_nz(t[1], h):, :] = 0
            blk[:, _nz(t[0], w):] = 0
    return c


def _oracle(oracle, coeff, pred, desc, offs, total):
    out = np.full(total, 0xABCD, np.uint16)
    oracle.or_ReconBatch(ptr(coeff), ptr(pred), ptr(offs), ptr(desc), 0, len(desc), ptr(out))
    return out


def _run(desc, offs, coeff, pred, fill=0xABCD):
    import torch
    lib = capi.load()
    s = ctypes.c_void_p()
    assert lib.xg_session_create(ctypes.byref(s), 16) == 0
    try:
        dc = torch.from_numpy(coeff).cuda()
        dp = torch.from_numpy(pred.view(np.int16)).cuda()
        dr = torch.full((coeff.size,), fill - 65536 if fill > 32767 else fill, dtype=torch.int16, device="cuda")
        assert lib.xg_recon_plan(s, ptr(offs), ptr(desc), len(desc)) == 0
        assert lib.xg_recon_run(s, dc.data_ptr(), dp.data_ptr(), dr.data_ptr()) == 0
        assert lib.xg_session_sync(s) == 0
        return dr.cpu().numpy().view(np.uint16)
    finally:
        lib.xg_session_destroy(s)


def _covered(desc, offs, total):
    cov = np.zeros(total, bool)
    for d, o in zip(desc.tolist(), offs.tolist()):
        cov[o:o + (1 << (d & 15)) * (1 << ((d >> 4) & 15))] = True
    return cov


def _check(got, exp, desc, offs, total):
    exp = exp.copy()
    exp[~_covered(desc, offs, total)] = 0xABCD      # samples between blocks stay as they were
    bad = np.nonzero(got != exp)[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[0]}"


def _all_shapes(reps):
    shapes = []
    for lw in range(1, 7):
        for lh in range(1, 7):
            for m in (0, 1, 2, 3, 4, 5):
                if m >= 2 and not (2 <= lw <= 5 and 2 <= lh <= 5):
                    continue
                shapes += [(lw, lh, m, 0)] * reps
                if m == 0:
                    shapes += [(lw, lh, 0, 1)] * reps
    return shapes


@pytest.mark.parametrize("tight", [False, True])
def test_every_shape_and_kernel(oracle, tight):
    rng = np.random.default_rng(51 + tight)
    desc, offs, total = _layout(_all_shapes(3), tight)
    coeff = _coeffs(rng, desc, offs, total)
    pred = rand_pixels(rng, (total,))
    exp = _oracle(oracle, coeff, pred, desc, offs, total)
    got = _run(desc, offs, coeff, pred)
    _check(got, exp, desc, offs, total)


def test_full_groups_and_single_blocks(oracle):
    """65 blocks of one shape (a full group of 64 plus one), and lone blocks."""
    rng = np.random.default_rng(52)
    shapes = [(2, 2, 0, 0)] * 65 + [(1, 1, 0, 0)] * 65 + [(3, 3, 2, 0)] * 33 + [(6, 6, 0, 0)] + [(1, 6, 0, 0)]
    desc, offs, total = _layout(shapes)
    coeff = _coeffs(rng, desc, offs, total)
    pred = rand_pixels(rng, (total,))
    _check(_run(desc, offs, coeff, pred), _oracle(oracle, coeff, pred, desc, offs, total), desc, offs, total)


def test_large_shuffled_batch(oracle):
    """20k blocks of the transform workload's shape mix, mixed with skip and
    implicit-MTS blocks, placed in shuffled order."""
    rng = np.random.default_rng(53)
    d0, _, _ = wl.residual_blocks(n=20000, seed=17)
    m = (d0 >> 8) & 15
    extra = rng.random(d0.size)
    m = np.where((m == 0) & (extra < 0.15), 1, m)
    imp = ((m == 0) & (extra > 0.7)).astype(np.uint32)
    desc = (d0 & 0xFF) | (m << 8) | (imp << 12)
    order = rng.permutation(desc.size)
    sizes = (1 << (desc & 15).astype(np.int64)) * (1 << ((desc >> 4) & 15).astype(np.int64))
    offs = np.zeros(desc.size, np.int64)
    offs[order] = np.concatenate([[0], np.cumsum(sizes[order])[:-1]])
    total = int(sizes.sum())
    coeff = _coeffs(rng, desc, offs, total, garbage=False)
    pred = rand_pixels(rng, (total,))
    _check(_run(desc, offs, coeff, pred), _oracle(oracle, coeff, pred, desc, offs, total), desc, offs, total)


def test_matches_per_call_idct_add(oracle):
    """The batched form equals the per-call Xg266IDctAdd slot on the same TUs."""
    lib = capi.load()
    rng = np.random.default_rng(54)
    shapes = [(3, 4, 0, 1), (5, 5, 3, 0), (6, 5, 0, 0), (2, 2, 1, 0), (4, 3, 4, 0)]
    desc, offs, total = _layout(shapes)
    coeff = _coeffs(rng, desc, offs, total, garbage=False)
    pred = rand_pixels(rng, (total,))
    got = _run(desc, offs, coeff, pred)
    for (lw, lh, m, imp), o in zip(shapes, offs.tolist()):
        w, h = 1 << lw, 1 << lh
        c32 = coeff[o:o + w * h].astype(np.int32)
        one = np.zeros(w * h, np.uint16)
        assert lib.Xg266IDctAdd(ptr(c32), w, ptr(pred, o), w, ptr(one), w, w, h, m, imp, imp) == 0
        assert np.array_equal(one, got[o:o + w * h]), (w, h, m)


def test_plan_rejects_bad_descriptors():
    lib = capi.load()
    s = ctypes.c_void_p()
    assert lib.xg_session_create(ctypes.byref(s), 16) == 0
    try:
        good = np.array([2 | (2 << 4)], np.uint32)
        offs = np.array([0], np.int64)
        assert lib.xg_recon_plan(s, ptr(offs), ptr(good), 1) == 0
        for d, o in ((0 | (2 << 4), 0), (7 | (2 << 4), 0), (6 | (2 << 4) | (2 << 8), 0),
                     (2 | (2 << 4) | (6 << 8), 0), (2 | (2 << 4) | (1 << 13), 0), (3 | (3 << 4), 4),
                     (2 | (2 << 4), -8)):
            assert lib.xg_recon_plan(s, ptr(np.array([o], np.int64)), ptr(np.array([d], np.uint32)), 1) \
                == capi.XG_ERR_ARG, hex(d)
        # a failed plan leaves none behind
        assert lib.xg_recon_run(s, 16, 16, 16) == capi.XG_ERR_STATE
        lib.xg_clear_error()
    finally:
        lib.xg_session_destroy(s)
