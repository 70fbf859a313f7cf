"""Batched forward transform on int8 MFMA against the CPU oracle's integer
matrix product: every shape 4..64 x 4..64, DCT-II and the four DST-VII/DCT-VIII
combinations, residual extremes, zero-out regions; bit-exact int32."""
import numpy as np
import pytest

from tests.helpers import ptr
from xin26x_amd import capi, workloads as wl

pytestmark = pytest.mark.gpu


def _oracle_blocks(oracle, desc, offs, resid):
    out = np.zeros(resid.size, np.int32)
    for d, o in zip(desc.tolist(), offs.tolist()):
        w, h, mts = 1 << (d & 15), 1 << ((d >> 4) & 15), (d >> 8) & 15
        r32 = resid[o:o + w * h].astype(np.int32)
        oracle.or_FDctWxH(ptr(r32), w, ptr(out, o), w, w, h, mts, 0, 0, 10)
    return out


def _run(desc, offs, resid):
    import ctypes
    import torch
    lib = capi.load()
    s = ctypes.c_void_p()
    assert lib.xg_session_create(ctypes.byref(s), 16) == 0
    try:
        dres = torch.from_numpy(resid).cuda()
        dco = torch.full((resid.size,), -12345, dtype=torch.int32, device="cuda")
        assert lib.xg_fdct_plan(s, ptr(offs), ptr(desc), len(desc)) == 0
        assert lib.xg_fdct_run(s, dres.data_ptr(), dco.data_ptr()) == 0
        assert lib.xg_session_sync(s) == 0
        return dco.cpu().numpy()
    finally:
        lib.xg_session_destroy(s)


def _layout(shapes, tight=False):
    """Blocks back to back, each offset rounded up to a multiple of 8 samples,
    or with `tight` to the least the batched entry accepts: min(8, width)."""
    desc = np.array([lw | (lh << 4) | (m << 8) for (lw, lh, m) in shapes], np.uint32)
    offs, o = [], 0
    for (lw, lh, _) in shapes:
        a = min(8, 1 << lw) if tight else 8
        o = (o + a - 1) // a * a
        offs.append(o)
        o += (1 << lw) * (1 << lh)
    return desc, np.array(offs, np.int64), o


def test_every_shape_and_type(oracle):
    rng = np.random.default_rng(41)
    shapes = []
    for lw in range(1, 7):
        for lh in range(1, 7):
            for m in (0, 2, 3, 4, 5):
                if m and not (2 <= lw <= 5 and 2 <= lh <= 5):
                    continue
                shapes += [(lw, lh, m)] * 3
    desc, offs, total = _layout(shapes)
    resid = rng.integers(-1023, 1024, total).astype(np.int16)
    # extreme blocks: all +1023 / all -1023 / alternating
    for k, (o, d) in enumerate(zip(offs[:60], desc[:60])):
        n = (1 << (d & 15)) * (1 << ((d >> 4) & 15))
        if k % 3 == 0:
            resid[o:o + n] = 1023
        elif k % 3 == 1:
            resid[o:o + n] = -1023
        else:
            resid[o:o + n] = np.where(np.arange(n) % 2, 1023, -1023)
    exp = _oracle_blocks(oracle, desc, offs, resid)
    got = _run(desc, offs, resid)
    # samples between blocks belong to nobody: the transform leaves them as they were
    covered = np.zeros(total, bool)
    for d, o in zip(desc.tolist(), offs.tolist()):
        covered[o:o + (1 << (d & 15)) * (1 << ((d >> 4) & 15))] = True
    exp[~covered] = -12345
    bad = np.nonzero(got != exp)[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[0]}"


def test_least_aligned_offsets(oracle):
    """Narrow blocks at offsets that are multiples of min(8, width) only, mixed
    with wide ones; many blocks per shape so groups fill up."""
    rng = np.random.default_rng(43)
    kinds = [(1, 1, 0), (1, 6, 0), (2, 2, 2), (2, 6, 0), (1, 3, 0), (6, 1, 0), (3, 6, 0), (5, 5, 3), (6, 6, 0)]
    shapes = [kinds[k] for k in rng.integers(0, len(kinds), 700)]
    desc, offs, total = _layout(shapes, tight=True)
    assert np.any(offs % 8 != 0)
    resid = rng.integers(-1023, 1024, total).astype(np.int16)
    exp = _oracle_blocks(oracle, desc, offs, resid)
    got = _run(desc, offs, resid)
    covered = np.zeros(total, bool)
    for d, o in zip(desc.tolist(), offs.tolist()):
        covered[o:o + (1 << (d & 15)) * (1 << ((d >> 4) & 15))] = True
    exp[~covered] = -12345
    bad = np.nonzero(got != exp)[0]
    assert bad.size == 0, f"{bad.size} mismatches, first at {bad[0]}"


def test_workload_sample(oracle):
    """The 200k-block workload's generator; parity on its first 20k blocks."""
    desc, offs, resid = wl.residual_blocks(n=20000)
    exp = _oracle_blocks(oracle, desc, offs, resid)
    got = _run(desc, offs, resid)
    assert np.array_equal(got, exp)


def test_plan_rejects_bad_descriptors():
    import ctypes
    lib = capi.load()
    s = ctypes.c_void_p()
    assert lib.xg_session_create(ctypes.byref(s), 16) == 0
    try:
        for d in (6 | (6 << 4) | (2 << 8), 7 | (2 << 4), 2 | (2 << 4) | (1 << 8)):
            desc = np.array([d], np.uint32)
            offs = np.zeros(1, np.int64)
            assert lib.xg_fdct_plan(s, ptr(offs), ptr(desc), 1) == capi.XG_ERR_ARG
        # rows move as whole vectors: an 8-wide block must start on a multiple of 8 samples
        desc = np.array([3 | (3 << 4)], np.uint32)
        assert lib.xg_fdct_plan(s, ptr(np.array([4], np.int64)), ptr(desc), 1) == capi.XG_ERR_ARG
        assert lib.xg_fdct_plan(s, ptr(np.array([8], np.int64)), ptr(desc), 1) == 0
    finally:
        lib.xg_clear_error()
        lib.xg_session_destroy(s)
