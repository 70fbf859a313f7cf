"""Per-call distortion and transform primitives (host pointers, synchronous)
against the CPU oracle on seeded inputs.  Bit-exact."""
import ctypes

import numpy as np
import pytest

from tests.helpers import ptr, rand_pixels
from xin26x_amd import capi

pytestmark = pytest.mark.gpu

SIZES = (1, 2, 4, 8, 16, 32, 64, 128)


@pytest.fixture(scope="module")
def lib():
    l = capi.load()
    l.xg_clear_error()
    yield l
    assert l.xg_last_error() == 0


def test_sad_all_width_classes(lib, oracle):
    rng = np.random.default_rng(11)
    plane_a = rand_pixels(rng, (140, 150))
    plane_b = rand_pixels(rng, (140, 150))
    for w in SIZES:
        for h in SIZES:
            oa, ob = int(rng.integers(0, 150 - w)), int(rng.integers(0, 150 - w))
            ya, yb = int(rng.integers(0, 140 - h)), int(rng.integers(0, 140 - h))
            out = np.zeros(1, np.uint32)
            lib.Xg266ComputeSad(ptr(plane_a, ya * 150 + oa), 150, ptr(plane_b, yb * 150 + ob), 150, w, h, ptr(out))
            exp = oracle.or_ComputeSad(ptr(plane_a, ya * 150 + oa), 150, ptr(plane_b, yb * 150 + ob), 150, w, h)
            assert int(out[0]) == exp, (w, h)


def test_sad_extremes(lib):
    a = np.full((128, 128), 1023, np.uint16)
    b = np.zeros((128, 128), np.uint16)
    out = np.zeros(1, np.uint32)
    lib.Xg266ComputeSad(ptr(a), 128, ptr(b), 128, 128, 128, ptr(out))
    assert int(out[0]) == 128 * 128 * 1023
    lib.Xg266ComputeSad(ptr(a), 128, ptr(a), 128, 64, 32, ptr(out))
    assert int(out[0]) == 0


def test_satd_all_even_sizes(lib, oracle):
    rng = np.random.default_rng(12)
    plane_a = rand_pixels(rng, (140, 150))
    plane_b = rand_pixels(rng, (140, 150))
    for w in SIZES[1:]:
        for h in SIZES[1:]:
            out = np.zeros(1, np.uint32)
            lib.Xg266ComputeSatd(ptr(plane_a, 3 * 150 + 5), 150, ptr(plane_b, 7 * 150 + 1), 150, w, h, ptr(out))
            exp = oracle.or_ComputeSatd(ptr(plane_a, 3 * 150 + 5), 150, ptr(plane_b, 7 * 150 + 1), 150, w, h)
            assert int(out[0]) == exp, (w, h)


def test_sadxn_and_hadamard(lib, oracle):
    rng = np.random.default_rng(13)
    plane = rand_pixels(rng, (200, 200))
    inp = rand_pixels(rng, (64, 64))
    for (w, h) in ((4, 4), (8, 16), (16, 16), (32, 8), (64, 64), (128, 32)):
        inp2 = rand_pixels(rng, (h, w))
        offs = [int(o) for o in rng.integers(0, 200 - 130, 8)]
        rows = [int(o) for o in rng.integers(0, 200 - 66, 8)]
        refs = [plane.ctypes.data + 2 * (rows[k] * 200 + offs[k]) for k in range(8)]
        exp = [oracle.or_ComputeSad(ptr(inp2), w, ctypes.c_void_p(r), 200, w, h) for r in refs]
        arr = (ctypes.c_void_p * 8)(*refs)
        out = np.zeros(8, np.uint32)
        lib.Xg266ComputeSadx8(ptr(inp2), w, ctypes.cast(arr, ctypes.c_void_p), 200, w, h, ptr(out))
        assert out.tolist() == exp
        out5 = np.zeros(5, np.uint32)
        lib.Xg266ComputeSadx5(ptr(inp2), w, *[ctypes.c_void_p(r) for r in refs[:5]], 200, w, h, ptr(out5))
        assert out5.tolist() == exp[:5]
        out3 = np.zeros(3, np.uint32)
        lib.Xg266ComputeSadx3(ptr(inp2), w, *[ctypes.c_void_p(r) for r in refs[:3]], 200, w, h, ptr(out3))
        assert out3.tolist() == exp[:3]
    for k in range(6):
        blk = rand_pixels(rng, (8, 8)) if k else np.full((8, 8), 1023, np.uint16)
        out = np.zeros(1, np.uint32)
        lib.Xg266ComputeHada8x8(ptr(blk), 8, ptr(out))
        assert int(out[0]) == oracle.or_ComputeHada8x8(ptr(blk), 8)


FDCT = [("Xg26xFDct2P2", 0, 2), ("Xg26xFDct2P4", 0, 4), ("Xg26xFDct2P8", 0, 8), ("Xg26xFDct2P16", 0, 16),
        ("Xg26xFDct2P32", 0, 32), ("Xg26xFDct2P64", 0, 64), ("Xg26xFDst7P4", 2, 4), ("Xg26xFDst7P8", 2, 8),
        ("Xg26xFDst7P16", 2, 16), ("Xg26xFDst7P32", 2, 32), ("Xg26xFDct8P4", 1, 4), ("Xg26xFDct8P8", 1, 8),
        ("Xg26xFDct8P16", 1, 16), ("Xg26xFDct8P32", 1, 32)]


@pytest.mark.parametrize("name,typ,n", FDCT)
def test_fdct1d(lib, oracle, name, typ, n):
    rng = np.random.default_rng(n * 7 + typ)
    for line, shift, amp in ((1, 3, 1023), (n, n.bit_length(), 1023), (37, 12, 46000), (64, 7, 46000)):
        inp = rng.integers(-amp, amp + 1, size=(line, n + 3)).astype(np.int32)
        got = np.full((n, line + 5), -7, np.int32)
        exp = np.full((n, line + 5), -7, np.int32)
        getattr(lib, name)(ptr(inp), n + 3, ptr(got), line + 5, line, shift)
        oracle.or_FDct1d(typ, n, ptr(inp), n + 3, ptr(exp), line + 5, line, shift)
        assert np.array_equal(got, exp), (name, line, shift)


def test_fdct2d_and_dispatcher(lib, oracle):
    rng = np.random.default_rng(21)
    for w in (2, 4, 8, 16, 32, 64):
        for h in (2, 4, 8, 16, 32, 64):
            res = rng.integers(-1023, 1024, size=(h, w + 2)).astype(np.int32)
            s1, s2 = w.bit_length() - 1 + 1, h.bit_length() - 1 + 6
            got = np.zeros((h, w + 1), np.int32)
            exp = np.zeros((h, w + 1), np.int32)
            tmp = np.zeros(w * h, np.int32)
            lib.Xg26xFDct2WxH(ptr(res), w + 2, ptr(got), w + 1, w, h, ptr(tmp), s1, s2)
            oracle.or_FDct2WxH(ptr(res), w + 2, ptr(exp), w + 1, w, h, ptr(tmp), s1, s2)
            assert np.array_equal(got, exp), (w, h)
            for mts in (0, 2, 3, 4, 5):
                for intra, en in ((0, 0), (1, 1)):
                    if mts and (w < 4 or h < 4 or w > 32 or h > 32):
                        continue
                    if intra and en and mts == 0 and (w < 4 or h < 4):
                        continue
                    got[:] = 0
                    exp[:] = 0
                    rc = lib.Xg266FDctWxH(ptr(res), w + 2, ptr(got), w + 1, w, h, mts, intra, en)
                    assert rc == 0
                    oracle.or_FDctWxH(ptr(res), w + 2, ptr(exp), w + 1, w, h, mts, intra, en, 10)
                    assert np.array_equal(got, exp), (w, h, mts, intra)


IDCT = [(n.replace("FD", "ID"), t, k) for (n, t, k) in FDCT]


def _coeffs(rng, h, w, amp):
    """Dequantiser-range coefficients: a sparse low-frequency corner with a
    few values at the int16 limits."""
    c = np.zeros((h, w), np.int64)
    c[:] = rng.integers(-amp, amp + 1, size=(h, w))
    c[rng.random((h, w)) < 0.5] = 0
    flat = c.reshape(-1)
    flat[rng.integers(0, flat.size, max(1, flat.size // 32))] = 32767
    flat[rng.integers(0, flat.size, max(1, flat.size // 32))] = -32768
    return c.astype(np.int32)


@pytest.mark.parametrize("name,typ,n", IDCT)
def test_idct1d(lib, oracle, name, typ, n):
    rng = np.random.default_rng(n * 11 + typ)
    for line, skip, shift, amp in ((1, 0, 7, 300), (n, 0, 7, 32768), (37, 5, 10, 4000), (64, 32, 0, 200),
                                   (8, 8, 7, 50)):
        inp = _coeffs(rng, n, line + 3, amp)
        got = np.full((line, n + 5), -7, np.int32)
        exp = np.full((line, n + 5), -7, np.int32)
        getattr(lib, name)(ptr(inp), line + 3, ptr(got), n + 5, line, skip, shift)
        oracle.or_IDct1d(typ, n, ptr(inp), line + 3, ptr(exp), n + 5, line, skip, shift)
        assert np.array_equal(got, exp), (name, line, skip, shift)


def test_idct2d_skip_and_dispatcher(lib, oracle):
    rng = np.random.default_rng(23)
    for w in (2, 4, 8, 16, 32, 64):
        for h in (2, 4, 8, 16, 32, 64):
            c = _coeffs(rng, h, w + 2, 2000)
            got = np.full((h, w + 1), -7, np.int32)
            exp = np.full((h, w + 1), -7, np.int32)
            tmp = np.zeros(w * h, np.int32)
            lib.Xg266IDct2WxH(ptr(c), w + 2, ptr(got), w + 1, w, h, ptr(tmp), 7, 10)
            oracle.or_IDct2WxH(ptr(c), w + 2, ptr(exp), w + 1, w, h, ptr(tmp), 7, 10)
            assert np.array_equal(got, exp), (w, h)
            for mts in (0, 1, 2, 3, 4, 5):
                for intra, en in ((0, 0), (1, 1)):
                    if mts > 1 and (w < 4 or h < 4 or w > 32 or h > 32):
                        continue
                    if intra and en and mts == 0 and (w < 4 or h < 4):
                        continue
                    got[:] = -7
                    exp[:] = -7
                    rc = lib.Xg266IDctWxH(ptr(c), w + 2, ptr(got), w + 1, w, h, mts, intra, en)
                    assert rc == 0
                    assert oracle.or_IDctWxH(ptr(c), w + 2, ptr(exp), w + 1, w, h, mts, intra, en, 10) == 0
                    assert np.array_equal(got, exp), (w, h, mts, intra)
    for w in (1, 2, 128):
        for h in (1, 16, 128):
            c = rng.integers(-2 ** 31, 2 ** 31, size=(h, w), dtype=np.int64).astype(np.int32)
            got = np.zeros((h, w), np.int32)
            exp = np.zeros((h, w), np.int32)
            lib.Xg266ISkipWxH(ptr(c), w, ptr(got), w, w, h)
            oracle.or_ISkipWxH(ptr(c), w, ptr(exp), w, w, h)
            assert np.array_equal(got, exp), (w, h)
    c = np.zeros((64, 64), np.int32)
    for (w, h, mts) in ((1, 16, 0), (16, 1, 2), (64, 8, 2), (8, 8, 6), (128, 8, 0)):
        assert lib.Xg266IDctWxH(ptr(c), 64, ptr(c), 64, w, h, mts, 0, 0) == capi.XG_ERR_ARG
        lib.xg_clear_error()


def test_block_recon_and_idct_add(lib, oracle):
    rng = np.random.default_rng(24)
    pred = rand_pixels(rng, (64, 70))
    for w in (2, 4, 8, 16, 32, 64):
        for h in (2, 4, 8, 16, 32, 64):
            res = rng.integers(-1500, 1500, size=(h, w + 3)).astype(np.int32)
            got = np.full((h, w + 1), 7, np.uint16)
            exp = np.full((h, w + 1), 7, np.uint16)
            lib.Xg26xBlockRecon(ptr(res), w + 3, ptr(pred), 70, ptr(got), w + 1, w, h)
            oracle.or_BlockRecon(ptr(res), w + 3, ptr(pred), 70, ptr(exp), w + 1, w, h, 10)
            assert np.array_equal(got, exp), (w, h)
            c = _coeffs(rng, h, w, 3000)
            for mts in (0, 1, 3):
                if mts == 3 and (w < 4 or h < 4 or w > 32 or h > 32):
                    continue
                got[:] = 7
                exp[:] = 7
                assert lib.Xg266IDctAdd(ptr(c), w, ptr(pred), 70, ptr(got), w + 1, w, h, mts, 1, 1) == 0
                assert oracle.or_IDctAdd(ptr(c), w, ptr(pred), 70, ptr(exp), w + 1, w, h, mts, 1, 1, 10) == 0
                assert np.array_equal(got, exp), (w, h, mts)


def test_block_sub(lib, oracle):
    rng = np.random.default_rng(22)
    a = rand_pixels(rng, (64, 70))
    b = rand_pixels(rng, (64, 70))
    got = np.zeros((64, 64), np.int32)
    exp = np.zeros((64, 64), np.int32)
    lib.Xg26xBlockSubForDct(ptr(a), 70, ptr(b), 70, ptr(got), 64, 32, 16)
    oracle.or_BlockSubForDct(ptr(a), 70, ptr(b), 70, ptr(exp), 64, 32, 16)
    assert np.array_equal(got, exp)


def test_avg_distortions(lib, oracle):
    """Averaged SAD / SATD of bi-prediction (every width class, stride-128
    prediction buffers as the reference passes them) and the x8 form."""
    rng = np.random.default_rng(14)
    plane = rand_pixels(rng, (140, 150))
    pa = rand_pixels(rng, (128, 128))
    pb = rand_pixels(rng, (128, 128))
    for w in SIZES[2:]:
        for h in SIZES[2:]:
            out = np.zeros(1, np.uint32)
            args = (ptr(plane, 5 * 150 + 3), 150, ptr(pa), ptr(pb), 128, w, h)
            lib.Xg266ComputeAvgSad(*args, ptr(out))
            assert int(out[0]) == oracle.or_ComputeAvgSad(*args), ("sad", w, h)
            lib.Xg266ComputeAvgSatd(*args, ptr(out))
            assert int(out[0]) == oracle.or_ComputeAvgSatd(*args), ("satd", w, h)
    refs_a = [int(pa.ctypes.data + 2 * (r * 128 + c)) for r, c in rng.integers(0, 60, size=(8, 2))]
    refs_b = [int(pb.ctypes.data + 2 * (r * 128 + c)) for r, c in rng.integers(0, 60, size=(8, 2))]
    arr = lambda r: ctypes.cast((ctypes.c_void_p * 8)(*r), ctypes.c_void_p)
    for (w, h) in ((8, 8), (16, 32), (64, 64), (4, 4)):
        got, exp = np.zeros(8, np.uint32), np.zeros(8, np.uint32)
        lib.Xg266ComputeAvgSadx8(ptr(plane), 150, arr(refs_a), arr(refs_b), 128, w, h, ptr(got))
        oracle.or_ComputeAvgSadx8(ptr(plane), 150, arr(refs_a), arr(refs_b), 128, w, h, ptr(exp))
        assert got.tolist() == exp.tolist(), (w, h)


def test_quant_inv_quant_reference_golden(lib):
    """Xg266QuantInvQuant against the reference's own Xin266QuantInvQuant
    outputs (tests/golden/quant_golden.npz): q, r and the coefficient-group map."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import quant_cases as qc
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "quant_golden.npz"))
    t = np.ascontiguousarray(g["t"])
    q, r, maps = qc.replay(qc.resolve(lib, "Xg266QuantInvQuant"), g["cases"], t)
    assert np.array_equal(q, g["q"]) and np.array_equal(r, g["r"])
    assert np.array_equal(maps, g["maps"])


def test_func_table(lib):
    t = capi.XgFuncStruct()
    assert lib.Xg266FuncInit(ctypes.byref(t), 1) == 0
    addr = lambda name: ctypes.cast(getattr(lib, name), ctypes.c_void_p).value
    assert t.pfXinComputeSad[4] == addr("Xg266ComputeSad")
    assert t.pfXinComputeSatd[0] is None and t.pfXinComputeSatd[1] is None
    assert t.pfXinComputeDist[3] == addr("Xg266ComputeSatd")
    # [!!fx][!!fy][lgW], flattened: index = fx*16 + fy*8 + lgW
    assert t.pfXinLumaInterp[0 * 16 + 0 * 8 + 3] == addr("Xg266InterpCopy")
    assert t.pfXinLumaInterp[1 * 16 + 0 * 8 + 3] == addr("Xg266LumaInterpHor")
    assert t.pfXinLumaInterp[0 * 16 + 1 * 8 + 3] == addr("Xg266LumaInterpVet")
    assert t.pfXinLumaInterp[1 * 16 + 1 * 8 + 3] == addr("Xg266LumaInterpHorVet")
    assert t.pfXinFor1dDct[0 * 8 + 6] == addr("Xg26xFDct2P64")
    assert t.pfXinFor1dDct[2 * 8 + 2] == addr("Xg26xFDst7P4")
    assert t.pfXinFor1dDct[1 * 8 + 6] is None
    assert t.pfXinComputeAvgSad[0] is None and t.pfXinComputeAvgSatd[1] is None
    assert t.pfXinComputeAvgSad[2] == addr("Xg266ComputeAvgSad")
    assert t.pfXinComputeAvgDist[4] == addr("Xg266ComputeAvgSatd")
    assert t.pfXinComputeAvgSadx8[0] == addr("Xg266ComputeAvgSadx8")
    assert t.pfXinQuantInvQuant[6] == addr("Xg266QuantInvQuant") and t.pfXinQuantInvQuant[7] is None
    assert t.pfXinInv1dDct[0 * 8 + 6] == addr("Xg26xIDct2P64") and t.pfXinInv1dDct[0 * 8 + 0] is None
    assert t.pfXinInv1dDct[1 * 8 + 5] == addr("Xg26xIDct8P32") and t.pfXinInv1dDct[1 * 8 + 6] is None
    assert t.pfXinInv1dDct[2 * 8 + 2] == addr("Xg26xIDst7P4") and t.pfXinInv1dDct[2 * 8 + 1] is None
    assert t.pfXinInv2dDct2[3 * 8 + 5] == addr("Xg266IDct2WxH") and t.pfXinInv2dDct2[7 * 8 + 3] is None
    assert t.pfXinInv2dSkip[7] == addr("Xg266ISkipWxH") and t.pfXinInv2dSkip[0] == addr("Xg266ISkipWxH")
    assert t.pfXinBlockAddForDct[1] == addr("Xg26xBlockRecon") and t.pfXinBlockAddForDct[6] is not None
    assert t.pfXinBlockAddForDct[0] is None and t.pfXinBlockAddForDct[7] is None
    assert lib.Xg266FuncInit(ctypes.byref(t), 0) == 0
    assert t.pfXinComputeDist[3] == addr("Xg266ComputeSad")
    assert t.pfXinComputeAvgDist[4] == addr("Xg266ComputeAvgSad")
