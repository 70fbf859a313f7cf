"""HIP interpolation (per-call C-ABI) against golden vectors written by the
reference C: every luma phase pair x filterIndex, every chroma phase pair,
every table width class 1..128, S16 variants, clipping extremes, bi-average
and BCW weighting.  Bit-exact."""
import numpy as np
import pytest

import golden_cases as gc
from xin26x_amd import capi

pytestmark = pytest.mark.gpu


def test_capi_interp_matches_reference_golden(golden):
    lib = capi.load()
    lib.xg_clear_error()
    planes = (golden["src_u16"], golden["src_s16"], golden["src2_s16"])
    cases = golden["cases"]
    out_u, out_s, offs = gc.replay(gc.resolver(lib, 2), cases, planes)
    assert lib.xg_last_error() == 0
    bad = np.nonzero((out_u != golden["out_u16"]) | (out_s != golden["out_s16"]))[0]
    if bad.size:
        i = int(np.searchsorted(offs, bad[0], side="right") - 1)
        raise AssertionError(f"{bad.size} mismatching samples; first in case {i}: "
                             f"{gc.KEYS[int(cases[i][0])]} params={cases[i].tolist()}")


def test_capi_interp_rejects_oversize():
    lib = capi.load()
    lib.xg_clear_error()
    src = np.zeros((300, 300), np.uint16)
    dst = np.zeros((300, 300), np.uint16)
    lib.Xg266LumaInterpHorVet(src.ctypes.data + 2 * (4 * 300 + 4), 300, dst.ctypes.data, 300, 0x55, 129, 8, 0)
    assert lib.xg_last_error() == capi.XG_ERR_ARG
    lib.xg_clear_error()


def test_capi_interp_out_of_range_filter_index_is_default():
    """Any filterIndex other than 1 (alternative half-pel) or 2 (4x4 filter)
    selects the default filter (h266_core/source/h266_luma_mc.c:117-128),
    including values past INT32_MAX."""
    lib = capi.load()
    lib.xg_clear_error()
    rng = np.random.default_rng(3)
    src = rng.integers(0, 1024, (40, 40)).astype(np.uint16)
    p = src.ctypes.data + 2 * (8 * 40 + 8)
    want = np.zeros((16, 16), np.uint16)
    lib.Xg266LumaInterpHorVet(p, 40, want.ctypes.data, 16, (8 << 4) | 8, 16, 16, 0)
    for fi in (3, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF):
        got = np.zeros((16, 16), np.uint16)
        lib.Xg266LumaInterpHorVet(p, 40, got.ctypes.data, 16, (8 << 4) | 8, 16, 16, fi)
        assert np.array_equal(got, want), fi
        s16 = np.zeros((16, 16), np.int16)
        w16 = np.zeros((16, 16), np.int16)
        lib.Xg266LumaInterpHorVetU8S16(p, 40, w16.ctypes.data, 16, (8 << 4) | 8, 16, 16, 0)
        lib.Xg266LumaInterpHorVetU8S16(p, 40, s16.ctypes.data, 16, (8 << 4) | 8, 16, 16, fi)
        assert np.array_equal(s16, w16), fi
    assert lib.xg_last_error() == 0
