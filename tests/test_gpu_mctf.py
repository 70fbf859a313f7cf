"""MCTF motion estimation on the GPU (xg_mctf_subsample, xg_mctf_me_level,
xg_mctf_me) against the oracle (or_MctfSubsample, or_MctfMeLevel), bit-exact:
subsampled pictures with their padding, every level of the hierarchy on the
oracle's own level pictures (candidates from the lower level, the integer
search, the fractional rounds), the whole frame search, and the padding
rule."""
import ctypes

import numpy as np
import pytest

from tests.helpers import ptr
from tests.test_oracle_drivers import MSLACK, oracle_mctf_levels
from xin26x_amd import capi
from xin26x_amd.device import DevicePicture

pytestmark = pytest.mark.gpu
W, H, PAD = 640, 360, 80


def _scene(seed=31):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H + 32, 0:W + 32]
    tex = 512 + 230 * np.sin(xx / 6.0 + yy / 13.0) + 140 * np.cos(yy / 5.0 - xx / 17.0) + rng.normal(0, 30, xx.shape)
    ref = np.clip(tex[16:16 + H, 16:16 + W], 0, 1023)
    cur = np.clip(tex[16 + 5:16 + 5 + H, 16 - 9:16 - 9 + W], 0, 1023)            # moved by (9, -5)
    half = 0.5 * (tex[16 + 2:16 + 2 + H, 16 + 3:16 + 3 + W] + tex[16 + 2:16 + 2 + H, 16 + 4:16 + 4 + W])
    cur[32:128, 32:192] = np.clip(half, 0, 1023)[32:128, 32:192]                  # fractional motion
    cur = np.clip(cur + rng.normal(0, 4, cur.shape), 0, 1023)
    cur[200:296, 400:560] = ref[200:296, 400:560]                                  # still
    ref[180:320, 100:300] = 300                                                    # flat: ties
    cur[200:296, 120:280] = 300
    return np.ascontiguousarray(cur.astype(np.uint16)), np.ascontiguousarray(ref.astype(np.uint16))


class Dev:
    """A device copy of a (possibly padded) host plane as an xg_plane."""

    def __init__(self, torch, arr, pad, w, h):
        self.t = torch.from_numpy(np.ascontiguousarray(arr).view(np.int16)).cuda()
        stride = arr.shape[1]
        off = (arr.shape[0] - h) // 2
        self.plane = capi.XgPlane(self.t.data_ptr() + 2 * (off * stride + off), stride, w, h, pad)


@pytest.fixture(scope="module")
def env():
    import torch
    lib = capi.load()
    s = ctypes.c_void_p()
    assert lib.xg_session_create(ctypes.byref(s), 1) == 0
    yield lib, s, torch
    lib.xg_session_destroy(s)


def _level(lib, s, torch, c, rp, pad, bs, rng, lower, factor, frac):
    h, w = c.shape
    dc = Dev(torch, c, 0, w, h)
    ps = (rp.shape[0] - h) // 2
    dr = Dev(torch, rp, pad, w, h)
    dr.plane.pad = pad
    nbx, nby = w // bs, h // bs
    mv = torch.zeros(2 * nbx * nby, dtype=torch.int32, device="cuda")
    err = torch.zeros(nbx * nby, dtype=torch.int32, device="cuda")
    low = None if lower is None else torch.from_numpy(np.ascontiguousarray(lower.astype(np.int32))).cuda()
    torch.cuda.synchronize()               # torch's fills and copies are done before the session stream reads
    assert lib.xg_mctf_me_level(s, ctypes.byref(dc.plane), ctypes.byref(dr.plane), bs, rng,
                                None if low is None else ctypes.c_void_p(low.data_ptr()),
                                0 if lower is None else lower.shape[1], 0 if lower is None else lower.shape[0],
                                factor, frac, ctypes.c_void_p(mv.data_ptr()), ctypes.c_void_p(err.data_ptr())) == 0
    assert lib.xg_session_sync(s) == 0
    assert ps >= pad
    return mv.cpu().numpy().reshape(nby, nbx, 2), err.cpu().numpy().view(np.uint32).reshape(nby, nbx)


def test_mctf_subsample_matches_oracle(oracle, env):
    lib, s, torch = env
    cur, ref = _scene()
    for pad in (0, 16, 32):
        src = Dev(torch, cur, 0, W, H)
        out = torch.zeros((H // 2 + 2 * pad) * (W // 2 + 2 * pad), dtype=torch.int16, device="cuda")
        stride = W // 2 + 2 * pad
        dst = capi.XgPlane(out.data_ptr() + 2 * (pad * stride + pad), stride, W // 2, H // 2, pad)
        torch.cuda.synchronize()
        assert lib.xg_mctf_subsample(s, ctypes.byref(src.plane), ctypes.byref(dst)) == 0
        assert lib.xg_session_sync(s) == 0
        exp = np.zeros((H // 2 + 2 * pad, stride), np.uint16)
        oracle.or_MctfSubsample(ptr(cur), W, W, H, ptr(exp, pad * stride + pad), stride, pad)
        assert np.array_equal(out.cpu().numpy().view(np.uint16).reshape(exp.shape), exp), pad


def test_mctf_levels_match_oracle(oracle, env):
    """Each level on the oracle's level pictures and lower fields."""
    lib, s, torch = env
    cur, ref = _scene()
    lv = oracle_mctf_levels(oracle, cur, ref, pads=(16, 32, PAD))
    n_frac = 0
    for name in ("l4", "l2", "l1", "l0"):
        c, rp, pad, bs, rng, lower, factor, frac, mv, err = lv[name]
        gmv, gerr = _level(lib, s, torch, c, rp, pad, bs, rng, lower, factor, frac)
        assert np.array_equal(gmv, mv), name
        assert np.array_equal(gerr, err), name
        n_frac += int(np.count_nonzero((mv & 15) != 0)) if frac else 0
        print(f"mctf level {name}: {mv.shape[1]}x{mv.shape[0]} blocks of {bs}, bit-exact", flush=True)
    assert n_frac > 100


def test_mctf_frame_matches_oracle(oracle, env):
    """xg_mctf_me on the frame pair: the oracle's hierarchy with the same
    paddings (16 and 32 for the subsampled references, the picture's 80)."""
    lib, s, torch = env
    cur, ref = _scene()
    lv = oracle_mctf_levels(oracle, cur, ref, pads=(16, 32, PAD))
    z = [np.zeros((H // 2, W // 2), np.uint16)] * 2
    dc, dr = DevicePicture([cur] + z, pad=0), DevicePicture([ref] + z, pad=PAD)
    nbx, nby = W // 8, H // 8
    mv = torch.zeros(2 * nbx * nby, dtype=torch.int32, device="cuda")
    err = torch.zeros(nbx * nby, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    for _ in range(2):                                   # buffers reused on the second call
        assert lib.xg_mctf_me(s, ctypes.byref(dc.pic.plane[0]), ctypes.byref(dr.pic.plane[0]),
                              ctypes.c_void_p(mv.data_ptr()), ctypes.c_void_p(err.data_ptr())) == 0
        assert lib.xg_session_sync(s) == 0
        assert np.array_equal(mv.cpu().numpy().reshape(nby, nbx, 2), lv["l0"][8])
        assert np.array_equal(err.cpu().numpy().view(np.uint32).reshape(nby, nbx), lv["l0"][9])
    # the pan is found: most 8x8 blocks outside the special regions move by (9, -5) pel
    m = lv["l0"][8]
    assert np.count_nonzero((m[..., 0] == -144) & (m[..., 1] == 80)) > nbx * nby // 3


def test_mctf_fractional_candidates_and_16x16_rounds(oracle, env):
    """Paths the frame search does not take: lower-level MVs with fractional
    parts (the finest level's own output used as the candidate field, so each
    candidate is costed through the staged two-pass filter), and the
    fractional rounds on 16x16 blocks (one position per lane, no shared first
    pass); both against the oracle on the same inputs."""
    lib, s, torch = env
    cur, ref = _scene()
    lv = oracle_mctf_levels(oracle, cur, ref, pads=(16, 32, PAD))
    rp, mv0 = lv["l0"][1], lv["l0"][8]
    assert np.count_nonzero(mv0 & 15) > 100                    # fractional candidates
    for bs, rng, lower, factor in ((8, 1, mv0, 1), (16, 2, lv["l2"][8], 2)):
        h, w = cur.shape
        nbx, nby = w // bs, h // bs
        emv, eerr = np.zeros(2 * nbx * nby, np.int32), np.zeros(nbx * nby, np.uint32)
        low = np.ascontiguousarray(lower.astype(np.int32))
        ps = (rp.shape[0] - h) // 2
        oracle.or_MctfMeLevel(ptr(cur), w, ptr(rp, ps * rp.shape[1] + ps), rp.shape[1], w, h, PAD, bs, rng, ptr(low),
                              low.shape[1], low.shape[0], factor, 1, 0, nby, ptr(emv), ptr(eerr))
        gmv, gerr = _level(lib, s, torch, cur, rp, PAD, bs, rng, lower, factor, 1)
        assert np.array_equal(gmv.reshape(-1), emv), bs
        assert np.array_equal(gerr.reshape(-1), eerr), bs
        print(f"mctf bs {bs}: {nbx}x{nby} blocks with fractional rounds, bit-exact", flush=True)


def test_mctf_padding_rule(oracle, env):
    """Texture in the padding, the current picture moved 6 right: the left
    blocks' true match reads past a padding of 8 and is skipped, as in the
    oracle."""
    lib, s, torch = env
    rng = np.random.default_rng(8)
    pad, w, h = 8, 64, 32
    buf = rng.integers(0, 1024, (h + 2 * (pad + MSLACK), w + 2 * (pad + MSLACK))).astype(np.uint16)
    o = pad + MSLACK
    c = np.ascontiguousarray(buf[o:o + h, o - 6:o - 6 + w])
    emv, eerr = np.zeros(2 * 8, np.int32), np.zeros(8, np.uint32)
    oracle.or_MctfMeLevel(ptr(c), w, ptr(buf, o * buf.shape[1] + o), buf.shape[1], w, h, pad, 16, 8, None, 0, 0, 1, 0,
                          0, h // 16, ptr(emv), ptr(eerr))
    gmv, gerr = _level(lib, s, torch, c, buf, pad, 16, 8, None, 1, 0)
    assert np.array_equal(gmv.reshape(-1), emv) and np.array_equal(gerr.reshape(-1), eerr)
    assert tuple(gmv[0, 0]) != (-96, 0)


def test_mctf_rejects_bad_arguments(env):
    lib, s, torch = env
    t = torch.zeros(64 * 64, dtype=torch.int16, device="cuda")
    p = capi.XgPlane(t.data_ptr(), 64, 64, 64, 0)
    out = torch.zeros(128, dtype=torch.int32, device="cuda")
    q = ctypes.c_void_p(out.data_ptr())
    assert lib.xg_mctf_me_level(s, ctypes.byref(p), ctypes.byref(p), 12, 4, None, 0, 0, 1, 0, q, q) == capi.XG_ERR_ARG
    assert lib.xg_mctf_me_level(s, ctypes.byref(p), ctypes.byref(p), 16, 9, None, 0, 0, 1, 0, q, q) == capi.XG_ERR_ARG
    half = capi.XgPlane(t.data_ptr(), 64, 30, 32, 0)
    assert lib.xg_mctf_subsample(s, ctypes.byref(p), ctypes.byref(half)) == capi.XG_ERR_ARG
    lib.xg_clear_error()
