"""The no-allocation lookahead entry (xg_lookahead_me_static) against the
session entry it mirrors (xg_lookahead_me): bit-identical MVs, costs and SADs
on two calls in a row, the second on a smaller picture, on a caller's stream;
inputs past the static scratch are refused before anything is enqueued."""
import ctypes

import numpy as np
import pytest

from xin26x_amd import capi, workloads as wl
from xin26x_amd.device import DevicePicture

pytestmark = pytest.mark.gpu


def _planes(cur, ref, pad):
    dc = DevicePicture([cur, cur[::2, ::2].copy(), cur[::2, ::2].copy()], pad=0)
    dr = DevicePicture([ref, ref[::2, ::2].copy(), ref[::2, ::2].copy()], pad=pad)
    return dc, dr


def _outputs(torch, nb):
    return (torch.zeros(2 * nb, dtype=torch.int16, device="cuda"),
            torch.zeros(nb, dtype=torch.int32, device="cuda"),
            torch.zeros(nb, dtype=torch.int32, device="cuda"))


def test_lookahead_static_matches_session_entry():
    import torch
    lib = capi.load()
    lib.xg_clear_error()
    fp = wl.frame_pair_1080p()
    small_cur, small_ref = wl.frame_pair_416x240()
    cases = [(fp.cur[0], fp.ref[0], 16, 32, 80), (small_cur, small_ref, 8, 24, 64)]
    s = ctypes.c_void_p()
    assert lib.xg_session_create(ctypes.byref(s), 1024) == 0
    stream = torch.cuda.Stream()
    try:
        for call, (cur, ref, unit, rng_, pad) in enumerate(cases):
            H, W = cur.shape
            print(f"no-alloc lookahead call {call}: {W}x{H}, unit {unit}, range {rng_}, pad {pad}", flush=True)
            dc, dr = _planes(cur, ref, pad)
            nb = (W // 2 // unit) * (H // 2 // unit)
            want = _outputs(torch, nb)
            got = _outputs(torch, nb)
            for t in got:
                t.fill_(-1)
            rc = lib.xg_lookahead_me(s, ctypes.byref(dc.pic.plane[0]), ctypes.byref(dr.pic.plane[0]), unit, rng_,
                                     1024, *(t.data_ptr() for t in want))
            assert rc == 0 and lib.xg_session_sync(s) == 0
            torch.cuda.synchronize()
            rc = lib.xg_lookahead_me_static(ctypes.c_void_p(stream.cuda_stream), ctypes.byref(dc.pic.plane[0]),
                                            ctypes.byref(dr.pic.plane[0]), unit, rng_, 1024,
                                            *(t.data_ptr() for t in got))
            assert rc == 0, f"xg_lookahead_me_static failed with {rc}"
            stream.synchronize()
            for w, g in zip(want, got):
                assert torch.equal(w, g), f"call {call}: static-scratch outputs differ from the session entry's"
            assert lib.xg_last_error() == 0
    finally:
        lib.xg_session_destroy(s)


def test_lookahead_static_refuses_oversize():
    import torch
    lib = capi.load()
    lib.xg_clear_error()
    buf = torch.zeros(64, dtype=torch.int16, device="cuda")
    # a 3840x2160 pair padded by 1000 needs more than the static scratch; the
    # sizes are checked on the host before anything is enqueued
    cur = capi.XgPlane(buf.data_ptr(), 3840, 3840, 2160, 0)
    ref = capi.XgPlane(buf.data_ptr(), 3840 + 2000, 3840, 2160, 1000)
    need = 2 * (1920 * 1080 + (1920 + 1000) * (1080 + 1000))
    assert need > capi.XG_LOOKAHEAD_STATIC_MAX_BYTES
    mv = torch.zeros(2, dtype=torch.int16, device="cuda")
    rc = lib.xg_lookahead_me_static(None, ctypes.byref(cur), ctypes.byref(ref), 16, 32, 1024, mv.data_ptr(),
                                    mv.data_ptr(), mv.data_ptr())
    assert rc == capi.XG_ERR_ARG and lib.xg_last_error() == capi.XG_ERR_ARG
    lib.xg_clear_error()
    torch.cuda.synchronize()
