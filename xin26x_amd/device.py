"""Device pictures for the batched entry points.

torch allocates the HBM buffers (plumbing only); every computation runs in
libxg266.so.  Planes are stored edge-padded, as the reference pads its
reference pictures (h266_encoder/source/h266_encoder_create.c:1065-1066).
"""
import ctypes

import numpy as np

from .capi import XgPicture, XgPlane

CTU_PAD = 64 + 16        # ctuSize + 16 luma samples


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("a HIP device is required (torch.cuda.is_available() is False)")
    return torch


class DevicePicture:
    """Three padded planes in HBM and the matching xg_picture."""

    def __init__(self, planes, pad=CTU_PAD, device="cuda"):
        torch = _torch()
        self.tensors = []
        self.pic = XgPicture()
        self.pad = pad
        for p, a in enumerate(planes):
            pp = pad if p == 0 else pad // 2
            if pp:
                a2 = np.pad(a, pp, mode="edge")
            else:
                a2 = a
            t = torch.from_numpy(np.ascontiguousarray(a2).view(np.int16)).to(device)
            stride = a2.shape[1]
            ptr = t.data_ptr() + 2 * (pp * stride + pp)
            self.pic.plane[p] = XgPlane(ptr, stride, a.shape[1], a.shape[0], pp)
            self.tensors.append(t)

    @classmethod
    def empty(cls, width, height, device="cuda"):
        z = [np.zeros((height, width), np.uint16), np.zeros((height // 2, width // 2), np.uint16),
             np.zeros((height // 2, width // 2), np.uint16)]
        return cls(z, pad=0, device=device)

    def planes(self):
        """Unpadded planes back on the host (uint16)."""
        out = []
        for p, t in enumerate(self.tensors):
            pl = self.pic.plane[p]
            a = t.cpu().numpy().view(np.uint16)
            out.append(a[pl.pad:pl.pad + pl.height, pl.pad:pl.pad + pl.width].copy())
        return out

    def ref(self):
        return ctypes.byref(self.pic)


BLOCK_DTYPE = np.dtype([("x", "<i2"), ("y", "<i2"), ("w", "u1"), ("h", "u1"), ("pic", "u1"),
                        ("ref", "u1"), ("mvx", "<i4"), ("mvy", "<i4"), ("mvpx", "<i4"), ("mvpy", "<i4")])


def blocks_to_struct(blocks, pic=0, ref=1):
    """(n, 8) int array [x, y, w, h, mvx, mvy, mvpx, mvpy] -> numpy array laid
    out as xg_block (24 bytes each); pass arr.ctypes.data to the C-ABI."""
    b = np.asarray(blocks, dtype=np.int64)
    arr = np.zeros(len(b), BLOCK_DTYPE)
    for i, name in enumerate(("x", "y", "w", "h")):
        arr[name] = b[:, i]
    arr["pic"] = pic
    arr["ref"] = ref
    for i, name in enumerate(("mvx", "mvy", "mvpx", "mvpy")):
        arr[name] = b[:, 4 + i]
    return arr


def refine_cfg(offsets, lam, ctu=64, use_chroma=1):
    from .capi import XgRefineCfg
    c = XgRefineCfg()
    c.ncand = len(offsets)
    for i, (dx, dy) in enumerate(offsets):
        c.offx[i] = int(dx)
        c.offy[i] = int(dy)
    c.lambda_ = lam
    c.ctuSize = ctu
    c.useChroma = use_chroma
    return c
