"""Small helpers shared by the tests."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def ptr(a, off=0):
    """Address of a[off] (flat); the pointer object keeps `a` alive, so a
    temporary array passed inline outlives the call it is passed to."""
    p = ctypes.c_void_p(a.ctypes.data + off * a.itemsize)
    p._keep = a
    return p


def header_functions():
    """Names of every function include/xg266.h declares."""
    txt = open(os.path.join(ROOT, "include", "xg266.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set()
    for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ ]*?[\s\*]+([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt, re.M):
        line = txt[txt.rfind("\n", 0, m.start()) + 1:m.end()]
        if line.lstrip().startswith("typedef"):
            continue
        names.add(m.group(1))
    return sorted(names)


def rand_pixels(rng, shape, extremes=True):
    a = rng.integers(0, 1024, size=shape)
    if extremes and a.size > 8:
        flat = a.reshape(-1)
        k = max(1, flat.size // 16)
        flat[rng.integers(0, flat.size, k)] = 0
        flat[rng.integers(0, flat.size, k)] = 1023
    return a.astype(np.uint16)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
