"""MI355X (gfx950) VVC block-matching library: SAD/SATD, 8-tap/4-tap
interpolation, integer and fractional motion search, forward transforms.

The compute lives in libxg266.so (HIP, include/xg266.h); this package holds
its ctypes bindings (capi), device-picture plumbing (device), the synthetic
workloads (workloads) and the build script (build).
"""
__all__ = ["capi", "device", "workloads", "build"]
