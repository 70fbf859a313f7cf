"""Write tests/golden/interp_golden.npz from the xin26x reference C.

Run in the build container only (the reference sources are not on the GPU
box):  make -C oracle ref && python tests/golden/make_golden.py

oracle/_ref/libxin_ref.so is the reference's own 10-bit build of
h266_core/source/h266_luma_mc.c and h266_chroma_mc.c (recipe: oracle/Makefile).
The fixture holds only data: the seeded input planes, the case table and the
reference's outputs.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import golden_cases as gc  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libxin_ref.so"))
    planes = gc.make_planes()
    cases = gc.build_cases()
    out_u, out_s, offs = gc.replay(gc.resolver(lib, 0), cases, planes)
    np.savez_compressed(os.path.join(HERE, "interp_golden.npz"),
                        src_u16=planes[0], src_s16=planes[1], src2_s16=planes[2],
                        cases=cases, out_u16=out_u, out_s16=out_s, offsets=offs)
    print("cases", len(cases), "outputs", int(offs[-1]))


if __name__ == "__main__":
    main()
