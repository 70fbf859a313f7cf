"""Write tests/golden/bili_golden.npz from the xin26x reference C.

Run in the build container only (the reference sources are not on the GPU
box):  make -C oracle ref && python tests/golden/make_bili_golden.py

oracle/_ref/libxin_ref.so holds the reference's own 10-bit build of
h266_core/source/h266_bilinear_filter.c (recipe: oracle/Makefile).  The
fixture holds only data: the seeded input plane, the case table and the
reference's outputs.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import bili_cases as bc  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libxin_ref.so"))
    plane = bc.make_plane()
    cases = bc.build_cases()
    out, offs = bc.replay(bc.resolve(lib, "Xin266"), cases, plane)
    np.savez_compressed(os.path.join(HERE, "bili_golden.npz"), plane=plane, cases=cases, out=out, offsets=offs)
    print("cases", len(cases), "outputs", int(offs[-1]))


if __name__ == "__main__":
    main()
