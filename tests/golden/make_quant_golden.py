"""Write tests/golden/quant_golden.npz from the xin26x reference C.

Run in the build container only (the reference sources are not on the GPU
box):  make -C oracle ref && python tests/golden/make_quant_golden.py

oracle/_ref/libxin_ref.so holds the reference's own 10-bit build of
h266_encoder/source/h266_quant_inv_quant.c (recipe: oracle/Makefile).  The
fixture holds only data: the case table, the seeded inputs and the
reference's outputs.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import quant_cases as qc  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libxin_ref.so"))
    cases, total = qc.build_cases()
    t = qc.make_inputs(total)
    q, r, maps = qc.replay(qc.resolve(lib, "Xin266QuantInvQuant"), cases, t)
    np.savez_compressed(os.path.join(HERE, "quant_golden.npz"), cases=cases, t=t, q=q, r=r, maps=maps)
    print("cases", len(cases), "coefficients", total, "non-zero maps", int((maps != 0).sum()))


if __name__ == "__main__":
    main()
