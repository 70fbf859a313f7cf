"""Write tests/golden/convention_golden.npz from the oracle (see convention_cases.py):
python tests/golden/make_convention_golden.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)
import convention_cases as cc  # noqa: E402
from tests.oracle_lib import load_oracle  # noqa: E402


def main():
    out = cc.compute(load_oracle())
    np.savez_compressed(os.path.join(HERE, "convention_golden.npz"), **out)
    print({k: v.size for k, v in out.items()})


if __name__ == "__main__":
    main()
