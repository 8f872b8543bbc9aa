"""Writes the oracle's histograms for test cases too slow to run live in the
GPU tests (tests/golden/oracle_out/*.txt, one integer count per line).

  2ani_bpc_d150_s250_b50_{DD,RR,DR}.txt
      PCF 2ani -bpc -d 150 -s 250 -b 50 on the reference's sample catalogs
      data.dat / rand0.dat (the configuration of the reference's own CPU
      program cpp/2PCFanisotropic/BPC/main.cpp:24-26; 2*dmax > box, so pairs
      count under their regular and their projected images).

Run from the repository root: python tests/golden/make_oracle_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

OUT = os.path.join(HERE, "oracle_out")


def main():
    os.makedirs(OUT, exist_ok=True)
    d = np.loadtxt(os.path.join(HERE, "catalogs", "data.dat"), dtype=np.float32)
    r = np.loadtxt(os.path.join(HERE, "catalogs", "rand0.dat"), dtype=np.float32)
    for name, a, b in (("DD", d, None), ("RR", r, None), ("DR", d, r)):
        c, _ = O.pairs(a, b, ani=True, bpc=True, box=250.0, bins=50, dmax=150.0)
        path = os.path.join(OUT, f"2ani_bpc_d150_s250_b50_{name}.txt")
        np.savetxt(path, c, fmt="%d")
        print(path, int(c.sum()), flush=True)


if __name__ == "__main__":
    main()
