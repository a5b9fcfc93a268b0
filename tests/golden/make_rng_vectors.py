"""Write tests/golden/rngs_c_vectors.npz from the reference's own RNGs.c.

The reference generator (src_random/RNGs.c) is compiled from the reference
tree by ``make -C oracle ref`` into oracle/_ref/librngs_ref.so; this script
records its outputs so the oracle's generator stays pinned where that build
is absent.  Contents:
  raw_<seed>:  first 2000 outputs of genrandbase() after srandgen(seed)
  normal_77:   first 5000 outputs of genrandn() after srandgen(77)
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

SEEDS = (1, 5489, 12345, 2 ** 31 + 7)


def main():
    ref = O.ref_rng()
    if ref is None:
        raise SystemExit("oracle/_ref/librngs_ref.so is missing: run `make -C oracle ref` with the reference tree")
    out = {}
    for s in SEEDS:
        ref.srandgen(s)
        out[f"raw_{s}"] = np.array([ref.genrandbase() & 0xFFFFFFFF for _ in range(2000)], dtype=np.uint32)
    ref.srandgen(77)
    v = C.c_double()
    nrm = []
    for _ in range(5000):
        ref.genrandn(C.byref(v))
        nrm.append(v.value)
    out["normal_77"] = np.array(nrm)
    np.savez_compressed(os.path.join(HERE, "rngs_c_vectors.npz"), **out)


if __name__ == "__main__":
    main()
