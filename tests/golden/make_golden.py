"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

The reference (Julia) cannot run in this environment and its own tests hold
no golden vectors, so these fixtures are regression pins of the oracle: inputs
(seeded SplitMix64 complex Gaussian fields and Fourier Green data built from
the oracle's test tensor models) with the outputs of the oracle's restatement
of the reference's host matvec.  The oracle itself is pinned by
tests/test_oracle.py.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import __graft_entry__ as ge  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    orc = ge.load_oracle()
    h = 1.0 / 32
    # self 8^3, face-pair model (the reference's formulation, model 2 rules)
    N = (8, 8, 8)
    toe = orc.toe_model(N, (h, h, h), model=2, qn=10)
    crc = orc.crc_self(N, toe)
    fur = orc.fur_from_crc(N, N, (1, 1, 1), crc)
    x = orc.complex_gaussian(N + (3,), seed=11)
    y = orc.matvec(N, N, (1, 1, 1), fur, x)
    np.savez_compressed(os.path.join(HERE, "self_8.npz"), cells=np.array(N), h=h, sym=np.array((1, 1, 1)),
                        toe=toe, fur=fur, x=x, y=y)
    # external 4x6x2 -> 6x8x4 at a separation, collocation model
    S, T = (4, 6, 2), (6, 8, 4)
    delta = (0.3, -0.2, 0.5)
    crc = orc.crc_ext(S, T, (h, h, h), delta, model=0)
    fur = orc.fur_from_crc(S, T, (0, 0, 0), crc)
    x = orc.complex_gaussian(S + (3,), seed=12)
    y = orc.matvec(S, T, (0, 0, 0), fur, x)
    np.savez_compressed(os.path.join(HERE, "ext_4x6x2_6x8x4.npz"), src=np.array(S), trg=np.array(T), h=h,
                        delta=np.array(delta), sym=np.array((0, 0, 0)), fur=fur, x=x, y=y)
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
