"""Writes the heat delta1 golden fixtures from the reference's own CPU step.

Runs cpu_fdm_heat_diffuse1 (heat-3D.cpp:48-62), compiled from the reference
sources by oracle/Makefile into oracle/_ref/libheat3d_ref.so, on padded blocks
whose one-cell frame is zero (R=1, one block: the reference's update region
[1, N-2] is then exactly the owned grid), with the reference's constants
Alpha=Dt=0.1, Dx=Dy=1, Dz=0.5.  Inputs come from the shared SplitMix64 uniform
generator.  Usage: python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import _oracle as O  # noqa: E402

CASES = [((9, 10, 12), 7, 0), ((16, 16, 16), 25, 1), ((5, 17, 8), 11, 2)]


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libheat3d_ref.so missing: make -C oracle ref")
    for shape, steps, seed in CASES:
        owned = O.uniform(shape, seed=seed)
        u = O.frame(owned, 1)
        out = u.copy()
        for _ in range(steps):
            out = O.reference_block_step(out, 0)
        name = "heat_delta1_ref_%dx%dx%d_s%d.npz" % (shape[2], shape[1], shape[0], steps)
        np.savez_compressed(os.path.join(HERE, name), input=u, output=out, steps=steps,
                            seed=seed, sha256=hashlib.sha256(out.tobytes()).hexdigest())
        print("wrote", name)


if __name__ == "__main__":
    main()
