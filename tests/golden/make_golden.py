"""Writes the golden fixtures under tests/golden/ with the CPU oracle.

config1_rk4.npz: config 1 of the benchmark set -- star + 256 planetesimals
(seed 1, the disk generator's defaults), classical RK4 with fixed
h = 0.1 day (0.1 k in tau units) for 10^4 steps, FP64, reference GPU force
formula.  Stores the initial and final positions/velocities and the final time.

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "red.cuda_amd"))
import oracle_lib as O  # noqa: E402

K = 0.01720209895


def config1(steps=10000):
    r, v, p, b, nbt = O.make_disk(256, 1)
    d = O.ODisk(r, v, p, b, nbt)
    it = O.OInteg(d, 2, 0.1 * K, adaptive=False)
    for _ in range(steps):
        it.step()
    t, rr, vv, _, _ = d.state()
    return r, v, rr, vv, t


if __name__ == "__main__":
    r0, v0, r1, v1, t = config1()
    def xyz(a):
        return np.stack([a["x"], a["y"], a["z"]], 1)
    np.savez(os.path.join(HERE, "config1_rk4.npz"), r0=xyz(r0), v0=xyz(v0), r=xyz(r1), v=xyz(v1), t=np.float64(t))
    print("t =", t)
