"""Writes the committed fixtures of tests/golden from the oracle.

python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from voidray_amd import workloads as W  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def c2_small():
    num_tris, seed, num_rays, res = 4096, 2, 8192, 128
    scene = W.random_triangle_scene(num_tris, seed=seed)
    o, d = W.microbench_rays(num_rays, res=res)
    r = oracle.OracleScene(scene).trace(o, d)
    np.savez_compressed(os.path.join(HERE, "c2_small.npz"), num_tris=num_tris, seed=seed, num_rays=num_rays,
                        res=res, origins=o, directions=d, prim=r["prim"], t=r["t"], u=r["u"], v=r["v"])


if __name__ == "__main__":
    c2_small()
    print("ok")
