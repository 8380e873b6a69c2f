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


C2_FULL = dict(num_tris=1 << 20, seed=2, num_rays=1 << 22, res=2048, block=1 << 16)


def c2_full_digests(prim, t, u, v, block):
    """Per-block SHA-256 of the closest-hit answers (prim, then the bits of t,
    u and v, each little-endian), so a mismatch names the rays it is in."""
    import hashlib
    out = []
    for b0 in range(0, len(prim), block):
        h = hashlib.sha256()
        for a in (prim, t, u, v):
            h.update(np.ascontiguousarray(a[b0:b0 + block]).view(np.uint32).astype("<u4").tobytes())
        out.append(h.hexdigest())
    return out


def c2_full_inputs():
    """The traversal microbenchmark at full size: 1,048,576 random triangles
    and 4,194,304 rays, with SHA-256 of the generated inputs."""
    import hashlib
    cfg = C2_FULL
    scene = W.random_triangle_scene(cfg["num_tris"], seed=cfg["seed"])
    o, d = W.microbench_rays(cfg["num_rays"], res=cfg["res"])
    m = scene.meshes[0]
    ins = dict(positions=hashlib.sha256(m.positions.astype("<f4").tobytes()).hexdigest(),
               origins=hashlib.sha256(o.astype("<f4").tobytes()).hexdigest(),
               directions=hashlib.sha256(d.astype("<f4").tobytes()).hexdigest())
    return scene, o, d, ins


def c2_full():
    """Oracle answers for every ray of the full microbenchmark, kept as block
    digests (the answers themselves are 64 MB)."""
    import json
    cfg = C2_FULL
    scene, o, d, ins = c2_full_inputs()
    r = oracle.OracleScene(scene).trace(o, d)
    doc = dict(cfg, inputs=ins, hits=int(np.sum(r["prim"] != 0xFFFFFFFF)),
               blocks=c2_full_digests(r["prim"], r["t"], r["u"], r["v"], cfg["block"]))
    with open(os.path.join(HERE, "c2_full_digests.json"), "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    which = sys.argv[1:] or ["c2_small", "c2_full"]
    for name in which:
        globals()[name]()
    print("ok")
