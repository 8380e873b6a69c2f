"""Ray-query scenes shared by the GPU and host traversal tests."""
import numpy as np

from voidray_amd.scene import Materials, Mesh, Scene


def edge_case_scene():
    """Axis-aligned directions (zero components), rays starting on triangle
    planes, a degenerate triangle, a coplanar fan with shared edges (equal-t
    ties) and a second surface.  Returns (scene, origins, directions)."""
    pos = []
    idx = []
    g = np.linspace(0, 1, 9, dtype=np.float32)
    for i in range(8):
        for j in range(8):
            a = [g[i], g[j], 0.5]
            b = [g[i + 1], g[j], 0.5]
            c = [g[i + 1], g[j + 1], 0.5]
            e = [g[i], g[j + 1], 0.5]
            k = len(pos)
            pos += [a, b, c, e]
            idx += [[k, k + 1, k + 2], [k, k + 2, k + 3]]
    k = len(pos)
    pos += [[0.2, 0.2, 0.2], [0.2, 0.2, 0.2], [0.3, 0.3, 0.3]]  # degenerate
    idx += [[k, k + 1, k + 2]]
    scene = Scene.empty()
    s = scene.add_mesh(Mesh(np.asarray(pos, np.float32), np.asarray(idx, np.uint32)))
    scene.add_object(scene.add_material(Materials.lambertian((0.5, 0.5, 0.5))), s)
    q = scene.add_mesh(Mesh(np.asarray([[0, 0, 0.7], [1, 0, 0.7], [0, 1, 0.7]], np.float32), [[0, 1, 2]]))
    scene.add_object(0, q)
    pts = np.stack(np.meshgrid(g, g), -1).reshape(-1, 2)
    o = np.concatenate([np.c_[pts, np.zeros(len(pts))], np.c_[pts, np.full(len(pts), 0.5)],
                        np.c_[pts + 0.01, np.ones(len(pts))]]).astype(np.float32)
    d = np.concatenate([np.tile([0, 0, 1], (len(pts), 1)), np.tile([0, 0, 1], (len(pts), 1)),
                        np.tile([0, 0, -1], (len(pts), 1))]).astype(np.float32)
    return scene, o, d


def compare_hits(ref, got, label):
    """Bit-exact: prim everywhere, t/u/v wherever the reference hits."""
    miss = ref["prim"] == 0xFFFFFFFF
    bad = np.nonzero(ref["prim"] != got["prim"])[0]
    assert len(bad) == 0, f"{label}: {len(bad)} prim mismatches, first {bad[:8]}"
    h = ~miss
    for k in ("t", "u", "v"):
        diff = np.nonzero(np.asarray(ref[k], np.float32)[h].view(np.uint32) != got[k][h].view(np.uint32))[0]
        assert len(diff) == 0, f"{label}: {len(diff)} {k} mismatches"
