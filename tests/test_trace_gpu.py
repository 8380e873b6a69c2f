"""Closest-hit / any-hit parity of the HIP engines against the oracle.

The bar is bit-exact: prim_id, t, u and v must equal the reference
restatement's values for every ray.  Reference behaviour pinned here:
SceneAcceleration::hit (core/scene.rs:182-185) with t_min = 1e-5,
BvhNode::hit/merge_optionals (core/bvh.rs:132-181), Triangle::hit
(core/mesh.rs:144-188), the brute-force small-mesh loop (mesh.rs:125-138).
"""
import numpy as np
import pytest

from tests.conftest import announce
from voidray_amd import _native as N
from voidray_amd import workloads as W
from voidray_amd.scene import Materials, Mesh, Scene

pytestmark = pytest.mark.gpu


def _compare(ref, got, label):
    miss = ref["prim"] == 0xFFFFFFFF
    bad = np.nonzero(ref["prim"] != got["prim"])[0]
    assert len(bad) == 0, f"{label}: {len(bad)} prim mismatches, first {bad[:8]}"
    h = ~miss
    for k in ("t", "u", "v"):
        diff = np.nonzero(ref[k][h].view(np.uint32) != got[k][h].view(np.uint32))[0]
        assert len(diff) == 0, f"{label}: {len(diff)} {k} mismatches"


def _scene_cases():
    yield "random4k", W.random_triangle_scene(4096, seed=11), W.microbench_rays(16384, res=128)
    room = W.cornell_room()
    rng = np.random.default_rng(5)
    o = rng.uniform(0.05, 0.95, (8192, 3)).astype(np.float32)
    d = W._normalize_rows(rng.normal(size=(8192, 3)).astype(np.float32))
    yield "cornell", room, (o, d)


@pytest.mark.parametrize("case", ["random4k", "cornell"])
def test_closest_hit_matches_oracle(case, gpu_ctx, oracle_lib):
    name, scene, (o, d) = next(c for c in _scene_cases() if c[0] == case)
    ref = oracle_lib.OracleScene(scene).trace(o, d)
    announce(f"trace wide closest {name} n={len(o)}")
    acc = scene.build_acceleration(gpu_ctx)
    got = acc.hit(o, d)
    _compare(ref, got, f"{name}/wide")
    announce(f"trace exact closest {name} n={len(o)}")
    got2 = acc.hit(o, d, engine=N.VR_ENGINE_REFERENCE)
    _compare(ref, got2, f"{name}/exact")
    announce(f"trace any {name}")
    anyh = acc.hit(o, d, any_hit=True)["hit"]
    assert np.array_equal(anyh.astype(bool), ref["prim"] != 0xFFFFFFFF)


def test_golden_fixture(gpu_ctx):
    fx = np.load("tests/golden/c2_small.npz")
    scene = W.random_triangle_scene(int(fx["num_tris"]), seed=int(fx["seed"]))
    o, d = W.microbench_rays(int(fx["num_rays"]), res=int(fx["res"]))
    assert np.array_equal(o, fx["origins"]) and np.array_equal(d, fx["directions"])
    announce("trace golden c2_small")
    got = scene.build_acceleration(gpu_ctx).hit(o, d)
    _compare(dict(prim=fx["prim"], t=fx["t"], u=fx["u"], v=fx["v"]), got, "golden")


def test_edge_cases(gpu_ctx, oracle_lib):
    # axis-aligned directions (zero components), rays from triangle planes,
    # degenerate triangles, a coplanar fan with shared edges (equal-t ties)
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
    ref = oracle_lib.OracleScene(scene).trace(o, d)
    announce(f"trace edge cases n={len(o)}")
    acc = scene.build_acceleration(gpu_ctx)
    _compare(ref, acc.hit(o, d), "edges/wide")
    _compare(ref, acc.hit(o, d, engine=N.VR_ENGINE_REFERENCE), "edges/exact")


def test_empty_scene_and_zero_rays(gpu_ctx):
    scene = Scene.empty()
    announce("trace empty scene")
    acc = scene.build_acceleration(gpu_ctx)
    o = np.zeros((5, 3), np.float32)
    d = np.tile([0, 0, 1], (5, 1)).astype(np.float32)
    r = acc.hit(o, d)
    assert np.all(r["prim"] == 0xFFFFFFFF)
    r = acc.hit(o[:0], d[:0])
    assert len(r["prim"]) == 0


def test_full_size_microbench(gpu_ctx, oracle_lib):
    """1,048,576 random triangles, 4,194,304 rays: wide engine == exact engine on
    every ray; a 4096-ray sample == the CPU oracle."""
    scene = W.random_triangle_scene(1 << 20)
    o, d = W.microbench_rays(1 << 22)
    announce("trace full size: upload 1M tris")
    acc = scene.build_acceleration(gpu_ctx)
    announce("trace full size: wide closest 4M rays")
    wide = acc.hit(o, d)
    fb = acc.last_fallbacks()
    announce(f"trace full size: exact engine 4M rays (wide fallbacks={fb})")
    exact = acc.hit(o, d, engine=N.VR_ENGINE_REFERENCE)
    _compare(exact, wide, "full/wide-vs-exact")
    assert fb < len(o) * 1e-3
    sel = np.random.default_rng(0).choice(len(o), 4096, replace=False)
    ref = oracle_lib.OracleScene(scene).trace(o[sel], d[sel])
    _compare(ref, {k: v[sel] for k, v in wide.items()}, "full/sample-vs-oracle")
    announce("trace full size: any-hit 4M rays")
    anyh = acc.hit(o, d, any_hit=True)["hit"]
    assert np.array_equal(anyh.astype(bool), wide["prim"] != 0xFFFFFFFF)
