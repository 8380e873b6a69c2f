"""The traversal routines and the BVH build, checked on the host (no GPU).

`vr_trace_host` builds the scene exactly as the device upload does (SAH 4-wide
BVH with spatial splits, reference trees, ancestor tables) and runs the rays
through the kernels' own traversal code on the CPU.  The bar is the GPU tests'
bar: prim, t, u, v bit-identical to the oracle (SceneAcceleration::hit,
core/scene.rs:182-185; BvhNode::hit/merge_optionals, core/bvh.rs:132-181;
Triangle::hit, core/mesh.rs:144-188).
"""
import numpy as np

from tests.trace_cases import compare_hits, default_box_scene, edge_case_scene
from voidray_amd import _native as N
from voidray_amd import workloads as W
from voidray_amd.device import trace_host
from voidray_amd.scene import Scene


def _check_all(scene, o, d, oracle_lib, label, build_flags=0):
    ref = oracle_lib.OracleScene(scene).trace(o, d)
    wide = trace_host(scene, o, d, build_flags=build_flags)
    compare_hits(ref, wide, f"{label}/wide")
    exact = trace_host(scene, o, d, engine=N.VR_ENGINE_REFERENCE)
    compare_hits(ref, exact, f"{label}/exact")
    anyh = trace_host(scene, o, d, any_hit=True, build_flags=build_flags)
    assert np.array_equal(anyh["hit"].astype(bool), ref["prim"] != 0xFFFFFFFF), f"{label}/any"
    return wide


def test_spatial_split_bvh_is_exact(oracle_lib):
    # long thin randomly oriented triangles: the build cuts many of them, so
    # leaves hold clipped pieces and triangles appear in several leaves
    scene = W.random_triangle_scene(20000, seed=5)
    o, d = W.microbench_rays(8192, res=64)
    wide = _check_all(scene, o, d, oracle_lib, "soup20k")
    st = wide["stats"]
    assert st["num_triangles"] == 20000
    assert 1.2 * st["num_triangles"] < st["num_tri_records"] <= 8 * st["num_triangles"], st
    assert wide["fallbacks"] < 8


def test_leaf_layouts(oracle_lib):
    # leaves holding the records themselves (the default) or 4-byte references
    # into one record per triangle give the same answers
    scene = W.random_triangle_scene(20000, seed=5)
    o, d = W.microbench_rays(4096, res=64)
    for flags, refs in ((N.VR_BUILD_LEAF_RECORDS, 0), (N.VR_BUILD_LEAF_REFS, 1)):
        st = _check_all(scene, o, d, oracle_lib, f"soup20k/layout{flags}", build_flags=flags)["stats"]
        assert st["leaf_refs"] == refs, st
    assert trace_host(scene, o[:1], d[:1])["stats"]["leaf_refs"] == 0
    _check_all(W.cornell_room(), o, d, oracle_lib, "cornell/refs", build_flags=N.VR_BUILD_LEAF_REFS)


def test_multi_surface_scene(oracle_lib):
    # cornell room: many surfaces (top-level tree, per-surface ancestor paths)
    # plus a 1k-triangle sphere mesh
    room = W.cornell_room()
    rng = np.random.default_rng(5)
    o = rng.uniform(0.05, 0.95, (4096, 3)).astype(np.float32)
    d = W._normalize_rows(rng.normal(size=(4096, 3)).astype(np.float32))
    _check_all(room, o, d, oracle_lib, "cornell")


def test_edge_cases(oracle_lib):
    scene, o, d = edge_case_scene()
    _check_all(scene, o, d, oracle_lib, "edges")


def test_default_boxes_beside_flat_boxes(oracle_lib):
    scene, o, d = default_box_scene()
    _check_all(scene, o, d, oracle_lib, "default-boxes")


def test_golden_fixture():
    fx = np.load("tests/golden/c2_small.npz")
    scene = W.random_triangle_scene(int(fx["num_tris"]), seed=int(fx["seed"]))
    o, d = W.microbench_rays(int(fx["num_rays"]), res=int(fx["res"]))
    got = trace_host(scene, o, d)
    compare_hits(dict(prim=fx["prim"], t=fx["t"], u=fx["u"], v=fx["v"]), got, "golden/host")


def test_empty_scene():
    o = np.zeros((5, 3), np.float32)
    d = np.tile([0, 0, 1], (5, 1)).astype(np.float32)
    r = trace_host(Scene.empty(), o, d)
    assert np.all(r["prim"] == 0xFFFFFFFF)
    assert trace_host(Scene.empty(), o[:0], d[:0])["prim"].shape == (0,)


def test_long_needles(oracle_lib):
    # 3,000 long thin triangles (length 0.6 in a unit box, random directions):
    # heavy box overlap; the tree holds at most 8 references per triangle
    rng = np.random.default_rng(3)
    n = 3000
    a = rng.uniform(0, 1, (n, 3))
    dv = rng.normal(size=(n, 3))
    dv /= np.linalg.norm(dv, axis=1, keepdims=True)
    pos = np.stack([a, a + 0.6 * dv, a + 2e-3 * rng.normal(size=(n, 3))], 1).reshape(-1, 3).astype(np.float32)
    from voidray_amd.scene import Materials, Mesh
    scene = Scene.empty()
    s = scene.add_mesh(Mesh(pos, np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)))
    scene.add_object(scene.add_material(Materials.lambertian((0.5, 0.5, 0.5))), s)
    o = rng.uniform(-0.2, 1.2, (4096, 3)).astype(np.float32)
    d = W._normalize_rows(rng.normal(size=(4096, 3)).astype(np.float32))
    st = _check_all(scene, o, d, oracle_lib, "needles")["stats"]
    assert n <= st["num_tri_records"] <= 8 * n, st


def test_analytic_surfaces(oracle_lib):
    # spheres and ground planes (surfaces.rs) mixed with meshes in the top-level
    # tree; rays from inside and outside the spheres, grazing the plane
    scene = W.analytic_and_meshes()
    rng = np.random.default_rng(8)
    o = rng.uniform(-1.5, 1.5, (6000, 3)).astype(np.float32)
    d = W._normalize_rows(rng.normal(size=(6000, 3)).astype(np.float32))
    d[:500, 1] = 0.0  # parallel to the ground plane
    d[:500] = W._normalize_rows(d[:500])
    wide = _check_all(scene, o, d, oracle_lib, "analytic+meshes")
    assert wide["stats"]["num_triangles"] == scene.num_triangles


def test_analytic_spheres_demo(oracle_lib):
    # the reference's spheres example: every surface analytic, no triangles
    scene = W.analytic_spheres()
    rng = np.random.default_rng(9)
    o = np.tile(np.float32([0.7166, 2.8803, -9.2992]), (4096, 1)) + rng.normal(scale=0.1, size=(4096, 3)).astype(np.float32)
    d = W._normalize_rows((np.float32([0.8673, 0.9557, 0.2095]) - o + rng.normal(scale=2.0, size=(4096, 3))).astype(np.float32))
    wide = _check_all(scene, o, d, oracle_lib, "spheres-demo")
    hits = wide["prim"] != 0xFFFFFFFF
    assert hits.mean() > 0.5
    assert wide["stats"]["num_triangles"] == 0


def test_cornell_demo_scale(oracle_lib):
    # the reference's cornell example at its own 555 scale: zero-normal quads
    # and two concentric analytic spheres 0.1 apart
    scene = W.cornell_demo()
    rng = np.random.default_rng(10)
    o = rng.uniform(5.0, 550.0, (4096, 3)).astype(np.float32)
    o[:1024] = np.float32([278.0, 278.0, -800.0])
    d = W._normalize_rows(rng.normal(size=(4096, 3)).astype(np.float32))
    d[:1024] = W._normalize_rows((np.float32([277.5, 100.0, 277.5]) + rng.normal(scale=60.0, size=(1024, 3)) - o[:1024]).astype(np.float32))
    wide = _check_all(scene, o, d, oracle_lib, "cornell-demo")
    assert (wide["prim"] != 0xFFFFFFFF).mean() > 0.8  # the room is open at the front
