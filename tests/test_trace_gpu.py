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
from tests.trace_cases import compare_hits, default_box_scene, edge_case_scene
from voidray_amd import _native as N
from voidray_amd import workloads as W
from voidray_amd.scene import Scene

pytestmark = pytest.mark.gpu


def _compare(ref, got, label):
    compare_hits(ref, got, label)


def _scene_cases():
    yield "random4k", W.random_triangle_scene(4096, seed=11), W.microbench_rays(16384, res=128)
    room = W.cornell_room()
    rng = np.random.default_rng(5)
    o = rng.uniform(0.05, 0.95, (8192, 3)).astype(np.float32)
    d = W._normalize_rows(rng.normal(size=(8192, 3)).astype(np.float32))
    yield "cornell", room, (o, d)
    # analytic spheres / ground planes mixed with meshes (surfaces.rs)
    o = rng.uniform(-1.5, 1.5, (8192, 3)).astype(np.float32)
    d = W._normalize_rows(rng.normal(size=(8192, 3)).astype(np.float32))
    d[:512, 1] = 0.0
    d[:512] = W._normalize_rows(d[:512])
    yield "analytic", W.analytic_and_meshes(), (o, d)


@pytest.mark.parametrize("case", ["random4k", "cornell", "analytic"])
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


def test_mid_fixture(gpu_ctx):
    """65,536 triangles, 262,144 rays (tests/golden/c2_mid_digests.json): the
    wide and exact engines give the oracle's digests on every block; the
    any-hit bits are the closest hits' and their count the fixture's."""
    import json
    from tests.golden.make_fixtures import c2_full_digests, c2_mid_inputs
    with open("tests/golden/c2_mid_digests.json") as f:
        fx = json.load(f)
    scene, o, d = c2_mid_inputs()
    acc = scene.build_acceleration(gpu_ctx)
    for engine in (N.VR_ENGINE_WIDE, N.VR_ENGINE_REFERENCE):
        announce(f"trace mid fixture engine {engine}")
        r = acc.hit(o, d, engine=engine)
        got = c2_full_digests(r["prim"], r["t"], r["u"], r["v"], fx["block"])
        bad = [k for k, (a, b) in enumerate(zip(got, fx["blocks"])) if a != b]
        assert not bad, f"engine {engine}: blocks {bad} differ"
    anyh = acc.hit(o, d, any_hit=True)["hit"].astype(bool)
    assert np.array_equal(anyh, r["prim"] != 0xFFFFFFFF) and int(anyh.sum()) == fx["hits"]


def test_edge_cases(gpu_ctx, oracle_lib):
    scene, o, d = edge_case_scene()
    ref = oracle_lib.OracleScene(scene).trace(o, d)
    announce(f"trace edge cases n={len(o)}")
    acc = scene.build_acceleration(gpu_ctx)
    _compare(ref, acc.hit(o, d), "edges/wide")
    _compare(ref, acc.hit(o, d, engine=N.VR_ENGINE_REFERENCE), "edges/exact")


def test_default_boxes_beside_flat_boxes(gpu_ctx, oracle_lib):
    """Empty surfaces (default boxes, which AABB::hit passes unconditionally)
    next to flat sibling boxes (which it rejects): both engines equal the
    oracle bit for bit.  The default-box test once compiled to a
    wave-uniform value (one lane's answer for all), which lanes testing the
    two kinds of box in one wave would get wrong."""
    scene, o, d = default_box_scene()
    ref = oracle_lib.OracleScene(scene).trace(o, d)
    acc = scene.build_acceleration(gpu_ctx)
    for engine in (N.VR_ENGINE_WIDE, N.VR_ENGINE_REFERENCE):
        announce(f"trace default boxes beside flat boxes, engine {engine}, n={len(o)}")
        _compare(ref, acc.hit(o, d, engine=engine), f"default-boxes/{engine}")


def test_trace_timing_only_on_request(gpu_ctx):
    """VR_TRACE_TIMING: vr_trace_last_timing reports the wide kernel's device
    time only after a trace that asked for it, and timing changes no hit."""
    import torch
    scene = W.random_triangle_scene(4096, seed=9)
    o, d = W.microbench_rays(8192, res=64)
    acc = scene.build_acceleration(gpu_ctx)
    announce("trace timing on request")
    to, td = torch.from_numpy(o).cuda(), torch.from_numpy(d).cuda()
    outs = []
    for timing in (False, True):
        prim = torch.empty(len(o), dtype=torch.int32, device="cuda")
        acc.trace_device(to.data_ptr(), td.data_ptr(), len(o), d_prim=prim.data_ptr(), timing=timing)
        torch.cuda.synchronize()
        if timing:
            assert acc.last_wide_ms() > 0.0
        else:
            with pytest.raises(N.VoidrayError):
                acc.last_wide_ms()
        outs.append(prim.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])


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


def test_full_size_microbench(gpu_ctx):
    """1,048,576 random triangles, 4,194,304 rays: every ray's closest hit
    (prim, t, u, v bits) equals the oracle's, compared through the committed
    per-block digests of the oracle's answers (tests/golden/c2_full_digests.json,
    written by tests/golden/make_fixtures.py); the exact engine agrees too, and
    the any-hit bit is set exactly where a closest hit exists."""
    import json
    from tests.golden.make_fixtures import c2_full_digests, c2_full_inputs
    with open("tests/golden/c2_full_digests.json") as f:
        fx = json.load(f)
    scene, o, d, ins = c2_full_inputs()
    assert ins == fx["inputs"], "the generators gave other inputs than the fixture's"
    announce("trace full size: upload 1M tris")
    acc = scene.build_acceleration(gpu_ctx)
    announce("trace full size: wide closest 4M rays")
    wide = acc.hit(o, d)
    fb = acc.last_fallbacks()
    got = c2_full_digests(wide["prim"], wide["t"], wide["u"], wide["v"], fx["block"])
    bad = [k for k, (a, b) in enumerate(zip(got, fx["blocks"])) if a != b]
    assert not bad, f"closest hits differ from the oracle in ray blocks {bad[:8]} of {len(got)}"
    assert int(np.sum(wide["prim"] != 0xFFFFFFFF)) == fx["hits"]
    assert fb < len(o) * 1e-3
    announce(f"trace full size: exact engine 4M rays (wide fallbacks={fb})")
    exact = acc.hit(o, d, engine=N.VR_ENGINE_REFERENCE)
    _compare(exact, wide, "full/wide-vs-exact")
    announce("trace full size: any-hit 4M rays")
    anyh = acc.hit(o, d, any_hit=True)["hit"]
    assert np.array_equal(anyh.astype(bool), wide["prim"] != 0xFFFFFFFF)


def test_equal_t_ties(gpu_ctx):
    """Shared edges and vertices, coincident triangles and spheres: both
    engines return the smallest (surface, triangle) handle among equal-t hits,
    as merge_optionals does (bvh.rs:143-181, ray.rs:52-62), checked against a
    numpy brute force that shares no code with the library or the oracle."""
    from tests.test_ties import SCENES
    from tests.trace_cases import bruteforce_lowest_handle
    for name in sorted(SCENES):
        sc, o, d = SCENES[name]
        ref = bruteforce_lowest_handle(sc, o, d)
        announce(f"trace ties {name} n={len(o)}")
        acc = sc.build_acceleration(gpu_ctx)
        for engine in (N.VR_ENGINE_WIDE, N.VR_ENGINE_REFERENCE):
            got = acc.hit(o, d, engine=engine)
            bad = np.nonzero(ref["prim"] != got["prim"])[0]
            assert len(bad) == 0, f"{name}/engine{engine}: {len(bad)} rays differ, first {bad[:8]}"
            hit = ref["prim"] != 0xFFFFFFFF
            assert np.array_equal(ref["t"][hit].view(np.uint32), got["t"][hit].view(np.uint32))


def test_leaf_reference_layout(gpu_ctx, oracle_lib):
    """Leaves holding 4-byte references into one record per triangle
    (VR_BUILD_LEAF_REFS): wide == exact on every ray, a sample == the oracle."""
    scene = W.random_triangle_scene(200000, seed=7)
    o, d = W.microbench_rays(1 << 18, res=512, seed=8)
    announce("trace leaf references: 200k tris, 262k rays")
    acc = scene.build_acceleration(gpu_ctx, build_flags=N.VR_BUILD_LEAF_REFS)
    assert acc.stats()["leaf_refs"] == 1
    wide = acc.hit(o, d)
    exact = acc.hit(o, d, engine=N.VR_ENGINE_REFERENCE)
    _compare(exact, wide, "refs/wide-vs-exact")
    sel = np.random.default_rng(1).choice(len(o), 2048, replace=False)
    ref = oracle_lib.OracleScene(scene).trace(o[sel], d[sel])
    _compare(ref, {k: v[sel] for k, v in wide.items()}, "refs/sample-vs-oracle")
    anyh = acc.hit(o, d, any_hit=True)["hit"]
    assert np.array_equal(anyh.astype(bool), wide["prim"] != 0xFFFFFFFF)


def test_watertight_engine(gpu_ctx):
    """VR_ENGINE_WATERTIGHT on the GPU gives the host run of the same
    traversal code bit for bit (closest and any hit), loses no ray through a
    closed sphere, and renders only under the extended estimator."""
    from tests.test_watertight import _scene, _sphere_targets
    from voidray_amd.device import make_render_desc, trace_host
    from voidray_amd.scene import RenderSettings
    mesh = W.uv_sphere((0.1, -0.2, 0.3), 1.0, 48, 24)
    sphere = _scene(mesh)
    c = np.array([0.1, -0.2, 0.3])
    tgt = _sphere_targets(mesh)
    src = c + np.array([0.13, -0.07, 0.05])
    o = np.tile(src, (len(tgt), 1)).astype(np.float32)
    d = W._normalize_rows((tgt - src).astype(np.float32))
    cases = [("sphere", sphere, o, d), ("soup", W.random_triangle_scene(20000, seed=5)) + W.microbench_rays(8192, res=64)]
    for name, scene, oo, dd in cases:
        announce(f"watertight {name} n={len(oo)}")
        acc = scene.build_acceleration(gpu_ctx)
        got = acc.hit(oo, dd, engine=N.VR_ENGINE_WATERTIGHT)
        ref = trace_host(scene, oo, dd, engine=N.VR_ENGINE_WATERTIGHT)
        _compare(ref, got, f"watertight/{name}")
        anyh = acc.hit(oo, dd, any_hit=True, engine=N.VR_ENGINE_WATERTIGHT)["hit"]
        assert np.array_equal(anyh.astype(bool), got["prim"] != 0xFFFFFFFF)
        if name == "sphere":
            assert np.all(got["prim"] != 0xFFFFFFFF)
    # renders: extended estimator only
    scene = W.terrain_spheres(grid=24, n_spheres=6, sphere_res=(16, 8), tex_size=64, env_size=(64, 32), sun_peak=200.0)
    acc = scene.build_acceleration(gpu_ctx)
    st = RenderSettings(total_samples=16, max_bounces=8, firefly_clamp=float("inf"))
    ext = N.VR_RENDER_ENV_IS | N.VR_RENDER_RUSSIAN_ROULETTE | N.VR_RENDER_ALPHA_PER_CALL
    imgs = {}
    for eng in (N.VR_ENGINE_WIDE, N.VR_ENGINE_WATERTIGHT):
        img = np.zeros((24, 32, 4), np.float32)
        acc.render(img, make_render_desc(32, 24, 16, 16, st, seed=3, flags=ext, rr_depth=4, engine=eng))
        imgs[eng] = img
    x, r = imgs[N.VR_ENGINE_WATERTIGHT][..., :3], imgs[N.VR_ENGINE_WIDE][..., :3]
    assert np.isfinite(x).all()
    relmse = float(np.mean((x - r) ** 2 / (r * r + 0.01)))
    print(f"[watertight] render vs Moller-Trumbore, same streams: relMSE {relmse:.3e}", flush=True)
    assert relmse < 1e-3
    with pytest.raises(N.VoidrayError, match="watertight"):
        acc.render(np.zeros((4, 4, 4), np.float32),
                   make_render_desc(4, 4, 1, 1, RenderSettings(total_samples=1), engine=N.VR_ENGINE_WATERTIGHT))
