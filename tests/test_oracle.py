"""Pins the oracle (CPU restatement of voidray's path) against things that do
not come from the oracle itself: published known-answer vectors, closed forms,
brute force, and an independent re-implementation of the median-split tree.

The reference ships no tests for this path (its only unit test is
voidray_renderer/src/util/vector.rs:36-43), so these checks are the pin.
"""
import math

import numpy as np
import pytest

import oracle
from voidray_amd import workloads as W
from voidray_amd import _native as N
from voidray_amd.device import make_render_desc
from voidray_amd.scene import Camera, Environments, Materials, Mesh, RenderSettings, Scene, SampleType

f32 = np.float32


# ---- Philox4x32-10: Random123 known-answer vectors (kat_vectors) -------------
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, expect):
    assert list(oracle.philox(ctr, key)) == expect


def test_path_stream_layout():
    # word k of a path is lane k % 4 of Philox(counter=(k/4, sample, pixel, request), key=seed)
    w = oracle.path_words(0x1234_5678_9abc_def0, 3, 17, 5, 12)
    for k in range(12):
        blk = oracle.philox([k // 4, 5, 17, 3], [0x9abcdef0, 0x12345678])
        assert w[k] == blk[k % 4]


# ---- Möller–Trumbore / slab test: closed-form hits ---------------------------
def _one_tri_scene(p0, p1, p2, normals=None):
    s = Scene.empty()
    m = Mesh(np.array([p0, p1, p2], f32), [[0, 1, 2]], normals)
    s.add_object(s.add_material(Materials.lambertian((0.5, 0.5, 0.5))), s.add_mesh(m))
    return s


def test_triangle_closed_form_hit():
    # unit right triangle in z = 2, ray along +z: t = 2, u = x, v = y (mesh.rs:144-188)
    sc = oracle.OracleScene(_one_tri_scene([0, 0, 2], [1, 0, 2], [0, 1, 2]))
    o = np.array([[0.25, 0.5, 0.0], [0.1, 0.1, 0.0], [0.8, 0.8, 0.0], [0.25, 0.5, 3.0]], f32)
    d = np.array([[0, 0, 1]] * 4, f32)
    r = sc.trace(o, d)
    assert list(r["prim"][:2]) == [0, 0] and r["prim"][2] == 0xFFFFFFFF and r["prim"][3] == 0xFFFFFFFF
    assert r["t"][0] == 2.0 and r["u"][0] == 0.25 and r["v"][0] == 0.5


def test_determinant_epsilon_is_tmin():
    # a = e1 . (d x e2); with |a| < t_min the triangle is invisible (mesh.rs:156)
    s = 3.0e-3  # a = s*s = 9e-6 < 1e-5
    sc = oracle.OracleScene(_one_tri_scene([0, 0, 1], [s, 0, 1], [0, s, 1]))
    r = sc.trace(np.array([[s / 4, s / 4, 0]], f32), np.array([[0, 0, 1]], f32))
    assert r["prim"][0] == 0xFFFFFFFF
    s = 4.0e-3  # a = 1.6e-5 > 1e-5
    sc = oracle.OracleScene(_one_tri_scene([0, 0, 1], [s, 0, 1], [0, s, 1]))
    r = sc.trace(np.array([[s / 4, s / 4, 0]], f32), np.array([[0, 0, 1]], f32))
    assert r["prim"][0] == 0


def test_edge_and_vertex_hits_are_inclusive():
    sc = oracle.OracleScene(_one_tri_scene([0, 0, 1], [1, 0, 1], [0, 1, 1]))
    o = np.array([[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [0.0, 0.5, 0.0], [0.5, 0.5, 0.0]], f32)
    r = sc.trace(o, np.array([[0, 0, 1]] * 4, f32))
    assert np.all(r["prim"] == 0)


def test_behind_origin_rejected():
    sc = oracle.OracleScene(_one_tri_scene([0, 0, -1], [1, 0, -1], [0, 1, -1]))
    r = sc.trace(np.array([[0.2, 0.2, 0.0]], f32), np.array([[0, 0, 1]], f32))
    assert r["prim"][0] == 0xFFFFFFFF


# ---- median-split tree: invariants + independent build -----------------------
def _py_median_split(boxes):
    """Independent recursive restatement of BvhNode::from_list (bvh.rs:48-130):
    returns the in-order leaf sequence."""
    def key(v):
        # f32::total_cmp: negative floats have all bits but the sign flipped
        b = int(np.float32(v).view(np.int32))
        return b ^ 0x7FFFFFFF if b < 0 else b

    def build(objs):
        if len(objs) == 1:
            return [objs[0]]
        c = [((boxes[o][0] + boxes[o][1]) / f32(2)).astype(f32) for o in objs]
        cmin = np.min(c, axis=0)
        cmax = np.max(c, axis=0)
        sp = (cmax - cmin).astype(f32)
        if sp[0] > sp[1] and sp[0] > sp[2]:
            ax = 0
        elif sp[1] > sp[0] and sp[1] > sp[2]:
            ax = 1
        else:
            ax = 2
        order = sorted(range(len(objs)), key=lambda i: (key(c[i][ax]), i))
        objs = [objs[i] for i in order]
        h = len(objs) // 2
        return build(objs[:h]) + build(objs[h:])
    return build(list(range(len(boxes))))


def _tri_boxes(mesh):
    P = mesh.positions[mesh.indices]  # (T, 3, 3)
    lo = P.min(axis=1).astype(f32)
    hi = P.max(axis=1).astype(f32)
    eps = f32(0.001)
    c = ((lo + hi) / f32(2)).astype(f32)
    dim = (hi - lo).astype(f32)
    m = dim < eps
    lo = np.where(m, (c - eps).astype(f32), lo)
    hi = np.where(m, (c + eps).astype(f32), hi)
    return [(lo[i], hi[i]) for i in range(len(lo))]


def test_median_split_matches_independent_build():
    mesh = W.random_triangles(300, seed=21)
    sc = Scene.empty()
    sc.add_object(sc.add_material(Materials.lambertian((1, 1, 1))), sc.add_mesh(mesh))
    got = list(oracle.OracleScene(sc).leaf_order(0))
    assert got == _py_median_split(_tri_boxes(mesh))


def test_median_split_invariants():
    mesh = W.random_triangles(1000, seed=22)
    sc = Scene.empty()
    sc.add_object(sc.add_material(Materials.lambertian((1, 1, 1))), sc.add_mesh(mesh))
    orc = oracle.OracleScene(sc)
    boxes, links = orc.mesh_tree(0)
    tb = _tri_boxes(mesh)
    leaves = 0

    def walk(i):
        nonlocal leaves
        l, r, obj = links[i]
        if obj >= 0:
            leaves += 1
            return tb[obj][0], tb[obj][1], 1
        a = walk(l)
        b = walk(r)
        lo = np.minimum(a[0], b[0])
        hi = np.maximum(a[1], b[1])
        assert np.array_equal(boxes[i, :3], lo) and np.array_equal(boxes[i, 3:], hi)  # union box
        assert a[2] == (a[2] + b[2]) // 2  # left = floor(n / 2)
        return lo, hi, a[2] + b[2]
    walk(0)
    assert leaves == 1000
    assert sorted(orc.leaf_order(0)) == list(range(1000))


def test_traversal_equals_brute_force_with_handle_ties():
    # the oracle's brute force resolves equal t by the rank it is given:
    # rank = -handle is the reference's rule (smallest triangle wins)
    scene = W.random_triangle_scene(3000, seed=23)
    orc = oracle.OracleScene(scene)
    rank = np.arange(3000, dtype=np.uint32)[::-1].copy()
    o, d = W.microbench_rays(4000, res=64)
    a = orc.trace(o, d)
    b = orc.trace_bruteforce(o, d, rank)
    assert np.array_equal(a["prim"], b["prim"]) and np.array_equal(a["t"], b["t"])


def test_equal_t_tie_keeps_smallest_triangle():
    # six coincident triangles: every leaf returns the same t, and
    # merge_optionals compares (HitRecord, handle) tuples whose HitRecord part
    # orders by t alone (bvh.rs:164-181, ray.rs:52-62), so the smallest
    # triangle index wins.  The brute-force small-mesh loop keeps the earlier
    # triangle with its strict `>` (mesh.rs:131): the same answer.
    tri = [[0, 0, 1], [1, 0, 1], [0, 1, 1]]
    for copies in (6, 3):  # 6: the mesh has a tree (> 4 tris); 3: brute force
        pos = np.array(tri * copies, f32)
        idx = np.arange(3 * copies, dtype=np.uint32).reshape(copies, 3)
        sc = Scene.empty()
        sc.add_object(sc.add_material(Materials.lambertian((1, 1, 1))), sc.add_mesh(Mesh(pos, idx)))
        orc = oracle.OracleScene(sc)
        r = orc.trace(np.array([[0.2, 0.2, 0]], f32), np.array([[0, 0, 1]], f32))
        assert r["prim"][0] == 0 and r["t"][0] == 1.0, copies


def test_flat_split_box_hides_surfaces():
    # quirk: two coplanar quads that are top-level siblings give a flat union box
    # (scene.rs:77 uses unexpanded mesh bounds) and the slab test rejects every ray
    sc = Scene.empty()
    from voidray_amd.scene import Surfaces
    for x0 in (0.0, 2.0):
        q = Surfaces.quad((x0, 0, 1), (x0 + 1, 0, 1), (x0 + 1, 1, 1), (x0, 1, 1))
        sc.add_object(sc.add_material(Materials.lambertian((1, 1, 1))), sc.add_mesh(q))
    r = oracle.OracleScene(sc).trace(np.array([[0.5, 0.5, 0]], f32), np.array([[0, 0, 1]], f32))
    assert r["prim"][0] == 0xFFFFFFFF


def test_golden_fixture_reproduces():
    fx = np.load("tests/golden/c2_small.npz")
    scene = W.random_triangle_scene(int(fx["num_tris"]), seed=int(fx["seed"]))
    o, d = W.microbench_rays(int(fx["num_rays"]), res=int(fx["res"]))
    assert np.array_equal(o, fx["origins"]) and np.array_equal(d, fx["directions"])
    r = oracle.OracleScene(scene).trace(o, d)
    for k in ("prim", "t", "u", "v"):
        assert np.array_equal(r[k], fx[k])


# ---- distributions ------------------------------------------------------------
def test_unit_sphere_is_uniform():
    s = oracle.unit_sphere_samples(5, 200000).astype(np.float64)
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-6)
    assert np.all(np.abs(s.mean(0)) < 0.01)
    assert np.allclose((s * s).mean(0), 1 / 3, atol=0.01)
    # z is uniform on [-1, 1] for the uniform sphere (Archimedes)
    hist, _ = np.histogram(s[:, 2], bins=20, range=(-1, 1))
    exp = len(s) / 20
    chi2 = float(((hist - exp) ** 2 / exp).sum())
    assert chi2 < 45.0  # 19 dof, p ~ 1e-3


# ---- integrator special cases ---------------------------------------------------
def test_uniform_env_no_geometry_returns_env():
    sc = Scene.empty()
    sc.environment = Environments.uniform((0.3, 0.6, 5.0))
    img, _, _ = oracle.OracleScene(sc).render(make_render_desc(4, 4, 3, 3, RenderSettings(total_samples=3)))
    # 3 samples of the constant, times 1/3
    exp = (f32(0.3) + f32(0.3) + f32(0.3)) * (f32(1) / f32(3))
    assert np.all(img[..., 0] == exp)


def test_emissive_clamp_and_max_depth():
    sc = W.cornell_room()
    for mb in (1,):
        img, _, st = oracle.OracleScene(sc).render(
            make_render_desc(16, 16, 1, 1, RenderSettings(total_samples=1, max_bounces=mb)))
        # with one bounce only directly visible emitters contribute, clamped to 3
        vals = np.unique(img[..., :3])
        assert set(vals.tolist()) <= {0.0, 3.0}


def test_fp64_oracle_agrees_with_f32_on_average():
    sc = W.cornell_room()
    desc = make_render_desc(24, 24, 16, 16, RenderSettings(total_samples=16, max_bounces=8), seed=5)
    a, _, _ = oracle.OracleScene(sc).render(desc)
    b, _, _ = oracle.OracleScene(sc, "f64").render(desc)
    assert abs(float(a[..., :3].mean()) - float(b[..., :3].mean())) < 0.02 * float(b[..., :3].mean())


# ---- textures / env lookups: hand-built images with known bilinear answers ------
def _tex_scene(img, sample_type):
    sc = Scene.empty()
    t = sc.add_image_texture(img, sample_type)
    m = Mesh(np.array([[0, 0, 1], [1, 0, 1], [0, 1, 1]], f32), [[0, 1, 2]],
             np.tile([0, 0, -1], (3, 1)).astype(f32), np.array([[0, 0], [1, 0], [0, 1]], f32))
    sc.add_object(sc.add_material(Materials.lambertian_texture_no_normal(t)), sc.add_mesh(m))
    return sc


def _tex(img, sample_type):
    import ctypes as C
    img = np.ascontiguousarray(img, f32)
    t = N.vr_texture(img.shape[1], img.shape[0], N.VR_TEX_RGB32F, sample_type, img.ctypes.data)
    def sample(u, v):
        out = np.zeros(3, f32)
        oracle.lib().orc_tex_sample_f32(C.byref(t), C.c_float(u), C.c_float(v), out.ctypes.data_as(C.c_void_p))
        return out
    return sample


def test_texture_known_values():
    # texel colours: texel k = (k, 10k, 100k) in a 2 (wide) x 2 image, row-major
    img = np.array([[[0, 0, 0], [1, 10, 100]], [[2, 20, 200], [3, 30, 300]]], f32)
    near = _tex(img, SampleType.Nearest)
    bil = _tex(img, SampleType.Bilinear)
    # x = u*W, y = (1 - v)*H  (texture.rs:91-92): (u, v) = (0.25, 0.75) -> (0.5, 0.5): texel 0
    assert np.array_equal(near(0.25, 0.75), img[0, 0])
    # (0.75, 0.25) -> (1.5, 1.5): texel 3
    assert np.array_equal(near(0.75, 0.25), img[1, 1])
    # negative u wraps: u -= trunc(u) - 1  -> -0.25 becomes 0.75
    assert np.array_equal(near(-0.25, 0.25), img[1, 1])
    # v = 1.0 -> v % 1 = 0 -> y = H clamps to the last row
    assert np.array_equal(near(0.25, 1.0), img[1, 0])
    # bilinear at (x, y) = (0.5, 0.5): x0 = y0 = 0, a = 0.5: mean of texels 0,1,2,3
    assert np.allclose(bil(0.25, 0.75), img.reshape(-1, 3).mean(0), rtol=1e-6)
    # bilinear at the last column: x0 + 1 spills into the next row (index % len)
    # (x, y) = (1.5, 0.5): taps 1, 2 (spill), 3, 4 % 4 = 0 with weights 1/4 each
    exp = (img[0, 1] + img[1, 0] + img[1, 1] + img[0, 0]) / 4
    assert np.allclose(bil(0.75, 0.75), exp, rtol=1e-6)


def test_environment_lookup_known_values():
    import ctypes as C
    W_, H_ = 4, 2
    img = np.arange(W_ * H_ * 3, dtype=f32).reshape(H_, W_, 3)
    def env(d):
        out = np.zeros(3, f32)
        dd = np.asarray(d, f32)
        oracle.lib().orc_env_sample_f32(W_, H_, img.ctypes.data_as(C.c_void_p), dd.ctypes.data_as(C.c_void_p),
                                        out.ctypes.data_as(C.c_void_p))
        return out
    # straight down: theta = acos(1) = 0 -> y = H - 1 = 1; phi = atan2(-0, 0) + pi = pi -> x = W/2 = 2
    assert np.array_equal(env([0, -1, 0]), img[1, 2])
    # along +x: theta = pi/2 -> y = 1 - 1 = 0; phi = atan2(0, 1) + pi = pi -> x = 2; texel (2, 0)
    assert np.allclose(env([1, 0, 0]), img[0, 2], atol=1e-5)


def test_tonemap_curves_fp64():
    # each published curve evaluated in FP64 at fixed inputs
    c = np.array([[[0.2, 0.5, 1.5, 0.0], [2.0, 0.01, 0.7, 0.0]]], f32)
    def run(mode, gamma=1.0, exposure=0.0, scale=1.0):
        return oracle.tonemap(c, N.vr_post_desc(2, 1, scale, gamma, exposure, mode, 0))
    x = c[..., :3].astype(np.float64)
    # None, gamma 1: identity; alpha = 1
    assert np.allclose(run(0)[..., :3], x) and np.all(run(0)[..., 3] == 1)
    # Filmic (Hejl-Burgess-Dawson)
    y = np.maximum(0, x - 0.004)
    assert np.allclose(run(3)[..., :3], (y * (6.2 * y + .5)) / (y * (6.2 * y + 1.7) + 0.06), rtol=1e-6)
    # Uncharted 2 (Hable), exposure bias 2, W = 11.2
    A, B, C, D, E, F = 0.15, 0.5, 0.1, 0.2, 0.02, 0.3
    h = lambda v: ((v * (A * v + C * B) + D * E) / (v * (A * v + B) + D * F)) - E / F
    assert np.allclose(run(4)[..., :3], h(2 * x) / h(11.2), rtol=1e-5)
    # Reinhard with white = 2 on luma
    l = x @ np.array([0.2126, 0.7152, 0.0722])
    tl = l * (1 + l / 4) / (1 + l)
    assert np.allclose(run(2)[..., :3], x * (tl / l)[..., None], rtol=1e-5)
    # ACES (Narkowicz/Hill fit with the 3x3 input/output matrices)
    M1 = np.array([[0.59719, 0.35458, 0.04823], [0.076, 0.90834, 0.01566], [0.0284, 0.13383, 0.83777]])
    M2 = np.array([[1.60475, -0.53108, -0.07367], [-0.10208, 1.10813, -0.00605], [-0.00327, -0.07276, 1.07602]])
    v = x @ M1.T
    v = (v * (v + 0.0245786) - 0.000090537) / (v * (0.983729 * v + 0.432951) + 0.238081)
    assert np.allclose(run(1)[..., :3], v @ M2.T, rtol=1e-5, atol=1e-6)
    # gamma and exposure: pow(c * scale * 2^e, 1/gamma)
    assert np.allclose(run(0, gamma=2.2, exposure=1.0, scale=0.5)[..., :3], (x * 0.5 * 2.0) ** (1 / 2.2), rtol=1e-5)


def test_legacy_viewer_tonemaps_fp64():
    # voidray/src/shaders/viewport_fragment.glsl + tonemapping.glsl, FP64 closed forms
    c = np.array([[[0.2, 0.5, 1.5, 0.8], [2.0, 0.01, 0.7, 0.25]]], f32)
    def run(mode, gamma=1.0, exposure=1.0, scale=1.0, keep_alpha=0):
        return oracle.tonemap(c, N.vr_post_desc(2, 1, scale, gamma, exposure, mode, 0, keep_alpha))
    x = c[..., :3].astype(np.float64)
    a = c[..., 3].astype(np.float64)
    # linear exposure and scale, gamma, alpha 1 unless kept (then scale * alpha)
    out = run(N.VR_TONEMAP_LEGACY_NONE, gamma=2.2, exposure=3.0, scale=0.5)
    assert np.allclose(out[..., :3], (x * 1.5) ** (1 / 2.2), rtol=1e-6) and np.all(out[..., 3] == 1)
    out = run(N.VR_TONEMAP_LEGACY_NONE, scale=0.5, keep_alpha=1)
    assert np.allclose(out[..., 3], a * 0.5, rtol=1e-7)
    # simple Reinhard with its own exposure 1.5
    assert np.allclose(run(N.VR_TONEMAP_LEGACY_SIMPLE_REINHARD)[..., :3], x * 1.5 / (1 + x / 1.5), rtol=1e-6)
    # luma Reinhard
    l = x @ np.array([0.2126, 0.7152, 0.0722])
    assert np.allclose(run(N.VR_TONEMAP_LEGACY_LUMA_REINHARD)[..., :3], x * ((l / (1 + l)) / l)[..., None], rtol=1e-6)
    # white-preserving luma Reinhard, white = 2
    tl = l * (1 + l / 4) / (1 + l)
    assert np.allclose(run(N.VR_TONEMAP_LEGACY_LUMA_WHITE_PRESERVING_REINHARD)[..., :3], x * (tl / l)[..., None],
                       rtol=1e-6)
    # RomBinDaHouse
    assert np.allclose(run(N.VR_TONEMAP_LEGACY_ROM_BIN_DA_HOUSE)[..., :3], np.exp(-1 / (2.72 * x + 0.15)), rtol=1e-6)
    # Filmic ignores the gamma setting
    y = np.maximum(0, x - 0.004)
    assert np.allclose(run(N.VR_TONEMAP_LEGACY_FILMIC, gamma=2.2)[..., :3],
                       (y * (6.2 * y + .5)) / (y * (6.2 * y + 1.7) + 0.06), rtol=1e-6)
    # ACES and Uncharted 2 are the curves of the current pass; exposure 1 here = 2^0 there
    assert np.array_equal(run(N.VR_TONEMAP_LEGACY_SIMPLE_ACES)[..., :3], run_new(c, N.VR_TONEMAP_ACES)[..., :3])
    assert np.array_equal(run(N.VR_TONEMAP_LEGACY_UNCHARTED2, gamma=2.2)[..., :3],
                          run_new(c, N.VR_TONEMAP_UNCHARTED2, gamma=2.2)[..., :3])


def run_new(c, mode, gamma=1.0):
    return oracle.tonemap(c, N.vr_post_desc(c.shape[1], c.shape[0], 1.0, gamma, 0.0, mode, 0))


def test_camera_look_at_is_orthonormal():
    cam = Camera.look_at((1, 2, 3), (0, 0, 0), (0, 1, 0), 0.7)
    assert abs(float(np.dot(cam.direction, cam.up))) < 1e-6
    assert abs(float(np.linalg.norm(cam.direction)) - 1) < 1e-6
    assert abs(float(np.linalg.norm(cam.up)) - 1) < 1e-6


def _analytic_scene(*surfaces):
    from voidray_amd.scene import Surfaces  # noqa: F401
    scene = Scene.empty()
    m = scene.add_material(Materials.lambertian((0.5, 0.5, 0.5)))
    for s in surfaces:
        scene.add_object(m, scene.add_analytic_surface(s))
    return scene


def test_sphere_closed_form():
    # Sphere::hit (surfaces.rs:46-78): nearest root >= t_min, else the far one
    from voidray_amd.scene import Surfaces
    orc = oracle.OracleScene(_analytic_scene(Surfaces.sphere((0.0, 0.0, 0.0), 1.0)))
    o = np.float32([[0, 5, 0], [0, 0, 0], [0, 5, 0], [3, 0.5, 0], [0, 0, 5]])
    d = np.float32([[0, -1, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, 0, 1]])
    r = orc.trace(o, d)
    assert r["prim"].tolist() == [0, 0, 0xFFFFFFFF, 0, 0xFFFFFFFF]
    assert r["t"][0] == 4.0                       # outside: near root
    assert r["t"][1] == 1.0                       # inside: the far root
    assert abs(r["t"][3] - (3.0 - math.sqrt(0.75))) < 1e-6
    assert r["u"][0] == 0.0 and r["v"][0] == 0.0  # uv (0, 0)


def test_ground_plane_closed_form():
    # GroundPlane::hit (surfaces.rs:88-104): t = (h - o.y) / d.y, strict bounds, uv = (x, z)
    from voidray_amd.scene import Surfaces
    orc = oracle.OracleScene(_analytic_scene(Surfaces.ground_plane(0.5)))
    s = np.float32(1 / math.sqrt(2))
    o = np.float32([[1, 2.5, -3], [0, 0.5, 0], [0, 1, 0], [0, 2, 0], [0, 0.5, 0]])
    d = np.float32([[0, -1, 0], [0, -1, 0], [1, 0, 0], [s, -s, 0], [1, 0, 0]])
    r = orc.trace(o, d)
    assert r["prim"].tolist() == [0, 0xFFFFFFFF, 0xFFFFFFFF, 0, 0xFFFFFFFF]
    assert r["t"][0] == 2.0 and r["u"][0] == 1.0 and r["v"][0] == -3.0
    # on the plane, moving down: t = 0 <= t_min -> miss; parallel rays miss (t = inf, or 0/0)
    assert abs(r["t"][3] - 1.5 / s) < 1e-5 and abs(r["u"][3] - 1.5) < 1e-5


def test_coincident_spheres_tie_keeps_smallest_surface():
    # two identical spheres give equal t; the top-level merge compares
    # (HitRecord, surface index) tuples (scene.rs:182-185, bvh.rs:164-181), so
    # surface 0 wins whichever leaf comes later in the tree
    from voidray_amd.scene import Surfaces
    orc = oracle.OracleScene(_analytic_scene(Surfaces.sphere((0.0, 0.0, 0.0), 1.0),
                                             Surfaces.sphere((0.0, 0.0, 0.0), 1.0)))
    r = orc.trace(np.float32([[0, 5, 0], [0.3, 0.2, 4]]), np.float32([[0, -1, 0], [0, 0, -1]]))
    assert r["prim"].tolist() == [0, 0] and r["t"][0] == 4.0


def test_microfacet_bsdf_closed_form():
    # MicrofacetBSDF::bsdf (microfacet.rs:120-213) at normal incidence:
    # h = n, D = 1/(pi m^2), Schlick F = F0 = ((ior-1)/(ior+1))^2, G = 1:
    # Cook-Torrance specular F0 / (4 pi m^2) plus Lambert (1 - F0) c / pi
    z = np.array([0.0, 0.0, 1.0])
    c = np.array([0.8, 0.5, 0.2])
    f0 = (0.5 / 2.5) ** 2
    for m in (0.5, 0.2):
        f = oracle.microfacet_bsdf(z, z, z, index=1.5, roughness=m, color=c)
        expect = f0 / (4 * math.pi * m * m) + (1 - f0) * c / math.pi
        assert np.allclose(f, expect, rtol=1e-12)
    # metallic = 1: F0 = colour, no diffuse energy left at normal incidence
    f = oracle.microfacet_bsdf(z, z, z, index=1.5, roughness=0.5, metallic=1.0, color=c)
    assert np.allclose(f, c / (4 * math.pi * 0.25) + (1 - c) * c / math.pi, rtol=1e-12)
    # opaque: nothing below the surface
    assert np.all(oracle.microfacet_bsdf(z, z, -z, roughness=0.5) == 0)
    # straight-through transmission (documented BTDF, eta_o = 1, eta_i = 1.5):
    # h = -n, |h.wi||h.wo|/(|n.wi||n.wo|) = 1, (1 - F0) D G / (1.5 - 1)^2 = 4 D (1 - F0)
    f = oracle.microfacet_bsdf(z, z, -z, index=1.5, roughness=0.5, transparent=True, color=c)
    assert np.allclose(f, 4 * (1 / (math.pi * 0.25)) * (1 - f0) * c, rtol=1e-12)


def test_microfacet_sampling_density():
    # MicrofacetBSDF::sample (microfacet.rs:215-312), opaque, normal +z (where
    # the reference's local_to_world is a rotation about the normal): the
    # returned pdf is the density of the sampled directions, so
    # E[1{wi.n > 0} / pdf] = 2 pi
    z = np.array([0.0, 0.0, 1.0])
    wo = np.array([0.3, 0.1, 0.9])
    wo /= np.linalg.norm(wo)
    for rough in (0.3, 0.7):
        wi, pdf = oracle.microfacet_sample(z, wo, 200000, roughness=rough, color=(0.6, 0.6, 0.6))
        assert np.all(pdf > 0)
        est = np.mean((wi[:, 2] > 0) / pdf)
        assert abs(est / (2 * math.pi) - 1) < 0.02, (rough, est)
