"""The watertight engine (VR_ENGINE_WATERTIGHT: the wide BVH with the
ray/triangle test of Woop, Benthin & Wald, JCGT 2013).  No counterpart in the
reference, whose Triangle::hit is Moller-Trumbore (core/mesh.rs:144-188);
renders use it only under the extended estimator.

Pins, on the host (the kernels' traversal code run by vr_trace_host):
* no ray slips through a closed mesh: rays from inside a UV sphere through
  every vertex and every edge midpoint, and rays from outside aimed at them,
  all hit (Moller-Trumbore, for comparison, loses some of these);
* rays through the interior shared vertices and edges of a coplanar fan all
  hit, straight down and tilted;
* elsewhere it agrees with Moller-Trumbore: the same triangle wherever both
  hit, t within 1e-5 relative, and no ray that Moller-Trumbore hits is lost.
"""
import numpy as np
import pytest

from tests.trace_cases import edge_case_scene
from voidray_amd import _native as N
from voidray_amd import workloads as W
from voidray_amd.device import trace_host
from voidray_amd.scene import Materials, Mesh, Scene

f32 = np.float32


def _scene(mesh):
    sc = Scene.empty()
    sc.add_object(sc.add_material(Materials.lambertian((0.5, 0.5, 0.5))), sc.add_mesh(mesh))
    return sc


def _sphere_targets(mesh):
    """Every vertex and every edge midpoint of the mesh (f32)."""
    P = mesh.positions.astype(np.float64)
    I = mesh.indices
    edges = np.concatenate([I[:, [0, 1]], I[:, [1, 2]], I[:, [2, 0]]])
    mids = (P[edges[:, 0]] + P[edges[:, 1]]) / 2
    return np.concatenate([P, mids])


def _leaks(sc, o, d, engine):
    r = trace_host(sc, o, d, engine=engine)
    return int(np.sum(r["prim"] == 0xFFFFFFFF)), r


def test_no_leaks_through_a_closed_sphere():
    mesh = W.uv_sphere((0.1, -0.2, 0.3), 1.0, 48, 24)
    sc = _scene(mesh)
    c = np.array([0.1, -0.2, 0.3])
    tgt = _sphere_targets(mesh)
    # from inside (the centre and an off-centre point) through each target
    inner = []
    for src in (c, c + np.array([0.13, -0.07, 0.05])):
        dd = tgt - src
        inner.append((np.tile(src, (len(tgt), 1)), dd / np.linalg.norm(dd, axis=1, keepdims=True)))
    # from outside, entering through each target: from a point out along the
    # outward direction at the target, tilted by up to ~30 degrees, so the
    # ray crosses the surface there (not a tangent that only grazes a vertex)
    rng = np.random.default_rng(3)
    out_dir = (tgt - c) / np.linalg.norm(tgt - c, axis=1, keepdims=True)
    far = tgt + 2.0 * W._normalize_rows(out_dir + 0.5 * rng.uniform(-1, 1, (len(tgt), 3)))
    dd = tgt - far
    outer = (far, dd / np.linalg.norm(dd, axis=1, keepdims=True))
    mt_total = 0
    for o, d in inner + [outer]:
        o, d = o.astype(f32), W._normalize_rows(d.astype(f32))
        leaks, _ = _leaks(sc, o, d, N.VR_ENGINE_WATERTIGHT)
        assert leaks == 0, leaks
        mt_total += _leaks(sc, o, d, N.VR_ENGINE_WIDE)[0]
    # the pin has teeth: Moller-Trumbore loses hundreds of these rays
    assert mt_total > 100, mt_total
    print(f"[watertight] Moller-Trumbore leaks on the same rays: {mt_total}")


def test_shared_vertices_and_edges_of_a_fan_hit():
    scene, _, _ = edge_case_scene()
    fan = _scene(scene.meshes[0])
    g = np.linspace(0, 1, 9, dtype=f32)
    pts = np.stack(np.meshgrid(g, g), -1).reshape(-1, 2)
    mids = (pts[:, None, :] + np.array([[0.0625, 0.0], [0.0, 0.0625], [0.0625, 0.0625]], f32)[None]).reshape(-1, 2)
    inside = np.concatenate([pts, mids])
    # interior vertices and edges only: on the fan's outer boundary a tilted
    # ray rounded past the edge rightly misses
    inside = inside[np.all((inside > 0) & (inside < 1), axis=1)]
    rng = np.random.default_rng(4)
    for tilt in (0.0, 0.3):
        o = np.c_[inside, np.zeros(len(inside))].astype(f32)
        d = np.tile(f32([0, 0, 1]), (len(o), 1)) + tilt * rng.normal(size=(len(o), 3)).astype(f32) * f32([1, 1, 0])
        d = W._normalize_rows(d.astype(f32))
        # aim through the target point on z = 0.5
        o = (np.c_[inside, np.full(len(inside), 0.5)] - 0.5 / d[:, 2:3] * d).astype(f32)
        r = trace_host(fan, o, d, engine=N.VR_ENGINE_WATERTIGHT)
        assert np.all(r["prim"] != 0xFFFFFFFF), int(np.sum(r["prim"] == 0xFFFFFFFF))


def test_agrees_with_moller_trumbore_away_from_edges():
    scene = W.random_triangle_scene(20000, seed=5)
    o, d = W.microbench_rays(8192, res=64)
    wt = trace_host(scene, o, d, engine=N.VR_ENGINE_WATERTIGHT)
    mt = trace_host(scene, o, d, engine=N.VR_ENGINE_WIDE)
    both = (wt["prim"] != 0xFFFFFFFF) & (mt["prim"] != 0xFFFFFFFF)
    same = both & (wt["prim"] == mt["prim"])
    # Moller-Trumbore's determinant epsilon (t_min = 1e-5) hides small
    # triangles at grazing angles that the watertight test keeps: those rays
    # may hit a nearer triangle; otherwise the answers are the same
    assert same.sum() >= 0.97 * both.sum()
    assert np.sum((mt["prim"] != 0xFFFFFFFF) & (wt["prim"] == 0xFFFFFFFF)) <= 2
    rel = np.abs(wt["t"][same] - mt["t"][same]) / mt["t"][same]
    assert rel.max() <= 1e-5
    # barycentrics of these long thin triangles are ill-conditioned: the two
    # tests' different rounding shows at ~1e-4 (measured 7.5e-5)
    assert np.allclose(wt["u"][same], mt["u"][same], atol=2e-4) and np.allclose(wt["v"][same], mt["v"][same], atol=2e-4)


def test_render_rejects_watertight_with_the_reference_estimator():
    from voidray_amd.device import make_render_desc
    from voidray_amd.scene import RenderSettings
    d = make_render_desc(8, 8, 1, 1, RenderSettings(total_samples=1), engine=N.VR_ENGINE_WATERTIGHT)
    assert d.engine == N.VR_ENGINE_WATERTIGHT  # the check itself runs in the library (GPU test)
