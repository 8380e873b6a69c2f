"""Equal-t ties, checked against a brute force that shares no code with the
oracle or the library.

The reference's BvhNode::hit returns Option<(HitRecord, usize)> and
merge_optionals keeps `left` iff left < right (bvh.rs:143-181).  HitRecord's
PartialEq/PartialOrd look at t only (ray.rs:52-62), so Rust's lexicographic
tuple order falls through to the handle on equal t: the smaller handle wins.
The handle is the triangle index inside a mesh (mesh.rs:269-279) and the
surface index at the top level (scene.rs:72-92, 182-185); the brute-force loop
of small meshes keeps the earlier triangle (mesh.rs:131).  The winner is the
lexicographic minimum of (t, surface, triangle) over the primitives the
reference reaches.

`bruteforce_lowest_handle` (tests/trace_cases.py) computes that minimum in
numpy over every primitive, with no boxes.  On the scenes below every
primitive a ray hits is also reached through the reference trees (no box
rejects it), so the two must agree ray for ray.
"""
import numpy as np
import pytest

import oracle
from tests.trace_cases import bruteforce_lowest_handle, compare_hits, edge_case_scene
from voidray_amd import _native as N
from voidray_amd.device import trace_host
from voidray_amd.scene import Materials, Mesh, Scene, Surfaces

f32 = np.float32


def _strip(n, reverse, z=1.0, x0=0):
    """n unit squares along x from x0, two triangles each, at height z.
    Neighbours share edges, so a ray through a shared edge hits both with the
    same t.  `reverse` lists the squares right to left: triangle index order
    is then the opposite of the median split's in-order leaf order.  One
    unreferenced vertex half a unit above gives the mesh's bounds (which
    count every vertex, mesh.rs:92-104) depth, so a top-level union of
    coplanar strips is not a flat box that the slab test rejects."""
    pos, idx = [], []
    xs = range(n - 1, -1, -1) if reverse else range(n)
    for x in xs:
        k = len(pos)
        x = x + x0
        pos += [[x, 0, z], [x + 1, 0, z], [x + 1, 1, z], [x, 1, z]]
        idx += [[k, k + 1, k + 2], [k, k + 2, k + 3]]
    pos.append([x0, 0, z + 0.5])
    return Mesh(np.asarray(pos, f32), np.asarray(idx, np.uint32))


def _strip_rays(n):
    # through every shared vertical edge (x = 1..n-1), every diagonal and
    # every square's interior, straight down the z axis
    xs = np.concatenate([np.arange(1, n, dtype=f32), np.arange(n, dtype=f32) + f32(0.5),
                         np.arange(n, dtype=f32) + f32(0.25)])
    ys = np.concatenate([np.full(n - 1, 0.5, f32), np.full(n, 0.5, f32), np.full(n, 0.75, f32)])
    o = np.stack([xs, ys, np.zeros_like(xs)], 1).astype(f32)
    d = np.tile(f32([0, 0, 1]), (len(o), 1))
    return o, d


def _scenes():
    out = {}
    for reverse in (False, True):
        sc = Scene.empty()
        m = sc.add_material(Materials.lambertian((0.5, 0.5, 0.5)))
        sc.add_object(m, sc.add_mesh(_strip(12, reverse)))
        out[f"strip-reverse{int(reverse)}"] = (sc, *_strip_rays(12))
    # two coplanar surfaces sharing the edge x = 6: a tie across surfaces
    # goes to surface 0 whichever side of the top-level tree it sits on
    for first_right in (False, True):
        sc = Scene.empty()
        m = sc.add_material(Materials.lambertian((0.5, 0.5, 0.5)))
        left, right = _strip(6, False), _strip(6, False, x0=6)
        for mesh in ((right, left) if first_right else (left, right)):
            sc.add_object(m, sc.add_mesh(mesh))
        # a third, small (brute-force) mesh: a coincident copy of one square
        sc.add_object(m, sc.add_mesh(Mesh(f32([[2, 0, 1], [3, 0, 1], [3, 1, 1], [2, 1, 1], [2, 0, 1.5]]),
                                          [[0, 1, 2], [0, 2, 3]])))
        out[f"two-surfaces-right{int(first_right)}"] = (sc, *_strip_rays(12))
    # coincident spheres, and a sphere coincident with nothing
    sc = Scene.empty()
    m = sc.add_material(Materials.lambertian((0.5, 0.5, 0.5)))
    for c in ((0, 0, 0), (0, 0, 0), (3, 0, 0), (0, 0, 0)):
        sc.add_object(m, sc.add_analytic_surface(Surfaces.sphere(c, 1.0)))
    rng = np.random.default_rng(1)
    o = np.concatenate([np.tile(f32([0, 5, 0]), (64, 1)) + rng.normal(0, 0.3, (64, 3)).astype(f32),
                        np.tile(f32([3, 5, 0]), (16, 1))]).astype(f32)
    d = np.tile(f32([0, -1, 0]), (len(o), 1))
    out["coincident-spheres"] = (sc, o, d)
    sc, o, d = edge_case_scene()
    out["edge-case-scene"] = (sc, o, d)
    fan = Scene.empty()
    fan.add_object(fan.add_material(Materials.lambertian((0.5, 0.5, 0.5))), fan.add_mesh(sc.meshes[0]))
    out["edge-case-fan"] = (fan, o, d)
    return out


SCENES = _scenes()


@pytest.mark.parametrize("name", sorted(SCENES))
def test_oracle_matches_lowest_handle_bruteforce(name):
    sc, o, d = SCENES[name]
    ref = bruteforce_lowest_handle(sc, o, d)
    got = oracle.OracleScene(sc).trace(o, d)
    bad = np.nonzero(ref["prim"] != got["prim"])[0]
    assert len(bad) == 0, f"{name}: {len(bad)} of {len(o)} rays differ, first {bad[:8]}"
    hit = ref["prim"] != 0xFFFFFFFF
    assert np.array_equal(ref["t"][hit].view(np.uint32), got["t"][hit].view(np.uint32))


@pytest.mark.parametrize("name", sorted(SCENES))
def test_host_engines_match_lowest_handle_bruteforce(name):
    # the library's own traversal code, run on the host over the uploaded layout
    sc, o, d = SCENES[name]
    ref = bruteforce_lowest_handle(sc, o, d)
    for engine in (N.VR_ENGINE_WIDE, N.VR_ENGINE_REFERENCE):
        got = trace_host(sc, o, d, engine=engine)
        bad = np.nonzero(ref["prim"] != got["prim"])[0]
        assert len(bad) == 0, f"{name}/engine{engine}: {len(bad)} rays differ, first {bad[:8]}"


def test_ties_exercise_the_rule():
    # the scenes must actually contain ties that the old in-order rule (the
    # later leaf of the median split wins) would settle differently; count
    # them with the oracle's brute force under that rule
    changed = 0
    for name, (sc, o, d) in SCENES.items():
        orc = oracle.OracleScene(sc)
        if any(s & N.VR_SURFACE_ANALYTIC for s in sc.surfaces) or len(sc.surfaces) > 1:
            continue
        order = orc.leaf_order(0) if sc.meshes[0].num_triangles > 4 else np.arange(sc.meshes[0].num_triangles)
        rank = np.zeros(len(order), np.uint32)
        rank[order] = np.arange(len(order))
        late = orc.trace_bruteforce(o, d, rank)
        changed += int(np.sum(late["prim"] != bruteforce_lowest_handle(sc, o, d)["prim"]))
    assert changed >= 150, changed


def test_edge_case_scene_exact_answers():
    # rays through the shared vertex (0.25, 0.25) of the 8x8 fan at z = 0.5:
    # eight triangles meet there, and the smallest index among them wins
    sc, _, _ = edge_case_scene()
    o = f32([[0.25, 0.25, 0.0]])
    d = f32([[0, 0, 1]])
    got = oracle.OracleScene(sc).trace(o, d)
    m = sc.meshes[0]
    P = m.positions[m.indices]
    lo, hi = P.min(1), P.max(1)
    touching = np.nonzero((lo[:, 0] <= 0.25) & (hi[:, 0] >= 0.25) & (lo[:, 1] <= 0.25) & (hi[:, 1] >= 0.25)
                          & (lo[:, 2] == 0.5))[0]
    assert len(touching) >= 6
    assert got["prim"][0] == touching.min() and got["t"][0] == 0.5
