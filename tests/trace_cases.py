"""Ray-query scenes shared by the GPU and host traversal tests."""
import numpy as np

from voidray_amd.scene import Materials, Mesh, Scene, Surfaces


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


def default_box_scene():
    """Empty surfaces (AABB::default() boxes in the top-level tree) beside
    axis-aligned quads in the plane z = 0.5 (siblings with a flat union box,
    which the reference's slab test rejects for every ray not in the plane:
    aabb.rs:109,126,143) and a 3,000-triangle soup.  Rays: straight up and
    down through the quads, plus random ones.  Within one wave of the exact
    engine, lanes then test default boxes and flat boxes side by side.
    Returns (scene, origins, directions)."""
    from voidray_amd import workloads as W
    f32 = np.float32
    sc = Scene.empty()
    m = sc.add_material(Materials.lambertian((0.5, 0.5, 0.5)))
    empty = Mesh(np.zeros((0, 3), f32), np.zeros((0, 3), np.uint32))
    for k in range(12):
        if k % 3 == 0:
            sc.add_object(m, sc.add_mesh(empty))
        x0 = f32(k * 0.25)
        pos = np.array([[x0, 0, 0.5], [x0 + 0.2, 0, 0.5], [x0 + 0.2, 1, 0.5], [x0, 1, 0.5]], f32)
        sc.add_object(m, sc.add_mesh(Mesh(pos, np.array([[0, 1, 2], [0, 2, 3]], np.uint32))))
    sc.add_object(m, sc.add_mesh(W.random_triangle_scene(3000, seed=13).meshes[0]))
    g = np.linspace(0.01, 2.99, 96, dtype=f32)
    h = np.linspace(0.01, 0.99, 32, dtype=f32)
    pts = np.stack(np.meshgrid(g, h), -1).reshape(-1, 2)
    rng = np.random.default_rng(5)
    o = np.concatenate([np.c_[pts, np.zeros(len(pts))], np.c_[pts, np.ones(len(pts))],
                        rng.uniform(0, 1, (4096, 3)) * [3, 1, 1]]).astype(f32)
    d = np.concatenate([np.tile([0, 0, 1], (len(pts), 1)), np.tile([0, 0, -1], (len(pts), 1)),
                        W._normalize_rows(rng.normal(size=(4096, 3)))]).astype(f32)
    return sc, o, d


def compare_hits(ref, got, label):
    """Bit-exact: prim everywhere, t/u/v wherever the reference hits."""
    miss = ref["prim"] == 0xFFFFFFFF
    bad = np.nonzero(ref["prim"] != got["prim"])[0]
    assert len(bad) == 0, f"{label}: {len(bad)} prim mismatches, first {bad[:8]}"
    h = ~miss
    for k in ("t", "u", "v"):
        diff = np.nonzero(np.asarray(ref[k], np.float32)[h].view(np.uint32) != got[k][h].view(np.uint32))[0]
        assert len(diff) == 0, f"{label}: {len(diff)} {k} mismatches"


def _dot(a, b):
    # (x*x' + y*y') + z*z', every operation rounded to f32 (numpy float32 ufuncs)
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)


def mt_numpy(o, d, p0, p1, p2, tmin=np.float32(1e-5)):
    """Triangle::hit (voidray_renderer/src/core/mesh.rs:144-188) for every
    (ray, triangle) pair, written independently of the oracle: o, d (R, 3),
    p0..p2 (T, 3), all float32.  Returns (hit (R, T) bool, t (R, T))."""
    o = o[:, None, :]
    d = d[:, None, :]
    e1 = (p1 - p0)[None]
    e2 = (p2 - p0)[None]
    h = _cross(d, e2)
    a = _dot(e1, h)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        ok = ~((a > -tmin) & (a < tmin))
        f = np.float32(1.0) / a
        s = o - p0[None]
        u = f * _dot(s, h)
        ok &= ~((u < 0) | (u > 1))
        q = _cross(s, e1)
        v = f * _dot(d, q)
        ok &= ~((v < 0) | (u + v > 1))
        t = f * _dot(e2, q)
        ok &= t > tmin
    return ok, t


def sphere_numpy(o, d, c, r, tmin=np.float32(1e-5)):
    """Sphere::hit (voidray_common/src/surfaces.rs:46-78) with t_max = inf, f32."""
    c = np.asarray(c, np.float32)
    r = np.float32(r)
    oc = o - c
    aa = _dot(d, d)
    half_b = _dot(oc, d)
    cc = _dot(oc, oc) - r * r
    disc = half_b * half_b - aa * cc
    with np.errstate(invalid="ignore"):
        sq = np.sqrt(disc)
        root = (-half_b - sq) / aa
        far = (root < tmin) | (np.float32(np.inf) < root)
        root = np.where(far, (-half_b + sq) / aa, root)
        ok = (disc >= 0) & ~((root < tmin) | (np.float32(np.inf) < root))
    return ok, root


def bruteforce_lowest_handle(scene, o, d):
    """Closest hit over every primitive with no boxes, ties to the smallest
    (surface, triangle) handle pair: the order merge_optionals' tuple compare
    gives (bvh.rs:164-181 with ray.rs:52-62) and the brute-force mesh loop's
    strict `>` (mesh.rs:131).  Returns prim (global id, surface by surface),
    t, and the surface index of the winner."""
    from voidray_amd import _native as N
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    n = len(o)
    best_t = np.full(n, np.inf, np.float32)
    best_p = np.full(n, 0xFFFFFFFF, np.uint32)
    best_s = np.full(n, -1, np.int64)
    base = 0
    for si, sm in enumerate(scene.surfaces):
        if sm & N.VR_SURFACE_ANALYTIC:
            a = scene.analytic[sm & ~N.VR_SURFACE_ANALYTIC]
            assert a.kind == N.VR_ANALYTIC_SPHERE, "brute force covers spheres and meshes"
            ok, t = sphere_numpy(o, d, a.center, a.radius)
            hit = ok[:, None]
            tt = t[:, None]
            ids = np.array([base], np.uint32)
        else:
            m = scene.meshes[sm]
            P = m.positions[m.indices]
            hit, tt = mt_numpy(o, d, P[:, 0], P[:, 1], P[:, 2])
            ids = base + np.arange(len(m.indices), dtype=np.uint32)
        # candidates in handle order: a strict `<` keeps the first of equal t
        for k in range(hit.shape[1]):
            take = hit[:, k] & (tt[:, k] < best_t)
            best_t = np.where(take, tt[:, k], best_t)
            best_p = np.where(take, ids[k], best_p)
            best_s = np.where(take, si, best_s)
        base += len(ids)
    best_t = np.where(best_p == 0xFFFFFFFF, np.float32(0), best_t)
    return dict(prim=best_p, t=best_t, surface=best_s)


def _rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def instanced_scene(n_instances=24, seed=11):
    """Instances of two meshes (a UV sphere with vertex normals, a closed
    six-face box) under random rotations, non-uniform scales (up to 3:1) and,
    for every fifth one, a mirror (det < 0); beside them an ordinary mesh
    surface and an analytic sphere.  Returns (scene, origins, directions) with
    rays from random outside points aimed at random points of random placed
    triangles (most rays hit an instance)."""
    from voidray_amd import workloads as W
    rng = np.random.default_rng(seed)
    scene = Scene.empty()
    mat = scene.add_material(Materials.lambertian((0.5, 0.5, 0.5)))
    sphere = scene.add_instanced_mesh(W.uv_sphere((0.0, 0.0, 0.0), 1.0, 24, 12))
    c = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], np.float32)
    faces = [[0, 1, 3, 2], [4, 6, 7, 5], [0, 4, 5, 1], [2, 3, 7, 6], [0, 2, 6, 4], [1, 5, 7, 3]]
    box_idx = np.array([[f[0], f[1], f[2], f[0], f[2], f[3]] for f in faces], np.uint32).reshape(-1, 3)
    box = scene.add_instanced_mesh(Mesh(c, box_idx))
    plain = scene.add_mesh(W.uv_sphere((0.0, -4.0, 0.0), 0.7, 16, 8))
    scene.add_object(mat, plain)
    scene.add_object(mat, scene.add_analytic_surface(Surfaces.sphere((4.0, 4.0, -4.0), 0.8)))
    for k in range(n_instances):
        R = _rotation(rng)
        S = np.diag(rng.uniform(0.25, 0.75, 3))
        if k % 5 == 4:
            S[0, 0] = -S[0, 0]
        M = R @ S
        t = rng.uniform(-4, 4, 3)
        xf = np.c_[M, t].astype(np.float32)
        scene.add_object(mat, scene.add_instance(sphere if k % 2 == 0 else box, xf))
    flat = scene.flattened()
    # targets: random points of random triangles of the placed surfaces
    tris = []
    for sm in flat.surfaces:
        if sm & 0x80000000:
            continue
        m = flat.meshes[sm]
        tris.append(m.positions[m.indices].astype(np.float64))
    tris = np.concatenate(tris)
    n = 6000
    pick = tris[rng.integers(0, len(tris), n)]
    b = rng.dirichlet(np.ones(3), n)
    tgt = np.einsum("nk,nkc->nc", b, pick)
    src = rng.normal(size=(n, 3))
    src = 12.0 * src / np.linalg.norm(src, axis=1, keepdims=True)
    d = tgt - src
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return scene, src.astype(np.float32), W._normalize_rows(d.astype(np.float32))
