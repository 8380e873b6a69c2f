"""Deterministic synthetic workloads (numpy Philox streams).

* cornell_room      — 512x512, 16 spp, depth 8: the box room of
                      voidray_app/src/examples/cornell.rs:20-55 scaled by 1/555
                      (10 axis-aligned quads), the emissive ceiling quad
                      (radiance 15) and a 32x16 UV sphere; Lambertian albedos
                      U(0.2, 0.8) per object; no environment.
* random_triangles  — traversal microbench: triangles with centres U[0,1]^3
                      and log-uniform edge lengths in [1e-3, 1e-1]; rays half
                      coherent (2048^2 pinhole, pixel centres with x+y even)
                      and half incoherent (uniform origins in the cube,
                      uniform sphere directions).
* terrain_spheres   — 1920x1080 path-integration scene: fBm heightfield,
                      200 UV spheres, mixed materials, procedural textures and
                      HDR sky with a sun lobe.
"""
import numpy as np

from .scene import (Camera, Environments, Materials, Mesh, Scene, SampleType, degrees_to_radians, _normalize)

f32 = np.float32


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def _normalize_rows(v):
    v = v.astype(f32)
    d = ((v[:, 0] * v[:, 0]).astype(f32) + (v[:, 1] * v[:, 1]).astype(f32)).astype(f32)
    d = (d + (v[:, 2] * v[:, 2]).astype(f32)).astype(f32)
    inv = (f32(1.0) / np.sqrt(d).astype(f32)).astype(f32)
    return (v * inv[:, None]).astype(f32)


def uv_sphere(center, radius, slices=32, stacks=16):
    """UV sphere: (stacks+1) x (slices+1) vertices with seam duplicates; pole
    rows produce one triangle per cell (2*slices*(stacks-1) triangles)."""
    c = np.asarray(center, np.float64)
    i = np.arange(stacks + 1)[:, None]
    j = np.arange(slices + 1)[None, :]
    th = np.pi * i / stacks
    ph = 2 * np.pi * j / slices
    n = np.stack([np.sin(th) * np.cos(ph), np.cos(th) * np.ones_like(ph), np.sin(th) * np.sin(ph)], -1).reshape(-1, 3)
    pos = (c + radius * n).astype(f32)
    uv = np.stack([(j / slices) * np.ones_like(th), 1 - (i / stacks) * np.ones_like(ph)], -1).reshape(-1, 2)
    tris = []
    W = slices + 1
    for a in range(stacks):
        for b in range(slices):
            v00, v01, v10, v11 = a * W + b, a * W + b + 1, (a + 1) * W + b, (a + 1) * W + b + 1
            if a != 0:
                tris.append((v00, v10, v01))
            if a != stacks - 1:
                tris.append((v01, v10, v11))
    return Mesh(pos, np.asarray(tris, np.uint32), n.astype(f32), uv.astype(f32))


def _quad_mesh(p0, p1, p2, p3):
    """A quad with real normals and uvs (as an OBJ-loaded quad would have)."""
    P = np.asarray([p0, p1, p2, p3], np.float64)
    n = np.cross(P[1] - P[0], P[3] - P[0])
    n = n / np.linalg.norm(n)
    uv = np.asarray([[0, 0], [0, 1], [1, 1], [1, 0]], f32)
    return Mesh(P.astype(f32), [0, 1, 2, 2, 0, 3], np.tile(n, (4, 1)).astype(f32), uv)


def cornell_room(seed=1, sphere=True):
    """Unit-scale cornell room (cornell.rs geometry / 555)."""
    s = 1.0 / 555.0
    q = lambda *pts: _quad_mesh(*[np.asarray(p) * s for p in pts])
    quads = [
        q((0, 0, 0), (0, 0, 555), (555, 0, 555), (555, 0, 0)),          # floor
        q((0, 0, 0), (0, 555, 0), (0, 555, 555), (0, 0, 555)),          # left wall
        q((555, 0, 0), (555, 0, 555), (555, 555, 555), (555, 555, 0)),  # right wall
        q((0, 0, 555), (0, 555, 555), (555, 555, 555), (555, 0, 555)),  # back wall
        q((0, 555, 0), (555, 555, 0), (555, 555, 555), (0, 555, 555)),  # ceiling
    ]
    # a 5-face block (no bottom) standing on the floor
    x0, x1, z0, z1, h = 130, 295, 65, 230, 165
    quads += [
        q((x0, h, z0), (x0, h, z1), (x1, h, z1), (x1, h, z0)),   # top
        q((x0, 0, z0), (x1, 0, z0), (x1, h, z0), (x0, h, z0)),   # front
        q((x0, 0, z1), (x0, h, z1), (x1, h, z1), (x1, 0, z1)),   # back
        q((x0, 0, z0), (x0, h, z0), (x0, h, z1), (x0, 0, z1)),   # left
        q((x1, 0, z0), (x1, 0, z1), (x1, h, z1), (x1, h, z0)),   # right
    ]
    light = q((213, 554, 227), (343, 554, 227), (343, 554, 332), (213, 554, 332))
    rng = _rng(seed)
    scene = Scene.empty()
    meshes = quads + [light]
    if sphere:
        meshes.append(uv_sphere((555 / 2 * s + 0.12, 100 * s, 555 / 2 * s + 0.1), 100 * s))
    for k, m in enumerate(meshes):
        surf = scene.add_mesh(m)
        if m is light:
            mat = scene.add_material(Materials.emissive(15.0))
        else:
            mat = scene.add_material(Materials.lambertian(rng.uniform(0.2, 0.8, 3)))
        scene.add_object(mat, surf)
    scene.camera = Camera.look_at((278 * s, 278 * s, -800 * s), (278 * s, 278 * s, 0.0), (0.0, 1.0, 0.0),
                                  degrees_to_radians(40.0))
    return scene


def random_triangles(num_tris, seed=2):
    """Triangles with centres U[0,1]^3 and two log-uniform edges in [1e-3, 1e-1]."""
    rng = _rng(seed)
    c = rng.random((num_tris, 3))
    L = np.exp(rng.uniform(np.log(1e-3), np.log(1e-1), (num_tris, 2)))
    dirs = rng.normal(size=(num_tris, 2, 3))
    dirs /= np.linalg.norm(dirs, axis=-1, keepdims=True)
    e1 = dirs[:, 0] * L[:, :1]
    e2 = dirs[:, 1] * L[:, 1:]
    p0 = c - (e1 + e2) / 3
    pos = np.stack([p0, p0 + e1, p0 + e2], 1).reshape(-1, 3).astype(f32)
    idx = np.arange(3 * num_tris, dtype=np.uint32).reshape(-1, 3)
    return Mesh(pos, idx)


def random_triangle_scene(num_tris, seed=2):
    scene = Scene.empty()
    surf = scene.add_mesh(random_triangles(num_tris, seed))
    mat = scene.add_material(Materials.lambertian((0.5, 0.5, 0.5)))
    scene.add_object(mat, surf)
    return scene


def pinhole_rays(res, count, eye=(0.5, 0.5, -1.5), center=(0.5, 0.5, 0.5), fov_deg=45.0):
    """Primary rays through pixel centres with (x + y) even (first `count` of
    them in scanline order), with CameraAcceleration::cast_ray's formula
    (camera.rs:69-82) and the NDC mapping of iterative.rs:30-33, in f32."""
    cam = Camera.look_at(eye, center, (0.0, 1.0, 0.0), degrees_to_radians(fov_deg))
    dcam = f32(f32(1.0) / f32(np.tan(f32(cam.fov / f32(2.0)))))
    # the reference cross product in f32
    dd, up = cam.direction, cam.up
    right = np.array([f32(dd[1] * up[2]) - f32(dd[2] * up[1]), f32(dd[2] * up[0]) - f32(dd[0] * up[2]),
                      f32(dd[0] * up[1]) - f32(dd[1] * up[0])], f32)
    right = _normalize(right)
    idx = np.arange(res * res, dtype=np.int64)
    x = idx % res
    y = idx // res
    sel = ((x + y) % 2) == 0
    x, y = x[sel][:count], y[sel][:count]
    d = f32(res)
    xf = ((2 * x + 1).astype(f32) - d) / d
    yf = ((2 * (res - y) - 1).astype(f32) - d) / d
    v = (dcam * cam.direction).astype(f32)[None, :] + (xf[:, None] * right[None, :]).astype(f32)
    v = (v.astype(f32) + (yf[:, None].astype(f32) * up[None, :]).astype(f32)).astype(f32)
    v = _normalize_rows(_normalize_rows(v))
    o = np.tile(cam.eye, (len(v), 1)).astype(f32)
    return o, v


def incoherent_rays(count, seed=3):
    rng = _rng(seed)
    o = rng.random((count, 3)).astype(f32)
    d = rng.normal(size=(count, 3))
    d = _normalize_rows((d / np.linalg.norm(d, axis=1, keepdims=True)).astype(f32))
    return o, d


def microbench_rays(num_rays, res=2048, seed=3):
    """Half coherent primary rays, half incoherent rays (concatenated)."""
    half = num_rays // 2
    o1, d1 = pinhole_rays(res, half)
    o2, d2 = incoherent_rays(num_rays - half, seed)
    return np.concatenate([o1, o2]), np.concatenate([d1, d2])
