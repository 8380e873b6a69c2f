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
* serving_*         — batched serving: 64 requests (resolutions from 256^2 to
                      1920x1080, 1-32 spp, own seeds) over eight scenes of 10k
                      to 2M triangles from the two generators above.
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


def _sphere_topology(slices, stacks):
    W = slices + 1
    a, b = np.meshgrid(np.arange(stacks), np.arange(slices), indexing="ij")
    v00, v01, v10, v11 = a * W + b, a * W + b + 1, (a + 1) * W + b, (a + 1) * W + b + 1
    t1 = np.stack([v00, v10, v01], -1)[1:].reshape(-1, 3)        # skipped on the top pole row
    t2 = np.stack([v01, v10, v11], -1)[:-1].reshape(-1, 3)       # skipped on the bottom pole row
    # interleave per cell in row order: (t1, t2) of cell (a, b), as a scanline mesh writer would
    rows = []
    for r in range(stacks):
        if r == 0:
            rows.append(t2[:slices])
        elif r == stacks - 1:
            rows.append(t1[(r - 1) * slices:r * slices])
        else:
            x = np.empty((2 * slices, 3), np.int64)
            x[0::2] = t1[(r - 1) * slices:r * slices]
            x[1::2] = t2[r * slices:(r + 1) * slices]
            rows.append(x)
    return np.concatenate(rows).astype(np.uint32)


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
    return Mesh(pos, _sphere_topology(slices, stacks), n.astype(f32), uv.astype(f32))


def fbm_heightfield(n, octaves=6, seed=4, amplitude=0.25):
    """(n+1) x (n+1) heights over [-1, 1]^2: value-noise fBm, octave o on a
    (2^(o+2)+1)^2 lattice with amplitude * 0.5^o, bilinear interpolation."""
    rng = _rng(seed)
    u = np.linspace(0.0, 1.0, n + 1)
    h = np.zeros((n + 1, n + 1))
    for o in range(octaves):
        m = 2 ** (o + 2)
        g = rng.random((m + 1, m + 1)) * 2 - 1
        x = u * m
        x0 = np.minimum(x.astype(int), m - 1)
        fx = x - x0
        fx = fx * fx * (3 - 2 * fx)
        a = g[x0][:, x0] * (1 - fx)[None, :] + g[x0][:, x0 + 1] * fx[None, :]
        b = g[x0 + 1][:, x0] * (1 - fx)[None, :] + g[x0 + 1][:, x0 + 1] * fx[None, :]
        h += amplitude * 0.5 ** o * (a * (1 - fx)[:, None] + b * fx[:, None])
    return h


def heightfield_mesh(h, uv_repeat=4.0, half=1.0):
    n = h.shape[0] - 1
    x = np.linspace(-half, half, n + 1)
    X, Z = np.meshgrid(x, x, indexing="xy")  # row = z, col = x
    pos = np.stack([X, h, Z], -1).reshape(-1, 3)
    dx = 2.0 * half / n
    X, Z = X / half, Z / half
    gz, gx = np.gradient(h, dx, dx)
    nrm = np.stack([-gx, np.ones_like(h), -gz], -1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    uv = np.stack([(X + 1) / 2 * uv_repeat, (Z + 1) / 2 * uv_repeat], -1).reshape(-1, 2)
    r, c = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    v00 = r * (n + 1) + c
    v01, v10, v11 = v00 + 1, v00 + n + 1, v00 + n + 2
    tris = np.stack([np.stack([v00, v10, v11], -1), np.stack([v00, v11, v01], -1)], 2).reshape(-1, 3)
    return Mesh(pos.astype(f32), tris.astype(np.uint32), nrm.reshape(-1, 3).astype(f32), uv.astype(f32))


def terrain_spheres(grid=1024, n_spheres=200, sphere_res=(64, 32), tex_size=2048, env_size=(2048, 1024),
                    sun_peak=1e4, seed=4, scale=10.0, instanced=False):
    """fBm-displaced heightfield (2*grid^2 triangles), n_spheres UV spheres with
    materials 1/3 Lambertian, 1/3 metal (fuzz U(0,1)), 1/3 dielectric 1.5; the
    terrain is Lambertian with procedural RGBA8 albedo and normal textures; an
    RGB f32 sky with a sun lobe of peak `sun_peak`.

    instanced=True: the spheres are instances of one unit UV sphere (scale r,
    translation to the centre), the same random draws; such a scene renders
    with the extended estimator only (the reference has no instancing).

    The scene spans [-scale, scale]^2.  At unit scale the 1024-grid triangles
    (edges ~2e-3) would all be invisible: Triangle::hit uses t_min = 1e-5 as
    its determinant epsilon (mesh.rs:156), and |e1 . (d x e2)| <= |e1||e2|."""
    rng = _rng(seed + 1)
    h = fbm_heightfield(grid, seed=seed) * scale
    scene = Scene.empty()
    alb = scene.add_image_texture(procedural_texture(tex_size, seed + 2, "albedo"), SampleType.Bilinear)
    nrm = scene.add_image_texture(procedural_texture(tex_size, seed + 3, "normal"), SampleType.Bilinear)
    s = scene.add_mesh(heightfield_mesh(h, half=scale))
    scene.add_object(scene.add_material(Materials.lambertian_texture(alb, nrm)), s)
    n = h.shape[0] - 1
    unit = scene.add_instanced_mesh(uv_sphere((0.0, 0.0, 0.0), 1.0, *sphere_res)) if instanced else None
    for k in range(n_spheres):
        x, z = rng.uniform(-0.85, 0.85, 2)
        r = rng.uniform(0.02, 0.06)
        ix, iz = int((x + 1) / 2 * n), int((z + 1) / 2 * n)
        x, z, r = x * scale, z * scale, r * scale
        y = h[iz, ix] + 0.8 * r
        kind = k % 3
        if kind == 0:
            mat = Materials.lambertian(rng.uniform(0.2, 0.8, 3))
        elif kind == 1:
            mat = Materials.metal(rng.uniform(0.5, 0.95, 3), rng.uniform(0.0, 1.0))
        else:
            mat = Materials.dielectric(1.5)
        if instanced:
            s = scene.add_instance(unit, [[r, 0, 0, x], [0, r, 0, y], [0, 0, r, z]])
        else:
            s = scene.add_mesh(uv_sphere((x, y, z), r, *sphere_res))
        scene.add_object(scene.add_material(mat), s)
    scene.environment = Environments.hdri(procedural_sky(*env_size, seed=seed + 2, sun_peak=sun_peak))
    scene.camera = Camera.look_at((0.0, 0.9 * scale, 2.3 * scale), (0.0, 0.0, 0.1 * scale), (0.0, 1.0, 0.0),
                                  degrees_to_radians(50.0))
    return scene


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


def procedural_texture(size, seed, kind="albedo"):
    """RGBA8 texture: smooth value noise (albedo) or a perturbed +Y normal map."""
    rng = _rng(seed)
    g = rng.random((9, 9, 3))
    yy, xx = np.meshgrid(np.linspace(0, 8, size, endpoint=False), np.linspace(0, 8, size, endpoint=False), indexing="ij")
    x0, y0 = xx.astype(int), yy.astype(int)
    fx, fy = (xx - x0)[..., None], (yy - y0)[..., None]
    v = (g[y0, x0] * (1 - fx) * (1 - fy) + g[y0, x0 + 1] * fx * (1 - fy) + g[y0 + 1, x0] * (1 - fx) * fy +
         g[y0 + 1, x0 + 1] * fx * fy)
    if kind == "normal":
        n = np.stack([(v[..., 0] - 0.5) * 0.4, np.ones_like(v[..., 0]), (v[..., 1] - 0.5) * 0.4], -1)
        n /= np.linalg.norm(n, axis=-1, keepdims=True)
        v = n * 0.5 + 0.5
    rgba = np.concatenate([v, np.ones_like(v[..., :1])], -1)
    return (np.clip(rgba, 0, 1) * 255 + 0.5).astype(np.uint8)


def procedural_sky(width, height, seed=6, sun_peak=1e4, sun_dir=(0.3, 0.8, 0.5), sun_sharpness=2000.0):
    """RGB f32 equirect sky: log-normal-ish gradient plus a sun lobe, laid out
    for HDRIEnvironment's lookup (row y at polar angle pi*(H-1-y)/H from -Y)."""
    rng = _rng(seed)
    y = np.arange(height)[:, None] + 0.5
    x = np.arange(width)[None, :] + 0.5
    th = np.pi * (height - 1 - y + 0.5) / height
    ph = 2 * np.pi * x / width
    d = np.stack([np.sin(th) * np.cos(ph - np.pi) * np.ones_like(ph), -np.cos(th) * np.ones_like(ph),
                  -np.sin(th) * np.sin(ph - np.pi) * np.ones_like(ph)], -1)
    up = np.clip(d[..., 1], -1, 1)
    base = np.exp(rng.normal(0.0, 0.15, (height, width)))[..., None]
    sky = (0.3 + 0.7 * np.clip(up, 0, 1))[..., None] * np.array([0.35, 0.55, 1.0]) * base
    ground = (np.clip(-up, 0, 1) > 0)[..., None] * np.array([0.25, 0.22, 0.2]) * 0.5
    s = np.asarray(sun_dir, np.float64)
    s /= np.linalg.norm(s)
    sun = sun_peak * np.exp(sun_sharpness * (np.sum(d * s, -1) - 1.0))[..., None] * np.array([1.0, 0.9, 0.7])
    return np.ascontiguousarray((sky + ground + sun).astype(f32))


def material_spheres(seed=9, env="hdri", dof=True, tex_size=64, sky=(64, 32)):
    """Small scene exercising every material, textures, the HDRI and depth of field."""
    rng = _rng(seed)
    scene = Scene.empty()
    alb = scene.add_image_texture(procedural_texture(tex_size, seed, "albedo"), SampleType.Bilinear)
    nrm = scene.add_image_texture(procedural_texture(tex_size, seed + 1, "normal"), SampleType.Nearest)
    g = 2.0
    ground = Mesh(np.array([[-g, 0, -g], [-g, 0, g], [g, 0, g], [g, 0, -g]], f32), [[0, 1, 2], [2, 0, 3]],
                  np.tile([0.0, 1.0, 0.0], (4, 1)).astype(f32),
                  np.array([[0, 0], [0, 2.5], [2.5, 2.5], [2.5, 0]], f32))
    s = scene.add_mesh(ground)
    scene.add_object(scene.add_material(Materials.lambertian_texture(alb, nrm)), s)
    mats = [Materials.metal((0.8, 0.6, 0.3), 0.3), Materials.dielectric(1.5), Materials.lambertian((0.7, 0.2, 0.2)),
            Materials.lambertian_texture_no_normal(alb), Materials.metal((0.9, 0.9, 0.9), 0.0),
            Materials.lambertian_bsdf((0.5, 0.8, 0.5))]
    for i, m in enumerate(mats):
        c = (-1.25 + 0.5 * i, 0.25 + 0.05 * rng.random(), 0.2 * rng.standard_normal())
        s = scene.add_mesh(uv_sphere(c, 0.22, 24, 12))
        scene.add_object(scene.add_material(m), s)
    light = Mesh(np.array([[-0.3, 1.5, -0.3], [0.3, 1.5, -0.3], [0.3, 1.5, 0.3], [-0.3, 1.5, 0.3]], f32),
                 [[0, 1, 2], [2, 3, 0]])
    s = scene.add_mesh(light)
    scene.add_object(scene.add_material(Materials.emissive(4.0)), s)
    if env == "hdri":
        scene.environment = Environments.hdri(procedural_sky(*sky, sun_peak=50.0, sun_sharpness=40.0))
    elif env == "uniform":
        scene.environment = Environments.uniform((0.4, 0.5, 0.7))
    scene.camera = Camera.look_at((0.0, 1.0, 3.5), (0.0, 0.2, 0.0), (0.0, 1.0, 0.0), degrees_to_radians(45.0))
    if dof:
        scene.camera.dof = (0.03, (0.0, 0.3, 0.0))
    return scene


def cornell_zero_normal_quads(seed=1):
    """The cornell room built with Surfaces::quad (Vertex::position: zero normals)."""
    from .scene import Surfaces
    s = 1.0 / 555.0
    q = lambda *pts: Surfaces.quad(*[np.asarray(p, np.float64) * s for p in pts])
    meshes = [q((0, 0, 0), (0, 0, 555), (555, 0, 555), (555, 0, 0)),
              q((0, 0, 0), (0, 555, 0), (0, 555, 555), (0, 0, 555)),
              q((555, 0, 0), (555, 0, 555), (555, 555, 555), (555, 555, 0)),
              q((0, 0, 555), (0, 555, 555), (555, 555, 555), (555, 0, 555)),
              q((0, 555, 0), (555, 555, 0), (555, 555, 555), (0, 555, 555)),
              q((213, 554, 227), (213, 554, 332), (343, 554, 332), (343, 554, 227))]
    rng = _rng(seed)
    scene = Scene.empty()
    for k, m in enumerate(meshes):
        surf = scene.add_mesh(m)
        mat = Materials.emissive(15.0) if k == len(meshes) - 1 else Materials.lambertian(rng.uniform(0.2, 0.8, 3))
        scene.add_object(scene.add_material(mat), surf)
    scene.camera = Camera.look_at((278 * s, 278 * s, -800 * s), (278 * s, 278 * s, 0.0), (0.0, 1.0, 0.0),
                                  degrees_to_radians(40.0))
    return scene


# -- batched serving: 64 requests over a pool of scenes -------------------------
SERVING_RESOLUTIONS = ((256, 256), (512, 512), (1280, 720), (1920, 1080))


def lit_random_triangle_scene(num_tris, seed=2, sky=(512, 256)):
    """The traversal-microbench geometry with mixed materials, a camera in front
    of the unit cube and an HDR sky, so that it can be rendered."""
    rng = _rng(seed + 100)
    scene = Scene.empty()
    tris = random_triangles(num_tris, seed)
    third = num_tris // 3
    parts = [(0, third), (third, 2 * third), (2 * third, num_tris)]
    mats = [Materials.lambertian(rng.uniform(0.2, 0.8, 3)),
            Materials.metal(rng.uniform(0.5, 0.95, 3), float(rng.uniform(0.0, 1.0))), Materials.dielectric(1.5)]
    for (a, b), m in zip(parts, mats):
        if b <= a:
            continue
        pos = tris.positions[3 * a:3 * b]
        idx = np.arange(3 * (b - a), dtype=np.uint32).reshape(-1, 3)
        # vertex normals = face normals (a mesh without normals has zero normals, which shade to NaN)
        p = pos.reshape(-1, 3, 3).astype(np.float64)
        fn = _normalize_rows(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]))
        nrm = np.repeat(fn, 3, axis=0).astype(f32)
        scene.add_object(scene.add_material(m), scene.add_mesh(Mesh(pos, idx, nrm)))
    scene.environment = Environments.hdri(procedural_sky(*sky, seed=seed + 5, sun_peak=1e3, sun_sharpness=500.0))
    scene.camera = Camera.look_at((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), degrees_to_radians(45.0))
    return scene


def serving_scene_pool(seed=7):
    """Eight scenes of 10k to 2M triangles from the microbench and terrain
    generators (smaller textures and skies than the 1080p terrain scene)."""
    specs = [("tris", 10_000), ("terrain", (72, 20, (16, 8))), ("tris", 100_000), ("terrain", (224, 50, (32, 16))),
             ("tris", 400_000), ("terrain", (448, 100, (32, 16))), ("tris", 1_000_000),
             ("terrain", (990, 200, (16, 8)))]
    pool = []
    for k, (kind, arg) in enumerate(specs):
        if kind == "tris":
            pool.append(lit_random_triangle_scene(arg, seed=seed + k))
        else:
            grid, ns, res = arg
            pool.append(terrain_spheres(grid=grid, n_spheres=ns, sphere_res=res, tex_size=512, env_size=(512, 256),
                                        sun_peak=1e4, seed=seed + k))
    return pool


def serving_requests(n=64, pool_size=8, seed=7):
    """n requests: (scene index, width, height, spp, seed).  Resolution uniform
    over SERVING_RESOLUTIONS, spp uniform in 1..32, scene uniform over the
    pool, request i seeded with seed + i."""
    rng = _rng(seed)
    out = []
    for i in range(n):
        w, h = SERVING_RESOLUTIONS[int(rng.integers(len(SERVING_RESOLUTIONS)))]
        out.append((int(rng.integers(pool_size)), w, h, int(rng.integers(1, 33)), seed + i))
    return out


def _hex(c):
    """hex_color (util/color.rs): 0xRRGGBB -> RGB in [0, 1]."""
    return ((c >> 16) & 255) / 255.0, ((c >> 8) & 255) / 255.0, (c & 255) / 255.0


def analytic_spheres(tex_size=64, sky=(64, 32)):
    """The reference's spheres example (voidray_app/src/examples/spheres.rs:5-60):
    analytic spheres (metals, glass, a glass shell around a diffuse core, a red
    emissive sphere) over an analytic ground plane with albedo and normal
    textures, DOF camera, HDRI.  The image assets (wood .tif, indoor .exr) are
    replaced by procedural textures and sky of the same roles."""
    from .scene import Surfaces
    scene = Scene.empty()
    red = scene.add_material(Materials.metal(_hex(0xE78999), 0.05))
    yellow = scene.add_material(Materials.dielectric(1.5))
    green = scene.add_material(Materials.metal(_hex(0xB3E7AA), 0.1))
    blue = scene.add_material(Materials.metal(_hex(0x7CA3E7), 0.01))
    grey = scene.add_material(Materials.metal(_hex(0xAAAAAA), 0.01))
    diffuse = scene.add_material(Materials.lambertian(_hex(0x7CA3E7)))
    light_mtl = scene.add_material(Materials.colored_emissive(_hex(0xFF0F0F), 200.0))
    for pos, mtl in (((0.5, 1.0, 4.0), red), ((3.15, 1.5, -0.7), yellow), ((0.6, 0.6, -2.0), green),
                     ((-1.7, 1.1, -0.2), blue), ((1.2, 0.5, 0.4), grey)):
        scene.add_object(mtl, scene.add_analytic_surface(Surfaces.sphere(pos, pos[1])))
    glass_2 = scene.add_material(Materials.dielectric(1.5))
    sph = scene.add_analytic_surface(Surfaces.sphere((1.5, 0.7, -3.1), 0.7))
    sph_inner = scene.add_analytic_surface(Surfaces.sphere((1.5, 0.695, -3.1), 0.695))
    scene.add_object(glass_2, sph)
    scene.add_object(diffuse, sph_inner)
    scene.add_object(light_mtl, scene.add_analytic_surface(Surfaces.sphere((1.2, 8.0, -1.5), 2.0)))
    scene.camera = Camera.look_at((0.7166, 2.8803, -9.2992), (0.8673, 0.9557, 0.2095), (0.0, 1.0, 0.0), 0.6911)
    scene.camera.dof = (0.12, (0.1, 0.6, -2.0))
    alb = scene.add_image_texture(procedural_texture(tex_size, 21, "albedo"), SampleType.Bilinear)
    nrm = scene.add_image_texture(procedural_texture(tex_size, 22, "normal"), SampleType.Nearest)
    gnd = scene.add_analytic_surface(Surfaces.ground_plane(0.0))
    scene.add_object(scene.add_material(Materials.lambertian_texture(alb, nrm)), gnd)
    scene.environment = Environments.hdri(procedural_sky(*sky, sun_peak=50.0, sun_sharpness=40.0))
    return scene


def analytic_and_meshes(seed=13):
    """Analytic spheres and a ground plane interleaved with triangle meshes in
    the surface list (so the top-level tree mixes both), including a sphere
    coinciding with a UV-sphere mesh (equal-t ties across surface kinds)."""
    from .scene import Surfaces
    rng = _rng(seed)
    scene = Scene.empty()
    mat = scene.add_material(Materials.lambertian((0.6, 0.6, 0.6)))
    scene.add_object(mat, scene.add_analytic_surface(Surfaces.ground_plane(-0.25)))
    for i in range(6):
        c = tuple(float(x) for x in rng.uniform(-1, 1, 3))
        if i % 2:
            scene.add_object(mat, scene.add_mesh(uv_sphere(c, 0.3, 16, 8)))
        else:
            scene.add_object(mat, scene.add_analytic_surface(Surfaces.sphere(c, 0.3)))
    scene.add_object(mat, scene.add_mesh(random_triangles(2000, seed)))
    scene.add_object(mat, scene.add_analytic_surface(Surfaces.sphere((0.0, 2.0, 0.0), 0.5)))
    return scene


def microfacet_spheres(sky=(64, 32)):
    """Every MicrofacetBSDF constructor (voidray_common/src/microfacet.rs:29-102)
    on analytic spheres: clear and tinted glass at several roughnesses, rough
    specular and metallic spheres, a diffuse one and a `light` one (whose
    emittance the reference never reads), over a Lambertian ground plane,
    lit by an HDRI.  The reference passes the ray's own direction as wo, so
    the opaque constructors shade black from outside; the glass ones
    transmit and reflect."""
    from .scene import MicrofacetBSDF, Surfaces
    scene = Scene.empty()
    mats = [MicrofacetBSDF.clear(1.5, 0.05), MicrofacetBSDF.clear(1.33, 0.4),
            MicrofacetBSDF.transparent(_hex(0x7CA3E7), 1.5, 0.2), MicrofacetBSDF.transparent(_hex(0xE78999), 2.0, 0.8),
            MicrofacetBSDF.specular(_hex(0xB3E7AA), 0.3), MicrofacetBSDF.metallic(_hex(0xE7C27C), 0.15),
            MicrofacetBSDF.diffuse(_hex(0xAAAAAA)), MicrofacetBSDF.light((1.0, 0.9, 0.8), 5.0)]
    for i, m in enumerate(mats):
        x = -2.1 + 0.6 * i
        scene.add_object(scene.add_material(m), scene.add_analytic_surface(Surfaces.sphere((x, 0.3, 0.0), 0.28)))
    # a glass shell around a coloured glass core
    scene.add_object(scene.add_material(MicrofacetBSDF.clear(1.5, 0.1)),
                     scene.add_analytic_surface(Surfaces.sphere((0.0, 1.0, 0.6), 0.5)))
    scene.add_object(scene.add_material(MicrofacetBSDF.transparent(_hex(0xB3E7AA), 1.2, 0.3)),
                     scene.add_analytic_surface(Surfaces.sphere((0.0, 1.0, 0.6), 0.3)))
    scene.add_object(scene.add_material(Materials.lambertian((0.6, 0.55, 0.5))),
                     scene.add_analytic_surface(Surfaces.ground_plane(0.0)))
    scene.camera = Camera.look_at((0.0, 1.2, -3.5), (0.0, 0.5, 0.0), (0.0, 1.0, 0.0), 0.9)
    scene.environment = Environments.hdri(procedural_sky(*sky, sun_peak=50.0, sun_sharpness=40.0))
    return scene


def cornell_demo():
    """The reference's cornell example (voidray_app/src/examples/cornell.rs:5-78)
    as written: Surfaces::quad walls at the 555 scale (vertex normals zero), its
    albedos, an emissive quad of 15, and an analytic glass sphere (IOR 1.33)
    around a blue Lambertian core 0.1 smaller."""
    from .scene import Surfaces
    scene = Scene.empty()
    red = scene.add_material(Materials.lambertian((0.65, 0.05, 0.05)))
    white = scene.add_material(Materials.lambertian((0.73, 0.73, 0.73)))
    green = scene.add_material(Materials.lambertian((0.12, 0.45, 0.15)))
    light = scene.add_material(Materials.emissive(15.0))
    q = lambda *pts: scene.add_mesh(Surfaces.quad(*[np.asarray(p, np.float64) for p in pts]))
    floor = q((0, 0, 0), (0, 0, 555), (555, 0, 555), (555, 0, 0))
    red_wall = q((0, 0, 0), (0, 0, 555), (0, 555, 555), (0, 555, 0))
    green_wall = q((555, 0, 0), (555, 0, 555), (555, 555, 555), (555, 555, 0))
    back_wall = q((0, 0, 555), (555, 0, 555), (555, 555, 555), (0, 555, 555))
    ceil = q((0, 555, 0), (0, 555, 555), (555, 555, 555), (555, 555, 0))
    light_plane = q((213, 554, 227), (213, 554, 332), (343, 554, 332), (343, 554, 227))
    for mat, surf in ((white, floor), (green, green_wall), (red, red_wall), (white, back_wall), (white, ceil),
                      (light, light_plane)):
        scene.add_object(mat, surf)
    sph = scene.add_analytic_surface(Surfaces.sphere((555.0 / 2.0, 100.0, 555.0 / 2.0), 100.0))
    glass = scene.add_material(Materials.dielectric(1.33))
    sph_inner = scene.add_analytic_surface(Surfaces.sphere((555.0 / 2.0, 100.0, 555.0 / 2.0), 99.9))
    glass_inner = scene.add_material(Materials.lambertian(_hex(0x0F1BF0)))
    scene.add_object(glass, sph)
    scene.add_object(glass_inner, sph_inner)
    scene.camera = Camera.look_at((278.0, 278.0, -800.0), (278.0, 278.0, 0.0), (0.0, 1.0, 0.0),
                                  degrees_to_radians(40.0))
    return scene


def uv_test_texture(size=64):
    """An 8x8 checker of distinct colours, standing in for assets/uv_test.png."""
    y, x = np.mgrid[0:size, 0:size]
    cx, cy = (x * 8) // size, (y * 8) // size
    img = np.zeros((size, size, 3), np.float32)
    img[..., 0] = (cx + 1) / 8.0
    img[..., 1] = (cy + 1) / 8.0
    img[..., 2] = np.where((cx + cy) % 2 == 0, 0.9, 0.2)
    return img


def material_demo(tex_size=64, sky=(64, 32)):
    """The reference's material example (voidray_app/src/examples/material.rs:3-30):
    a LambertianBSDF sphere on an analytic ground plane with a nearest-sampled
    texture and no normal map, the default camera moved to (0.5, 0.6, 5), an
    HDRI; the image assets are replaced by a procedural checker and sky."""
    from .scene import Surfaces
    scene = Scene.empty()
    ground = scene.add_analytic_surface(Surfaces.ground_plane(0.0))
    uv = scene.add_image_texture(uv_test_texture(tex_size), SampleType.Nearest)
    scene.add_object(scene.add_material(Materials.lambertian_texture_no_normal(uv)), ground)
    sphere = scene.add_analytic_surface(Surfaces.sphere((0.0, 0.5, 0.0), 0.5))
    scene.add_object(scene.add_material(Materials.lambertian_bsdf(_hex(0xFF3030))), sphere)
    scene.camera.eye = np.asarray((0.5, 0.6, 5.0), np.float32)
    scene.environment = Environments.hdri(procedural_sky(*sky, sun_peak=50.0, sun_sharpness=40.0))
    return scene
