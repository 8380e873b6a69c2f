"""Host-side mirror of voidray_renderer's scene API over the C ABI.

Names, argument meaning and behaviour follow the reference:

* ``Scene.empty / add_material / add_mesh / add_object / add_image_texture``
  — voidray_renderer/src/core/scene.rs:94-161
* ``Materials.*`` — voidray_common/src/simple.rs:17-58
* ``Surfaces.quad`` — voidray_common/src/surfaces.rs:18-28
* ``Mesh.from_buffers`` — voidray_renderer/src/core/mesh.rs:76-116
* ``Environments.uniform / hdri`` — voidray_common/src/environments.rs:9-17
* ``Camera.look_at`` — voidray_renderer/src/core/camera.rs:26-36
* ``RenderSettings / ColorManagementSettings / Tonemap`` — core/settings.rs
* ``Scene.build_acceleration`` — core/scene.rs:163-179 (uploads to the GPU)
* ``iterative_render`` — render/iterative.rs:11-55

Image and mesh data are given as arrays (the reference reads files through
the image and obj crates; their decoders are not part of this path).
"""
import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _native as N

f32 = np.float32


def _f3(v):
    return np.asarray(v, dtype=f32).reshape(3)


def _dot(a, b):
    # cgmath dot: (x*x' + y*y') + z*z', all in f32
    return f32(f32(f32(a[0] * b[0]) + f32(a[1] * b[1])) + f32(a[2] * b[2]))


def _normalize(v):
    m = f32(np.sqrt(_dot(v, v)))
    return (v * f32(f32(1.0) / m)).astype(f32)


# ---- materials ---------------------------------------------------------------

@dataclass
class Material:
    kind: int
    color: tuple = (0.0, 0.0, 0.0)
    param: float = 0.0
    albedo_texture: int = -1
    normal_texture: int = -1
    index: float = 1.5          # MicrofacetBSDF parameters (microfacet.rs:9-27)
    roughness: float = 1.0
    metallic: float = 0.0
    emittance: float = 0.0
    transparent: bool = False


class MicrofacetBSDF:
    """Constructors of voidray_common/src/microfacet.rs:29-102 (f32 parameters)."""

    @staticmethod
    def _make(color, index, roughness, metallic, emittance, transparent):
        return Material(N.VR_MAT_MICROFACET, tuple(_f3(color)), 0.0, -1, -1, float(f32(index)),
                        float(f32(roughness)), float(f32(metallic)), float(f32(emittance)), bool(transparent))

    @staticmethod
    def diffuse(color):
        return MicrofacetBSDF._make(color, 1.5, 1.0, 0.0, 0.0, False)

    @staticmethod
    def specular(color, roughness):
        return MicrofacetBSDF._make(color, 1.5, roughness, 0.0, 0.0, False)

    @staticmethod
    def clear(index, roughness):
        return MicrofacetBSDF._make((1.0, 1.0, 1.0), index, roughness, 0.0, 0.0, True)

    @staticmethod
    def transparent(color, index, roughness):
        return MicrofacetBSDF._make(color, index, roughness, 0.0, 0.0, True)

    @staticmethod
    def metallic(color, roughness):
        return MicrofacetBSDF._make(color, 1.5, roughness, 1.0, 0.0, False)

    @staticmethod
    def light(color, emittance):
        return MicrofacetBSDF._make(color, 1.0, 1.0, 0.0, emittance, False)


class Materials:
    """voidray_common/src/simple.rs:17-58."""

    @staticmethod
    def lambertian(albedo):
        return Material(N.VR_MAT_LAMBERTIAN, tuple(_f3(albedo)))

    @staticmethod
    def lambertian_bsdf(albedo):
        return Material(N.VR_MAT_LAMBERTIAN_BSDF, tuple(_f3(albedo)))

    @staticmethod
    def lambertian_texture_no_normal(albedo):
        return Material(N.VR_MAT_LAMBERTIAN, (0.0, 0.0, 0.0), 0.0, int(albedo), -1)

    @staticmethod
    def lambertian_texture(albedo, normal):
        return Material(N.VR_MAT_LAMBERTIAN, (0.0, 0.0, 0.0), 0.0, int(albedo), int(normal))

    @staticmethod
    def metal(albedo, fuzz):
        return Material(N.VR_MAT_METAL, tuple(_f3(albedo)), float(f32(fuzz)))

    @staticmethod
    def dielectric(ir):
        return Material(N.VR_MAT_DIELECTRIC, (1.0, 1.0, 1.0), float(f32(ir)))

    @staticmethod
    def emissive(strength):
        # Emission::new(Color(1,1,1), strength) stores color * strength (simple.rs:167-173)
        return Materials.colored_emissive((1.0, 1.0, 1.0), strength)

    @staticmethod
    def colored_emissive(color, strength):
        c = (_f3(color) * f32(strength)).astype(f32)
        return Material(N.VR_MAT_EMISSION, tuple(c))


# ---- geometry ----------------------------------------------------------------

def instance_points(xf, p):
    """world = M p + c for a 3x4 row-major transform, in f32 with the
    library's operation order: ((m0 x + m1 y) + m2 z) + c."""
    xf = np.asarray(xf, f32).reshape(3, 4)
    p = np.asarray(p, f32)
    out = np.empty_like(p)
    for r in range(3):
        out[:, r] = ((xf[r, 0] * p[:, 0] + xf[r, 1] * p[:, 1]) + xf[r, 2] * p[:, 2]) + xf[r, 3]
    return out


def instance_normal_matrix(xf):
    """(M^-1)^T of the 3x3 part, computed in f64 and rounded to f32 once."""
    m = np.asarray(xf, np.float64).reshape(3, 4)[:, :3]
    return np.linalg.inv(m).T.astype(f32)


def instance_normals(xf, n):
    """n' = (M^-1)^T n in f32 with the library's operation order (not renormalised)."""
    nm = instance_normal_matrix(xf)
    n = np.asarray(n, f32)
    out = np.empty_like(n)
    for r in range(3):
        out[:, r] = (nm[r, 0] * n[:, 0] + nm[r, 1] * n[:, 1]) + nm[r, 2] * n[:, 2]
    return out


class Mesh:
    """Triangle mesh (core/mesh.rs).  normals/uvs None = Vertex::position."""

    def __init__(self, positions, indices, normals=None, uvs=None):
        self.positions = np.ascontiguousarray(positions, dtype=f32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        self.normals = None if normals is None else np.ascontiguousarray(normals, dtype=f32).reshape(-1, 3)
        self.uvs = None if uvs is None else np.ascontiguousarray(uvs, dtype=f32).reshape(-1, 2)
        n = len(self.positions)
        if self.normals is not None and len(self.normals) != n:
            raise ValueError("normals must match positions")
        if self.uvs is not None and len(self.uvs) != n:
            raise ValueError("uvs must match positions")
        if self.indices.size and int(self.indices.max()) >= n:
            raise ValueError("index out of range")

    @staticmethod
    def from_buffers(positions, indices, normals=None, uvs=None):
        return Mesh(positions, indices, normals, uvs)

    @staticmethod
    def from_file(path):
        """Mesh::from_file (mesh.rs:46-74): a Wavefront OBJ read by the library's
        loader (obj-rs 0.7 semantics for TexturedVertex meshes)."""
        m = N.vr_mesh()
        N.check(N.lib().vr_mesh_load_obj(os.fsencode(path), C.byref(m)))
        try:
            nv, nt = m.num_vertices, m.num_triangles
            arr = lambda ptr, n, ct: np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), (n,)).copy() if n else \
                np.zeros(0, np.float32 if ct is C.c_float else np.uint32)
            return Mesh(arr(m.positions, 3 * nv, C.c_float), arr(m.indices, 3 * nt, C.c_uint32),
                        arr(m.normals, 3 * nv, C.c_float), arr(m.uvs, 2 * nv, C.c_float))
        finally:
            N.lib().vr_mesh_release(C.byref(m))

    @property
    def num_triangles(self):
        return len(self.indices)


@dataclass
class AnalyticSurface:
    """Sphere (surfaces.rs:31-79) or GroundPlane (surfaces.rs:81-114), f32 parameters."""
    kind: int
    center: tuple = (0.0, 0.0, 0.0)
    radius: float = 0.0
    height: float = 0.0


class Surfaces:
    @staticmethod
    def sphere(center, radius):
        """surfaces.rs:10-12."""
        return AnalyticSurface(N.VR_ANALYTIC_SPHERE, tuple(float(x) for x in _f3(center)), float(f32(radius)))

    @staticmethod
    def ground_plane(height):
        """surfaces.rs:14-16."""
        return AnalyticSurface(N.VR_ANALYTIC_GROUND_PLANE, height=float(f32(height)))

    @staticmethod
    def quad(q1, q2, q3, q4):
        """surfaces.rs:18-28: Vertex::position vertices, indices [0,1,2, 2,0,3]."""
        return Mesh(np.stack([_f3(q1), _f3(q2), _f3(q3), _f3(q4)]), [0, 1, 2, 2, 0, 3])


# ---- textures / environments --------------------------------------------------

class SampleType:
    Nearest = N.VR_SAMPLE_NEAREST
    Bilinear = N.VR_SAMPLE_BILINEAR


@dataclass
class Texture:
    data: np.ndarray  # (H, W, 4) uint8 or (H, W, 3) float32
    sample_type: int = SampleType.Bilinear

    def __post_init__(self):
        if self.data.dtype == np.uint8:
            assert self.data.ndim == 3 and self.data.shape[2] == 4
            self.format = N.VR_TEX_RGBA8
        else:
            self.data = np.ascontiguousarray(self.data, dtype=f32)
            assert self.data.ndim == 3 and self.data.shape[2] == 3
            self.format = N.VR_TEX_RGB32F
        self.data = np.ascontiguousarray(self.data)


@dataclass
class Environment:
    kind: int
    color: tuple = (0.0, 0.0, 0.0)
    image: Optional[np.ndarray] = None  # (H, W, 3) float32


class Environments:
    @staticmethod
    def uniform(color):
        return Environment(N.VR_ENV_UNIFORM, tuple(_f3(color)))

    @staticmethod
    def hdri(image):
        """environments.rs:14-16: an (H, W, 3) array, or an image file read as
        image::open(path).to_rgb32f() (OpenEXR, PNG, TIFF, ...)."""
        if isinstance(image, (str, os.PathLike)):
            from .imageio import load_image
            image = load_image(image, want_float=True)
        return Environment(N.VR_ENV_HDRI, (0.0, 0.0, 0.0), np.ascontiguousarray(image, dtype=f32))


# ---- camera --------------------------------------------------------------------

@dataclass
class Camera:
    eye: np.ndarray
    direction: np.ndarray
    up: np.ndarray
    fov: float
    dof: Optional[tuple] = None  # (aperture, focal_point)

    @staticmethod
    def look_at(eye, center, up, fov):
        """camera.rs:26-36, evaluated in f32 with the reference operation order."""
        eye, center, up = _f3(eye), _f3(center), _f3(up)
        direction = _normalize((center - eye).astype(f32))
        up = _normalize((up - (_dot(up, direction) * direction).astype(f32)).astype(f32))
        return Camera(eye, direction, up, float(f32(fov)))


def degrees_to_radians(deg):
    """math.rs:30-33 in f32: degrees * PI / 180."""
    return float(f32(f32(f32(deg) * f32(np.pi)) / f32(180.0)))


# ---- settings --------------------------------------------------------------------

class RenderMode:
    Full = N.VR_MODE_FULL
    Normal = N.VR_MODE_NORMAL


@dataclass
class RenderSettings:
    """core/settings.rs:24-34 defaults."""
    total_samples: int = 100
    update_frequency: float = 0.1
    render_mode: int = RenderMode.Full
    firefly_clamp: float = 3.0
    max_bounces: int = 10


class Tonemap:
    None_ = N.VR_TONEMAP_NONE
    ACES = N.VR_TONEMAP_ACES
    Reinhard = N.VR_TONEMAP_REINHARD
    Filmic = N.VR_TONEMAP_FILMIC
    Uncharted2 = N.VR_TONEMAP_UNCHARTED2


class LegacyTonemap:
    """The legacy viewer's Tonemap (voidray/src/pipeline.rs:59-97), for vr_post_desc.tonemap."""
    None_ = N.VR_TONEMAP_LEGACY_NONE
    SimpleACES = N.VR_TONEMAP_LEGACY_SIMPLE_ACES
    SimpleReinhard = N.VR_TONEMAP_LEGACY_SIMPLE_REINHARD
    LumaReinhard = N.VR_TONEMAP_LEGACY_LUMA_REINHARD
    LumaWhitePreservingReinhard = N.VR_TONEMAP_LEGACY_LUMA_WHITE_PRESERVING_REINHARD
    RomBinDaHouse = N.VR_TONEMAP_LEGACY_ROM_BIN_DA_HOUSE
    Filmic = N.VR_TONEMAP_LEGACY_FILMIC
    Uncharted2 = N.VR_TONEMAP_LEGACY_UNCHARTED2


@dataclass
class ColorManagementSettings:
    """core/settings.rs:65-74 defaults."""
    tonemap: int = Tonemap.None_
    gamma: float = 2.2
    exposure: float = 0.0
    transparent: bool = True


@dataclass
class Settings:
    render: RenderSettings = field(default_factory=RenderSettings)
    color_management: ColorManagementSettings = field(default_factory=ColorManagementSettings)


# ---- scene -------------------------------------------------------------------------

class Scene:
    """core/scene.rs:36-161.  Handles are plain indices."""

    def __init__(self):
        self.camera = Camera.look_at((1.0, 0.0, 10.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0),
                                     float(f32(f32(np.pi) / f32(6.0))))
        self.materials: List[Material] = []
        self.objects: List[tuple] = []
        self.surfaces: List[int] = []  # surface -> mesh index, or VR_SURFACE_ANALYTIC | analytic index
        self.meshes: List[Mesh] = []
        self.analytic: List[AnalyticSurface] = []
        self.instances: List[tuple] = []  # (mesh index, 3x4 row-major float32 transform)
        self.textures: List[Texture] = []
        self.environment: Optional[Environment] = None

    @staticmethod
    def empty():
        return Scene()

    def add_material(self, material):
        self.materials.append(material)
        return len(self.materials) - 1

    def add_mesh(self, mesh):
        self.meshes.append(mesh)
        self.surfaces.append(len(self.meshes) - 1)
        return len(self.surfaces) - 1

    def add_analytic_surface(self, surface):
        """scene.rs:118-124."""
        self.analytic.append(surface)
        self.surfaces.append(N.VR_SURFACE_ANALYTIC | (len(self.analytic) - 1))
        return len(self.surfaces) - 1

    def add_instanced_mesh(self, mesh):
        """A mesh that only instances place in the scene (not a surface itself).
        Returns its mesh index for add_instance.  No counterpart in the reference."""
        self.meshes.append(mesh)
        return len(self.meshes) - 1

    def add_instance(self, mesh_index, transform):
        """Surface = mesh `mesh_index` under the affine map world = M p + c,
        `transform` a 3x4 (row-major [M | c]).  All instances of a mesh share
        one acceleration structure.  No counterpart in the reference: scenes
        with instances render with the extended estimator only."""
        xf = np.ascontiguousarray(transform, dtype=f32).reshape(3, 4)
        self.instances.append((int(mesh_index), xf))
        self.surfaces.append(N.VR_SURFACE_INSTANCE | (len(self.instances) - 1))
        return len(self.surfaces) - 1

    def flattened(self):
        """The same scene with every instance replaced by its own world-space
        mesh (what the reference, which has no instancing, would be given):
        positions p' = M p + c, normals n' = (M^-1)^T n, in f32 with the
        library's operation order; meshes keep their order, the instanced ones
        are appended."""
        import copy
        out = copy.copy(self)
        out.meshes = list(self.meshes)
        out.surfaces = list(self.surfaces)
        out.instances = []
        for i, sm in enumerate(self.surfaces):
            if not (sm & N.VR_SURFACE_INSTANCE) or (sm & N.VR_SURFACE_ANALYTIC):
                continue
            mi, xf = self.instances[sm & ~N.VR_SURFACE_INSTANCE]
            m = self.meshes[mi]
            out.meshes.append(Mesh(instance_points(xf, m.positions), m.indices,
                                   None if m.normals is None else instance_normals(xf, m.normals), m.uvs))
            out.surfaces[i] = len(out.meshes) - 1
        return out

    def add_object(self, material, surface):
        self.objects.append((int(surface), int(material)))
        return len(self.objects) - 1

    def add_image_texture(self, data, sample_type=SampleType.Bilinear):
        """scene.rs:154-160: an array ((H, W, 4) uint8 or (H, W, 3) float32) or an
        image file path, read as image::open(path).to_rgb32f()."""
        if isinstance(data, (str, os.PathLike)):
            from .imageio import load_image
            data = load_image(data)
        self.textures.append(Texture(np.asarray(data), sample_type))
        return len(self.textures) - 1

    def add_mesh_from_file(self, path):
        """scene.rs:137-144."""
        return self.add_mesh(Mesh.from_file(path))

    @property
    def num_triangles(self):
        n = 0
        for m in self.surfaces:
            if m & N.VR_SURFACE_ANALYTIC:
                continue
            if m & N.VR_SURFACE_INSTANCE:
                m = self.instances[m & ~N.VR_SURFACE_INSTANCE][0]
            n += self.meshes[m].num_triangles
        return n

    def to_desc(self, build_flags=0):
        """Flat C description; returns (desc, keepalive)."""
        keep = []

        def ptr(a):
            if a is None:
                return None
            keep.append(a)
            return a.ctypes.data

        meshes = (N.vr_mesh * max(1, len(self.meshes)))()
        for i, m in enumerate(self.meshes):
            meshes[i] = N.vr_mesh(len(m.positions), len(m.indices), ptr(m.positions), ptr(m.normals),
                                  ptr(m.uvs), ptr(m.indices))
        surf = np.asarray(self.surfaces, dtype=np.uint32)
        objs = (N.vr_object * max(1, len(self.objects)))()
        for i, (s, m) in enumerate(self.objects):
            objs[i] = N.vr_object(s, m)
        mats = (N.vr_material * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            mats[i] = N.vr_material(m.kind, N.F3(*m.color), m.param, m.albedo_texture, m.normal_texture,
                                    m.index, m.roughness, m.metallic, m.emittance, int(m.transparent))
        texs = (N.vr_texture * max(1, len(self.textures)))()
        for i, t in enumerate(self.textures):
            texs[i] = N.vr_texture(t.data.shape[1], t.data.shape[0], t.format, t.sample_type, ptr(t.data))
        env = N.vr_environment()
        e = self.environment
        if e is None:
            env.kind = N.VR_ENV_NONE
        else:
            env.kind = e.kind
            env.color = N.F3(*e.color)
            if e.image is not None:
                env.width, env.height = e.image.shape[1], e.image.shape[0]
                env.data = ptr(e.image)
        cam = N.vr_camera()
        c = self.camera
        cam.eye = N.F3(*c.eye)
        cam.direction = N.F3(*c.direction)
        cam.up = N.F3(*c.up)
        cam.fov = c.fov
        if c.dof is not None:
            cam.has_dof = 1
            cam.aperture = c.dof[0]
            cam.focal_point = N.F3(*_f3(c.dof[1]))
        ana = (N.vr_analytic * max(1, len(self.analytic)))()
        for i, a in enumerate(self.analytic):
            ana[i] = N.vr_analytic(a.kind, N.F3(*a.center), a.radius, a.height)
        inst = (N.vr_instance * max(1, len(self.instances)))()
        for i, (mi, xf) in enumerate(self.instances):
            inst[i] = N.vr_instance(mi, (C.c_float * 12)(*np.asarray(xf, f32).reshape(12).tolist()))
        desc = N.vr_scene_desc(len(self.meshes), meshes, len(self.surfaces), ptr(surf), len(self.objects), objs,
                               len(self.materials), mats, len(self.textures), texs, env, cam, build_flags,
                               len(self.analytic), ana, len(self.instances), inst)
        keep += [meshes, objs, mats, texs, ana, inst]
        return desc, keep

    def build_acceleration(self, ctx=None, build_flags=0):
        """Scene::build_acceleration: flatten, build the BVHs and upload to HBM."""
        from .device import Context, SceneAcceleration
        ctx = ctx or Context.default()
        return SceneAcceleration(ctx, self, build_flags)
