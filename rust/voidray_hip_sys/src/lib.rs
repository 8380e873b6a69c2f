//! Raw bindings to `libvoidray_hip.so`: a mechanical mirror of
//! `include/voidray_hip.h` (every struct with its field order, every
//! constant, every entry point).  voidray's own crates call it through two
//! additive hooks: `SceneAcceleration::export_flat` builds a `vr_scene_desc`
//! and `iterative_render_gpu` replaces `iterative_render`
//! (voidray_renderer/src/render/iterative.rs:11-55) at
//! voidray_renderer/src/render/renderer.rs:67 and :87.
//!
//! Every function returns an `i32` status (`VR_OK` or a negative `VR_ERR_*`);
//! `vr_last_error` gives the failing call's message.  A safe wrapper maps a
//! nonzero status to `panic!`, as the reference panics on its failures.
#![allow(non_camel_case_types)]
#![no_std]

use core::ffi::{c_char, c_int, c_void};

// ---- opaque handles ----------------------------------------------------------
#[repr(C)]
pub struct vr_ctx {
    _private: [u8; 0],
}
#[repr(C)]
pub struct vr_scene {
    _private: [u8; 0],
}
#[repr(C)]
pub struct vr_cache {
    _private: [u8; 0],
}

// ---- status codes ------------------------------------------------------------
pub const VR_OK: i32 = 0;
pub const VR_ERR_INVALID: i32 = -1;
pub const VR_ERR_HIP: i32 = -2;
pub const VR_ERR_OOM: i32 = -3;
pub const VR_ERR_UNSUPPORTED: i32 = -4;

// ---- scene description -------------------------------------------------------
/// Mesh::from_buffers / Mesh::from_file (core/mesh.rs:46-116); normals and uvs
/// may be null (Vertex::position: zero normals and uvs).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vr_mesh {
    pub num_vertices: u32,
    pub num_triangles: u32,
    pub positions: *const f32,
    pub normals: *const f32,
    pub uvs: *const f32,
    pub indices: *const u32,
}

pub const VR_MAT_LAMBERTIAN: i32 = 0;
pub const VR_MAT_METAL: i32 = 1;
pub const VR_MAT_DIELECTRIC: i32 = 2;
pub const VR_MAT_EMISSION: i32 = 3;
pub const VR_MAT_LAMBERTIAN_BSDF: i32 = 4;
pub const VR_MAT_MICROFACET: i32 = 5;

/// Materials of voidray_common/src/simple.rs:17-58 and MicrofacetBSDF
/// (voidray_common/src/microfacet.rs:9-27).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vr_material {
    pub kind: i32,
    pub color: [f32; 3],
    pub param: f32,
    pub albedo_texture: i32,
    pub normal_texture: i32,
    pub index: f32,
    pub roughness: f32,
    pub metallic: f32,
    pub emittance: f32,
    pub transparent: i32,
}

pub const VR_TEX_RGB32F: i32 = 0;
pub const VR_TEX_RGBA8: i32 = 1;
pub const VR_SAMPLE_NEAREST: i32 = 0;
pub const VR_SAMPLE_BILINEAR: i32 = 1;

/// ImageTexture (core/texture.rs:28-99).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vr_texture {
    pub width: u32,
    pub height: u32,
    pub format: i32,
    pub sample_type: i32,
    pub data: *const c_void,
}

pub const VR_ENV_NONE: i32 = 0;
pub const VR_ENV_UNIFORM: i32 = 1;
pub const VR_ENV_HDRI: i32 = 2;

/// UniformEnvironment / HDRIEnvironment (voidray_common/src/environments.rs:19-91).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vr_environment {
    pub kind: i32,
    pub color: [f32; 3],
    pub width: u32,
    pub height: u32,
    pub data: *const f32,
}

/// Camera (core/camera.rs:5-54).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vr_camera {
    pub eye: [f32; 3],
    pub direction: [f32; 3],
    pub up: [f32; 3],
    pub fov: f32,
    pub has_dof: i32,
    pub aperture: f32,
    pub focal_point: [f32; 3],
}

/// Object = (surface, material) (core/scene.rs:30-34).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vr_object {
    pub surface: u32,
    pub material: u32,
}

pub const VR_ANALYTIC_SPHERE: i32 = 0;
pub const VR_ANALYTIC_GROUND_PLANE: i32 = 1;

/// Sphere / GroundPlane (voidray_common/src/surfaces.rs:31-114).
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vr_analytic {
    pub kind: i32,
    pub center: [f32; 3],
    pub radius: f32,
    pub height: f32,
}

pub const VR_SURFACE_ANALYTIC: u32 = 0x8000_0000;
pub const VR_SURFACE_INSTANCE: u32 = 0x4000_0000;

/// An instance of a mesh under an affine map (row-major 3x4, object to world);
/// no counterpart in the reference.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vr_instance {
    pub mesh: u32,
    pub transform: [f32; 12],
}

/// Whole scene (core/scene.rs:36-44); the material of surface i is objects[i].material.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vr_scene_desc {
    pub num_meshes: u32,
    pub meshes: *const vr_mesh,
    pub num_surfaces: u32,
    pub surface_mesh: *const u32,
    pub num_objects: u32,
    pub objects: *const vr_object,
    pub num_materials: u32,
    pub materials: *const vr_material,
    pub num_textures: u32,
    pub textures: *const vr_texture,
    pub environment: vr_environment,
    pub camera: vr_camera,
    pub build_flags: u32,
    pub num_analytic: u32,
    pub analytic: *const vr_analytic,
    pub num_instances: u32,
    pub instances: *const vr_instance,
}

pub const VR_BUILD_DEFAULT: u32 = 0;
pub const VR_BUILD_NO_ENV_CDF: u32 = 1;
pub const VR_BUILD_LEAF_RECORDS: u32 = 2;
pub const VR_BUILD_LEAF_REFS: u32 = 4;

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct vr_scene_stats {
    pub num_triangles: u32,
    pub num_surfaces: u32,
    pub num_qnodes: u32,
    pub num_ref_nodes: u32,
    pub qbvh_depth: u32,
    pub max_leaf_size: u32,
    pub build_ref_seconds: f64,
    pub build_sah_seconds: f64,
    pub upload_seconds: f64,
    pub device_bytes: u64,
    pub num_tri_records: u32,
    pub leaf_refs: u32,
    /// Build self-check: wide-node child bounds decoded as the kernels decode them ...
    pub qbox_checked: u64,
    /// ... of which did not contain the exact box they stand for (0 for a correct build).
    pub qbox_not_containing: u64,
}

// ---- ray queries (SceneAcceleration::hit, core/scene.rs:182-185) ------------
pub const VR_NOALLOC_MAX_RAYS: u32 = 1_048_576;
pub const VR_NOALLOC_SCRATCH_BYTES: u32 = 75_501_568;
pub const VR_TRACE_CLOSEST: i32 = 0;
pub const VR_TRACE_ANY: i32 = 1;
pub const VR_TRACE_TIMING: i32 = 0x100;
pub const VR_ENGINE_WIDE: i32 = 0;
pub const VR_ENGINE_REFERENCE: i32 = 1;
pub const VR_ENGINE_WATERTIGHT: i32 = 2;

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vr_trace_desc {
    pub num_rays: u32,
    pub t_min: f32,
    pub mode: i32,
    pub engine: i32,
}

// ---- rendering (render/iterative.rs:11-55, core/tracer.rs:9-56) --------------
pub const VR_MODE_FULL: i32 = 0;
pub const VR_MODE_NORMAL: i32 = 1;
pub const VR_RENDER_ROWS_BY_HEIGHT: u32 = 1;
pub const VR_RENDER_ENV_IS: u32 = 2;
pub const VR_RENDER_RUSSIAN_ROULETTE: u32 = 4;
pub const VR_RENDER_ALPHA_PER_CALL: u32 = 8;
pub const VR_RENDER_OVERWRITE: u32 = 16;

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vr_render_desc {
    pub width: u32,
    pub height: u32,
    pub samples: u32,
    pub total_samples: u32,
    pub sample_offset: u32,
    pub max_bounces: u32,
    pub firefly_clamp: f32,
    pub render_mode: i32,
    pub flags: u32,
    pub rr_depth: u32,
    pub seed: u64,
    pub request: u32,
    pub engine: u32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct vr_render_stats {
    pub segments: u64,
    pub samples: u64,
    pub env_misses: u64,
    pub fallback_rays: u64,
    pub capped_loops: u64,
    pub kernel_ms: f64,
    pub textured_hits: u64,
    pub env_is_samples: u64,
}

// ---- batched serving (no counterpart; replaces one RenderThread::one_shot per request) ----
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vr_request {
    pub scene: *mut vr_scene,
    pub desc: vr_render_desc,
    pub d_rgba: *mut f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct vr_tuning {
    pub chunk_paths: u32,
    pub block: u32,
}

pub const VR_BATCH_GRAPH: u32 = 1;
pub const VR_BATCH_TIMESTAMPS: u32 = 2;
pub const VR_BATCH_SHORTEST_FIRST: u32 = 4;

// ---- learned radiance cache (no counterpart) -----------------------------------
pub const VR_CACHE_BF16: u32 = 0;
pub const VR_CACHE_FP32: u32 = 1;
pub const VR_CACHE_LEVELS: usize = 16;
pub const VR_CACHE_MLP_PARAMS: usize = 21504;

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vr_cache_desc {
    pub log2_table: u32,
    pub base_resolution: u32,
    pub max_resolution: u32,
    pub precision: u32,
    pub bbox_lo: [f32; 3],
    pub bbox_hi: [f32; 3],
    pub lr_grid: f32,
    pub lr_mlp: f32,
    pub beta1: f32,
    pub beta2: f32,
    pub eps: f32,
    pub seed: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct vr_cache_info {
    pub levels: u32,
    pub table_size: u32,
    pub precision: u32,
    pub pad: u32,
    pub resolution: [u32; VR_CACHE_LEVELS],
    pub origin: [f32; 3],
    pub inv_extent: f32,
    pub grid_params: u64,
    pub steps: u64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct vr_cache_frame {
    pub train_paths: u32,
    pub train_vertices: u32,
    pub cache_depth: u32,
    pub train_steps: u32,
    pub frame: u32,
    pub query_roughness: f32,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct vr_cache_stats {
    pub loss: f64,
    pub train_samples: u64,
    pub queries: u64,
    pub train_ms: f64,
    pub render_ms: f64,
}

// ---- post-processing (shaders/post_process.glsl:23-49) ----------------------------
pub const VR_TONEMAP_NONE: i32 = 0;
pub const VR_TONEMAP_ACES: i32 = 1;
pub const VR_TONEMAP_REINHARD: i32 = 2;
pub const VR_TONEMAP_FILMIC: i32 = 3;
pub const VR_TONEMAP_UNCHARTED2: i32 = 4;
// the legacy viewer's set (voidray/src/pipeline.rs:86-97, Tonemap::to_int + 16)
pub const VR_TONEMAP_LEGACY_NONE: i32 = 16;
pub const VR_TONEMAP_LEGACY_SIMPLE_ACES: i32 = 17;
pub const VR_TONEMAP_LEGACY_SIMPLE_REINHARD: i32 = 18;
pub const VR_TONEMAP_LEGACY_LUMA_REINHARD: i32 = 19;
pub const VR_TONEMAP_LEGACY_LUMA_WHITE_PRESERVING_REINHARD: i32 = 20;
pub const VR_TONEMAP_LEGACY_ROM_BIN_DA_HOUSE: i32 = 21;
pub const VR_TONEMAP_LEGACY_FILMIC: i32 = 22;
pub const VR_TONEMAP_LEGACY_UNCHARTED2: i32 = 23;

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct vr_post_desc {
    pub width: u32,
    pub height: u32,
    pub scale: f32,
    pub gamma: f32,
    pub exposure: f32,
    pub tonemap: i32,
    pub clip_to_1024: u32,
    pub keep_alpha: u32,
}

#[link(name = "voidray_hip")]
extern "C" {
    // files (Mesh::from_file without obj-rs; host only)
    pub fn vr_mesh_load_obj(path: *const c_char, out: *mut vr_mesh) -> c_int;
    pub fn vr_mesh_release(mesh: *mut vr_mesh) -> c_int;

    // context and scene
    pub fn vr_create(device: c_int, out: *mut *mut vr_ctx) -> c_int;
    pub fn vr_destroy(ctx: *mut vr_ctx) -> c_int;
    pub fn vr_last_error() -> *const c_char;
    pub fn vr_stream(ctx: *mut vr_ctx) -> *mut c_void;
    pub fn vr_scene_upload(ctx: *mut vr_ctx, desc: *const vr_scene_desc, out: *mut *mut vr_scene) -> c_int;
    pub fn vr_scene_release(scene: *mut vr_scene) -> c_int;
    pub fn vr_scene_get_stats(scene: *const vr_scene, out: *mut vr_scene_stats) -> c_int;

    // ray queries
    pub fn vr_trace(scene: *mut vr_scene, desc: *const vr_trace_desc, origins: *const f32,
                    directions: *const f32, prim_id: *mut u32, t: *mut f32, u: *mut f32, v: *mut f32,
                    hit: *mut u8) -> c_int;
    pub fn vr_trace_device_async(scene: *mut vr_scene, desc: *const vr_trace_desc, d_origins: *const f32,
                                 d_directions: *const f32, d_prim_id: *mut u32, d_t: *mut f32, d_u: *mut f32,
                                 d_v: *mut f32, d_hit: *mut u8, stream: *mut c_void) -> c_int;
    // no allocation: scratch is one static device array per GPU (callers under
    // graph capture or in serving loops that cannot hipMalloc)
    pub fn vr_trace_device_noalloc_async(scene: *mut vr_scene, desc: *const vr_trace_desc,
                                         d_origins: *const f32, d_directions: *const f32, d_prim_id: *mut u32,
                                         d_t: *mut f32, d_u: *mut f32, d_v: *mut f32, d_hit: *mut u8,
                                         stream: *mut c_void) -> c_int;
    pub fn vr_trace_host(scene: *const vr_scene_desc, desc: *const vr_trace_desc, origins: *const f32,
                         directions: *const f32, prim_id: *mut u32, t: *mut f32, u: *mut f32, v: *mut f32,
                         hit: *mut u8, stats: *mut vr_scene_stats, fallbacks: *mut u32) -> c_int;
    pub fn vr_trace_last_fallbacks(scene: *mut vr_scene, count: *mut u32) -> c_int;
    pub fn vr_trace_last_timing(scene: *mut vr_scene, wide_ms: *mut f32) -> c_int;

    // rendering
    pub fn vr_render(scene: *mut vr_scene, desc: *const vr_render_desc, rgba_accum: *mut f32,
                     stats: *mut vr_render_stats) -> c_int;
    pub fn vr_render_device_async(scene: *mut vr_scene, desc: *const vr_render_desc, d_rgba_accum: *mut f32,
                                  stats: *mut vr_render_stats, stream: *mut c_void) -> c_int;
    pub fn vr_render_last_stats(scene: *mut vr_scene, stats: *mut vr_render_stats) -> c_int;

    // batched serving
    pub fn vr_render_batch_async(ctx: *mut vr_ctx, requests: *const vr_request, count: u32,
                                 tuning: *const vr_tuning, flags: u32, stream: *mut c_void) -> c_int;
    pub fn vr_batch_relaunch_async(ctx: *mut vr_ctx, stream: *mut c_void) -> c_int;
    pub fn vr_batch_latencies(ctx: *mut vr_ctx, ms: *mut f64, count: u32) -> c_int;
    pub fn vr_batch_last_stats(ctx: *mut vr_ctx, stats: *mut vr_render_stats) -> c_int;
    pub fn vr_batch_autotune(ctx: *mut vr_ctx, requests: *const vr_request, count: u32, flags: u32, reps: u32,
                             stream: *mut c_void, best: *mut vr_tuning, best_ms: *mut f32) -> c_int;

    // learned radiance cache
    pub fn vr_cache_create(ctx: *mut vr_ctx, desc: *const vr_cache_desc, out: *mut *mut vr_cache) -> c_int;
    pub fn vr_cache_destroy(cache: *mut vr_cache) -> c_int;
    pub fn vr_cache_get_info(cache: *const vr_cache, info: *mut vr_cache_info) -> c_int;
    pub fn vr_cache_get_params(cache: *mut vr_cache, mlp: *mut f32, grid: *mut f32) -> c_int;
    pub fn vr_cache_set_params(cache: *mut vr_cache, mlp: *const f32, grid: *const f32) -> c_int;
    pub fn vr_cache_infer_async(cache: *mut vr_cache, n: u32, d_pos: *const f32, d_dir: *const f32,
                                d_nrm: *const f32, d_rough: *const f32, d_out: *mut f32,
                                stream: *mut c_void) -> c_int;
    pub fn vr_cache_gradients_async(cache: *mut vr_cache, n: u32, d_pos: *const f32, d_dir: *const f32,
                                    d_nrm: *const f32, d_rough: *const f32, d_target: *const f32,
                                    d_mlp_grad: *mut f32, d_grid_grad: *mut f32, d_loss: *mut f32,
                                    stream: *mut c_void) -> c_int;
    pub fn vr_cache_train_async(cache: *mut vr_cache, n: u32, d_pos: *const f32, d_dir: *const f32,
                                d_nrm: *const f32, d_rough: *const f32, d_target: *const f32,
                                stream: *mut c_void) -> c_int;
    pub fn vr_render_cached_async(scene: *mut vr_scene, cache: *mut vr_cache, desc: *const vr_render_desc,
                                  frame: *const vr_cache_frame, d_rgba: *mut f32, stream: *mut c_void) -> c_int;
    pub fn vr_cache_last_stats(cache: *mut vr_cache, stats: *mut vr_cache_stats) -> c_int;

    // post-processing
    pub fn vr_tonemap(ctx: *mut vr_ctx, desc: *const vr_post_desc, src_rgba: *const f32,
                      dst_rgba: *mut f32) -> c_int;
    pub fn vr_tonemap_device_async(ctx: *mut vr_ctx, desc: *const vr_post_desc, d_src: *const f32,
                                   d_dst: *mut f32, stream: *mut c_void) -> c_int;
}
