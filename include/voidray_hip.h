/*
 * voidray_hip.h — C ABI of the MI355X path-tracing core (libvoidray_hip.so).
 *
 * Every entry point is plain C: POD structs, raw pointers, sizes and an int
 * status (0 = ok, < 0 = error code; the message is in vr_last_error()).  No
 * call unwinds across the ABI.  Calls are synchronous at return unless the
 * name ends in _async (then they are ordered on the stream passed in).
 *
 * Each declaration names the reference item it replaces (paths relative to
 * the voidray repository).
 */
#ifndef VOIDRAY_HIP_H
#define VOIDRAY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VR_OK 0
#define VR_ERR_INVALID (-1)     /* bad argument / scene description        */
#define VR_ERR_HIP (-2)         /* HIP runtime failure                     */
#define VR_ERR_OOM (-3)         /* device allocation failed                */
#define VR_ERR_UNSUPPORTED (-4) /* feature not available in this build     */

typedef struct vr_ctx vr_ctx;
typedef struct vr_scene vr_scene;

/* ---- scene description ------------------------------------------------ */

/* One triangle mesh in world space.  Mirrors Mesh::from_buffers /
 * Mesh::from_file (voidray_renderer/src/core/mesh.rs:46-116).  normals/uvs
 * may be NULL: that is the Vertex::position constructor (mesh.rs:26-33),
 * i.e. zero normals and zero uvs. */
typedef struct vr_mesh {
    uint32_t num_vertices;
    uint32_t num_triangles;
    const float* positions; /* 3 * num_vertices */
    const float* normals;   /* 3 * num_vertices or NULL */
    const float* uvs;       /* 2 * num_vertices or NULL */
    const uint32_t* indices;/* 3 * num_triangles */
} vr_mesh;

/* Mesh::from_file (mesh.rs:46-74): reads a Wavefront OBJ as obj-rs 0.7 does
 * for Obj<TexturedVertex, u32> (triangles with position/texture/normal
 * references; vertices = distinct reference triples in order of first use).
 * On success *out holds library-allocated arrays (normals and uvs set), to be
 * given back with vr_mesh_release.  Errors (unreadable file, untriangulated
 * faces, missing uv or normal references, indices out of range) return
 * VR_ERR_INVALID where the reference panics.  Host only; no GPU needed. */
int vr_mesh_load_obj(const char* path, vr_mesh* out);
int vr_mesh_release(vr_mesh* mesh);

/* Materials of voidray_common/src/simple.rs:17-58. */
#define VR_MAT_LAMBERTIAN 0      /* Lambertian::scatter  simple.rs:102-132 */
#define VR_MAT_METAL 1           /* Metal::scatter       simple.rs:140-160 */
#define VR_MAT_DIELECTRIC 2      /* Dielectric::scatter  simple.rs:200-232 */
#define VR_MAT_EMISSION 3        /* Emission::scatter    simple.rs:175-185 */
#define VR_MAT_LAMBERTIAN_BSDF 4 /* LambertianBSDF via traits.rs:19-41      */
#define VR_MAT_MICROFACET 5      /* MicrofacetBSDF (voidray_common/src/microfacet.rs) via traits.rs:19-41 */

typedef struct vr_material {
    int32_t kind;
    float color[3];         /* albedo, or emission colour * strength      */
    float param;            /* Metal fuzz, Dielectric ir                  */
    int32_t albedo_texture; /* Lambertian texture albedo, -1 = none       */
    int32_t normal_texture; /* Lambertian normal map, -1 = none           */
    /* MicrofacetBSDF (microfacet.rs:9-27); emittance is stored but, as in the
     * reference's BSDFMaterial::scatter, never emitted */
    float index, roughness, metallic, emittance;
    int32_t transparent;
} vr_material;

#define VR_TEX_RGB32F 0   /* image.to_rgb32f() layout (texture.rs:37)   */
#define VR_TEX_RGBA8 1    /* 8-bit unorm, converted as x / 255          */
#define VR_SAMPLE_NEAREST 0
#define VR_SAMPLE_BILINEAR 1

/* ImageTexture (voidray_renderer/src/core/texture.rs:28-99). */
typedef struct vr_texture {
    uint32_t width, height;
    int32_t format;       /* VR_TEX_* */
    int32_t sample_type;  /* VR_SAMPLE_* */
    const void* data;     /* row-major, width*height texels */
} vr_texture;

#define VR_ENV_NONE 0
#define VR_ENV_UNIFORM 1 /* UniformEnvironment  environments.rs:19-33 */
#define VR_ENV_HDRI 2    /* HDRIEnvironment     environments.rs:35-91 */

typedef struct vr_environment {
    int32_t kind;
    float color[3];       /* uniform colour */
    uint32_t width, height;
    const float* data;    /* HDRI: RGB f32, width*height*3 */
} vr_environment;

/* Camera (voidray_renderer/src/core/camera.rs:5-54): eye, unit direction,
 * orthonormalised up, fov along the longer image side, optional depth of
 * field (aperture, focal point). */
typedef struct vr_camera {
    float eye[3];
    float direction[3];
    float up[3];
    float fov;
    int32_t has_dof;
    float aperture;
    float focal_point[3];
} vr_camera;

/* Object = (surface, material) (core/scene.rs:30-34). */
typedef struct vr_object {
    uint32_t surface;
    uint32_t material;
} vr_object;

/* Analytic surface (voidray_common/src/surfaces.rs:31-114): a sphere
 * (center, radius) or a ground plane y = height.  A ray query that returns one
 * reports its surface's single prim id and, in u and v, the hit record's uv:
 * (0, 0) on a sphere, (x, z) of the hit point on a ground plane. */
#define VR_ANALYTIC_SPHERE 0
#define VR_ANALYTIC_GROUND_PLANE 1
typedef struct vr_analytic {
    int32_t kind;      /* VR_ANALYTIC_* */
    float center[3];
    float radius;
    float height;
} vr_analytic;

/* surface_mesh[i] with this bit set: surface i is analytic[surface_mesh[i] & ~bit]
 * (Scene::add_analytic_surface, scene.rs:118-124) */
#define VR_SURFACE_ANALYTIC 0x80000000u

/* An instance: mesh `mesh` placed by the affine map world = M p + c, with M
 * the row-major 3x3 part of `transform` (elements 0-2, 4-6, 8-10) and c its
 * last column (3, 7, 11).  surface_mesh[i] = VR_SURFACE_INSTANCE | k makes
 * surface i instance k.  All instances of a mesh share one acceleration
 * structure in object space; the ray enters it through the inverse map, so t
 * keeps its world-space meaning; normals go through the inverse transpose.
 * No counterpart in the reference (every surface there is its own mesh):
 * scenes with instances render with the extended estimator only (the
 * reference estimator refuses them).  Every engine traces them: inside an
 * instance the wide engines take every candidate the conservative boxes
 * reach (no reference reach test), the exact engine walks the mesh's
 * median-split tree with the mapped ray.  Prim ids count an instance's
 * triangles like a mesh surface's.  The transform must be finite and
 * invertible (VR_ERR_INVALID otherwise). */
#define VR_SURFACE_INSTANCE 0x40000000u
typedef struct vr_instance {
    uint32_t mesh;
    float transform[12];
} vr_instance;

/* Whole scene (core/scene.rs:36-44).  Surface i is mesh surface_mesh[i]; as in
 * SceneAcceleration::hit (scene.rs:182-185) the material of a hit on surface
 * i is objects[i].material, so objects must have at least num_surfaces
 * entries. */
typedef struct vr_scene_desc {
    uint32_t num_meshes;
    const vr_mesh* meshes;
    uint32_t num_surfaces;
    const uint32_t* surface_mesh;
    uint32_t num_objects;
    const vr_object* objects;
    uint32_t num_materials;
    const vr_material* materials;
    uint32_t num_textures;
    const vr_texture* textures;
    vr_environment environment;
    vr_camera camera;
    uint32_t build_flags; /* VR_BUILD_* */
    uint32_t num_analytic;
    const vr_analytic* analytic;
    uint32_t num_instances;
    const vr_instance* instances;
} vr_scene_desc;

#define VR_BUILD_DEFAULT 0
#define VR_BUILD_NO_ENV_CDF 1u /* skip the environment importance-sampling tables */
/* Leaf layout of the wide BVH.  Spatial splits put a triangle into every leaf
 * holding a piece of it.  By default each leaf holds its 48-byte triangle
 * records (the fastest layout).  VR_BUILD_LEAF_REFS makes leaves hold 4-byte
 * references into one record per triangle instead: on the 1M-triangle
 * random soup (5.7 records per triangle) that is a third less device memory
 * for ~5% lower closest-hit throughput.  Same hits either way. */
#define VR_BUILD_LEAF_RECORDS 2u /* the default, stated explicitly */
#define VR_BUILD_LEAF_REFS 4u

typedef struct vr_scene_stats {
    uint32_t num_triangles;
    uint32_t num_surfaces;
    uint32_t num_qnodes;        /* wide quantised nodes */
    uint32_t num_ref_nodes;     /* flattened reference median-split nodes */
    uint32_t qbvh_depth;
    uint32_t max_leaf_size;
    double build_ref_seconds;   /* median-split trees (winner confirmation + exact engine) */
    double build_sah_seconds;   /* SAH build + collapse + quantisation   */
    double upload_seconds;
    uint64_t device_bytes;
    uint32_t num_tri_records;   /* triangle records in the wide BVH's leaves: >= num_triangles,
                                   a triangle cut by spatial splits has one per leaf holding a piece */
    uint32_t leaf_refs;         /* 1: leaves hold references into one record per triangle */
    uint64_t qbox_checked;      /* build self-check: (child, axis) bounds of the wide nodes, decoded as the
                                   kernels decode them, compared with the exact boxes they stand for */
    uint64_t qbox_not_containing; /* ... of which did not contain the exact box (0 for a correct build) */
} vr_scene_stats;

/* ---- context ---------------------------------------------------------- */

/* Replaces the rayon pool + RenderThread set-up (render/renderer.rs:35-60).
 * One context per host thread; not re-entrant. */
int vr_create(int device, vr_ctx** out);
int vr_destroy(vr_ctx* ctx);
/* Thread-local message of the last failing call on this thread. */
const char* vr_last_error(void);
/* Device stream used by the synchronous entry points (hipStream_t). */
void* vr_stream(vr_ctx* ctx);

/* Scene::build_acceleration (core/scene.rs:163-179) + Mesh::from_buffers
 * BVH build (mesh.rs:111-114): flattens the scene, builds the reference
 * median-split trees (bvh.rs:48-130) and the SAH quantised wide BVH, and
 * uploads everything to HBM.  All input pointers are borrowed for the call. */
int vr_scene_upload(vr_ctx* ctx, const vr_scene_desc* desc, vr_scene** out);
int vr_scene_release(vr_scene* scene);
int vr_scene_get_stats(const vr_scene* scene, vr_scene_stats* out);

/* ---- ray queries (SceneAcceleration::hit, core/scene.rs:182-185) -------- */

#define VR_TRACE_CLOSEST 0
#define VR_TRACE_ANY 1
/* OR-ed into vr_trace_desc.mode: record HIP events around the wide-engine
 * kernel so that vr_trace_last_timing can report its device time (two event
 * records per call, a few microseconds; off by default). */
#define VR_TRACE_TIMING 0x100
/* traversal engines */
#define VR_ENGINE_WIDE 0      /* quantised BVH4 + robust-winner check + exact fallback */
#define VR_ENGINE_REFERENCE 1 /* flattened reference trees, both children, no pruning  */
/* The wide BVH with the watertight ray/triangle test (Woop, Benthin & Wald,
 * JCGT 2013) in place of Moller-Trumbore: no ray slips between the triangles
 * of a closed mesh.  No counterpart in the reference (Triangle::hit,
 * mesh.rs:144-188, is Moller-Trumbore): hits are not the reference's, so
 * renders accept it only with the extended estimator (VR_RENDER_ENV_IS or
 * VR_RENDER_RUSSIAN_ROULETTE). */
#define VR_ENGINE_WATERTIGHT 2

typedef struct vr_trace_desc {
    uint32_t num_rays;
    float t_min;          /* SceneAcceleration uses 0.00001 (scene.rs:183) */
    int32_t mode;         /* VR_TRACE_CLOSEST or VR_TRACE_ANY, optionally | VR_TRACE_TIMING */
    int32_t engine;       /* VR_ENGINE_* */
} vr_trace_desc;

/* Closest-hit: prim_id = global triangle index (surfaces in order, triangles
 * in face order, one prim per analytic surface) or 0xFFFFFFFF on a miss; t, u,
 * v = Möller–Trumbore values of Triangle::hit (mesh.rs:144-188), or t and the
 * hit record's uv of an analytic surface (see vr_analytic).  Any-hit: hit[i] = 1 iff the reference
 * query returns Some, prim_id/t/u/v may be NULL.
 * Host pointers: origins/directions are 3*n floats (xyz per ray, directions
 * already normalised as Ray::new does, ray.rs:12-17). */
int vr_trace(vr_scene* scene, const vr_trace_desc* desc,
             const float* origins, const float* directions,
             uint32_t* prim_id, float* t, float* u, float* v, uint8_t* hit);
/* Same with device pointers, enqueued on `stream` (hipStream_t, NULL = the
 * context stream).  Any pointer among the outputs may be NULL.
 * One stream per scene: every trace and render call on a scene uses that
 * scene's device scratch (ray-fetch and fallback counters, fallback lists,
 * traversal spill stacks), so calls on one scene that may overlap (other
 * streams, threads) corrupt each other's results; enqueue them on one stream
 * or synchronise between streams.  Different scenes are independent.  (The
 * reference's SceneAcceleration::hit is &self: its callers may share one
 * across threads; here each thread needs its own vr_scene or a common
 * stream.) */
int vr_trace_device_async(vr_scene* scene, const vr_trace_desc* desc,
                          const float* d_origins, const float* d_directions,
                          uint32_t* d_prim_id, float* d_t, float* d_u, float* d_v,
                          uint8_t* d_hit, void* stream);
/* vr_trace_device_async without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The kernels' scratch (counter block, fallback and overflow lists, traversal
 * spill stacks: VR_NOALLOC_SCRATCH_BYTES) is one static device array per GPU,
 * set up by vr_create on that GPU and held until the process ends.  At most
 * VR_NOALLOC_MAX_RAYS rays per call (more: VR_ERR_INVALID, nothing enqueued).
 * Every call on one GPU shares that array, so calls that overlap on one GPU
 * (other streams, threads or contexts) race: enqueue them on one stream.
 * Results equal vr_trace_device_async's; vr_trace_last_fallbacks and
 * vr_trace_last_timing do not see these calls. */
#define VR_NOALLOC_MAX_RAYS 1048576u
#define VR_NOALLOC_SCRATCH_BYTES 75501568u /* 4 KiB + 2 x 4 MiB + 64 MiB */
int vr_trace_device_noalloc_async(vr_scene* scene, const vr_trace_desc* desc,
                                  const float* d_origins, const float* d_directions,
                                  uint32_t* d_prim_id, float* d_t, float* d_u, float* d_v,
                                  uint8_t* d_hit, void* stream);
/* The same queries answered on the host, without a GPU: the scene is built
 * into host memory exactly as vr_scene_upload builds it, and the rays run
 * through the same traversal routines the kernels use (wide engine with its
 * exact-engine fallback, or the exact engine), single-threaded.  For checking
 * the BVH build and the traversal where no GPU is present.  stats and
 * fallbacks (rays the wide engine handed to the exact engine) may be NULL.
 * No counterpart in the reference. */
int vr_trace_host(const vr_scene_desc* scene, const vr_trace_desc* desc,
                  const float* origins, const float* directions,
                  uint32_t* prim_id, float* t, float* u, float* v, uint8_t* hit,
                  vr_scene_stats* stats, uint32_t* fallbacks);
/* Rays of the last trace that needed the exact fallback engine (call after
 * the stream of that trace has been synchronised). */
int vr_trace_last_fallbacks(vr_scene* scene, uint32_t* count);
/* Device time (ms, HIP events on the trace's stream) of the wide-engine kernel
 * of the last vr_trace / vr_trace_device_async, which must have carried
 * VR_TRACE_TIMING in its mode; waits for that kernel. */
int vr_trace_last_timing(vr_scene* scene, float* wide_ms);

/* ---- rendering (render/iterative.rs:11-55, core/tracer.rs:9-56) ------- */

#define VR_MODE_FULL 0    /* RenderMode::Full   */
#define VR_MODE_NORMAL 1  /* RenderMode::Normal */

#define VR_RENDER_ROWS_BY_HEIGHT 1u /* reproduce y = i / height (iterative.rs:26) */
#define VR_RENDER_ENV_IS 2u         /* environment importance sampling + MIS     */
#define VR_RENDER_RUSSIAN_ROULETTE 4u
#define VR_RENDER_ALPHA_PER_CALL 8u /* alpha += 1 per call, as iterative.rs:51   */
#define VR_RENDER_OVERWRITE 16u     /* store the frame instead of adding it (alpha = 1 with ALPHA_PER_CALL) */

typedef struct vr_render_desc {
    uint32_t width, height;
    uint32_t samples;        /* samples traced by this call (iterative_render `samples`) */
    uint32_t total_samples;  /* RenderSettings::total_samples (scale 1/total)          */
    uint32_t sample_offset;  /* index of the first sample of this call in the stream  */
    uint32_t max_bounces;    /* RenderSettings::max_bounces */
    float firefly_clamp;     /* RenderSettings::firefly_clamp (inf = off) */
    int32_t render_mode;     /* VR_MODE_* */
    uint32_t flags;          /* VR_RENDER_* */
    uint32_t rr_depth;       /* Russian roulette from this depth (flag RR) */
    uint64_t seed;           /* Philox4x32-10 key */
    uint32_t request;        /* request id (third counter word) */
    uint32_t engine;         /* VR_ENGINE_* for extension rays */
} vr_render_desc;

typedef struct vr_render_stats {
    uint64_t segments;        /* rays traced (all bounces) */
    uint64_t samples;         /* camera samples */
    uint64_t env_misses;
    uint64_t fallback_rays;   /* rays re-traced by the exact engine */
    uint64_t capped_loops;    /* rejection loops that hit their cap */
    double kernel_ms;         /* device time of the render passes */
    uint64_t textured_hits;   /* hits on a material with an albedo or normal texture */
    uint64_t env_is_samples;  /* environment importance samples (extended estimator) */
} vr_render_stats;

/* Adds (sum over `samples` of L) * (1/total_samples) into rgba_accum (host,
 * width*height*4 f32, additive, like iterative.rs:45-51). */
int vr_render(vr_scene* scene, const vr_render_desc* desc, float* rgba_accum,
              vr_render_stats* stats);
/* Same into a device accumulator, enqueued on `stream` (NULL = context
 * stream); `stats` is not filled (use vr_render_last_stats). */
int vr_render_device_async(vr_scene* scene, const vr_render_desc* desc,
                           float* d_rgba_accum, vr_render_stats* stats, void* stream);
/* Counters of the last render or batch on the scene's context (waits for its
 * statistics read-back). */
int vr_render_last_stats(vr_scene* scene, vr_render_stats* stats);

/* ---- batched serving ----------------------------------------------------
 * The reference serves one frame at a time: RenderThread::one_shot
 * (render/renderer.rs:67-107) calls iterative_render once per request.  These
 * entry points serve many requests (different scenes, sizes, spp, seeds) in
 * one launch: requests on the same scene share wavefront chunks, the whole
 * batch is enqueued without host synchronisation and can be captured into a
 * hipGraph and relaunched.  Each request's image equals its standalone
 * vr_render_device_async result.  Render calls on one context share its
 * scratch: enqueue them on one stream (or synchronise between streams). */

typedef struct vr_request {
    vr_scene* scene;
    vr_render_desc desc;
    float* d_rgba;            /* device target, width*height*4 f32 */
} vr_request;

typedef struct vr_tuning {
    uint32_t chunk_paths;     /* paths per wavefront chunk (queue length, at most 2^27);
                                 0 = 2^27, halved while the batch's path state would
                                 exceed 48 GiB */
    uint32_t block;           /* workgroup size of raygen/shade/accumulate; 0 = 1024 */
} vr_tuning;

#define VR_BATCH_GRAPH 1u          /* capture into a hipGraph and launch it (relaunchable) */
#define VR_BATCH_TIMESTAMPS 2u     /* record per-request completion times */
#define VR_BATCH_SHORTEST_FIRST 4u /* run scene groups in ascending order of paths */

/* Enqueues the batch on `stream` (NULL = context stream).  A captured graph
 * stays valid until the next render call on the context. */
int vr_render_batch_async(vr_ctx* ctx, const vr_request* requests, uint32_t count,
                          const vr_tuning* tuning, uint32_t flags, void* stream);
/* Launches the last captured batch again (same targets, same random streams). */
int vr_batch_relaunch_async(vr_ctx* ctx, void* stream);
/* Per-request completion times of the last VR_BATCH_TIMESTAMPS batch, in ms
 * after the batch started on the device (waits for the batch). */
int vr_batch_latencies(vr_ctx* ctx, double* ms, uint32_t count);
/* Statistics of the last batch (waits for it). */
int vr_batch_last_stats(vr_ctx* ctx, vr_render_stats* stats);
/* Times every candidate (chunk_paths x block) on the batch as a captured
 * graph, `reps` timed launches each, and returns the fastest.  The targets
 * are written during tuning: clear them afterwards. */
int vr_batch_autotune(vr_ctx* ctx, const vr_request* requests, uint32_t count, uint32_t flags,
                      uint32_t reps, void* stream, vr_tuning* best, float* best_ms);

/* ---- learned radiance cache ---------------------------------------------
 * No counterpart in the reference (its integrator traces every path to the
 * end, core/tracer.rs:9-56).  A multiresolution hash grid on position (16
 * levels x 2^log2_table entries x 2 FP16 features) plus one-blob encodings of
 * the outgoing direction, shading normal and roughness feed an MLP
 * 128 -> 64 -> 64 -> 64 -> 64 -> 3 (ReLU, no biases) that predicts outgoing
 * radiance.  BF16 weights and activations on MFMA with FP32 accumulation
 * (precision VR_CACHE_FP32: an all-FP32 debug path on the vector ALUs).
 * Trained online with Adam on a relative L2 loss against path-suffix radiance
 * produced by the tracer itself; queried in place of the rest of a path. */

typedef struct vr_cache vr_cache;

#define VR_CACHE_BF16 0
#define VR_CACHE_FP32 1
#define VR_CACHE_LEVELS 16
#define VR_CACHE_MLP_PARAMS 21504 /* 64x128 + 3 x 64x64 + 16x64 (rows 3..15 of the last layer unused) */

typedef struct vr_cache_desc {
    uint32_t log2_table;      /* entries per level = 2^log2_table (19 for the full cache) */
    uint32_t base_resolution; /* coarsest grid resolution (cells per axis) */
    uint32_t max_resolution;  /* finest grid resolution */
    uint32_t precision;       /* VR_CACHE_* */
    float bbox_lo[3], bbox_hi[3]; /* positions are normalised by the largest extent of this box */
    float lr_grid, lr_mlp;    /* Adam learning rates */
    float beta1, beta2, eps;
    uint64_t seed;            /* parameter initialisation */
} vr_cache_desc;

typedef struct vr_cache_info {
    uint32_t levels, table_size, precision, pad;
    uint32_t resolution[VR_CACHE_LEVELS];
    float origin[3];
    float inv_extent;         /* p' = (p - origin) * inv_extent, clamped to [0, 1] */
    uint64_t grid_params;     /* levels * table_size * 2 */
    uint64_t steps;           /* Adam steps taken */
} vr_cache_info;

typedef struct vr_cache_frame {
    uint32_t train_paths;     /* expected training paths per frame, drawn at random among its paths */
    uint32_t train_vertices;  /* vertices per training path recorded as samples (<= 8) */
    uint32_t cache_depth;     /* other paths query the cache at this vertex (2 = after bounce 2) ... */
    uint32_t train_steps;     /* Adam steps per frame (<= 32), each on one random mini-batch of the frame's samples */
    uint32_t frame;           /* frame index: varies the choice of training paths and mini-batches */
    float query_roughness;    /* ... or at the first vertex from cache_depth on whose roughness input (1 diffuse,
                                 the fuzz of Metal, 0 Dielectric) is at least this; 0: at cache_depth always */
} vr_cache_frame;

typedef struct vr_cache_stats {
    double loss;              /* mean relative L2 loss of the last training step */
    uint64_t train_samples;   /* samples of the last frame */
    uint64_t queries;         /* cache queries of the last frame */
    double train_ms, render_ms;
} vr_cache_stats;

int vr_cache_create(vr_ctx* ctx, const vr_cache_desc* desc, vr_cache** out);
int vr_cache_destroy(vr_cache* cache);
int vr_cache_get_info(const vr_cache* cache, vr_cache_info* info);
/* Parameters as the forward pass uses them (BF16 weights / FP16 grid, as f32;
 * FP32 precision: the master copies).  mlp: VR_CACHE_MLP_PARAMS floats, rows
 * [out][in] per layer; grid: [level][entry][feature]. */
int vr_cache_get_params(vr_cache* cache, float* mlp, float* grid);
/* Sets the master copies (and their rounded forward copies); resets Adam. */
int vr_cache_set_params(vr_cache* cache, const float* mlp, const float* grid);
/* Device arrays: pos, dir (outgoing, unit), nrm (unit) n*3 f32, rough n f32.
 * out: n*3 f32 predicted radiance. */
int vr_cache_infer_async(vr_cache* cache, uint32_t n, const float* d_pos, const float* d_dir,
                         const float* d_nrm, const float* d_rough, float* d_out, void* stream);
/* Gradients of the mean relative L2 loss sum_c (y_c - t_c)^2 / (y_c^2 + 0.01)
 * / (3n) (denominator not differentiated) w.r.t. the forward parameters,
 * without an update.  d_mlp_grad: VR_CACHE_MLP_PARAMS f32; d_grid_grad:
 * grid_params f32; d_loss: 1 f32.  Any output may be NULL. */
int vr_cache_gradients_async(vr_cache* cache, uint32_t n, const float* d_pos, const float* d_dir,
                             const float* d_nrm, const float* d_rough, const float* d_target,
                             float* d_mlp_grad, float* d_grid_grad, float* d_loss, void* stream);
/* One Adam step on the given samples. */
int vr_cache_train_async(vr_cache* cache, uint32_t n, const float* d_pos, const float* d_dir,
                         const float* d_nrm, const float* d_rough, const float* d_target, void* stream);
/* One frame of the cached renderer (extended estimator: flags ENV_IS/RR):
 * desc->samples paths per pixel; each stops at vertex `cache_depth` and adds
 * throughput x cache(vertex), except ~train_paths randomly chosen training
 * paths that run to the end like the plain tracer's and record their first
 * `train_vertices` vertices with their path-suffix radiance as samples.  All
 * paths accumulate into d_rgba like vr_render_device_async; then
 * `train_steps` Adam steps on the frame's samples. */
int vr_render_cached_async(vr_scene* scene, vr_cache* cache, const vr_render_desc* desc,
                           const vr_cache_frame* frame, float* d_rgba, void* stream);
/* Statistics of the last frame (waits for it). */
int vr_cache_last_stats(vr_cache* cache, vr_cache_stats* stats);

/* ---- post-processing (shaders/post_process.glsl:23-49) ----------------- */

#define VR_TONEMAP_NONE 0
#define VR_TONEMAP_ACES 1
#define VR_TONEMAP_REINHARD 2
#define VR_TONEMAP_FILMIC 3
#define VR_TONEMAP_UNCHARTED2 4

/* The legacy viewer's pass (voidray/src/shaders/viewport_fragment.glsl with
 * voidray/src/shaders/tonemapping.glsl; ids are Tonemap::to_int,
 * voidray/src/pipeline.rs:86-97, plus 16).  It differs from the pass above:
 * exposure is a linear factor, not a power of two; Filmic forces gamma = 1;
 * alpha is scale * source alpha when keep_alpha is set (the viewer's
 * `transparent`), else 1. */
#define VR_TONEMAP_LEGACY_NONE 16
#define VR_TONEMAP_LEGACY_SIMPLE_ACES 17
#define VR_TONEMAP_LEGACY_SIMPLE_REINHARD 18
#define VR_TONEMAP_LEGACY_LUMA_REINHARD 19
#define VR_TONEMAP_LEGACY_LUMA_WHITE_PRESERVING_REINHARD 20
#define VR_TONEMAP_LEGACY_ROM_BIN_DA_HOUSE 21
#define VR_TONEMAP_LEGACY_FILMIC 22
#define VR_TONEMAP_LEGACY_UNCHARTED2 23

typedef struct vr_post_desc {
    uint32_t width, height;
    float scale;     /* total/current samples (voidray_app/src/main.rs:75-79) */
    float gamma;
    float exposure;
    int32_t tonemap; /* VR_TONEMAP_* (Tonemap::as_i32, settings.rs:44-54) */
    uint32_t clip_to_1024; /* reproduce the fixed 1024x1024 dispatch (post_process.rs:74) */
    uint32_t keep_alpha;   /* legacy modes only: output alpha = scale * source alpha */
} vr_post_desc;

int vr_tonemap(vr_ctx* ctx, const vr_post_desc* desc, const float* src_rgba, float* dst_rgba);
int vr_tonemap_device_async(vr_ctx* ctx, const vr_post_desc* desc, const float* d_src,
                            float* d_dst, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* VOIDRAY_HIP_H */
