// Internal layout of a scene in HBM and the host-side build products.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/voidray_hip.h"
#include "vrmath.h"

namespace vr {

// ---- errors ---------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
#define VR_HIP_CHECK(expr)                                                                              \
    do {                                                                                                \
        hipError_t _e = (expr);                                                                         \
        if (_e != hipSuccess)                                                                           \
            return ::vr::fail(VR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));           \
    } while (0)

// ---- wide quantised BVH node (64 B, one cache line pair of dwordx4 loads) --
// Child boxes are stored as 8-bit offsets on a per-node, per-axis power-of-two
// grid anchored at `origin`, rounded outward (lo floored, hi ceiled), so the
// decoded box always contains the exact (padded) child box.
struct alignas(16) QNode {
    float ox, oy, oz;
    uint8_t ex, ey, ez, nchild;  // biased exponents: scale = 2^(e - 127)
    uint32_t child[4];           // internal: node index; leaf: kLeafBit | (count-1) << 27 | first
    uint8_t qlo[3][4];           // [axis][child]
    uint8_t qhi[3][4];
    uint32_t pad[2];
};
static_assert(sizeof(QNode) == 64, "QNode must be 64 bytes");
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kInvalidChild = 0xFFFFFFFFu;
constexpr uint32_t kMaxLeaf = 3;  // triangles per leaf (count-1 fits 3 bits; 4 measured 0.7% slower)

// Triangle record in leaf order (48 B): world-space vertices exactly as the
// reference stores them and the global prim id.  Prim ids number the
// triangles surface by surface (surface index, then triangle index within the
// mesh), so the smaller prim id is the smaller (surface, triangle) handle pair
// that merge_optionals keeps on equal t (bvh.rs:164-181, ray.rs:52-62).
struct alignas(16) TriRec {
    float p0[3], p1[3], p2[3];
    uint32_t prim;
    uint32_t reserved;
    uint32_t flags;  // bit 0: the owning mesh has a reference BVH (box tests apply); bits 1..31: surface
};
static_assert(sizeof(TriRec) == 48, "TriRec must be 48 bytes");

// Flattened reference median-split node (bvh.rs:40-45), pre-order: the left
// child is the next node.  Leaf iff obj >= 0 (triangle index within the mesh,
// or surface index in the top-level tree).
struct alignas(16) RefNode {
    float lo[3], hi[3];
    int32_t right;
    int32_t obj;
};
static_assert(sizeof(RefNode) == 32, "RefNode must be 32 bytes");

struct MeshInfo {
    int32_t root;          // index into ref_nodes, -1: small mesh (brute force)
    uint32_t num_tris;
    uint32_t tri_base;     // first triangle of the mesh in the concatenated mesh arrays
    uint32_t vert_base;    // first vertex of the mesh in the concatenated vertex arrays
};

// An instance (vr_instance): the placement of a mesh, its inverse, the
// inverse transpose for normals, and the object-space wide BVH its rays walk.
// In the top-level wide BVH an instance is one leaf entry: a TriRec whose p0
// and p1 hold its world box and whose p2.x holds kInstanceMarker (a
// signalling-NaN pattern no vertex keeps: the build quiets such vertices), p2.y
// the instance index.
struct alignas(16) InstanceDev {
    float xf[12];          // object -> world, row-major 3x4
    float inv[12];         // world -> object
    float nrm[9];          // (M^-1)^T, row-major
    uint32_t mesh, surface, prim_base, blas_root;  // blas_root: QNode index
    float face_sign;       // sign of det M: the transformed winding's face normal is face_sign (M^-1)^T n
    float det_inv;         // 1 / |det M|: scales the ray/triangle determinant threshold in object space
    uint32_t pad;
};
static_assert(sizeof(InstanceDev) == 160, "InstanceDev layout");
constexpr uint32_t kInstanceMarker = 0x7FA1B2C3u;

struct SurfaceInfo {
    uint32_t mesh;         // mesh index (analytic surfaces: kNoMesh)
    uint32_t prim_base;    // global prim id of the surface's first triangle
    uint32_t material;     // objects[surface].material (scene.rs:184)
    uint32_t path_off;     // top-level ancestors: indices into top_path
    uint32_t path_len;
    int32_t analytic;      // index into SceneView::analytic, -1 for a mesh surface
    int32_t instance;      // index into SceneView::instances, -1 otherwise (mesh: the instanced mesh)
};
constexpr uint32_t kNoMesh = 0xFFFFFFFFu;

// Analytic surface (voidray_common/src/surfaces.rs:31-114) with its surface
// index and prim id.
struct alignas(16) AnalyticDev {
    int32_t kind;          // VR_ANALYTIC_*
    float c[3];            // sphere centre
    float r, h;            // sphere radius, ground-plane height
    uint32_t surface, prim;
};

struct GpuMaterial {
    int32_t kind;
    float color[3];
    float param;
    int32_t albedo_tex;
    int32_t normal_tex;
    int32_t transparent;        // MicrofacetBSDF
    float index, roughness, metallic;
    int32_t pad;
};

struct GpuTexture {
    uint32_t w, h;
    int32_t format, sample_type;
    uint64_t offset;  // byte offset into tex_data
};

// Device view of a scene, passed by value to kernels.
struct SceneView {
    const QNode* qnodes;
    const TriRec* tris;
    const uint32_t* tri_ref;      // per leaf entry: index into tris (nullptr: tris are in leaf order)
    const RefNode* ref_nodes;     // all mesh trees
    const RefNode* top_nodes;     // top-level tree over surfaces
    const uint32_t* top_path;     // per-surface ancestor lists (top_nodes indices)
    const MeshInfo* meshes;
    const SurfaceInfo* surfaces;
    const uint32_t* mesh_idx;     // per mesh triangle: 3 vertex indices (mesh-local)
    const float* face_n;          // per mesh triangle: face normal (mesh.rs:80-84)
    const float* vpos;            // concatenated vertices
    const float* vnrm;
    const float* vuv;
    const uint32_t* prim_surface; // global prim -> surface
    const int32_t* ref_parent;    // per ref node: parent ref node, -1 at a mesh root
    const int32_t* prim_anc;      // global prim -> parent of its leaf in the mesh tree (-1: none)
    const AnalyticDev* analytic;  // analytic surfaces, in surface order
    uint32_t num_analytic;
    const InstanceDev* instances;
    uint32_t num_instances;
    const GpuMaterial* materials;
    const GpuTexture* textures;
    const uint8_t* tex_data;
    const float* env;             // HDRI RGB f32
    const float* env_marg_cdf;    // env importance sampling tables (f32 copies of FP64 prefix sums)
    const float* env_cond_cdf;
    const float* env_pdf;         // per texel pdf (solid-angle measure factor folded in at use)
    const uint32_t* env_marg_guide;  // guide tables of the CDFs: marginal (env_gm + 1 entries) and
    const uint32_t* env_cond_guide;  // per row (env_gc + 1 entries each); see build_guide
    uint32_t env_gm, env_gc;
    uint32_t num_qnodes, num_tris, num_surfaces, top_root_valid;
    int32_t env_kind;
    float env_color[3];
    uint32_t env_w, env_h;
    float env_total;              // sum of luminance * sin(theta) weights (FP64 computed)
    // CameraAcceleration (camera.rs:57-67)
    float cam_o[3], cam_dir[3], cam_right[3], cam_up[3];
    float cam_d;
    int32_t dof;
    float aperture, focal_length;
};

}  // namespace vr

namespace vr {
struct RenderState;
}

struct vr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    vr::RenderState* render = nullptr;  // batch planning, render scratch, captured graph (render.hip)
};

struct vr_scene {
    vr_ctx* ctx = nullptr;
    vr::SceneView view{};
    std::vector<void*> allocations;
    vr_scene_stats stats{};
    // scratch for traces (grown on demand)
    uint32_t* fallback_list = nullptr;
    uint32_t* counters = nullptr;       // trace counter block (64 words; layout in trace.hip)
    uint32_t* host_counters = nullptr;  // pinned: [0] fallbacks, [1] overflows of the last vr_trace*
    uint32_t fallback_cap = 0;
    u32x2* spill = nullptr;             // per-lane spill stacks of the persistent traversal
    uint32_t spill_lanes = 0;
    uint32_t persist_blocks[8] = {};  // resident blocks of the persistent kernels [any + 2 * watertight + 4 * instances]
    hipEvent_t trace_ev[2] = {nullptr, nullptr};  // around the wide kernel of the last vr_trace*
    bool trace_timed = false;
    uint32_t last_fallbacks = 0;
    uint32_t num_prims = 0;
    bool has_microfacet = false;  // a material of kind VR_MAT_MICROFACET (the shading kernel's variant)
};

namespace vr {
// Host-side build products of a scene (the arrays that go to HBM).  `view`
// points into these vectors, so the traversal routines can also run on the
// host over the same data.
struct HostScene {
    std::vector<QNode> qnodes;
    std::vector<TriRec> tris;
    std::vector<uint32_t> tri_ref;  // leaf order -> tris index; empty: tris are in leaf order
    std::vector<RefNode> ref_nodes, top_nodes;
    std::vector<uint32_t> top_path, mesh_idx, prim_surface;
    std::vector<int32_t> ref_parent, prim_anc, mesh_leaf_parent;
    std::vector<MeshInfo> meshes;
    std::vector<SurfaceInfo> surfaces;
    std::vector<AnalyticDev> analytic;
    std::vector<InstanceDev> instances;
    std::vector<float> face_n, vpos, vnrm, vuv, env, env_marg, env_cond, env_pdf;
    std::vector<uint32_t> env_marg_guide, env_cond_guide;
    std::vector<GpuMaterial> materials;
    std::vector<GpuTexture> textures;
    std::vector<uint8_t> tex_data;
    SceneView view{};
    vr_scene_stats stats{};
};
// scene.cpp
int build_host_scene(const vr_scene_desc* desc, HostScene& h);
int build_scene(vr_ctx* ctx, const vr_scene_desc* desc, vr_scene* s);
void free_scene(vr_scene* s);
// trace.hip
struct TraceJob {
    const float* o = nullptr;       // 3 floats per ray
    const float* d = nullptr;
    uint32_t n_max = 0;             // rays (or queue capacity)
    const uint32_t* queue = nullptr;  // optional indirection
    const uint32_t* count = nullptr;  // optional device-side count
    float tmin = 1e-5f, tmax = kInf;
    bool any = false;
    int engine = VR_ENGINE_WIDE;
    uint32_t* prim = nullptr;
    float *t = nullptr, *u = nullptr, *v = nullptr;
    uint8_t* hit = nullptr;
    uint64_t* fallback_accum = nullptr;  // optional: += number of exact-engine rays
    bool timed = false;                  // record trace_ev[0..1] around the wide kernel
};
int launch_trace(vr_scene* s, const TraceJob& j, hipStream_t stream);
int trace_device(vr_scene* s, const vr_trace_desc* desc, const float* d_o, const float* d_d, uint32_t* d_prim,
                 float* d_t, float* d_u, float* d_v, uint8_t* d_hit, hipStream_t stream);
int ensure_trace_scratch(vr_scene* s, uint32_t n);
int noalloc_scratch(int device, void** out);
int trace_device_noalloc(vr_scene* s, const vr_trace_desc* desc, const float* d_o, const float* d_d,
                         uint32_t* d_prim, float* d_t, float* d_u, float* d_v, uint8_t* d_hit, hipStream_t stream);
}  // namespace vr
