// Wavefront path tracer: the MI355X form of iterative_render
// (voidray_renderer/src/render/iterative.rs:11-55) and trace_ray /
// trace_ray_internal (core/tracer.rs:9-56), for one request or a batch.
//
// A batch is a list of requests (scene, image size, spp, seed, settings, target).
// Requests on the same scene form a group; a group's paths (pixel-major,
// sample-minor, requests concatenated) are cut into chunks of whole pixels, at
// most `chunk_paths` paths each, so small requests share a wavefront.  Per chunk:
//   raygen   one path per (request, pixel, sample): jitter, camera ray (camera.rs:69-82)
//   for depth = 0 .. max over the chunk's requests of max_bounces - 1:
//     extend   closest hit of every live path (wide engine + exact fallback)
//     shade    hit record, material scatter, termination; survivors are
//              appended to the next queue with one atomic per block
//     shadow   (extended estimator) any-hit test of environment samples
//   accumulate  per pixel, samples summed in index order, * 1/total_samples,
//              added into the request's RGBA f32 target (iterative.rs:45-51)
// No host synchronisation inside: queue lengths live on the device, so a whole
// batch is captured once into a hipGraph and launched (and relaunched) as one.
//
// Reference estimator (parity): the per-level firefly clamp of tracer.rs:50 is
// evaluated exactly as the recursion does it.  Each path keeps its per-level
// attenuations; when it terminates with value X at level m the shade kernel
// folds x = min(a_j * x, clamp) for j = m-1 .. 0, the same sequence of f32
// operations as the recursive return path.
//
// Extended estimator (flags ENV_IS / RUSSIAN_ROULETTE, firefly clamp off):
// forward throughput, environment importance sampling with MIS (power
// heuristic) at Lambertian vertices, Russian roulette from rr_depth.
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>

#include "nrc.h"
#include "render.h"
#include "shade.h"

namespace vr {

// Paths per wavefront chunk.  Every bounce of a chunk is a few dependent
// launches over the chunk's live paths; once Russian roulette and escapes
// thin them out, a small chunk leaves the traversal kernels' persistent waves
// mostly idle.  On the 1080p x 64 spp terrain frame (132.7 M paths): 2^22
// paths per chunk 394 ms, 2^24 215 ms, 2^26 (two chunks) 146 ms, 2^27 (one
// chunk) 137 ms; about 133 B of path state per path with the extended
// estimator, 272 B with the reference estimator's 16-level attenuation stack:
// 18 or 36 GB of the GPU's 288 GB at 2^27.  Scratch is sized by the largest
// chunk a call actually uses, so small frames allocate little.
constexpr uint32_t kDefaultChunkPaths = 1u << 27;
constexpr uint32_t kMaxChunkPaths = 1u << 27;
// The default chunk is halved until the batch's path state fits this budget
// (the reference estimator's attenuation stack grows with max_bounces: at the
// 64-bounce limit a 2^27-path chunk would hold 110 GB).
constexpr size_t kDefaultScratchBudget = 48ull << 30;
constexpr uint32_t kMaxBouncesLimit = 64;
// Most threads a grid-stride kernel over a chunk launches: one round of the
// shading kernel's resident waves (256 CUs x 16 waves at its 128 VGPRs = 4,096
// waves), e.g. 256 blocks of 1,024 threads.  With blocks of 256, 1,024 blocks
// beat 2,048 on the render legs (cornell 5%, the 1080p frames 0.5-0.7%), and
// 768 or 512 left CUs idle (up to 11% slower).
constexpr uint32_t kShadeThreads = 4096u * 64u;

// Blocks of a grid-stride kernel over at most n items.
static inline uint32_t grid_for(uint32_t n, uint32_t block) {
    return std::max(1u, std::min(kShadeThreads / block, (n + block - 1) / block));
}

// render counters (u32 words per chunk)
enum : int {
    kCntActive = 0,                                  // [kMaxBouncesLimit + 1] live paths per depth
    kCntShadow = kCntActive + kMaxBouncesLimit + 1,  // [kMaxBouncesLimit] shadow rays per depth
    kCntCapped = kCntShadow + kMaxBouncesLimit,
    // counters every shading wave adds to, striped over kStatStripes 128-byte
    // lines (stripe = block index mod kStatStripes) so that the waves' atomics
    // do not all meet on one address; k_stats_chunk sums the stripes
    kCntStriped = (kCntCapped + 1 + 31) & ~31,
    kStatStripes = 64,
    kStripeWords = 32,
    kStripeMiss = 0,      // environment misses
    kStripeTextured = 1,  // hits on a material with a texture
    kStripeEnvIS = 2,     // environment importance samples (CDF lookups) for next-event estimation
    kCntWords = kCntStriped + kStatStripes * kStripeWords
};

// radiance-cache mode of a request: paths stop at the query vertex and take
// the cache's answer, except a random subset of training paths that run to the
// end and record their first vertices as training samples
enum : uint32_t { kCacheNone = 0, kCacheOn = 1 };

struct ReqDev {
    float* rgba;
    uint64_t seed;
    uint32_t W, H, spp, total, sample_offset, max_bounces, flags, rr_depth, request;
    int32_t mode;
    float clamp;
    uint32_t ext;
    uint32_t npix;            // W*H
    uint32_t cmode;           // kCache*
    uint32_t cache_depth;     // query vertex
    uint32_t train_vertices;  // recorded vertices per training path
    uint32_t train_threshold; // a path trains iff train_hash < threshold (threshold / 2^32 = fraction)
    uint32_t frame;
    float query_roughness;    // query at the first vertex from cache_depth on with cache_roughness >= this
};

struct SegDev {
    uint32_t req, pix0, npix, path_base, pix_base;
};

struct RenderArgs {
    SceneView sv;
    const ReqDev* reqs;
    const SegDev* segs;
    uint32_t nseg, npix, npaths, cap;
    uint32_t* path_req;
    uint32_t* path_pix;
    uint32_t* path_smp;
    float* ray_o;
    float* ray_d;
    uint32_t* hit_prim;
    float *hit_t, *hit_u, *hit_v;
    uint32_t* rng_n;
    float* att;  // reference estimator: [max_bounces][cap][3]
    float* L;    // [cap][3]
    float* A;    // extended: throughput [cap][3]
    float* pdf;  // extended: pdf of the last diffuse scatter (0 = camera / specular)
    float *sh_o, *sh_d, *sh_c;
    uint8_t* sh_hit;
    uint32_t* queue[2];
    uint32_t* sh_queue;
    uint32_t* cnt;
    NrcSamples tr, qr;  // radiance cache: training samples, queries
};

__device__ __forceinline__ vf3 ld3(const float* b, uint32_t p) {
    return mk3(b[3 * (size_t)p], b[3 * (size_t)p + 1], b[3 * (size_t)p + 2]);
}
__device__ __forceinline__ void st3(float* b, uint32_t p, vf3 v) {
    b[3 * (size_t)p] = v.x;
    b[3 * (size_t)p + 1] = v.y;
    b[3 * (size_t)p + 2] = v.z;
}

// Largest segment k with key(k) <= x (segments sorted by key, key(0) = 0).
template <class Key>
__device__ __forceinline__ uint32_t seg_find(const SegDev* s, uint32_t n, uint32_t x, Key key) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (key(s[mid]) <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Appends to two queues with one atomic per queue per block (every wave of
// the block must call it, the same number of times): the waves' counts meet
// in LDS, one lane adds the block's total to each counter, and each wave
// writes at the block's base plus the counts of the waves before it.
// Every shading wave appending to the same two counters directly made those
// two addresses a serialisation point.
__device__ __forceinline__ void block_append2(bool want0, bool want1, uint32_t value, uint32_t* q0, uint32_t* c0,
                                              uint32_t* q1, uint32_t* c1) {
    __shared__ uint32_t s_n[2][16];
    __shared__ uint32_t s_base[2];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
    const uint64_t m0 = __ballot(want0), m1 = __ballot(want1);
    if (lane == 0) {
        s_n[0][wave] = (uint32_t)__popcll(m0);
        s_n[1][wave] = (uint32_t)__popcll(m1);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < nw; ++w) tot += s_n[threadIdx.x][w];
        s_base[threadIdx.x] = tot ? atomicAdd(threadIdx.x ? c1 : c0, tot) : 0u;
    }
    __syncthreads();
    uint32_t off0 = s_base[0], off1 = s_base[1];
    for (uint32_t w = 0; w < wave; ++w) {
        off0 += s_n[0][w];
        off1 += s_n[1][w];
    }
    const uint64_t below = (1ull << lane) - 1ull;
    if (want0) q0[off0 + (uint32_t)__popcll(m0 & below)] = value;
    if (want1) q1[off1 + (uint32_t)__popcll(m1 & below)] = value;
    __syncthreads();  // s_n and s_base are rewritten by the next call
}

// Slots in two sample buffers for the lanes that want one, one atomic per
// buffer per block (every wave of the block must call it, the same number of
// times); UINT32_MAX for lanes that want none.
__device__ __forceinline__ void block_alloc2(bool want0, bool want1, uint32_t* c0, uint32_t* c1, uint32_t& slot0,
                                             uint32_t& slot1) {
    __shared__ uint32_t s_n[2][16];
    __shared__ uint32_t s_base[2];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
    const uint64_t m0 = __ballot(want0), m1 = __ballot(want1);
    if (lane == 0) {
        s_n[0][wave] = (uint32_t)__popcll(m0);
        s_n[1][wave] = (uint32_t)__popcll(m1);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < nw; ++w) tot += s_n[threadIdx.x][w];
        s_base[threadIdx.x] = tot ? atomicAdd(threadIdx.x ? c1 : c0, tot) : 0u;
    }
    __syncthreads();
    uint32_t off0 = s_base[0], off1 = s_base[1];
    for (uint32_t w = 0; w < wave; ++w) {
        off0 += s_n[0][w];
        off1 += s_n[1][w];
    }
    const uint64_t below = (1ull << lane) - 1ull;
    slot0 = want0 ? off0 + (uint32_t)__popcll(m0 & below) : 0xFFFFFFFFu;
    slot1 = want1 ? off1 + (uint32_t)__popcll(m1 & below) : 0xFFFFFFFFu;
    __syncthreads();  // s_n and s_base are rewritten by the next call
}

// Training-path selection: a 32-bit mix of (pixel, sample, request, frame)
// (the murmur3 finaliser), compared with the request's threshold.
__device__ __forceinline__ uint32_t train_hash(uint32_t pixel, uint32_t sample, uint32_t request, uint32_t frame) {
    uint32_t h = pixel * 0x9E3779B1u ^ (sample + 0x7F4A7C15u) * 0x85EBCA77u ^ request * 0xC2B2AE3Du ^ frame * 0x27D4EB2Fu;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

// Roughness input of the radiance cache: 1 for diffuse and emissive
// surfaces, the fuzz of Metal, 0 for Dielectric.
__device__ __forceinline__ float cache_roughness(const GpuMaterial& m) {
    if (m.kind == VR_MAT_METAL) return fminf(fmaxf(m.param, 0.f), 1.f);
    if (m.kind == VR_MAT_DIELECTRIC) return 0.f;
    return 1.f;
}

__global__ void k_render_zero(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_raygen(RenderArgs a) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.npaths) return;
    const SceneView& sv = a.sv;
    uint32_t k = seg_find(a.segs, a.nseg, p, [](const SegDev& s) { return s.path_base; });
    const SegDev sg = a.segs[k];
    const ReqDev& r = a.reqs[sg.req];
    uint32_t local = p - sg.path_base;
    uint32_t i = sg.pix0 + local / r.spp;
    uint32_t s = r.sample_offset + local % r.spp;
    a.path_req[p] = sg.req;
    a.path_pix[p] = i;
    a.path_smp[p] = s;
    uint32_t W = r.W, H = r.H;
    uint32_t x = i % W;
    uint32_t y = (r.flags & VR_RENDER_ROWS_BY_HEIGHT) ? i / H : i / W;
    float dd = (float)(W > H ? W : H);
    float xf = ((float)(2u * x + 1u) - (float)W) / dd;
    float yf = ((float)(uint32_t)(2u * (uint32_t)(H - y) - 1u) - (float)H) / dd;
    PathRng rng;
    rng.init(r.seed, r.request, i, s, 0);
    float lo = -1.0f / dd, hi = 1.0f / dd;
    float dx = rng.uniform() * (hi - lo) + lo;
    float dy = rng.uniform() * (hi - lo) + lo;
    // CameraAcceleration::cast_ray (camera.rs:69-82)
    vf3 right = mk3(sv.cam_right[0], sv.cam_right[1], sv.cam_right[2]);
    vf3 up = mk3(sv.cam_up[0], sv.cam_up[1], sv.cam_up[2]);
    vf3 origin = mk3(sv.cam_o[0], sv.cam_o[1], sv.cam_o[2]);
    vf3 nd = (sv.cam_d * mk3(sv.cam_dir[0], sv.cam_dir[1], sv.cam_dir[2]) + (xf + dx) * right) + (yf + dy) * up;
    uint32_t capped = 0;
    if (sv.dof) {
        vf3 fp = origin + normalize3(nd) * sv.focal_length;
        vf2 dk = unit_disc(rng, capped);
        origin = origin + (dk.x * right + dk.y * up) * sv.aperture;
        nd = fp - origin;
    }
    vf3 dir = normalize3(normalize3(nd));  // Ray::new(origin, new_dir.normalize())
    st3(a.ray_o, p, origin);
    st3(a.ray_d, p, dir);
    a.rng_n[p] = rng.n;
    st3(a.L, p, mk3(0.f, 0.f, 0.f));
    if (r.ext) {
        st3(a.A, p, mk3(1.f, 1.f, 1.f));
        a.pdf[p] = 0.f;
    }
    if (capped) atomicAdd(a.cnt + kCntCapped, capped);
}

// Folds the terminal value x at level m through the stored attenuations
// (the return path of trace_ray_internal, tracer.rs:44-50).
__device__ __forceinline__ vf3 fold_levels(const RenderArgs& a, uint32_t p, uint32_t m, vf3 x, float c) {
    for (uint32_t j = m; j-- > 0;) {
        vf3 aj = ld3(a.att + (size_t)j * 3 * a.cap, p);
        x = mk3(0.f, 0.f, 0.f) + vminc(mulv(aj, x), c);
    }
    return x;
}

// Register budget: with the default bounds (blocks of up to 1024 threads) the
// compiler holds this kernel at 128 VGPRs, 4 waves per SIMD (64 bytes of
// scratch per lane).  Measured on the terrain frame against other budgets
// (while the microfacet material's transcendentals were still
// inlined): 233 VGPRs without spills at 2 waves 9% slower, 168 at 3 waves 3%
// slower, 96 at 5 and 80 at 6 waves (more spills) 7% and 13% slower.
//
// EST: the estimators present in the launch (0: both, 1: reference only, 2:
// extended only); CACHE: radiance-cache requests present; MICRO: the scene
// has MicrofacetBSDF materials.  A launch compiles out what it does not run,
// which frees registers for what it does.
template <int EST, bool CACHE, bool MICRO>
__global__ void k_shade(RenderArgs a, uint32_t depth) {
    const SceneView& sv = a.sv;
    const uint32_t n = depth == 0 ? a.npaths : a.cnt[kCntActive + depth];
    const uint32_t* qin = depth == 0 ? nullptr : a.queue[depth & 1];
    uint32_t* qout = a.queue[(depth + 1) & 1];
    const uint32_t stride = gridDim.x * blockDim.x;
    // the same trip count for every thread of the block, so that all of them
    // reach the block-wide appends
    for (uint32_t kb = blockIdx.x * blockDim.x; kb < n; kb += stride) {
        uint32_t k = kb + threadIdx.x;
        bool act = k < n;
        uint32_t p = 0;
        bool cont = false, shadow = false, miss = false, textured = false, env_is = false;
        uint32_t rec = 0;  // radiance cache: 1 = training sample, 2 = query
        vf3 rec_pos{}, rec_dir{}, rec_n{}, rec_A{}, rec_L{};
        float rec_rough = 0.f;
        uint32_t capped = 0;
        if (act) {
            p = qin ? qin[k] : k;
            const ReqDev& r = a.reqs[a.path_req[p]];
            const bool ext = EST == 2 || (EST == 0 && r.ext);
            const float c = r.clamp;
            const bool last = depth + 1 >= r.max_bounces;
            PathRng rng;
            rng.init(r.seed, r.request, a.path_pix[p], a.path_smp[p], a.rng_n[p]);
            vf3 o = ld3(a.ray_o, p), d = ld3(a.ray_d, p);
            uint32_t prim = a.hit_prim[p];
            if (depth >= r.max_bounces) {
                // max_bounces == 0: trace_ray_internal returns BLACK without a query
                st3(a.L, p, mk3(0.f, 0.f, 0.f));
            } else if (prim == 0xFFFFFFFFu) {
                miss = true;
                vf3 env = sv.env_kind == VR_ENV_NONE ? mk3(0.f, 0.f, 0.f) : env_lookup(sv, d);
                if (!ext) {
                    st3(a.L, p, fold_levels(a, p, depth, env, c));
                } else {
                    float w = 1.f;
                    float pb = a.pdf[p];
                    if ((r.flags & VR_RENDER_ENV_IS) && pb > 0.f && sv.env_kind == VR_ENV_HDRI) {
                        float pe = env_pdf_dir(sv, d);
                        w = (pb * pb) / (pb * pb + pe * pe);
                    }
                    vf3 L = ld3(a.L, p), A = ld3(a.A, p);
                    st3(a.L, p, L + mulv(A, env) * w);
                }
            } else {
                HitRec hit = make_hit(sv, o, d, prim, a.hit_t[p], a.hit_u[p], a.hit_v[p]);
                GpuMaterial m = sv.materials[sv.surfaces[hit.surface].material];
                textured = m.albedo_tex >= 0 || m.normal_tex >= 0;
                if (!ext) {
                    if (r.mode == VR_MODE_NORMAL) {
                        vf3 nn = normalize3(hit.normal);
                        vf3 att = mk3(0.5f * nn.x + 0.5f, 0.5f * nn.y + 0.5f, 0.5f * nn.z + 0.5f);
                        st3(a.L, p, fold_levels(a, p, depth, mk3(0.f, 0.f, 0.f) + vminc(att, c), c));
                    } else {
                        ScatterOut sc = scatter<MICRO>(sv, m, d, hit, rng, capped, false);
                        if (!sc.has_ray) {
                            st3(a.L, p, fold_levels(a, p, depth, mk3(0.f, 0.f, 0.f) + vminc(sc.att, c), c));
                        } else {
                            st3(a.att + (size_t)depth * 3 * a.cap, p, sc.att);
                            if (last) {
                                st3(a.L, p, fold_levels(a, p, depth + 1, mk3(0.f, 0.f, 0.f), c));
                            } else {
                                st3(a.ray_o, p, hit.point);
                                st3(a.ray_d, p, sc.dir);
                                cont = true;
                            }
                        }
                    }
                } else {
                    vf3 A = ld3(a.A, p), L = ld3(a.L, p);
                    const bool train_path =
                        CACHE && r.cmode == kCacheOn &&
                        train_hash(a.path_pix[p], a.path_smp[p], r.request, r.frame) < r.train_threshold;
                    const bool query = CACHE && r.cmode == kCacheOn && !train_path && depth >= r.cache_depth &&
                                       cache_roughness(m) >= r.query_roughness;
                    if (query || (train_path && depth < r.train_vertices)) {
                        // the vertex, its outgoing direction and the path state before its contributions
                        vf3 ns = m.normal_tex >= 0 ? tex_sample(sv, m.normal_tex, hit.uv.x, hit.uv.y) : hit.normal;
                        rec = query ? 2u : 1u;
                        rec_pos = hit.point;
                        rec_dir = -d;
                        rec_n = normalize3(ns);
                        rec_rough = cache_roughness(m);
                        rec_A = A;
                        rec_L = L;
                    }
                    if (query) {
                        // the path ends: the cache supplies the radiance leaving this vertex
                    } else if (m.kind == VR_MAT_EMISSION) {
                        st3(a.L, p, L + mulv(A, mk3(m.color[0], m.color[1], m.color[2])));
                    } else {
                        if (m.kind == VR_MAT_LAMBERTIAN && (r.flags & VR_RENDER_ENV_IS) &&
                            sv.env_kind == VR_ENV_HDRI) {
                            // next-event estimation towards the environment
                            float u1 = rng.uniform(), u2 = rng.uniform();
                            vf3 ns = m.normal_tex >= 0 ? tex_sample(sv, m.normal_tex, hit.uv.x, hit.uv.y) : hit.normal;
                            ns = normalize3(ns);
                            float pe = 0.f;
                            vf3 wi = env_sample_dir(sv, u1, u2, pe);
                            env_is = true;
                            float cw = dot3(ns, wi);
                            if (cw > 0.f && pe > 0.f) {
                                float pb = cw / kPi;
                                float w = (pe * pe) / (pe * pe + pb * pb);
                                vf3 albedo = m.albedo_tex >= 0 ? tex_sample(sv, m.albedo_tex, hit.uv.x, hit.uv.y)
                                                               : mk3(m.color[0], m.color[1], m.color[2]);
                                vf3 contrib = mulv(mulv(A, albedo), env_lookup(sv, wi)) * (pb * w / pe);
                                st3(a.sh_o, p, hit.point);
                                st3(a.sh_d, p, wi);
                                st3(a.sh_c, p, contrib);
                                shadow = true;
                            }
                        }
                        ScatterOut sc = scatter<MICRO>(sv, m, d, hit, rng, capped, true);
                        if (sc.has_ray && !last) {
                            A = mulv(A, sc.att);
                            float pb = 0.f;
                            if (sc.diffuse) pb = fmaxf(0.f, dot3(sc.lobe, sc.dir)) / kPi;
                            // a non-finite throughput (a zero BSDF value over a sampling pdf
                            // that underflowed: 0 * inf) ends the path instead of reaching
                            // the pixel as NaN; the reference estimator's clamp handles it
                            bool alive = isfinite(A.x) && isfinite(A.y) && isfinite(A.z);
                            if (alive && (r.flags & VR_RENDER_RUSSIAN_ROULETTE) && depth >= r.rr_depth) {
                                float q = fminf(0.95f, fmaxf(A.x, fmaxf(A.y, A.z)));
                                if (rng.uniform() >= q) alive = false;
                                else A = A / q;
                            }
                            if (alive) {
                                st3(a.A, p, A);
                                a.pdf[p] = pb;
                                st3(a.ray_o, p, hit.point);
                                st3(a.ray_d, p, sc.dir);
                                cont = true;
                            }
                        }
                    }
                }
            }
            a.rng_n[p] = rng.n;
        }
        if (CACHE) {
            uint32_t slot_t = 0xFFFFFFFFu, slot_q = 0xFFFFFFFFu;
            if (a.tr.count)  // the same for the whole grid: a batch with radiance-cache requests
                block_alloc2(rec == 1u, rec == 2u, a.tr.count, a.qr.count, slot_t, slot_q);
            if (rec) {
                const bool t = rec == 1u;
                const uint32_t slot = t ? slot_t : slot_q;
                if (slot < (t ? a.tr.cap : a.qr.cap)) {
                    st3(t ? a.tr.pos : a.qr.pos, slot, rec_pos);
                    st3(t ? a.tr.dir : a.qr.dir, slot, rec_dir);
                    st3(t ? a.tr.nrm : a.qr.nrm, slot, rec_n);
                    (t ? a.tr.rough : a.qr.rough)[slot] = rec_rough;
                    (t ? a.tr.path : a.qr.path)[slot] = p;
                    st3(t ? a.tr.A : a.qr.A, slot, rec_A);
                    st3(t ? a.tr.L : a.qr.L, slot, rec_L);
                }
            }
        }
        block_append2(cont, shadow, p, qout, a.cnt + kCntActive + depth + 1, a.sh_queue, a.cnt + kCntShadow + depth);
        uint64_t mm = __ballot(miss);
        uint64_t tm = __ballot(textured);
        uint64_t em = __ballot(env_is);
        if ((threadIdx.x & 63u) == 0) {
            uint32_t* stripe = a.cnt + kCntStriped + (blockIdx.x % kStatStripes) * kStripeWords;
            if (mm) atomicAdd(stripe + kStripeMiss, (uint32_t)__popcll(mm));
            if (tm) atomicAdd(stripe + kStripeTextured, (uint32_t)__popcll(tm));
            if (em) atomicAdd(stripe + kStripeEnvIS, (uint32_t)__popcll(em));
        }
        if (capped) atomicAdd(a.cnt + kCntCapped, capped);
    }
}

__global__ void k_shadow_resolve(RenderArgs a, uint32_t depth) {
    const uint32_t n = a.cnt[kCntShadow + depth];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        uint32_t p = a.sh_queue[k];
        if (a.sh_hit[p] == 0) st3(a.L, p, ld3(a.L, p) + ld3(a.sh_c, p));
    }
}

// Cache queries of the chunk: L[path] += A * max(cache output, 0).
__global__ void k_cache_apply(RenderArgs a) {
    const uint32_t n = min(*a.qr.count, a.qr.cap);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t p = a.qr.path[k];
        vf3 y = ld3(a.qr.target, k);
        y = mk3(fmaxf(y.x, 0.f), fmaxf(y.y, 0.f), fmaxf(y.z, 0.f));
        st3(a.L, p, ld3(a.L, p) + mulv(ld3(a.qr.A, k), y));
    }
}

// Training targets: the radiance a path gathered from vertex k on, divided by
// the throughput that arrived at vertex k, (L_final - L_k) / A_k per channel.
// Samples of this chunk: [count[1], count[0]) (count[1] marks the chunk start).
__global__ void k_cache_targets(RenderArgs a) {
    const uint32_t n = min(a.tr.count[0], a.tr.cap);
    const uint32_t k0 = min(a.tr.count[1], n);
    for (uint32_t k = k0 + blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t p = a.tr.path[k];
        vf3 lf = ld3(a.L, p), lk = ld3(a.tr.L, k), ak = ld3(a.tr.A, k);
        float t[3];
        for (int c = 0; c < 3; ++c) {
            float v = comp(ak, c) > 1e-20f ? (comp(lf, c) - comp(lk, c)) / comp(ak, c) : 0.f;
            t[c] = (v == v && fabsf(v) < 3.0e38f) ? fmaxf(v, 0.f) : 0.f;
        }
        st3(a.tr.target, k, mk3(t[0], t[1], t[2]));
    }
}

__global__ void k_cache_mark(uint32_t* count) {
    if (threadIdx.x == 0 && blockIdx.x == 0) count[1] = count[0];
}

// Per pixel: sum of the samples in index order, scaled by 1/total_samples
// and added into the request's target (iterative.rs:45-51).
__global__ void k_accumulate(RenderArgs a) {
    uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.npix) return;
    uint32_t k = seg_find(a.segs, a.nseg, q, [](const SegDev& s) { return s.pix_base; });
    const SegDev sg = a.segs[k];
    const ReqDev& r = a.reqs[sg.req];
    uint32_t local = q - sg.pix_base;
    uint32_t base = sg.path_base + local * r.spp;
    vf3 color = mk3(0.f, 0.f, 0.f);
    for (uint32_t s = 0; s < r.spp; ++s) color = color + ld3(a.L, base + s);
    color = color * (1.0f / (float)r.total);
    float* t = r.rgba + 4 * (size_t)(sg.pix0 + local);
    if (r.flags & VR_RENDER_OVERWRITE) {
        t[0] = color.x;
        t[1] = color.y;
        t[2] = color.z;
        if (r.flags & VR_RENDER_ALPHA_PER_CALL) t[3] = 1.0f;
    } else {
        t[0] += color.x;
        t[1] += color.y;
        t[2] += color.z;
        if (r.flags & VR_RENDER_ALPHA_PER_CALL) t[3] += 1.0f;
    }
}

// Adds one chunk's counters into the 64-bit statistics
// [segments, samples, env misses, exact-engine rays, capped loops, textured hits, env IS samples].
__global__ void k_stats_chunk(const uint32_t* cnt, uint64_t* stats, uint32_t npaths, uint32_t max_depth) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t seg = max_depth ? npaths : 0;
    for (uint32_t k = 1; k < max_depth; ++k) seg += cnt[kCntActive + k];
    for (uint32_t k = 0; k < max_depth; ++k) seg += cnt[kCntShadow + k];
    stats[0] += seg;
    stats[1] += npaths;
    uint64_t miss = 0, tex = 0, env = 0;
    for (int s = 0; s < kStatStripes; ++s) {
        const uint32_t* stripe = cnt + kCntStriped + s * kStripeWords;
        miss += stripe[kStripeMiss];
        tex += stripe[kStripeTextured];
        env += stripe[kStripeEnvIS];
    }
    stats[2] += miss;
    stats[4] += cnt[kCntCapped];
    stats[5] += tex;
    stats[6] += env;
}

__global__ void k_stats_zero(uint64_t* stats) {
    if (threadIdx.x < 8) stats[threadIdx.x] = 0;
}

// Device wall-clock stamps (constant-rate counter) for per-request latency:
// stamps[0] = start of the batch, stamps[1 + r] = completion of request r.
__global__ void k_stamp_start(uint64_t* stamps) {
    if (threadIdx.x == 0 && blockIdx.x == 0) stamps[0] = wall_clock64();
}
__global__ void k_stamp_done(uint64_t* stamps, const uint32_t* reqs, uint32_t n) {
    if (blockIdx.x != 0) return;
    uint64_t t = wall_clock64();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) stamps[1 + reqs[i]] = t;
}

// ---------------------------------------------------------------------------
// host side: planning, enqueue, graph capture, autotuning
// ---------------------------------------------------------------------------
namespace {
template <class T> T* carve(char*& cur, size_t count) {
    T* p = (T*)cur;
    cur += (count * sizeof(T) + 255) & ~(size_t)255;
    return p;
}
size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t scratch_bytes(uint32_t cap, uint32_t att_levels, bool ext) {
    size_t b = 0;
    auto add = [&](size_t bytes) { b += align256(bytes); };
    for (int k = 0; k < 3; ++k) add(4ull * cap);  // path req/pix/sample
    add(12ull * cap);                              // ray_o
    add(12ull * cap);                              // ray_d
    for (int k = 0; k < 4; ++k) add(4ull * cap);  // hit prim/t/u/v
    add(4ull * cap);                               // rng
    if (att_levels) add(12ull * cap * att_levels); // attenuation stack
    add(12ull * cap);                              // L
    if (ext) {
        add(12ull * cap);                          // A
        add(4ull * cap);                           // pdf
        for (int k = 0; k < 3; ++k) add(12ull * cap);  // shadow o/d/c
        add(1ull * cap);                           // shadow hit
        add(4ull * cap);                           // shadow queue
    }
    add(4ull * cap);                               // queue 0
    add(4ull * cap);                               // queue 1
    add(4ull * kCntWords);
    return b;
}

bool is_ext(const vr_render_desc& d) { return (d.flags & (VR_RENDER_ENV_IS | VR_RENDER_RUSSIAN_ROULETTE)) != 0; }

int validate(const vr_scene* s, const vr_render_desc& d, uint32_t chunk_paths) {
    if (d.total_samples == 0) return fail(VR_ERR_INVALID, "total_samples must be > 0");
    if (d.max_bounces > kMaxBouncesLimit) return fail(VR_ERR_INVALID, "max_bounces above 64");
    if ((uint64_t)d.width * d.height >= (1ull << 32)) return fail(VR_ERR_INVALID, "image too large");
    if (d.engine != VR_ENGINE_WIDE && d.engine != VR_ENGINE_REFERENCE && d.engine != VR_ENGINE_WATERTIGHT)
        return fail(VR_ERR_INVALID, "bad engine");
    bool ext = is_ext(d);
    if (s->view.num_instances && !ext)
        return fail(VR_ERR_INVALID, "scenes with instances render with the extended estimator only "
                                    "(flags ENV_IS and/or RUSSIAN_ROULETTE): the reference has no instancing");
    if (d.engine == VR_ENGINE_WATERTIGHT && !ext)
        return fail(VR_ERR_INVALID, "the watertight engine renders with the extended estimator only "
                                    "(flags ENV_IS and/or RUSSIAN_ROULETTE): its hits are not the reference's");
    if (ext && d.render_mode != VR_MODE_FULL)
        return fail(VR_ERR_INVALID, "the extended estimator renders RenderMode::Full only");
    if (ext && !(d.firefly_clamp == kInf))
        return fail(VR_ERR_INVALID, "the extended estimator requires firefly_clamp = inf (unbiased)");
    if ((d.flags & VR_RENDER_ENV_IS) && s->view.env_kind == VR_ENV_HDRI && !s->view.env_marg_cdf)
        return fail(VR_ERR_INVALID, "scene was uploaded without environment sampling tables");
    if (d.samples > chunk_paths) return fail(VR_ERR_INVALID, "samples per call above the chunk size");
    return 0;
}

struct Chunk {
    uint32_t seg_off, nseg, npix, npaths;
    uint32_t done_off, ndone;  // requests whose last pixel is rendered by this chunk
};

struct Group {
    vr_scene* scene = nullptr;
    uint32_t engine = 0;
    bool any_ext = false, any_ref = false, any_env_is = false;
    bool any_cached = false;  // radiance-cache requests (queries and training samples)
    uint32_t att_levels = 0;  // max bounces among reference-estimator requests
    uint32_t max_depth = 0;
    uint64_t paths = 0;
    std::vector<uint32_t> reqs;  // batch indices
    std::vector<Chunk> chunks;
};
}  // namespace

// Per-context render state: the planned batch, its device tables and scratch,
// statistics and the captured graph.
struct RenderState {
    std::vector<ReqDev> reqs;
    std::vector<SegDev> segs;
    std::vector<uint32_t> done_ids;
    std::vector<Group> groups;
    uint32_t cap = 0;
    vr_tuning tune{};
    vr_cache* cache = nullptr;  // radiance cache of the batch being planned (cached frames only)
    bool stamps = false;
    // device table: [reqs][segs][done ids][stamps]; staged through pinned host memory
    void* table = nullptr;
    size_t table_bytes = 0;
    void* table_host = nullptr;
    size_t table_host_bytes = 0;
    hipEvent_t table_ev = nullptr;  // the last table upload has been read
    ReqDev* d_reqs = nullptr;
    SegDev* d_segs = nullptr;
    uint32_t* d_done = nullptr;
    uint64_t* d_stamps = nullptr;
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    uint64_t* stats_dev = nullptr;
    uint64_t* stats_host = nullptr;  // pinned copy, valid once done_ev has completed
    uint64_t* stamps_host = nullptr;
    uint32_t stamps_host_n = 0;
    uint32_t last_n = 0;
    hipEvent_t done_ev = nullptr;  // after the read-back of the last batch
    bool done_recorded = false;
    hipGraphExec_t exec = nullptr;
};

static int ensure_state(vr_ctx* ctx) {
    if (!ctx->render) ctx->render = new RenderState();
    RenderState& R = *ctx->render;
    if (!R.stats_dev) {
        VR_HIP_CHECK(hipMalloc((void**)&R.stats_dev, 8 * sizeof(uint64_t)));
        VR_HIP_CHECK(hipHostMalloc((void**)&R.stats_host, 8 * sizeof(uint64_t), hipHostMallocDefault));
        for (int k = 0; k < 8; ++k) R.stats_host[k] = 0;
    }
    if (!R.table_ev) VR_HIP_CHECK(hipEventCreateWithFlags(&R.table_ev, hipEventDisableTiming));
    if (!R.done_ev) VR_HIP_CHECK(hipEventCreateWithFlags(&R.done_ev, hipEventDisableTiming));
    return 0;
}

static void drop_graph(RenderState& R) {
    if (R.exec) (void)hipGraphExecDestroy(R.exec);
    R.exec = nullptr;
}

// Plans the chunks of a batch and allocates everything it needs (never inside
// a graph capture).
static int plan_batch(vr_ctx* ctx, const vr_request* items, uint32_t n, const vr_tuning& tune, uint32_t flags,
                      hipStream_t st, const CacheReq* creq) {
    RenderState& R = *ctx->render;
    uint32_t chunk_paths = tune.chunk_paths ? tune.chunk_paths : kDefaultChunkPaths;
    if (chunk_paths > kMaxChunkPaths) return fail(VR_ERR_INVALID, "chunk_paths above 2^27");
    // 1,024 threads: on the render legs 4% faster than 256 on cornell, 0.6-1.2%
    // on the 1080p frames (fewer blocks, one queue atomic per block), 128 slower
    uint32_t block = tune.block ? tune.block : 1024;
    if (block % 64 != 0 || block > 1024) return fail(VR_ERR_INVALID, "block must be a multiple of 64 up to 1024");
    drop_graph(R);
    R.reqs.clear();
    R.segs.clear();
    R.done_ids.clear();
    R.groups.clear();
    R.cap = 0;
    R.tune = tune;
    R.tune.chunk_paths = chunk_paths;
    R.tune.block = block;
    // group key: scene, engine and, when shortest-first, the request's size
    // class (paths in [4^c, 4^(c+1))), so that small requests on every scene
    // run before large ones instead of waiting for their scene's whole group
    std::map<std::tuple<vr_scene*, uint32_t, uint32_t>, size_t> gidx;
    const bool sjf = (flags & VR_BATCH_SHORTEST_FIRST) != 0;
    for (uint32_t i = 0; i < n; ++i) {
        const vr_request& it = items[i];
        if (!it.scene || !it.d_rgba) return fail(VR_ERR_INVALID, "request without scene or target");
        if (it.scene->ctx != ctx) return fail(VR_ERR_INVALID, "request scene belongs to another context");
        const vr_render_desc& d = it.desc;
        int rc = validate(it.scene, d, chunk_paths);
        if (rc) return rc;
        ReqDev r{};
        r.rgba = it.d_rgba;
        r.seed = d.seed;
        r.W = d.width;
        r.H = d.height;
        r.spp = d.samples;
        r.total = d.total_samples;
        r.sample_offset = d.sample_offset;
        r.max_bounces = d.max_bounces;
        r.flags = d.flags;
        r.rr_depth = d.rr_depth;
        r.request = d.request;
        r.mode = d.render_mode;
        r.clamp = d.firefly_clamp;
        r.ext = is_ext(d) ? 1u : 0u;
        r.npix = d.width * d.height;
        if (creq) {
            r.cmode = creq[i].cmode;
            r.cache_depth = creq[i].cache_depth;
            r.query_roughness = creq[i].query_roughness;
            r.train_vertices = creq[i].train_vertices;
            r.train_threshold = creq[i].train_threshold;
            r.frame = creq[i].frame;
        }
        R.reqs.push_back(r);
        uint64_t rp = (uint64_t)r.npix * d.samples;
        uint32_t size_class = 0;
        if (sjf)
            while (rp >> (2 * (size_class + 1))) ++size_class;
        auto key = std::make_tuple(it.scene, d.engine, size_class + (r.cmode << 16));
        auto f = gidx.find(key);
        if (f == gidx.end()) {
            Group g;
            g.scene = it.scene;
            g.engine = d.engine;
            f = gidx.emplace(key, R.groups.size()).first;
            R.groups.push_back(std::move(g));
        }
        Group& g = R.groups[f->second];
        g.reqs.push_back(i);
        g.any_ext |= r.ext != 0;
        g.any_ref |= r.ext == 0;
        g.any_env_is |= r.ext && (d.flags & VR_RENDER_ENV_IS) && it.scene->view.env_kind == VR_ENV_HDRI;
        if (!r.ext) g.att_levels = std::max(g.att_levels, d.max_bounces);
        g.max_depth = std::max(g.max_depth, d.max_bounces);
        g.paths += rp;
        g.any_cached |= r.cmode == kCacheOn;
    }
    if (!tune.chunk_paths) {
        uint32_t att = 0, max_spp = 1;
        bool ext = false;
        for (const Group& g : R.groups) {
            att = std::max(att, g.att_levels);
            ext |= g.any_ext;
        }
        for (const ReqDev& r : R.reqs) max_spp = std::max(max_spp, r.spp);
        while (chunk_paths > 1u && chunk_paths / 2 >= max_spp && scratch_bytes(chunk_paths, att, ext) > kDefaultScratchBudget)
            chunk_paths /= 2;
        R.tune.chunk_paths = chunk_paths;
    }
    if (flags & VR_BATCH_SHORTEST_FIRST) {
        std::stable_sort(R.groups.begin(), R.groups.end(),
                         [](const Group& a, const Group& b) { return a.paths < b.paths; });
        auto paths = [&](uint32_t i) { return (uint64_t)R.reqs[i].npix * R.reqs[i].spp; };
        for (Group& g : R.groups)
            std::stable_sort(g.reqs.begin(), g.reqs.end(),
                             [&](uint32_t a, uint32_t b) { return paths(a) < paths(b); });
    }
    for (Group& g : R.groups) {
        // cut the group's pixels into chunks of <= chunk_paths paths
        std::vector<uint32_t> last_chunk;
        Chunk cur{(uint32_t)R.segs.size(), 0, 0, 0, 0, 0};
        auto close = [&]() {
            g.chunks.push_back(cur);
            cur = Chunk{(uint32_t)R.segs.size(), 0, 0, 0, 0, 0};
        };
        std::vector<std::pair<uint32_t, uint32_t>> finish;  // (request, chunk index)
        for (uint32_t ri : g.reqs) {
            const ReqDev& r = R.reqs[ri];
            uint64_t npx = r.npix;
            if (r.spp == 0 || npx == 0) {
                finish.emplace_back(ri, UINT32_MAX);  // nothing to render
                continue;
            }
            uint64_t done = 0;
            while (done < npx) {
                uint32_t room_px = (chunk_paths - cur.npaths) / r.spp;
                if (room_px == 0) {
                    close();
                    continue;
                }
                uint32_t take = (uint32_t)std::min<uint64_t>(room_px, npx - done);
                R.segs.push_back(SegDev{ri, (uint32_t)done, take, cur.npaths, cur.npix});
                ++cur.nseg;
                cur.npix += take;
                cur.npaths += take * r.spp;
                done += take;
            }
            finish.emplace_back(ri, (uint32_t)g.chunks.size());
        }
        if (cur.npaths) close();
        // completion lists: requests finishing in chunk c (empty requests with the first chunk)
        for (uint32_t c = 0; c < g.chunks.size(); ++c) {
            g.chunks[c].done_off = (uint32_t)R.done_ids.size();
            for (auto& fr : finish)
                if (fr.second == c || (c == 0 && fr.second == UINT32_MAX)) R.done_ids.push_back(fr.first);
            g.chunks[c].ndone = (uint32_t)R.done_ids.size() - g.chunks[c].done_off;
        }
        if (g.chunks.empty()) {
            // a group of empty requests: stamp them at the group's position
            Chunk c{(uint32_t)R.segs.size(), 0, 0, 0, (uint32_t)R.done_ids.size(), 0};
            for (auto& fr : finish) R.done_ids.push_back(fr.first);
            c.ndone = (uint32_t)R.done_ids.size() - c.done_off;
            g.chunks.push_back(c);
        }
        for (auto& c : g.chunks) R.cap = std::max(R.cap, c.npaths);
    }
    // device table
    size_t b_reqs = align256(R.reqs.size() * sizeof(ReqDev));
    size_t b_segs = align256(R.segs.size() * sizeof(SegDev));
    size_t b_done = align256(R.done_ids.size() * sizeof(uint32_t));
    size_t b_stamps = align256((R.reqs.size() + 1) * sizeof(uint64_t));
    size_t table = b_reqs + b_segs + b_done + b_stamps;
    bool grow_table = R.table_bytes < table;
    size_t need = 0;
    for (auto& g : R.groups) need = std::max(need, scratch_bytes(std::max(R.cap, 1u), g.att_levels, g.any_ext));
    bool grow_scratch = R.scratch_bytes < need;
    if (grow_table || grow_scratch) VR_HIP_CHECK(hipStreamSynchronize(st));  // buffers may be in use
    if (grow_table) {
        if (R.table) VR_HIP_CHECK(hipFree(R.table));
        R.table = nullptr;
        R.table_bytes = 0;
        if (hipMalloc(&R.table, table) != hipSuccess) return fail(VR_ERR_OOM, "request table allocation failed");
        R.table_bytes = table;
    }
    if (grow_scratch) {
        if (R.scratch) VR_HIP_CHECK(hipFree(R.scratch));
        R.scratch = nullptr;
        R.scratch_bytes = 0;
        if (hipMalloc(&R.scratch, need) != hipSuccess) return fail(VR_ERR_OOM, "render scratch allocation failed");
        R.scratch_bytes = need;
    }
    char* tb = (char*)R.table;
    R.d_reqs = (ReqDev*)tb;
    R.d_segs = (SegDev*)(tb + b_reqs);
    R.d_done = (uint32_t*)(tb + b_reqs + b_segs);
    R.d_stamps = (uint64_t*)(tb + b_reqs + b_segs + b_done);
    if (R.stamps_host_n < R.reqs.size() + 1) {
        if (R.stamps_host) VR_HIP_CHECK(hipHostFree(R.stamps_host));
        R.stamps_host = nullptr;
        R.stamps_host_n = 0;
        VR_HIP_CHECK(hipHostMalloc((void**)&R.stamps_host, (R.reqs.size() + 1) * sizeof(uint64_t),
                                   hipHostMallocDefault));
        R.stamps_host_n = (uint32_t)R.reqs.size() + 1;
    }
    for (auto& g : R.groups) {
        int rc = ensure_trace_scratch(g.scene, std::max(R.cap, 1u));
        if (rc) return rc;
    }
    // upload the table through pinned staging; the previous upload must have been read first
    size_t host_bytes = b_reqs + b_segs + b_done;
    if (R.table_host_bytes < host_bytes) {
        VR_HIP_CHECK(hipEventSynchronize(R.table_ev));
        if (R.table_host) VR_HIP_CHECK(hipHostFree(R.table_host));
        R.table_host = nullptr;
        R.table_host_bytes = 0;
        VR_HIP_CHECK(hipHostMalloc(&R.table_host, host_bytes, hipHostMallocDefault));
        R.table_host_bytes = host_bytes;
    }
    VR_HIP_CHECK(hipEventSynchronize(R.table_ev));
    char* hb = (char*)R.table_host;
    if (!R.reqs.empty()) std::memcpy(hb, R.reqs.data(), R.reqs.size() * sizeof(ReqDev));
    if (!R.segs.empty()) std::memcpy(hb + b_reqs, R.segs.data(), R.segs.size() * sizeof(SegDev));
    if (!R.done_ids.empty()) std::memcpy(hb + b_reqs + b_segs, R.done_ids.data(), R.done_ids.size() * 4);
    if (host_bytes) VR_HIP_CHECK(hipMemcpyAsync(R.table, hb, host_bytes, hipMemcpyHostToDevice, st));
    VR_HIP_CHECK(hipEventRecord(R.table_ev, st));
    R.last_n = n;
    return 0;
}

// Enqueues the kernels of the planned batch on `st`: no allocation, no copy,
// no host synchronisation, so the sequence can be captured into a graph.
static int enqueue_batch(vr_ctx* ctx, hipStream_t st) {
    RenderState& R = *ctx->render;
    const uint32_t sb = R.tune.block;
    const uint32_t cap = std::max(R.cap, 1u);
    hipLaunchKernelGGL(k_stats_zero, dim3(1), dim3(64), 0, st, R.stats_dev);
    if (R.stamps) hipLaunchKernelGGL(k_stamp_start, dim3(1), dim3(64), 0, st, R.d_stamps);
    for (const Group& g : R.groups) {
        vr_scene* s = g.scene;
        RenderArgs a{};
        a.sv = s->view;
        a.reqs = R.d_reqs;
        a.cap = cap;
        char* cur = (char*)R.scratch;
        a.path_req = carve<uint32_t>(cur, cap);
        a.path_pix = carve<uint32_t>(cur, cap);
        a.path_smp = carve<uint32_t>(cur, cap);
        a.ray_o = carve<float>(cur, 3ull * cap);
        a.ray_d = carve<float>(cur, 3ull * cap);
        a.hit_prim = carve<uint32_t>(cur, cap);
        a.hit_t = carve<float>(cur, cap);
        a.hit_u = carve<float>(cur, cap);
        a.hit_v = carve<float>(cur, cap);
        a.rng_n = carve<uint32_t>(cur, cap);
        a.att = g.att_levels ? carve<float>(cur, 3ull * cap * g.att_levels) : nullptr;
        a.L = carve<float>(cur, 3ull * cap);
        if (g.any_ext) {
            a.A = carve<float>(cur, 3ull * cap);
            a.pdf = carve<float>(cur, cap);
            a.sh_o = carve<float>(cur, 3ull * cap);
            a.sh_d = carve<float>(cur, 3ull * cap);
            a.sh_c = carve<float>(cur, 3ull * cap);
            a.sh_hit = carve<uint8_t>(cur, cap);
            a.sh_queue = carve<uint32_t>(cur, cap);
        }
        a.queue[0] = carve<uint32_t>(cur, cap);
        a.queue[1] = carve<uint32_t>(cur, cap);
        a.cnt = carve<uint32_t>(cur, kCntWords);
        if (g.any_cached) {
            if (!R.cache) return fail(VR_ERR_INVALID, "radiance-cache request without a cache");
            a.tr = R.cache->tr;
            a.qr = R.cache->qr;
            // training samples accumulate over the group's chunks
            hipLaunchKernelGGL(k_render_zero, dim3(1), dim3(64), 0, st, a.tr.count, 2u);
        }
        for (const Chunk& c : g.chunks) {
            if (c.npaths) {
                a.segs = R.d_segs + c.seg_off;
                a.nseg = c.nseg;
                a.npix = c.npix;
                a.npaths = c.npaths;
                hipLaunchKernelGGL(k_render_zero, dim3((kCntWords + 255) / 256), dim3(256), 0, st, a.cnt,
                                   (uint32_t)kCntWords);
                if (g.any_cached) {
                    hipLaunchKernelGGL(k_cache_mark, dim3(1), dim3(64), 0, st, a.tr.count);
                    hipLaunchKernelGGL(k_render_zero, dim3(1), dim3(64), 0, st, a.qr.count, 1u);
                }
                hipLaunchKernelGGL(k_raygen, dim3((c.npaths + sb - 1) / sb), dim3(sb), 0, st, a);
                for (uint32_t depth = 0; depth < g.max_depth; ++depth) {
                    TraceJob j;
                    j.o = a.ray_o;
                    j.d = a.ray_d;
                    j.n_max = c.npaths;
                    j.queue = depth == 0 ? nullptr : a.queue[depth & 1];
                    j.count = depth == 0 ? nullptr : a.cnt + kCntActive + depth;
                    j.tmin = 0.00001f;  // SceneAcceleration::hit (scene.rs:183)
                    j.any = false;
                    j.engine = (int)g.engine;
                    j.prim = a.hit_prim;
                    j.t = a.hit_t;
                    j.u = a.hit_u;
                    j.v = a.hit_v;
                    j.fallback_accum = R.stats_dev + 3;
                    int rc = launch_trace(s, j, st);
                    if (rc) return rc;
                    // [estimators][cache][microfacet materials]
                    void (*const shades[3][2][2])(RenderArgs, uint32_t) = {
                        {{k_shade<0, false, false>, k_shade<0, false, true>},
                         {k_shade<0, true, false>, k_shade<0, true, true>}},
                        {{k_shade<1, false, false>, k_shade<1, false, true>},
                         {k_shade<0, true, false>, k_shade<0, true, true>}},  // (cache requests are extended)
                        {{k_shade<2, false, false>, k_shade<2, false, true>},
                         {k_shade<2, true, false>, k_shade<2, true, true>}}};
                    const int est = g.any_ext && g.any_ref ? 0 : g.any_ext ? 2 : 1;
                    void (*shade)(RenderArgs, uint32_t) = shades[est][g.any_cached][s->has_microfacet];
                    hipLaunchKernelGGL(shade, dim3(grid_for(c.npaths, sb)), dim3(sb), 0, st, a, depth);
                    if (g.any_env_is) {
                        TraceJob sj;
                        sj.o = a.sh_o;
                        sj.d = a.sh_d;
                        sj.n_max = c.npaths;
                        sj.queue = a.sh_queue;
                        sj.count = a.cnt + kCntShadow + depth;
                        sj.tmin = 0.00001f;
                        sj.tmax = kInf;
                        sj.any = true;
                        sj.engine = (int)g.engine;
                        sj.hit = a.sh_hit;
                        sj.fallback_accum = R.stats_dev + 3;
                        rc = launch_trace(s, sj, st);
                        if (rc) return rc;
                        hipLaunchKernelGGL(k_shadow_resolve, dim3(grid_for(c.npaths, sb)), dim3(sb), 0, st, a,
                                           depth);
                    }
                }
                if (g.any_cached) {
                    int rc = nrc_infer(R.cache, a.qr.cap, a.qr.count, a.qr.pos, a.qr.dir, a.qr.nrm, a.qr.rough,
                                       a.qr.target, st);
                    if (rc) return rc;
                    hipLaunchKernelGGL(k_cache_apply, dim3(grid_for(c.npaths, sb)), dim3(sb), 0, st, a);
                    hipLaunchKernelGGL(k_cache_targets, dim3(grid_for(c.npaths, sb)), dim3(sb), 0, st, a);
                }
                hipLaunchKernelGGL(k_accumulate, dim3((c.npix + sb - 1) / sb), dim3(sb), 0, st, a);
                hipLaunchKernelGGL(k_stats_chunk, dim3(1), dim3(64), 0, st, (const uint32_t*)a.cnt, R.stats_dev,
                                   c.npaths, g.max_depth);
            }
            if (R.stamps && c.ndone)
                hipLaunchKernelGGL(k_stamp_done, dim3(1), dim3(64), 0, st, R.d_stamps,
                                   (const uint32_t*)(R.d_done + c.done_off), c.ndone);
        }
        VR_HIP_CHECK(hipGetLastError());
    }
    return 0;
}

// Copies statistics (and stamps) to pinned host memory after the batch.
static int enqueue_readback(vr_ctx* ctx, hipStream_t st) {
    RenderState& R = *ctx->render;
    VR_HIP_CHECK(hipMemcpyAsync(R.stats_host, R.stats_dev, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    if (R.stamps)
        VR_HIP_CHECK(hipMemcpyAsync(R.stamps_host, R.d_stamps, (R.reqs.size() + 1) * sizeof(uint64_t),
                                    hipMemcpyDeviceToHost, st));
    VR_HIP_CHECK(hipEventRecord(R.done_ev, st));
    R.done_recorded = true;
    return 0;
}

int render_batch(vr_ctx* ctx, const vr_request* items, uint32_t n, const vr_tuning* tuning, uint32_t flags,
                 hipStream_t st, const CacheReq* creq, vr_cache* cache) {
    int rc = ensure_state(ctx);
    if (rc) return rc;
    RenderState& R = *ctx->render;
    vr_tuning tune = tuning ? *tuning : vr_tuning{};
    R.cache = cache;
    rc = plan_batch(ctx, items, n, tune, flags, st, creq);
    if (rc) return rc;
    R.stamps = (flags & VR_BATCH_TIMESTAMPS) != 0;
    if (!(flags & VR_BATCH_GRAPH)) {
        rc = enqueue_batch(ctx, st);
        if (rc) return rc;
        return enqueue_readback(ctx, st);
    }
    // capture the kernels of the whole batch into one graph, then launch it
    hipGraph_t graph = nullptr;
    VR_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    rc = enqueue_batch(ctx, st);
    hipError_t ce = hipStreamEndCapture(st, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ce != hipSuccess) return fail(VR_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    hipError_t ie = hipGraphInstantiate(&R.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        R.exec = nullptr;
        return fail(VR_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    VR_HIP_CHECK(hipGraphLaunch(R.exec, st));
    return enqueue_readback(ctx, st);
}

int relaunch_batch(vr_ctx* ctx, hipStream_t st) {
    if (!ctx->render || !ctx->render->exec) return fail(VR_ERR_INVALID, "no captured batch to relaunch");
    VR_HIP_CHECK(hipGraphLaunch(ctx->render->exec, st));
    return enqueue_readback(ctx, st);
}

int render_device(vr_scene* s, const vr_render_desc* desc, float* d_rgba, hipStream_t st) {
    vr_request r{s, *desc, d_rgba};
    return render_batch(s->ctx, &r, 1, nullptr, 0, st);
}

// One frame of the cached renderer: the frame's paths stop at the query
// vertex and take the cache's answer, except ~train_paths randomly chosen
// training paths that run to the end (and count in the image like any path)
// and record their first vertices as samples; then Adam steps on the samples.
int render_cached(vr_scene* s, vr_cache* c, const vr_render_desc* desc, const vr_cache_frame* f, float* d_rgba,
                  hipStream_t st) {
    vr_ctx* ctx = s->ctx;
    if (c->ctx != ctx) return fail(VR_ERR_INVALID, "cache and scene belong to different contexts");
    if (!is_ext(*desc))
        return fail(VR_ERR_INVALID, "the cached renderer runs the extended estimator (flags ENV_IS and/or RUSSIAN_ROULETTE)");
    if (f->train_vertices == 0 || f->train_vertices > kNrcMaxTrainVertices)
        return fail(VR_ERR_INVALID, "train_vertices must be in [1, 8]");
    if (f->cache_depth >= desc->max_bounces) return fail(VR_ERR_INVALID, "cache_depth must be below max_bounces");
    if (f->train_steps > kNrcMaxParts) return fail(VR_ERR_INVALID, "train_steps must be at most 32");
    const uint64_t npx = (uint64_t)desc->width * desc->height;
    if (npx == 0 || desc->samples == 0) return 0;
    if (npx >= (1ull << 32)) return fail(VR_ERR_INVALID, "image too large");
    const uint64_t paths = npx * desc->samples;
    const uint64_t train_paths = std::min<uint64_t>(f->train_paths, paths);
    // room for twice the expected number of samples; a frame's excess is dropped
    const uint64_t train_cap = std::max<uint64_t>(64, 2 * train_paths * f->train_vertices);
    if (train_cap > (1ull << 28)) return fail(VR_ERR_INVALID, "too many training samples per frame");
    const uint32_t qcap = (uint32_t)std::min<uint64_t>(paths, kDefaultChunkPaths);
    int rc = ensure_state(ctx);
    if (!rc) rc = nrc_ensure_frame_buffers(c, (uint32_t)train_cap, qcap);
    if (rc) return rc;
    CacheReq cr{};
    cr.cmode = kCacheOn;
    cr.cache_depth = f->cache_depth;
    cr.train_vertices = f->train_vertices;
    cr.train_threshold = (uint32_t)std::min<double>(4294967295.0, (double)train_paths / (double)paths * 4294967296.0);
    cr.frame = f->frame;
    cr.query_roughness = f->query_roughness;
    VR_HIP_CHECK(hipEventRecord(c->ev[0], st));
    vr_request rq{s, *desc, d_rgba};
    rc = render_batch(ctx, &rq, 1, nullptr, 0, st, &cr, c);
    if (rc) return rc;
    VR_HIP_CHECK(hipEventRecord(c->ev[1], st));
    // one Adam step per mini-batch: the frame's samples, reordered into
    // train_steps random mini-batches (Morton order inside each), in slices
    if (f->train_steps && train_paths) rc = nrc_order_samples(c, f->train_steps, f->frame, st);
    for (uint32_t k = 0; k < f->train_steps && train_paths && !rc; ++k)
        rc = nrc_train_step(c, c->trs.cap, c->trs.count, c->trs.pos, c->trs.dir, c->trs.nrm, c->trs.rough,
                            c->trs.target, true, st, k, f->train_steps);
    if (rc) return rc;
    VR_HIP_CHECK(hipEventRecord(c->ev[2], st));
    c->frame_timed = true;
    return nrc_frame_readback(c, st);
}

// Statistics of the last render or batch on this context (waits for it).
int render_finish_stats(vr_ctx* ctx, vr_render_stats* stats) {
    double ms = stats->kernel_ms;
    *stats = vr_render_stats{};
    stats->kernel_ms = ms;
    if (!ctx->render || !ctx->render->done_recorded) return 0;
    RenderState& R = *ctx->render;
    VR_HIP_CHECK(hipEventSynchronize(R.done_ev));
    const uint64_t* h = R.stats_host;
    stats->segments = h[0];
    stats->samples = h[1];
    stats->env_misses = h[2];
    stats->fallback_rays = h[3];
    stats->capped_loops = h[4];
    stats->textured_hits = h[5];
    stats->env_is_samples = h[6];
    return 0;
}

int batch_latencies(vr_ctx* ctx, double* ms, uint32_t n) {
    if (!ctx->render || !ctx->render->stamps || !ctx->render->done_recorded)
        return fail(VR_ERR_INVALID, "the last batch was not time-stamped");
    RenderState& R = *ctx->render;
    if (n > R.reqs.size()) return fail(VR_ERR_INVALID, "more latencies requested than the batch has requests");
    VR_HIP_CHECK(hipEventSynchronize(R.done_ev));
    int rate_khz = 0;
    VR_HIP_CHECK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device));
    if (rate_khz <= 0) return fail(VR_ERR_HIP, "wall clock rate unavailable");
    for (uint32_t i = 0; i < n; ++i)
        ms[i] = (double)(int64_t)(R.stamps_host[1 + i] - R.stamps_host[0]) / (double)rate_khz;
    return 0;
}

// Times every candidate tuning on the batch (captured graph, `reps` timed
// relaunches after one warm-up launch) and returns the fastest.
int autotune_batch(vr_ctx* ctx, const vr_request* items, uint32_t n, uint32_t flags, uint32_t reps,
                   hipStream_t st, vr_tuning* best, float* best_ms) {
    static const uint32_t chunks[] = {1u << 22, 1u << 23, 1u << 24, 1u << 25, 1u << 26, 1u << 27};
    static const uint32_t blocks[] = {128, 256, 512, 1024};
    uint32_t max_spp = 1;
    for (uint32_t i = 0; i < n; ++i) max_spp = std::max(max_spp, items[i].desc.samples);
    hipEvent_t e0, e1;
    VR_HIP_CHECK(hipEventCreate(&e0));
    VR_HIP_CHECK(hipEventCreate(&e1));
    float bms = -1.f;
    vr_tuning bt{};
    int rc = 0;
    for (uint32_t c : chunks) {
        if (c < max_spp) continue;
        for (uint32_t b : blocks) {
            vr_tuning t{c, b};
            rc = render_batch(ctx, items, n, &t, (flags | VR_BATCH_GRAPH) & ~VR_BATCH_TIMESTAMPS, st);
            if (rc) break;
            float ms_min = -1.f;
            for (uint32_t k = 0; k < std::max(reps, 1u) && !rc; ++k) {
                if (hipEventRecord(e0, st) != hipSuccess) rc = fail(VR_ERR_HIP, "hipEventRecord failed");
                if (!rc) rc = relaunch_batch(ctx, st);
                if (!rc && hipEventRecord(e1, st) != hipSuccess) rc = fail(VR_ERR_HIP, "hipEventRecord failed");
                float ms = 0.f;
                if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                    rc = fail(VR_ERR_HIP, "event timing failed");
                if (!rc && (ms_min < 0.f || ms < ms_min)) ms_min = ms;
            }
            if (rc) break;
            if (bms < 0.f || ms_min < bms) {
                bms = ms_min;
                bt = t;
            }
        }
        if (rc) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (bms < 0.f) return fail(VR_ERR_INVALID, "no candidate tuning fits the batch");
    *best = bt;
    if (best_ms) *best_ms = bms;
    return 0;
}

void free_render_state(vr_ctx* ctx) {
    RenderState* R = ctx->render;
    if (!R) return;
    drop_graph(*R);
    if (R->table) (void)hipFree(R->table);
    if (R->table_host) (void)hipHostFree(R->table_host);
    if (R->scratch) (void)hipFree(R->scratch);
    if (R->stats_dev) (void)hipFree(R->stats_dev);
    if (R->stats_host) (void)hipHostFree(R->stats_host);
    if (R->stamps_host) (void)hipHostFree(R->stamps_host);
    if (R->table_ev) (void)hipEventDestroy(R->table_ev);
    if (R->done_ev) (void)hipEventDestroy(R->done_ev);
    delete R;
    ctx->render = nullptr;
}

}  // namespace vr
