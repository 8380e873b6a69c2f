// Wavefront path tracer: the MI355X form of iterative_render
// (voidray_renderer/src/render/iterative.rs:11-55) and trace_ray /
// trace_ray_internal (core/tracer.rs:9-56).
//
// A render call is split into chunks of whole pixels (all their samples), at
// most kChunkPaths paths each.  Per chunk:
//   raygen   one path per (pixel, sample): jitter, camera ray (camera.rs:69-82)
//   for depth = 0 .. max_bounces-1:
//     extend   closest hit of every live path (wide engine + exact fallback)
//     shade    hit record, material scatter, termination; survivors are
//              appended to the next queue with one atomic per wave
//     shadow   (extended estimator) any-hit test of environment samples
//   accumulate  per pixel, samples summed in index order, * 1/total_samples,
//              added into the RGBA f32 target (iterative.rs:45-51)
//
// Reference estimator (parity): the per-level firefly clamp of tracer.rs:50 is
// evaluated exactly as the recursion does it.  Each path keeps its per-level
// attenuations; when it terminates with value X at level m the shade kernel
// folds x = min(a_j * x, clamp) for j = m-1 .. 0, which is the same sequence of
// f32 operations as the recursive return path.
//
// Extended estimator (flags ENV_IS / RUSSIAN_ROULETTE, firefly clamp off):
// forward throughput, environment importance sampling with MIS (power
// heuristic) at Lambertian vertices, Russian roulette from rr_depth.
#include "render.h"
#include "shade.h"

namespace vr {

constexpr uint32_t kChunkPaths = 1u << 22;
constexpr int kShadeBlock = 256;
constexpr uint32_t kShadeGrid = 2048;
constexpr uint32_t kMaxBouncesLimit = 64;

// render counters (u32 words in the scratch counter block)
enum : int {
    kCntActive = 0,                          // [kMaxBouncesLimit + 1] live paths per depth
    kCntShadow = kCntActive + kMaxBouncesLimit + 1,  // [kMaxBouncesLimit] shadow rays per depth
    kCntMiss = kCntShadow + kMaxBouncesLimit,
    kCntCapped,
    kCntWords
};

struct RenderArgs {
    SceneView sv;
    vr_render_desc d;
    uint32_t pix0, npix, npaths, cap;
    bool ext;
    float* ray_o;
    float* ray_d;
    uint32_t* hit_prim;
    float *hit_t, *hit_u, *hit_v;
    uint32_t* rng_n;
    float* att;  // parity: [max_bounces][cap][3]
    float* L;    // [cap][3]
    float* A;    // extended: throughput [cap][3]
    float* pdf;  // extended: pdf of the last diffuse scatter (0 = camera / specular)
    float *sh_o, *sh_d, *sh_c;
    uint8_t* sh_hit;
    uint32_t* queue[2];
    uint32_t* sh_queue;
    uint32_t* cnt;
    float* rgba;
};

__device__ __forceinline__ void path_ids(const RenderArgs& a, uint32_t p, uint32_t& pixel, uint32_t& sample) {
    uint32_t spp = a.d.samples;
    pixel = a.pix0 + p / spp;
    sample = a.d.sample_offset + p % spp;
}

__device__ __forceinline__ vf3 ld3(const float* b, uint32_t p) { return mk3(b[3 * (size_t)p], b[3 * (size_t)p + 1], b[3 * (size_t)p + 2]); }
__device__ __forceinline__ void st3(float* b, uint32_t p, vf3 v) {
    b[3 * (size_t)p] = v.x;
    b[3 * (size_t)p + 1] = v.y;
    b[3 * (size_t)p + 2] = v.z;
}

// Appends `want` lanes to a queue: one atomic per wave.
__device__ __forceinline__ void wave_append(bool want, uint32_t value, uint32_t* queue, uint32_t* counter) {
    uint64_t mask = __ballot(want);
    if (mask == 0) return;
    int lane = threadIdx.x & 63;
    int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (want) queue[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = value;
}

__global__ void k_render_zero(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ __launch_bounds__(kShadeBlock) void k_raygen(RenderArgs a) {
    uint32_t p = blockIdx.x * kShadeBlock + threadIdx.x;
    if (p >= a.npaths) return;
    const SceneView& sv = a.sv;
    uint32_t i, s;
    path_ids(a, p, i, s);
    uint32_t W = a.d.width, H = a.d.height;
    uint32_t x = i % W;
    uint32_t y = (a.d.flags & VR_RENDER_ROWS_BY_HEIGHT) ? i / H : i / W;
    float dd = (float)(W > H ? W : H);
    float xf = ((float)(2u * x + 1u) - (float)W) / dd;
    float yf = ((float)(uint32_t)(2u * (uint32_t)(H - y) - 1u) - (float)H) / dd;
    PathRng rng;
    rng.init(a.d.seed, a.d.request, i, s, 0);
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
    if (a.ext) {
        st3(a.A, p, mk3(1.f, 1.f, 1.f));
        a.pdf[p] = 0.f;
    }
    if (capped) atomicAdd(a.cnt + kCntCapped, capped);
}

// Folds the terminal value x at level m through the stored attenuations
// (the return path of trace_ray_internal, tracer.rs:44-50).
__device__ __forceinline__ vf3 fold_levels(const RenderArgs& a, uint32_t p, uint32_t m, vf3 x) {
    float c = a.d.firefly_clamp;
    for (uint32_t j = m; j-- > 0;) {
        vf3 aj = ld3(a.att + (size_t)j * 3 * a.cap, p);
        x = mk3(0.f, 0.f, 0.f) + vminc(mulv(aj, x), c);
    }
    return x;
}

__global__ __launch_bounds__(kShadeBlock) void k_shade(RenderArgs a, uint32_t depth) {
    const SceneView& sv = a.sv;
    const uint32_t n = depth == 0 ? a.npaths : a.cnt[kCntActive + depth];
    const uint32_t* qin = depth == 0 ? nullptr : a.queue[depth & 1];
    uint32_t* qout = a.queue[(depth + 1) & 1];
    const float c = a.d.firefly_clamp;
    const bool last = depth + 1 >= a.d.max_bounces;
    const uint32_t stride = gridDim.x * kShadeBlock;
    // uniform trip count per wave so that every lane reaches the wave-wide appends
    uint32_t k0 = blockIdx.x * kShadeBlock + (threadIdx.x & ~63u);
    for (uint32_t kb = k0; kb < n; kb += stride) {
        uint32_t k = kb + (threadIdx.x & 63u);
        bool act = k < n;
        uint32_t p = 0;
        bool cont = false, shadow = false, miss = false;
        uint32_t capped = 0;
        if (act) {
            p = qin ? qin[k] : k;
            uint32_t pixel, sample;
            path_ids(a, p, pixel, sample);
            PathRng rng;
            rng.init(a.d.seed, a.d.request, pixel, sample, a.rng_n[p]);
            vf3 o = ld3(a.ray_o, p), d = ld3(a.ray_d, p);
            uint32_t prim = a.hit_prim[p];
            if (prim == 0xFFFFFFFFu) {
                miss = true;
                vf3 env = sv.env_kind == VR_ENV_NONE ? mk3(0.f, 0.f, 0.f) : env_lookup(sv, d);
                if (!a.ext) {
                    st3(a.L, p, fold_levels(a, p, depth, env));
                } else {
                    float w = 1.f;
                    float pb = a.pdf[p];
                    if ((a.d.flags & VR_RENDER_ENV_IS) && pb > 0.f && sv.env_kind == VR_ENV_HDRI) {
                        float pe = env_pdf_dir(sv, d);
                        w = (pb * pb) / (pb * pb + pe * pe);
                    }
                    vf3 L = ld3(a.L, p), A = ld3(a.A, p);
                    st3(a.L, p, L + mulv(A, env) * w);
                }
            } else {
                HitRec hit = make_hit(sv, o, d, prim, a.hit_t[p], a.hit_u[p], a.hit_v[p]);
                GpuMaterial m = sv.materials[sv.surfaces[hit.surface].material];
                if (!a.ext) {
                    if (a.d.render_mode == VR_MODE_NORMAL) {
                        vf3 nn = normalize3(hit.normal);
                        vf3 att = mk3(0.5f * nn.x + 0.5f, 0.5f * nn.y + 0.5f, 0.5f * nn.z + 0.5f);
                        st3(a.L, p, fold_levels(a, p, depth, mk3(0.f, 0.f, 0.f) + vminc(att, c)));
                    } else {
                        ScatterOut sc = scatter(sv, m, d, hit, rng, capped, false);
                        if (!sc.has_ray) {
                            st3(a.L, p, fold_levels(a, p, depth, mk3(0.f, 0.f, 0.f) + vminc(sc.att, c)));
                        } else {
                            st3(a.att + (size_t)depth * 3 * a.cap, p, sc.att);
                            if (last) {
                                st3(a.L, p, fold_levels(a, p, depth + 1, mk3(0.f, 0.f, 0.f)));
                            } else {
                                st3(a.ray_o, p, hit.point);
                                st3(a.ray_d, p, sc.dir);
                                cont = true;
                            }
                        }
                    }
                } else {
                    vf3 A = ld3(a.A, p), L = ld3(a.L, p);
                    if (m.kind == VR_MAT_EMISSION) {
                        st3(a.L, p, L + mulv(A, mk3(m.color[0], m.color[1], m.color[2])));
                    } else {
                        if (m.kind == VR_MAT_LAMBERTIAN && (a.d.flags & VR_RENDER_ENV_IS) &&
                            sv.env_kind == VR_ENV_HDRI) {
                            // next-event estimation towards the environment
                            float u1 = rng.uniform(), u2 = rng.uniform();
                            vf3 ns = m.normal_tex >= 0 ? tex_sample(sv, m.normal_tex, hit.uv.x, hit.uv.y) : hit.normal;
                            ns = normalize3(ns);
                            float pe = 0.f;
                            vf3 wi = env_sample_dir(sv, u1, u2, pe);
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
                        ScatterOut sc = scatter(sv, m, d, hit, rng, capped, true);
                        if (sc.has_ray && !last) {
                            A = mulv(A, sc.att);
                            float pb = 0.f;
                            if (sc.diffuse) pb = fmaxf(0.f, dot3(sc.lobe, sc.dir)) / kPi;
                            bool alive = true;
                            if ((a.d.flags & VR_RENDER_RUSSIAN_ROULETTE) && depth >= a.d.rr_depth) {
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
        wave_append(cont, p, qout, a.cnt + kCntActive + depth + 1);
        wave_append(shadow, p, a.sh_queue, a.cnt + kCntShadow + depth);
        uint64_t mm = __ballot(miss);
        if ((threadIdx.x & 63u) == 0 && mm) atomicAdd(a.cnt + kCntMiss, (uint32_t)__popcll(mm));
        if (capped) atomicAdd(a.cnt + kCntCapped, capped);
    }
}

// Adds one chunk's counters into the 64-bit render statistics
// [segments, samples, env misses, exact-engine rays, capped loops].
__global__ void k_stats_chunk(const uint32_t* cnt, uint64_t* stats, uint32_t npaths, uint32_t max_bounces) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t seg = npaths;
    for (uint32_t k = 1; k < max_bounces; ++k) seg += cnt[kCntActive + k];
    for (uint32_t k = 0; k < max_bounces; ++k) seg += cnt[kCntShadow + k];
    stats[0] += seg;
    stats[1] += npaths;
    stats[2] += cnt[kCntMiss];
    stats[4] += cnt[kCntCapped];
}

__global__ void k_stats_zero(uint64_t* stats) {
    if (threadIdx.x < 8) stats[threadIdx.x] = 0;
}

__global__ __launch_bounds__(kShadeBlock) void k_shadow_resolve(RenderArgs a, uint32_t depth) {
    const uint32_t n = a.cnt[kCntShadow + depth];
    for (uint32_t k = blockIdx.x * kShadeBlock + threadIdx.x; k < n; k += gridDim.x * kShadeBlock) {
        uint32_t p = a.sh_queue[k];
        if (a.sh_hit[p] == 0) st3(a.L, p, ld3(a.L, p) + ld3(a.sh_c, p));
    }
}

// Per pixel: sum of the samples in index order, scaled by 1/total_samples
// and added into the target (iterative.rs:45-51).
__global__ __launch_bounds__(kShadeBlock) void k_accumulate(RenderArgs a) {
    uint32_t px = blockIdx.x * kShadeBlock + threadIdx.x;
    if (px >= a.npix) return;
    uint32_t spp = a.d.samples;
    vf3 color = mk3(0.f, 0.f, 0.f);
    for (uint32_t s = 0; s < spp; ++s) color = color + ld3(a.L, px * spp + s);
    color = color * (1.0f / (float)a.d.total_samples);
    float* t = a.rgba + 4 * (size_t)(a.pix0 + px);
    t[0] += color.x;
    t[1] += color.y;
    t[2] += color.z;
    if (a.d.flags & VR_RENDER_ALPHA_PER_CALL) t[3] += 1.0f;
}

namespace {
template <class T> T* carve(char*& cur, size_t count) {
    T* p = (T*)cur;
    cur += (count * sizeof(T) + 255) & ~(size_t)255;
    return p;
}
size_t scratch_bytes(uint32_t cap, uint32_t max_bounces, bool ext) {
    size_t b = 0;
    auto add = [&](size_t bytes) { b += (bytes + 255) & ~(size_t)255; };
    add(12ull * cap);                       // ray_o
    add(12ull * cap);                       // ray_d
    add(4ull * cap * 4);                    // hit prim/t/u/v
    add(4ull * cap);                        // rng
    add(12ull * cap * (ext ? 0 : max_bounces));  // att stack
    add(12ull * cap);                       // L
    if (ext) {
        add(12ull * cap);                   // A
        add(4ull * cap);                    // pdf
        add(36ull * cap);                   // shadow o/d/c
        add(1ull * cap);                    // shadow hit
        add(4ull * cap);                    // shadow queue
    }
    add(4ull * cap * 2);                    // queues
    add(4ull * kCntWords);
    return b;
}
}  // namespace

int render_device(vr_scene* s, const vr_render_desc* desc, float* d_rgba, vr_render_stats* stats, hipStream_t st) {
    const vr_render_desc& d = *desc;
    if (d.width == 0 || d.height == 0) return 0;
    if (d.samples == 0) return 0;
    if (d.total_samples == 0) return fail(VR_ERR_INVALID, "total_samples must be > 0");
    if (d.max_bounces > kMaxBouncesLimit) return fail(VR_ERR_INVALID, "max_bounces above 64");
    if ((uint64_t)d.width * d.height >= (1ull << 32)) return fail(VR_ERR_INVALID, "image too large");
    if (d.engine != VR_ENGINE_WIDE && d.engine != VR_ENGINE_REFERENCE) return fail(VR_ERR_INVALID, "bad engine");
    bool ext = (d.flags & (VR_RENDER_ENV_IS | VR_RENDER_RUSSIAN_ROULETTE)) != 0;
    if (ext && d.render_mode != VR_MODE_FULL)
        return fail(VR_ERR_INVALID, "the extended estimator renders RenderMode::Full only");
    if (ext && !(d.firefly_clamp == kInf))
        return fail(VR_ERR_INVALID, "the extended estimator requires firefly_clamp = inf (unbiased)");
    if ((d.flags & VR_RENDER_ENV_IS) && s->view.env_kind == VR_ENV_HDRI && !s->view.env_marg_cdf)
        return fail(VR_ERR_INVALID, "scene was uploaded without environment sampling tables");
    if (d.samples > kChunkPaths) return fail(VR_ERR_INVALID, "samples per call above 4194304");
    uint64_t total_px = (uint64_t)d.width * d.height;
    uint32_t chunk_px = (uint32_t)std::min<uint64_t>(total_px, kChunkPaths / d.samples);
    uint32_t cap = chunk_px * d.samples;
    size_t need = scratch_bytes(cap, d.max_bounces, ext);
    if (s->render_scratch_bytes < need) {
        if (s->render_scratch) VR_HIP_CHECK(hipFree(s->render_scratch));
        s->render_scratch = nullptr;
        s->render_scratch_bytes = 0;
        if (hipMalloc(&s->render_scratch, need) != hipSuccess) return fail(VR_ERR_OOM, "render scratch allocation failed");
        s->render_scratch_bytes = need;
    }
    if (!s->render_stats_dev) {
        VR_HIP_CHECK(hipMalloc((void**)&s->render_stats_dev, 8 * sizeof(uint64_t)));
        VR_HIP_CHECK(hipHostMalloc((void**)&s->render_stats_host, 8 * sizeof(uint64_t), hipHostMallocDefault));
        for (int k = 0; k < 8; ++k) s->render_stats_host[k] = 0;
    }
    RenderArgs a{};
    a.sv = s->view;
    a.d = d;
    a.cap = cap;
    a.ext = ext;
    a.rgba = d_rgba;
    char* cur = (char*)s->render_scratch;
    a.ray_o = carve<float>(cur, 3ull * cap);
    a.ray_d = carve<float>(cur, 3ull * cap);
    a.hit_prim = carve<uint32_t>(cur, cap);
    a.hit_t = carve<float>(cur, cap);
    a.hit_u = carve<float>(cur, cap);
    a.hit_v = carve<float>(cur, cap);
    a.rng_n = carve<uint32_t>(cur, cap);
    a.att = ext ? nullptr : carve<float>(cur, 3ull * cap * d.max_bounces);
    a.L = carve<float>(cur, 3ull * cap);
    if (ext) {
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
    hipLaunchKernelGGL(k_stats_zero, dim3(1), dim3(64), 0, st, s->render_stats_dev);
    for (uint64_t pix0 = 0; pix0 < total_px; pix0 += chunk_px) {
        a.pix0 = (uint32_t)pix0;
        a.npix = (uint32_t)std::min<uint64_t>(chunk_px, total_px - pix0);
        a.npaths = a.npix * d.samples;
        hipLaunchKernelGGL(k_render_zero, dim3((kCntWords + 255) / 256), dim3(256), 0, st, a.cnt, (uint32_t)kCntWords);
        hipLaunchKernelGGL(k_raygen, dim3((a.npaths + kShadeBlock - 1) / kShadeBlock), dim3(kShadeBlock), 0, st, a);
        for (uint32_t depth = 0; depth < d.max_bounces; ++depth) {
            TraceJob j;
            j.o = a.ray_o;
            j.d = a.ray_d;
            j.n_max = a.npaths;
            j.queue = depth == 0 ? nullptr : a.queue[depth & 1];
            j.count = depth == 0 ? nullptr : a.cnt + kCntActive + depth;
            j.tmin = 0.00001f;  // SceneAcceleration::hit (scene.rs:183)
            j.any = false;
            j.engine = (int)d.engine;
            j.prim = a.hit_prim;
            j.t = a.hit_t;
            j.u = a.hit_u;
            j.v = a.hit_v;
            j.fallback_accum = s->render_stats_dev + 3;
            int rc = launch_trace(s, j, st);
            if (rc) return rc;
            hipLaunchKernelGGL(k_shade, dim3(kShadeGrid), dim3(kShadeBlock), 0, st, a, depth);
            if (ext && (d.flags & VR_RENDER_ENV_IS)) {
                TraceJob sj;
                sj.o = a.sh_o;
                sj.d = a.sh_d;
                sj.n_max = a.npaths;
                sj.queue = a.sh_queue;
                sj.count = a.cnt + kCntShadow + depth;
                sj.tmin = 0.00001f;
                sj.tmax = kInf;
                sj.any = true;
                sj.engine = (int)d.engine;
                sj.hit = a.sh_hit;
                sj.fallback_accum = s->render_stats_dev + 3;
                rc = launch_trace(s, sj, st);
                if (rc) return rc;
                hipLaunchKernelGGL(k_shadow_resolve, dim3(kShadeGrid), dim3(kShadeBlock), 0, st, a, depth);
            }
        }
        hipLaunchKernelGGL(k_accumulate, dim3((a.npix + kShadeBlock - 1) / kShadeBlock), dim3(kShadeBlock), 0, st, a);
        hipLaunchKernelGGL(k_stats_chunk, dim3(1), dim3(64), 0, st, (const uint32_t*)a.cnt, s->render_stats_dev,
                           a.npaths, d.max_bounces);
        VR_HIP_CHECK(hipGetLastError());
    }
    VR_HIP_CHECK(hipMemcpyAsync(s->render_stats_host, s->render_stats_dev, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    (void)stats;
    return 0;
}

// Statistics of the last render on this scene; valid once its stream has completed.
void render_finish_stats(vr_scene* s, vr_render_stats* stats) {
    const uint64_t* h = s->render_stats_host;
    double ms = stats->kernel_ms;
    *stats = vr_render_stats{};
    stats->kernel_ms = ms;
    if (!h) return;
    stats->segments = h[0];
    stats->samples = h[1];
    stats->env_misses = h[2];
    stats->fallback_rays = h[3];
    stats->capped_loops = h[4];
}

}  // namespace vr
