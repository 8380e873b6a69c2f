// Ray-query kernels: the batch form of SceneAcceleration::hit
// (voidray_renderer/src/core/scene.rs:182-185) over rays resident in HBM.
#include "internal.h"
#include "traverse.h"

namespace vr {

constexpr int kTraceBlock = 128;  // two wave64s
constexpr int kTraceStack = 32;   // LDS stack entries per lane (8 B each)

struct TraceArgs {
    SceneView sv;
    const float* o;
    const float* d;
    uint32_t n;
    float tmin;
    int any;
    uint32_t* prim;
    float* t;
    float* u;
    float* v;
    uint8_t* hit;
    uint32_t* fb_list;
    uint32_t* fb_count;
};

__device__ __forceinline__ void write_out(const TraceArgs& a, uint32_t i, const HitOut& h) {
    if (a.prim) a.prim[i] = h.prim;
    if (a.t) a.t[i] = h.t;
    if (a.u) a.u[i] = h.u;
    if (a.v) a.v[i] = h.v;
    if (a.hit) a.hit[i] = h.prim != 0xFFFFFFFFu ? 1 : 0;
}

template <bool ANY>
__global__ __launch_bounds__(kTraceBlock) void k_trace_wide(TraceArgs a) {
    __shared__ u32x2 stk[kTraceStack * kTraceBlock];
    uint32_t i = blockIdx.x * kTraceBlock + threadIdx.x;
    if (i >= a.n) return;
    vf3 o = mk3(a.o[3 * i], a.o[3 * i + 1], a.o[3 * i + 2]);
    vf3 d = mk3(a.d[3 * i], a.d[3 * i + 1], a.d[3 * i + 2]);
    HitOut h;
    int rc = trace_wide<kTraceStack, ANY>(a.sv, o, d, a.tmin, kInf, stk + threadIdx.x, kTraceBlock, h);
    if (rc != kWideDone) {
        uint32_t slot = atomicAdd(a.fb_count, 1u);
        a.fb_list[slot] = i;
        return;
    }
    write_out(a, i, h);
}

// Exact engine over a ray list (fallbacks) or over all rays (list == nullptr).
__global__ __launch_bounds__(64) void k_trace_exact(TraceArgs a, const uint32_t* list, const uint32_t* count) {
    uint32_t n = list ? *count : a.n;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        uint32_t i = list ? list[j] : j;
        vf3 o = mk3(a.o[3 * i], a.o[3 * i + 1], a.o[3 * i + 2]);
        vf3 d = mk3(a.d[3 * i], a.d[3 * i + 1], a.d[3 * i + 2]);
        HitOut h;
        trace_exact(a.sv, o, d, a.tmin, h);
        write_out(a, i, h);
    }
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

int ensure_trace_scratch(vr_scene* s, uint32_t n) {
    if (!s->counters) {
        VR_HIP_CHECK(hipMalloc((void**)&s->counters, 64 * sizeof(uint32_t)));
        VR_HIP_CHECK(hipHostMalloc((void**)&s->host_counters, 64 * sizeof(uint32_t), hipHostMallocDefault));
        for (int k = 0; k < 64; ++k) s->host_counters[k] = 0;
    }
    if (s->fallback_cap < n) {
        if (s->fallback_list) VR_HIP_CHECK(hipFree(s->fallback_list));
        s->fallback_list = nullptr;
        if (hipMalloc((void**)&s->fallback_list, (size_t)n * sizeof(uint32_t)) != hipSuccess)
            return fail(VR_ERR_OOM, "fallback list allocation failed");
        s->fallback_cap = n;
    }
    return 0;
}

int trace_device(vr_scene* s, const vr_trace_desc* desc, const float* d_o, const float* d_d, uint32_t* d_prim,
                 float* d_t, float* d_u, float* d_v, uint8_t* d_hit, hipStream_t stream) {
    uint32_t n = desc->num_rays;
    if (n == 0) return 0;
    if (!d_o || !d_d) return fail(VR_ERR_INVALID, "null ray buffers");
    if (desc->mode != VR_TRACE_CLOSEST && desc->mode != VR_TRACE_ANY) return fail(VR_ERR_INVALID, "bad trace mode");
    if (desc->engine != VR_ENGINE_WIDE && desc->engine != VR_ENGINE_REFERENCE)
        return fail(VR_ERR_INVALID, "bad engine");
    int rc = ensure_trace_scratch(s, n);
    if (rc) return rc;
    TraceArgs a{s->view, d_o, d_d, n, desc->t_min, desc->mode == VR_TRACE_ANY ? 1 : 0, d_prim, d_t, d_u, d_v, d_hit,
                s->fallback_list, s->counters};
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, s->counters, 64u);
    if (desc->engine == VR_ENGINE_WIDE) {
        dim3 grid((n + kTraceBlock - 1) / kTraceBlock);
        if (a.any)
            hipLaunchKernelGGL(k_trace_wide<true>, grid, dim3(kTraceBlock), 0, stream, a);
        else
            hipLaunchKernelGGL(k_trace_wide<false>, grid, dim3(kTraceBlock), 0, stream, a);
        // fallbacks: persistent grid reads the device-side count (no host sync)
        hipLaunchKernelGGL(k_trace_exact, dim3(1024), dim3(64), 0, stream, a, (const uint32_t*)s->fallback_list,
                           (const uint32_t*)s->counters);
    } else {
        uint32_t blocks = std::min<uint32_t>((n + 63) / 64, 65535u);
        hipLaunchKernelGGL(k_trace_exact, dim3(blocks), dim3(64), 0, stream, a, (const uint32_t*)nullptr,
                           (const uint32_t*)nullptr);
    }
    VR_HIP_CHECK(hipGetLastError());
    VR_HIP_CHECK(hipMemcpyAsync(s->host_counters, s->counters, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    return 0;
}

}  // namespace vr
