// Ray-query kernels: the batch form of SceneAcceleration::hit
// (voidray_renderer/src/core/scene.rs:182-185) over rays resident in HBM.
#include <mutex>
#include <string>

#include "internal.h"
#include "traverse.h"

namespace vr {

constexpr int kWave = 64;        // gfx950 wavefront size; one wave per block
// LDS stack entries per lane (8 B each): 5 KB per wave, so 32 waves per CU
// (160 KB of LDS) fit, as many as the closest-hit kernel's 64 VGPRs allow
// (8 per SIMD).  Measured against 12 entries (6 waves): microbench 0.8%
// faster, terrain frame 2.4% faster, the cornell room (long bounce chains,
// deeper stacks) 5% slower; 14 entries 4% slower.
constexpr int kLdsStack = 10;
constexpr int kLdsStackAny = 12;  // any hit: 6, 8, 10 and 11 entries each measured slower (0.4-3.8%)
constexpr int kSpillCap = 48;    // global spill entries per lane beyond the LDS part
constexpr int kRefill = 40;      // idle lanes that trigger a wave-wide refill (closest hit; 24, 32, 48 and 56 measured slower)
constexpr int kRefillAny = 40;   // the same for any-hit queries (shorter rays: refill less often)
constexpr int kNodeSteps = 4;    // most node-phase iterations between two leaf phases

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
    uint32_t* next;       // partition counters of the dynamic ray fetch (kFetchStride apart)
    u32x2* spill;         // [spill_cap][spill_lanes]
    uint32_t spill_lanes;
    const uint32_t* queue;  // optional: item k traces ray queue[k] (else ray k)
    const uint32_t* count;  // optional: device-side item count (else n)
    float tmax;             // any-hit: hits at t >= tmax do not count
    int spill_cap;          // global spill entries per lane; a deeper stack hands the ray to the exact engine
};

__device__ __forceinline__ void write_out(const TraceArgs& a, uint32_t i, const HitOut& h) {
    if (a.prim) a.prim[i] = h.prim;
    if (a.t) a.t[i] = h.t;
    if (a.u) a.u[i] = h.u;
    if (a.v) a.v[i] = h.v;
    if (a.hit) a.hit[i] = h.prim != 0xFFFFFFFFu ? 1 : 0;
}

// Words of a scene's trace counter block.  The wide kernel appends to the
// fallback list and fetches rays through the partition counters; the
// cooperative exact kernel appends overflows, counts finished blocks, and its
// last block publishes the launch's counts and clears the working words, so
// every launch starts from zero without a reset launch (or a memset node in a
// captured graph).
enum : int {
    kCtrFallbacks = 0,      // rays handed from the wide engine to the exact engine
    kCtrOverflow = 2,       // rays whose frontier outgrew the cooperative LDS stack
    kCtrBlocksDone = 3,     // finished blocks of the cooperative exact kernel
    kCtrLastFallbacks = 8,  // published: fallbacks of the last launch
    kCtrLastOverflow = 9,   // published: overflows of the last launch
    kCtrFetch = 64,         // first ray-fetch partition counter
    kCtrWords = 1024,
};
// Dynamic ray fetch draws from kFetchParts partitions of the ray range, each
// with its own counter on its own 128-byte line: with ~20 traversal steps per
// ray, thousands of waves refilling through one counter serialise on that
// address.  A wave starts on partition blockIdx % kFetchParts and moves on
// when it runs dry.  Workgroups are dealt round-robin to the 8 XCDs, so a
// partition's first waves all run on one XCD (blockIdx % 8) and adjacent rays
// share that XCD's L2: 32 partitions (four per XCD) cost the coherent rays 13%.
constexpr int kFetchParts = 16;  // 8 and 32 measured: equal and 4% slower
constexpr int kFetchStride = 32;  // words between two partition counters
static_assert(kCtrFetch + kFetchParts * kFetchStride <= kCtrWords, "counter block too small");
static_assert(kFetchParts <= 64, "one lane clears each partition counter");

// Persistent wide-engine kernel: every lane runs its own ray as a resumable
// state machine; each iteration is a node phase (up to kNodeSteps inner nodes)
// followed by a leaf phase (one leaf's triangles); whenever kRefill lanes of
// the wave are idle the wave fetches that many new rays with one atomic
// (dynamic ray fetch), so long incoherent rays do not leave the wave idle.
// A refill costs an atomic round trip plus the new rays' loads, and with a
// spatial-split BVH a ray needs only ~20 steps, so refills wait until a
// good part of the wave is idle (40 of 64 lanes, for closest and any hit)
// and draw from partitioned counters (below).
template <bool ANY, bool WT, bool INST>
__device__ __forceinline__ void trace_persist(const TraceArgs& a) {
    constexpr int LDS = ANY ? kLdsStackAny : kLdsStack;
    __shared__ u32x2 lds[LDS * kWave];
    const int lane = threadIdx.x;
    const uint32_t glane = blockIdx.x * kWave + lane;
    const WideStack ws{lds + lane, kWave, a.spill + glane, (int)a.spill_lanes, a.spill_cap};
    const uint32_t kNone = 0xFFFFFFFFu;
    const uint32_t n = a.count ? *a.count : a.n;
    WideState st;
    uint32_t ray = kNone;
    bool exhausted = false;
    uint32_t part = blockIdx.x % kFetchParts;  // wave-uniform
    int dry = 0;                               // partitions found empty so far
    for (;;) {
        bool idle = ray == kNone;
        uint64_t idle_mask = __ballot(idle);
        int nidle = __popcll(idle_mask);
        if (!exhausted && nidle >= (ANY ? kRefillAny : kRefill)) {
            // take nidle rays from the current partition, moving on while it is empty
            uint32_t base = 0, hi = 0;
            for (;;) {
                const uint32_t lo = (uint32_t)((uint64_t)n * part / kFetchParts);
                hi = (uint32_t)((uint64_t)n * (part + 1) / kFetchParts);
                uint32_t got = 0;
                if (lane == 0) got = atomicAdd(a.next + part * kFetchStride, (uint32_t)nidle);
                base = lo + __shfl(got, 0);
                if (base < hi) break;
                if (++dry == kFetchParts) {
                    exhausted = true;
                    break;
                }
                part = part + 1 == (uint32_t)kFetchParts ? 0u : part + 1;
            }
            if (idle && !exhausted) {
                uint32_t k = base + (uint32_t)__popcll(idle_mask & ((1ull << lane) - 1ull));
                if (k < hi) {
                    uint32_t r = a.queue ? a.queue[k] : k;
                    ray = r;
                    vf3 o = mk3(a.o[3 * r], a.o[3 * r + 1], a.o[3 * r + 2]);
                    vf3 d = mk3(a.d[3 * r], a.d[3 * r + 1], a.d[3 * r + 2]);
                    wide_init<WT>(st, o, d, a.sv.num_qnodes > 0);
                    if (a.sv.num_analytic) wide_seed_analytic<ANY, WT>(a.sv, st, a.tmin, a.tmax);
                }
            }
        }
        if (exhausted && __ballot(ray != kNone) == 0ull) break;
        // Node phase (while-while traversal with leaf postponement, Aila &
        // Laine 2009): lanes walk inner nodes together; a lane that reaches a
        // leaf parks it and keeps walking until it meets a second leaf; the
        // phase ends once no lane is still looking for its first leaf.  Node
        // and triangle work then run in separate passes instead of splitting
        // every iteration between the two.
        // The lane's state as one word: 0 walking, 1 done, 2 handed to the
        // exact engine, 3 no ray.  (As three bools they were three lane
        // masks in scalar registers, which spilled: 3% slower.)
        uint32_t status = ray != kNone ? 0u : 3u;
        uint32_t leaf = kNone;
        for (int it = 0; it < kNodeSteps; ++it) {
            const bool searching = status == 0u && leaf == kNone;
            if (__ballot(searching) == 0ull) break;
            if (status != 0u) continue;
            if (!st.have) {
                if (!wide_pop_next<LDS, ANY>(st, ws)) {
                    if (leaf == kNone) status = 1u;  // stack empty, nothing parked: the ray is finished
                    continue;
                }
            }
            if (st.cur & kLeafBit) {
                if (leaf == kNone) {  // park the first leaf, keep walking
                    leaf = st.cur;
                    st.have = false;
                }
                continue;  // a second leaf waits in st.cur
            }
            if (INST && (st.cur & kInstEntry)) {
                // entering or leaving an instance changes the ray: a parked
                // leaf is tested with the current one first (the entry waits)
                if (leaf != kNone) continue;
                st.have = false;
                if (st.cur == kInstExit) {  // back to the world ray, reloaded
                    wide_instance_exit<WT>(st, mk3(a.o[3 * ray], a.o[3 * ray + 1], a.o[3 * ray + 2]),
                                           mk3(a.d[3 * ray], a.d[3 * ray + 1], a.d[3 * ray + 2]));
                } else if (wide_instance_enter<LDS, WT>(a.sv, st, ws, st.cur, a.tmin) == kStepFallback) {
                    status = 2u;
                }
                continue;
            }
            st.have = false;
            if (wide_node<LDS, ANY>(a.sv, st, ws, st.cur, a.tmin, a.tmax) == kStepFallback) status = 2u;
        }
        // Leaf phase
        if (leaf != kNone && status != 2u) {
            const int rc = wide_leaf<LDS, ANY, WT, INST>(a.sv, st, ws, leaf, a.tmin, a.tmax);
            if (rc == kStepDone) status = 1u;
            else if (rc == kStepFallback) status = 2u;
        }
        if (status == 0u && !st.have && st.sp == 0) status = 1u;
        if (status == 1u || status == 2u) {
            HitOut h;
            int rc = status == 2u ? kStepFallback : wide_finish<ANY, WT, INST>(a.sv, st, a.tmin, h);
            if (rc == kStepFallback) {
                uint32_t slot = atomicAdd(a.fb_count, 1u);
                a.fb_list[slot] = ray;
            } else {
                write_out(a, ray, h);
            }
            ray = kNone;
        }
    }
}

// Closest hit: registers held to what 8 waves per SIMD allow (64 VGPRs, no
// scratch; the 5 KB LDS ring allows 8 as well).  Any hit keeps the registers
// it wants (72 VGPRs, 7 waves; its 6 KB ring allows 6.7 per SIMD): held to 64
// it spills.
template <bool ANY, bool WT = false, bool INST = false>
__global__ void k_trace_persist(TraceArgs a);
template <>
__global__ __launch_bounds__(kWave, 8) void k_trace_persist<false, false, false>(TraceArgs a) {
    trace_persist<false, false, false>(a);
}
template <>
__global__ __launch_bounds__(kWave) void k_trace_persist<true, false, false>(TraceArgs a) {
    trace_persist<true, false, false>(a);
}
// watertight engine (extended estimator only)
template <>
__global__ __launch_bounds__(kWave, 8) void k_trace_persist<false, true, false>(TraceArgs a) {
    trace_persist<false, true, false>(a);
}
template <>
__global__ __launch_bounds__(kWave) void k_trace_persist<true, true, false>(TraceArgs a) {
    trace_persist<true, true, false>(a);
}
// scenes with instances (the instance bookkeeping needs a few more
// registers: held to 8 waves per SIMD the closest-hit kernels spill)
template <>
__global__ __launch_bounds__(kWave, 7) void k_trace_persist<false, false, true>(TraceArgs a) {
    trace_persist<false, false, true>(a);
}
template <>
__global__ __launch_bounds__(kWave, 6) void k_trace_persist<true, false, true>(TraceArgs a) {
    trace_persist<true, false, true>(a);
}
template <>
__global__ __launch_bounds__(kWave, 6) void k_trace_persist<false, true, true>(TraceArgs a) {
    trace_persist<false, true, true>(a);
}
template <>
__global__ __launch_bounds__(kWave) void k_trace_persist<true, true, true>(TraceArgs a) {
    trace_persist<true, true, true>(a);
}

__device__ __forceinline__ uint32_t load_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Serial exact engine for one ray, with its stacks in the caller's LDS.
template <bool WT>
__device__ __forceinline__ void exact_serial(const TraceArgs& a, uint32_t i, int32_t* tstack, int32_t* mstack, int stride) {
    vf3 o = mk3(a.o[3 * i], a.o[3 * i + 1], a.o[3 * i + 2]);
    vf3 d = mk3(a.d[3 * i], a.d[3 * i + 1], a.d[3 * i + 2]);
    HitOut h;
    trace_exact_stk<WT>(a.sv, o, d, a.tmin, h, tstack, mstack, stride);
    if (a.any && h.prim != 0xFFFFFFFFu && !(h.t < a.tmax)) h.prim = 0xFFFFFFFFu;
    write_out(a, i, h);
}

// Wave-cooperative exact engine: one wave per listed ray.  The reference
// traversal visits every node whose ancestors all pass the slab test, so the
// candidate set does not depend on visiting order; the wave pops up to 64
// pending nodes at once from an LDS stack, tests them in parallel and pushes
// the children of passing nodes.  The winner is the lexicographic minimum of
// (t, prim) over the wave, identical to trace_exact.  A ray whose frontier
// outgrows the LDS stack is listed for the serial engine, which the last block
// to finish runs over the list, one ray per lane.  `list` == nullptr: every
// ray (or every queued ray) is traced; otherwise the wide engine's fallbacks.
constexpr int kCoopStack = 6144;
static_assert(kCoopStack * 2 >= 2 * 64 * kWave, "the serial stacks reuse the cooperative stack");
enum : uint32_t { kCoopTop = 0u, kCoopMesh = 1u, kCoopTri = 2u };

template <bool WT>
__global__ __launch_bounds__(kWave) void k_trace_exact_wave(TraceArgs a, const uint32_t* list, uint32_t* ctr,
                                                            uint32_t* serial_list, uint64_t* accum) {
    __shared__ u32x2 stk[kCoopStack];
    const int lane = threadIdx.x;
    const SceneView& sv = a.sv;
    // (agent-scope loads: a select between a device count and the by-value
    // argument would otherwise be folded into one load through a pointer to a
    // private copy of the arguments)
    uint32_t n = a.n;
    if (list) n = load_agent(&ctr[kCtrFallbacks]);
    else if (a.count) n = load_agent(a.count);
    // Blocks without a ray leave at once; the others (at least block 0) take
    // part in the exit count.  A block that leaves may read the count after
    // the reset, but then it reads 0 and leaves all the same: the reset only
    // happens once every taking-part block has read the count.
    const uint32_t parts = n < gridDim.x ? (n > 1u ? n : 1u) : gridDim.x;
    if (blockIdx.x >= parts) return;
    for (uint32_t j = blockIdx.x; j < n; j += gridDim.x) {
        uint32_t ray = list ? list[j] : (a.queue ? a.queue[j] : j);
        vf3 o = mk3(a.o[3 * ray], a.o[3 * ray + 1], a.o[3 * ray + 2]);
        vf3 d = mk3(a.d[3 * ray], a.d[3 * ray + 1], a.d[3 * ray + 2]);
        float bt = kInf, bu = 0.f, bv = 0.f;
        uint32_t bprim = 0xFFFFFFFFu;
        WTRay wt{};
        if (WT) wt = wt_prepare(d);
        int sp = sv.top_root_valid ? 1 : 0;  // wave-uniform
        if (lane == 0) stk[0] = u32x2{kCoopTop << 30, 0u};
        bool overflow = false;
        __syncthreads();
        while (sp > 0) {
            int take = sp < kWave ? sp : kWave;
            u32x2 e = u32x2{0u, 0u};
            bool act = lane < take;
            if (act) e = stk[sp - take + lane];
            __syncthreads();
            sp -= take;
            // entries to push: two tree children (e0 = right, e1 = left), one
            // mesh root, or a run of np small-mesh triangles (e0.x + k)
            u32x2 e0 = u32x2{0u, 0u}, e1 = u32x2{0u, 0u};
            int np = 0;
            bool run = false;
            if (act) {
                uint32_t kind = e.x >> 30, idx = e.x & 0x3FFFFFFFu, surf = e.y;
                if (kind == kCoopTop) {
                    RefNode nd = sv.top_nodes[idx];
                    if (nd.obj >= 0) {
                        uint32_t s2 = (uint32_t)nd.obj;
                        const SurfaceInfo& si2 = sv.surfaces[s2];
                        if (si2.analytic >= 0) {  // analytic surface: its hit, in this lane
                            float t, u, v;
                            if (analytic_hit(sv.analytic[si2.analytic], o, d, a.tmin, t, u, v) &&
                                (t < bt || (t == bt && si2.prim_base < bprim))) {
                                bt = t;
                                bu = u;
                                bv = v;
                                bprim = si2.prim_base;
                            }
                        } else {
                            MeshInfo mi = sv.meshes[si2.mesh];
                            if (mi.root >= 0) {
                                e0 = u32x2{(kCoopMesh << 30) | (uint32_t)mi.root, s2};
                                np = 1;
                            } else {
                                e0 = u32x2{kCoopTri << 30, s2};
                                np = (int)(mi.num_tris < 4u ? mi.num_tris : 4u);
                                run = true;
                            }
                        }
                    } else {
                        Box3 b{mk3(nd.lo[0], nd.lo[1], nd.lo[2]), mk3(nd.hi[0], nd.hi[1], nd.hi[2])};
                        if (WT ? wt_slab_hit(b, o, d, a.tmin, kInf) : ref_slab_hit(b, o, d, a.tmin, kInf)) {
                            e0 = u32x2{(kCoopTop << 30) | (uint32_t)nd.right, 0u};
                            e1 = u32x2{(kCoopTop << 30) | (idx + 1), 0u};
                            np = 2;
                        }
                    }
                } else {
                    SurfaceInfo si = sv.surfaces[surf];
                    MeshInfo mi = sv.meshes[si.mesh];
                    // an instance's mesh is walked with the ray in its object space
                    vf3 mo = o, md = d;
                    WTRay mwt = wt;
                    float meps = -1.0f;
                    if (si.instance >= 0) {
                        instance_ray(sv.instances[si.instance], o, d, mo, md);
                        if (WT) mwt = wt_prepare(md);
                        meps = a.tmin * sv.instances[si.instance].det_inv;
                    }
                    int32_t tri = -1;
                    if (kind == kCoopMesh) {
                        RefNode nd = sv.ref_nodes[idx];
                        if (nd.obj >= 0) {
                            tri = nd.obj;
                        } else {
                            Box3 b{mk3(nd.lo[0], nd.lo[1], nd.lo[2]), mk3(nd.hi[0], nd.hi[1], nd.hi[2])};
                            if (WT ? wt_slab_hit(b, mo, md, a.tmin, kInf) : ref_slab_hit(b, mo, md, a.tmin, kInf)) {
                                e0 = u32x2{(kCoopMesh << 30) | (uint32_t)nd.right, surf};
                                e1 = u32x2{(kCoopMesh << 30) | (idx + 1), surf};
                                np = 2;
                            }
                        }
                    } else {
                        tri = (int32_t)idx;
                    }
                    if (tri >= 0) {
                        const uint32_t* ix = sv.mesh_idx + 3 * (size_t)(mi.tri_base + (uint32_t)tri);
                        const float* vp = sv.vpos + 3 * (size_t)mi.vert_base;
                        vf3 p0 = mk3(vp[3 * ix[0]], vp[3 * ix[0] + 1], vp[3 * ix[0] + 2]);
                        vf3 p1 = mk3(vp[3 * ix[1]], vp[3 * ix[1] + 1], vp[3 * ix[1] + 2]);
                        vf3 p2 = mk3(vp[3 * ix[2]], vp[3 * ix[2] + 1], vp[3 * ix[2] + 2]);
                        MTResult r;
                        if (WT ? wt_hit(mwt, mo, p0, p1, p2, a.tmin, r)
                               : mt_hit(mo, md, p0, p1, p2, a.tmin, r, meps)) {
                            const uint32_t prim = si.prim_base + (uint32_t)tri;
                            if (r.t < bt || (r.t == bt && prim < bprim)) {
                                bt = r.t;
                                bu = r.u;
                                bv = r.v;
                                bprim = prim;
                            }
                        }
                    }
                }
            }
            // wave exclusive scan of push counts
            int incl = np;
#pragma unroll
            for (int off = 1; off < kWave; off <<= 1) {
                int y = __shfl_up(incl, off);
                if (lane >= off) incl += y;
            }
            int total = __shfl(incl, kWave - 1);
            int base = sp + incl - np;
            if (sp + total > kCoopStack) {
                overflow = true;
                break;
            }
            for (int k = 0; k < np; ++k)
                stk[base + k] = run ? u32x2{e0.x + (uint32_t)k, e0.y} : (k == 0 ? e0 : e1);
            sp += total;
            __syncthreads();
        }
        __syncthreads();
        if (overflow) {
            if (lane == 0) {
                uint32_t slot = atomicAdd(&ctr[kCtrOverflow], 1u);
                serial_list[slot] = ray;
            }
            continue;
        }
        // wave argmin of (t, prim)
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            float ot = __shfl_xor(bt, off);
            float ou = __shfl_xor(bu, off), ov = __shfl_xor(bv, off);
            uint32_t oprim = __shfl_xor(bprim, off);
            bool take_other = oprim != 0xFFFFFFFFu &&
                              (bprim == 0xFFFFFFFFu || ot < bt || (ot == bt && oprim < bprim));
            if (take_other) {
                bt = ot;
                bu = ou;
                bv = ov;
                bprim = oprim;
            }
        }
        if (a.any && bprim != 0xFFFFFFFFu && !(bt < a.tmax)) bprim = 0xFFFFFFFFu;
        if (lane == 0) {
            HitOut h;
            h.prim = bprim;
            h.t = bprim == 0xFFFFFFFFu ? 0.f : bt;
            h.u = bprim == 0xFFFFFFFFu ? 0.f : bu;
            h.v = bprim == 0xFFFFFFFFu ? 0.f : bv;
            write_out(a, ray, h);
        }
    }
    // count this block out; the last one runs the overflow list and resets
    // (one same-address atomic per taking-part block, not per block)
    __threadfence();
    uint32_t ticket = 0;
    if (lane == 0) ticket = atomicAdd(&ctr[kCtrBlocksDone], 1u);
    ticket = __shfl(ticket, 0);
    if (ticket != parts - 1) return;
    __threadfence();
    const uint32_t ns = load_agent(&ctr[kCtrOverflow]);
    if (ns) {  // the cooperative stack is free now: 64 + 64 stack entries per lane
        int32_t* lds32 = (int32_t*)stk;
        for (uint32_t k = lane; k < ns; k += kWave)
            exact_serial<WT>(a, load_agent(&serial_list[k]), lds32 + lane, lds32 + 64 * kWave + lane, kWave);
    }
    if (lane == 0) {
        uint32_t fb = list ? load_agent(&ctr[kCtrFallbacks]) : 0u;
        if (accum) *accum += fb;
        ctr[kCtrLastFallbacks] = fb;
        ctr[kCtrLastOverflow] = ns;
        ctr[kCtrFallbacks] = 0u;
        ctr[kCtrOverflow] = 0u;
        ctr[kCtrBlocksDone] = 0u;
    }
    if (lane < kFetchParts) ctr[kCtrFetch + lane * kFetchStride] = 0u;
}

// The persistent kernel for a query: [any + 2 * watertight + 4 * instances].
static void (*const kPersistKernels[8])(TraceArgs) = {
    k_trace_persist<false, false, false>, k_trace_persist<true, false, false>, k_trace_persist<false, true, false>,
    k_trace_persist<true, true, false>,   k_trace_persist<false, false, true>, k_trace_persist<true, false, true>,
    k_trace_persist<false, true, true>,   k_trace_persist<true, true, true>};

static int persist_grid(vr_scene* s, bool any, bool wt, uint32_t& blocks) {
    const bool inst = s->view.num_instances > 0;
    const int k = (any ? 1 : 0) + (wt ? 2 : 0) + (inst ? 4 : 0);
    if (s->persist_blocks[k] == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        VR_HIP_CHECK(hipGetDevice(&dev));
        VR_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        void (*const* kern)(TraceArgs) = kPersistKernels;
        VR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern[k], kWave, 0));
        s->persist_blocks[k] = (uint32_t)std::max(1, cus * std::max(1, per_cu));
    }
    blocks = s->persist_blocks[k];
    return 0;
}

int ensure_trace_scratch(vr_scene* s, uint32_t n) {
    if (!s->spill) {
        uint32_t b[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; ++k) {
            int rc = persist_grid(s, k & 1, k & 2, b[k]);
            if (rc) return rc;
        }
        s->spill_lanes = std::max(std::max(b[0], b[1]), std::max(b[2], b[3])) * kWave;
        size_t bytes = (size_t)kSpillCap * s->spill_lanes * sizeof(u32x2);
        if (hipMalloc((void**)&s->spill, bytes) != hipSuccess) return fail(VR_ERR_OOM, "spill stack allocation failed");
    }
    if (!s->counters) {
        VR_HIP_CHECK(hipMalloc((void**)&s->counters, kCtrWords * sizeof(uint32_t)));
        VR_HIP_CHECK(hipHostMalloc((void**)&s->host_counters, 64 * sizeof(uint32_t), hipHostMallocDefault));
        for (int k = 0; k < 64; ++k) s->host_counters[k] = 0;
        // starts at zero; from then on every launch leaves it at zero
        std::vector<uint32_t> zeros(kCtrWords, 0u);
        VR_HIP_CHECK(hipMemcpy(s->counters, zeros.data(), kCtrWords * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    // two lists of n entries: wide-engine fallbacks, then the serial-exact overflow list
    if (s->fallback_cap < 2 * n) {
        if (s->fallback_list) VR_HIP_CHECK(hipFree(s->fallback_list));
        s->fallback_list = nullptr;
        if (hipMalloc((void**)&s->fallback_list, 2 * (size_t)n * sizeof(uint32_t)) != hipSuccess)
            return fail(VR_ERR_OOM, "fallback list allocation failed");
        s->fallback_cap = 2 * n;
    }
    return 0;
}

// Device scratch of one trace launch: the counter block (kCtrWords), the
// wide engine's fallback list and the exact engine's overflow list (one entry
// per ray each), and the per-lane spill stacks of the persistent kernels.
struct TraceScratch {
    uint32_t* counters;
    uint32_t* fb_list;
    uint32_t* serial_list;
    u32x2* spill;
    uint32_t spill_lanes;   // lanes the spill stacks cover (the grid may use fewer)
    int spill_cap;
};

static int launch_trace_with(vr_scene* s, const TraceJob& j, const TraceScratch& sc, hipStream_t stream) {
    TraceArgs a{s->view, j.o, j.d, j.n_max, j.tmin, j.any ? 1 : 0, j.prim, j.t, j.u, j.v, j.hit,
                sc.fb_list, sc.counters, sc.counters + kCtrFetch, sc.spill, sc.spill_lanes, j.queue, j.count, j.tmax,
                sc.spill_cap};
    const bool wt = j.engine == VR_ENGINE_WATERTIGHT;
    uint64_t* accum = j.engine != VR_ENGINE_REFERENCE ? j.fallback_accum : nullptr;
    if (j.engine != VR_ENGINE_REFERENCE) {
        uint32_t blocks = 0;
        int rc = persist_grid(s, j.any, wt, blocks);
        if (rc) return rc;
        blocks = std::min<uint32_t>(blocks, sc.spill_lanes / kWave);
        blocks = std::min<uint32_t>(blocks, (j.n_max + kWave - 1) / kWave);
        void (*persist)(TraceArgs) =
            kPersistKernels[(j.any ? 1 : 0) + (wt ? 2 : 0) + (s->view.num_instances > 0 ? 4 : 0)];
        void (*exact)(TraceArgs, const uint32_t*, uint32_t*, uint32_t*, uint64_t*) =
            wt ? k_trace_exact_wave<true> : k_trace_exact_wave<false>;
        if (j.timed) VR_HIP_CHECK(hipEventRecord(s->trace_ev[0], stream));
        hipLaunchKernelGGL(persist, dim3(blocks), dim3(kWave), 0, stream, a);
        if (j.timed) VR_HIP_CHECK(hipEventRecord(s->trace_ev[1], stream));
        // fallbacks: one wave per ray; the grid reads the device-side count (no host sync)
        hipLaunchKernelGGL(exact, dim3(1024), dim3(kWave), 0, stream, a, (const uint32_t*)sc.fb_list, sc.counters,
                           sc.serial_list, accum);
    } else {
        uint32_t blocks = std::min<uint32_t>(j.n_max, 65535u);
        void (*exact)(TraceArgs, const uint32_t*, uint32_t*, uint32_t*, uint64_t*) = k_trace_exact_wave<false>;
        hipLaunchKernelGGL(exact, dim3(blocks), dim3(kWave), 0, stream, a, (const uint32_t*)nullptr, sc.counters,
                           sc.serial_list, accum);
    }
    VR_HIP_CHECK(hipGetLastError());
    return 0;
}

int launch_trace(vr_scene* s, const TraceJob& j, hipStream_t stream) {
    if (j.n_max == 0) return 0;
    int rc = ensure_trace_scratch(s, j.n_max);
    if (rc) return rc;
    TraceScratch sc{s->counters, s->fallback_list, s->fallback_list + s->fallback_cap / 2, s->spill, s->spill_lanes,
                    kSpillCap};
    return launch_trace_with(s, j, sc, stream);
}

static int check_trace_desc(const vr_trace_desc* desc, const float* d_o, const float* d_d) {
    if (!d_o || !d_d) return fail(VR_ERR_INVALID, "null ray buffers");
    const int32_t m = desc->mode & ~VR_TRACE_TIMING;
    if (m != VR_TRACE_CLOSEST && m != VR_TRACE_ANY) return fail(VR_ERR_INVALID, "bad trace mode");
    if (desc->engine != VR_ENGINE_WIDE && desc->engine != VR_ENGINE_REFERENCE && desc->engine != VR_ENGINE_WATERTIGHT)
        return fail(VR_ERR_INVALID, "bad engine");
    return 0;
}

int trace_device(vr_scene* s, const vr_trace_desc* desc, const float* d_o, const float* d_d, uint32_t* d_prim,
                 float* d_t, float* d_u, float* d_v, uint8_t* d_hit, hipStream_t stream) {
    uint32_t n = desc->num_rays;
    if (n == 0) return 0;
    int rc = check_trace_desc(desc, d_o, d_d);
    if (rc) return rc;
    TraceJob j;
    j.o = d_o;
    j.d = d_d;
    j.n_max = n;
    j.tmin = desc->t_min;
    j.tmax = kInf;
    j.any = (desc->mode & ~VR_TRACE_TIMING) == VR_TRACE_ANY;
    j.engine = desc->engine;
    j.prim = d_prim;
    j.t = d_t;
    j.u = d_u;
    j.v = d_v;
    j.hit = d_hit;
    j.timed = (desc->mode & VR_TRACE_TIMING) && desc->engine != VR_ENGINE_REFERENCE;
    if (j.timed && !s->trace_ev[0]) {
        VR_HIP_CHECK(hipEventCreate(&s->trace_ev[0]));
        VR_HIP_CHECK(hipEventCreate(&s->trace_ev[1]));
    }
    s->trace_timed = j.timed;
    rc = launch_trace(s, j, stream);
    if (rc) return rc;
    VR_HIP_CHECK(hipMemcpyAsync(s->host_counters, s->counters + kCtrLastFallbacks, 2 * sizeof(uint32_t),
                                hipMemcpyDeviceToHost, stream));
    return 0;
}

// ---- ray queries without allocation --------------------------------------
// vr_trace_device_noalloc_async keeps all of its device scratch in one
// file-scope array per GPU, laid out as
//   [counter block: kCtrWords u32][fallback list: kNoAllocMaxRays u32]
//   [overflow list: kNoAllocMaxRays u32][spill stacks: kNoAllocSpill x kNoAllocLanes x 8 B]
// The persistent grid is capped at kNoAllocMaxBlocks one-wave blocks (32 per
// CU on 256 CUs: every resident wave), and a lane's traversal stack may spill
// kNoAllocSpill entries past its LDS ring; a deeper stack sends the ray to the
// exact engine, which gives the same answer.
constexpr uint32_t kNoAllocMaxRays = VR_NOALLOC_MAX_RAYS;
constexpr uint32_t kNoAllocMaxBlocks = 8192;
constexpr uint32_t kNoAllocLanes = kNoAllocMaxBlocks * kWave;
constexpr int kNoAllocSpill = 16;
constexpr size_t kNoAllocCtrBytes = (size_t)kCtrWords * sizeof(uint32_t);
constexpr size_t kNoAllocListBytes = (size_t)kNoAllocMaxRays * sizeof(uint32_t);
constexpr size_t kNoAllocSpillBytes = (size_t)kNoAllocSpill * kNoAllocLanes * sizeof(u32x2);
constexpr size_t kNoAllocBytes = kNoAllocCtrBytes + 2 * kNoAllocListBytes + kNoAllocSpillBytes;
static_assert(kNoAllocBytes == VR_NOALLOC_SCRATCH_BYTES, "scratch size differs from the header's");
static_assert(kNoAllocCtrBytes % 256 == 0 && kNoAllocListBytes % 256 == 0, "scratch parts stay 256-byte aligned");

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

__global__ void k_clear_words(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// Device address of `scratch` on `device`, looked up once per device (with
// that device current) and kept: in host code the array's name is only a
// host-side stand-in, never a device address.
int noalloc_scratch(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return fail(VR_ERR_INVALID, "device index out of range for the trace scratch");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        VR_HIP_CHECK(hipGetDevice(&cur));
        if (cur != device) VR_HIP_CHECK(hipSetDevice(device));
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e == hipSuccess && p) {
            // the counter block must start at zero; every launch leaves it so
            hipLaunchKernelGGL(k_clear_words, dim3(4), dim3(256), 0, nullptr, (uint32_t*)p, (uint32_t)kCtrWords);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipDeviceSynchronize();
        }
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return fail(VR_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return 0;
}

int trace_device_noalloc(vr_scene* s, const vr_trace_desc* desc, const float* d_o, const float* d_d,
                         uint32_t* d_prim, float* d_t, float* d_u, float* d_v, uint8_t* d_hit, hipStream_t stream) {
    const uint32_t n = desc->num_rays;
    if (n > kNoAllocMaxRays)
        return fail(VR_ERR_INVALID, "vr_trace_device_noalloc_async traces at most VR_NOALLOC_MAX_RAYS (1,048,576) "
                                    "rays per call; split the batch");
    if (n == 0) return 0;
    int rc = check_trace_desc(desc, d_o, d_d);
    if (rc) return rc;
    void* base = nullptr;
    rc = noalloc_scratch(s->ctx->device, &base);
    if (rc) return rc;
    uint8_t* b = (uint8_t*)base;
    TraceScratch sc{(uint32_t*)b, (uint32_t*)(b + kNoAllocCtrBytes),
                    (uint32_t*)(b + kNoAllocCtrBytes + kNoAllocListBytes),
                    (u32x2*)(b + kNoAllocCtrBytes + 2 * kNoAllocListBytes), kNoAllocLanes, kNoAllocSpill};
    // whatever an earlier call on this GPU left, this call starts from zero
    hipLaunchKernelGGL(k_clear_words, dim3(4), dim3(256), 0, stream, sc.counters, (uint32_t)kCtrWords);
    TraceJob j;
    j.o = d_o;
    j.d = d_d;
    j.n_max = n;
    j.tmin = desc->t_min;
    j.tmax = kInf;
    j.any = (desc->mode & ~VR_TRACE_TIMING) == VR_TRACE_ANY;
    j.engine = desc->engine;
    j.prim = d_prim;
    j.t = d_t;
    j.u = d_u;
    j.v = d_v;
    j.hit = d_hit;
    j.timed = false;
    return launch_trace_with(s, j, sc, stream);
}

}  // namespace vr
