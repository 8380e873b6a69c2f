// Learned radiance cache: internal declarations shared by nrc.hip (network)
// and render.hip (training-sample recording, cache queries).
#pragma once
#include <vector>

#include "internal.h"

namespace vr {

constexpr int kNrcLevels = VR_CACHE_LEVELS;
constexpr int kNrcIn = 128;   // encoded input width
constexpr int kNrcHid = 64;   // hidden width
constexpr int kNrcOut = 16;   // output rows (3 used)
constexpr int kNrcParams = VR_CACHE_MLP_PARAMS;
// layer l weights start at kNrcOff[l], rows [out][in]
constexpr int kNrcOff0 = 0, kNrcOff1 = 8192, kNrcOff2 = 12288, kNrcOff3 = 16384, kNrcOff4 = 20480;
constexpr uint32_t kNrcMaxTrainVertices = 8;

// Device view of the network, passed by value to kernels.
struct NrcView {
    const __bf16* w_bf;      // forward weights (BF16 copy of the master)
    const float* w_f;        // FP32 master weights
    const _Float16* grid_h;  // forward grid (FP16 copy) [level][entry][2]
    const float* grid_f;     // FP32 master grid
    uint32_t table_mask;     // entries per level - 1
    uint32_t log2_table;
    uint32_t fp32;           // 1: all-FP32 debug path
    uint32_t pad;
    uint32_t res[kNrcLevels];
    uint32_t dense[kNrcLevels];  // (res + 1)^3 <= entries: direct indexing instead of the hash
    float origin[3];
    float inv_extent;
};

// SoA sample arrays (training samples or cache queries).
struct NrcSamples {
    float* pos;    // [cap][3]
    float* dir;    // [cap][3] outgoing direction (towards the previous vertex)
    float* nrm;    // [cap][3] shading normal
    float* rough;  // [cap]
    float* target; // [cap][3] training: path-suffix radiance; queries: cache output
    uint32_t* path;   // path slot in the render scratch
    float* A;      // [cap][3] throughput arriving at the vertex
    float* L;      // [cap][3] training: radiance gathered before the vertex
    uint32_t* count;  // device count
    uint32_t cap;
};

}  // namespace vr

struct vr_cache {
    vr_ctx* ctx = nullptr;
    vr_cache_desc desc{};
    vr_cache_info info{};
    vr::NrcView view{};
    float *w_master = nullptr, *w_m = nullptr, *w_v = nullptr, *w_grad = nullptr;
    __bf16* w_bf = nullptr;
    float *g_master = nullptr, *g_m = nullptr, *g_v = nullptr, *g_grad = nullptr;
    _Float16* g_h = nullptr;
    // training scratch: transposed inputs/activations/deltas, split-K partials
    uint32_t train_cap = 0;
    __bf16 *xT = nullptr, *hT = nullptr, *dT = nullptr;
    float* part = nullptr;
    float* loss = nullptr;       // [0] sum of loss terms, [1] mean loss of the last step
    uint32_t* counts = nullptr;
    // frame buffers: training samples and cache queries
    vr::NrcSamples tr{}, qr{};
    // the frame's training samples reordered for its Adam steps: each sample
    // draws a mini-batch at random, and inside a mini-batch samples follow the
    // Morton order of their positions (pos, dir, nrm, rough, target only)
    vr::NrcSamples trs{};
    uint32_t *sort_keys = nullptr, *sort_vals = nullptr;  // [2][cap] each: input and output of the sort
    void* sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
    std::vector<void*> frame_allocs;
    // pinned read-back: [0] train count, [1] query count, [2] mean loss (f32 bits)
    uint32_t* host_rb = nullptr;
    hipEvent_t done = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};  // frame: start, paths done, training done
    bool done_recorded = false, frame_timed = false;
};

namespace vr {
int nrc_infer(vr_cache* c, uint32_t n, const uint32_t* d_n, const float* pos, const float* dir, const float* nrm,
              const float* rough, float* out, hipStream_t st);
int nrc_frame_readback(vr_cache* c, hipStream_t st);
// Loss gradients into the cache's gradient buffers (zero on entry, zero again
// after an update); update = Adam step.  Mini-batch `part` of `parts` equal
// slices of the samples.
int nrc_train_step(vr_cache* c, uint32_t n, const uint32_t* d_n, const float* pos, const float* dir, const float* nrm,
                   const float* rough, const float* target, bool update, hipStream_t st, uint32_t part = 0,
                   uint32_t parts = 1);
int nrc_ensure_frame_buffers(vr_cache* c, uint32_t train_cap, uint32_t query_cap);
// Fills c->trs from c->tr: sample i goes to mini-batch hash(i, frame) % parts,
// mini-batches in order, Morton order of the position inside each.  Equal
// slices of the result are then independent random mini-batches whose
// neighbouring samples share hash-grid cells.  parts <= kNrcMaxParts.
constexpr uint32_t kNrcMaxParts = 32;
int nrc_order_samples(vr_cache* c, uint32_t parts, uint32_t frame, hipStream_t st);

}  // namespace vr
