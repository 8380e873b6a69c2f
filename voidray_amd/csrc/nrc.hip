// Learned radiance cache: multiresolution hash-grid encoding + 5-layer MLP,
// trained online.  The MLP is the one dense contraction of the path tracer and
// runs on MFMA: v_mfma_f32_16x16x32_bf16 for 32-deep products, 16x16x16 for the
// 16-wide output layer's backward, FP32 accumulation.
//
// Operand layouts (checked on gfx950 with an in-bounds probe): for
// 16x16x32, lane l holds A[row l%16][k 8*(l/16) .. +7] and B[k 8*(l/16) .. +7]
// [col l%16]; the 16x16 result lands as D[4*(l/16) + i][l%16], i = 0..3.
// Layers are computed transposed, D = W . X (units x samples), so a layer's
// result in lane l is units 4*(l/16)+i of sample l%16 of each 16-unit tile.
// Feeding that straight back as the next B operand only needs the K order of
// a 32-deep step to be {tile 2s: 4q+0..3, tile 2s+1: 4q+0..3}; the weight (A)
// loads use the same permuted order, so activations never leave registers.
//
// Encoding (Mueller et al. 2022 hash grid, one-blob of Mueller et al. 2019):
//   [0, 32)   16 levels x 2 features, trilinear over the level's cell; level
//             resolution res_l = floor(base * b^l), b = (max/base)^(1/15);
//             dense indexing when (res+1)^3 <= entries, else the spatial hash
//             x ^ y*2654435761 ^ z*805459861 mod entries
//   [32, 64)  outgoing direction: one-blob(acos(d.y)/pi), one-blob(atan2(d.z, d.x)/2pi + 1/2), 16 bins
//   [64, 96)  shading normal, same
//   [96, 104) roughness, 8 bins; one-blob bin b of v: exp(-(v*B - (b + 1/2))^2 / 2)
//   [104]     1 (the only bias term); [105, 128) zero
// Network: x -> relu(W0 x) -> relu(W1 .) -> relu(W2 .) -> relu(W3 .) -> W4 . (3 of 16 rows).
// Loss: mean over samples and channels of (y - t)^2 / (y^2 + 0.01), the
// denominator held constant (relative L2 of Mueller et al. 2021).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>

#include "nrc.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));

namespace vr {
namespace {

constexpr float kInvPi = 0.31830988618379067f;
constexpr float kInv2Pi = 0.15915494309189535f;
constexpr int kWgradBlocks = 64;
constexpr int kFwdBlock = 256;

// LDS copy of the weights with padded rows (bank-conflict-free 16-byte reads)
constexpr int kS0 = 136, kSH = 72;  // row strides (bf16) of layer 0 and of the 64-wide layers
constexpr int kL0 = 0, kL1 = 64 * kS0, kL2 = kL1 + 64 * kSH, kL3 = kL2 + 64 * kSH, kL4 = kL3 + 64 * kSH;
constexpr int kLdsW = kL4 + 16 * kSH;

__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.f), 1.f); }

struct EncIn {
    float p[3];
    float dth, dph, nth, nph, rough;
};

__device__ __forceinline__ EncIn enc_prepare(const NrcView& v, const float* pos, const float* dir, const float* nrm,
                                             const float* rough, uint32_t s) {
    EncIn e;
    for (int a = 0; a < 3; ++a) e.p[a] = clamp01((pos[3 * (size_t)s + a] - v.origin[a]) * v.inv_extent);
    const float* d = dir + 3 * (size_t)s;
    const float* n = nrm + 3 * (size_t)s;
    e.dth = acosf(fminf(fmaxf(d[1], -1.f), 1.f)) * kInvPi;
    e.dph = atan2f(d[2], d[0]) * kInv2Pi + 0.5f;
    e.nth = acosf(fminf(fmaxf(n[1], -1.f), 1.f)) * kInvPi;
    e.nph = atan2f(n[2], n[0]) * kInv2Pi + 0.5f;
    e.rough = clamp01(rough[s]);
    return e;
}

__device__ __forceinline__ float blob(float v, int bin, float nb) {
    float d = v * nb - ((float)bin + 0.5f);
    return expf(-0.5f * (d * d));
}

// Input feature k in [32, 128).
__device__ __forceinline__ float blob_feature(const EncIn& e, int k) {
    if (k < 48) return blob(e.dth, k - 32, 16.f);
    if (k < 64) return blob(e.dph, k - 48, 16.f);
    if (k < 80) return blob(e.nth, k - 64, 16.f);
    if (k < 96) return blob(e.nph, k - 80, 16.f);
    if (k < 104) return blob(e.rough, k - 96, 8.f);
    return k == 104 ? 1.f : 0.f;
}

// The 8 grid entries (level base included) and trilinear weights of level l.
__device__ __forceinline__ void level_corners(const NrcView& v, int l, const float p[3], uint32_t idx[8], float w[8]) {
    const uint32_t r = v.res[l];
    const float fr = (float)r;
    uint32_t c[3];
    float f[3];
    for (int a = 0; a < 3; ++a) {
        float g = p[a] * fr;
        uint32_t ci = (uint32_t)floorf(g);
        if (ci > r - 1) ci = r - 1;
        c[a] = ci;
        f[a] = g - (float)ci;
    }
    const uint32_t base = (uint32_t)l << v.log2_table;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        uint32_t x = c[0] + (k & 1), y = c[1] + ((k >> 1) & 1), z = c[2] + ((k >> 2) & 1);
        uint32_t i;
        if (v.dense[l]) i = x + (r + 1) * (y + (r + 1) * z);
        else i = (x ^ (y * 2654435761u) ^ (z * 805459861u)) & v.table_mask;
        idx[k] = base + i;
        float wx = (k & 1) ? f[0] : 1.f - f[0];
        float wy = (k & 2) ? f[1] : 1.f - f[1];
        float wz = (k & 4) ? f[2] : 1.f - f[2];
        w[k] = (wx * wy) * wz;
    }
}

__device__ __forceinline__ void level_feature(const NrcView& v, int l, const float p[3], float& f0, float& f1) {
    uint32_t idx[8];
    float w[8];
    level_corners(v, l, p, idx, w);
    f0 = 0.f;
    f1 = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        float g0, g1;
        if (v.fp32) {
            g0 = v.grid_f[2 * (size_t)idx[k]];
            g1 = v.grid_f[2 * (size_t)idx[k] + 1];
        } else {
            uint32_t bits = ((const uint32_t*)v.grid_h)[idx[k]];
            g0 = (float)__builtin_bit_cast(_Float16, (uint16_t)(bits & 0xFFFFu));
            g1 = (float)__builtin_bit_cast(_Float16, (uint16_t)(bits >> 16));
        }
        f0 += w[k] * g0;
        f1 += w[k] * g1;
    }
}

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma16(s16x4 a, s16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ short bfbits(__bf16 x) { return __builtin_bit_cast(short, x); }
__device__ __forceinline__ bf16x8 cat8(bf16x4 a, bf16x4 b) {
    bf16x8 r;
    r[0] = a[0]; r[1] = a[1]; r[2] = a[2]; r[3] = a[3];
    r[4] = b[0]; r[5] = b[1]; r[6] = b[2]; r[7] = b[3];
    return r;
}

// One 64-wide layer forward: acc[t] = units 16t + 4q + i of this lane's sample.
__device__ __forceinline__ void fwd_hidden(const __bf16* Wl, int stride, int tiles, const bf16x4 hin[4], f32x4* acc,
                                           int col, int q) {
    for (int t = 0; t < tiles; ++t) {
        acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const __bf16* row = Wl + (16 * t + col) * stride + 32 * s + 4 * q;
            bf16x8 a = cat8(*(const bf16x4*)row, *(const bf16x4*)(row + 16));
            acc[t] = mfma32(a, cat8(hin[2 * s], hin[2 * s + 1]), acc[t]);
        }
    }
}

// Backward through a 64-wide layer: dh[t] = (W^T delta) for input units 16t + 4q + i.
__device__ __forceinline__ void bwd_hidden(const __bf16* Wl, int stride, int tiles, const bf16x4 delta[4],
                                           f32x4* dh, int col, int q) {
    for (int t = 0; t < tiles; ++t) {
        dh[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            bf16x8 a;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                a[j] = Wl[(32 * s + 4 * q + j) * stride + 16 * t + col];
                a[4 + j] = Wl[(32 * s + 16 + 4 * q + j) * stride + 16 * t + col];
            }
            dh[t] = mfma32(a, cat8(delta[2 * s], delta[2 * s + 1]), dh[t]);
        }
    }
}

__device__ __forceinline__ void relu_pack(const f32x4* acc, bf16x4* h) {
    for (int t = 0; t < 4; ++t)
        for (int i = 0; i < 4; ++i) h[t][i] = (__bf16)fmaxf(acc[t][i], 0.f);
}

__device__ __forceinline__ void mask_pack(const f32x4* dh, const bf16x4* h, bf16x4* d) {
    for (int t = 0; t < 4; ++t)
        for (int i = 0; i < 4; ++i) d[t][i] = (__bf16)((float)h[t][i] > 0.f ? dh[t][i] : 0.f);
}

// Transposed store of a 64-unit register tile set: rows row0 + 16t + 4q + i, column s.
__device__ __forceinline__ void store_rows(__bf16* base, size_t cs, uint32_t s, int row0, const bf16x4* h, int q) {
    for (int t = 0; t < 4; ++t)
        for (int i = 0; i < 4; ++i) base[(size_t)(row0 + 16 * t + 4 * q + i) * cs + s] = h[t][i];
}

struct MlpArgs {
    NrcView v;
    uint32_t n;
    const uint32_t* d_n;
    const float *pos, *dir, *nrm, *rough;
    float* out;           // inference: [n][3]
    const float* target;  // training: [n][3]
    float* loss;          // training: += sum of loss terms
    __bf16 *xT, *hT, *dT; // training: transposed [rows][cap]
    uint32_t cap;
    float* grid_grad;
    uint32_t part, parts;  // training mini-batch: slice part of parts of the samples
};

// Mini-batch `part` of `parts`: samples [begin, begin + count) of total.
__device__ __forceinline__ void slice_of(uint32_t total, uint32_t part, uint32_t parts, uint32_t& begin,
                                         uint32_t& count) {
    const uint64_t b = (uint64_t)total * part / parts, e = (uint64_t)total * (part + 1) / parts;
    begin = (uint32_t)b;
    count = (uint32_t)(e - b);
}

// Copies the weights into LDS with padded rows.
__device__ __forceinline__ void load_weights(const __bf16* src, __bf16* W) {
    for (int c = threadIdx.x; c < kNrcParams / 8; c += blockDim.x) {
        int e = c * 8;
        int dst;
        if (e < kNrcOff1) dst = kL0 + (e / 128) * kS0 + (e % 128);
        else if (e < kNrcOff2) dst = kL1 + ((e - kNrcOff1) / 64) * kSH + ((e - kNrcOff1) % 64);
        else if (e < kNrcOff3) dst = kL2 + ((e - kNrcOff2) / 64) * kSH + ((e - kNrcOff2) % 64);
        else if (e < kNrcOff4) dst = kL3 + ((e - kNrcOff3) / 64) * kSH + ((e - kNrcOff3) % 64);
        else dst = kL4 + ((e - kNrcOff4) / 64) * kSH + ((e - kNrcOff4) % 64);
        *(u32x4*)(W + dst) = ((const u32x4*)src)[c];
    }
}

// Fused encode + MLP forward (+ loss, backward, grid scatter, transposed
// stores for the weight gradients when TRAIN).  One wave = 16 samples per
// iteration; all 64 lanes take part in every MFMA.
template <bool TRAIN>
__device__ __forceinline__ void nrc_mlp_body(const MlpArgs& a, __bf16* W) {
    load_weights(a.v.w_bf, W);
    __syncthreads();
    const uint32_t total = a.d_n ? min(*a.d_n, a.n) : a.n;  // a device count never exceeds the capacity a.n
    uint32_t begin = 0, n = total;
    if (TRAIN) slice_of(total, a.part, a.parts, begin, n);
    const float* pos = a.pos + 3 * (size_t)begin;
    const float* dir = a.dir + 3 * (size_t)begin;
    const float* nrm = a.nrm + 3 * (size_t)begin;
    const float* rough = a.rough + begin;
    const float* target = TRAIN ? a.target + 3 * (size_t)begin : nullptr;
    const int lane = threadIdx.x & 63, col = lane & 15, q = lane >> 4, wave = threadIdx.x >> 6;
    const uint32_t ntiles = TRAIN ? ((n + 31) / 32) * 2 : (n + 15) / 16;
    const float scale = n ? 1.f / (3.f * (float)n) : 0.f;
    const int waves = (int)(blockDim.x >> 6);
    for (uint32_t tile = blockIdx.x * waves + wave; tile < ntiles; tile += gridDim.x * waves) {
        const uint32_t s = tile * 16 + col;
        const bool valid = s < n;
        EncIn e{};
        if (valid) e = enc_prepare(a.v, pos, dir, nrm, rough, s);
        bf16x8 xb[4];
        for (int li = 0; li < 4; ++li) {
            float f0 = 0.f, f1 = 0.f;
            if (valid) level_feature(a.v, 4 * q + li, e.p, f0, f1);
            xb[0][2 * li] = (__bf16)f0;
            xb[0][2 * li + 1] = (__bf16)f1;
        }
        for (int sp = 1; sp < 4; ++sp)
            for (int j = 0; j < 8; ++j) xb[sp][j] = (__bf16)(valid ? blob_feature(e, 32 * sp + 8 * q + j) : 0.f);
        // layer 0: 128 -> 64
        f32x4 acc[4];
        for (int t = 0; t < 4; ++t) {
            acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
            for (int sp = 0; sp < 4; ++sp)
                acc[t] = mfma32(*(const bf16x8*)(W + kL0 + (16 * t + col) * kS0 + 32 * sp + 8 * q), xb[sp], acc[t]);
        }
        bf16x4 h1[4], h2[4], h3[4], h4[4];
        relu_pack(acc, h1);
        fwd_hidden(W + kL1, kSH, 4, h1, acc, col, q);
        relu_pack(acc, h2);
        fwd_hidden(W + kL2, kSH, 4, h2, acc, col, q);
        relu_pack(acc, h3);
        fwd_hidden(W + kL3, kSH, 4, h3, acc, col, q);
        relu_pack(acc, h4);
        f32x4 y;
        fwd_hidden(W + kL4, kSH, 1, h4, &y, col, q);
        if (!TRAIN) {
            if (q == 0 && valid) {
                a.out[3 * (size_t)s] = y[0];
                a.out[3 * (size_t)s + 1] = y[1];
                a.out[3 * (size_t)s + 2] = y[2];
            }
            continue;
        }
        // ---- loss and its gradient (outputs 0..2 live in lanes q == 0) ----
        float g[4] = {0.f, 0.f, 0.f, 0.f};
        float lsum = 0.f;
        if (q == 0 && valid) {
            for (int i = 0; i < 3; ++i) {
                float t = target[3 * (size_t)s + i];
                float den = y[i] * y[i] + 0.01f;
                float diff = y[i] - t;
                lsum += diff * diff / den;
                g[i] = 2.f * diff / den * scale;
            }
        }
        for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
        if (lane == 0 && lsum != 0.f) unsafeAtomicAdd(a.loss, lsum);
        const size_t cs = a.cap;
        for (int sp = 0; sp < 4; ++sp)
            for (int j = 0; j < 8; ++j) a.xT[(size_t)(32 * sp + 8 * q + j) * cs + s] = xb[sp][j];
        store_rows(a.hT, cs, s, 0, h1, q);
        store_rows(a.hT, cs, s, 64, h2, q);
        store_rows(a.hT, cs, s, 128, h3, q);
        store_rows(a.hT, cs, s, 192, h4, q);
        // delta_4 (16 rows)
        bf16x4 d4;
        for (int i = 0; i < 4; ++i) {
            d4[i] = (__bf16)g[i];
            a.dT[(size_t)(256 + 4 * q + i) * cs + s] = d4[i];
        }
        // dh4 = W4^T delta_4 (K = 16: 16x16x16 MFMA, delta already in B layout)
        f32x4 dh[4];
        const s16x4 bq = {bfbits(d4[0]), bfbits(d4[1]), bfbits(d4[2]), bfbits(d4[3])};
        for (int t = 0; t < 4; ++t) {
            s16x4 wa;
            for (int j = 0; j < 4; ++j) wa[j] = bfbits(W[kL4 + (4 * q + j) * kSH + 16 * t + col]);
            dh[t] = mfma16(wa, bq, f32x4{0.f, 0.f, 0.f, 0.f});
        }
        bf16x4 d3[4], d2[4], d1[4], d0[4];
        mask_pack(dh, h4, d3);
        bwd_hidden(W + kL3, kSH, 4, d3, dh, col, q);
        mask_pack(dh, h3, d2);
        bwd_hidden(W + kL2, kSH, 4, d2, dh, col, q);
        mask_pack(dh, h2, d1);
        bwd_hidden(W + kL1, kSH, 4, d1, dh, col, q);
        mask_pack(dh, h1, d0);
        store_rows(a.dT, cs, s, 0, d0, q);
        store_rows(a.dT, cs, s, 64, d1, q);
        store_rows(a.dT, cs, s, 128, d2, q);
        store_rows(a.dT, cs, s, 192, d3, q);
        // d(grid features) = W0^T delta_0, input units 0..31.  The 16 lanes
        // of one q hold the same level for 16 samples; when all 16 touch the
        // same grid entry (neighbouring samples in a coarse cell: the
        // training samples come in Morton order), their contributions are
        // summed across the lanes and one lane adds the sum, instead of 16
        // same-address atomics.
        f32x4 dx[2];
        bwd_hidden(W + kL0, kS0, 2, d0, dx, col, q);
        const int lead = lane & ~15;
        for (int t = 0; t < 2; ++t)
            for (int lv = 0; lv < 2; ++lv) {
                const int l = 8 * t + 2 * q + lv;
                const float g0 = valid ? dx[t][2 * lv] : 0.f, g1 = valid ? dx[t][2 * lv + 1] : 0.f;
                uint32_t idx[8];
                float w[8];
                level_corners(a.v, l, e.p, idx, w);
                for (int k = 0; k < 8; ++k) {
                    float c0 = w[k] * g0, c1 = w[k] * g1;
                    const uint32_t li = (uint32_t)__shfl((int)idx[k], lead);
                    // an invalid lane contributes zero wherever it is counted
                    const uint64_t same = __ballot(!valid || idx[k] == li);
                    const bool merged = ((same >> lead) & 0xFFFFull) == 0xFFFFull;
                    if (__ballot(merged)) {  // some group of the wave merges: every lane takes part in the sums
                        for (int off = 8; off >= 1; off >>= 1) {
                            c0 += __shfl_xor(c0, off);
                            c1 += __shfl_xor(c1, off);
                        }
                    }
                    if (merged) {
                        if (lane == lead && valid) {  // an invalid leader: the whole tile is past the count
                            unsafeAtomicAdd(a.grid_grad + 2 * (size_t)li, c0);
                            unsafeAtomicAdd(a.grid_grad + 2 * (size_t)li + 1, c1);
                        }
                    } else if (valid) {
                        unsafeAtomicAdd(a.grid_grad + 2 * (size_t)idx[k], c0);
                        unsafeAtomicAdd(a.grid_grad + 2 * (size_t)idx[k] + 1, c1);
                    }
                }
            }
    }
}

// Inference: the weights end up hoisted into registers; two waves per SIMD
// still fit and hide the grid gathers better than one.
__global__ __launch_bounds__(kFwdBlock, 2) void k_nrc_infer(MlpArgs a) {
    __shared__ __attribute__((aligned(16))) __bf16 W[kLdsW];
    nrc_mlp_body<false>(a, W);
}
__global__ __launch_bounds__(kFwdBlock) void k_nrc_train(MlpArgs a) {
    __shared__ __attribute__((aligned(16))) __bf16 W[kLdsW];
    nrc_mlp_body<true>(a, W);
}

// dW_L = sum over samples of delta_L (x) input_L, split over blocks along the
// samples; blockIdx.y = layer.  Partials are reduced by k_nrc_reduce.
__global__ __launch_bounds__(256) void k_nrc_wgrad(const __bf16* xT, const __bf16* hT, const __bf16* dT, uint32_t cap,
                                                   const uint32_t* d_n, uint32_t n_host, uint32_t mb, uint32_t mbs,
                                                   float* part) {
    const int layer = blockIdx.y;
    uint32_t begin = 0, n = 0;
    slice_of(d_n ? min(*d_n, n_host) : n_host, mb, mbs, begin, n);
    const uint32_t npad = (n + 31) & ~31u;
    const int M = layer == 4 ? 16 : 64, K = layer == 0 ? 128 : 64;
    const int offs[5] = {kNrcOff0, kNrcOff1, kNrcOff2, kNrcOff3, kNrcOff4};
    const __bf16* dl = dT + (size_t)layer * 64 * cap;
    const __bf16* in = layer == 0 ? xT : hT + (size_t)(layer - 1) * 64 * cap;
    const int tk = K / 16, ntile = (M / 16) * tk;
    const int lane = threadIdx.x & 63, col = lane & 15, q = lane >> 4, wave = threadIdx.x >> 6;
    const uint32_t steps = npad / 32;
    const uint32_t per = (steps + gridDim.x - 1) / gridDim.x;
    const uint32_t s_begin = blockIdx.x * per * 32;
    const uint32_t s_end = min(npad, s_begin + per * 32);
    f32x4 acc[8];
    for (int k = 0; k < 8; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (uint32_t s0 = s_begin; s0 < s_end; s0 += 32) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int tile = wave + 4 * k;
            if (tile < ntile) {
                const int ta = tile / tk, tb = tile % tk;
                bf16x8 av = *(const bf16x8*)(dl + (size_t)(16 * ta + col) * cap + s0 + 8 * q);
                bf16x8 bv = *(const bf16x8*)(in + (size_t)(16 * tb + col) * cap + s0 + 8 * q);
                acc[k] = mfma32(av, bv, acc[k]);
            }
        }
    }
    float* out = part + (size_t)blockIdx.x * kNrcParams + offs[layer];
    for (int k = 0; k < 8; ++k) {
        const int tile = wave + 4 * k;
        if (tile < ntile) {
            const int ta = tile / tk, tb = tile % tk;
            for (int i = 0; i < 4; ++i) out[(16 * ta + 4 * q + i) * K + 16 * tb + col] = acc[k][i];
        }
    }
}

__global__ void k_nrc_reduce(const float* part, uint32_t blocks, float* grad) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)kNrcParams) return;
    float s = 0.f;
    for (uint32_t b = 0; b < blocks; ++b) s += part[(size_t)b * kNrcParams + i];
    grad[i] = s;
}

// All-FP32 debug path: one thread per sample, weights and grid from the FP32
// masters, gradients by atomics.
__global__ void k_nrc_f32(MlpArgs a, int train, float* w_grad) {
    const uint32_t total = a.d_n ? min(*a.d_n, a.n) : a.n;  // a device count never exceeds the capacity a.n
    uint32_t begin = 0, n = total;
    if (train) slice_of(total, a.part, a.parts, begin, n);
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const NrcView& v = a.v;
    a.pos += 3 * (size_t)begin;
    a.dir += 3 * (size_t)begin;
    a.nrm += 3 * (size_t)begin;
    a.rough += begin;
    if (train) a.target += 3 * (size_t)begin;
    const float* Wf = v.w_f;
    EncIn e = enc_prepare(v, a.pos, a.dir, a.nrm, a.rough, s);
    float x[kNrcIn];
    for (int l = 0; l < kNrcLevels; ++l) level_feature(v, l, e.p, x[2 * l], x[2 * l + 1]);
    for (int k = 32; k < kNrcIn; ++k) x[k] = blob_feature(e, k);
    float h[4][kNrcHid];
    for (int o = 0; o < kNrcHid; ++o) {
        float acc = 0.f;
        for (int k = 0; k < kNrcIn; ++k) acc += Wf[kNrcOff0 + o * kNrcIn + k] * x[k];
        h[0][o] = fmaxf(acc, 0.f);
    }
    const int offs[5] = {kNrcOff0, kNrcOff1, kNrcOff2, kNrcOff3, kNrcOff4};
    for (int L = 1; L < 4; ++L)
        for (int o = 0; o < kNrcHid; ++o) {
            float acc = 0.f;
            for (int k = 0; k < kNrcHid; ++k) acc += Wf[offs[L] + o * kNrcHid + k] * h[L - 1][k];
            h[L][o] = fmaxf(acc, 0.f);
        }
    float y[3];
    for (int o = 0; o < 3; ++o) {
        float acc = 0.f;
        for (int k = 0; k < kNrcHid; ++k) acc += Wf[kNrcOff4 + o * kNrcHid + k] * h[3][k];
        y[o] = acc;
    }
    if (!train) {
        for (int o = 0; o < 3; ++o) a.out[3 * (size_t)s + o] = y[o];
        return;
    }
    const float scale = 1.f / (3.f * (float)n);
    float g[3], lsum = 0.f;
    for (int i = 0; i < 3; ++i) {
        float t = a.target[3 * (size_t)s + i];
        float den = y[i] * y[i] + 0.01f;
        float diff = y[i] - t;
        lsum += diff * diff / den;
        g[i] = 2.f * diff / den * scale;
    }
    unsafeAtomicAdd(a.loss, lsum);
    float d[kNrcHid], dn[kNrcHid];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < kNrcHid; ++k) unsafeAtomicAdd(w_grad + kNrcOff4 + i * kNrcHid + k, g[i] * h[3][k]);
    for (int k = 0; k < kNrcHid; ++k) {
        float acc = 0.f;
        for (int i = 0; i < 3; ++i) acc += Wf[kNrcOff4 + i * kNrcHid + k] * g[i];
        d[k] = h[3][k] > 0.f ? acc : 0.f;  // delta_3
    }
    for (int L = 3; L >= 1; --L) {
        for (int o = 0; o < kNrcHid; ++o)
            if (d[o] != 0.f)
                for (int k = 0; k < kNrcHid; ++k) unsafeAtomicAdd(w_grad + offs[L] + o * kNrcHid + k, d[o] * h[L - 1][k]);
        for (int k = 0; k < kNrcHid; ++k) {
            float acc = 0.f;
            for (int o = 0; o < kNrcHid; ++o) acc += Wf[offs[L] + o * kNrcHid + k] * d[o];
            dn[k] = h[L - 1][k] > 0.f ? acc : 0.f;
        }
        for (int k = 0; k < kNrcHid; ++k) d[k] = dn[k];
    }
    for (int o = 0; o < kNrcHid; ++o)
        if (d[o] != 0.f)
            for (int k = 0; k < kNrcIn; ++k) unsafeAtomicAdd(w_grad + kNrcOff0 + o * kNrcIn + k, d[o] * x[k]);
    for (int l = 0; l < kNrcLevels; ++l) {
        float g0 = 0.f, g1 = 0.f;
        for (int o = 0; o < kNrcHid; ++o) {
            g0 += Wf[kNrcOff0 + o * kNrcIn + 2 * l] * d[o];
            g1 += Wf[kNrcOff0 + o * kNrcIn + 2 * l + 1] * d[o];
        }
        uint32_t idx[8];
        float w[8];
        level_corners(v, l, e.p, idx, w);
        for (int k = 0; k < 8; ++k) {
            unsafeAtomicAdd(a.grid_grad + 2 * (size_t)idx[k], w[k] * g0);
            unsafeAtomicAdd(a.grid_grad + 2 * (size_t)idx[k] + 1, w[k] * g1);
        }
    }
}

// Sort keys of a frame's training samples: bits 27..31 the sample's random
// mini-batch (hash of its index and the frame, uniform over `parts`), bits
// 0..26 the Morton code of its position on a 512^3 grid over the cache box;
// samples past the count sort last.
__device__ __forceinline__ uint32_t spread3(uint32_t x) {  // 9 bits -> every third bit
    x &= 0x1FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
__global__ void k_nrc_sort_keys(NrcView v, const float* pos, const uint32_t* count, uint32_t cap, uint32_t parts,
                                uint32_t frame, uint32_t* keys, uint32_t* vals) {
    const uint32_t n = min(*count, cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) {
        uint32_t key = 0xFFFFFFFFu;
        if (i < n) {
            uint32_t h = i * 0x9E3779B1u ^ (frame + 0x7F4A7C15u) * 0x85EBCA77u;
            h ^= h >> 15;
            h *= 0x2C1B3C6Du;
            h ^= h >> 12;
            h *= 0x297A2D39u;
            h ^= h >> 15;
            const uint32_t mb = (uint32_t)(((uint64_t)h * parts) >> 32);
            uint32_t m = 0;
            for (int a = 0; a < 3; ++a) {
                const float p = clamp01((pos[3 * (size_t)i + a] - v.origin[a]) * v.inv_extent);
                m |= spread3(min((uint32_t)(p * 512.f), 511u)) << a;
            }
            key = (mb << 27) | m;
        }
        keys[i] = key;
        vals[i] = i;
    }
}

// dst[k] = src[order[k]] for the first count samples (the fields training reads).
__global__ void k_nrc_gather(NrcSamples src, NrcSamples dst, const uint32_t* order, uint32_t cap) {
    const uint32_t n = min(*src.count, cap);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const size_t i = order[k];
        for (int a = 0; a < 3; ++a) {
            dst.pos[3 * (size_t)k + a] = src.pos[3 * i + a];
            dst.dir[3 * (size_t)k + a] = src.dir[3 * i + a];
            dst.nrm[3 * (size_t)k + a] = src.nrm[3 * i + a];
            dst.target[3 * (size_t)k + a] = src.target[3 * i + a];
        }
        dst.rough[k] = src.rough[i];
    }
}

// Adam (Kingma & Ba) with the bias corrections folded into lr_t; clears the gradient.
// lazy: entries with a zero gradient are left untouched (their moments are not decayed).
__global__ void k_adam(float* p, float* g, float* m, float* v, uint64_t n, float lr_t, float b1, float b2, float eps,
                       __bf16* out_bf, _Float16* out_h, int lazy) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        float gi = g[i];
        if (lazy && gi == 0.f) continue;
        float mi = b1 * m[i] + (1.f - b1) * gi;
        float vi = b2 * v[i] + (1.f - b2) * (gi * gi);
        float pi = p[i] - lr_t * mi / (sqrtf(vi) + eps);
        m[i] = mi;
        v[i] = vi;
        p[i] = pi;
        g[i] = 0.f;
        if (out_bf) out_bf[i] = (__bf16)pi;
        if (out_h) out_h[i] = (_Float16)pi;
    }
}

// The same lazy Adam over the hash grid, four entries per thread (n % 4 == 0,
// 16-byte aligned arrays): most of the 2^24 grid parameters are untouched in
// a step, so the pass is a read of the gradients; the arithmetic per touched
// entry is k_adam's.
__global__ void k_adam_grid4(float* p, float* g, float* m, float* v, uint64_t n4, float lr_t, float b1, float b2,
                             float eps, _Float16* out_h) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * blockDim.x) {
        const f32x4 g4 = ((const f32x4*)g)[i];
        if (g4.x == 0.f && g4.y == 0.f && g4.z == 0.f && g4.w == 0.f) continue;
        f32x4 p4 = ((f32x4*)p)[i], m4 = ((f32x4*)m)[i], v4 = ((f32x4*)v)[i];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float gi = g4[k];
            if (gi == 0.f) continue;
            const float mi = b1 * m4[k] + (1.f - b1) * gi;
            const float vi = b2 * v4[k] + (1.f - b2) * (gi * gi);
            p4[k] = p4[k] - lr_t * mi / (sqrtf(vi) + eps);
            m4[k] = mi;
            v4[k] = vi;
            out_h[4 * i + k] = (_Float16)p4[k];
        }
        ((f32x4*)p)[i] = p4;
        ((f32x4*)m)[i] = m4;
        ((f32x4*)v)[i] = v4;
        ((f32x4*)g)[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
}

__global__ void k_round_copy(const float* p, uint64_t n, __bf16* out_bf, _Float16* out_h) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (out_bf) out_bf[i] = (__bf16)p[i];
        if (out_h) out_h[i] = (_Float16)p[i];
    }
}

__global__ void k_fill_f32(float* p, uint64_t n, float value) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// loss[1] = loss[0] / (3 n)
__global__ void k_loss_finish(float* loss, const uint32_t* d_n, uint32_t n_host, uint32_t mb, uint32_t mbs) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        uint32_t begin = 0, n = 0;
        slice_of(d_n ? min(*d_n, n_host) : n_host, mb, mbs, begin, n);
        loss[1] = n ? loss[0] / (3.f * (float)n) : 0.f;
    }
}

__global__ void k_set_u32(uint32_t* p, uint32_t value) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = value;
}

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

uint32_t grid_launch(uint64_t n) { return (uint32_t)std::min<uint64_t>((n + 255) / 256, 8192); }

int fill(float* p, uint64_t n, float value, hipStream_t st) {
    if (n) hipLaunchKernelGGL(k_fill_f32, dim3(grid_launch(n)), dim3(256), 0, st, p, n, value);
    VR_HIP_CHECK(hipGetLastError());
    return 0;
}

int ensure_train_scratch(vr_cache* c, uint32_t n) {
    uint32_t need = (n + 31) & ~31u;
    if (need <= c->train_cap) return 0;
    VR_HIP_CHECK(hipDeviceSynchronize());
    if (c->xT) VR_HIP_CHECK(hipFree(c->xT));
    if (c->hT) VR_HIP_CHECK(hipFree(c->hT));
    if (c->dT) VR_HIP_CHECK(hipFree(c->dT));
    c->xT = c->hT = c->dT = nullptr;
    c->train_cap = 0;
    if (hipMalloc((void**)&c->xT, (size_t)kNrcIn * need * 2) != hipSuccess ||
        hipMalloc((void**)&c->hT, (size_t)4 * kNrcHid * need * 2) != hipSuccess ||
        hipMalloc((void**)&c->dT, (size_t)(4 * kNrcHid + kNrcOut) * need * 2) != hipSuccess)
        return fail(VR_ERR_OOM, "radiance cache training scratch allocation failed");
    c->train_cap = need;
    return 0;
}

MlpArgs make_args(vr_cache* c, uint32_t n, const uint32_t* d_n, const float* pos, const float* dir, const float* nrm,
                  const float* rough) {
    MlpArgs a{};
    a.v = c->view;
    a.n = n;
    a.d_n = d_n;
    a.pos = pos;
    a.dir = dir;
    a.nrm = nrm;
    a.rough = rough;
    return a;
}

uint32_t mlp_blocks(uint32_t n) {
    uint32_t tiles = ((n + 31) / 32) * 2;
    return std::max<uint32_t>(1, std::min<uint32_t>(2048, (tiles + 3) / 4));
}
}  // namespace

int nrc_infer(vr_cache* c, uint32_t n, const uint32_t* d_n, const float* pos, const float* dir, const float* nrm,
              const float* rough, float* out, hipStream_t st) {
    if (n == 0) return 0;
    MlpArgs a = make_args(c, n, d_n, pos, dir, nrm, rough);
    a.out = out;
    if (c->view.fp32)
        hipLaunchKernelGGL(k_nrc_f32, dim3((n + 127) / 128), dim3(128), 0, st, a, 0, (float*)nullptr);
    else
        hipLaunchKernelGGL(k_nrc_infer, dim3(mlp_blocks(n)), dim3(kFwdBlock), 0, st, a);
    VR_HIP_CHECK(hipGetLastError());
    return 0;
}

int nrc_train_step(vr_cache* c, uint32_t n, const uint32_t* d_n, const float* pos, const float* dir, const float* nrm,
                   const float* rough, const float* target, bool update, hipStream_t st, uint32_t part,
                   uint32_t parts) {
    if (parts == 0 || part >= parts) return fail(VR_ERR_INVALID, "bad mini-batch index");
    int rc = ensure_train_scratch(c, std::max(n, 1u));
    if (rc) return rc;
    rc = fill(c->loss, 2, 0.f, st);
    if (rc) return rc;
    MlpArgs a = make_args(c, n, d_n, pos, dir, nrm, rough);
    a.target = target;
    a.loss = c->loss;
    a.grid_grad = c->g_grad;
    a.part = part;
    a.parts = parts;
    if (n) {
        if (c->view.fp32) {
            hipLaunchKernelGGL(k_nrc_f32, dim3((n + 127) / 128), dim3(128), 0, st, a, 1, c->w_grad);
        } else {
            a.xT = c->xT;
            a.hT = c->hT;
            a.dT = c->dT;
            a.cap = c->train_cap;
            // a mini-batch holds at most ceil(n / parts) samples: no blocks beyond those
            hipLaunchKernelGGL(k_nrc_train, dim3(mlp_blocks((n + parts - 1) / parts)), dim3(kFwdBlock), 0, st, a);
            hipLaunchKernelGGL(k_nrc_wgrad, dim3(kWgradBlocks, 5), dim3(256), 0, st, (const __bf16*)c->xT,
                               (const __bf16*)c->hT, (const __bf16*)c->dT, c->train_cap, d_n, n, part, parts,
                               c->part);
            hipLaunchKernelGGL(k_nrc_reduce, dim3((kNrcParams + 255) / 256), dim3(256), 0, st, (const float*)c->part,
                               (uint32_t)kWgradBlocks, c->w_grad);
        }
    }
    hipLaunchKernelGGL(k_loss_finish, dim3(1), dim3(64), 0, st, c->loss, d_n, n, part, parts);
    VR_HIP_CHECK(hipGetLastError());
    if (!update) return 0;
    c->info.steps += 1;
    const double t = (double)c->info.steps;
    const double b1 = c->desc.beta1, b2 = c->desc.beta2;
    const double corr = std::sqrt(1.0 - std::pow(b2, t)) / (1.0 - std::pow(b1, t));
    const float lrw = (float)(c->desc.lr_mlp * corr), lrg = (float)(c->desc.lr_grid * corr);
    hipLaunchKernelGGL(k_adam, dim3(grid_launch(kNrcParams)), dim3(256), 0, st, c->w_master, c->w_grad, c->w_m,
                       c->w_v, (uint64_t)kNrcParams, lrw, (float)b1, (float)b2, c->desc.eps, c->w_bf,
                       (_Float16*)nullptr, 0);
    static_assert(kNrcLevels * 2 % 4 == 0, "grid parameters come in multiples of 4");
    hipLaunchKernelGGL(k_adam_grid4, dim3(grid_launch(c->info.grid_params / 4)), dim3(256), 0, st, c->g_master,
                       c->g_grad, c->g_m, c->g_v, c->info.grid_params / 4, lrg, (float)b1, (float)b2, c->desc.eps,
                       c->g_h);
    VR_HIP_CHECK(hipGetLastError());
    return 0;
}

int nrc_ensure_frame_buffers(vr_cache* c, uint32_t train_cap, uint32_t query_cap) {
    auto grow = [&](NrcSamples& S, uint32_t cap) -> int {
        if (S.cap >= cap) return 0;
        VR_HIP_CHECK(hipDeviceSynchronize());
        NrcSamples T{};
        size_t n = std::max(cap, 1u);
        auto alloc = [&](void** p, size_t bytes) -> bool {
            if (hipMalloc(p, bytes) != hipSuccess) return false;
            c->frame_allocs.push_back(*p);
            return true;
        };
        bool ok = alloc((void**)&T.pos, 12 * n) && alloc((void**)&T.dir, 12 * n) && alloc((void**)&T.nrm, 12 * n) &&
                  alloc((void**)&T.rough, 4 * n) && alloc((void**)&T.target, 12 * n) &&
                  alloc((void**)&T.path, 4 * n) && alloc((void**)&T.A, 12 * n) && alloc((void**)&T.L, 12 * n) &&
                  alloc((void**)&T.count, 256);
        if (!ok) return fail(VR_ERR_OOM, "radiance cache sample buffers allocation failed");
        T.cap = cap;
        S = T;  // the old buffers stay in frame_allocs until the cache is destroyed
        return 0;
    };
    const uint32_t old_cap = c->tr.cap;
    int rc = grow(c->tr, train_cap);
    if (rc) return rc;
    if (c->tr.cap != old_cap) {
        // sorted copies of the training samples, the sort's keys and values and its scratch
        const size_t n = std::max(c->tr.cap, 1u);
        NrcSamples T{};
        void* p[7] = {};
        const size_t bytes[7] = {12 * n, 12 * n, 12 * n, 4 * n, 12 * n, 8 * n, 8 * n};
        for (int k = 0; k < 7; ++k) {
            if (hipMalloc(&p[k], bytes[k]) != hipSuccess) return fail(VR_ERR_OOM, "radiance cache sort buffers");
            c->frame_allocs.push_back(p[k]);
        }
        T.pos = (float*)p[0];
        T.dir = (float*)p[1];
        T.nrm = (float*)p[2];
        T.rough = (float*)p[3];
        T.target = (float*)p[4];
        T.count = c->tr.count;
        T.cap = c->tr.cap;
        c->trs = T;
        c->sort_keys = (uint32_t*)p[5];
        c->sort_vals = (uint32_t*)p[6];
        size_t temp = 0;
        VR_HIP_CHECK(rocprim::radix_sort_pairs(nullptr, temp, c->sort_keys, c->sort_keys + n, c->sort_vals,
                                               c->sort_vals + n, c->tr.cap, 0, 32, (hipStream_t)0));
        if (hipMalloc(&c->sort_temp, std::max<size_t>(temp, 256)) != hipSuccess)
            return fail(VR_ERR_OOM, "radiance cache sort scratch");
        c->frame_allocs.push_back(c->sort_temp);
        c->sort_temp_bytes = std::max<size_t>(temp, 256);
    }
    rc = ensure_train_scratch(c, train_cap);
    if (rc) return rc;
    return grow(c->qr, query_cap);
}

int nrc_order_samples(vr_cache* c, uint32_t parts, uint32_t frame, hipStream_t st) {
    if (parts == 0 || parts > kNrcMaxParts) return fail(VR_ERR_INVALID, "train_steps must be in [1, 32]");
    const uint32_t cap = c->tr.cap;
    if (cap == 0 || !c->sort_keys) return fail(VR_ERR_INVALID, "training sample buffers not allocated");
    const size_t n = cap;
    hipLaunchKernelGGL(k_nrc_sort_keys, dim3(grid_launch(cap)), dim3(256), 0, st, c->view, (const float*)c->tr.pos,
                       (const uint32_t*)c->tr.count, cap, parts, frame, c->sort_keys, c->sort_vals);
    VR_HIP_CHECK(hipGetLastError());
    size_t temp = c->sort_temp_bytes;
    VR_HIP_CHECK(rocprim::radix_sort_pairs(c->sort_temp, temp, c->sort_keys, c->sort_keys + n, c->sort_vals,
                                           c->sort_vals + n, cap, 0, 32, st));
    hipLaunchKernelGGL(k_nrc_gather, dim3(grid_launch(cap)), dim3(256), 0, st, c->tr, c->trs,
                       (const uint32_t*)(c->sort_vals + n), cap);
    VR_HIP_CHECK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
static void free_cache(vr_cache* c) {
    void* ptrs[] = {c->w_master, c->w_m, c->w_v, c->w_grad, c->w_bf, c->g_master, c->g_m, c->g_v, c->g_grad, c->g_h,
                    c->xT, c->hT, c->dT, c->part, c->loss, c->counts};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (void* p : c->frame_allocs) (void)hipFree(p);
    if (c->host_rb) (void)hipHostFree(c->host_rb);
    if (c->done) (void)hipEventDestroy(c->done);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
}

static int create_cache(vr_ctx* ctx, const vr_cache_desc* d, vr_cache** out) {
    if (d->log2_table < 8 || d->log2_table > 24) return fail(VR_ERR_INVALID, "log2_table must be in [8, 24]");
    if (d->base_resolution < 2 || d->max_resolution < d->base_resolution || d->max_resolution > (1u << 20))
        return fail(VR_ERR_INVALID, "bad grid resolutions");
    if (d->precision != VR_CACHE_BF16 && d->precision != VR_CACHE_FP32) return fail(VR_ERR_INVALID, "bad precision");
    float ext = 0.f;
    for (int a = 0; a < 3; ++a) ext = std::max(ext, d->bbox_hi[a] - d->bbox_lo[a]);
    if (!(ext > 0.f) || !std::isfinite(ext)) return fail(VR_ERR_INVALID, "bounding box must have a positive extent");
    auto* c = new vr_cache();
    c->ctx = ctx;
    c->desc = *d;
    vr_cache_info& I = c->info;
    I.levels = kNrcLevels;
    I.table_size = 1u << d->log2_table;
    I.precision = d->precision;
    const double b = std::exp((std::log((double)d->max_resolution) - std::log((double)d->base_resolution)) /
                              (double)(kNrcLevels - 1));
    NrcView& v = c->view;
    for (int l = 0; l < kNrcLevels; ++l) {
        uint32_t r = (uint32_t)std::floor((double)d->base_resolution * std::pow(b, (double)l) + 1e-6);
        I.resolution[l] = r;
        v.res[l] = r;
        v.dense[l] = (uint64_t)(r + 1) * (r + 1) * (r + 1) <= I.table_size ? 1u : 0u;
    }
    for (int a = 0; a < 3; ++a) {
        I.origin[a] = d->bbox_lo[a];
        v.origin[a] = d->bbox_lo[a];
    }
    I.inv_extent = 1.0f / ext;
    v.inv_extent = I.inv_extent;
    v.table_mask = I.table_size - 1;
    v.log2_table = d->log2_table;
    v.fp32 = d->precision == VR_CACHE_FP32 ? 1u : 0u;
    I.grid_params = (uint64_t)kNrcLevels * I.table_size * 2;
    const uint64_t G = I.grid_params;
    bool ok = hipMalloc((void**)&c->w_master, 4 * kNrcParams) == hipSuccess &&
              hipMalloc((void**)&c->w_m, 4 * kNrcParams) == hipSuccess &&
              hipMalloc((void**)&c->w_v, 4 * kNrcParams) == hipSuccess &&
              hipMalloc((void**)&c->w_grad, 4 * kNrcParams) == hipSuccess &&
              hipMalloc((void**)&c->w_bf, 2 * kNrcParams) == hipSuccess &&
              hipMalloc((void**)&c->g_master, 4 * G) == hipSuccess && hipMalloc((void**)&c->g_m, 4 * G) == hipSuccess &&
              hipMalloc((void**)&c->g_v, 4 * G) == hipSuccess && hipMalloc((void**)&c->g_grad, 4 * G) == hipSuccess &&
              hipMalloc((void**)&c->g_h, 2 * G) == hipSuccess &&
              hipMalloc((void**)&c->part, 4ull * kWgradBlocks * kNrcParams) == hipSuccess &&
              hipMalloc((void**)&c->loss, 256) == hipSuccess && hipMalloc((void**)&c->counts, 256) == hipSuccess &&
              hipHostMalloc((void**)&c->host_rb, 256, hipHostMallocDefault) == hipSuccess &&
              hipEventCreateWithFlags(&c->done, hipEventDisableTiming) == hipSuccess &&
              hipEventCreate(&c->ev[0]) == hipSuccess && hipEventCreate(&c->ev[1]) == hipSuccess &&
              hipEventCreate(&c->ev[2]) == hipSuccess;
    if (!ok) {
        free_cache(c);
        delete c;
        return fail(VR_ERR_OOM, "radiance cache allocation failed");
    }
    std::memset(c->host_rb, 0, 256);
    v.w_bf = c->w_bf;
    v.w_f = c->w_master;
    v.grid_h = c->g_h;
    v.grid_f = c->g_master;
    // initialisation: Glorot-uniform weights, grid U(-1e-4, 1e-4)
    uint64_t s = d->seed ^ 0x5EEDCAC4Eull;
    std::vector<float> w(kNrcParams, 0.f);
    const int outs[5] = {64, 64, 64, 64, 16}, ins[5] = {128, 64, 64, 64, 64};
    const int offs[5] = {kNrcOff0, kNrcOff1, kNrcOff2, kNrcOff3, kNrcOff4};
    for (int L = 0; L < 5; ++L) {
        const int rows = L == 4 ? 3 : outs[L];
        const double lim = std::sqrt(6.0 / (double)(ins[L] + (L == 4 ? 3 : outs[L])));
        for (int o = 0; o < rows; ++o)
            for (int k = 0; k < ins[L]; ++k) w[offs[L] + o * ins[L] + k] = (float)((2.0 * uniform01(s) - 1.0) * lim);
    }
    std::vector<float> g(G);
    for (uint64_t i = 0; i < G; ++i) g[i] = (float)((2.0 * uniform01(s) - 1.0) * 1e-4);
    hipStream_t st = ctx->stream;
    int rc = 0;
    if (hipMemcpy(c->w_master, w.data(), 4 * kNrcParams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->g_master, g.data(), 4 * G, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(VR_ERR_HIP, "radiance cache upload failed");
    float* zero[] = {c->w_m, c->w_v, c->w_grad, c->g_m, c->g_v, c->g_grad};
    uint64_t zn[] = {kNrcParams, kNrcParams, kNrcParams, G, G, G};
    for (int k = 0; k < 6 && !rc; ++k) rc = fill(zero[k], zn[k], 0.f, st);
    if (!rc) {
        hipLaunchKernelGGL(k_round_copy, dim3(grid_launch(kNrcParams)), dim3(256), 0, st, (const float*)c->w_master,
                           (uint64_t)kNrcParams, c->w_bf, (_Float16*)nullptr);
        hipLaunchKernelGGL(k_round_copy, dim3(grid_launch(G)), dim3(256), 0, st, (const float*)c->g_master, G,
                           (__bf16*)nullptr, c->g_h);
        if (hipStreamSynchronize(st) != hipSuccess) rc = fail(VR_ERR_HIP, "radiance cache initialisation failed");
    }
    if (rc) {
        free_cache(c);
        delete c;
        return rc;
    }
    *out = c;
    return 0;
}

}  // namespace vr

using namespace vr;

namespace {
template <class F> int guard(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(VR_ERR_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(VR_ERR_INVALID, e.what());
    }
}
hipStream_t pick(vr_cache* c, void* stream) { return stream ? (hipStream_t)stream : c->ctx->stream; }
}  // namespace

extern "C" {

int vr_cache_create(vr_ctx* ctx, const vr_cache_desc* desc, vr_cache** out) {
    return guard([&] {
        if (!ctx || !desc || !out) return fail(VR_ERR_INVALID, "null argument");
        VR_HIP_CHECK(hipSetDevice(ctx->device));
        return create_cache(ctx, desc, out);
    });
}

int vr_cache_destroy(vr_cache* c) {
    if (!c) return 0;
    (void)hipSetDevice(c->ctx->device);
    (void)hipDeviceSynchronize();
    free_cache(c);
    delete c;
    return 0;
}

int vr_cache_get_info(const vr_cache* c, vr_cache_info* info) {
    if (!c || !info) return fail(VR_ERR_INVALID, "null argument");
    *info = c->info;
    return 0;
}

int vr_cache_get_params(vr_cache* c, float* mlp, float* grid) {
    return guard([&] {
        if (!c) return fail(VR_ERR_INVALID, "null argument");
        VR_HIP_CHECK(hipDeviceSynchronize());
        const uint64_t G = c->info.grid_params;
        if (mlp) {
            if (c->view.fp32) {
                VR_HIP_CHECK(hipMemcpy(mlp, c->w_master, 4 * kNrcParams, hipMemcpyDeviceToHost));
            } else {
                std::vector<uint16_t> b(kNrcParams);
                VR_HIP_CHECK(hipMemcpy(b.data(), c->w_bf, 2 * kNrcParams, hipMemcpyDeviceToHost));
                for (int i = 0; i < kNrcParams; ++i) {
                    uint32_t u = (uint32_t)b[i] << 16;
                    std::memcpy(&mlp[i], &u, 4);
                }
            }
        }
        if (grid) {
            if (c->view.fp32) {
                VR_HIP_CHECK(hipMemcpy(grid, c->g_master, 4 * G, hipMemcpyDeviceToHost));
            } else {
                std::vector<_Float16> h(G);
                VR_HIP_CHECK(hipMemcpy(h.data(), c->g_h, 2 * G, hipMemcpyDeviceToHost));
                for (uint64_t i = 0; i < G; ++i) grid[i] = (float)h[i];
            }
        }
        return 0;
    });
}

int vr_cache_set_params(vr_cache* c, const float* mlp, const float* grid) {
    return guard([&] {
        if (!c) return fail(VR_ERR_INVALID, "null argument");
        VR_HIP_CHECK(hipDeviceSynchronize());
        const uint64_t G = c->info.grid_params;
        hipStream_t st = c->ctx->stream;
        if (mlp) VR_HIP_CHECK(hipMemcpy(c->w_master, mlp, 4 * kNrcParams, hipMemcpyHostToDevice));
        if (grid) VR_HIP_CHECK(hipMemcpy(c->g_master, grid, 4 * G, hipMemcpyHostToDevice));
        float* zero[] = {c->w_m, c->w_v, c->w_grad, c->g_m, c->g_v, c->g_grad};
        uint64_t zn[] = {kNrcParams, kNrcParams, kNrcParams, G, G, G};
        for (int k = 0; k < 6; ++k) {
            int rc = fill(zero[k], zn[k], 0.f, st);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_round_copy, dim3(grid_launch(kNrcParams)), dim3(256), 0, st, (const float*)c->w_master,
                           (uint64_t)kNrcParams, c->w_bf, (_Float16*)nullptr);
        hipLaunchKernelGGL(k_round_copy, dim3(grid_launch(G)), dim3(256), 0, st, (const float*)c->g_master, G,
                           (__bf16*)nullptr, c->g_h);
        VR_HIP_CHECK(hipStreamSynchronize(st));
        c->info.steps = 0;
        return 0;
    });
}

int vr_cache_infer_async(vr_cache* c, uint32_t n, const float* d_pos, const float* d_dir, const float* d_nrm,
                         const float* d_rough, float* d_out, void* stream) {
    return guard([&] {
        if (!c || (n && (!d_pos || !d_dir || !d_nrm || !d_rough || !d_out))) return fail(VR_ERR_INVALID, "null argument");
        return nrc_infer(c, n, nullptr, d_pos, d_dir, d_nrm, d_rough, d_out, pick(c, stream));
    });
}

int vr_cache_gradients_async(vr_cache* c, uint32_t n, const float* d_pos, const float* d_dir, const float* d_nrm,
                             const float* d_rough, const float* d_target, float* d_mlp_grad, float* d_grid_grad,
                             float* d_loss, void* stream) {
    return guard([&] {
        if (!c || (n && (!d_pos || !d_dir || !d_nrm || !d_rough || !d_target)))
            return fail(VR_ERR_INVALID, "null argument");
        hipStream_t st = pick(c, stream);
        int rc = nrc_train_step(c, n, nullptr, d_pos, d_dir, d_nrm, d_rough, d_target, false, st, 0, 1);
        if (rc) return rc;
        const uint64_t G = c->info.grid_params;
        if (d_mlp_grad) VR_HIP_CHECK(hipMemcpyAsync(d_mlp_grad, c->w_grad, 4 * kNrcParams, hipMemcpyDeviceToDevice, st));
        if (d_grid_grad) VR_HIP_CHECK(hipMemcpyAsync(d_grid_grad, c->g_grad, 4 * G, hipMemcpyDeviceToDevice, st));
        if (d_loss) VR_HIP_CHECK(hipMemcpyAsync(d_loss, c->loss + 1, 4, hipMemcpyDeviceToDevice, st));
        rc = fill(c->w_grad, kNrcParams, 0.f, st);
        if (!rc) rc = fill(c->g_grad, G, 0.f, st);
        return rc;
    });
}

int vr_cache_train_async(vr_cache* c, uint32_t n, const float* d_pos, const float* d_dir, const float* d_nrm,
                         const float* d_rough, const float* d_target, void* stream) {
    return guard([&] {
        if (!c || (n && (!d_pos || !d_dir || !d_nrm || !d_rough || !d_target)))
            return fail(VR_ERR_INVALID, "null argument");
        hipStream_t st = pick(c, stream);
        int rc = nrc_train_step(c, n, nullptr, d_pos, d_dir, d_nrm, d_rough, d_target, true, st, 0, 1);
        if (rc) return rc;
        VR_HIP_CHECK(hipMemcpyAsync(c->host_rb + 2, c->loss + 1, 4, hipMemcpyDeviceToHost, st));
        VR_HIP_CHECK(hipEventRecord(c->done, st));
        c->done_recorded = true;
        c->frame_timed = false;
        return 0;
    });
}

int vr_cache_last_stats(vr_cache* c, vr_cache_stats* stats) {
    return guard([&] {
        if (!c || !stats) return fail(VR_ERR_INVALID, "null argument");
        *stats = vr_cache_stats{};
        if (!c->done_recorded) return 0;
        VR_HIP_CHECK(hipEventSynchronize(c->done));
        float l;
        std::memcpy(&l, &c->host_rb[2], 4);
        stats->loss = l;
        stats->train_samples = c->host_rb[0];
        stats->queries = c->host_rb[1];
        if (c->frame_timed) {
            float t = 0.f, r = 0.f;
            VR_HIP_CHECK(hipEventElapsedTime(&r, c->ev[0], c->ev[1]));  // the frame's paths
            VR_HIP_CHECK(hipEventElapsedTime(&t, c->ev[1], c->ev[2]));  // Adam steps on its samples
            stats->train_ms = t;
            stats->render_ms = r;
        }
        return 0;
    });
}

}  // extern "C"

namespace vr {
// Read-back of the frame's counts and loss, recorded by the cached renderer.
int nrc_frame_readback(vr_cache* c, hipStream_t st) {
    if (c->tr.count) VR_HIP_CHECK(hipMemcpyAsync(c->host_rb, c->tr.count, 4, hipMemcpyDeviceToHost, st));
    if (c->qr.count) VR_HIP_CHECK(hipMemcpyAsync(c->host_rb + 1, c->qr.count, 4, hipMemcpyDeviceToHost, st));
    VR_HIP_CHECK(hipMemcpyAsync(c->host_rb + 2, c->loss + 1, 4, hipMemcpyDeviceToHost, st));
    VR_HIP_CHECK(hipEventRecord(c->done, st));
    c->done_recorded = true;
    return 0;
}
}  // namespace vr
