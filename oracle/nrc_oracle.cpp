// ORACLE — test infrastructure only.  FP64 restatement of the learned radiance
// cache's network: hash-grid + one-blob encoding, the 128-64-64-64-64-3 ReLU
// MLP, the relative L2 loss and its gradients.  Used by tests/ to check the
// library's BF16/MFMA and FP32 kernels; never linked into the product.
//
// The cache has no counterpart in voidray; its definition (Mueller et al.
// 2022 multiresolution hash encoding, one-blob encoding of Mueller et al.
// 2019, relative L2 loss of Mueller et al. 2021) is restated here from those
// papers and the library's published interface (include/voidray_hip.h):
//   grid level l: resolution res[l]; p' = clamp((p - origin) * inv_extent, 0, 1)
//     (f32, as the index computation must match), g = p' * res (f32),
//     cell = min(floor(g), res - 1), frac = g - cell; corner (x, y, z) index:
//     dense  x + (res+1) * (y + (res+1) * z)   if (res+1)^3 <= entries
//     hashed (x ^ y * 2654435761 ^ z * 805459861) mod entries   (u32 arithmetic)
//     feature f = sum over corners of trilinear weight * grid[l][index][f]
//   one-blob bin b of v with B bins: exp(-(v*B - (b + 1/2))^2 / 2)
//   inputs: 32 grid features, direction (acos(d.y)/pi, atan2(d.z, d.x)/2pi + 1/2)
//     x 16 bins, normal likewise, roughness x 8 bins, 1, zeros to 128
//   loss: mean over samples and 3 channels of (y - t)^2 / (y^2 + 0.01), the
//     denominator treated as a constant
// Everything after the f32 index computation is double.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int L = 16, IN = 128, HID = 64;
constexpr int OFF[5] = {0, 8192, 12288, 16384, 20480};
constexpr int NPARAM = 21504;

struct Layout {
    const uint32_t* res;
    uint32_t log2_table;
    const float* origin;
    float inv_extent;
};

struct Corners {
    uint32_t idx[8];
    double w[8];
};

float clamp01f(float x) { return x < 0.f ? 0.f : (x > 1.f ? 1.f : x); }

// Rounding model of the BF16 kernels: value -> f32 -> bf16 (round to nearest even).
double bf16r(double v) {
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

void corners(const Layout& lay, int l, const float pn[3], Corners& c) {
    const uint32_t r = lay.res[l];
    const uint32_t T = 1u << lay.log2_table;
    const bool dense = (uint64_t)(r + 1) * (r + 1) * (r + 1) <= T;
    uint32_t ci[3];
    double fr[3];
    for (int a = 0; a < 3; ++a) {
        float g = pn[a] * (float)r;
        uint32_t k = (uint32_t)std::floor(g);
        if (k > r - 1) k = r - 1;
        ci[a] = k;
        fr[a] = (double)(g - (float)k);
    }
    for (int k = 0; k < 8; ++k) {
        uint32_t x = ci[0] + (k & 1), y = ci[1] + ((k >> 1) & 1), z = ci[2] + ((k >> 2) & 1);
        uint32_t i = dense ? x + (r + 1) * (y + (r + 1) * z) : ((x ^ (y * 2654435761u) ^ (z * 805459861u)) & (T - 1));
        c.idx[k] = ((uint32_t)l << lay.log2_table) + i;
        double wx = (k & 1) ? fr[0] : 1.0 - fr[0];
        double wy = (k & 2) ? fr[1] : 1.0 - fr[1];
        double wz = (k & 4) ? fr[2] : 1.0 - fr[2];
        c.w[k] = wx * wy * wz;
    }
}

double blob(double v, int b, double nb) {
    double d = v * nb - (b + 0.5);
    return std::exp(-0.5 * d * d);
}

struct Sample {
    double x[IN];
    Corners cs[L];
};

void encode(const Layout& lay, const float* p, const float* d, const float* n, float rough, const double* grid,
            Sample& s, int bf16 = 0) {
    float pn[3];
    for (int a = 0; a < 3; ++a) pn[a] = clamp01f((p[a] - lay.origin[a]) * lay.inv_extent);
    for (int l = 0; l < L; ++l) {
        corners(lay, l, pn, s.cs[l]);
        double f0 = 0, f1 = 0;
        for (int k = 0; k < 8; ++k) {
            f0 += s.cs[l].w[k] * grid[2 * (size_t)s.cs[l].idx[k]];
            f1 += s.cs[l].w[k] * grid[2 * (size_t)s.cs[l].idx[k] + 1];
        }
        s.x[2 * l] = f0;
        s.x[2 * l + 1] = f1;
    }
    auto clamp1 = [](double v) { return v < -1 ? -1.0 : (v > 1 ? 1.0 : v); };
    const double pi = 3.14159265358979323846;
    double dth = std::acos(clamp1(d[1])) / pi, dph = std::atan2((double)d[2], (double)d[0]) / (2 * pi) + 0.5;
    double nth = std::acos(clamp1(n[1])) / pi, nph = std::atan2((double)n[2], (double)n[0]) / (2 * pi) + 0.5;
    double r = rough < 0.f ? 0.0 : (rough > 1.f ? 1.0 : (double)rough);
    for (int b = 0; b < 16; ++b) {
        s.x[32 + b] = blob(dth, b, 16);
        s.x[48 + b] = blob(dph, b, 16);
        s.x[64 + b] = blob(nth, b, 16);
        s.x[80 + b] = blob(nph, b, 16);
    }
    for (int b = 0; b < 8; ++b) s.x[96 + b] = blob(r, b, 8);
    s.x[104] = 1.0;
    for (int k = 105; k < IN; ++k) s.x[k] = 0.0;
    if (bf16)
        for (int k = 0; k < IN; ++k) s.x[k] = bf16r(s.x[k]);
}

// h[0..3] post-ReLU activations, y[3].  bf16: activations rounded as the
// BF16 kernels round them (the inputs x are rounded by encode).
void forward(const double* W, const Sample& s, double h[4][HID], double y[3], int bf16 = 0) {
    for (int o = 0; o < HID; ++o) {
        double a = 0;
        for (int k = 0; k < IN; ++k) a += W[OFF[0] + o * IN + k] * s.x[k];
        h[0][o] = a > 0 ? (bf16 ? bf16r(a) : a) : 0;
    }
    for (int Lr = 1; Lr < 4; ++Lr)
        for (int o = 0; o < HID; ++o) {
            double a = 0;
            for (int k = 0; k < HID; ++k) a += W[OFF[Lr] + o * HID + k] * h[Lr - 1][k];
            h[Lr][o] = a > 0 ? (bf16 ? bf16r(a) : a) : 0;
        }
    for (int o = 0; o < 3; ++o) {
        double a = 0;
        for (int k = 0; k < HID; ++k) a += W[OFF[4] + o * HID + k] * h[3][k];
        y[o] = a;
    }
}

}  // namespace

extern "C" {

// out: n*3 predicted radiance.
// bf16 = 1: the rounding model of the BF16 kernels (inputs, activations and
// deltas rounded to bf16 where the kernels round them; FP64 accumulation).
int orc_nrc_forward(int n, const float* pos, const float* dir, const float* nrm, const float* rough,
                    const uint32_t* res, uint32_t log2_table, const float* origin, float inv_extent, const double* mlp,
                    const double* grid, double* out, int bf16) {
    Layout lay{res, log2_table, origin, inv_extent};
    std::vector<Sample> S(1);
    double h[4][HID];
    for (int i = 0; i < n; ++i) {
        encode(lay, pos + 3 * i, dir + 3 * i, nrm + 3 * i, rough[i], grid, S[0], bf16);
        forward(mlp, S[0], h, out + 3 * i, bf16);
    }
    return 0;
}

// Mean relative L2 loss and its gradients w.r.t. the MLP weights (NPARAM,
// rows 3..15 of the last layer stay zero) and the grid (dense, levels *
// entries * 2).  Returns the loss.  abs_mode = 1: the same recursion with
// every weight, input and loss derivative replaced by its absolute value
// (ReLU masks unchanged) — each entry is then the sum of the absolute values
// of the terms of that gradient entry, the scale of its rounding error.
double orc_nrc_gradients(int n, const float* pos, const float* dir, const float* nrm, const float* rough,
                         const float* target, const uint32_t* res, uint32_t log2_table, const float* origin,
                         float inv_extent, const double* W, const double* grid, double* gW, double* gG,
                         int abs_mode, int bf16) {
    auto A = [abs_mode](double v) { return abs_mode ? std::fabs(v) : v; };
    auto R = [bf16](double v) { return bf16 ? bf16r(v) : v; };
    Layout lay{res, log2_table, origin, inv_extent};
    const size_t G = (size_t)L << log2_table;
    std::memset(gW, 0, sizeof(double) * NPARAM);
    std::memset(gG, 0, sizeof(double) * G * 2);
    std::vector<Sample> S(1);
    Sample& s = S[0];
    double h[4][HID], y[3];
    double loss = 0;
    const double scale = 1.0 / (3.0 * n);
    for (int i = 0; i < n; ++i) {
        encode(lay, pos + 3 * i, dir + 3 * i, nrm + 3 * i, rough[i], grid, s, bf16);
        forward(W, s, h, y, bf16);
        double g[3];
        for (int c = 0; c < 3; ++c) {
            double t = target[3 * i + c];
            double den = y[c] * y[c] + 0.01;
            loss += (y[c] - t) * (y[c] - t) / den;
            g[c] = R(A(2.0 * (y[c] - t) / den * scale));
        }
        double d[HID], dn[HID];
        for (int c = 0; c < 3; ++c)
            for (int k = 0; k < HID; ++k) gW[OFF[4] + c * HID + k] += g[c] * h[3][k];
        for (int k = 0; k < HID; ++k) {
            double a = 0;
            for (int c = 0; c < 3; ++c) a += A(W[OFF[4] + c * HID + k]) * g[c];
            d[k] = h[3][k] > 0 ? R(a) : 0;
        }
        for (int Lr = 3; Lr >= 1; --Lr) {
            for (int o = 0; o < HID; ++o)
                for (int k = 0; k < HID; ++k) gW[OFF[Lr] + o * HID + k] += d[o] * h[Lr - 1][k];
            for (int k = 0; k < HID; ++k) {
                double a = 0;
                for (int o = 0; o < HID; ++o) a += A(W[OFF[Lr] + o * HID + k]) * d[o];
                dn[k] = h[Lr - 1][k] > 0 ? R(a) : 0;
            }
            std::memcpy(d, dn, sizeof d);
        }
        for (int o = 0; o < HID; ++o)
            for (int k = 0; k < IN; ++k) gW[OFF[0] + o * IN + k] += d[o] * A(s.x[k]);
        for (int l = 0; l < L; ++l) {
            double g0 = 0, g1 = 0;
            for (int o = 0; o < HID; ++o) {
                g0 += A(W[OFF[0] + o * IN + 2 * l]) * d[o];
                g1 += A(W[OFF[0] + o * IN + 2 * l + 1]) * d[o];
            }
            for (int k = 0; k < 8; ++k) {
                gG[2 * (size_t)s.cs[l].idx[k]] += s.cs[l].w[k] * g0;
                gG[2 * (size_t)s.cs[l].idx[k] + 1] += s.cs[l].w[k] * g1;
            }
        }
    }
    return loss * scale;
}

// The 8 corner indices (level base included) and trilinear weights of one level.
int orc_nrc_corners(const float* pos, int level, const uint32_t* res, uint32_t log2_table, const float* origin,
                    float inv_extent, uint32_t* idx, double* w) {
    Layout lay{res, log2_table, origin, inv_extent};
    float pn[3];
    for (int a = 0; a < 3; ++a) pn[a] = clamp01f((pos[a] - origin[a]) * inv_extent);
    Corners c;
    corners(lay, level, pn, c);
    for (int k = 0; k < 8; ++k) {
        idx[k] = c.idx[k];
        w[k] = c.w[k];
    }
    return 0;
}

}  // extern "C"
