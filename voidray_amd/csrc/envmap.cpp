#include "envmap.h"

#include <algorithm>
#include <cmath>

namespace vr {

namespace {
// Bilinear lookup of environments.rs:57-76 in double (weights only).
double lum_at(const float* rgb, unsigned W, unsigned H, double x, double y) {
    size_t len = (size_t)W * H;
    auto sat = [](double v, size_t hi) -> size_t {
        if (!(v >= 0)) return 0;
        size_t k = (size_t)v;
        return std::min(k, hi);
    };
    size_t x0 = sat(x, W - 1), y0 = sat(y, H - 1);
    double ax = x - (double)x0, ay = y - (double)y0;
    auto L = [&](size_t i) {
        const float* p = rgb + 3 * (i % len);
        return 0.2126 * p[0] + 0.7152 * p[1] + 0.0722 * p[2];
    };
    double a = L(y0 * W + x0) * (1 - ax) + L(y0 * W + x0 + 1) * ax;
    double b = L((y0 + 1) * W + x0) * (1 - ax) + L((y0 + 1) * W + x0 + 1) * ax;
    return a * (1 - ay) + b * ay;
}

void weights(const float* rgb, unsigned W, unsigned H, std::vector<double>& w) {
    w.resize((size_t)W * H);
    double mean = 0;
    for (unsigned r = 0; r < H; ++r) {
        double yc = (double)r - 0.5;
        double theta = M_PI * ((double)H - 1.0 - yc) / (double)H;
        double st = std::sin(theta);
        for (unsigned i = 0; i < W; ++i) {
            double l = std::max(0.0, lum_at(rgb, W, H, i + 0.5, yc));
            if (!std::isfinite(l)) l = 0;
            w[(size_t)r * W + i] = l * st;
            mean += l * st;
        }
    }
    mean /= (double)W * H;
    double floor_w = std::max(mean * 1e-3, 1e-12);
    for (auto& x : w) x += floor_w;
}
}  // namespace

void env_tables_fp64(const float* rgb, unsigned W, unsigned H, std::vector<double>& marg, std::vector<double>& cond) {
    std::vector<double> w;
    weights(rgb, W, H, w);
    marg.assign(H + 1, 0.0);
    cond.assign((size_t)H * (W + 1), 0.0);
    for (unsigned r = 0; r < H; ++r) {
        double* c = &cond[(size_t)r * (W + 1)];
        for (unsigned i = 0; i < W; ++i) c[i + 1] = c[i] + w[(size_t)r * W + i];
        marg[r + 1] = marg[r] + c[W];
    }
}

double build_env_tables(const float* rgb, unsigned W, unsigned H, std::vector<float>& marg, std::vector<float>& cond,
                        std::vector<float>& pdf) {
    std::vector<double> m64, c64, w;
    env_tables_fp64(rgb, W, H, m64, c64);
    weights(rgb, W, H, w);
    double total = m64[H];
    marg.resize(H + 1);
    cond.resize((size_t)H * (W + 1));
    pdf.resize((size_t)W * H);
    for (unsigned r = 0; r <= H; ++r) marg[r] = (float)(m64[r] / total);
    marg[H] = 1.0f;
    for (unsigned r = 0; r < H; ++r) {
        const double* c = &c64[(size_t)r * (W + 1)];
        float* o = &cond[(size_t)r * (W + 1)];
        for (unsigned i = 0; i <= W; ++i) o[i] = (float)(c[i] / c[W]);
        o[W] = 1.0f;
        for (unsigned i = 0; i < W; ++i) pdf[(size_t)r * W + i] = (float)(w[(size_t)r * W + i] / total);
    }
    return total;
}

uint32_t guide_size(uint32_t n) {
    uint32_t g = 1;
    while (g < n) g <<= 1;
    return g;
}

void build_guide(const float* cdf, uint32_t n, uint32_t G, uint32_t* guide) {
    uint32_t k = 0;
    for (uint32_t j = 0; j <= G; ++j) {
        const float t = (float)j / (float)G;  // exact: G is a power of two
        while (k + 1 < n && cdf[k + 1] <= t) ++k;
        guide[j] = k;
    }
}

}  // namespace vr
