// Test harness for voidray_amd/csrc/f32math.h: evaluates its functions over
// an array on the host and on the device (the same header compiled for both),
// and, for comparison, the host's libm f32 functions (glibc's expf, logf,
// atanf, sinf, cosf: what the reference's f32::exp etc. call) and the f64
// libm functions rounded once to f32.
//
// fn: 0 exp, 1 ln, 2 atan, 3 sin, 4 cos.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <mutex>
#include <thread>
#include <vector>

#include "../../voidray_amd/csrc/f32math.h"

namespace {

__host__ __device__ float eval_cr(int fn, float x) {
    using namespace vr::f32m;
    float s, c;
    switch (fn) {
        case 0: return exp_cr(x);
        case 1: return ln_cr(x);
        case 2: return atan_cr(x);
        case 3: sincos_cr(x, s, c); return s;
        default: sincos_cr(x, s, c); return c;
    }
}

__global__ void k_pow(const float* x, const float* y, float* z, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = vr::f32m::pow_cr(x[i], y[i]);
}

__global__ void k_eval(int fn, const float* x, float* y, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = eval_cr(fn, x[i]);
}

bool valid(int fn, const float* x, const float* y) { return fn >= 0 && fn <= 4 && x && y; }

float eval_libm(int fn, float v) {
    return fn == 0 ? expf(v) : fn == 1 ? logf(v) : fn == 2 ? atanf(v) : fn == 3 ? sinf(v) : cosf(v);
}

float eval_f64(int fn, float x) {
    const double v = x;
    return (float)(fn == 0 ? exp(v) : fn == 1 ? log(v) : fn == 2 ? atan(v) : fn == 3 ? sin(v) : cos(v));
}

bool same(float a, float b) {
    uint32_t ua, ub;
    memcpy(&ua, &a, 4);
    memcpy(&ub, &b, 4);
    return ua == ub || (a != a && b != b);
}

}  // namespace

extern "C" {

int f32m_eval_host(int fn, const float* x, float* y, uint64_t n) {
    if (!valid(fn, x, y)) return -1;
    for (uint64_t i = 0; i < n; ++i) y[i] = eval_cr(fn, x[i]);
    return 0;
}

int f32m_eval_libm(int fn, const float* x, float* y, uint64_t n) {
    if (!valid(fn, x, y)) return -1;
    for (uint64_t i = 0; i < n; ++i) y[i] = eval_libm(fn, x[i]);
    return 0;
}

int f32m_eval_f64(int fn, const float* x, float* y, uint64_t n) {
    if (!valid(fn, x, y)) return -1;
    for (uint64_t i = 0; i < n; ++i) y[i] = eval_f64(fn, x[i]);
    return 0;
}

// Copies x to the device, evaluates there, copies the results back.
// Returns 0 or the HIP error code.
int f32m_eval_device(int fn, const float* x, float* y, uint64_t n) {
    if (!valid(fn, x, y)) return -1;
    if (n == 0) return 0;
    float *dx = nullptr, *dy = nullptr;
    hipError_t e = hipMalloc(&dx, n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dy, n * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(dx, x, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const unsigned blocks = (unsigned)((n + 255) / 256);
        hipLaunchKernelGGL(k_eval, dim3(blocks), dim3(256), 0, 0, fn, dx, dy, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(y, dy, n * sizeof(float), hipMemcpyDeviceToHost);
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    return (int)e;
}

// Every float whose bit pattern lies in [lo, hi] (host, `threads` threads):
// counts[0] = inputs, counts[1] = results that differ from the f64 function
// rounded once, counts[2] = results that differ from the f32 libm function;
// the first `cap` inputs of counts[1] are written to `out` (in no fixed order).
int f32m_sweep(int fn, uint32_t lo, uint32_t hi, int threads, uint64_t* counts, uint32_t* out, uint64_t cap) {
    if (fn < 0 || fn > 4 || lo > hi || threads < 1 || !counts) return -1;
    std::mutex mu;
    counts[0] = counts[1] = counts[2] = 0;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        pool.emplace_back([&, t] {
            uint64_t n = 0, dlibm = 0;
            std::vector<uint32_t> bad;
            for (uint64_t u = (uint64_t)lo + t; u <= hi; u += (uint64_t)threads) {
                float x;
                const uint32_t b = (uint32_t)u;
                memcpy(&x, &b, 4);
                const float y = eval_cr(fn, x);
                ++n;
                if (!same(y, eval_f64(fn, x))) bad.push_back(b);
                if (!same(y, eval_libm(fn, x))) ++dlibm;
            }
            std::lock_guard<std::mutex> lock(mu);
            for (uint32_t b : bad)
                if (out && counts[1] < cap) out[counts[1]++] = b;
                else ++counts[1];
            counts[0] += n;
            counts[2] += dlibm;
        });
    }
    for (auto& th : pool) th.join();
    return 0;
}

// pow over pairs: where 0 = the routine on the host, 1 = on the device,
// 2 = f64 pow rounded once, 3 = the f32 libm powf.  Returns 0 or an error.
int f32m_pow(int where, const float* x, const float* y, float* z, uint64_t n) {
    if (where < 0 || where > 3 || !x || !y || !z) return -1;
    if (where == 1) {
        if (n == 0) return 0;
        float *dx = nullptr, *dy = nullptr, *dz = nullptr;
        hipError_t e = hipMalloc(&dx, n * sizeof(float));
        if (e == hipSuccess) e = hipMalloc(&dy, n * sizeof(float));
        if (e == hipSuccess) e = hipMalloc(&dz, n * sizeof(float));
        if (e == hipSuccess) e = hipMemcpy(dx, x, n * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dy, y, n * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_pow, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dx, dy, dz, n);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpy(z, dz, n * sizeof(float), hipMemcpyDeviceToHost);
        if (dx) (void)hipFree(dx);
        if (dy) (void)hipFree(dy);
        if (dz) (void)hipFree(dz);
        return (int)e;
    }
    for (uint64_t i = 0; i < n; ++i)
        z[i] = where == 0 ? vr::f32m::pow_cr(x[i], y[i])
               : where == 2 ? (float)pow((double)x[i], (double)y[i])
                            : powf(x[i], y[i]);
    return 0;
}

}  // extern "C"
