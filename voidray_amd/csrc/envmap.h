// Environment importance-sampling tables (no counterpart in the reference,
// whose HDRIEnvironment is lookup-only: voidray_common/src/environments.rs:80-90).
//
// Sampling domain: the continuous lookup coordinates of HDRIEnvironment::sample,
// x = phi / (2 pi) * W in [0, W) and y = (H - 1) - theta / pi * H in [-1, H - 1).
// Cell (i, r) covers x in [i, i+1), y in [r-1, r); its weight is
// luminance(lookup at the cell centre) * sin(theta_centre) + floor, so the pdf is
// positive everywhere.  Prefix sums are FP64; the device gets f32 copies of the
// normalised CDFs and the per-cell pdf (in the xy measure).
#pragma once
#include <cstdint>
#include <vector>

namespace vr {
// Returns the FP64 total weight.  marg: H+1 entries; cond: H*(W+1); pdf: H*W.
double build_env_tables(const float* rgb, unsigned W, unsigned H, std::vector<float>& marg, std::vector<float>& cond,
                        std::vector<float>& pdf);
// Guide table of a CDF with n cells (cdf[0] = 0, non-decreasing): for a power
// of two G >= n, guide[j] (j = 0..G) is the largest k < n with cdf[k] <= j / G.
// A sample u in [j/G, (j+1)/G) then lies in cells guide[j]..guide[j+1], so the
// binary search over that range returns the same cell as over the whole CDF.
uint32_t guide_size(uint32_t n);
void build_guide(const float* cdf, uint32_t n, uint32_t G, uint32_t* guide);
// FP64 reference CDFs (unnormalised prefix sums) for error reporting.
void env_tables_fp64(const float* rgb, unsigned W, unsigned H, std::vector<double>& marg, std::vector<double>& cond);
}  // namespace vr
