// Plain-data types of the learned surrogate shared by the context and the kernels.
#pragma once

namespace mappdmd {

constexpr int kSurOut = 6;
// r bins of the residual envelope: the networks learn (quadrature - baseline) / s(r),
// with log2 s(r) interpolated linearly between bin centres
constexpr int kEnvBins = 32;

// trained box and standardisation of the surrogate networks
struct SurNorm {
    float r_lo, r_hi, a_lo, a_hi;  // trained box
    float r_c, r_s, a_c, a_s;      // input standardisation
    float y_mean[4][kSurOut], y_std[4][kSurOut];
};

}  // namespace mappdmd
