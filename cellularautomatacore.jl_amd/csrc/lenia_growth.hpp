// lenia_growth.hpp -- local_rule(::Lenia) (src/lenia.jl:42-47) on one cell.
#pragma once
#include "device.hpp"

// G(U) = 2 exp(-(U - mu)^2 / (2 sigma^2)) - 1 and clamp(cell + dt G(U), 0, 1), in
// FP64 for Lenia() (then rounded to Float32) and in FP32 for
// Lenia(; precision=Float32).
CAH_DEV float lenia_new64(float c, double U, double mu, double sigma, double dt) {
    const double d = U - mu;
    const double g = 2.0 * exp(-(d * d) / (2.0 * (sigma * sigma))) - 1.0;
    double v = (double)c + dt * g;
    v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
    return (float)v;
}

CAH_DEV float lenia_new32(float c, float U, float mu, float sigma, float dt) {
    const float d = U - mu;
    const float g = 2.0f * expf(-(d * d) / (2.0f * (sigma * sigma))) - 1.0f;
    float v = c + dt * g;
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    return v;
}

// The growth of the MFMA field forms, in FP32 with the hardware exponential and
// the scale nk = -1 / (2 sigma^2) made once per launch.  Its error (below 1e-7
// per step) is far below the error of the FP16/BF16 field it is applied to.
CAH_DEV float lenia_fast32(float c, float U, float mu, float nk, float dt) {
    const float d = U - mu;
    const float g = 2.0f * __expf(d * d * nk) - 1.0f;
    const float v = c + dt * g;
    return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
}
