// am_mlp.hpp - argument blocks and launchers of the learned warm start.
#pragma once

#include <algorithm>

#include "am_internal.hpp"

namespace am {

constexpr int kMlpIn = 24;
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 3;
// flat FP32 parameter vector: W1 (24x64) b1 (64) W2 (64x64) b2 W3 (64x64) b3 W4 (64x3) b4 (3)
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + 2 * (kMlpHidden * kMlpHidden + kMlpHidden) +
                           kMlpHidden * kMlpOut + kMlpOut;

struct FeatureArgs {
    BodiesView B;
    ContactsView C;
    int n;
    double g;
    double dt;
    float* out;     // n x 24
    double* scale;  // n
};

// Device weights: padded BF16 copies (row-major [in][out]: 32x64, 64x64, 64x64,
// 64x16) and FP32 biases (padded to 64 / 16).
struct MlpWeights {
    const unsigned short* w1_bf;
    const unsigned short* w2_bf;
    const unsigned short* w3_bf;
    const unsigned short* w4_bf;
    const float* b1;
    const float* b2;
    const float* b3;
    const float* b4;
};

// FP32 gradients in the same padded layout.
struct MlpGrads {
    float* w1;
    float* w2;
    float* w3;
    float* w4;
    float* b1;
    float* b2;
    float* b3;
    float* b4;
};

// padded sizes
constexpr int kPadW1 = 32 * 64, kPadW2 = 64 * 64, kPadW3 = 64 * 64, kPadW4 = 64 * 16;
constexpr int kPadB1 = 64, kPadB2 = 64, kPadB3 = 64, kPadB4 = 16;
constexpr int kPadTotal = kPadW1 + kPadW2 + kPadW3 + kPadW4 + kPadB1 + kPadB2 + kPadB3 + kPadB4;

void launch_features(hipStream_t s, const FeatureArgs& F);
void launch_mlp_forward(hipStream_t s, const MlpWeights& W, const float* X, int n, float* Y);
void launch_mlp_backward(hipStream_t s, const MlpWeights& W, const float* X, const float* T, const int* idx, int n,
                         float inv_total, const MlpGrads& G, float* loss);
void launch_adam(hipStream_t s, float* p, float* g, float* m, float* v, unsigned short* bf, int n, float lr, float b1,
                 float b2, float eps, int step, const unsigned char* mask);
void launch_to_bf16(hipStream_t s, const float* p, unsigned short* bf, int n);
void launch_apply_prediction(hipStream_t s, ContactsView C, BodiesView B, const float* Y, const double* scale, int n,
                             int fric_ovr, double fric_val, double* lam_out);
void launch_targets(hipStream_t s, const double* lam, const double* scale, const float* X, int n, float* T);
void launch_start_residual(hipStream_t s, const double* rec, const int* meta, const double* dv, const double* minv,
                           const double* jinv, int n, double omega, double compliance, unsigned long long* out);
void launch_rec_lambda(hipStream_t s, double* rec, const int* order, const double* lam, int n);

}  // namespace am
