// am_mlp.hip - learned warm start: per-contact features, a 24-64-64-64-3 ReLU
// MLP with BF16 weights on the CDNA matrix cores (v_mfma_f32_16x16x16_bf16,
// FP32 accumulation), its backward pass and an Adam update, and the
// start-quality residual that decides between the learned and the plain start.
//
// Feature vector (24 FP32, dimensionless), per contact with solve bodies a, b:
//   0-2 normal, 3-5 tangent 1 (world; tangent 2 = n x t1 adds nothing),
//   6 1 when the plain warm start found no previous contact (new this step),
//   7-8 0,
//   9-11 constraint-space velocity J (v_a, w_a, v_b, w_b) / (g dt),
//   12 violation / (g dt^2),
//   13 minv_a / (minv_a + minv_b), 14 log10((minv_a + minv_b) * 1 kg) / 3,
//   15-17 diag(jinv_a) / trace(jinv_a), 18-20 same for b (0 when pinned),
//   21-23 slog(previous-frame lambda / S),
// with S = g dt / D_nn (D_nn = J_n M^-1 J_n^T) and slog(x) = sign(x) ln(1 + |x|),
// which keeps the impulses of a 100:1 mass-ratio pile in a range BF16 holds.
// The network predicts the change of slog(lambda / S) from the plain start,
// bounded by kMaxLogStep.
#include "am_mlp.hpp"

namespace am {

typedef short short4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// features
AMD_INLINE double slog(double x) { return x < 0 ? -log1p(-x) : log1p(x); }
AMD_INLINE double sexp(double z) { return z < 0 ? -expm1(-z) : expm1(z); }
constexpr double kMaxLogStep = 2.0;  // largest change of slog(lambda / S) a prediction makes

__global__ void k_features(FeatureArgs F) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= F.n) return;
    const BodiesView& B = F.B;
    const ContactsView& C = F.C;
    int b1 = C.b1[k], b2 = C.b2[k];
    int a = B.parent[b1] >= 0 ? B.parent[b1] : b1;
    int b = B.parent[b2] >= 0 ? B.parent[b2] : b2;
    const double* J = &C.J[(size_t)k * 36];
    float* f = &F.out[(size_t)k * kMlpIn];
    double gdt = F.g * F.dt;
    // frame (jb.v of the normal and first tangent rows)
    for (int r = 0; r < 2; r++)
        for (int i = 0; i < 3; i++) f[r * 3 + i] = (float)J[r * 12 + 6 + i];
    f[6] = C.is_new[k] ? 1.f : 0.f;
    f[7] = 0.f;
    f[8] = 0.f;
    d3 va = ld3(&B.v[a * 3]), wa = ld3(&B.w[a * 3]), vb = ld3(&B.v[b * 3]), wb = ld3(&B.w[b * 3]);
    for (int r = 0; r < 3; r++) {
        const double* row = J + r * 12;
        double u = dot(ld3(row), va) + dot(ld3(row + 3), wa) + dot(ld3(row + 6), vb) + dot(ld3(row + 9), wb);
        f[9 + r] = (float)(u / gdt);
    }
    f[12] = (float)(C.viol[k] / (gdt * F.dt));
    double ma = B.minv[a], mb = B.minv[b];
    double msum = ma + mb;
    f[13] = (float)(msum > 0 ? ma / msum : 0.5);
    f[14] = (float)(log10(msum + 1e-12) / 3.0);
    const int ids[2] = {a, b};
    for (int s = 0; s < 2; s++) {
        const double* Ji = &B.jinv[ids[s] * 9];
        double tr = Ji[0] + Ji[4] + Ji[8];
        for (int i = 0; i < 3; i++) f[15 + s * 3 + i] = (float)(tr > 0 ? Ji[i * 4] / tr : 0.0);
    }
    // D_nn (Contact.computeJMinvJt normal row)
    d3 t1 = mul(ldm(&B.jinv[a * 9]), ld3(J + 3));
    d3 t2 = mul(ldm(&B.jinv[b * 9]), ld3(J + 9));
    double D = ma * dot(ld3(J), ld3(J)) + dot(ld3(J + 3), t1) + mb * dot(ld3(J + 6), ld3(J + 6)) + dot(ld3(J + 9), t2);
    double S = D > 0 ? gdt / D : 1.0;
    F.scale[k] = S;
    for (int r = 0; r < 3; r++) f[21 + r] = (float)slog(C.lam[k * 3 + r] / S);
}

void launch_features(hipStream_t s, const FeatureArgs& F) {
    if (F.n > 0) hipLaunchKernelGGL(k_features, dim3((F.n + 127) / 128), dim3(128), 0, s, F);
}

// ---------------------------------------------------------------------------
// BF16 helpers (round to nearest even)
AMD_INLINE unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
AMD_INLINE float bf2f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

// Layer geometry: padded input 32, hidden 64, padded output 16.
constexpr int kIn = 32, kH = 64, kOut = 16;
// bf16 weights in LDS, row-major [in][out]
struct MlpLds {
    unsigned short w1[kIn * kH];
    unsigned short w2[kH * kH];
    unsigned short w3[kH * kH];
    unsigned short w4[kH * kOut];
    float b1[kH], b2[kH], b3[kH], b4[kOut];
};

// One 16x16 output tile C += A(16 x K) * B(K x 16) with v_mfma_f32_16x16x16_bf16.
// Lane l holds A[l%16][k0+4(l/16)+j], B[k0+4(l/16)+j][l%16] and returns
// C[4(l/16)+j][l%16] for j = 0..3.  A and B are read from LDS through strides
// (sa_r, sa_k) and (sb_k, sb_n), so transposed operands are just strides.
AMD_INLINE float4v mfma_tile(const unsigned short* A, int sa_r, int sa_k, const unsigned short* Bm, int sb_k,
                             int sb_n, int K, float4v c) {
    const int l = threadIdx.x & 63;
    const int row = l & 15, kq = (l >> 4) * 4;
    for (int k0 = 0; k0 < K; k0 += 16) {
        short4v a, b;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            a[j] = (short)A[row * sa_r + (k0 + kq + j) * sa_k];
            b[j] = (short)Bm[(k0 + kq + j) * sb_k + row * sb_n];
        }
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
    return c;
}

AMD_INLINE void load_weights(const MlpWeights& W, MlpLds& L) {
    for (int i = threadIdx.x; i < kIn * kH; i += blockDim.x) L.w1[i] = W.w1_bf[i];
    for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) {
        L.w2[i] = W.w2_bf[i];
        L.w3[i] = W.w3_bf[i];
    }
    for (int i = threadIdx.x; i < kH * kOut; i += blockDim.x) L.w4[i] = W.w4_bf[i];
    for (int i = threadIdx.x; i < kH; i += blockDim.x) {
        L.b1[i] = W.b1[i];
        L.b2[i] = W.b2[i];
        L.b3[i] = W.b3[i];
    }
    for (int i = threadIdx.x; i < kOut; i += blockDim.x) L.b4[i] = W.b4[i];
    __syncthreads();
}

// Per-wave activation buffers: x (16 x 32), h1..h3 (16 x 64), all bf16.
struct WaveAct {
    unsigned short x[16 * kIn];
    unsigned short h1[16 * kH];
    unsigned short h2[16 * kH];
    unsigned short h3[16 * kH];
    unsigned short dy[16 * kOut];
    unsigned short d3_[16 * kH];
    unsigned short d2_[16 * kH];
    unsigned short d1_[16 * kH];
};

constexpr int kMlpWaves = 4;

// dense layer: out(16 x 64) = relu(in(16 x K) W(K x 64) + b), written as bf16
AMD_INLINE void dense_relu(const unsigned short* in, int K, const unsigned short* W, const float* bias,
                           unsigned short* out) {
    const int l = threadIdx.x & 63;
    for (int t = 0; t < kH / 16; t++) {
        float4v c = {0, 0, 0, 0};
        c = mfma_tile(in, K, 1, W + t * 16, kH, 1, K, c);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int r = (l >> 4) * 4 + j, col = t * 16 + (l & 15);
            float v = c[j] + bias[col];
            out[r * kH + col] = f2bf(v > 0.f ? v : 0.f);
        }
    }
}

AMD_INLINE void load_x_tile(const float* X, int row0, int n, unsigned short* x) {
    const int l = threadIdx.x & 63;
    for (int i = l; i < 16 * kIn; i += 64) {
        int r = i / kIn, cidx = i % kIn;
        float v = (row0 + r < n && cidx < kMlpIn) ? X[(size_t)(row0 + r) * kMlpIn + cidx] : 0.f;
        x[i] = f2bf(v);
    }
}

// forward: Y[n x 3] = mlp(X[n x 24])
__global__ __launch_bounds__(kMlpWaves * 64) void k_mlp_forward(MlpWeights W, const float* X, int n, float* Y) {
    __shared__ MlpLds L;
    __shared__ WaveAct act[kMlpWaves];
    load_weights(W, L);
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    WaveAct& A = act[wave];
    const int tiles = (n + 15) / 16;
    for (int tile = blockIdx.x * kMlpWaves + wave; tile < tiles; tile += gridDim.x * kMlpWaves) {
        const int row0 = tile * 16;
        load_x_tile(X, row0, n, A.x);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        dense_relu(A.x, kIn, L.w1, L.b1, A.h1);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        dense_relu(A.h1, kH, L.w2, L.b2, A.h2);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        dense_relu(A.h2, kH, L.w3, L.b3, A.h3);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        float4v c = {0, 0, 0, 0};
        c = mfma_tile(A.h3, kH, 1, L.w4, kOut, 1, kH, c);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int r = (l >> 4) * 4 + j, col = l & 15;
            if (col < kMlpOut && row0 + r < n) Y[(size_t)(row0 + r) * kMlpOut + col] = c[j] + L.b4[col];
        }
        __builtin_amdgcn_wave_barrier();
    }
}

void launch_mlp_forward(hipStream_t s, const MlpWeights& W, const float* X, int n, float* Y) {
    if (n <= 0) return;
    int tiles = (n + 15) / 16;
    int blocks = std::min((tiles + kMlpWaves - 1) / kMlpWaves, 1024);
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(kMlpWaves * 64), 0, s, W, X, n, Y);
}

// ---------------------------------------------------------------------------
// backward: loss = sum over rows of |y - t|^2 / n_total; gradients of every
// weight and bias accumulated (FP32 atomics) into G.  rows = idx[0..n) of X/T.
AMD_INLINE void grad_tile_atomic(const float4v& c, float* G, int ld, int r0, int c0, int rows, int cols) {
    const int l = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        int r = r0 + (l >> 4) * 4 + j, col = c0 + (l & 15);
        if (r < rows && col < cols) atomicAdd(&G[r * ld + col], c[j]);
    }
}

// dh (16 x 64) = (dnext (16 x kn) W^T) * relu'(h), W is [64][wn]; then
// dW (kin x 64) += in^T dh and db += column sums of dh.
AMD_INLINE void back_layer(const unsigned short* dnext, int kn, const unsigned short* W, int wn,
                           const unsigned short* h, const unsigned short* in, int kin, unsigned short* dh, float* gw,
                           float* gb, int gw_rows, int row0, int n) {
    const int l = threadIdx.x & 63;
    for (int t = 0; t < kH / 16; t++) {
        float4v c = {0, 0, 0, 0};
        c = mfma_tile(dnext, kn, 1, W + t * 16 * wn, 1, wn, kn, c);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int r = (l >> 4) * 4 + j, col = t * 16 + (l & 15);
            float hv = bf2f(h[r * kH + col]);
            float g = hv > 0.f ? c[j] : 0.f;
            dh[r * kH + col] = f2bf(g);
            if (row0 + r < n) atomicAdd(&gb[col], g);
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    for (int m = 0; m < kin / 16; m++)
        for (int t = 0; t < kH / 16; t++) {
            float4v c = {0, 0, 0, 0};
            c = mfma_tile(in + m * 16, 1, kin, dh + t * 16, kH, 1, 16, c);
            grad_tile_atomic(c, gw, kH, m * 16, t * 16, gw_rows, kH);
        }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(kMlpWaves * 64) void k_mlp_backward(MlpWeights W, const float* X, const float* T,
                                                                 const int* idx, int n, float inv_total,
                                                                 MlpGrads G, float* loss) {
    __shared__ MlpLds L;
    __shared__ WaveAct act[kMlpWaves];
    load_weights(W, L);
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    WaveAct& A = act[wave];
    const int tiles = (n + 15) / 16;
    float my_loss = 0.f;
    for (int tile = blockIdx.x * kMlpWaves + wave; tile < tiles; tile += gridDim.x * kMlpWaves) {
        const int row0 = tile * 16;
        for (int i = l; i < 16 * kIn; i += 64) {
            int r = i / kIn, cidx = i % kIn;
            float v = 0.f;
            if (row0 + r < n && cidx < kMlpIn) v = X[(size_t)idx[row0 + r] * kMlpIn + cidx];
            A.x[i] = f2bf(v);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        dense_relu(A.x, kIn, L.w1, L.b1, A.h1);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        dense_relu(A.h1, kH, L.w2, L.b2, A.h2);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        dense_relu(A.h2, kH, L.w3, L.b3, A.h3);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // output layer and dL/dy = 2 (y - t) / n_total
        {
            float4v c = {0, 0, 0, 0};
            c = mfma_tile(A.h3, kH, 1, L.w4, kOut, 1, kH, c);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                int r = (l >> 4) * 4 + j, col = l & 15;
                float g = 0.f;
                if (col < kMlpOut && row0 + r < n) {
                    float y = c[j] + L.b4[col];
                    float t = T[(size_t)idx[row0 + r] * kMlpOut + col];
                    float e = y - t;
                    my_loss += e * e * inv_total;
                    g = 2.f * e * inv_total;
                    atomicAdd(&G.b4[col], g);
                }
                A.dy[r * kOut + col] = f2bf(g);
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // dW4 (64 x 16) += h3^T (64 x 16 rows) dy (16 x 16)
        for (int m = 0; m < kH / 16; m++) {
            float4v c = {0, 0, 0, 0};
            c = mfma_tile(A.h3 + m * 16, 1, kH, A.dy, kOut, 1, 16, c);
            grad_tile_atomic(c, G.w4, kOut, m * 16, 0, kH, kMlpOut);
        }
        // back through layers 3, 2, 1: dh = (d_next W^T) * relu'(h)
        back_layer(A.dy, kOut, L.w4, kOut, A.h3, A.h2, kH, A.d3_, G.w3, G.b3, kH, row0, n);
        back_layer(A.d3_, kH, L.w3, kH, A.h2, A.h1, kH, A.d2_, G.w2, G.b2, kH, row0, n);
        back_layer(A.d2_, kH, L.w2, kH, A.h1, A.x, kIn, A.d1_, G.w1, G.b1, kMlpIn, row0, n);
        __builtin_amdgcn_wave_barrier();
    }
    // loss reduction
    for (int o = 32; o > 0; o >>= 1) my_loss += __shfl_xor(my_loss, o, 64);
    if (l == 0 && loss) atomicAdd(loss, my_loss);
}

void launch_mlp_backward(hipStream_t s, const MlpWeights& W, const float* X, const float* T, const int* idx, int n,
                         float inv_total, const MlpGrads& G, float* loss) {
    if (n <= 0) return;
    int tiles = (n + 15) / 16;
    int blocks = std::min((tiles + kMlpWaves - 1) / kMlpWaves, 512);
    hipLaunchKernelGGL(k_mlp_backward, dim3(blocks), dim3(kMlpWaves * 64), 0, s, W, X, T, idx, n, inv_total, G, loss);
}

// ---------------------------------------------------------------------------
// Adam on the FP32 master weights, then BF16 copies; gradients cleared.
__global__ void k_adam(float* p, float* g, float* m, float* v, unsigned short* bf, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2, const unsigned char* mask) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    g[i] = 0.f;
    if (mask && !mask[i]) {
        if (bf) bf[i] = f2bf(p[i]);
        return;
    }
    float mi = b1 * m[i] + (1.f - b1) * gi;
    float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (bf) bf[i] = f2bf(pi);
}

void launch_adam(hipStream_t s, float* p, float* g, float* m, float* v, unsigned short* bf, int n, float lr, float b1,
                 float b2, float eps, int step, const unsigned char* mask) {
    if (n <= 0) return;
    float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, p, g, m, v, bf, n, lr, b1, b2, eps, c1, c2,
                       mask);
}

__global__ void k_to_bf16(const float* p, unsigned short* bf, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bf[i] = f2bf(p[i]);
}
void launch_to_bf16(hipStream_t s, const float* p, unsigned short* bf, int n) {
    if (n > 0) hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, s, p, bf, n);
}

// ---------------------------------------------------------------------------
// Predicted start: for a contact the plain warm start found no previous
// contact for (new this step, lambda_plain = 0), the network predicts the
// change y of slog(lambda / S), clamped to +-kMaxLogStep, so lambda =
// S sexp(slog(lambda_plain / S) + y), projected onto the friction box
// (lambda_n >= 0, |lambda_t| <= mu lambda_n).  A contact that persists keeps
// its previous-frame lambda: on a resting pile that start is already close
// to converged, closer than the network's per-contact accuracy.
AMD_INLINE double predicted(double lam_plain, double S, float y) {
    const double z = slog(lam_plain / S) + fmin(fmax((double)y, -kMaxLogStep), kMaxLogStep);
    return S * sexp(z);
}
__global__ void k_apply_prediction(ContactsView C, const BodiesView B, const float* Y, const double* scale, int n,
                                   int fric_ovr, double fric_val, double* lam_out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (!C.is_new[k]) {
#pragma unroll
        for (int r = 0; r < 3; r++) lam_out[k * 3 + r] = C.lam[k * 3 + r];
        return;
    }
    double S = scale[k];
    double ln = fmax(0.0, predicted(C.lam[k * 3 + 0], S, Y[k * 3 + 0]));
    double mu;
    if (fric_ovr) {
        mu = fric_val;
    } else {
        double f1 = B.desc[C.b1[k]].friction, f2 = B.desc[C.b2[k]].friction;
        if (f1 < 0.2 || f2 < 0.2) mu = fmin(f1, f2);
        else if (f1 > 1. || f2 > 1.) mu = fmax(f1, f2);
        else mu = (f1 + f2) / 2.;
    }
    double lim = mu * ln;
    double l1 = fmin(fmax(predicted(C.lam[k * 3 + 1], S, Y[k * 3 + 1]), -lim), lim);
    double l2 = fmin(fmax(predicted(C.lam[k * 3 + 2], S, Y[k * 3 + 2]), -lim), lim);
    lam_out[k * 3 + 0] = ln;
    lam_out[k * 3 + 1] = l1;
    lam_out[k * 3 + 2] = l2;
}

void launch_apply_prediction(hipStream_t s, ContactsView C, BodiesView B, const float* Y, const double* scale, int n,
                             int fric_ovr, double fric_val, double* lam_out) {
    if (n > 0)
        hipLaunchKernelGGL(k_apply_prediction, dim3((n + 255) / 256), dim3(256), 0, s, C, B, Y, scale, n, fric_ovr,
                           fric_val, lam_out);
}

// Targets: the change of slog(lambda / S) from the plain start to the
// converged solve (features 21..23 hold the plain start's), clamped as the
// prediction is.
__global__ void k_targets(const double* lam, const double* scale, const float* X, int n, float* T) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    for (int r = 0; r < 3; r++) {
        const double d = slog(lam[k * 3 + r] / scale[k]) - (double)X[(size_t)k * kMlpIn + 21 + r];
        T[k * 3 + r] = (float)fmin(fmax(d, -kMaxLogStep), kMaxLogStep);
    }
}
void launch_targets(hipStream_t s, const double* lam, const double* scale, const float* X, int n, float* T) {
    if (n > 0) hipLaunchKernelGGL(k_targets, dim3((n + 255) / 256), dim3(256), 0, s, lam, scale, X, n, T);
}

// Start quality: with deltaV = M^-1 J^T lambda already applied (warm-start
// pass of the sweep kernel), the largest change one projected Gauss-Seidel
// row update would make, evaluated for all rows at once.
__global__ void k_start_residual(const double* rec, const int* meta, const double* dv, const double* minv,
                                 const double* jinv, int n, double omega, double compliance,
                                 unsigned long long* out) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    double r_max = 0;
    if (k < n) {
        const double* R = &rec[(size_t)k * kRecDoubles];
        int4 m = reinterpret_cast<const int4*>(meta)[k];
        const double* da = &dv[m.x * 6];
        const double* db = &dv[m.y * 6];
        double lam[3] = {R[kRecLambda], R[kRecLambda + 1], R[kRecLambda + 2]};
        for (int r = 0; r < 3; r++) {
            const double* q = R + r * kRecRow;
            double jd = (-q[3] * da[0] + -q[4] * da[1] + -q[5] * da[2]) + (q[0] * da[3] + q[1] * da[4] + q[2] * da[5]) +
                        (q[3] * db[0] + q[4] * db[1] + q[5] * db[2]) + (q[6] * db[3] + q[7] * db[4] + q[8] * db[5]);
            double D = R[kRecD + r];
            double l = (D * lam[r] - omega * (R[kRecB + r] + jd)) / (D + compliance);
            if (m.z) {
                if (r == 0) l = fmax(0.0, l);
                else l = fmin(fmax(l, -R[kRecMu] * lam[0]), R[kRecMu] * lam[0]);
            }
            r_max = fmax(r_max, fabs(l - lam[r]));
        }
    }
    for (int o = 32; o > 0; o >>= 1) r_max = fmax(r_max, __shfl_xor(r_max, o, 64));
    if ((threadIdx.x & 63) == 0 && r_max > 0) atomicMax(out, (unsigned long long)__double_as_longlong(r_max));
    (void)minv;
    (void)jinv;
}

void launch_start_residual(hipStream_t s, const double* rec, const int* meta, const double* dv, const double* minv,
                           const double* jinv, int n, double omega, double compliance, unsigned long long* out) {
    if (n > 0)
        hipLaunchKernelGGL(k_start_residual, dim3((n + 255) / 256), dim3(256), 0, s, rec, meta, dv, minv, jinv, n,
                           omega, compliance, out);
}

// lambda of the records (sweep order) <-> a list-order lambda array
__global__ void k_rec_lambda(double* rec, const int* order, const double* lam, int n, int to_rec) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int c = order[k];
    double* R = &rec[(size_t)k * kRecDoubles + kRecLambda];
    for (int r = 0; r < 3; r++) {
        if (to_rec) R[r] = lam[c * 3 + r];
    }
}
void launch_rec_lambda(hipStream_t s, double* rec, const int* order, const double* lam, int n) {
    if (n > 0) hipLaunchKernelGGL(k_rec_lambda, dim3((n + 255) / 256), dim3(256), 0, s, rec, order, lam, n, 1);
}

}  // namespace am
