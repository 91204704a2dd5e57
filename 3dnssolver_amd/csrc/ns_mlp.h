// ns_mlp.h -- the learned WENO-5 weights on the GPU: an 8-32-32-3 MLP (GELU, softmax) evaluated
// by one wavefront for 64 samples at a time with BF16 matrix instructions
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation).  Shared by the face-flux kernels (inference
// inside the reconstruction) and the training kernels (ns_mlp.hip).  No reference counterpart.
//
// Matrix layout of v_mfma_f32_16x16x32_bf16 (D = A B + C, A 16x32, B 32x16): lane l, with
// g = l / 16 and c = l % 16, holds A[c][8g .. 8g+7], B[8g .. 8g+7][c] and D[4g .. 4g+3][c].
//
// Samples are the columns (N): 64 samples = 4 column tiles of 16.  Hidden units are the rows.
// The D fragment of a layer, two row tiles of 16 units, gives lane l the units
//   {4g + i, 16 + 4g + i : i = 0..3} of sample c of its column tile,
// which is exactly the B fragment the next layer needs when its K index 8g + j stands for the
// unit perm(g, j) = (j < 4) ? 4g + j : 16 + 4g + j - 4.  The packed A operands store the
// weight columns in that order, so activations never move between lanes between layers.
#pragma once
#include <hip/hip_runtime.h>

namespace nsgpu {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float  f32x4 __attribute__((ext_vector_type(4)));

constexpr int MLP_H      = 32;
constexpr int MLP_NPARAM = 32 * 8 + 32 + 32 * 32 + 32 + 3 * 32 + 3;
// parameter offsets of the flat FP32 layout: W1[32][8], b1[32], W2[32][32], b2[32], W3[3][32], b3[3]
constexpr int MLP_W1 = 0, MLP_B1 = 256, MLP_W2 = 288, MLP_B2 = 1312, MLP_W3 = 1344, MLP_B3 = 1440;

// regime thresholds on (max beta + EPS) / (min beta + EPS) of the normalised stencil
constexpr float MLP_EPS    = 1e-6f;
constexpr float MLP_SMOOTH = 3.0f;
constexpr float MLP_FLAG   = 1e3f;
// a stencil whose range is at most MLP_RANGE of its largest magnitude counts as constant (smooth
// regime, classic weights): its normalised shape is round-off or noise, which the network never
// saw in training (2^-13, exact in float and double)
constexpr float MLP_RANGE  = 1.0f / 8192.0f;

// every lane's A fragments (bf16, already in the layouts above) and the FP32 biases
struct MlpPacked {
    bf16x8 A1[2][64];   // W1 rows 16m + c, K = input features (8 real, rows g > 0 zero)
    bf16x8 A2[2][64];   // W2 rows 16m + c, K = hidden-1 units in perm order
    bf16x8 A3[64];      // W3 rows c < 3, K = hidden-2 units in perm order
    bf16x8 A3T[2][64];  // W3^T rows 16m + c (hidden-2 units), K = outputs (g = 0, j < 3)
    bf16x8 A2T[2][64];  // W2^T rows 16m + c (hidden-1 units), K = hidden-2 units in perm order
    float  b1[32], b2[32], b3[4];
};

__device__ __forceinline__ int mlp_perm(int g, int j) { return (j < 4) ? 4 * g + j : 12 + 4 * g + j; }

// GELU, tanh form (Hendrycks & Gimpel 2016), and its derivative; tanh(y) = 1 - 2 / (e^{2y} + 1)
// GELU, tanh form: 0.5 x (1 + tanh(y)) = x s, s = 1 / (1 + exp(-2y)), y = sqrt(2/pi) (x +
// 0.044715 x^3).  exp2 and the reciprocal are the hardware approximations (about 1 ulp of
// float), far below the bfloat16 rounding the activations get next.
__device__ __forceinline__ float mlp_sig(float x) {
    constexpr float K = -2.0f * 0.7978845608028654f * 1.4426950408889634f;  // -2 sqrt(2/pi) log2 e
    const float     e = __builtin_amdgcn_exp2f(x * fmaf(0.044715f * K, x * x, K));
    return __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float mlp_gelu(float x) { return x * mlp_sig(x); }
// d/dx of x s: s + 2 x s (1 - s) dy/dx
__device__ __forceinline__ float mlp_gelu_d(float x) {
    const float s = mlp_sig(x);
    return s + 2.0f * x * s * (1.0f - s) * 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x * x);
}

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// bias fragment of row tile m: rows 4g + i
__device__ __forceinline__ f32x4 mlp_bias(const float* b, int m, int g) {
    return f32x4{b[16 * m + 4 * g], b[16 * m + 4 * g + 1], b[16 * m + 4 * g + 2],
                 b[16 * m + 4 * g + 3]};
}

// B fragment of the next layer from the two row tiles of this one
__device__ __forceinline__ bf16x8 mlp_join(const f32x4& lo, const f32x4& hi) {
    bf16x8 r;
#pragma unroll
    for (int i = 0; i < 4; i++) r[i] = (__bf16)lo[i], r[4 + i] = (__bf16)hi[i];
    return r;
}

// Forward pass of 64 samples held by one wavefront.  bx[t]: the B fragment of the input of column
// tile t (lane l with g = 0 holds the 8 features of sample 16t + c, other lanes zero).  Outputs
// the logit fragments z3[t] (lanes g = 0 hold logits 0..2 of sample 16t + c) and, when the
// caller trains, the pre-activations z1[m][t], z2[m][t] and the activation fragments.
struct MlpFwd {
    f32x4  z1[2][4], z2[2][4], z3[4];
    bf16x8 h1[4], h2[4];  // B fragments of hidden layers 1 and 2 (perm order)
};

__device__ __forceinline__ void mlp_forward_wave(const MlpPacked* __restrict__ P, const bf16x8 bx[4],
                                                 MlpFwd& F) {
    const int l = threadIdx.x & 63, g = l >> 4;
    const bf16x8 a10 = P->A1[0][l], a11 = P->A1[1][l];
    const f32x4  c10 = mlp_bias(P->b1, 0, g), c11 = mlp_bias(P->b1, 1, g);
#pragma unroll
    for (int t = 0; t < 4; t++) {
        F.z1[0][t] = mfma(a10, bx[t], c10);
        F.z1[1][t] = mfma(a11, bx[t], c11);
        f32x4 h0, h1;
#pragma unroll
        for (int i = 0; i < 4; i++) h0[i] = mlp_gelu(F.z1[0][t][i]), h1[i] = mlp_gelu(F.z1[1][t][i]);
        F.h1[t] = mlp_join(h0, h1);
    }
    const bf16x8 a20 = P->A2[0][l], a21 = P->A2[1][l];
    const f32x4  c20 = mlp_bias(P->b2, 0, g), c21 = mlp_bias(P->b2, 1, g);
#pragma unroll
    for (int t = 0; t < 4; t++) {
        F.z2[0][t] = mfma(a20, F.h1[t], c20);
        F.z2[1][t] = mfma(a21, F.h1[t], c21);
        f32x4 h0, h1;
#pragma unroll
        for (int i = 0; i < 4; i++) h0[i] = mlp_gelu(F.z2[0][t][i]), h1[i] = mlp_gelu(F.z2[1][t][i]);
        F.h2[t] = mlp_join(h0, h1);
    }
    const bf16x8 a3 = P->A3[l];
    const f32x4  c3 = (g == 0) ? f32x4{P->b3[0], P->b3[1], P->b3[2], 0.0f} : f32x4{0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < 4; t++) F.z3[t] = mfma(a3, F.h2[t], c3);
}

// softmax of three logits
__device__ __forceinline__ void mlp_softmax(float z0, float z1, float z2, float w[3]) {
    float m  = fmaxf(z0, fmaxf(z1, z2));
    float e0 = __expf(z0 - m), e1 = __expf(z1 - m), e2 = __expf(z2 - m);
    float r  = __builtin_amdgcn_rcpf(e0 + e1 + e2);
    w[0] = e0 * r, w[1] = e1 * r, w[2] = e2 * r;
}

// share of the network in the reconstruction as a function of the indicator ratio: rises from
// 0 at SMOOTH to 1 at 2 SMOOTH, falls from 1 at FLAG / 4 to 0 at FLAG (no jump in the weights
// where a face changes regime)
__device__ __forceinline__ float mlp_blend(float ratio) {
    float up   = fminf(1.0f, (ratio - MLP_SMOOTH) / MLP_SMOOTH);
    float down = fminf(1.0f, (MLP_FLAG - ratio) / (0.75f * MLP_FLAG));
    return fmaxf(0.0f, up) * fmaxf(0.0f, down);
}

// range-normalised stencil and smoothness indicators (features 0..7) of a left-biased stencil,
// the regime (0 smooth, 1 network, 2 flagged discontinuity) and the indicator ratio
template <class T>
__device__ __forceinline__ int mlp_features(const T v[5], float x[8], float* ratio_out = nullptr) {
    T m = v[0], M = v[0];
#pragma unroll
    for (int k = 1; k < 5; k++) m = v[k] < m ? v[k] : m, M = v[k] > M ? v[k] : M;
    const T    r    = M - m;
    const T    am   = (m < T(0)) ? -m : m, aM = (M < T(0)) ? -M : M;
    const bool vary = r > T(MLP_RANGE) * (am > aM ? am : aM);
    if constexpr (sizeof(T) == sizeof(float)) {  // single precision: one reciprocal
        const float ir = __builtin_amdgcn_rcpf(r);
#pragma unroll
        for (int k = 0; k < 5; k++) x[k] = vary ? (v[k] - m) * ir : 0.0f;
    } else {
#pragma unroll
        for (int k = 0; k < 5; k++) x[k] = vary ? (float)((v[k] - m) / r) : 0.0f;
    }
    float t1, t2;
    t1   = x[0] - 2.f * x[1] + x[2];
    t2   = x[0] - 4.f * x[1] + 3.f * x[2];
    x[5] = (13.f / 12.f) * t1 * t1 + 0.25f * t2 * t2;
    t1   = x[1] - 2.f * x[2] + x[3];
    t2   = x[1] - x[3];
    x[6] = (13.f / 12.f) * t1 * t1 + 0.25f * t2 * t2;
    t1   = x[2] - 2.f * x[3] + x[4];
    t2   = 3.f * x[2] - 4.f * x[3] + x[4];
    x[7] = (13.f / 12.f) * t1 * t1 + 0.25f * t2 * t2;
    float bmin = fminf(x[5], fminf(x[6], x[7])), bmax = fmaxf(x[5], fmaxf(x[6], x[7]));
    float ratio = (bmax + MLP_EPS) / (bmin + MLP_EPS);
    if (ratio_out) *ratio_out = ratio;
    return ratio < MLP_SMOOTH ? 0 : (ratio > MLP_FLAG ? 2 : 1);
}

// Network weights of the 64 samples held one per lane (features x of this lane's sample; lanes
// that do not need the network pass anything finite).  Every lane of the wavefront must call it
// (matrix instructions need the whole wave).  Returns the softmax weights of this lane's sample.
__device__ __forceinline__ void mlp_weights_lane(const MlpPacked* __restrict__ P, const float x[8],
                                                 float w[3]) {
    const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
    // features as bf16 pairs; lane c of column tile t needs those of lane 16t + c
    unsigned pk[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        __bf16 lo = (__bf16)x[2 * k], hi = (__bf16)x[2 * k + 1];
        pk[k] = (unsigned)__builtin_bit_cast(unsigned short, lo) |
                ((unsigned)__builtin_bit_cast(unsigned short, hi) << 16);
    }
    bf16x8 bx[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        unsigned q[4];
#pragma unroll
        for (int k = 0; k < 4; k++) q[k] = (unsigned)__shfl((int)pk[k], 16 * t + c, 64);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            unsigned v = (g == 0) ? q[k] : 0u;
            bx[t][2 * k]     = __builtin_bit_cast(__bf16, (unsigned short)(v & 0xffffu));
            bx[t][2 * k + 1] = __builtin_bit_cast(__bf16, (unsigned short)(v >> 16));
        }
    }
    MlpFwd F;
    mlp_forward_wave(P, bx, F);
    // the logits of sample l sit in lane c (g = 0) of column tile l / 16
    float z[3];
#pragma unroll
    for (int o = 0; o < 3; o++) {
        float v0 = __shfl(F.z3[0][o], c, 64), v1 = __shfl(F.z3[1][o], c, 64);
        float v2 = __shfl(F.z3[2][o], c, 64), v3 = __shfl(F.z3[3][o], c, 64);
        z[o]     = (g == 0) ? v0 : (g == 1) ? v1 : (g == 2) ? v2 : v3;
    }
    mlp_softmax(z[0], z[1], z[2], w);
}

}  // namespace nsgpu
