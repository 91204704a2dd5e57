// ns_mlp.hip -- training and batch inference of the learned WENO-5 weights (ns_mlp.h) with BF16
// matrix instructions, and their C ABI (include/nsgpu.h).  No reference counterpart.
//
// Training minimises the mean over samples of (sum_k w_k q_k - t)^2, w = softmax(MLP(x)), q the
// three ENO candidates of the normalised stencil, t the normalised target face value.  One
// wavefront runs forward and backward passes for 64 samples at a time: every matrix product
// (the three layers, the two back-propagations through W3^T and W2^T and the three weight
// gradients) is a v_mfma_f32_16x16x32_bf16; the weight gradients take the samples as their K
// dimension through a transposed copy of the activations in LDS.  Partial gradients of each
// wavefront are summed in a fixed order (deterministic), then an Adam step updates the FP32
// master parameters and re-packs the BF16 operands.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/nsgpu.h"
#include "ns_mlp.h"

using namespace nsgpu;

namespace {

// ------------------------------------------------------------------------------------------
// packing of the FP32 master parameters into the lane fragments
// ------------------------------------------------------------------------------------------
__global__ void k_mlp_pack(const float* __restrict__ p, MlpPacked* __restrict__ P) {
    const int l = threadIdx.x, g = l >> 4, c = l & 15;  // 64 threads
#pragma unroll
    for (int m = 0; m < 2; m++) {
        bf16x8 a1, a2, a3t, a2t;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int u = 16 * m + c;
            a1[j]  = (__bf16)((g == 0) ? p[MLP_W1 + u * 8 + j] : 0.0f);
            a2[j]  = (__bf16)p[MLP_W2 + u * 32 + mlp_perm(g, j)];
            a3t[j] = (__bf16)((g == 0 && j < 3) ? p[MLP_W3 + j * 32 + u] : 0.0f);
            a2t[j] = (__bf16)p[MLP_W2 + mlp_perm(g, j) * 32 + u];
        }
        P->A1[m][l] = a1, P->A2[m][l] = a2, P->A3T[m][l] = a3t, P->A2T[m][l] = a2t;
    }
    bf16x8 a3;
#pragma unroll
    for (int j = 0; j < 8; j++) a3[j] = (__bf16)((c < 3) ? p[MLP_W3 + c * 32 + mlp_perm(g, j)] : 0.0f);
    P->A3[l] = a3;
    if (l < 32) P->b1[l] = p[MLP_B1 + l], P->b2[l] = p[MLP_B2 + l];
    if (l < 4) P->b3[l] = (l < 3) ? p[MLP_B3 + l] : 0.0f;
}

// B fragment of the input layer for column tile t: lanes g = 0 hold the features of sample
// base + 16t + c (zero past n), other lanes zero
__device__ __forceinline__ bf16x8 load_bx(const float* __restrict__ X, long n, long base, int t) {
    const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
    const long s = base + 16 * t + c;
    bf16x8     b;
#pragma unroll
    for (int j = 0; j < 8; j++) b[j] = (__bf16)((g == 0 && s < n) ? X[s * 8 + j] : 0.0f);
    return b;
}

// ------------------------------------------------------------------------------------------
// batch inference: W[s][0..2] = softmax weights of the features X[s][0..7]
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_mlp_infer(const MlpPacked* __restrict__ P, long n,
                                                  const float* __restrict__ X,
                                                  float* __restrict__ W) {
    const int  l = threadIdx.x, g = l >> 4, c = l & 15;
    const long base = (long)blockIdx.x * 64;
    bf16x8     bx[4];
#pragma unroll
    for (int t = 0; t < 4; t++) bx[t] = load_bx(X, n, base, t);
    MlpFwd F;
    mlp_forward_wave(P, bx, F);
    if (g != 0) return;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const long s = base + 16 * t + c;
        if (s >= n) continue;
        float w[3];
        mlp_softmax(F.z3[t][0], F.z3[t][1], F.z3[t][2], w);
        W[s * 3] = w[0], W[s * 3 + 1] = w[1], W[s * 3 + 2] = w[2];
    }
}

// ------------------------------------------------------------------------------------------
// forward + backward over samples [0, n) (one wavefront per block, `iters` groups of 64
// samples per block); partial[blockIdx][MLP_NPARAM + 1] = gradient of the summed loss, loss sum
// ------------------------------------------------------------------------------------------
constexpr int LDS_ROW = 64 + 8;  // bf16 per transposed row (64 samples + pad)

__global__ __launch_bounds__(64) void k_mlp_grad(const MlpPacked* __restrict__ P, long n,
                                                 int iters, const float* __restrict__ X,
                                                 const float* __restrict__ Q,
                                                 const float* __restrict__ T,
                                                 float* __restrict__ partial) {
    // transposed activations [unit][sample] of one group of 64 samples
    __shared__ __bf16 xT[16][LDS_ROW], h1T[32][LDS_ROW], h2T[32][LDS_ROW];
    __shared__ __bf16 d1T[32][LDS_ROW], d2T[32][LDS_ROW], d3T[16][LDS_ROW];
    const int l = threadIdx.x, g = l >> 4, c = l & 15;
    // rows that stay zero: features 8..15, outputs 3..15
    for (int i = l; i < 8 * LDS_ROW; i += 64) xT[8 + i / LDS_ROW][i % LDS_ROW] = (__bf16)0.0f;
    for (int i = l; i < 13 * LDS_ROW; i += 64) d3T[3 + i / LDS_ROW][i % LDS_ROW] = (__bf16)0.0f;
    f32x4 gW1[2] = {}, gW2[2][2] = {}, gW3[2] = {};
    float gb1[8] = {}, gb2[8] = {}, gb3[3] = {}, lsum = 0.0f;
    const bf16x8 a3t0 = P->A3T[0][l], a3t1 = P->A3T[1][l], a2t0 = P->A2T[0][l], a2t1 = P->A2T[1][l];
    for (int it = 0; it < iters; it++) {
        const long base = ((long)blockIdx.x * iters + it) * 64;
        if (base >= n) break;
        bf16x8 bx[4];
#pragma unroll
        for (int t = 0; t < 4; t++) bx[t] = load_bx(X, n, base, t);
        MlpFwd F;
        mlp_forward_wave(P, bx, F);
        // loss and dz3 = w (dw - w.dw), dw_k = 2 e q_k (lanes g = 0; zero elsewhere / past n)
        bf16x8 bd3[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const long s  = base + 16 * t + c;
            float      d3[3] = {0.f, 0.f, 0.f};
            if (g == 0 && s < n) {
                float w[3];
                mlp_softmax(F.z3[t][0], F.z3[t][1], F.z3[t][2], w);
                const float q0 = Q[s * 3], q1 = Q[s * 3 + 1], q2 = Q[s * 3 + 2];
                const float e   = w[0] * q0 + w[1] * q1 + w[2] * q2 - T[s];
                lsum += e * e;
                const float dw0 = 2.f * e * q0, dw1 = 2.f * e * q1, dw2 = 2.f * e * q2;
                const float dot = w[0] * dw0 + w[1] * dw1 + w[2] * dw2;
                d3[0] = w[0] * (dw0 - dot), d3[1] = w[1] * (dw1 - dot), d3[2] = w[2] * (dw2 - dot);
            }
#pragma unroll
            for (int j = 0; j < 8; j++) bd3[t][j] = (__bf16)((j < 3) ? d3[j] : 0.0f);
            if (g == 0) {
#pragma unroll
                for (int o = 0; o < 3; o++) {
                    gb3[o] += (float)bd3[t][o];
                    d3T[o][16 * t + c] = bd3[t][o];
                }
#pragma unroll
                for (int j = 0; j < 8; j++) xT[j][16 * t + c] = bx[t][j];
            }
        }
        // back through W3^T and the second GELU: dz2 = (W3^T dz3) * gelu'(z2)
        bf16x8 bd2[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const f32x4 zero = {0, 0, 0, 0};
            f32x4       d0 = mfma(a3t0, bd3[t], zero), d1 = mfma(a3t1, bd3[t], zero);
            f32x4       e0, e1;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                e0[i] = d0[i] * mlp_gelu_d(F.z2[0][t][i]);
                e1[i] = d1[i] * mlp_gelu_d(F.z2[1][t][i]);
            }
            bd2[t] = mlp_join(e0, e1);
        }
        // back through W2^T and the first GELU: dz1 = (W2^T dz2) * gelu'(z1)
        bf16x8 bd1[4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const f32x4 zero = {0, 0, 0, 0};
            f32x4       d0 = mfma(a2t0, bd2[t], zero), d1 = mfma(a2t1, bd2[t], zero);
            f32x4       e0, e1;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                e0[i] = d0[i] * mlp_gelu_d(F.z1[0][t][i]);
                e1[i] = d1[i] * mlp_gelu_d(F.z1[1][t][i]);
            }
            bd1[t] = mlp_join(e0, e1);
        }
        // transposed copies [unit][sample] and the bias gradients (sums of the rounded dz)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int s = 16 * t + c;
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int u = mlp_perm(g, j);
                h1T[u][s]   = F.h1[t][j];
                h2T[u][s]   = F.h2[t][j];
                d1T[u][s]   = bd1[t][j];
                d2T[u][s]   = bd2[t][j];
                gb1[j] += (float)bd1[t][j];
                gb2[j] += (float)bd2[t][j];
            }
        }
        __syncthreads();
        // weight gradients: rows = this layer's units, columns = inputs, K = 64 samples
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            const int k0 = 32 * ks + 8 * g;
            auto      row = [&](const __bf16(*A)[LDS_ROW], int r) {
                return *reinterpret_cast<const bf16x8*>(&A[r][k0]);
            };
            const bf16x8 x_b = row(xT, c);
            const bf16x8 h1b0 = row(h1T, c), h1b1 = row(h1T, 16 + c);
            const bf16x8 h2b0 = row(h2T, c), h2b1 = row(h2T, 16 + c);
#pragma unroll
            for (int m = 0; m < 2; m++) {
                const bf16x8 d1a = row(d1T, 16 * m + c), d2a = row(d2T, 16 * m + c);
                gW1[m]    = mfma(d1a, x_b, gW1[m]);
                gW2[m][0] = mfma(d2a, h1b0, gW2[m][0]);
                gW2[m][1] = mfma(d2a, h1b1, gW2[m][1]);
            }
            const bf16x8 d3a = row(d3T, c);
            gW3[0] = mfma(d3a, h2b0, gW3[0]);
            gW3[1] = mfma(d3a, h2b1, gW3[1]);
        }
        __syncthreads();
    }
    // bias sums over the 16 lanes sharing g (different samples); loss over all lanes
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            gb1[j] += __shfl_xor(gb1[j], o, 64);
            gb2[j] += __shfl_xor(gb2[j], o, 64);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) gb3[k] += __shfl_xor(gb3[k], o, 64);
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) lsum += __shfl_xor(lsum, o, 64);
    float* out = partial + (long)blockIdx.x * (MLP_NPARAM + 1);
    // D fragments: rows 4g + i of the row tile, column c of the column tile
#pragma unroll
    for (int m = 0; m < 2; m++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int u = 16 * m + 4 * g + i;
            if (c < 8) out[MLP_W1 + u * 8 + c] = gW1[m][i];
            out[MLP_W2 + u * 32 + c]      = gW2[m][0][i];
            out[MLP_W2 + u * 32 + 16 + c] = gW2[m][1][i];
        }
    if (g == 0)
#pragma unroll
        for (int i = 0; i < 3; i++) {
            out[MLP_W3 + i * 32 + c]      = gW3[0][i];
            out[MLP_W3 + i * 32 + 16 + c] = gW3[1][i];
        }
    if (c == 0) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            out[MLP_B1 + mlp_perm(g, j)] = gb1[j];
            out[MLP_B2 + mlp_perm(g, j)] = gb2[j];
        }
        if (g == 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) out[MLP_B3 + k] = gb3[k];
            out[MLP_NPARAM] = lsum;
        }
    }
}

// fixed-order sum of the partials, scaled by 1/n: grad[i] (i < NPARAM), loss in grad[NPARAM]
__global__ void k_mlp_reduce(const float* __restrict__ partial, int nblocks, float inv_n,
                             float* __restrict__ grad) {
    const int i = threadIdx.x + blockIdx.x * blockDim.x;
    if (i > MLP_NPARAM) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; b++) s += partial[(long)b * (MLP_NPARAM + 1) + i];
    grad[i] = (float)(s * inv_n);
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters; step t >= 1
__global__ void k_mlp_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ grad, float lr, float b1, float b2,
                           float eps, int t) {
    const int i = threadIdx.x + blockIdx.x * blockDim.x;
    if (i >= MLP_NPARAM) return;
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi, v[i] = vi;
    const float mh = mi / (1.f - powf(b1, (float)t)), vh = vi / (1.f - powf(b2, (float)t));
    p[i] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void k_copy_loss(const float* __restrict__ grad, float* __restrict__ hist, int k) {
    hist[k] = grad[MLP_NPARAM];
}

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            fprintf(stderr, "nsgpu: %s failed: %s\n", #x, hipGetErrorString(e_));             \
            return NSGPU_ERR_HIP;                                                              \
        }                                                                                      \
    } while (0)

// device buffers of one MLP call; freed by the destructor
struct Bufs {
    std::vector<void*> p;
    ~Bufs() {
        for (void* q : p) hipFree(q);
    }
    template <class T>
    int get(T** out, size_t bytes) {
        void* q = nullptr;
        if (hipMalloc(&q, bytes) != hipSuccess) return NSGPU_ERR_HIP;
        p.push_back(q);
        *out = (T*)q;
        return NSGPU_OK;
    }
};

int device_ok(int device) {
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0) return NSGPU_ERR_NODEVICE;
    if (device < 0 || device >= nd) return NSGPU_ERR_NODEVICE;
    HIPCHK(hipSetDevice(device));
    return NSGPU_OK;
}

constexpr int GRAD_ITERS = 4;  // groups of 64 samples per wavefront in k_mlp_grad

int grad_launch(const MlpPacked* P, long n, const float* X, const float* Q, const float* T,
                float* partial, float* grad, hipStream_t s) {
    const long nb = (n + 64L * GRAD_ITERS - 1) / (64L * GRAD_ITERS);
    k_mlp_grad<<<(unsigned)nb, 64, 0, s>>>(P, n, GRAD_ITERS, X, Q, T, partial);
    k_mlp_reduce<<<(MLP_NPARAM + 1 + 255) / 256, 256, 0, s>>>(partial, (int)nb, 1.0f / (float)n,
                                                               grad);
    return NSGPU_OK;
}

// ------------------------------------------------------------------------------------------
// allocation-free loss gradient (nsgpu_mlp_loss_grad_noalloc): the packed network and the
// per-workgroup partial gradients live in one file-scope device array, sized for at most
// NSGPU_MLP_NOALLOC_MAX_N samples (1024 workgroups of 256 samples: 5.9 MB + the 9.5 KB network).
// Each GPU holds its own copy until the process ends; its address is looked up once per device.
// k_mlp_pack rewrites the whole network and k_mlp_grad every entry of the partial rows the
// reduction reads, so no part of the array has to start at zero.
// ------------------------------------------------------------------------------------------
constexpr long   NOALLOC_BLOCKS  = (NSGPU_MLP_NOALLOC_MAX_N + 64L * GRAD_ITERS - 1) / (64L * GRAD_ITERS);
constexpr size_t NOALLOC_NET     = (sizeof(MlpPacked) + 255) / 256 * 256;
constexpr size_t NOALLOC_BYTES   = NOALLOC_NET + sizeof(float) * (MLP_NPARAM + 1) * NOALLOC_BLOCKS;
constexpr int    NOALLOC_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

std::mutex     g_scratch_mu;
unsigned char* g_scratch_ptr[NOALLOC_DEVICES] = {};

// the device address of `scratch` on the current device `device` (looked up on first use)
int scratch_address(int device, unsigned char** out) {
    if (device >= NOALLOC_DEVICES) return NSGPU_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[device];
    return NSGPU_OK;
}

}  // namespace

namespace nsgpu {
// packs FP32 parameters (device pointer) into `P` on stream s (used by nsgpu_set_mlp)
void mlp_pack(const float* params, MlpPacked* P, hipStream_t s) {
    k_mlp_pack<<<1, 64, 0, s>>>(params, P);
}
}  // namespace nsgpu

extern "C" {

int nsgpu_mlp_forward(int device, const float* params, long n, const float* X, float* W) {
    if (!params || !X || !W || n < 1 || n > (1L << 31)) return NSGPU_ERR_INVALID;
    int rc = device_ok(device);
    if (rc) return rc;
    Bufs       B;
    float *    dp = nullptr, *dX = nullptr, *dW = nullptr;
    MlpPacked* P  = nullptr;
    if ((rc = B.get(&dp, sizeof(float) * MLP_NPARAM)) || (rc = B.get(&P, sizeof(MlpPacked))) ||
        (rc = B.get(&dX, sizeof(float) * 8 * n)) || (rc = B.get(&dW, sizeof(float) * 3 * n)))
        return rc;
    HIPCHK(hipMemcpy(dp, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dX, X, sizeof(float) * 8 * n, hipMemcpyHostToDevice));
    k_mlp_pack<<<1, 64>>>(dp, P);
    k_mlp_infer<<<(unsigned)((n + 63) / 64), 64>>>(P, n, dX, dW);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(W, dW, sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
    return NSGPU_OK;
}

int nsgpu_mlp_loss_grad(int device, const float* params, long n, const float* X, const float* Q,
                        const float* T, double* loss, float* grad) {
    if (!params || !X || !Q || !T || n < 1 || n > (1L << 31)) return NSGPU_ERR_INVALID;
    int rc = device_ok(device);
    if (rc) return rc;
    Bufs       B;
    float *    dp = nullptr, *dX = nullptr, *dQ = nullptr, *dT = nullptr, *dpart = nullptr,
          *dg = nullptr;
    MlpPacked* P  = nullptr;
    const long nb = (n + 64L * GRAD_ITERS - 1) / (64L * GRAD_ITERS);
    if ((rc = B.get(&dp, sizeof(float) * MLP_NPARAM)) || (rc = B.get(&P, sizeof(MlpPacked))) ||
        (rc = B.get(&dX, sizeof(float) * 8 * n)) || (rc = B.get(&dQ, sizeof(float) * 3 * n)) ||
        (rc = B.get(&dT, sizeof(float) * n)) ||
        (rc = B.get(&dpart, sizeof(float) * (MLP_NPARAM + 1) * nb)) ||
        (rc = B.get(&dg, sizeof(float) * (MLP_NPARAM + 1))))
        return rc;
    HIPCHK(hipMemcpy(dp, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dX, X, sizeof(float) * 8 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dQ, Q, sizeof(float) * 3 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dT, T, sizeof(float) * n, hipMemcpyHostToDevice));
    k_mlp_pack<<<1, 64>>>(dp, P);
    grad_launch(P, n, dX, dQ, dT, dpart, dg, nullptr);
    HIPCHK(hipGetLastError());
    std::vector<float> h(MLP_NPARAM + 1);
    HIPCHK(hipMemcpy(h.data(), dg, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
    if (grad)
        for (int i = 0; i < MLP_NPARAM; i++) grad[i] = h[i];
    if (loss) *loss = h[MLP_NPARAM];
    return NSGPU_OK;
}

int nsgpu_mlp_train(int device, float* params, long n, const float* X, const float* Q,
                    const float* T, int steps, long batch, float lr, float* loss_hist) {
    if (!params || !X || !Q || !T || n < 1 || n > (1L << 31) || steps < 0 || batch < 64 ||
        !(lr > 0.0f))
        return NSGPU_ERR_INVALID;
    int rc = device_ok(device);
    if (rc) return rc;
    if (batch > n) batch = n;
    Bufs       B;
    float *    dp = nullptr, *dm = nullptr, *dv = nullptr, *dX = nullptr, *dQ = nullptr,
          *dT = nullptr, *dpart = nullptr, *dg = nullptr, *dh = nullptr;
    MlpPacked* P  = nullptr;
    const long nb = (batch + 64L * GRAD_ITERS - 1) / (64L * GRAD_ITERS);
    if ((rc = B.get(&dp, sizeof(float) * MLP_NPARAM)) || (rc = B.get(&dm, sizeof(float) * MLP_NPARAM)) ||
        (rc = B.get(&dv, sizeof(float) * MLP_NPARAM)) || (rc = B.get(&P, sizeof(MlpPacked))) ||
        (rc = B.get(&dX, sizeof(float) * 8 * n)) || (rc = B.get(&dQ, sizeof(float) * 3 * n)) ||
        (rc = B.get(&dT, sizeof(float) * n)) ||
        (rc = B.get(&dpart, sizeof(float) * (MLP_NPARAM + 1) * nb)) ||
        (rc = B.get(&dg, sizeof(float) * (MLP_NPARAM + 1))) ||
        (rc = B.get(&dh, sizeof(float) * (steps > 0 ? steps : 1))))
        return rc;
    HIPCHK(hipMemcpy(dp, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(dm, 0, sizeof(float) * MLP_NPARAM));
    HIPCHK(hipMemset(dv, 0, sizeof(float) * MLP_NPARAM));
    HIPCHK(hipMemcpy(dX, X, sizeof(float) * 8 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dQ, Q, sizeof(float) * 3 * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dT, T, sizeof(float) * n, hipMemcpyHostToDevice));
    HIPCHK(hipDeviceSynchronize());
    // mini-batches are consecutive windows of the (caller-shuffled) samples, wrapping around
    long off = 0;
    for (int k = 0; k < steps; k++) {
        long nbat = batch;
        if (off + nbat > n) off = 0;
        k_mlp_pack<<<1, 64>>>(dp, P);
        grad_launch(P, nbat, dX + off * 8, dQ + off * 3, dT + off, dpart, dg, nullptr);
        k_mlp_adam<<<(MLP_NPARAM + 255) / 256, 256>>>(dp, dm, dv, dg, lr, 0.9f, 0.999f, 1e-8f,
                                                      k + 1);
        k_copy_loss<<<1, 1>>>(dg, dh, k);
        off += nbat;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(params, dp, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
    if (loss_hist && steps > 0)
        HIPCHK(hipMemcpy(loss_hist, dh, sizeof(float) * steps, hipMemcpyDeviceToHost));
    return NSGPU_OK;
}

int nsgpu_mlp_loss_grad_noalloc(int device, const float* params, long n, const float* X,
                                const float* Q, const float* T, float* out, void* stream) {
    if (!params || !X || !Q || !T || !out || n < 1) return NSGPU_ERR_INVALID;
    if (n > NSGPU_MLP_NOALLOC_MAX_N) {
        fprintf(stderr, "nsgpu: nsgpu_mlp_loss_grad_noalloc: n = %ld exceeds the maximum of %ld "
                        "samples of its fixed device scratch (use nsgpu_mlp_loss_grad)\n",
                n, (long)NSGPU_MLP_NOALLOC_MAX_N);
        return NSGPU_ERR_INVALID;
    }
    int rc = device_ok(device);
    if (rc) return rc;
    unsigned char* base = nullptr;
    if ((rc = scratch_address(device, &base))) return rc;
    MlpPacked*        P       = reinterpret_cast<MlpPacked*>(base);
    float*            partial = reinterpret_cast<float*>(base + NOALLOC_NET);
    const hipStream_t s       = static_cast<hipStream_t>(stream);
    k_mlp_pack<<<1, 64, 0, s>>>(params, P);
    grad_launch(P, n, X, Q, T, partial, out, s);
    HIPCHK(hipGetLastError());
    return NSGPU_OK;
}

}  // extern "C"
