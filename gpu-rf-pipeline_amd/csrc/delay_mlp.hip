// Delay-proposal network: 256 -> 256 -> 256 -> 64 multilayer perceptron with exact GELU
// activations and a softmax over 64 delay bins, trained on the GPU with its own forward,
// backward and Adam kernels. No counterpart in the reference.
//
// Numerics: FP32 master parameters; every GEMM operand (inputs, activations, weights,
// gradients) is rounded to BF16 (round to nearest even) as it is loaded into the matrix cores
// (v_mfma_f32_32x32x16_bf16), products accumulate in FP32; biases, GELU, softmax, the loss and
// Adam run in FP32.
//
// Parameter layout (floats): W1[256][256], b1[256], W2[256][256], b2[256], W3[64][256], b3[64]
// (148 032 values); layer l computes z = W_l h + b_l on a row vector h.

#include "philox.hpp"
#include "rf_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>

namespace {

constexpr int kIn = 256;
constexpr int kHidden = 256;
constexpr int kOut = 64;
constexpr long long kW1 = 0;
constexpr long long kB1 = kW1 + kHidden * kIn;
constexpr long long kW2 = kB1 + kHidden;
constexpr long long kB2 = kW2 + kHidden * kHidden;
constexpr long long kW3 = kB2 + kHidden;
constexpr long long kB3 = kW3 + kOut * kHidden;
constexpr long long kParams = kB3 + kOut;  // 148 032

}  // namespace

struct rf_delay_mlp {
    int max_batch = 0;
    float* params = nullptr;  // kParams
    float* grads = nullptr;   // kParams
    float* adam_m = nullptr;  // kParams
    float* adam_v = nullptr;  // kParams
    float* z1 = nullptr;      // max_batch x 256 (pre-activations)
    float* h1 = nullptr;
    float* z2 = nullptr;
    float* h2 = nullptr;
    float* z3 = nullptr;      // max_batch x 64 logits
    float* probs = nullptr;   // max_batch x 64
    float* dz = nullptr;      // max_batch x 256 scratch gradient
    float* dh = nullptr;      // max_batch x 256 scratch gradient
    int steps = 0;            // Adam step count
};

namespace rf {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Strided matrix view: element (r, c) at p[r * sr + c * sc].
struct View {
    const float* p;
    long long sr;
    long long sc;
};

enum Epilogue : int {
    kStore = 0,     // C = A B
    kBias = 1,      // Z = A B + bias (and H = Z)
    kBiasGelu = 2,  // Z = A B + bias, H = gelu(Z)
};

__device__ __forceinline__ float gelu(float z)
{
    return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_grad(float z)
{
    // d/dz [z Phi(z)] = Phi(z) + z phi(z)
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

// Eight consecutive k values of row `row` of a view (k-contiguous when VEC: two 16-byte
// loads), rows >= rows and depths >= K read as zero without leaving the buffer (indices are
// clamped, values zeroed by select, no branch per element).
template <bool VEC>
__device__ __forceinline__ void load8(const View& v, int row, int rows, int k, int K, float* out)
{
    const bool row_ok = row < rows;
    const int rc = row_ok ? row : rows - 1;
    if (VEC && k + 8 <= K) {
        const float4* p = reinterpret_cast<const float4*>(v.p + rc * v.sr + k);
        const float4 lo = p[0];
        const float4 hi = p[1];
        const float z = row_ok ? 1.0f : 0.0f;
        out[0] = lo.x * z; out[1] = lo.y * z; out[2] = lo.z * z; out[3] = lo.w * z;
        out[4] = hi.x * z; out[5] = hi.y * z; out[6] = hi.z * z; out[7] = hi.w * z;
        return;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const bool ok = row_ok && k + q < K;
        const int kc = k + q < K ? k + q : K - 1;
        const float val = v.p[rc * v.sr + kc * v.sc];
        out[q] = ok ? val : 0.0f;
    }
}

// C[M x N] = A[M x K] B[K x N]: one 32 x 32 tile of C per workgroup of four waves, which split
// the K/16 matrix-core steps between them (wave w takes steps w, w + 4, ...; two steps' loads
// in flight per wave) and add their partial tiles through LDS in wave order, so the result does
// not depend on timing. Grid (N / 32, ceil(M / 32)). Rows m >= M and depths k >= K read as
// zero; N is a multiple of 32. VEC_A / VEC_B: the operand is contiguous along k (A: sc == 1,
// B viewed as rows n: sr == 1).
// MFMA 32x32x16 operand layout: lane l holds A(l % 32, 8 (l / 32) + q) and
// B(8 (l / 32) + q, l % 32) for q < 8; accumulator v of lane l is C(row, l % 32) with
// row = (v % 4) + 8 (v / 4) + 4 (l / 32).
constexpr int kGemmWaves = 4;

template <bool VEC_A, bool VEC_B>
__device__ __forceinline__ void gemm_step(const View& a, const View& bt, int am, int bn, int M,
                                          int N, int k, int K, bf16x8& av, bf16x8& bv)
{
    float x[8];
    float y[8];
    load8<VEC_A>(a, am, M, k, K, x);
    load8<VEC_B>(bt, bn, N, k, K, y);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        av[q] = static_cast<__bf16>(x[q]);
        bv[q] = static_cast<__bf16>(y[q]);
    }
}

template <bool VEC_A, bool VEC_B>
__global__ __launch_bounds__(64 * kGemmWaves) void mlp_gemm_kernel(
    View a, View b, int M, int K, int N, int epilogue, const float* __restrict__ bias,
    float* __restrict__ c, float* __restrict__ h)
{
    __shared__ float partial[kGemmWaves - 1][16][64];
    const int l = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int m0 = blockIdx.y * 32;
    const int n0 = blockIdx.x * 32;
    const int i = l & 31;
    const int kq = 8 * (l >> 5);
    const int am = m0 + i;
    // B as rows n (k along the row): element (k, n) at p[k sr + n sc].
    const View bt{b.p, b.sc, b.sr};
    f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        acc[v] = 0.0f;
    }
    const int steps = (K + 15) >> 4;
    for (int s = w; s < steps; s += 2 * kGemmWaves) {
        bf16x8 av0, bv0, av1, bv1;
        const bool second = s + kGemmWaves < steps;
        gemm_step<VEC_A, VEC_B>(a, bt, am, n0 + i, M, N, 16 * s + kq, K, av0, bv0);
        if (second) {
            gemm_step<VEC_A, VEC_B>(a, bt, am, n0 + i, M, N, 16 * (s + kGemmWaves) + kq, K, av1,
                                    bv1);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av0, bv0, acc, 0, 0, 0);
        if (second) {
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av1, bv1, acc, 0, 0, 0);
        }
    }
    if (w > 0) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            partial[w - 1][v][l] = acc[v];
        }
    }
    __syncthreads();
    if (w > 0) {
        return;
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
#pragma unroll
        for (int u = 0; u < kGemmWaves - 1; ++u) {
            acc[v] += partial[u][v][l];
        }
    }
    const int col = n0 + i;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = m0 + (v & 3) + 8 * (v >> 2) + 4 * (l >> 5);
        if (row >= M) {
            continue;
        }
        float z = acc[v];
        const long long o = static_cast<long long>(row) * N + col;
        if (epilogue == kStore) {
            c[o] = z;
            continue;
        }
        z += bias[col];
        c[o] = z;
        h[o] = epilogue == kBiasGelu ? gelu(z) : z;
    }
}

void gemm(View a, View b, int M, int K, int N, int epilogue, const float* bias, float* c,
          float* h, hipStream_t stream)
{
    const dim3 grid(static_cast<unsigned>(N / 32), static_cast<unsigned>(ceil_div(M, 32)));
    const unsigned threads = 64 * kGemmWaves;
    // 16-byte loads need k-contiguous operands at 16-byte aligned rows (strides multiple of 4)
    const bool va = a.sc == 1 && a.sr % 4 == 0;
    const bool vb = b.sr == 1 && b.sc % 4 == 0;
    if (va && vb) {
        mlp_gemm_kernel<true, true><<<grid, threads, 0, stream>>>(a, b, M, K, N, epilogue, bias,
                                                                 c, h);
    } else if (va) {
        mlp_gemm_kernel<true, false><<<grid, threads, 0, stream>>>(a, b, M, K, N, epilogue, bias,
                                                                  c, h);
    } else if (vb) {
        mlp_gemm_kernel<false, true><<<grid, threads, 0, stream>>>(a, b, M, K, N, epilogue, bias,
                                                                  c, h);
    } else {
        mlp_gemm_kernel<false, false><<<grid, threads, 0, stream>>>(a, b, M, K, N, epilogue,
                                                                   bias, c, h);
    }
    check_launch("mlp_gemm_kernel");
}

// Softmax over 64 logits per row (one wave per row); with labels, also the mean cross-entropy
// (atomically summed into *loss, which the caller clears) and dZ3 = (P - onehot) / B.
__global__ __launch_bounds__(64) void mlp_softmax_kernel(const float* __restrict__ z3,
                                                         float* __restrict__ probs,
                                                         const int* __restrict__ labels,
                                                         float* __restrict__ dz3,
                                                         float* __restrict__ loss, int batch)
{
    const int row = blockIdx.x;
    const int j = threadIdx.x;
    const float z = z3[static_cast<long long>(row) * kOut + j];
    float mx = z;
    for (int off = 32; off > 0; off >>= 1) {
        mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    }
    const float e = expf(z - mx);
    float sum = e;
    for (int off = 32; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off, 64);
    }
    const float p = e / sum;
    probs[static_cast<long long>(row) * kOut + j] = p;
    if (labels != nullptr) {
        const int label = labels[row];
        dz3[static_cast<long long>(row) * kOut + j] =
            (p - (j == label ? 1.0f : 0.0f)) / static_cast<float>(batch);
        if (j == label) {
            // -log p = log(sum) - (z - max)
            atomicAdd(loss, (logf(sum) - (z - mx)) / static_cast<float>(batch));
        }
    }
}

// dZ = dH * gelu'(Z), elementwise.
__global__ void mlp_gelu_backward_kernel(const float* __restrict__ dh, const float* __restrict__ z,
                                         float* __restrict__ dz, long long count)
{
    for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < count;
         e += static_cast<long long>(gridDim.x) * blockDim.x) {
        dz[e] = dh[e] * gelu_grad(z[e]);
    }
}

// db[j] = sum over rows of dZ[row][j] (one lane per column, rows in order).
__global__ void mlp_bias_grad_kernel(const float* __restrict__ dz, int rows, int cols,
                                     float* __restrict__ db)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) {
        return;
    }
    float s = 0.0f;
    for (int r = 0; r < rows; ++r) {
        s += dz[static_cast<long long>(r) * cols + j];
    }
    db[j] = s;
}

__global__ void mlp_fill_kernel(float* __restrict__ p, long long count, float value)
{
    for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < count;
         e += static_cast<long long>(gridDim.x) * blockDim.x) {
        p[e] = value;
    }
}

// He-uniform initialisation: W ~ U(-sqrt(6 / fan_in), sqrt(6 / fan_in)) from Philox4x32-10
// (key = seed, counter (index, 0x4D4C50, 0, 0), first word), biases 0.
__global__ void mlp_init_kernel(float* __restrict__ params, uint32_t k0, uint32_t k1)
{
    for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
         e < kParams; e += static_cast<long long>(gridDim.x) * blockDim.x) {
        const bool bias = (e >= kB1 && e < kW2) || (e >= kB2 && e < kW3) || e >= kB3;
        if (bias) {
            params[e] = 0.0f;
            continue;
        }
        const Philox4 x = philox4x32_10(Philox4{{static_cast<uint32_t>(e), 0x4D4C50u, 0u, 0u}},
                                        k0, k1);
        const float bound = sqrtf(6.0f / static_cast<float>(kIn));  // every fan-in is 256
        params[e] = bound * (2.0f * uniform24(x.v[0]) - 1.0f);
    }
}

// Adam (Kingma & Ba 2015) with bias correction: m = b1 m + (1 - b1) g,
// v = b2 v + (1 - b2) g^2, p -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps).
__global__ void mlp_adam_kernel(float* __restrict__ params, const float* __restrict__ grads,
                                float* __restrict__ m, float* __restrict__ v, float lr,
                                float corr1, float corr2)
{
    constexpr float b1 = 0.9f;
    constexpr float b2 = 0.999f;
    constexpr float eps = 1e-8f;
    for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
         e < kParams; e += static_cast<long long>(gridDim.x) * blockDim.x) {
        const float g = grads[e];
        const float mm = b1 * m[e] + (1.0f - b1) * g;
        const float vv = b2 * v[e] + (1.0f - b2) * g * g;
        m[e] = mm;
        v[e] = vv;
        params[e] -= lr * (mm / corr1) / (sqrtf(vv / corr2) + eps);
    }
}

unsigned elementwise_grid(long long count)
{
    return static_cast<unsigned>(std::min<long long>(ceil_div(count, 256), 4LL * cu_count()));
}

void forward(rf_delay_mlp& h, const float* x, int batch, hipStream_t stream)
{
    const float* p = h.params;
    // Z1 = X W1^T + b1: A(m, k) = X[m][k]; B(k, n) = W1[n][k]
    gemm(View{x, kIn, 1}, View{p + kW1, 1, kIn}, batch, kIn, kHidden, kBiasGelu, p + kB1, h.z1,
         h.h1, stream);
    gemm(View{h.h1, kHidden, 1}, View{p + kW2, 1, kHidden}, batch, kHidden, kHidden, kBiasGelu,
         p + kB2, h.z2, h.h2, stream);
    gemm(View{h.h2, kHidden, 1}, View{p + kW3, 1, kHidden}, batch, kHidden, kOut, kBias, p + kB3,
         h.z3, h.z3, stream);
}

// Gradients of the mean cross-entropy into h.grads, loss into *loss.
void backward(rf_delay_mlp& h, const float* x, const int* labels, int batch, float* loss,
              hipStream_t stream)
{
    const float* p = h.params;
    float* g = h.grads;
    mlp_fill_kernel<<<1, 64, 0, stream>>>(loss, 1, 0.0f);
    check_launch("mlp_fill_kernel");
    float* dz3 = h.dz;  // batch x 64
    mlp_softmax_kernel<<<static_cast<unsigned>(batch), 64, 0, stream>>>(h.z3, h.probs, labels,
                                                                        dz3, loss, batch);
    check_launch("mlp_softmax_kernel");
    // dW3[o][i] = sum_b dZ3[b][o] H2[b][i]: A(m=o, k=b) = dZ3[b][o]; B(k=b, n=i) = H2[b][i]
    gemm(View{dz3, 1, kOut}, View{h.h2, kHidden, 1}, kOut, batch, kHidden, kStore, nullptr,
         g + kW3, nullptr, stream);
    mlp_bias_grad_kernel<<<1, kOut, 0, stream>>>(dz3, batch, kOut, g + kB3);
    check_launch("mlp_bias_grad_kernel");
    // dH2[b][i] = sum_o dZ3[b][o] W3[o][i]
    gemm(View{dz3, kOut, 1}, View{p + kW3, kHidden, 1}, batch, kOut, kHidden, kStore, nullptr,
         h.dh, nullptr, stream);
    const long long acts = static_cast<long long>(batch) * kHidden;
    float* dz2 = h.dz;  // dz3 is no longer needed
    mlp_gelu_backward_kernel<<<elementwise_grid(acts), 256, 0, stream>>>(h.dh, h.z2, dz2, acts);
    check_launch("mlp_gelu_backward_kernel");
    gemm(View{dz2, 1, kHidden}, View{h.h1, kHidden, 1}, kHidden, batch, kHidden, kStore, nullptr,
         g + kW2, nullptr, stream);
    mlp_bias_grad_kernel<<<1, kHidden, 0, stream>>>(dz2, batch, kHidden, g + kB2);
    check_launch("mlp_bias_grad_kernel");
    gemm(View{dz2, kHidden, 1}, View{p + kW2, kHidden, 1}, batch, kHidden, kHidden, kStore,
         nullptr, h.dh, nullptr, stream);
    float* dz1 = h.dz;
    mlp_gelu_backward_kernel<<<elementwise_grid(acts), 256, 0, stream>>>(h.dh, h.z1, dz1, acts);
    check_launch("mlp_gelu_backward_kernel");
    gemm(View{dz1, 1, kHidden}, View{x, kIn, 1}, kHidden, batch, kIn, kStore, nullptr, g + kW1,
         nullptr, stream);
    mlp_bias_grad_kernel<<<1, kHidden, 0, stream>>>(dz1, batch, kHidden, g + kB1);
    check_launch("mlp_bias_grad_kernel");
}

void check_batch(const rf_delay_mlp* h, const float* x, int batch)
{
    if (h == nullptr || x == nullptr) fail(RF_INVALID_ARG, "delay mlp: NULL handle or features");
    if (batch < 1 || batch > h->max_batch) {
        fail(RF_LIMIT, "delay mlp: batch " + std::to_string(batch) + " outside 1.."
                           + std::to_string(h->max_batch) + " (the handle's max_batch)");
    }
}

void destroy_mlp(rf_delay_mlp* h)
{
    if (h == nullptr) {
        return;
    }
    for (float* p : {h->params, h->grads, h->adam_m, h->adam_v, h->z1, h->h1, h->z2, h->h2,
                     h->z3, h->probs, h->dz, h->dh}) {
        (void)hipFree(p);
    }
    delete h;
}

}  // namespace

// Library-internal entry points used by the delay search and the training driver.
void delay_mlp_forward(rf_delay_mlp& h, const float* x, int batch, float* probs,
                       hipStream_t stream)
{
    forward(h, x, batch, stream);
    mlp_softmax_kernel<<<static_cast<unsigned>(batch), 64, 0, stream>>>(h.z3, probs, nullptr,
                                                                        nullptr, nullptr, batch);
    check_launch("mlp_softmax_kernel");
}

void delay_mlp_train_step(rf_delay_mlp& h, const float* x, const int* labels, int batch,
                          float lr, float* loss, hipStream_t stream)
{
    forward(h, x, batch, stream);
    backward(h, x, labels, batch, loss, stream);
    h.steps += 1;
    const float corr1 = 1.0f - static_cast<float>(std::pow(0.9, h.steps));
    const float corr2 = 1.0f - static_cast<float>(std::pow(0.999, h.steps));
    mlp_adam_kernel<<<elementwise_grid(kParams), 256, 0, stream>>>(h.params, h.grads, h.adam_m,
                                                                   h.adam_v, lr, corr1, corr2);
    check_launch("mlp_adam_kernel");
}

int delay_mlp_max_batch(const rf_delay_mlp& h)
{
    return h.max_batch;
}

}  // namespace rf

using rf::guarded;

extern "C" {

rf_status rf_delay_mlp_create(int max_batch, unsigned long long seed, rf_delay_mlp** out)
{
    return guarded([&] {
        if (out == nullptr) rf::fail(RF_INVALID_ARG, "delay mlp: NULL output handle");
        *out = nullptr;
        if (max_batch < 1 || max_batch > (1 << 20)) {
            rf::fail(RF_INVALID_ARG, "delay mlp: max_batch must be 1..2^20");
        }
        rf_delay_mlp* h = new rf_delay_mlp;
        try {
            h->max_batch = max_batch;
            for (float** p : {&h->params, &h->grads, &h->adam_m, &h->adam_v}) {
                RF_HIP_CHECK(hipMalloc(p, kParams * sizeof(float)));
            }
            const std::size_t act = static_cast<std::size_t>(max_batch) * kHidden * sizeof(float);
            for (float** p : {&h->z1, &h->h1, &h->z2, &h->h2, &h->dz, &h->dh}) {
                RF_HIP_CHECK(hipMalloc(p, act));
            }
            const std::size_t out_bytes = static_cast<std::size_t>(max_batch) * kOut * sizeof(float);
            RF_HIP_CHECK(hipMalloc(&h->z3, out_bytes));
            RF_HIP_CHECK(hipMalloc(&h->probs, out_bytes));
            const unsigned grid = rf::elementwise_grid(kParams);
            rf::mlp_init_kernel<<<grid, 256>>>(h->params, static_cast<uint32_t>(seed),
                                               static_cast<uint32_t>(seed >> 32));
            rf::check_launch("mlp_init_kernel");
            rf::mlp_fill_kernel<<<grid, 256>>>(h->adam_m, kParams, 0.0f);
            rf::check_launch("mlp_fill_kernel");
            rf::mlp_fill_kernel<<<grid, 256>>>(h->adam_v, kParams, 0.0f);
            rf::check_launch("mlp_fill_kernel");
            RF_HIP_CHECK(hipDeviceSynchronize());
        } catch (...) {
            rf::destroy_mlp(h);
            throw;
        }
        *out = h;
    });
}

rf_status rf_delay_mlp_destroy(rf_delay_mlp* h)
{
    return guarded([&] { rf::destroy_mlp(h); });
}

rf_status rf_delay_mlp_params(rf_delay_mlp* h, float** params, size_t* count)
{
    return guarded([&] {
        if (h == nullptr || params == nullptr) rf::fail(RF_INVALID_ARG, "delay mlp: NULL");
        *params = h->params;
        if (count != nullptr) {
            *count = static_cast<size_t>(kParams);
        }
    });
}

rf_status rf_delay_mlp_forward(rf_delay_mlp* h, const float* features, int batch, float* probs,
                               rf_stream stream)
{
    return guarded([&] {
        rf::check_batch(h, features, batch);
        if (probs == nullptr) rf::fail(RF_INVALID_ARG, "delay mlp: NULL probabilities");
        rf::delay_mlp_forward(*h, features, batch, probs, rf::as_stream(stream));
    });
}

rf_status rf_delay_mlp_gradients(rf_delay_mlp* h, const float* features, const int* labels,
                                 int batch, float* grads, float* loss, rf_stream stream)
{
    return guarded([&] {
        rf::check_batch(h, features, batch);
        if (labels == nullptr || loss == nullptr) {
            rf::fail(RF_INVALID_ARG, "delay mlp: NULL labels or loss");
        }
        hipStream_t s = rf::as_stream(stream);
        rf::forward(*h, features, batch, s);
        rf::backward(*h, features, labels, batch, loss, s);
        if (grads != nullptr) {
            RF_HIP_CHECK(hipMemcpyAsync(grads, h->grads, kParams * sizeof(float),
                                        hipMemcpyDeviceToDevice, s));
        }
    });
}

rf_status rf_delay_mlp_train_step(rf_delay_mlp* h, const float* features, const int* labels,
                                  int batch, float learning_rate, float* loss, rf_stream stream)
{
    return guarded([&] {
        rf::check_batch(h, features, batch);
        if (labels == nullptr || loss == nullptr) {
            rf::fail(RF_INVALID_ARG, "delay mlp: NULL labels or loss");
        }
        if (!(learning_rate > 0.0f)) rf::fail(RF_INVALID_ARG, "delay mlp: learning rate <= 0");
        rf::delay_mlp_train_step(*h, features, labels, batch, learning_rate, loss,
                                 rf::as_stream(stream));
    });
}

}  // extern "C"
