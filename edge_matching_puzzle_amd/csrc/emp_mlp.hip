// emp_mlp.hip -- training and inference kernels of the candidate scorer.
//
// The dense layers run on the matrix cores: v_mfma_f32_16x16x16_bf16 takes a
// 16x16 BF16 tile of A and of B per wave (four values per lane) and
// accumulates a 16x16 FP32 tile.  Each wave owns one 16x16 output tile and
// walks K in steps of 16; epilogues fuse the bias, the ReLU and the ReLU mask
// of the backward pass and write the BF16 copy the next GEMM reads.
//
//   forward   Z1 = X W1 + b1, H1 = relu(Z1)            (MFMA, K = 128)
//             Z2 = H1 W2 + b2, H2 = relu(Z2)           (MFMA, K = 256)
//             s  = H2 . w3 + b3                        (one row per thread)
//   backward  ds  = 2 wt (s - y) / sum(wt)
//             dZ2 = (ds w3^T) * [H2 > 0]               (elementwise)
//             dW2 = H1^T dZ2, dH1 = dZ2 W2^T           (MFMA)
//             dZ1 = dH1 * [H1 > 0], dW1 = X^T dZ1      (MFMA)
//   update    Adam on the FP32 master weights, then fresh BF16 copies of
//             W1 and W2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>

#include "emp_mlp.h"

namespace emp {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

#define MLP_CHECK(call)                                                                   \
    do {                                                                                  \
        hipError_t err_ = (call);                                                         \
        if (err_ != hipSuccess)                                                           \
            throw DeviceError(std::string(#call) + ": " + hipGetErrorString(err_));       \
    } while (0)

// FP32 -> BF16 with round-to-nearest-even on the discarded 16 bits.
__host__ __device__ inline uint16_t to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

enum Epilogue { kStoreF32 = 0, kBiasReluBf16 = 1, kMaskBf16 = 2 };

struct GemmArgs {
    const uint16_t *A, *B;
    uint32_t M, N, K, lda, ldb;
    float *out_f32;            // kStoreF32
    uint16_t *out_bf16;        // kBiasReluBf16, kMaskBf16
    uint32_t ldc;
    const float *bias;         // kBiasReluBf16
    const uint16_t *mask;      // kMaskBf16: keep where mask > 0 (same shape as C)
};

// C (M x N) = op(A) op(B); op(A)(m, k) = TA ? A[k lda + m] : A[m lda + k],
// op(B)(k, n) = TB ? B[n ldb + k] : B[k ldb + n].  M, N, K multiples of 16.
// Block: 4 waves on 4 consecutive 16-column tiles of one 16-row strip.
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) gemm_kernel(GemmArgs g) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t n0 = (blockIdx.x * 4 + wave) * 16, m0 = blockIdx.y * 16;
    if (n0 >= g.N || m0 >= g.M) return;
    const uint32_t r = lane & 15, q = (lane >> 4) * 4;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (uint32_t k0 = 0; k0 < g.K; k0 += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint32_t k = k0 + q + j;
            a[j] = (short)(TA ? g.A[(size_t)k * g.lda + m0 + r] : g.A[(size_t)(m0 + r) * g.lda + k]);
            b[j] = (short)(TB ? g.B[(size_t)(n0 + r) * g.ldb + k] : g.B[(size_t)k * g.ldb + n0 + r]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    // lane holds C[m0 + q + j][n0 + r], j = 0..3
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const size_t at = (size_t)(m0 + q + j) * g.ldc + n0 + r;
        if (EPI == kStoreF32) g.out_f32[at] = acc[j];
        else if (EPI == kBiasReluBf16) g.out_bf16[at] = to_bf16(fmaxf(acc[j] + g.bias[n0 + r], 0.f));
        else g.out_bf16[at] = from_bf16(g.mask[at]) > 0.f ? to_bf16(acc[j]) : (uint16_t)0;
    }
}

template <bool TA, bool TB, int EPI>
void gemm(hipStream_t s, const GemmArgs &g) {
    const dim3 grid((g.N / 16 + 3) / 4, g.M / 16);
    hipLaunchKernelGGL((gemm_kernel<TA, TB, EPI>), grid, dim3(256), 0, s, g);
}

// Rows idx[r] of the feature matrix (BF16, kMlpInPad wide), targets (none
// when y is null) and unit weights; rows past `rows` are zero with weight 0.
__global__ void gather_kernel(const uint16_t *X, const float *y, const uint32_t *idx, uint32_t rows, uint32_t padded,
                              uint16_t *Xb, float *yb, float *wb) {
    const uint32_t r = blockIdx.x, f = threadIdx.x;
    if (r >= padded) return;
    const bool on = r < rows;
    const uint32_t src = on ? idx[r] : 0;
    Xb[(size_t)r * kMlpInPad + f] = on ? X[(size_t)src * kMlpInPad + f] : (uint16_t)0;
    if (f == 0) {
        yb[r] = on && y ? y[src] : 0.f;
        wb[r] = on ? 1.f : 0.f;
    }
}

// s = H2 . w3 + b3; ds = 2 wt (s - y) / wsum; loss += wt (s - y)^2 / wsum.
__global__ void out_layer_kernel(const uint16_t *H2, const float *params, const float *y, const float *wt, uint32_t rows,
                                 float wsum, float *score, float *ds, float *loss) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const float *w3 = params + kMlpOffW3;
    float s = params[kMlpOffB3];
    for (uint32_t n = 0; n < kMlpHidden; ++n) s += from_bf16(H2[(size_t)r * kMlpHidden + n]) * w3[n];
    score[r] = s;
    if (ds) {
        const float e = s - y[r];
        ds[r] = 2.f * wt[r] * e / wsum;
        atomicAdd(loss, wt[r] * e * e / wsum);
    }
}

// Gradients of w3 and b3: gw3[n] = sum_r H2[r][n] ds[r], gb3 = sum_r ds[r].
// Column reductions over the batch rows in two stages: block b sums rows
// [kSumRows * b, kSumRows * (b + 1)) of every column (thread n: column n)
// into partial[b][n], then one block adds the partials in block order --
// the same order on every run.
constexpr uint32_t kSumRows = 32;

// partial[b][n] = sum of D[r][n] (times ds[r] when ds is given); with ds,
// partial[b][kMlpHidden] = sum of ds[r].
__global__ void colsum_partial_kernel(const uint16_t *D, const float *ds, uint32_t rows, float *partial) {
    const uint32_t n = threadIdx.x, r0 = blockIdx.x * kSumRows;
    const uint32_t r1 = r0 + kSumRows < rows ? r0 + kSumRows : rows;
    float g = 0.f, gb = 0.f;
    for (uint32_t r = r0; r < r1; ++r) {
        const float w = ds ? ds[r] : 1.f;
        g += from_bf16(D[(size_t)r * kMlpHidden + n]) * w;
        gb += w;
    }
    partial[(size_t)blockIdx.x * (kMlpHidden + 1) + n] = g;
    if (n == 0) partial[(size_t)blockIdx.x * (kMlpHidden + 1) + kMlpHidden] = gb;
}

// grads[off + n] = sum over the blocks of partial[b][n]; with bias_off set,
// grads[bias_off] = sum of partial[b][kMlpHidden].
__global__ void colsum_final_kernel(const float *partial, uint32_t blocks, float *grads, uint32_t off,
                                    int32_t bias_off) {
    const uint32_t n = threadIdx.x;
    float g = 0.f, gb = 0.f;
    for (uint32_t b = 0; b < blocks; ++b) {
        g += partial[(size_t)b * (kMlpHidden + 1) + n];
        gb += partial[(size_t)b * (kMlpHidden + 1) + kMlpHidden];
    }
    grads[off + n] = g;
    if (bias_off >= 0 && n == 0) grads[bias_off] = gb;
}

// dZ2 = (ds w3^T) * [H2 > 0] in BF16.
__global__ void dz2_kernel(const uint16_t *H2, const float *ds, const float *params, uint32_t rows, uint16_t *dZ2) {
    const uint32_t r = blockIdx.x, n = threadIdx.x;
    if (r >= rows) return;
    const size_t at = (size_t)r * kMlpHidden + n;
    dZ2[at] = from_bf16(H2[at]) > 0.f ? to_bf16(ds[r] * params[kMlpOffW3 + n]) : (uint16_t)0;
}


__global__ void adam_kernel(float *p, const float *g, float *m, float *v, uint32_t n, float lr_t, float b1, float b2,
                            float eps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

// BF16 copies of W1 and W2 from the FP32 master.
__global__ void refresh_kernel(const float *params, uint16_t *W1b, uint16_t *W2b) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kMlpInPad * kMlpHidden) W1b[i] = to_bf16(params[kMlpOffW1 + i]);
    if (i < kMlpHidden * kMlpHidden) W2b[i] = to_bf16(params[kMlpOffW2 + i]);
}

struct Net {
    uint32_t cap;   // rows per batch buffer
    float *params, *grads, *m, *v;
    uint16_t *W1b, *W2b, *Xb, *H1, *H2, *dZ1, *dZ2;
    float *yb, *wb, *score, *ds, *loss, *partial;
    uint32_t *idx;

    Net(Workspace &ws, uint32_t rows) : cap(rows) {
        params = ws.get<float>("mlp_params", kMlpParams);
        grads = ws.get<float>("mlp_grads", kMlpParams);
        m = ws.get<float>("mlp_m", kMlpParams);
        v = ws.get<float>("mlp_v", kMlpParams);
        W1b = ws.get<uint16_t>("mlp_w1b", kMlpInPad * kMlpHidden);
        W2b = ws.get<uint16_t>("mlp_w2b", kMlpHidden * kMlpHidden);
        Xb = ws.get<uint16_t>("mlp_xb", (size_t)rows * kMlpInPad);
        H1 = ws.get<uint16_t>("mlp_h1", (size_t)rows * kMlpHidden);
        H2 = ws.get<uint16_t>("mlp_h2", (size_t)rows * kMlpHidden);
        dZ1 = ws.get<uint16_t>("mlp_dz1", (size_t)rows * kMlpHidden);
        dZ2 = ws.get<uint16_t>("mlp_dz2", (size_t)rows * kMlpHidden);
        yb = ws.get<float>("mlp_yb", rows);
        wb = ws.get<float>("mlp_wb", rows);
        score = ws.get<float>("mlp_score", rows);
        ds = ws.get<float>("mlp_ds", rows);
        loss = ws.get<float>("mlp_loss", 1);
        idx = ws.get<uint32_t>("mlp_idx", rows);
        partial = ws.get<float>("mlp_partial", (size_t)((rows + kSumRows - 1) / kSumRows) * (kMlpHidden + 1));
    }

    // grads[off + n] = column n of D summed over `rows` (weighted by ds, and
    // grads[bias_off] = the sum of ds, when ds is given)
    void colsum(hipStream_t s, const uint16_t *D, const float *w, uint32_t rows, uint32_t off, int32_t bias_off) {
        const uint32_t blocks = (rows + kSumRows - 1) / kSumRows;
        hipLaunchKernelGGL(colsum_partial_kernel, dim3(blocks), dim3(kMlpHidden), 0, s, D, w, rows, partial);
        hipLaunchKernelGGL(colsum_final_kernel, dim3(1), dim3(kMlpHidden), 0, s, partial, blocks, grads, off, bias_off);
    }

    void refresh(hipStream_t s) {
        hipLaunchKernelGGL(refresh_kernel, dim3((kMlpHidden * kMlpHidden + 255) / 256), dim3(256), 0, s, params, W1b, W2b);
    }

    // Forward over `rows` valid rows padded to `padded` (multiple of 16).
    void forward(hipStream_t s, uint32_t padded) {
        GemmArgs g{};
        g.A = Xb; g.B = W1b; g.M = padded; g.N = kMlpHidden; g.K = kMlpInPad; g.lda = kMlpInPad; g.ldb = kMlpHidden;
        g.out_bf16 = H1; g.ldc = kMlpHidden; g.bias = params + kMlpOffB1;
        gemm<false, false, kBiasReluBf16>(s, g);
        g.A = H1; g.B = W2b; g.K = kMlpHidden; g.lda = kMlpHidden; g.ldb = kMlpHidden;
        g.out_bf16 = H2; g.bias = params + kMlpOffB2;
        gemm<false, false, kBiasReluBf16>(s, g);
    }

    // Backward of the weighted MSE; grads filled; loss accumulated.
    void backward(hipStream_t s, uint32_t rows, uint32_t padded, float wsum) {
        hipLaunchKernelGGL(out_layer_kernel, dim3((padded + 255) / 256), dim3(256), 0, s, H2, params, yb, wb, padded,
                           wsum, score, ds, loss);
        colsum(s, H2, ds, padded, kMlpOffW3, (int32_t)kMlpOffB3);
        hipLaunchKernelGGL(dz2_kernel, dim3(padded), dim3(kMlpHidden), 0, s, H2, ds, params, padded, dZ2);
        colsum(s, dZ2, nullptr, padded, kMlpOffB2, -1);
        GemmArgs g{};
        // dW2 = H1^T dZ2
        g.A = H1; g.B = dZ2; g.M = kMlpHidden; g.N = kMlpHidden; g.K = padded; g.lda = kMlpHidden; g.ldb = kMlpHidden;
        g.out_f32 = grads + kMlpOffW2; g.ldc = kMlpHidden;
        gemm<true, false, kStoreF32>(s, g);
        // dZ1 = (dZ2 W2^T) * [H1 > 0]
        g = GemmArgs{};
        g.A = dZ2; g.B = W2b; g.M = padded; g.N = kMlpHidden; g.K = kMlpHidden; g.lda = kMlpHidden; g.ldb = kMlpHidden;
        g.out_bf16 = dZ1; g.ldc = kMlpHidden; g.mask = H1;
        gemm<false, true, kMaskBf16>(s, g);
        colsum(s, dZ1, nullptr, padded, kMlpOffB1, -1);
        // dW1 = X^T dZ1
        g = GemmArgs{};
        g.A = Xb; g.B = dZ1; g.M = kMlpInPad; g.N = kMlpHidden; g.K = padded; g.lda = kMlpInPad; g.ldb = kMlpHidden;
        g.out_f32 = grads + kMlpOffW1; g.ldc = kMlpHidden;
        gemm<true, false, kStoreF32>(s, g);
        (void)rows;
    }
};

// Features (n x kMlpIn FP32, host) -> device BF16 [n][kMlpInPad].
uint16_t *upload_features(hipStream_t s, Workspace &ws, const float *features, uint64_t n, const char *name) {
    std::vector<uint16_t> h((size_t)n * kMlpInPad, 0);
    for (uint64_t r = 0; r < n; ++r)
        for (uint32_t f = 0; f < kMlpIn; ++f) h[r * kMlpInPad + f] = to_bf16(features[r * kMlpIn + f]);
    uint16_t *d = ws.get<uint16_t>(name, std::max<size_t>(h.size(), 1));
    MLP_CHECK(hipMemcpyAsync(d, h.data(), h.size() * 2, hipMemcpyHostToDevice, s));
    MLP_CHECK(hipStreamSynchronize(s));
    return d;
}

uint32_t pad64(uint64_t n) { return (uint32_t)((n + 63) / 64 * 64); }

}  // namespace

void mlp_gradients_on_device(hipStream_t stream, Workspace &ws, const float *params, const float *features,
                             const float *targets, uint64_t n, float *grads, float &loss) {
    if (n == 0 || n > 65536) throw std::invalid_argument("gradient rows must be in 1..65536");
    const uint32_t padded = pad64(n);
    Net net(ws, padded);
    uint16_t *X = upload_features(stream, ws, features, n, "mlp_x_all");
    float *y = ws.get<float>("mlp_y_all", n);
    MLP_CHECK(hipMemcpyAsync(y, targets, n * 4, hipMemcpyHostToDevice, stream));
    MLP_CHECK(hipMemcpyAsync(net.params, params, kMlpParams * 4, hipMemcpyHostToDevice, stream));
    std::vector<uint32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    MLP_CHECK(hipMemcpyAsync(net.idx, idx.data(), n * 4, hipMemcpyHostToDevice, stream));
    const float zero = 0.f;
    MLP_CHECK(hipMemcpyAsync(net.loss, &zero, 4, hipMemcpyHostToDevice, stream));
    net.refresh(stream);
    hipLaunchKernelGGL(gather_kernel, dim3(padded), dim3(kMlpInPad), 0, stream, X, y, net.idx, (uint32_t)n, padded,
                       net.Xb, net.yb, net.wb);
    net.forward(stream, padded);
    net.backward(stream, (uint32_t)n, padded, (float)n);
    MLP_CHECK(hipGetLastError());
    MLP_CHECK(hipMemcpyAsync(grads, net.grads, kMlpParams * 4, hipMemcpyDeviceToHost, stream));
    MLP_CHECK(hipMemcpyAsync(&loss, net.loss, 4, hipMemcpyDeviceToHost, stream));
    MLP_CHECK(hipStreamSynchronize(stream));
}

void mlp_train_on_device(hipStream_t stream, Workspace &ws, const float *features, const float *targets, uint64_t n,
                         const MlpTrainConfig &cfg, float *params, std::vector<float> &loss_per_epoch) {
    if (n == 0 || n > (1ull << 31)) throw std::invalid_argument("training rows must be in 1..2^31");
    if (cfg.batch == 0 || cfg.batch % 64 || cfg.batch > 65536) throw std::invalid_argument("batch must be a multiple of 64, at most 65536");
    const uint32_t B = cfg.batch;
    Net net(ws, B);
    uint16_t *X = upload_features(stream, ws, features, n, "mlp_x_all");
    float *y = ws.get<float>("mlp_y_all", n);
    MLP_CHECK(hipMemcpyAsync(y, targets, n * 4, hipMemcpyHostToDevice, stream));
    MLP_CHECK(hipMemcpyAsync(net.params, params, kMlpParams * 4, hipMemcpyHostToDevice, stream));
    std::vector<float> zeros(kMlpParams, 0.f);
    MLP_CHECK(hipMemcpyAsync(net.m, zeros.data(), kMlpParams * 4, hipMemcpyHostToDevice, stream));
    MLP_CHECK(hipMemcpyAsync(net.v, zeros.data(), kMlpParams * 4, hipMemcpyHostToDevice, stream));
    net.refresh(stream);
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    uint64_t state = cfg.seed;
    auto next = [&]() { return splitmix64(state); };
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    uint64_t step = 0;
    loss_per_epoch.assign(cfg.epochs, 0.f);
    for (uint32_t e = 0; e < cfg.epochs; ++e) {
        for (uint64_t i = n; i > 1; --i) std::swap(order[i - 1], order[next() % i]);
        const float zero = 0.f;
        MLP_CHECK(hipMemcpyAsync(net.loss, &zero, 4, hipMemcpyHostToDevice, stream));
        uint64_t batches = 0;
        for (uint64_t at = 0; at < n; at += B) {
            const uint32_t rows = (uint32_t)std::min<uint64_t>(B, n - at);
            const uint32_t padded = pad64(rows);
            MLP_CHECK(hipMemcpyAsync(net.idx, order.data() + at, rows * 4, hipMemcpyHostToDevice, stream));
            hipLaunchKernelGGL(gather_kernel, dim3(padded), dim3(kMlpInPad), 0, stream, X, y, net.idx, rows, padded,
                               net.Xb, net.yb, net.wb);
            net.forward(stream, padded);
            net.backward(stream, rows, padded, (float)rows);
            ++step;
            const float lr_t = cfg.lr * std::sqrt(1.f - std::pow(b2, (float)step)) / (1.f - std::pow(b1, (float)step));
            hipLaunchKernelGGL(adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream, net.params, net.grads,
                               net.m, net.v, kMlpParams, lr_t, b1, b2, eps);
            net.refresh(stream);
            ++batches;
            // the index upload above reads host memory the next batch rewrites
            MLP_CHECK(hipStreamSynchronize(stream));
        }
        MLP_CHECK(hipGetLastError());
        float total = 0.f;
        MLP_CHECK(hipMemcpyAsync(&total, net.loss, 4, hipMemcpyDeviceToHost, stream));
        MLP_CHECK(hipStreamSynchronize(stream));
        loss_per_epoch[e] = total / (float)std::max<uint64_t>(batches, 1);
    }
    MLP_CHECK(hipMemcpyAsync(params, net.params, kMlpParams * 4, hipMemcpyDeviceToHost, stream));
    MLP_CHECK(hipStreamSynchronize(stream));
}

void mlp_score_on_device(hipStream_t stream, Workspace &ws, const float *params, const float *features, uint64_t n,
                         float *scores) {
    if (n == 0) return;
    const uint32_t B = 16384;
    Net net(ws, B);
    uint16_t *X = upload_features(stream, ws, features, n, "mlp_x_all");
    const float *y = nullptr;
    MLP_CHECK(hipMemcpyAsync(net.params, params, kMlpParams * 4, hipMemcpyHostToDevice, stream));
    net.refresh(stream);
    std::vector<uint32_t> idx(B);
    for (uint64_t at = 0; at < n; at += B) {
        const uint32_t rows = (uint32_t)std::min<uint64_t>(B, n - at);
        const uint32_t padded = pad64(rows);
        for (uint32_t r = 0; r < rows; ++r) idx[r] = (uint32_t)(at + r);
        MLP_CHECK(hipMemcpyAsync(net.idx, idx.data(), rows * 4, hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(gather_kernel, dim3(padded), dim3(kMlpInPad), 0, stream, X, y, net.idx, rows, padded,
                           net.Xb, net.yb, net.wb);
        net.forward(stream, padded);
        hipLaunchKernelGGL(out_layer_kernel, dim3((padded + 255) / 256), dim3(256), 0, stream, net.H2, net.params,
                           net.yb, net.wb, rows, 1.f, net.score, (float *)nullptr, (float *)nullptr);
        MLP_CHECK(hipGetLastError());
        MLP_CHECK(hipMemcpyAsync(scores + at, net.score, rows * 4, hipMemcpyDeviceToHost, stream));
        MLP_CHECK(hipStreamSynchronize(stream));
    }
}

}  // namespace emp
