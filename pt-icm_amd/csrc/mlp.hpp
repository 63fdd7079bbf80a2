// mlp.hpp -- the ladder surrogate: a small fully connected network trained
// and evaluated on the GPU with hand-written BF16 MFMA kernels.
//
// Layers: dims[0] -> dims[1] -> ... -> dims[L]; ReLU after every hidden
// layer; the output o is mapped to positive ladder increments y = softplus(o).
// GEMM operands are BF16 (round to nearest even), accumulation FP32; master
// weights, biases, Adam moments and gradients are FP32.
//
// Rounding points (the CPU oracle applies the same ones):
//   forward   X_0 = bf16(x);  Z_l = X_l . bf16(W_l) + b_l;
//             X_{l+1} = bf16(relu(Z_l)) for hidden layers;  o = Z_{L-1}
//   backward  G_{L-1} = dLoss/do (FP32);  dW_l = X_l^T . bf16(G_l);
//             db_l = column sums of G_l (FP32);
//             G_{l-1} = (bf16(G_l) . bf16(W_l)^T) * [Z_{l-1} > 0]
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "engine.hpp"

namespace ptx {

struct Mlp {
    int32_t nl = 0;               // number of weight layers
    std::vector<int32_t> dims;    // real widths, nl + 1 entries
    std::vector<int32_t> pdims;   // padded widths (multiples of 32)
    int32_t max_batch = 0;        // padded to a multiple of 32
    int64_t nparams = 0;          // real parameter count (W and b of every layer)
    int64_t adam_t = 0;
    hipStream_t stream = nullptr;

    // per layer l: W [pdims[l]][pdims[l+1]] FP32 + BF16 copy, b [pdims[l+1]]
    std::vector<DevBuf<float>> W, b, mW, vW, mb, vb, gW, gb;
    std::vector<DevBuf<uint16_t>> Wb;
    // activations: X[l] BF16 [max_batch][pdims[l]] (input of layer l),
    // Z[l] FP32 [max_batch][pdims[l+1]] (pre-activation of layer l)
    std::vector<DevBuf<uint16_t>> X;
    std::vector<DevBuf<float>> Z;
    DevBuf<float> G, G2;          // FP32 gradients w.r.t. pre-activations
    DevBuf<uint16_t> Gb;          // BF16 copy of G
    DevBuf<float> y, target, lossel, lossacc;
    DevBuf<uint8_t> mode;
    // resident training set
    DevBuf<uint16_t> dx;
    DevBuf<float> dt;
    DevBuf<uint8_t> dm;
    DevBuf<int32_t> idx;
    int64_t dn = 0;

    ~Mlp();
    Status create(const int32_t* d, int32_t ndims, int32_t maxb, uint64_t seed);
    Status set_params(const float* p, int64_t count);
    Status get_params(float* p, int64_t count);
    Status get_grads(float* p, int64_t count);
    Status forward(const float* x, int32_t batch, float* yout);
    Status loss_grad(const float* x, const float* t, const uint8_t* m, int32_t batch, double* loss);
    Status adam(float lr, float b1, float b2, float eps);
    Status train(const float* x, const float* t, const uint8_t* m, int64_t n, int32_t batch, int32_t epochs,
                 float lr, uint64_t seed, double* epoch_loss);

  private:
    Status upload_input(const float* x, int32_t batch);
    Status run_forward(int32_t bpad);
    Status run_backward(int32_t bpad, int32_t batch);
};

// Forward pass of the network given by dims and flat parameters, on a fixed
// device array (no allocation); y[batch][dims[ndims-1]].
Status mlp_forward_noalloc(const int32_t* dims, int32_t ndims, const float* params, int64_t count, const float* x,
                           int32_t batch, float* y);

// GEMM self-test: C[M][N] = A[M][K] . B[K][N], BF16 operands, FP32 result.
Status selftest_gemm_bf16(int32_t M, int32_t N, int32_t K, const uint16_t* A, const uint16_t* B, float* C);

}  // namespace ptx
