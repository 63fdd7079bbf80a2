// mlp.hip -- BF16 MFMA training and inference kernels of the ladder surrogate.
//
// One GEMM kernel serves the forward pass (X . W), the input gradient
// (G . W^T) and the weight gradient (X^T . G): operands are addressed through
// (row, k) strides, so the transposes are free.  Each wavefront computes a
// 32 x 32 tile with v_mfma_f32_32x32x16_bf16 (FP32 accumulate; a last step of
// v_mfma_f32_32x32x8_bf16 when K is 8 mod 16); a workgroup of four
// wavefronts covers 64 x 64.  Epilogues fuse bias + ReLU (+ the BF16 copy
// the next layer reads) and the ReLU mask of the backward pass.
#include "mlp.hpp"

#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>

namespace ptx {

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

enum : int32_t { kEpiPlain = 0, kEpiBias = 1, kEpiBiasRelu = 2, kEpiReluMask = 3 };

__device__ __forceinline__ uint16_t dev_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (uint16_t)(u >> 16);
}

uint16_t host_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

struct Gemm {
    const uint16_t* A;
    int64_t sai, sak;  // A(i, k) = A[i * sai + k * sak]
    const uint16_t* B;
    int64_t sbk, sbj;  // B(k, j) = B[k * sbk + j * sbj]
    int32_t M, N, K;   // multiples of 32, 32 and 8
    int32_t epi;
    const float* bias;   // [N]
    const float* zmask;  // kEpiReluMask: Z(i, j) = zmask[i * ldz + j]
    int32_t ldz;
    float* C;            // FP32 result, ldc
    int32_t ldc;
    uint16_t* Cb;        // optional BF16 copy, ldcb
    int32_t ldcb;
};

__global__ __launch_bounds__(256) void k_gemm_bf16(Gemm g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const int j0 = blockIdx.x * 64 + (wave & 1) * 32;
    if (i0 >= g.M || j0 >= g.N) return;
    const int r = lane & 31, hi = lane >> 5;
    v16f acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    // 16 k per instruction (v_mfma_f32_32x32x16_bf16): lane holds
    // A(i0 + r, k0 + 8 hi .. +7) and B(k0 + 8 hi .. +7, j0 + r)
    const uint16_t* ap = g.A + (int64_t)(i0 + r) * g.sai + (int64_t)(8 * hi) * g.sak;
    const uint16_t* bp = g.B + (int64_t)(j0 + r) * g.sbj + (int64_t)(8 * hi) * g.sbk;
    int k0 = 0;
    for (; k0 + 16 <= g.K; k0 += 16) {
        v8s a, b;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            a[t] = (short)ap[(int64_t)(k0 + t) * g.sak];
            b[t] = (short)bp[(int64_t)(k0 + t) * g.sbk];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), acc,
                                                      0, 0, 0);
    }
    if (k0 < g.K) {  // K = 8 mod 16: one step of 8 (v_mfma_f32_32x32x8_bf16: k0 + 4 hi .. +3 per lane)
        const uint16_t* a8 = g.A + (int64_t)(i0 + r) * g.sai + (int64_t)(k0 + 4 * hi) * g.sak;
        const uint16_t* b8 = g.B + (int64_t)(j0 + r) * g.sbj + (int64_t)(k0 + 4 * hi) * g.sbk;
        v4s a, b;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            a[t] = (short)a8[(int64_t)t * g.sak];
            b[t] = (short)b8[(int64_t)t * g.sbk];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    // result element v of the lane: row 8 (v / 4) + 4 (lane / 32) + v % 4, column lane % 32
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int i = i0 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
        const int j = j0 + r;
        float x = acc[v];
        if (g.epi == kEpiBias || g.epi == kEpiBiasRelu) x += g.bias[j];
        if (g.epi == kEpiReluMask && !(g.zmask[(int64_t)i * g.ldz + j] > 0.0f)) x = 0.0f;
        g.C[(int64_t)i * g.ldc + j] = x;
        if (g.Cb) g.Cb[(int64_t)i * g.ldcb + j] = dev_bf16(g.epi == kEpiBiasRelu ? fmaxf(x, 0.0f) : x);
    }
}

// Output head and loss.  y = softplus(o); mode 1: squared error to t;
// mode 2: one-sided penalty max(0, t - y)^2; mode 0: no loss.  Writes
// dLoss/do (scaled by inv_b) in FP32 and BF16.
__global__ void k_head(const float* __restrict__ Z, const float* __restrict__ t, const uint8_t* __restrict__ mode,
                       int32_t batch, int32_t bpad, int32_t n, float inv_b, float* __restrict__ y,
                       float* __restrict__ G, uint16_t* __restrict__ Gb, float* __restrict__ lossel) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)bpad * n) return;
    const int i = (int)(idx / n);
    const float o = Z[idx];
    const float e = expf(-fabsf(o));
    const float sp = fmaxf(o, 0.0f) + log1pf(e);         // softplus, stable for both signs
    const float sig = o >= 0.0f ? 1.0f / (1.0f + e) : e / (1.0f + e);
    y[idx] = sp;
    if (!mode) return;
    const int md = i < batch ? mode[idx] : 0;
    float l = 0.0f, dy = 0.0f;
    if (md == 1) {
        const float d = sp - t[idx];
        l = d * d;
        dy = 2.0f * d;
    } else if (md == 2) {
        const float d = t[idx] - sp;
        if (d > 0.0f) {
            l = d * d;
            dy = -2.0f * d;
        }
    }
    const float gr = dy * sig * inv_b;
    G[idx] = gr;
    Gb[idx] = dev_bf16(gr);
    lossel[idx] = l * inv_b;
}

__global__ void k_colsum(const float* __restrict__ G, int32_t ld, int32_t rows, int32_t n, float* __restrict__ out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float s = 0.0f;
    for (int i = 0; i < rows; ++i) s += G[(int64_t)i * ld + j];
    out[j] = s;
}

// Deterministic sum of n floats into *out (one workgroup of 256 threads).
__global__ __launch_bounds__(256) void k_sum(const float* __restrict__ x, int64_t n, float* __restrict__ out) {
    __shared__ float part[256];
    float s = 0.0f;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += x[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = part[0];
}

__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, uint16_t* __restrict__ wb, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = dev_bf16(wi);
}

__global__ void k_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ wb, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = dev_bf16(w[i]);
}

// Minibatch rows from the resident training set (rows >= batch are zero).
__global__ void k_gather_rows(const uint16_t* __restrict__ sx, const float* __restrict__ st,
                              const uint8_t* __restrict__ sm, const int32_t* __restrict__ idx, int32_t batch,
                              int32_t bpad, int32_t kx, int32_t nt, uint16_t* __restrict__ X, float* __restrict__ t,
                              uint8_t* __restrict__ m) {
    const int i = blockIdx.x;
    if (i >= bpad) return;
    const int64_t src = i < batch ? (int64_t)idx[i] : -1;
    for (int k = threadIdx.x; k < kx; k += blockDim.x) X[(int64_t)i * kx + k] = src >= 0 ? sx[src * kx + k] : 0;
    for (int k = threadIdx.x; k < nt; k += blockDim.x) {
        t[(int64_t)i * nt + k] = src >= 0 ? st[src * nt + k] : 0.0f;
        m[(int64_t)i * nt + k] = src >= 0 ? sm[src * nt + k] : 0;
    }
}

int32_t pad32(int32_t x) { return (x + 31) / 32 * 32; }

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

Status launch_gemm(const Gemm& g, hipStream_t s) {
    if (g.M % 32 || g.N % 32 || g.K % 8 || g.M <= 0 || g.N <= 0 || g.K <= 0)
        return Status::err(PTX_ERR_INVALID, "gemm shape must be multiples of (32, 32, 8)");
    hipLaunchKernelGGL(k_gemm_bf16, dim3((g.N + 63) / 64, (g.M + 63) / 64), dim3(256), 0, s, g);
    return hip_check(hipGetLastError(), "k_gemm_bf16");
}

}  // namespace

Mlp::~Mlp() {
    if (stream) (void)hipStreamDestroy(stream);
}

Status Mlp::create(const int32_t* d, int32_t ndims, int32_t maxb, uint64_t seed) {
    if (ndims < 2 || ndims > 16) return Status::err(PTX_ERR_INVALID, "an MLP needs 2..16 layer widths");
    if (maxb <= 0 || maxb > 65536) return Status::err(PTX_ERR_INVALID, "max_batch must be in 1..65536");
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count <= 0)
        return Status::err(PTX_ERR_DEVICE, "no HIP device");
    nl = ndims - 1;
    dims.assign(d, d + ndims);
    pdims.resize(ndims);
    for (int i = 0; i < ndims; ++i) {
        if (dims[i] <= 0 || dims[i] > 8192) return Status::err(PTX_ERR_INVALID, "layer widths must be in 1..8192");
        pdims[i] = pad32(dims[i]);
    }
    max_batch = pad32(maxb);
    PTX_TRY(hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate"));
    W.resize(nl); b.resize(nl); mW.resize(nl); vW.resize(nl); mb.resize(nl); vb.resize(nl);
    gW.resize(nl); gb.resize(nl); Wb.resize(nl); X.resize(nl); Z.resize(nl);
    nparams = 0;
    int32_t pmax = 0;
    for (int l = 0; l < nl; ++l) {
        const size_t nw = (size_t)pdims[l] * pdims[l + 1];
        PTX_TRY(W[l].alloc(nw)); PTX_TRY(mW[l].alloc(nw)); PTX_TRY(vW[l].alloc(nw)); PTX_TRY(gW[l].alloc(nw));
        PTX_TRY(Wb[l].alloc(nw));
        PTX_TRY(b[l].alloc(pdims[l + 1])); PTX_TRY(mb[l].alloc(pdims[l + 1])); PTX_TRY(vb[l].alloc(pdims[l + 1]));
        PTX_TRY(gb[l].alloc(pdims[l + 1]));
        PTX_TRY(X[l].alloc((size_t)max_batch * pdims[l]));
        PTX_TRY(Z[l].alloc((size_t)max_batch * pdims[l + 1]));
        nparams += (int64_t)dims[l] * dims[l + 1] + dims[l + 1];
        pmax = std::max(pmax, std::max(pdims[l], pdims[l + 1]));
    }
    PTX_TRY(G.alloc((size_t)max_batch * pmax)); PTX_TRY(G2.alloc((size_t)max_batch * pmax));
    PTX_TRY(Gb.alloc((size_t)2 * max_batch * pmax));
    const size_t nout = (size_t)max_batch * pdims[nl];
    PTX_TRY(y.alloc(nout)); PTX_TRY(target.alloc(nout)); PTX_TRY(lossel.alloc(nout)); PTX_TRY(mode.alloc(nout));
    PTX_TRY(lossacc.alloc(1));
    // He-uniform weights from a splitmix64 stream, zero biases.
    uint64_t s = seed;
    std::vector<float> host;
    for (int l = 0; l < nl; ++l) {
        const double lim = std::sqrt(6.0 / dims[l]);
        host.assign((size_t)pdims[l] * pdims[l + 1], 0.0f);
        for (int i = 0; i < dims[l]; ++i)
            for (int j = 0; j < dims[l + 1]; ++j) {
                const double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
                host[(size_t)i * pdims[l + 1] + j] = (float)((2.0 * u - 1.0) * lim);
            }
        PTX_TRY(W[l].upload(host.data(), host.size(), 0, stream));
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((host.size() + 255) / 256)), dim3(256), 0, stream, W[l].p,
                           Wb[l].p, (int64_t)host.size());
        PTX_TRY(hip_check(hipStreamSynchronize(stream), "mlp init"));
    }
    return Status::ok();
}

Status Mlp::set_params(const float* p, int64_t count) {
    if (count != nparams) return Status::err(PTX_ERR_INVALID, "parameter count mismatch");
    std::vector<float> host;
    int64_t o = 0;
    for (int l = 0; l < nl; ++l) {
        host.assign((size_t)pdims[l] * pdims[l + 1], 0.0f);
        for (int i = 0; i < dims[l]; ++i)
            for (int j = 0; j < dims[l + 1]; ++j) host[(size_t)i * pdims[l + 1] + j] = p[o++];
        PTX_TRY(W[l].upload(host.data(), host.size(), 0, stream));
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((host.size() + 255) / 256)), dim3(256), 0, stream, W[l].p,
                           Wb[l].p, (int64_t)host.size());
        std::vector<float> hb(pdims[l + 1], 0.0f);
        for (int j = 0; j < dims[l + 1]; ++j) hb[j] = p[o++];
        PTX_TRY(b[l].upload(hb.data(), hb.size(), 0, stream));
        PTX_TRY(hip_check(hipStreamSynchronize(stream), "set_params"));
    }
    // a new parameter set restarts the optimiser
    for (int l = 0; l < nl; ++l) {
        PTX_TRY(hip_check(hipMemsetAsync(mW[l].p, 0, mW[l].n * sizeof(float), stream), "memset"));
        PTX_TRY(hip_check(hipMemsetAsync(vW[l].p, 0, vW[l].n * sizeof(float), stream), "memset"));
        PTX_TRY(hip_check(hipMemsetAsync(mb[l].p, 0, mb[l].n * sizeof(float), stream), "memset"));
        PTX_TRY(hip_check(hipMemsetAsync(vb[l].p, 0, vb[l].n * sizeof(float), stream), "memset"));
    }
    adam_t = 0;
    return hip_check(hipStreamSynchronize(stream), "set_params");
}

static Status fetch(const std::vector<DevBuf<float>>& Wv, const std::vector<DevBuf<float>>& bv,
                    const std::vector<int32_t>& dims, const std::vector<int32_t>& pdims, int nl, hipStream_t s,
                    float* p) {
    int64_t o = 0;
    std::vector<float> host;
    for (int l = 0; l < nl; ++l) {
        host.resize((size_t)pdims[l] * pdims[l + 1]);
        PTX_TRY(Wv[l].download(host.data(), host.size(), 0, s));
        for (int i = 0; i < dims[l]; ++i)
            for (int j = 0; j < dims[l + 1]; ++j) p[o++] = host[(size_t)i * pdims[l + 1] + j];
        host.resize(pdims[l + 1]);
        PTX_TRY(bv[l].download(host.data(), host.size(), 0, s));
        for (int j = 0; j < dims[l + 1]; ++j) p[o++] = host[j];
    }
    return Status::ok();
}

Status Mlp::get_params(float* p, int64_t count) {
    if (count != nparams) return Status::err(PTX_ERR_INVALID, "parameter count mismatch");
    return fetch(W, b, dims, pdims, nl, stream, p);
}

Status Mlp::get_grads(float* p, int64_t count) {
    if (count != nparams) return Status::err(PTX_ERR_INVALID, "parameter count mismatch");
    return fetch(gW, gb, dims, pdims, nl, stream, p);
}

Status Mlp::upload_input(const float* x, int32_t batch) {
    const int32_t bpad = pad32(batch);
    std::vector<uint16_t> hx((size_t)bpad * pdims[0], 0);
    for (int i = 0; i < batch; ++i)
        for (int k = 0; k < dims[0]; ++k) hx[(size_t)i * pdims[0] + k] = host_bf16(x[(size_t)i * dims[0] + k]);
    PTX_TRY(X[0].upload(hx.data(), hx.size(), 0, stream));
    return hip_check(hipStreamSynchronize(stream), "upload");
}

Status Mlp::run_forward(int32_t bpad) {
    for (int l = 0; l < nl; ++l) {
        Gemm g{};
        g.A = X[l].p; g.sai = pdims[l]; g.sak = 1;
        g.B = Wb[l].p; g.sbk = pdims[l + 1]; g.sbj = 1;
        g.M = bpad; g.N = pdims[l + 1]; g.K = pdims[l];
        g.bias = b[l].p;
        g.C = Z[l].p; g.ldc = pdims[l + 1];
        if (l + 1 < nl) {
            g.epi = kEpiBiasRelu;
            g.Cb = X[l + 1].p; g.ldcb = pdims[l + 1];
        } else {
            g.epi = kEpiBias;
        }
        PTX_TRY(launch_gemm(g, stream));
    }
    return Status::ok();
}

Status Mlp::run_backward(int32_t bpad, int32_t batch) {
    (void)batch;
    float* Gc = G.p;   // FP32 gradient of the current layer's pre-activation
    float* Gn = G2.p;
    const size_t half = Gb.n / 2;
    uint16_t* Gbc = Gb.p;
    uint16_t* Gbn = Gb.p + half;
    for (int l = nl - 1; l >= 0; --l) {
        const int32_t din = pdims[l], dout = pdims[l + 1];
        // dW_l = X_l^T . bf16(G_l)
        Gemm gw{};
        gw.A = X[l].p; gw.sai = 1; gw.sak = din;
        gw.B = Gbc; gw.sbk = dout; gw.sbj = 1;
        gw.M = din; gw.N = dout; gw.K = bpad;
        gw.epi = kEpiPlain;
        gw.C = gW[l].p; gw.ldc = dout;
        PTX_TRY(launch_gemm(gw, stream));
        hipLaunchKernelGGL(k_colsum, dim3((dout + 255) / 256), dim3(256), 0, stream, Gc, dout, bpad, dout, gb[l].p);
        if (l > 0) {
            // G_{l-1} = (bf16(G_l) . bf16(W_l)^T) * [Z_{l-1} > 0]
            Gemm gx{};
            gx.A = Gbc; gx.sai = dout; gx.sak = 1;
            gx.B = Wb[l].p; gx.sbk = 1; gx.sbj = dout;
            gx.M = bpad; gx.N = din; gx.K = dout;
            gx.epi = kEpiReluMask;
            gx.zmask = Z[l - 1].p; gx.ldz = din;
            gx.C = Gn; gx.ldc = din;
            gx.Cb = Gbn; gx.ldcb = din;
            PTX_TRY(launch_gemm(gx, stream));
            std::swap(Gc, Gn);
            std::swap(Gbc, Gbn);
        }
    }
    return hip_check(hipGetLastError(), "mlp backward");
}

Status Mlp::forward(const float* x, int32_t batch, float* yout) {
    if (batch <= 0 || batch > max_batch) return Status::err(PTX_ERR_INVALID, "batch must be in 1..max_batch");
    const int32_t bpad = pad32(batch);
    PTX_TRY(upload_input(x, batch));
    PTX_TRY(run_forward(bpad));
    const int32_t n = pdims[nl];
    const int64_t tot = (int64_t)bpad * n;
    hipLaunchKernelGGL(k_head, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, Z[nl - 1].p, nullptr,
                       nullptr, batch, bpad, n, 1.0f, y.p, nullptr, nullptr, nullptr);
    std::vector<float> hy((size_t)tot);
    PTX_TRY(y.download(hy.data(), hy.size(), 0, stream));
    for (int i = 0; i < batch; ++i)
        for (int j = 0; j < dims[nl]; ++j) yout[(size_t)i * dims[nl] + j] = hy[(size_t)i * n + j];
    return Status::ok();
}

Status Mlp::loss_grad(const float* x, const float* t, const uint8_t* m, int32_t batch, double* loss) {
    if (batch <= 0 || batch > max_batch) return Status::err(PTX_ERR_INVALID, "batch must be in 1..max_batch");
    const int32_t bpad = pad32(batch);
    const int32_t n = pdims[nl];
    PTX_TRY(upload_input(x, batch));
    std::vector<float> ht((size_t)bpad * n, 0.0f);
    std::vector<uint8_t> hm((size_t)bpad * n, 0);
    for (int i = 0; i < batch; ++i)
        for (int j = 0; j < dims[nl]; ++j) {
            ht[(size_t)i * n + j] = t[(size_t)i * dims[nl] + j];
            hm[(size_t)i * n + j] = m[(size_t)i * dims[nl] + j];
        }
    PTX_TRY(target.upload(ht.data(), ht.size(), 0, stream));
    PTX_TRY(mode.upload(hm.data(), hm.size(), 0, stream));
    PTX_TRY(run_forward(bpad));
    const int64_t tot = (int64_t)bpad * n;
    hipLaunchKernelGGL(k_head, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, Z[nl - 1].p, target.p,
                       mode.p, batch, bpad, n, 1.0f / (float)batch, y.p, G.p, Gb.p, lossel.p);
    PTX_TRY(run_backward(bpad, batch));
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, stream, lossel.p, tot, lossacc.p);
    float l = 0.0f;
    PTX_TRY(lossacc.download(&l, 1, 0, stream));
    if (loss) *loss = l;
    return Status::ok();
}

Status Mlp::adam(float lr, float b1, float b2, float eps) {
    ++adam_t;
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)adam_t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)adam_t)));
    for (int l = 0; l < nl; ++l) {
        const int64_t nw = (int64_t)W[l].n, nb = (int64_t)b[l].n;
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, stream, W[l].p, mW[l].p,
                           vW[l].p, gW[l].p, Wb[l].p, nw, lr, b1, b2, eps, c1, c2);
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, stream, b[l].p, mb[l].p,
                           vb[l].p, gb[l].p, (uint16_t*)nullptr, nb, lr, b1, b2, eps, c1, c2);
    }
    return hip_check(hipGetLastError(), "k_adam");
}

Status Mlp::train(const float* x, const float* t, const uint8_t* m, int64_t n, int32_t batch, int32_t epochs,
                  float lr, uint64_t seed, double* epoch_loss) {
    if (batch <= 0 || batch > max_batch) return Status::err(PTX_ERR_INVALID, "batch must be in 1..max_batch");
    if (n < batch) return Status::err(PTX_ERR_INVALID, "fewer samples than one batch");
    if (epochs <= 0) return Status::err(PTX_ERR_INVALID, "epochs must be positive");
    const int32_t kx = pdims[0], nt = pdims[nl], bpad = pad32(batch);
    {
        std::vector<uint16_t> hx((size_t)n * kx, 0);
        std::vector<float> ht((size_t)n * nt, 0.0f);
        std::vector<uint8_t> hm((size_t)n * nt, 0);
        for (int64_t i = 0; i < n; ++i) {
            for (int k = 0; k < dims[0]; ++k) hx[(size_t)i * kx + k] = host_bf16(x[(size_t)i * dims[0] + k]);
            for (int j = 0; j < dims[nl]; ++j) {
                ht[(size_t)i * nt + j] = t[(size_t)i * dims[nl] + j];
                hm[(size_t)i * nt + j] = m[(size_t)i * dims[nl] + j];
            }
        }
        PTX_TRY(dx.alloc(hx.size())); PTX_TRY(dt.alloc(ht.size())); PTX_TRY(dm.alloc(hm.size()));
        PTX_TRY(dx.upload(hx.data(), hx.size(), 0, stream));
        PTX_TRY(dt.upload(ht.data(), ht.size(), 0, stream));
        PTX_TRY(dm.upload(hm.data(), hm.size(), 0, stream));
        PTX_TRY(hip_check(hipStreamSynchronize(stream), "train upload"));
    }
    const int64_t nb = n / batch;
    PTX_TRY(idx.alloc((size_t)n));
    DevBuf<float> acc;
    PTX_TRY(acc.alloc((size_t)nb));
    std::vector<int32_t> perm((size_t)n);
    std::vector<float> hl((size_t)nb);
    uint64_t s = seed;
    const int64_t tot = (int64_t)bpad * nt;
    for (int e = 0; e < epochs; ++e) {
        std::iota(perm.begin(), perm.end(), 0);
        for (int64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[splitmix64(s) % (uint64_t)(i + 1)]);
        PTX_TRY(idx.upload(perm.data(), perm.size(), 0, stream));
        for (int64_t bi = 0; bi < nb; ++bi) {
            hipLaunchKernelGGL(k_gather_rows, dim3(bpad), dim3(256), 0, stream, dx.p, dt.p, dm.p,
                               idx.p + bi * batch, batch, bpad, kx, nt, X[0].p, target.p, mode.p);
            PTX_TRY(run_forward(bpad));
            hipLaunchKernelGGL(k_head, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, Z[nl - 1].p,
                               target.p, mode.p, batch, bpad, nt, 1.0f / (float)batch, y.p, G.p, Gb.p, lossel.p);
            PTX_TRY(run_backward(bpad, batch));
            hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, stream, lossel.p, tot, acc.p + bi);
            PTX_TRY(adam(lr, 0.9f, 0.999f, 1e-8f));
        }
        PTX_TRY(acc.download(hl.data(), hl.size(), 0, stream));  // also keeps perm alive until the epoch ends
        double sum = 0.0;
        for (float v : hl) sum += v;
        if (epoch_loss) epoch_loss[e] = sum / (double)nb;
    }
    return Status::ok();
}

Status selftest_gemm_bf16(int32_t M, int32_t N, int32_t K, const uint16_t* A, const uint16_t* B, float* C) {
    if (M <= 0 || N <= 0 || K <= 0 || M % 32 || N % 32 || K % 8)
        return Status::err(PTX_ERR_INVALID, "gemm self-test shape must be multiples of (32, 32, 8)");
    DevBuf<uint16_t> a, bb;
    DevBuf<float> c;
    PTX_TRY(a.alloc((size_t)M * K));
    PTX_TRY(bb.alloc((size_t)K * N));
    PTX_TRY(c.alloc((size_t)M * N));
    PTX_TRY(a.upload(A, (size_t)M * K, 0, 0));
    PTX_TRY(bb.upload(B, (size_t)K * N, 0, 0));
    Gemm g{};
    g.A = a.p; g.sai = K; g.sak = 1;
    g.B = bb.p; g.sbk = N; g.sbj = 1;
    g.M = M; g.N = N; g.K = K;
    g.epi = kEpiPlain;
    g.C = c.p; g.ldc = N;
    PTX_TRY(launch_gemm(g, 0));
    return c.download(C, (size_t)M * N, 0, 0);
}

// ---------------------------------------------------------------------------
// Forward pass with no allocation (ptx_mlp_forward_noalloc): the same kernels
// and rounding points as Mlp::forward, on a file-scope device array instead
// of buffers the library allocates.  The host addresses the array only
// through the device pointer hipGetSymbolAddress returns (kept per device);
// its name in host code is a host-side stand-in, never a device address.
// Layout of one call, each part 256-byte aligned: the flat FP32 parameters
// as given; per layer the padded BF16 weights and padded FP32 biases; the
// FP32 input rows; per layer the BF16 input X_l and the FP32 pre-activation
// Z_l of the padded batch; the softplus outputs.  Every byte a kernel reads
// is written earlier in the same call (the pack kernels write the padding
// zeros), so nothing depends on what the previous call left there.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu[kMaxScratchDevices];  // one call at a time per device: they share the array
uint8_t* g_scratch_ptr[kMaxScratchDevices] = {};

// dst[i][j] (prow x pcol, BF16) = src[i][j] (rows x cols, FP32) inside, 0 in the padding
__global__ void k_pack_bf16(const float* __restrict__ src, int32_t rows, int32_t cols, int32_t prow, int32_t pcol,
                            uint16_t* __restrict__ dst) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)prow * pcol) return;
    const int32_t i = (int32_t)(idx / pcol), j = (int32_t)(idx % pcol);
    dst[idx] = (i < rows && j < cols) ? dev_bf16(src[(int64_t)i * cols + j]) : (uint16_t)0;
}

// dst[j] (pn) = src[j] (n) inside, 0 in the padding
__global__ void k_pack_f32(const float* __restrict__ src, int32_t n, int32_t pn, float* __restrict__ dst) {
    const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < pn) dst[j] = j < n ? src[j] : 0.0f;
}

size_t take(size_t& off, size_t bytes) {
    const size_t at = off;
    off += (bytes + 255) / 256 * 256;
    return at;
}

}  // namespace

Status mlp_forward_noalloc(const int32_t* dims, int32_t ndims, const float* params, int64_t count, const float* x,
                           int32_t batch, float* yout) {
    if (ndims < 2 || ndims > 16) return Status::err(PTX_ERR_INVALID, "an MLP needs 2..16 layer widths");
    if (batch <= 0) return Status::err(PTX_ERR_INVALID, "batch must be positive");
    const int nl = ndims - 1;
    int32_t pd[16];
    int64_t np = 0;
    for (int i = 0; i < ndims; ++i) {
        if (dims[i] <= 0 || dims[i] > 8192) return Status::err(PTX_ERR_INVALID, "layer widths must be in 1..8192");
        pd[i] = pad32(dims[i]);
        if (i > 0) np += (int64_t)dims[i - 1] * dims[i] + dims[i];
    }
    if (count != np) return Status::err(PTX_ERR_INVALID, "parameter count mismatch");
    if (batch > 65536) return Status::err(PTX_ERR_INVALID, "batch must be in 1..65536");
    const int32_t bpad = pad32(batch);
    // layout (offsets into the array)
    size_t off = 0;
    const size_t o_par = take(off, (size_t)np * 4);
    size_t o_w[16], o_b[16], o_x[16], o_z[16];
    for (int l = 0; l < nl; ++l) {
        o_w[l] = take(off, (size_t)pd[l] * pd[l + 1] * 2);
        o_b[l] = take(off, (size_t)pd[l + 1] * 4);
    }
    const size_t o_in = take(off, (size_t)batch * dims[0] * 4);
    for (int l = 0; l < nl; ++l) {
        o_x[l] = take(off, (size_t)bpad * pd[l] * 2);
        o_z[l] = take(off, (size_t)bpad * pd[l + 1] * 4);
    }
    const size_t o_y = take(off, (size_t)bpad * pd[nl] * 4);
    if (off > kNoAllocScratchBytes)
        return Status::err(PTX_ERR_UNSUPPORTED, "network and batch need " + std::to_string(off) +
                                                    " bytes of device scratch; the fixed array holds " +
                                                    std::to_string(kNoAllocScratchBytes));
    int dev = 0;
    PTX_TRY(hip_check(hipGetDevice(&dev), "hipGetDevice"));
    if (dev < 0 || dev >= kMaxScratchDevices) return Status::err(PTX_ERR_UNSUPPORTED, "device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu[dev]);
    if (!g_scratch_ptr[dev]) {  // once per device, with that device current
        void* p = nullptr;
        PTX_TRY(hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"));
        g_scratch_ptr[dev] = static_cast<uint8_t*>(p);
    }
    uint8_t* const base = g_scratch_ptr[dev];
    const hipStream_t s = 0;
    PTX_TRY(hip_check(hipMemcpy(base + o_par, params, (size_t)np * 4, hipMemcpyHostToDevice), "copy parameters"));
    PTX_TRY(hip_check(hipMemcpy(base + o_in, x, (size_t)batch * dims[0] * 4, hipMemcpyHostToDevice), "copy input"));
    const float* P = reinterpret_cast<const float*>(base + o_par);
    int64_t po = 0;
    for (int l = 0; l < nl; ++l) {
        const int64_t tw = (int64_t)pd[l] * pd[l + 1];
        hipLaunchKernelGGL(k_pack_bf16, dim3((unsigned)((tw + 255) / 256)), dim3(256), 0, s, P + po, dims[l],
                           dims[l + 1], pd[l], pd[l + 1], reinterpret_cast<uint16_t*>(base + o_w[l]));
        po += (int64_t)dims[l] * dims[l + 1];
        hipLaunchKernelGGL(k_pack_f32, dim3((unsigned)((pd[l + 1] + 255) / 256)), dim3(256), 0, s, P + po,
                           dims[l + 1], pd[l + 1], reinterpret_cast<float*>(base + o_b[l]));
        po += dims[l + 1];
    }
    const int64_t tx = (int64_t)bpad * pd[0];
    hipLaunchKernelGGL(k_pack_bf16, dim3((unsigned)((tx + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const float*>(base + o_in), batch, dims[0], bpad, pd[0],
                       reinterpret_cast<uint16_t*>(base + o_x[0]));
    PTX_TRY(hip_check(hipGetLastError(), "pack kernels"));
    for (int l = 0; l < nl; ++l) {  // the layers of Mlp::run_forward
        Gemm g{};
        g.A = reinterpret_cast<const uint16_t*>(base + o_x[l]); g.sai = pd[l]; g.sak = 1;
        g.B = reinterpret_cast<const uint16_t*>(base + o_w[l]); g.sbk = pd[l + 1]; g.sbj = 1;
        g.M = bpad; g.N = pd[l + 1]; g.K = pd[l];
        g.bias = reinterpret_cast<const float*>(base + o_b[l]);
        g.C = reinterpret_cast<float*>(base + o_z[l]); g.ldc = pd[l + 1];
        if (l + 1 < nl) {
            g.epi = kEpiBiasRelu;
            g.Cb = reinterpret_cast<uint16_t*>(base + o_x[l + 1]); g.ldcb = pd[l + 1];
        } else {
            g.epi = kEpiBias;
        }
        PTX_TRY(launch_gemm(g, s));
    }
    const int32_t n = pd[nl];
    const int64_t tot = (int64_t)bpad * n;
    hipLaunchKernelGGL(k_head, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const float*>(base + o_z[nl - 1]), nullptr, nullptr, batch, bpad, n, 1.0f,
                       reinterpret_cast<float*>(base + o_y), nullptr, nullptr, nullptr);
    PTX_TRY(hip_check(hipGetLastError(), "k_head"));
    // the rows of the batch, without the padding columns (a synchronous copy: the call returns with y written)
    return hip_check(hipMemcpy2D(yout, (size_t)dims[nl] * 4, base + o_y, (size_t)n * 4, (size_t)dims[nl] * 4,
                                 (size_t)batch, hipMemcpyDeviceToHost),
                     "copy output");
}

}  // namespace ptx
