// Region-scoring network of the learned proposal: a 3-layer perceptron
// in -> 256 -> 256 -> 1 (ReLU), FP16 weights, FP32 accumulation.
//
// Training keeps FP32 master weights and Adam moments; every forward and
// backward pass reads the FP16 copies, so the trained network is exactly the
// one inference runs.  Loss: mean squared error.  All kernels are this
// file's own (no BLAS); activations of a minibatch live in HBM between the
// passes.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "empi_hip.h"
#include "mlp_internal.hpp"

#define MLP_HIP(call)                                                                                    \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) {                                                                          \
            std::fprintf(stderr, "empi_hip mlp: %s failed: %s\n", #call, hipGetErrorString(e_));        \
            return EMPI_DEVICE;                                                                          \
        }                                                                                                \
    } while (0)

namespace {

constexpr int H = 256;       // hidden width
constexpr int ROWS = 32;     // rows per workgroup tile
constexpr int NT = 256;      // threads per workgroup (one per hidden unit)

// forward over a row tile: H1 = relu(W1 x + b1), H2 = relu(W2 H1 + b2), y = w3 H2 + b3
__global__ __launch_bounds__(NT) void k_mlp_forward(const float* __restrict__ X, int R, const int* __restrict__ Rdev,
                                                    int din,
                                                    const __half* __restrict__ W1, const float* __restrict__ b1,
                                                    const __half* __restrict__ W2, const float* __restrict__ b2,
                                                    const __half* __restrict__ w3, const float* __restrict__ b3,
                                                    float* __restrict__ y, float* __restrict__ H1out,
                                                    float* __restrict__ H2out) {
    extern __shared__ float sm[];
    float* xs = sm;                    // [ROWS][din]
    float* h1 = xs + ROWS * din;       // [ROWS][H]
    __shared__ float red[ROWS][NT / 64];
    const int j = threadIdx.x;
    const int r0 = blockIdx.x * ROWS;
    if (Rdev) R = min(R, *Rdev);
    if (r0 >= R) return;   // uniform over the workgroup
    for (int i = j; i < ROWS * din; i += NT) {
        const int r = i / din, c = i % din;
        xs[i] = r0 + r < R ? X[static_cast<long long>(r0 + r) * din + c] : 0.0f;
    }
    __syncthreads();
    float acc[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) acc[r] = b1[j];
    for (int i = 0; i < din; ++i) {
        const float w = __half2float(W1[j * din + i]);
#pragma unroll
        for (int r = 0; r < ROWS; ++r) acc[r] = fmaf(w, xs[r * din + i], acc[r]);
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const float v = fmaxf(acc[r], 0.0f);
        h1[r * H + j] = v;
        if (H1out && r0 + r < R) H1out[static_cast<long long>(r0 + r) * H + j] = v;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ROWS; ++r) acc[r] = b2[j];
    for (int i = 0; i < H; ++i) {
        const float w = __half2float(W2[j * H + i]);
#pragma unroll
        for (int r = 0; r < ROWS; ++r) acc[r] = fmaf(w, h1[r * H + i], acc[r]);
    }
    const float w3j = __half2float(w3[j]);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const float v = fmaxf(acc[r], 0.0f);
        if (H2out && r0 + r < R) H2out[static_cast<long long>(r0 + r) * H + j] = v;
        float p = v * w3j;
        for (int off = 32; off >= 1; off >>= 1) p += __shfl_xor(p, off);
        if ((j & 63) == 0) red[r][j >> 6] = p;
    }
    __syncthreads();
    if (j < ROWS && r0 + j < R) {
        float s = b3[0];
        for (int w = 0; w < NT / 64; ++w) s += red[j][w];
        y[r0 + j] = s;
    }
}

// Inference on the matrix cores: inputs and the first hidden layer rounded
// to FP16, FP16 weights, FP32 accumulation (v_mfma_f32_32x32x8f16).  One
// workgroup takes 32 rows; wave w computes hidden units [64 w, 64 w + 64) as
// two 32-column tiles.  Operand layout of the instruction (lane l, k < 4):
// A[l % 32][4 (l / 32) + k], B[4 (l / 32) + k][l % 32]; result j < 16 is
// D[8 (j / 4) + 4 (l / 32) + j % 4][l % 32].
typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));
constexpr int kMfmaRows = 32;
constexpr int kMfmaMaxBlocks = 2048;   // 8 per CU
constexpr int kMaxDin = 256;

__global__ __launch_bounds__(NT) void k_mlp_forward_mfma(const float* __restrict__ X, int R,
                                                         const int* __restrict__ Rdev, int din,
                                                         const __half* __restrict__ W1, const float* __restrict__ b1,
                                                         const __half* __restrict__ W2, const float* __restrict__ b2,
                                                         const __half* __restrict__ w3, const float* __restrict__ b3,
                                                         float* __restrict__ y) {
    __shared__ _Float16 xs[kMfmaRows][kMaxDin + 8];
    __shared__ _Float16 h1[kMfmaRows][H + 8];
    __shared__ float red[NT / 64][kMfmaRows];
    const int tid = threadIdx.x;
    if (Rdev) R = min(R, *Rdev);
    const int kp = (din + 7) / 8 * 8;
    // a bounded grid strides over the 32-row tiles: the launch is sized for
    // the most rows a pass can have, the device count is usually far lower
    for (int r0 = blockIdx.x * kMfmaRows; r0 < R; r0 += gridDim.x * kMfmaRows) {   // uniform per workgroup
        for (int i = tid; i < kMfmaRows * kp; i += NT) {
            const int r = i / kp, c = i % kp;
            xs[r][c] = static_cast<_Float16>(r0 + r < R && c < din ? X[static_cast<long long>(r0 + r) * din + c] : 0.0f);
        }
        __syncthreads();
        const int wave = tid >> 6, l = tid & 63;
        const int lr = l % 32, lk = 4 * (l / 32);
        const __half_raw* W1r = reinterpret_cast<const __half_raw*>(W1);
        const __half_raw* W2r = reinterpret_cast<const __half_raw*>(W2);
        auto hw = [](const __half_raw* w, long long i) {
            _Float16 v;
            __builtin_memcpy(&v, &w[i].data, sizeof v);
            return v;
        };
        // four consecutive weights in one 8-byte load (i % 4 == 0 in rows of a
        // multiple of 4 weights)
        auto hw4 = [](const __half_raw* w, long long i) {
            const uint2 raw = *reinterpret_cast<const uint2*>(w + i);
            half4_t v;
            __builtin_memcpy(&v, &raw, sizeof v);
            return v;
        };
        const bool din4 = din % 4 == 0;
        // layer 1
        for (int nt = 0; nt < 2; ++nt) {
            const int col = 64 * wave + 32 * nt + lr;
            float16_t acc;
            for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
            for (int k0 = 0; k0 < kp; k0 += 8) {
                half4_t a, b;
                for (int k = 0; k < 4; ++k) a[k] = xs[lr][k0 + lk + k];
                if (din4) {
                    if (k0 + lk < din) {
                        b = hw4(W1r, static_cast<long long>(col) * din + k0 + lk);
                    } else {
                        for (int k = 0; k < 4; ++k) b[k] = static_cast<_Float16>(0.0f);
                    }
                } else {
                    for (int k = 0; k < 4; ++k) {
                        const int kk = k0 + lk + k;
                        b[k] = kk < din ? hw(W1r, static_cast<long long>(col) * din + kk)
                                        : static_cast<_Float16>(0.0f);
                    }
                }
                acc = __builtin_amdgcn_mfma_f32_32x32x8f16(a, b, acc, 0, 0, 0);
            }
            const float bias = b1[col];
            for (int j = 0; j < 16; ++j) h1[8 * (j / 4) + lk + (j % 4)][col] = static_cast<_Float16>(fmaxf(acc[j] + bias, 0.0f));
        }
        __syncthreads();
        // layer 2 and the output dot product
        float part[16];
        for (int j = 0; j < 16; ++j) part[j] = 0.0f;
        for (int nt = 0; nt < 2; ++nt) {
            const int col = 64 * wave + 32 * nt + lr;
            float16_t acc;
            for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
            for (int k0 = 0; k0 < H; k0 += 8) {
                half4_t a, b;
                for (int k = 0; k < 4; ++k) a[k] = h1[lr][k0 + lk + k];
                b = hw4(W2r, static_cast<long long>(col) * H + k0 + lk);
                acc = __builtin_amdgcn_mfma_f32_32x32x8f16(a, b, acc, 0, 0, 0);
            }
            const float bias = b2[col], wo = __half2float(w3[col]);
            for (int j = 0; j < 16; ++j) part[j] = fmaf(fmaxf(acc[j] + bias, 0.0f), wo, part[j]);
        }
        // sum over the 32 columns held by the lanes of each half-wave
        for (int j = 0; j < 16; ++j) {
            float v = part[j];
            for (int off = 16; off >= 1; off >>= 1) v += __shfl_xor(v, off);
            part[j] = v;
        }
        if (lr == 0)
            for (int j = 0; j < 16; ++j) red[wave][8 * (j / 4) + lk + (j % 4)] = part[j];
        __syncthreads();
        if (tid < kMfmaRows && r0 + tid < R) {
            float v = b3[0];
            for (int w = 0; w < NT / 64; ++w) v += red[w][tid];
            y[r0 + tid] = v;
        }
        __syncthreads();
    }
}

// dy = 2 (y - t) / R and the summed squared error
__global__ void k_mlp_loss(const float* __restrict__ y, const float* __restrict__ t, int R, float* __restrict__ dy,
                           float* __restrict__ loss) {
    __shared__ float red[256];
    float s = 0.0f;
    for (int r = blockIdx.x * 256 + threadIdx.x; r < R; r += gridDim.x * 256) {
        const float d = y[r] - t[r];
        dy[r] = 2.0f * d / static_cast<float>(R);
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// dH2 = dy w3 [H2 > 0];  dH1 = (dH2 W2) [H1 > 0]   (row tiles)
__global__ __launch_bounds__(NT) void k_mlp_backward_data(const float* __restrict__ dy, int R,
                                                          const __half* __restrict__ W2,
                                                          const __half* __restrict__ w3,
                                                          const float* __restrict__ H1, const float* __restrict__ H2,
                                                          float* __restrict__ dH1, float* __restrict__ dH2) {
    __shared__ float g2[ROWS][H];
    const int j = threadIdx.x;
    const int r0 = blockIdx.x * ROWS;
    const float w3j = __half2float(w3[j]);
    for (int r = 0; r < ROWS; ++r) {
        float v = 0.0f;
        if (r0 + r < R) {
            const long long o = static_cast<long long>(r0 + r) * H + j;
            v = H2[o] > 0.0f ? dy[r0 + r] * w3j : 0.0f;
            dH2[o] = v;
        }
        g2[r][j] = v;
    }
    __syncthreads();
    float acc[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) acc[r] = 0.0f;
    for (int k = 0; k < H; ++k) {
        const float w = __half2float(W2[k * H + j]);   // W2[k][j]
#pragma unroll
        for (int r = 0; r < ROWS; ++r) acc[r] = fmaf(g2[r][k], w, acc[r]);
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        if (r0 + r < R) {
            const long long o = static_cast<long long>(r0 + r) * H + j;
            dH1[o] = H1[o] > 0.0f ? acc[r] : 0.0f;
        }
    }
}

// dW[j][i] = sum_r G[r][j] A[r][i], db[j] = sum_r G[r][j]; one workgroup per
// output row j, threads over i (A has width n, G has width H)
__global__ __launch_bounds__(NT) void k_mlp_weight_grad(const float* __restrict__ G, const float* __restrict__ A,
                                                        int R, int n, float* __restrict__ dW,
                                                        float* __restrict__ db) {
    const int j = blockIdx.x;
    __shared__ float gcol[256];
    float acc = 0.0f, bacc = 0.0f;
    for (int r0 = 0; r0 < R; r0 += 256) {
        const int rr = r0 + threadIdx.x;
        gcol[threadIdx.x] = rr < R ? G[static_cast<long long>(rr) * H + j] : 0.0f;
        __syncthreads();
        const int m = min(256, R - r0);
        if (threadIdx.x < n) {
            for (int r = 0; r < m; ++r) acc = fmaf(gcol[r], A[static_cast<long long>(r0 + r) * n + threadIdx.x], acc);
        }
        if (threadIdx.x == 0)
            for (int r = 0; r < m; ++r) bacc += gcol[r];
        __syncthreads();
    }
    if (threadIdx.x < n) dW[j * n + threadIdx.x] = acc;
    if (threadIdx.x == 0 && db) db[j] = bacc;
}

// dw3[j] = sum_r dy[r] H2[r][j]; db3 = sum_r dy[r]
__global__ __launch_bounds__(NT) void k_mlp_out_grad(const float* __restrict__ dy, const float* __restrict__ H2, int R,
                                                     float* __restrict__ dw3, float* __restrict__ db3) {
    const int j = threadIdx.x;
    float acc = 0.0f, b = 0.0f;
    for (int r = 0; r < R; ++r) {
        acc = fmaf(dy[r], H2[static_cast<long long>(r) * H + j], acc);
        b += dy[r];
    }
    dw3[j] = acc;
    if (j == 0) db3[0] = b;
}

// Adam on the FP32 masters, then the FP16 copies
__global__ void k_mlp_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, __half* __restrict__ wh, int n, float lr, float b1, float b2,
                           float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wh) wh[i] = __float2half(wi);
}

template <typename T>
struct Dev {
    T* p = nullptr;
    size_t n = 0;
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    empi_status alloc(size_t count) {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = count;
        if (!count) return EMPI_OK;
        hipError_t e = hipMalloc(&p, sizeof(T) * count);
        if (e != hipSuccess) {
            p = nullptr;
            return e == hipErrorOutOfMemory ? EMPI_NO_MEMORY : EMPI_DEVICE;
        }
        return EMPI_OK;
    }
};

}  // namespace

// Parameter block order: W1 [H][din], b1 [H], W2 [H][H], b2 [H], w3 [H], b3 [1]
struct empi_mlp {
    int din = 0;
    int step = 0;
    size_t nW1 = 0, nW2 = 0, nparams = 0;
    Dev<float> w, g, m, v;     // FP32 masters / gradients / Adam moments, one block
    Dev<__half> wh;            // FP16 copies of W1, W2, w3 (same offsets as in w)
    // minibatch scratch
    Dev<float> H1, H2, dH1, dH2, y, dy, loss;
    size_t rows = 0;
    hipStream_t stream = nullptr;

    size_t oW1() const { return 0; }
    size_t ob1() const { return nW1; }
    size_t oW2() const { return nW1 + H; }
    size_t ob2() const { return nW1 + H + nW2; }
    size_t ow3() const { return nW1 + 2 * H + nW2; }
    size_t ob3() const { return nW1 + 3 * H + nW2; }
};

namespace {

empi_status ensure_rows(empi_mlp* m, size_t R) {
    if (m->rows >= R) return EMPI_OK;
    empi_status st;
    if ((st = m->H1.alloc(R * H)) || (st = m->H2.alloc(R * H)) || (st = m->dH1.alloc(R * H)) ||
        (st = m->dH2.alloc(R * H)) || (st = m->y.alloc(R)) || (st = m->dy.alloc(R)))
        return st;
    m->rows = R;
    return EMPI_OK;
}

empi_status sync_half(empi_mlp* m) {
    std::vector<float> w(m->nparams);
    MLP_HIP(hipMemcpy(w.data(), m->w.p, sizeof(float) * m->nparams, hipMemcpyDeviceToHost));
    std::vector<__half> h(m->nparams);
    for (size_t i = 0; i < m->nparams; ++i) h[i] = __float2half(w[i]);
    MLP_HIP(hipMemcpy(m->wh.p, h.data(), sizeof(__half) * m->nparams, hipMemcpyHostToDevice));
    return EMPI_OK;
}

empi_status forward(empi_mlp* m, const float* X, int R, float* y, bool keep) {
    const int blocks = (R + ROWS - 1) / ROWS;
    const size_t shm = sizeof(float) * ROWS * (m->din + H);
    hipLaunchKernelGGL(k_mlp_forward, dim3(blocks), dim3(NT), shm, m->stream, X, R, nullptr, m->din, m->wh.p + m->oW1(),
                       m->w.p + m->ob1(), m->wh.p + m->oW2(), m->w.p + m->ob2(), m->wh.p + m->ow3(),
                       m->w.p + m->ob3(), y, keep ? m->H1.p : nullptr, keep ? m->H2.p : nullptr);
    MLP_HIP(hipGetLastError());
    return EMPI_OK;
}

// the matrix-core inference pass (FP16 inputs and hidden layer)
empi_status forward16(empi_mlp* m, const float* X, int R, const int* Rdev, float* y, hipStream_t s) {
    if (m->din > kMaxDin) return EMPI_UNSUPPORTED;
    const int blocks = std::min(kMfmaMaxBlocks, (R + kMfmaRows - 1) / kMfmaRows);
    hipLaunchKernelGGL(k_mlp_forward_mfma, dim3(blocks), dim3(NT), 0, s, X, R, Rdev, m->din, m->wh.p + m->oW1(),
                       m->w.p + m->ob1(), m->wh.p + m->oW2(), m->w.p + m->ob2(), m->wh.p + m->ow3(),
                       m->w.p + m->ob3(), y);
    MLP_HIP(hipGetLastError());
    return EMPI_OK;
}

empi_status gradients(empi_mlp* m, const float* X, const float* t, int R, float* loss_out) {
    empi_status st = ensure_rows(m, static_cast<size_t>(R));
    if (st) return st;
    if ((st = forward(m, X, R, m->y.p, true))) return st;
    MLP_HIP(hipMemsetAsync(m->loss.p, 0, sizeof(float), m->stream));
    hipLaunchKernelGGL(k_mlp_loss, dim3(std::min(1024, (R + 255) / 256)), dim3(256), 0, m->stream, m->y.p, t, R,
                       m->dy.p, m->loss.p);
    const int blocks = (R + ROWS - 1) / ROWS;
    hipLaunchKernelGGL(k_mlp_backward_data, dim3(blocks), dim3(NT), 0, m->stream, m->dy.p, R, m->wh.p + m->oW2(),
                       m->wh.p + m->ow3(), m->H1.p, m->H2.p, m->dH1.p, m->dH2.p);
    hipLaunchKernelGGL(k_mlp_out_grad, dim3(1), dim3(NT), 0, m->stream, m->dy.p, m->H2.p, R, m->g.p + m->ow3(),
                       m->g.p + m->ob3());
    hipLaunchKernelGGL(k_mlp_weight_grad, dim3(H), dim3(NT), 0, m->stream, m->dH2.p, m->H1.p, R, H,
                       m->g.p + m->oW2(), m->g.p + m->ob2());
    hipLaunchKernelGGL(k_mlp_weight_grad, dim3(H), dim3(NT), 0, m->stream, m->dH1.p, X, R, m->din,
                       m->g.p + m->oW1(), m->g.p + m->ob1());
    MLP_HIP(hipGetLastError());
    if (loss_out) {
        float l = 0.0f;
        MLP_HIP(hipMemcpyAsync(&l, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, m->stream));
        MLP_HIP(hipStreamSynchronize(m->stream));
        *loss_out = l / static_cast<float>(R);
    }
    return EMPI_OK;
}

}  // namespace

extern "C" {

empi_status empi_mlp_create(int input_dim, unsigned seed, empi_mlp** out) {
    if (!out || input_dim <= 0 || input_dim > 256) return EMPI_INVALID_ARGUMENT;
    *out = nullptr;
    empi_mlp* m = new (std::nothrow) empi_mlp();
    if (!m) return EMPI_NO_MEMORY;
    m->din = input_dim;
    m->nW1 = static_cast<size_t>(H) * input_dim;
    m->nW2 = static_cast<size_t>(H) * H;
    m->nparams = m->nW1 + m->nW2 + 3 * H + 1;
    empi_status st;
    if ((st = m->w.alloc(m->nparams)) || (st = m->g.alloc(m->nparams)) || (st = m->m.alloc(m->nparams)) ||
        (st = m->v.alloc(m->nparams)) || (st = m->wh.alloc(m->nparams)) || (st = m->loss.alloc(1))) {
        delete m;
        return st;
    }
    // He initialisation, zero biases
    std::mt19937 rng(seed);
    std::vector<float> w(m->nparams, 0.0f);
    std::normal_distribution<float> n1(0.0f, std::sqrt(2.0f / input_dim)), n2(0.0f, std::sqrt(2.0f / H));
    for (size_t i = 0; i < m->nW1; ++i) w[m->oW1() + i] = n1(rng);
    for (size_t i = 0; i < m->nW2; ++i) w[m->oW2() + i] = n2(rng);
    for (int i = 0; i < H; ++i) w[m->ow3() + i] = n2(rng) * 0.5f;
    if (hipMemcpy(m->w.p, w.data(), sizeof(float) * m->nparams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(m->m.p, 0, sizeof(float) * m->nparams) != hipSuccess ||
        hipMemset(m->v.p, 0, sizeof(float) * m->nparams) != hipSuccess ||
        hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
        delete m;
        return EMPI_DEVICE;
    }
    if ((st = sync_half(m))) {
        delete m;
        return st;
    }
    *out = m;
    return EMPI_OK;
}

void empi_mlp_destroy(empi_mlp* m) {
    if (!m) return;
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

size_t empi_mlp_param_count(const empi_mlp* m) { return m ? m->nparams : 0; }

empi_status empi_mlp_get_params(const empi_mlp* m, float* out) {
    if (!m || !out) return EMPI_INVALID_ARGUMENT;
    MLP_HIP(hipMemcpy(out, m->w.p, sizeof(float) * m->nparams, hipMemcpyDeviceToHost));
    return EMPI_OK;
}

empi_status empi_mlp_set_params(empi_mlp* m, const float* in) {
    if (!m || !in) return EMPI_INVALID_ARGUMENT;
    MLP_HIP(hipMemcpy(m->w.p, in, sizeof(float) * m->nparams, hipMemcpyHostToDevice));
    return sync_half(m);
}

empi_status empi_mlp_forward(empi_mlp* m, const float* X, int rows, float* y, int on_device) {
    if (!m || !X || !y || rows < 0) return EMPI_INVALID_ARGUMENT;
    if (rows == 0) return EMPI_OK;
    if (on_device) {
        empi_status st = forward(m, X, rows, y, false);
        if (st) return st;
        MLP_HIP(hipStreamSynchronize(m->stream));
        return EMPI_OK;
    }
    Dev<float> dx, dyv;
    empi_status st;
    if ((st = dx.alloc(static_cast<size_t>(rows) * m->din)) || (st = dyv.alloc(rows))) return st;
    MLP_HIP(hipMemcpy(dx.p, X, sizeof(float) * rows * m->din, hipMemcpyHostToDevice));
    if ((st = forward(m, dx.p, rows, dyv.p, false))) return st;
    MLP_HIP(hipStreamSynchronize(m->stream));
    MLP_HIP(hipMemcpy(y, dyv.p, sizeof(float) * rows, hipMemcpyDeviceToHost));
    return EMPI_OK;
}

empi_status empi_mlp_forward16(empi_mlp* m, const float* X, int rows, float* y, int on_device) {
    if (!m || !X || !y || rows < 0) return EMPI_INVALID_ARGUMENT;
    if (rows == 0) return EMPI_OK;
    empi_status st;
    if (on_device) {
        if ((st = forward16(m, X, rows, nullptr, y, m->stream))) return st;
        MLP_HIP(hipStreamSynchronize(m->stream));
        return EMPI_OK;
    }
    Dev<float> dx, dyv;
    if ((st = dx.alloc(static_cast<size_t>(rows) * m->din)) || (st = dyv.alloc(rows))) return st;
    MLP_HIP(hipMemcpy(dx.p, X, sizeof(float) * rows * m->din, hipMemcpyHostToDevice));
    if ((st = forward16(m, dx.p, rows, nullptr, dyv.p, m->stream))) return st;
    MLP_HIP(hipStreamSynchronize(m->stream));
    MLP_HIP(hipMemcpy(y, dyv.p, sizeof(float) * rows, hipMemcpyDeviceToHost));
    return EMPI_OK;
}

empi_status empi_mlp_gradients(empi_mlp* m, const float* X, const float* target, int rows, float* grads,
                               float* loss) {
    if (!m || !X || !target || rows <= 0) return EMPI_INVALID_ARGUMENT;
    Dev<float> dx, dt;
    empi_status st;
    if ((st = dx.alloc(static_cast<size_t>(rows) * m->din)) || (st = dt.alloc(rows))) return st;
    MLP_HIP(hipMemcpy(dx.p, X, sizeof(float) * rows * m->din, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(dt.p, target, sizeof(float) * rows, hipMemcpyHostToDevice));
    if ((st = gradients(m, dx.p, dt.p, rows, loss))) return st;
    MLP_HIP(hipStreamSynchronize(m->stream));
    if (grads) MLP_HIP(hipMemcpy(grads, m->g.p, sizeof(float) * m->nparams, hipMemcpyDeviceToHost));
    return EMPI_OK;
}

empi_status empi_mlp_train(empi_mlp* m, const float* X, const float* target, int rows, int on_device, int epochs,
                           int batch, float lr, unsigned seed, float* final_loss) {
    if (!m || !X || !target || rows <= 0 || epochs <= 0 || batch <= 0) return EMPI_INVALID_ARGUMENT;
    Dev<float> dx, dt, bx, bt;
    const float* Xd = X;
    const float* Td = target;
    empi_status st;
    if (!on_device) {
        if ((st = dx.alloc(static_cast<size_t>(rows) * m->din)) || (st = dt.alloc(rows))) return st;
        MLP_HIP(hipMemcpy(dx.p, X, sizeof(float) * rows * m->din, hipMemcpyHostToDevice));
        MLP_HIP(hipMemcpy(dt.p, target, sizeof(float) * rows, hipMemcpyHostToDevice));
        Xd = dx.p;
        Td = dt.p;
    }
    // shuffled minibatches: a host permutation of contiguous row blocks of
    // the batch size (rows are assumed pre-shuffled when generated)
    const int nb = (rows + batch - 1) / batch;
    std::mt19937 rng(seed);
    std::vector<int> order(nb);
    float last = 0.0f;
    for (int e = 0; e < epochs; ++e) {
        for (int i = 0; i < nb; ++i) order[i] = i;
        std::shuffle(order.begin(), order.end(), rng);
        double sum = 0.0;
        for (int bi = 0; bi < nb; ++bi) {
            const int r0 = order[bi] * batch;
            const int R = std::min(batch, rows - r0);
            float l = 0.0f;
            if ((st = gradients(m, Xd + static_cast<size_t>(r0) * m->din, Td + r0, R, &l))) return st;
            sum += l;
            ++m->step;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.0f - std::pow(b1, static_cast<float>(m->step));
            const float c2 = 1.0f - std::pow(b2, static_cast<float>(m->step));
            const int n = static_cast<int>(m->nparams);
            hipLaunchKernelGGL(k_mlp_adam, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->w.p, m->g.p, m->m.p,
                               m->v.p, m->wh.p, n, lr, b1, b2, 1e-8f, c1, c2);
            MLP_HIP(hipGetLastError());
        }
        last = static_cast<float>(sum / nb);
    }
    MLP_HIP(hipStreamSynchronize(m->stream));
    if (final_loss) *final_loss = last;
    return EMPI_OK;
}

}  // extern "C"

empi_status empi_mlp_forward_counted(empi_mlp* m, const float* X, const int* rows_dev, int max_rows, float* y,
                                     hipStream_t s) {
    if (!m || !X || !rows_dev || !y || max_rows < 0) return EMPI_INVALID_ARGUMENT;
    if (max_rows == 0) return EMPI_OK;
    return forward16(m, X, max_rows, rows_dev, y, s);
}

int empi_mlp_input_dim(const empi_mlp* m) { return m ? m->din : 0; }
