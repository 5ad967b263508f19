// The learned screen's network: ni inputs -> 3 hidden layers of 64 ReLU units -> one linear
// output, FP32.  Inference, training gradients and the Adam update are hand-written kernels.
//
// Layout of the parameter vector (mlp_offsets): W1 (64 x ni, row-major), b1 (64), W2 (64 x 64),
// b2 (64), W3 (64 x 64), b3 (64), w4 (64), b4 (1).
//
// Kernel shape: one 64-lane wave per workgroup, lane j = hidden unit j, a tile of kMlpTile samples
// per workgroup.  A lane holds its row of the layer's weight matrix in registers and walks the
// tile's samples; the activations of the tile live in LDS (one row of 64 per sample, padded to 65
// words so that the per-sample reads of the output layer, lane = sample, fall in different
// banks).  The inputs of a layer are read at wave-uniform LDS addresses (broadcast).
#include "internal.hpp"
#include "kernels.hpp"

namespace lsjk {

namespace {

constexpr int H = kMlpHidden;
constexpr int S = kMlpTile;
constexpr int R = H + 1;  // padded activation row

struct Off {
    int W1, b1, W2, b2, W3, b3, w4, b4, n;
};

__host__ __device__ inline Off mlp_offsets(int ni)
{
    Off o;
    o.W1 = 0;
    o.b1 = H * ni;
    o.W2 = o.b1 + H;
    o.b2 = o.W2 + H * H;
    o.W3 = o.b2 + H;
    o.b3 = o.W3 + H * H;
    o.w4 = o.b3 + H;
    o.b4 = o.w4 + H;
    o.n = o.b4 + 1;
    return o;
}

// Dense layer over the tile: out[s][j] = relu(b[j] + sum_k W[j][k] in[s][k]) for lane j.
template <int K>
__device__ __forceinline__ void dense_relu(const float* __restrict__ W, const float* __restrict__ b, const float* in,
                                           int ldin, float* out, int nk)
{
    const int j = threadIdx.x;
    float w[K];
#pragma unroll
    for (int k = 0; k < K; k++) w[k] = k < nk ? W[j * nk + k] : 0.0f;
    const float bj = b[j];
    for (int s = 0; s < S; s++) {
        float a = bj;
#pragma unroll
        for (int k = 0; k < K; k++) a = fmaf(w[k], in[s * ldin + k], a);
        out[s * R + j] = fmaxf(a, 0.0f);
    }
}

// Forward pass of a tile into LDS (x, h1, h2, h3); returns nothing, h3 holds the last hidden layer.
__device__ __forceinline__ void forward_tile(const float* __restrict__ p, const Off& o, int ni, const float* x,
                                             float* h1, float* h2, float* h3)
{
    dense_relu<kMlpMaxIn>(p + o.W1, p + o.b1, x, kMlpMaxIn, h1, ni);
    __syncthreads();
    dense_relu<H>(p + o.W2, p + o.b2, h1, R, h2, H);
    __syncthreads();
    dense_relu<H>(p + o.W3, p + o.b3, h2, R, h3, H);
    __syncthreads();
}

// Stage a tile of inputs: rows idx[s] (or s when idx is null) of X (n x ni), zero past n.
__device__ __forceinline__ void load_tile(const float* __restrict__ X, const int* __restrict__ idx, int n, int ni,
                                          int t0, float* x)
{
    for (int q = threadIdx.x; q < S * kMlpMaxIn; q += H) {
        const int s = q / kMlpMaxIn, k = q - s * kMlpMaxIn;
        const int t = t0 + s;
        float v = 0.0f;
        if (t < n && k < ni) v = X[(size_t)(idx ? idx[t] : t) * ni + k];
        x[q] = v;
    }
}

__global__ __launch_bounds__(64) void mlp_forward_kernel(const float* __restrict__ p, int ni, int n,
                                                         const float* __restrict__ X, float* __restrict__ out)
{
    __shared__ float x[S * kMlpMaxIn];
    __shared__ float h1[S * R], h2[S * R], h3[S * R];
    const Off o = mlp_offsets(ni);
    const int t0 = blockIdx.x * S;
    load_tile(X, nullptr, n, ni, t0, x);
    __syncthreads();
    forward_tile(p, o, ni, x, h1, h2, h3);
    const int s = threadIdx.x;  // lane = sample for the output layer
    if (t0 + s < n) {
        float f = p[o.b4];
        for (int j = 0; j < H; j++) f = fmaf(p[o.w4 + j], h3[s * R + j], f);
        out[t0 + s] = f;
    }
}

// Gradient of L = (1/nb) sum_s sw_s r_s^2 (wu if r_s < 0), r_s = f(x_s) - y_s, over the nb samples
// idx[0..nb) of (X, y); accumulated into grad (atomics), L into loss[0].
__global__ __launch_bounds__(64) void mlp_grad_kernel(const float* __restrict__ p, int ni, int nb,
                                                      const int* __restrict__ idx, const float* __restrict__ X,
                                                      const float* __restrict__ y, const float* __restrict__ sw,
                                                      float wu, float* __restrict__ grad, float* __restrict__ loss)
{
    __shared__ float x[S * kMlpMaxIn];
    __shared__ float h1[S * R], h2[S * R], h3[S * R];
    __shared__ float g[S];
    const Off o = mlp_offsets(ni);
    const int t0 = blockIdx.x * S;
    const int lane = threadIdx.x;
    load_tile(X, idx, nb, ni, t0, x);
    __syncthreads();
    forward_tile(p, o, ni, x, h1, h2, h3);
    // output, residual and dL/df per sample (lane = sample)
    {
        const int s = lane;
        float gs = 0.0f, ls = 0.0f;
        if (t0 + s < nb) {
            const int row = idx ? idx[t0 + s] : t0 + s;
            float f = p[o.b4];
            for (int j = 0; j < H; j++) f = fmaf(p[o.w4 + j], h3[s * R + j], f);
            const float r = f - y[row];
            const float wgt = (sw ? sw[row] : 1.0f) * (r < 0.0f ? wu : 1.0f) / (float)nb;
            ls = wgt * r * r;
            gs = 2.0f * wgt * r;
        }
        g[s] = gs;
        // wave sum of the loss terms
        for (int off = 32; off > 0; off >>= 1) ls += __shfl_xor(ls, off);
        if (lane == 0) atomicAdd(loss, ls);
    }
    __syncthreads();
    const int j = lane;
    // output layer: dw4[j], db4; delta3[s][j] = g[s] w4[j] [h3 > 0] (over h3)
    {
        const float w4j = p[o.w4 + j];
        float dw = 0.0f, db = 0.0f;
        for (int s = 0; s < S; s++) {
            const float h = h3[s * R + j];
            dw = fmaf(g[s], h, dw);
            db += g[s];
            h3[s * R + j] = h > 0.0f ? g[s] * w4j : 0.0f;
        }
        atomicAdd(grad + o.w4 + j, dw);
        if (j == 0) atomicAdd(grad + o.b4, db);
    }
    __syncthreads();
    // hidden layers 3 and 2: dW[j][k] = sum_s delta[s][j] in[s][k]; db[j] = sum_s delta[s][j];
    // delta_prev[s][k] = (sum_j W[j][k] delta[s][j]) [in[s][k] > 0], written over in[] (lane k)
    auto backward_hidden = [&](int oW, int ob, float* delta, float* in) {
        float acc[H];
#pragma unroll
        for (int k = 0; k < H; k++) acc[k] = 0.0f;
        float db = 0.0f;
        for (int s = 0; s < S; s++) {
            const float d = delta[s * R + j];
            db += d;
#pragma unroll
            for (int k = 0; k < H; k++) acc[k] = fmaf(d, in[s * R + k], acc[k]);
        }
#pragma unroll
        for (int k = 0; k < H; k++) atomicAdd(grad + oW + j * H + k, acc[k]);
        atomicAdd(grad + ob + j, db);
        // column k = lane of W
        const int k = lane;
        float wc[H];
#pragma unroll
        for (int q = 0; q < H; q++) wc[q] = p[oW + q * H + k];
        __syncthreads();
        for (int s = 0; s < S; s++) {
            float a = 0.0f;
#pragma unroll
            for (int q = 0; q < H; q++) a = fmaf(wc[q], delta[s * R + q], a);
            const float h = in[s * R + k];
            in[s * R + k] = h > 0.0f ? a : 0.0f;
        }
        __syncthreads();
    };
    backward_hidden(o.W3, o.b3, h3, h2);  // h2 now holds delta2
    backward_hidden(o.W2, o.b2, h2, h1);  // h1 now holds delta1
    // first layer: dW1[j][k] = sum_s delta1[s][j] x[s][k], db1[j]
    {
        float acc[kMlpMaxIn];
#pragma unroll
        for (int k = 0; k < kMlpMaxIn; k++) acc[k] = 0.0f;
        float db = 0.0f;
        for (int s = 0; s < S; s++) {
            const float d = h1[s * R + j];
            db += d;
#pragma unroll
            for (int k = 0; k < kMlpMaxIn; k++) acc[k] = fmaf(d, x[s * kMlpMaxIn + k], acc[k]);
        }
        for (int k = 0; k < ni; k++) atomicAdd(grad + o.W1 + j * ni + k, acc[k]);
        atomicAdd(grad + o.b1 + j, db);
    }
}

// Adam (Kingma & Ba 2015) with bias correction; step t >= 1.
__global__ void mlp_adam_kernel(int np, float* __restrict__ p, const float* __restrict__ grad, float* __restrict__ m,
                                float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float gi = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace

int mlp_nparam(int ni) { return mlp_offsets(ni).n; }

void launch_mlp_forward(const float* p, int ni, int n, const float* X, float* out, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + S - 1) / S)), dim3(H), 0, s, p, ni, n, X, out);
}

void launch_mlp_grad(const float* p, int ni, int nb, const int* idx, const float* X, const float* y, const float* sw,
                     float wu, float* grad, float* loss, hipStream_t s)
{
    if (nb <= 0) return;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3((unsigned)((nb + S - 1) / S)), dim3(H), 0, s, p, ni, nb, idx, X, y, sw,
                       wu, grad, loss);
}

void launch_mlp_adam(int np, float* p, const float* grad, float* m, float* v, float lr, float b1, float b2, float eps,
                     int t, hipStream_t s)
{
    const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, np, p, grad, m, v, lr, b1,
                       b2, eps, c1, c2);
}

}  // namespace lsjk
