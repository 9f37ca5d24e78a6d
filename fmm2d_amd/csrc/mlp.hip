// Parameter-proposer network: a 3-layer perceptron nin -> 64 -> 64 -> nout
// (tanh, tanh, linear), FP32, with its forward pass, the gradient of a
// weighted squared-error loss and an Adam step as device kernels.
//
// No reference counterpart: the reference chooses nterms and the leaf size by
// fixed rules (laplace/l2dterms.f, helmholtz/h2dterms.f, laplace/lndiv2d.f);
// the network proposes overrides that the drivers accept through
// fmm2d_opts.nterms / ndiv.
//
// Flat parameter layout (floats): W1[64][nin], b1[64], W2[64][64], b2[64],
// W3[nout][64], b3[nout].  Activations are row-major [batch][width].
// The network is small (about 5e3 + 64 nout parameters) and its batches are
// a few thousand rows: each layer is one launch with a thread per output
// element, which keeps the gradients deterministic (no atomics).
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/fmm2d_amd.h"
#include "devbuf.h"

namespace fmm2d {
namespace {

constexpr int kH = 64;

// y[b][o] = act(sum_i x[b][i] W[o][i] + bias[o]); act 1 = tanh, 0 = identity
__global__ void k_dense(int batch, int nin, int nout, const float* x, const float* W,
                        const float* bias, int act, float* y) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)batch * nout) return;
  const int b = (int)(e / nout), o = (int)(e % nout);
  const float* xr = x + (int64_t)b * nin;
  const float* w = W + (int64_t)o * nin;
  float s = bias[o];
  for (int i = 0; i < nin; ++i) s = fmaf(xr[i], w[i], s);
  y[e] = act ? tanhf(s) : s;
}

// loss = sum_b,o wt[b][o] (y - t)^2 / batch; dy = 2 wt (y - t) / batch
__global__ void k_sqerr(int batch, int nout, const float* y, const float* t, const float* wt,
                        float* dy, float* lossparts) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)batch * nout) return;
  const float r = y[e] - t[e];
  const float w = wt ? wt[e] : 1.0f;
  dy[e] = 2.0f * w * r / batch;
  lossparts[e] = w * r * r / batch;
}

// sum of n values by one block (deterministic order)
__global__ void k_sum(int64_t n, const float* v, float* out) {
  __shared__ float part[256];
  float s = 0.0f;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += v[i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) part[threadIdx.x] += part[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = part[0];
}

// dx[b][i] = (sum_o dy[b][o] W[o][i]) * (1 - x[b][i]^2)  (x = tanh output)
__global__ void k_back_input(int batch, int nin, int nout, const float* dy, const float* W,
                             const float* x, float* dx) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)batch * nin) return;
  const int b = (int)(e / nin), i = (int)(e % nin);
  const float* d = dy + (int64_t)b * nout;
  float s = 0.0f;
  for (int o = 0; o < nout; ++o) s = fmaf(d[o], W[(int64_t)o * nin + i], s);
  const float a = x[e];
  dx[e] = s * (1.0f - a * a);
}

// dW[o][i] = sum_b dy[b][o] x[b][i]; db[o] = sum_b dy[b][o]
__global__ void k_back_params(int batch, int nin, int nout, const float* dy, const float* x,
                              float* dW, float* db) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nout * (nin + 1)) return;
  const int o = (int)(e / (nin + 1)), i = (int)(e % (nin + 1));
  float s = 0.0f;
  if (i < nin)
    for (int b = 0; b < batch; ++b) s = fmaf(dy[(int64_t)b * nout + o], x[(int64_t)b * nin + i], s);
  else
    for (int b = 0; b < batch; ++b) s += dy[(int64_t)b * nout + o];
  if (i < nin) dW[(int64_t)o * nin + i] = s;
  else db[o] = s;
}

__global__ void k_adam(int64_t n, float* p, const float* g, float* m, float* v, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

struct Layout {
  int64_t W1, b1, W2, b2, W3, b3, total;
  Layout(int nin, int nout) {
    W1 = 0;
    b1 = W1 + (int64_t)kH * nin;
    W2 = b1 + kH;
    b2 = W2 + (int64_t)kH * kH;
    W3 = b2 + kH;
    b3 = W3 + (int64_t)nout * kH;
    total = b3 + nout;
  }
};

unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace fmm2d

using namespace fmm2d;

extern "C" {

int64_t fmm2d_mlp_nparams(int nin, int nout) {
  if (nin < 1 || nout < 1) return -1;
  return Layout(nin, nout).total;
}

int fmm2d_mlp_forward(int nin, int nout, const float* params, int batch, const float* x,
                      float* y, void* stream) {
  if (nin < 1 || nout < 1 || batch < 0 || !params || (batch > 0 && (!x || !y)))
    return FMM2D_ERR_ARG;
  if (batch == 0) return FMM2D_OK;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const Layout L(nin, nout);
  try {
    DBuf<float> h1((size_t)batch * kH, st), h2((size_t)batch * kH, st);
    hipLaunchKernelGGL(k_dense, dim3(blocks((int64_t)batch * kH)), dim3(256), 0, st, batch, nin,
                       kH, x, params + L.W1, params + L.b1, 1, h1.p);
    hipLaunchKernelGGL(k_dense, dim3(blocks((int64_t)batch * kH)), dim3(256), 0, st, batch, kH,
                       kH, h1.p, params + L.W2, params + L.b2, 1, h2.p);
    hipLaunchKernelGGL(k_dense, dim3(blocks((int64_t)batch * nout)), dim3(256), 0, st, batch, kH,
                       nout, h2.p, params + L.W3, params + L.b3, 0, y);
    hip_throw(hipGetLastError(), "mlp forward");
  } catch (const HipError& e) {
    return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return FMM2D_ERR_ALLOC;
  } catch (...) {
    return FMM2D_ERR_HIP;
  }
  return FMM2D_OK;
}

int fmm2d_mlp_loss_grad(int nin, int nout, const float* params, int batch, const float* x,
                        const float* target, const float* weight, float* grad, float* loss,
                        void* stream) {
  if (nin < 1 || nout < 1 || batch < 1 || !params || !x || !target || !grad || !loss)
    return FMM2D_ERR_ARG;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const Layout L(nin, nout);
  try {
    const size_t bh = (size_t)batch * kH, bo = (size_t)batch * nout;
    DBuf<float> h1(bh, st), h2(bh, st), y(bo, st), dy(bo, st), lp(bo, st), d2(bh, st), d1(bh, st);
    hipLaunchKernelGGL(k_dense, dim3(blocks(bh)), dim3(256), 0, st, batch, nin, kH, x,
                       params + L.W1, params + L.b1, 1, h1.p);
    hipLaunchKernelGGL(k_dense, dim3(blocks(bh)), dim3(256), 0, st, batch, kH, kH, h1.p,
                       params + L.W2, params + L.b2, 1, h2.p);
    hipLaunchKernelGGL(k_dense, dim3(blocks(bo)), dim3(256), 0, st, batch, kH, nout, h2.p,
                       params + L.W3, params + L.b3, 0, y.p);
    hipLaunchKernelGGL(k_sqerr, dim3(blocks(bo)), dim3(256), 0, st, batch, nout, y.p, target,
                       weight, dy.p, lp.p);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, st, (int64_t)bo, lp.p, loss);
    // layer 3
    hipLaunchKernelGGL(k_back_params, dim3(blocks((int64_t)nout * (kH + 1))), dim3(256), 0, st,
                       batch, kH, nout, dy.p, h2.p, grad + L.W3, grad + L.b3);
    hipLaunchKernelGGL(k_back_input, dim3(blocks(bh)), dim3(256), 0, st, batch, kH, nout, dy.p,
                       params + L.W3, h2.p, d2.p);
    // layer 2
    hipLaunchKernelGGL(k_back_params, dim3(blocks((int64_t)kH * (kH + 1))), dim3(256), 0, st,
                       batch, kH, kH, d2.p, h1.p, grad + L.W2, grad + L.b2);
    hipLaunchKernelGGL(k_back_input, dim3(blocks(bh)), dim3(256), 0, st, batch, kH, kH, d2.p,
                       params + L.W2, h1.p, d1.p);
    // layer 1
    hipLaunchKernelGGL(k_back_params, dim3(blocks((int64_t)kH * (nin + 1))), dim3(256), 0, st,
                       batch, nin, kH, d1.p, x, grad + L.W1, grad + L.b1);
    hip_throw(hipGetLastError(), "mlp loss/grad");
  } catch (const HipError& e) {
    return e.err == hipErrorOutOfMemory ? FMM2D_ERR_ALLOC : FMM2D_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return FMM2D_ERR_ALLOC;
  } catch (...) {
    return FMM2D_ERR_HIP;
  }
  return FMM2D_OK;
}

int fmm2d_mlp_adam(int64_t n, float* params, const float* grad, float* m, float* v, float lr,
                   float beta1, float beta2, float eps, int step, void* stream) {
  if (n < 1 || !params || !grad || !m || !v || step < 1) return FMM2D_ERR_ARG;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const float c1 = 1.0f - std::pow(beta1, (float)step), c2 = 1.0f - std::pow(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, st, n, params, grad, m, v, lr, beta1,
                     beta2, eps, c1, c2);
  return hipGetLastError() == hipSuccess ? FMM2D_OK : FMM2D_ERR_HIP;
}

}  // extern "C"
