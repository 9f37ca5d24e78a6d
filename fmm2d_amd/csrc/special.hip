// Batched evaluation of H0^(1), H1^(1) on the device (the C-ABI's
// fmm2d_hank103; reference common/hank103.f:116).
#include "../../include/fmm2d_amd.h"
#include "hankel.h"

using namespace fmm2d;

__global__ void hank103_kernel(const cplx* __restrict__ z, int64_t n, cplx* __restrict__ h0,
                               cplx* __restrict__ h1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cplx a, b;
  hank01(z[i], &a, &b);
  h0[i] = a;
  h1[i] = b;
}

extern "C" int fmm2d_hank103(const fmm2d_cplx* z, int64_t n, fmm2d_cplx* h0, fmm2d_cplx* h1,
                             void* stream) {
  if (n < 0 || (n > 0 && (!z || !h0 || !h1))) return FMM2D_ERR_ARG;
  if (n == 0) return FMM2D_OK;
  hipLaunchKernelGGL(hank103_kernel, dim3(ceil_div(n, 256)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), reinterpret_cast<const cplx*>(z), n,
                     reinterpret_cast<cplx*>(h0), reinterpret_cast<cplx*>(h1));
  FMM2D_HIP_TRY(hipGetLastError());
  return FMM2D_OK;
}
