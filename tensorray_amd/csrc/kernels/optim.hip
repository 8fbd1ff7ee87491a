// Optimizer step on device parameters: Adam (Kingma & Ba 2015) with the
// update order of the reference's C++ examples (Optimizer.h:66-123):
//   m = b1*m + (1-b1)*g;  v = b2*v + (1-b2)*g*g;
//   x -= lr * (m / bc1) / (sqrt(v / bc2) + eps),  bc_i = 1 - b_i^t.
// Non-finite gradient entries count as zero (the reference scrubs NaN
// gradients before its optimizer, renderer_single_view.py:82).
#include <hip/hip_runtime.h>

#include "../trc_internal.h"

namespace trc {

__global__ void k_adam(float* x, const float* g, float* m, float* v, uint64_t n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const float c1 = 1.0f - b1, c2 = 1.0f - b2;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    float gi = g[i];
    if (!(gi - gi == 0.0f)) gi = 0.0f;  // NaN or infinity
    const float mi = b1 * m[i] + c1 * gi;
    const float vi = b2 * v[i] + c2 * (gi * gi);
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / bc1;
    const float vh = vi / bc2;
    x[i] = x[i] - (lr * mh) / (sqrtf(vh) + eps);
  }
}

hipError_t launch_adam(float* x, const float* g, float* m, float* v, uint64_t n, float lr, float b1, float b2,
                       float eps, float bc1, float bc2, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const uint64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_adam, dim3((unsigned)(blocks < 8192 ? blocks : 8192)), dim3(256), 0, st, x, g, m, v, n, lr,
                     b1, b2, eps, bc1, bc2);
  return hipGetLastError();
}

}  // namespace trc
