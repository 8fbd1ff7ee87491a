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

namespace trc {

// ---------------------------------------------------------------------------
// Uniform Adam (the step of the reference's LargeSteps optimizer,
// pyTensorRay/optimizer.py:40-63, AdamUniform of Nicolet et al. 2021): Adam
// moments as above, but every coordinate is scaled by the one largest
// second moment, x -= lr / (bc1 * sqrt(max(v) / bc2)) * m.
// ---------------------------------------------------------------------------
__global__ void k_adam_uniform_moments(const float* g, float* m, float* v, uint64_t n, float b1, float b2,
                                       unsigned int* vmax_bits) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const float c1 = 1.0f - b1, c2 = 1.0f - b2;
  float local = 0.0f;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    float gi = g[i];
    if (!(gi - gi == 0.0f)) gi = 0.0f;
    m[i] = b1 * m[i] + c1 * gi;
    const float vi = b2 * v[i] + c2 * (gi * gi);
    v[i] = vi;
    local = fmaxf(local, vi);
  }
  // v >= 0, so the order of the float bit patterns is the order of the
  // values; one atomic per block (atomics on one word from every wave of a
  // large grid serialise)
  __shared__ float wmax[4];
  for (int o = 32; o > 0; o >>= 1) local = fmaxf(local, __shfl_xor(local, o));
  if ((threadIdx.x & 63u) == 0u) wmax[threadIdx.x >> 6] = local;
  __syncthreads();
  if (threadIdx.x == 0u) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) local = fmaxf(local, wmax[w]);
    atomicMax(vmax_bits, __float_as_uint(local));
  }
}

__global__ void k_adam_uniform_apply(float* x, const float* m, uint64_t n, float lr, float bc1, float bc2,
                                     const unsigned int* vmax_bits) {
  const float vmax = __uint_as_float(*vmax_bits);
  if (!(vmax > 0.0f)) return;
  const float scale = lr / (bc1 * sqrtf(vmax / bc2));
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = x[i] - scale * m[i];
}

__global__ void k_zero_u32(unsigned int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

hipError_t launch_adam_uniform(float* x, const float* g, float* m, float* v, uint64_t n, float lr, float b1, float b2,
                               float bc1, float bc2, unsigned int* vmax_bits, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, vmax_bits, 1);
  if (n == 0) return hipGetLastError();
  const uint64_t blocks = (n + 255) / 256;
  const dim3 grid((unsigned)(blocks < 8192 ? blocks : 8192));
  const dim3 grid_m((unsigned)(blocks < 1024 ? blocks : 1024));  // 256 threads = 4 waves per block
  hipLaunchKernelGGL(k_adam_uniform_moments, grid_m, dim3(256), 0, st, g, m, v, n, b1, b2, vmax_bits);
  hipLaunchKernelGGL(k_adam_uniform_apply, grid, dim3(256), 0, st, x, m, n, lr, bc1, bc2, vmax_bits);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Conjugate gradients for M x = b with M = I + lambda L symmetric positive
// definite (CSR), three right-hand sides at once in the [V][3] layout of the
// vertex buffers.  Scalars live on the device (CgScalars); every kernel
// returns at once when all three columns have converged, so the host can
// queue max_iter iterations without reading anything back.  A dot product
// is a partial per block (no atomics: the partials are summed in block order
// by every kernel that needs the total, so all of them see the same bits).
// ---------------------------------------------------------------------------
constexpr int kCgThreads = 256;

// The block's sum of (a, b, c) into part[blockIdx.x], in a fixed order.
__device__ inline void block_partial3(float a, float b, float c, float (*part)[3]) {
  __shared__ float w[kCgThreads / 64][3];
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
    c += __shfl_xor(c, o);
  }
  if ((threadIdx.x & 63u) == 0u) {
    w[threadIdx.x >> 6][0] = a;
    w[threadIdx.x >> 6][1] = b;
    w[threadIdx.x >> 6][2] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
    for (int k = 0; k < kCgThreads / 64; ++k) {
      s0 += w[k][0];
      s1 += w[k][1];
      s2 += w[k][2];
    }
    part[blockIdx.x][0] = s0;
    part[blockIdx.x][1] = s1;
    part[blockIdx.x][2] = s2;
  }
}

// Sum of part[0 .. nb) per column, the same bits in every block: the first
// wave sums a fixed subset per lane, then reduces across lanes.
__device__ inline void sum_partials3(const float (*part)[3], int nb, float out[3]) {
  __shared__ float tot[3];
  if (threadIdx.x < 64u) {
    float a = 0.0f, b = 0.0f, c = 0.0f;
    for (int k = (int)threadIdx.x; k < nb; k += 64) {
      a += part[k][0];
      b += part[k][1];
      c += part[k][2];
    }
    for (int o = 32; o > 0; o >>= 1) {
      a += __shfl_xor(a, o);
      b += __shfl_xor(b, o);
      c += __shfl_xor(c, o);
    }
    if (threadIdx.x == 0u) {
      tot[0] = a;
      tot[1] = b;
      tot[2] = c;
    }
  }
  __syncthreads();
  out[0] = tot[0];
  out[1] = tot[1];
  out[2] = tot[2];
}

__device__ inline bool cg_done(const CgScalars* s) { return s->active == 0u; }

// y = M x (and, with with_dot, the partials of x_c . y_c)
__global__ void __launch_bounds__(kCgThreads) k_csr_spmv3(CsrDev M, const float* x, float* y, CgScalars* s,
                                                          int use_flag, int with_dot) {
  if (use_flag && cg_done(s)) return;
  float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f;
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M.n; i += stride) {
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
    for (int k = M.row[i]; k < M.row[i + 1]; ++k) {
      const float w = M.val[k];
      const float* xj = x + 3 * (size_t)M.col[k];
      a0 = a0 + w * xj[0];
      a1 = a1 + w * xj[1];
      a2 = a2 + w * xj[2];
    }
    y[3 * (size_t)i] = a0;
    y[3 * (size_t)i + 1] = a1;
    y[3 * (size_t)i + 2] = a2;
    if (with_dot) {
      d0 = d0 + x[3 * (size_t)i] * a0;
      d1 = d1 + x[3 * (size_t)i + 1] * a1;
      d2 = d2 + x[3 * (size_t)i + 2] * a2;
    }
  }
  if (with_dot) block_partial3(d0, d1, d2, s->part_pap);
}

// r = b - Mx (Mx in r on entry), p = r; partials of r.r and b.b
__global__ void __launch_bounds__(kCgThreads) k_cg_init(const float* b, float* r, float* p, int n, CgScalars* s) {
  float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f, e0 = 0.0f, e1 = 0.0f, e2 = 0.0f;
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    for (int c = 0; c < 3; ++c) {
      const size_t j = 3 * (size_t)i + c;
      const float ri = b[j] - r[j];
      r[j] = ri;
      p[j] = ri;
    }
    d0 = d0 + r[3 * (size_t)i] * r[3 * (size_t)i];
    d1 = d1 + r[3 * (size_t)i + 1] * r[3 * (size_t)i + 1];
    d2 = d2 + r[3 * (size_t)i + 2] * r[3 * (size_t)i + 2];
    e0 = e0 + b[3 * (size_t)i] * b[3 * (size_t)i];
    e1 = e1 + b[3 * (size_t)i + 1] * b[3 * (size_t)i + 1];
    e2 = e2 + b[3 * (size_t)i + 2] * b[3 * (size_t)i + 2];
  }
  block_partial3(d0, d1, d2, s->part_rr);
  __syncthreads();  // block_partial3's LDS is reused
  block_partial3(e0, e1, e2, s->part_b);
}

// x += alpha p, r -= alpha Ap; partials of the new r.r   (alpha = rr / pAp)
__global__ void __launch_bounds__(kCgThreads) k_cg_update(float* x, float* r, const float* p, const float* ap, int n,
                                                          CgScalars* s) {
  if (cg_done(s)) return;
  float pap[3];
  sum_partials3(s->part_pap, (int)gridDim.x, pap);
  float alpha[3];
  for (int c = 0; c < 3; ++c) alpha[c] = ((s->active >> c) & 1u) && pap[c] > 0.0f ? s->rr[c] / pap[c] : 0.0f;
  float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f;
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    for (int c = 0; c < 3; ++c) {
      const size_t j = 3 * (size_t)i + c;
      x[j] = x[j] + alpha[c] * p[j];
      r[j] = r[j] - alpha[c] * ap[j];
    }
    d0 = d0 + r[3 * (size_t)i] * r[3 * (size_t)i];
    d1 = d1 + r[3 * (size_t)i + 1] * r[3 * (size_t)i + 1];
    d2 = d2 + r[3 * (size_t)i + 2] * r[3 * (size_t)i + 2];
  }
  block_partial3(d0, d1, d2, s->part_rr);
}

// p = r + beta p  (beta = rr_new / rr per column)
__global__ void __launch_bounds__(kCgThreads) k_cg_direction(const float* r, float* p, int n, CgScalars* s) {
  if (cg_done(s)) return;
  float rr_new[3];
  sum_partials3(s->part_rr, (int)gridDim.x, rr_new);
  float beta[3];
  for (int c = 0; c < 3; ++c) beta[c] = s->rr[c] > 0.0f ? rr_new[c] / s->rr[c] : 0.0f;
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    for (int c = 0; c < 3; ++c) {
      const size_t j = 3 * (size_t)i + c;
      p[j] = r[j] + beta[c] * p[j];
    }
}

// One block: rr = the summed r.r partials (rr0 = b.b at the start), retire
// converged columns.
__global__ void __launch_bounds__(kCgThreads) k_cg_scalars(CgScalars* s, float tol2, int first, int nb) {
  if (!first && cg_done(s)) return;
  float rr[3];
  sum_partials3(s->part_rr, nb, rr);
  if (first) {
    float b2[3];
    __syncthreads();  // sum_partials3's LDS is reused
    sum_partials3(s->part_b, nb, b2);
    if (threadIdx.x != 0u) return;
    s->active = 0u;
    for (int c = 0; c < 3; ++c) {
      s->rr[c] = rr[c];
      s->rr0[c] = b2[c];
      s->pap[c] = s->rr_new[c] = 0.0f;
      if (rr[c] > tol2 * b2[c] && rr[c] > 0.0f) s->active |= 1u << c;
    }
    s->iters = 0u;
    return;
  }
  if (threadIdx.x != 0u) return;
  s->iters += 1u;
  for (int c = 0; c < 3; ++c) {
    s->rr[c] = rr[c];
    s->rr_new[c] = rr[c];
    if (!(rr[c] > tol2 * s->rr0[c])) s->active &= ~(1u << c);
  }
}

__global__ void k_cg_reset(CgScalars* s) {
  for (int c = 0; c < 3; ++c) s->rr[c] = s->rr_new[c] = s->pap[c] = s->rr0[c] = 0.0f;
  s->active = 7u;
  s->iters = 0u;
}

static int cg_grid(int n) {
  const int g = (n + kCgThreads - 1) / kCgThreads;
  return g < 1 ? 1 : (g > kCgBlocks ? kCgBlocks : g);
}

hipError_t launch_csr_spmv3(const CsrDev& M, const float* x, float* y, hipStream_t st) {
  hipLaunchKernelGGL(k_csr_spmv3, dim3(cg_grid(M.n)), dim3(kCgThreads), 0, st, M, x, y, (CgScalars*)nullptr, 0, 0);
  return hipGetLastError();
}

hipError_t launch_cg(const CsrDev& M, const float* b, float* x, float* r, float* p, float* ap, CgScalars* s,
                     int max_iter, float tol, hipStream_t st) {
  const int n = M.n;
  const int nb = cg_grid(n);
  const dim3 g(nb), blk(kCgThreads);
  const float tol2 = tol * tol;
  hipLaunchKernelGGL(k_cg_reset, dim3(1), dim3(1), 0, st, s);
  hipLaunchKernelGGL(k_csr_spmv3, g, blk, 0, st, M, x, r, s, 0, 0);  // r = M x
  hipLaunchKernelGGL(k_cg_init, g, blk, 0, st, b, r, p, n, s);
  hipLaunchKernelGGL(k_cg_scalars, dim3(1), blk, 0, st, s, tol2, 1, nb);
  for (int it = 0; it < max_iter; ++it) {
    hipLaunchKernelGGL(k_csr_spmv3, g, blk, 0, st, M, p, ap, s, 1, 1);
    hipLaunchKernelGGL(k_cg_update, g, blk, 0, st, x, r, p, ap, n, s);
    hipLaunchKernelGGL(k_cg_direction, g, blk, 0, st, r, p, n, s);
    hipLaunchKernelGGL(k_cg_scalars, dim3(1), blk, 0, st, s, tol2, 0, nb);
  }
  return hipGetLastError();
}

}  // namespace trc
