// HIP kernels of the interior integrator and the ray-query entry point.
//
//   k_intersect     one ray per thread (trc_intersect)
//   k_render_paths  persistent megakernel: one path per thread iteration,
//                   per-path radiance to HBM (forward of PathTracer::Radiance)
//   k_film_reduce   deterministic per-pixel sum over samples
//   k_render_adj    persistent megakernel: replay + backward sweep per path
//
// All kernels use 128-thread blocks; the traversal short stack lives in LDS
// (kStackLds entries per thread, lane-interleaved) and spills to an HBM
// slice owned by the context only for unusually deep BVH paths.
#include <hip/hip_runtime.h>

#include "../trc_internal.h"
#include "adjoint.h"

namespace trc {

__device__ inline TraceCtx make_trace_ctx(uint32_t* lds, const LaunchCfg& L) {
  TraceCtx t;
  t.lds = lds + threadIdx.x;
  t.lds_stride = kBlock;
  t.lds_depth = kStackLds;
  const int gtid = blockIdx.x * kBlock + threadIdx.x;
  t.spill = L.spill + gtid;
  t.spill_stride = L.grid * kBlock;
  t.cap = L.cap;
  t.err = L.err;
  t.n_nodes = t.n_tris = t.n_queries = 0u;
  return t;
}

__global__ void __launch_bounds__(kBlock) k_intersect(SceneDev S, LaunchCfg L, const float* rays, int n,
                                                      int32_t* tri, float* uvt, int any_hit) {
  __shared__ uint32_t lds_stack[kStackLds * kBlock];
  TraceCtx tc = make_trace_ctx(lds_stack, L);
  const int stride = L.grid * kBlock;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const float* r = rays + 8 * (size_t)i;
    F3 o{r[0], r[1], r[2]}, d{r[4], r[5], r[6]};
    Hit h = any_hit ? trace_ray<true>(S, tc, o, d, r[3], r[7]) : trace_ray<false>(S, tc, o, d, r[3], r[7]);
    tri[i] = h.tri;
    uvt[3 * (size_t)i] = h.u;
    uvt[3 * (size_t)i + 1] = h.v;
    uvt[3 * (size_t)i + 2] = h.t;
  }
}

__global__ void __launch_bounds__(kBlock) k_render_paths(SceneDev S, CameraDev C, PathParams P, LaunchCfg L,
                                                         uint64_t npaths, float* paths) {
  __shared__ uint32_t lds_stack[kStackLds * kBlock];
  TraceCtx tc = make_trace_ctx(lds_stack, L);
  const uint64_t stride = (uint64_t)L.grid * kBlock;
  for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < npaths; p += stride) {
    F3 c = path_forward<false>(S, C, P, p, tc, nullptr);
    paths[3 * p] = c.x;
    paths[3 * p + 1] = c.y;
    paths[3 * p + 2] = c.z;
  }
}

// Work statistics of the forward paths: counters[0..3] += (paths, queries,
// node visits, triangle tests).  Used to compute algorithmic bytes.
__global__ void __launch_bounds__(kBlock) k_render_stats(SceneDev S, CameraDev C, PathParams P, LaunchCfg L,
                                                         uint64_t npaths, unsigned long long* counters) {
  __shared__ uint32_t lds_stack[kStackLds * kBlock];
  TraceCtx tc = make_trace_ctx(lds_stack, L);
  const uint64_t stride = (uint64_t)L.grid * kBlock;
  uint64_t n = 0;
  for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < npaths; p += stride) {
    F3 c = path_forward<false>(S, C, P, p, tc, nullptr);
    if (c.x == -1.0f) tc.n_queries++;  // keeps the path live
    ++n;
  }
  atomicAdd(&counters[0], (unsigned long long)n);
  atomicAdd(&counters[1], (unsigned long long)tc.n_queries);
  atomicAdd(&counters[2], (unsigned long long)tc.n_nodes);
  atomicAdd(&counters[3], (unsigned long long)tc.n_tris);
}

// film[c][pix] = (sum_s paths[s*WH + pix][c]) * (1/spp), summed in sample order.
__global__ void k_film_reduce(const float* paths, int spp, int wh, float* film) {
  int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= wh) return;
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  for (int s = 0; s < spp; ++s) {
    const float* p = paths + 3 * ((size_t)s * wh + pix);
    sx += p[0];
    sy += p[1];
    sz += p[2];
  }
  const float inv = 1.0f / (float)spp;
  film[pix] = sx * inv;
  film[wh + pix] = sy * inv;
  film[2 * wh + pix] = sz * inv;
}

__global__ void __launch_bounds__(kBlock) k_render_adj(SceneDev S, CameraDev C, PathParams P, LaunchCfg L,
                                                       uint64_t npaths, const float* dLdI, float* dpos,
                                                       float* dmat, float* adj_vn, float* dtex) {
  __shared__ uint32_t lds_stack[kStackLds * kBlock];
  __shared__ float mat_grad[kMaxMatGrad * kMatGradStride];
  for (int i = threadIdx.x; i < kMaxMatGrad * kMatGradStride; i += kBlock) mat_grad[i] = 0.0f;
  __syncthreads();
  TraceCtx tc = make_trace_ctx(lds_stack, L);
  const uint32_t WH = (uint32_t)C.width * (uint32_t)C.height;
  const float inv_spp = 1.0f / (float)P.spp;
  const uint64_t stride = (uint64_t)L.grid * kBlock;
  for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < npaths; p += stride) {
    const uint32_t pix = (uint32_t)(p % WH);
    F3 g{dLdI[pix] * inv_spp, dLdI[WH + pix] * inv_spp, dLdI[2 * WH + pix] * inv_spp};
    if (g.x == 0.0f && g.y == 0.0f && g.z == 0.0f) continue;
    PathRec R;
    path_forward<true>(S, C, P, p, tc, &R);
    path_backward(S, C, P, R, g, dpos, mat_grad, adj_vn, dtex);
  }
  __syncthreads();
  if (dmat) {
    const int n = (S.n_mats < kMaxMatGrad ? S.n_mats : kMaxMatGrad) * kMatGradStride;
    for (int i = threadIdx.x; i < n; i += kBlock)
      if (mat_grad[i] != 0.0f) atomicAdd(&dmat[i], mat_grad[i]);
  }
}

__global__ void __launch_bounds__(kBlock) k_vertex_normal_adj(SceneDev S, float* adj_vn, float* dpos) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  if (v < (uint32_t)S.n_verts) vertex_normal_backward(S, v, adj_vn, dpos);
}

hipError_t launch_intersect(const SceneDev& S, const LaunchCfg& L, const float* rays, int n, int32_t* tri,
                            float* uvt, int any_hit, hipStream_t st) {
  hipLaunchKernelGGL(k_intersect, dim3(L.grid), dim3(kBlock), 0, st, S, L, rays, n, tri, uvt, any_hit);
  return hipGetLastError();
}

hipError_t launch_render_paths(const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                               const LaunchCfg& L, float* paths, hipStream_t st) {
  PathParams P{seed, max_bounces, spp, 0u};
  uint64_t np = (uint64_t)spp * (uint64_t)C.width * (uint64_t)C.height;
  hipLaunchKernelGGL(k_render_paths, dim3(L.grid), dim3(kBlock), 0, st, S, C, P, L, np, paths);
  return hipGetLastError();
}

hipError_t launch_render_stats(const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                               const LaunchCfg& L, unsigned long long* counters, hipStream_t st) {
  PathParams P{seed, max_bounces, spp, 0u};
  uint64_t np = (uint64_t)spp * (uint64_t)C.width * (uint64_t)C.height;
  hipLaunchKernelGGL(k_render_stats, dim3(L.grid), dim3(kBlock), 0, st, S, C, P, L, np, counters);
  return hipGetLastError();
}

hipError_t launch_film_reduce(const float* paths, int spp, int wh, float* film, hipStream_t st) {
  hipLaunchKernelGGL(k_film_reduce, dim3((wh + 255) / 256), dim3(256), 0, st, paths, spp, wh, film);
  return hipGetLastError();
}

hipError_t launch_render_adj(const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                             const LaunchCfg& L, const float* dLdI, float* dpos, float* dmat, float* dtex,
                             float* vn_adj, hipStream_t st) {
  PathParams P{seed, max_bounces, spp, 0u};
  uint64_t np = (uint64_t)spp * (uint64_t)C.width * (uint64_t)C.height;
  hipLaunchKernelGGL(k_render_adj, dim3(L.grid), dim3(kBlock), 0, st, S, C, P, L, np, dLdI, dpos, dmat, vn_adj, dtex);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !vn_adj || S.n_verts == 0) return e;
  const int grid = (S.n_verts + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_vertex_normal_adj, dim3(grid), dim3(kBlock), 0, st, S, vn_adj, dpos);
  return hipGetLastError();
}

}  // namespace trc
