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
  t.leafq = lds + kStackLds * kBlock + threadIdx.x;
  t.spill = L.spill;
  t.spill_stride = L.grid * kBlock;
  t.cap = L.cap;
  t.n_nodes = t.n_tris = t.n_queries = 0u;
  t.slot_key = nullptr;
  t.slot_uv = nullptr;
  return t;
}

// The context of a kernel whose lds array also holds the wave-wide leaf
// tests' slots: kTraceLdsWords + kTraceSlotWords words per lane.
__device__ inline TraceCtx make_trace_ctx_wave(uint32_t* lds, const LaunchCfg& L) {
  TraceCtx t = make_trace_ctx(lds, L);
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(lds + kTraceLdsWords * kBlock);
  t.slot_key = keys + threadIdx.x;
  t.slot_uv = reinterpret_cast<float2*>(keys + kBlock) + threadIdx.x;
  return t;
}

__global__ void __launch_bounds__(kBlock) k_intersect(SceneDev S, LaunchCfg L, const float* rays, int n,
                                                      int32_t* tri, float* uvt, int any_hit) {
  __shared__ uint32_t lds_stack[kTraceLdsWords * kBlock];
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

template <bool kCkpt>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kFwdWaves))) k_render_paths(SceneDev S, CameraDev C, PathParams P, LaunchCfg L,
                                                         uint64_t first, uint64_t count, uint32_t tile, float* paths,
                                                         uint64_t paths_first, CkptDev ck) {
  __shared__ __align__(8) uint32_t lds_stack[(kTraceLdsWords + kTraceSlotWords) * kBlock];
  TraceCtx tc = make_trace_ctx_wave(lds_stack, L);
  // paths [first, first + count) in tiles of `tile` paths, dealt to the waves round-robin
  const uint64_t nwaves = (uint64_t)L.grid * (kBlock / 64);
  const uint64_t wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  const uint64_t ntiles = (count + tile - 1) / tile;
  TileSupply sup = TileSupply::make(wave, ntiles, nwaves, 0, first, first + count, tile);
  sup.ctr = L.tiles;
  render_paths_regen<kCkpt>(S, C, P, tc, sup, ck, [&](uint64_t p, F3 c) {
    const uint64_t q = p - paths_first;
    paths[3 * q] = c.x;
    paths[3 * q + 1] = c.y;
    paths[3 * q + 2] = c.z;
  });
}

// Work statistics of the forward paths: counters[0..3] += (paths, queries,
// node visits, triangle tests).  Used to compute algorithmic bytes.
__global__ void __launch_bounds__(kBlock) k_render_stats(SceneDev S, CameraDev C, PathParams P, LaunchCfg L,
                                                         uint64_t npaths, unsigned long long* counters) {
  __shared__ uint32_t lds_stack[kTraceLdsWords * kBlock];
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

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kAdjWaves))) k_render_adj(SceneDev S, CameraDev C, PathParams P, LaunchCfg L,
                                                       uint64_t npaths, const float* dLdI, float* dpos,
                                                       float* dmat, float* adj_vn, float* dtex, GradCopies gc) {
  if (gc.base) dpos = gc.base + (size_t)(blockIdx.x % gc.copies) * gc.n3;
  __shared__ uint32_t lds_stack[kTraceLdsWords * kBlock];
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
    GradSink sk{dpos, mat_grad, adj_vn, dtex};
    path_backward(S, C, P, R, g, sk);
  }
  __syncthreads();
  if (dmat) {
    const int n = (S.n_mats < kMaxMatGrad ? S.n_mats : kMaxMatGrad) * kMatGradStride;
    for (int i = threadIdx.x; i < n; i += kBlock)
      if (mat_grad[i] != 0.0f) atomicAdd(&dmat[i], mat_grad[i]);
  }
}

// Adjoint from a path checkpoint: the path is re-evaluated from its recorded
// hits (no traversal), then swept backwards as in k_render_adj.  Blocks take
// chunks of kAdjChunk consecutive paths from a counter (the next free chunk,
// so no block idles while another still has several) and order each chunk
// in LDS by path length (longest first, camera misses and zero dL/dI
// dropped) with a counting sort, so the waves sweep paths of equal length:
// on the metric scene 23 % of the camera rays miss and a wave of 64
// consecutive paths runs 2.6 vertices where its paths average 1.1.
// Measured there: 1.84 ms unsorted, 1.40 with 256-path chunks, 1.12 with
// 1,024, 1.04 with 1,024 fetched from the counter (2,048: 1.19).
constexpr int kAdjSortR = 8;                   // paths per thread per chunk
constexpr int kAdjChunk = kBlock * kAdjSortR;
constexpr int kLenBuckets = kMaxBounces + 2;   // K + 1 in 0 .. max_bounces + 1
constexpr uint64_t kAdjTail = 256;             // paths per chunk at the end of a launch

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kAdjWaves))) k_render_adj_ckpt(SceneDev S, CameraDev C, PathParams P, LaunchCfg L,
                                                            uint64_t first, uint64_t count, CkptDev ck,
                                                            const float* dLdI, float* dpos, float* dmat, float* adj_vn,
                                                            float* dtex, GradCopies gc, uint32_t* chunk_ctr) {
  if (gc.base) dpos = gc.base + (size_t)(blockIdx.x % gc.copies) * gc.n3;
  __shared__ float mat_grad[kMaxMatGrad * kMatGradStride];
  __shared__ uint32_t s_ch;
  __shared__ uint16_t order[kAdjChunk];
  __shared__ uint32_t bcount[kLenBuckets], bnext[kLenBuckets];
  for (int i = threadIdx.x; i < kMaxMatGrad * kMatGradStride; i += kBlock) mat_grad[i] = 0.0f;
  const uint32_t WH = (uint32_t)C.width * (uint32_t)C.height;
  const float inv_spp = 1.0f / (float)P.spp;
  const uint64_t end = first + count;
  // the last paths come in chunks of kAdjTail, so that the blocks finish
  // together: an eighth of the paths, at most two such chunks per block
  // (metric scene: 0.91 ms with 1,024-path chunks only, 0.86 with this tail)
  const uint64_t tail_max = (uint64_t)gridDim.x * 2u * kAdjTail;
  const uint64_t tail = count / 8u < tail_max ? count / 8u : tail_max;
  const uint64_t nbig = (count - tail) / kAdjChunk;
  const uint64_t nchunks = nbig + (count - nbig * kAdjChunk + kAdjTail - 1) / kAdjTail;
  for (;;) {
    if (threadIdx.x == 0) s_ch = atomicAdd(chunk_ctr, 1u);
    if (threadIdx.x < kLenBuckets) bcount[threadIdx.x] = 0u;
    __syncthreads();
    const uint64_t ch = s_ch;
    if (ch >= nchunks) break;  // s_ch is the block's: every thread leaves together
    const uint64_t base = ch < nbig ? first + ch * kAdjChunk : first + nbig * kAdjChunk + (ch - nbig) * kAdjTail;
    const uint64_t cend0 = base + (ch < nbig ? kAdjChunk : kAdjTail);
    const uint64_t cend = cend0 < end ? cend0 : end;
    uint32_t kb[kAdjSortR];
#pragma unroll
    for (int r = 0; r < kAdjSortR; ++r) {
      const uint64_t p = base + threadIdx.x + (uint64_t)r * kBlock;
      kb[r] = 0u;
      if (p < cend) {
        const uint32_t pix = (uint32_t)(p % WH);
        const bool zero = dLdI[pix] == 0.0f && dLdI[WH + pix] == 0.0f && dLdI[2 * WH + pix] == 0.0f;
        const uint32_t k1 = ck.meta[p] & 0xFFu;
        kb[r] = zero ? 0u : (k1 < (uint32_t)kLenBuckets ? k1 : (uint32_t)kLenBuckets - 1u);
        if (kb[r]) atomicAdd(&bcount[kb[r]], 1u);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t o = 0u;
      for (int b = kLenBuckets - 1; b >= 1; --b) {
        bnext[b] = o;
        o += bcount[b];
      }
      bcount[0] = o;  // paths to sweep in this chunk
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kAdjSortR; ++r)
      if (kb[r]) order[atomicAdd(&bnext[kb[r]], 1u)] = (uint16_t)(threadIdx.x + r * kBlock);
    __syncthreads();
    const uint32_t nsweep = bcount[0];
    for (uint32_t j = threadIdx.x; j < nsweep; j += kBlock) {
      const uint64_t p = base + order[j];
      const uint32_t pix = (uint32_t)(p % WH);
      F3 g{dLdI[pix] * inv_spp, dLdI[WH + pix] * inv_spp, dLdI[2 * WH + pix] * inv_spp};
      CkptSrc src{&ck, p, ck.meta[p]};
      PathRec R;
      path_forward<true>(S, C, P, p, src, &R);
      GradSink sk{dpos, mat_grad, adj_vn, dtex};
      path_backward(S, C, P, R, g, sk);
    }
    __syncthreads();  // order / bcount are reused by the next chunk
  }
  __syncthreads();
  if (dmat) {
    const int n = (S.n_mats < kMaxMatGrad ? S.n_mats : kMaxMatGrad) * kMatGradStride;
    for (int i = threadIdx.x; i < n; i += kBlock)
      if (mat_grad[i] != 0.0f) atomicAdd(&dmat[i], mat_grad[i]);
  }
}

// Derivative image of the interior integrator along one parameter tangent
// (the reference's export_deriv: GradiantImageHandler::AccumulateDeriv,
// Integrator.cpp:338-373, the forward-mode derivative of every path's
// contribution): each path is replayed once and swept backwards three
// times, with the seeds e_r, e_g, e_b / spp, into a TangentSink; the sums
// are the derivatives of the path's RGB contribution and go to
// dimg[c][pixel].  Non-finite derivatives are dropped, as the reference
// filters its derivative images (Integrator.cpp:364-367).
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kAdjWaves))) k_render_deriv(SceneDev S, CameraDev C, PathParams P, LaunchCfg L,
                                                         uint64_t npaths, const float* tpos, const float* tmat,
                                                         const float* tvn, float* dimg) {
  __shared__ uint32_t lds_stack[kTraceLdsWords * kBlock];
  TraceCtx tc = make_trace_ctx(lds_stack, L);
  const uint32_t WH = (uint32_t)C.width * (uint32_t)C.height;
  const float inv_spp = 1.0f / (float)P.spp;
  const uint64_t stride = (uint64_t)L.grid * kBlock;
  for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < npaths; p += stride) {
    const uint32_t pix = (uint32_t)(p % WH);
    PathRec R;
    path_forward<true>(S, C, P, p, tc, &R);
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
      const F3 g{c == 0 ? inv_spp : 0.0f, c == 1 ? inv_spp : 0.0f, c == 2 ? inv_spp : 0.0f};
      TangentSink sk{tpos, tmat, tvn, 0.0f};
      path_backward(S, C, P, R, g, sk);
      if (sk.acc != 0.0f && isfinite(sk.acc)) atomicAdd(&dimg[(size_t)c * WH + pix], sk.acc);
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_vertex_normal_tangent(SceneDev S, const float* tpos, float* tvn) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= (uint32_t)S.n_verts) return;
  const F3 d = vertex_normal_tangent(S, v, tpos);
  tvn[3 * v] = d.x;
  tvn[3 * v + 1] = d.y;
  tvn[3 * v + 2] = d.z;
}

__global__ void __launch_bounds__(kBlock) k_vertex_normal_adj(SceneDev S, float* adj_vn, float* dpos) {
  const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
  if (v < (uint32_t)S.n_verts) vertex_normal_backward(S, v, adj_vn, dpos);
}

// ---------------------------------------------------------------------------
// Privatized gradient copies
// ---------------------------------------------------------------------------
__global__ void k_fill_zero(float* p, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.0f;
}

__global__ void k_grad_reduce(float* base, uint32_t copies, uint32_t n3, float* dpos, uint32_t rezero) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += stride) {
    float s = 0.0f;
    for (uint32_t k = 0; k < copies; ++k) {
      float* q = base + (size_t)k * n3 + i;
      s += *q;
      if (rezero) *q = 0.0f;
    }
    if (s != 0.0f) dpos[i] += s;
  }
}

uint32_t grad_copies_for(uint64_t n3) {
  // above ~64k vertices the paths spread over enough vertices that atomics
  // rarely collide, and the copies would cost more to clear and reduce
  if (n3 == 0 || n3 > (192u << 10)) return 1;
  uint64_t k = (4u << 20) / n3;  // at most 16 MB of copies
  return (uint32_t)(k > 64 ? 64 : (k < 2 ? 1 : k));
}

hipError_t launch_fill_zero(float* p, uint64_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  uint64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_grad_reduce(const GradCopies& g, float* dpos, hipStream_t st) {
  const uint32_t blocks = (g.n3 + 255) / 256;
  hipLaunchKernelGGL(k_grad_reduce, dim3(blocks < 4096 ? blocks : 4096), dim3(256), 0, st, g.base, g.copies, g.n3,
                     dpos, g.keep);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched launches: the grid walks the batch's tiles (contiguous path ranges
// of one job); a wave takes tiles wave, wave + nwaves, ...  The job of a tile
// is wave-uniform, so its scene / camera descriptors are read once per tile
// into scalar registers.
// ---------------------------------------------------------------------------
template <class T>
__device__ inline T wave_uniform_load(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "word-sized struct");
  T out;
  const uint32_t* src = reinterpret_cast<const uint32_t*>(p);
  uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
  return out;
}

__device__ inline int job_of_tile(const JobDev* jobs, int njobs, uint64_t t) {
  int lo = 0, hi = njobs - 1;  // last job with tile0 <= t
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].tile0 <= t) lo = mid;
    else hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kFwdWaves))) k_render_batch_fwd(const JobDev* __restrict__ jobs, int njobs, uint64_t ntiles,
                                                             uint32_t tile, LaunchCfg L) {
  __shared__ __align__(8) uint32_t lds_stack[(kTraceLdsWords + kTraceSlotWords) * kBlock];
  TraceCtx tc = make_trace_ctx_wave(lds_stack, L);
  const uint64_t nwaves = (uint64_t)L.grid * (kBlock / 64);
  const uint64_t wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  for (uint64_t t = wave; t < ntiles;) {
    const int j = job_of_tile(jobs, njobs, t);
    // read in place from the (read-only) job table: copying the records
    // costs registers, and spill, in the path loop
    const SceneDev& S = jobs[j].S;
    const CameraDev& C = jobs[j].C;
    const PathParams P{jobs[j].seed, jobs[j].max_bounces, jobs[j].spp, 0u};
    const uint64_t first = jobs[j].first, count = jobs[j].count, t0 = jobs[j].tile0;
    const uint64_t npaths = (uint64_t)P.spp * (uint64_t)C.width * (uint64_t)C.height;
    const CkptDev ck = ckpt_view(jobs[j].ckpt, npaths, P.max_bounces);
    float* paths = jobs[j].paths;
    // the wave's tiles of this job, rendered without draining between them
    const uint64_t tend = t0 + (count + tile - 1) / tile;
    TileSupply sup = TileSupply::make(t, tend, nwaves, t0, first, first + count, tile);
    sup.ctr = L.tiles;
    render_paths_regen<true>(S, C, P, tc, sup, ck, [&](uint64_t p, F3 c) {
      paths[3 * p] = c.x;
      paths[3 * p + 1] = c.y;
      paths[3 * p + 2] = c.z;
    });
    t = sup.t;
  }
}

// Material gradients are summed per wave in LDS and flushed at the end of
// each tile (one job per tile).  Like k_render_adj_ckpt, each wave sorts
// its tile's paths by length in LDS, kAdjChunk paths at a time, before the
// sweep.
__device__ inline uint32_t adj_bucket(const CkptDev& ck, const float* dLdI, uint32_t WH, uint64_t p) {
  const uint32_t pix = (uint32_t)(p % WH);
  if (dLdI[pix] == 0.0f && dLdI[WH + pix] == 0.0f && dLdI[2 * WH + pix] == 0.0f) return 0u;
  const uint32_t k1 = ck.meta[p] & 0xFFu;
  return k1 < (uint32_t)kLenBuckets ? k1 : (uint32_t)kLenBuckets - 1u;
}

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kAdjWaves))) k_render_batch_adj(const JobDev* __restrict__ jobs, int njobs, uint64_t ntiles,
                                                             uint32_t tile, LaunchCfg L) {
  __shared__ float mat_lds[(kBlock / 64) * kMaxMatGrad * kMatGradStride];
  __shared__ uint16_t ord_lds[(kBlock / 64) * kAdjChunk];
  __shared__ uint32_t cnt_lds[(kBlock / 64) * 2 * kLenBuckets];
  const int lane = (int)(threadIdx.x & 63u);
  float* mat_grad = mat_lds + (threadIdx.x >> 6) * (kMaxMatGrad * kMatGradStride);
  uint16_t* order = ord_lds + (threadIdx.x >> 6) * kAdjChunk;
  uint32_t* bcount = cnt_lds + (threadIdx.x >> 6) * 2 * kLenBuckets;
  uint32_t* bnext = bcount + kLenBuckets;
  const uint64_t nwaves = (uint64_t)L.grid * (kBlock / 64);
  const uint64_t wave = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  // the wave's first tile is tile `wave`; later ones come from the tile
  // counter as the wave needs them (round-robin without one)
  for (uint64_t t = wave; t < ntiles;) {
    const int j = job_of_tile(jobs, njobs, t);
    // read in place from the (read-only) job table: copying the records
    // costs registers, and spill, in the path loop
    const SceneDev& S = jobs[j].S;
    const CameraDev& C = jobs[j].C;
    const PathParams P{jobs[j].seed, jobs[j].max_bounces, jobs[j].spp, 0u};
    const uint64_t first = jobs[j].first, count = jobs[j].count, t0 = jobs[j].tile0;
    const uint32_t WH = (uint32_t)C.width * (uint32_t)C.height;
    const uint64_t npaths = (uint64_t)P.spp * (uint64_t)WH;
    const CkptDev ck = ckpt_view(jobs[j].ckpt, npaths, P.max_bounces);
    const float* dLdI = jobs[j].dLdI;
    const GradCopies gc = wave_uniform_load(&jobs[j].gc);
    float* dpos = gc.base ? gc.base + (size_t)(blockIdx.x % gc.copies) * gc.n3 : jobs[j].dpos;
    float* dmat = jobs[j].dmat;
    float* dtex = jobs[j].dtex;
    float* vn = jobs[j].vn;
    const float inv_spp = 1.0f / (float)P.spp;
    for (int i = lane; i < kMaxMatGrad * kMatGradStride; i += 64) mat_grad[i] = 0.0f;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint64_t begin = first + (t - t0) * tile;
    const uint64_t end = begin + tile < first + count ? begin + tile : first + count;
    for (uint64_t sub = begin; sub < end; sub += kAdjChunk) {
      const uint64_t send = sub + kAdjChunk < end ? sub + kAdjChunk : end;
      if (lane < kLenBuckets) bcount[lane] = 0u;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      for (uint64_t p = sub + (uint64_t)lane; p < send; p += 64) {
        const uint32_t b = adj_bucket(ck, dLdI, WH, p);
        if (b) atomicAdd(&bcount[b], 1u);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      if (lane == 0) {
        uint32_t o = 0u;
        for (int b = kLenBuckets - 1; b >= 1; --b) {
          bnext[b] = o;
          o += bcount[b];
        }
        bcount[0] = o;
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      for (uint64_t p = sub + (uint64_t)lane; p < send; p += 64) {
        const uint32_t b = adj_bucket(ck, dLdI, WH, p);
        if (b) order[atomicAdd(&bnext[b], 1u)] = (uint16_t)(p - sub);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const uint32_t nsweep = bcount[0];
      for (uint32_t q = (uint32_t)lane; q < nsweep; q += 64) {
        const uint64_t p = sub + order[q];
        const uint32_t pix = (uint32_t)(p % WH);
        F3 g{dLdI[pix] * inv_spp, dLdI[WH + pix] * inv_spp, dLdI[2 * WH + pix] * inv_spp};
        CkptSrc src{&ck, p, ck.meta[p]};
        PathRec R;
        path_forward<true>(S, C, P, p, src, &R);
        GradSink sk{dpos, mat_grad, vn, dtex};
        path_backward(S, C, P, R, g, sk);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (dmat) {
      const int n = (S.n_mats < kMaxMatGrad ? S.n_mats : kMaxMatGrad) * kMatGradStride;
      for (int i = lane; i < n; i += 64)
        if (mat_grad[i] != 0.0f) atomicAdd(&dmat[i], mat_grad[i]);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (L.tiles) {
      uint32_t v = 0u;
      if (lane == 0) v = atomicAdd(L.tiles, 1u);
      t = nwaves + (uint64_t)(uint32_t)__shfl((int)v, 0);
    } else {
      t += nwaves;
    }
  }
}

hipError_t launch_render_batch(const JobDev* jobs_dev, const JobDev* jobs_host, int njobs, uint64_t ntiles,
                               uint32_t tile, int adjoint, const LaunchCfg& L, hipStream_t st) {
  hipError_t e = hipSuccess;
  // jobs large enough to fill the GPU on their own run through the
  // single-request kernels over their path range (no mixing of scenes in the
  // caches, no tile bookkeeping); the others share one tiled launch
  for (int j = 0; j < njobs && e == hipSuccess; ++j) {
    const JobDev& J = jobs_host[j];
    if (!J.solo) continue;
    if (!adjoint)
      e = launch_render_paths(J.S, J.C, J.seed, J.max_bounces, J.spp, L, J.paths, J.ckpt, st, J.first, J.count);
    else
      e = launch_render_adj_ckpt(J.S, J.C, J.seed, J.max_bounces, J.spp, L, J.ckpt, J.dLdI, J.dpos, J.dmat, J.dtex,
                                 J.vn, J.gc, st, J.first, J.count, /*fold=*/false);
  }
  if (e != hipSuccess) return e;
  if (!adjoint) {
    if (ntiles == 0) return hipSuccess;
    const LaunchCfg Lk = fit_grid(k_render_batch_fwd, L);
    if (Lk.tiles && (e = launch_fill_zero(reinterpret_cast<float*>(Lk.tiles), 1, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_render_batch_fwd, dim3(Lk.grid), dim3(kBlock), 0, st, jobs_dev, njobs, ntiles, tile, Lk);
    return hipGetLastError();
  }
  if (ntiles > 0) {
    const LaunchCfg Lk = fit_grid(k_render_batch_adj, L);
    if (Lk.tiles && (e = launch_fill_zero(reinterpret_cast<float*>(Lk.tiles), 1, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_render_batch_adj, dim3(Lk.grid), dim3(kBlock), 0, st, jobs_dev, njobs, ntiles, tile, Lk);
    e = hipGetLastError();
  }
  for (int j = 0; j < njobs && e == hipSuccess; ++j) {
    const JobDev& J = jobs_host[j];
    bool first_use = J.gc.base != nullptr;  // copies shared by jobs with one dpos are reduced once
    for (int q = 0; q < j && first_use; ++q) first_use = jobs_host[q].gc.base != J.gc.base;
    if (first_use) e = launch_grad_reduce(J.gc, J.dpos, st);
    if (e != hipSuccess) break;
    if (!J.vn || !J.S.vf_off || J.S.n_verts == 0) continue;
    bool vn_first = true;  // a scratch shared by several jobs is folded once
    for (int q = 0; q < j && vn_first; ++q) vn_first = jobs_host[q].vn != J.vn;
    if (!vn_first) continue;
    const int grid = (J.S.n_verts + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_vertex_normal_adj, dim3(grid), dim3(kBlock), 0, st, J.S, J.vn, J.dpos);
    e = hipGetLastError();
  }
  return e;
}

hipError_t launch_intersect(const SceneDev& S, const LaunchCfg& L, const float* rays, int n, int32_t* tri,
                            float* uvt, int any_hit, hipStream_t st) {
  const LaunchCfg Lk = fit_grid(k_intersect, L);
  hipLaunchKernelGGL(k_intersect, dim3(Lk.grid), dim3(kBlock), 0, st, S, Lk, rays, n, tri, uvt, any_hit);
  return hipGetLastError();
}

// About four tiles per wave of the grid, 64..4096 paths each.
// Paths per tile: with the tile counter, small tiles taken as waves need
// them (kDynTile); dealt statically, about four tiles per wave.
constexpr uint32_t kDynTile = 64;
static uint32_t forward_tile(uint64_t npaths, int grid, bool dynamic) {
  if (dynamic) return kDynTile;
  const uint64_t waves = (uint64_t)grid * (kBlock / 64);
  const uint64_t t = npaths / (waves * 4) / 64 * 64;
  return (uint32_t)(t < 64 ? 64 : (t > 4096 ? 4096 : t));
}

template <bool kCkpt>
static hipError_t launch_paths(const SceneDev& S, const CameraDev& C, const PathParams& P, const LaunchCfg& L,
                               uint64_t first, uint64_t count, float* paths, uint64_t paths_first, const CkptDev& ck,
                               hipStream_t st) {
  const LaunchCfg Lk = fit_grid(k_render_paths<kCkpt>, L);
  if (Lk.tiles) {
    const hipError_t e = launch_fill_zero(reinterpret_cast<float*>(Lk.tiles), 1, st);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_render_paths<kCkpt>), dim3(Lk.grid), dim3(kBlock), 0, st, S, C, P, Lk, first,
                     count, forward_tile(count, Lk.grid, Lk.tiles != nullptr), paths, paths_first, ck);
  return hipGetLastError();
}

hipError_t launch_render_paths(const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                               const LaunchCfg& L, float* paths, void* ckpt, hipStream_t st, uint64_t first,
                               uint64_t count, uint64_t paths_first) {
  PathParams P{seed, max_bounces, spp, 0u};
  const uint64_t np = (uint64_t)spp * (uint64_t)C.width * (uint64_t)C.height;
  if (count == 0) {
    first = 0;
    count = np;
  }
  const CkptDev ck = ckpt ? ckpt_view(ckpt, np, max_bounces) : CkptDev{};
  if (ckpt) return launch_paths<true>(S, C, P, L, first, count, paths, paths_first, ck, st);
  return launch_paths<false>(S, C, P, L, first, count, paths, paths_first, ck, st);
}

uint64_t checkpoint_bytes(uint64_t npaths, int max_bounces) { return ckpt_bytes(npaths, max_bounces); }

hipError_t launch_render_adj_ckpt(const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                                  const LaunchCfg& L, const void* ckpt, const float* dLdI, float* dpos, float* dmat,
                                  float* dtex, float* vn_adj, const GradCopies& gc, hipStream_t st,
                                  uint64_t first, uint64_t count, bool fold) {
  PathParams P{seed, max_bounces, spp, 0u};
  const uint64_t np = (uint64_t)spp * (uint64_t)C.width * (uint64_t)C.height;
  if (count == 0) {
    first = 0;
    count = np;
  }
  const CkptDev ck = ckpt_view(const_cast<void*>(ckpt), np, max_bounces);
  const LaunchCfg Lk = fit_grid(k_render_adj_ckpt, L);
  // this kernel traces no rays: word 0 of the stream's traversal spill
  // buffer serves as its chunk counter, cleared by a kernel on the stream
  if (!L.spill) return hipErrorInvalidValue;
  hipError_t e = launch_fill_zero((float*)L.spill, 1, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_render_adj_ckpt, dim3(Lk.grid), dim3(kBlock), 0, st, S, C, P, Lk, first, count, ck, dLdI,
                     dpos, dmat, vn_adj, dtex, gc, L.spill);
  e = hipGetLastError();
  if (!fold) return e;  // the caller folds the gradient copies and vertex normals
  if (e == hipSuccess && gc.base) e = launch_grad_reduce(gc, dpos, st);
  if (e != hipSuccess || !vn_adj || S.n_verts == 0) return e;
  const int grid = (S.n_verts + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_vertex_normal_adj, dim3(grid), dim3(kBlock), 0, st, S, vn_adj, dpos);
  return hipGetLastError();
}

hipError_t launch_render_stats(const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                               const LaunchCfg& L, unsigned long long* counters, hipStream_t st) {
  PathParams P{seed, max_bounces, spp, 0u};
  uint64_t np = (uint64_t)spp * (uint64_t)C.width * (uint64_t)C.height;
  const LaunchCfg Lk = fit_grid(k_render_stats, L);
  hipLaunchKernelGGL(k_render_stats, dim3(Lk.grid), dim3(kBlock), 0, st, S, C, P, Lk, np, counters);
  return hipGetLastError();
}

hipError_t launch_film_reduce(const float* paths, int spp, int wh, float* film, hipStream_t st) {
  hipLaunchKernelGGL(k_film_reduce, dim3((wh + 255) / 256), dim3(256), 0, st, paths, spp, wh, film);
  return hipGetLastError();
}

// k_film_reduce over samples [s0, s0 + ns) only: paths[s][pix] holds sample
// s0 + s.  The running per-pixel sums stay unscaled in film between chunks
// and the last chunk applies 1/spp, so the additions are the same, in the
// same order, as one k_film_reduce over all samples.
__global__ void k_film_accum(const float* paths, int ns, int wh, float* film, int first_chunk, int last_chunk,
                             int spp) {
  int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= wh) return;
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  if (!first_chunk) {
    sx = film[pix];
    sy = film[wh + pix];
    sz = film[2 * wh + pix];
  }
  for (int s = 0; s < ns; ++s) {
    const float* p = paths + 3 * ((size_t)s * wh + pix);
    sx += p[0];
    sy += p[1];
    sz += p[2];
  }
  if (last_chunk) {
    const float inv = 1.0f / (float)spp;
    sx *= inv;
    sy *= inv;
    sz *= inv;
  }
  film[pix] = sx;
  film[wh + pix] = sy;
  film[2 * wh + pix] = sz;
}

hipError_t launch_film_accum(const float* paths, int ns, int wh, float* film, bool first_chunk, bool last_chunk,
                             int spp, hipStream_t st) {
  hipLaunchKernelGGL(k_film_accum, dim3((wh + 255) / 256), dim3(256), 0, st, paths, ns, wh, film,
                     first_chunk ? 1 : 0, last_chunk ? 1 : 0, spp);
  return hipGetLastError();
}

// Scratch of trc_render_fwd_noalloc: the traversal spill slice and the
// per-path radiance of one chunk of samples.  Host code never names this
// array in a kernel argument or copy: it uses the address that
// noalloc_scratch_address returns for the current device.
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

size_t noalloc_scratch_bytes() { return kNoAllocScratchBytes; }

hipError_t noalloc_scratch_address(void** out) { return hipGetSymbolAddress(out, HIP_SYMBOL(scratch)); }

hipError_t launch_render_adj(const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                             const LaunchCfg& L, const float* dLdI, float* dpos, float* dmat, float* dtex,
                             float* vn_adj, const GradCopies& gc, hipStream_t st) {
  PathParams P{seed, max_bounces, spp, 0u};
  uint64_t np = (uint64_t)spp * (uint64_t)C.width * (uint64_t)C.height;
  const LaunchCfg Lk = fit_grid(k_render_adj, L);
  hipLaunchKernelGGL(k_render_adj, dim3(Lk.grid), dim3(kBlock), 0, st, S, C, P, Lk, np, dLdI, dpos, dmat, vn_adj, dtex,
                     gc);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && gc.base) e = launch_grad_reduce(gc, dpos, st);
  if (e != hipSuccess || !vn_adj || S.n_verts == 0) return e;
  const int grid = (S.n_verts + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_vertex_normal_adj, dim3(grid), dim3(kBlock), 0, st, S, vn_adj, dpos);
  return hipGetLastError();
}

hipError_t launch_render_deriv(const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                               const LaunchCfg& L, const float* tpos, const float* tmat, float* tvn, float* dimg,
                               hipStream_t st) {
  const PathParams P{seed, max_bounces, spp, 0u};
  const uint64_t np = (uint64_t)spp * (uint64_t)C.width * (uint64_t)C.height;
  if (np == 0) return hipSuccess;
  if (tvn) {
    const unsigned nb = (unsigned)((S.n_verts + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_vertex_normal_tangent, dim3(nb), dim3(kBlock), 0, st, S, tpos, tvn);
  }
  const LaunchCfg Lk = fit_grid(k_render_deriv, L);
  hipLaunchKernelGGL(k_render_deriv, dim3(Lk.grid), dim3(kBlock), 0, st, S, C, P, Lk, np, tpos, tmat,
                     (const float*)tvn, dimg);
  return hipGetLastError();
}

}  // namespace trc
