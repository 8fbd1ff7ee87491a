// Device-side scene update (trc_scene_set_vertices with refit) and the
// per-camera primary-edge list.
//
// After the vertex positions change, everything derived from them is
// recomputed where the scene lives, with the same float operations in the
// same order as the host code of the first configure
// (csrc/host/scene_host.cpp, csrc/host/bvh.cpp), so a refit scene is
// bit-identical to one the host would have produced:
//
//   k_face_normals    face normal and area per triangle
//                     (TriangleMesh::ComputeFaceNormalAndDistrib, Primitive.cpp:299-321)
//   k_vertex_normals  normalized mean of the incident smooth face normals, summed
//                     in triangle order (ComputeVertexNormal, Primitive.cpp:323-340)
//   k_light_table     emitter sampling weights and CDF (Scene::ConfigureLight,
//                     Scene.cpp:471-520; Distribution1D::SetFunction, Distribution.cpp:45-58)
//   k_edge_count / k_block_scan / k_edge_scatter
//                     order-preserving selection of edges (secondary edges:
//                     Edge.cpp:42-88; primary edges for an eye: Edge.cpp:313-348)
//   k_edge_len / k_cdf_local / k_cdf_offsets / k_cdf_final / k_edge_cdf_norm
//                     their length distribution (Edge.cpp:90-106, 350-367)
//   k_tri_records / k_refit_level / k_sah
//                     BVH refit level by level, deepest first, and the tree's
//                     SAH cost (the reference refits its GAS, Scene.cpp:368-469)
//
// The prefix sums over floats that define a CDF follow the host code's
// order (sequential for the emitter table, blocked for the edge lists): they
// are part of what the results must reproduce bit for bit.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "../trc_internal.h"

namespace trc {

namespace {

constexpr int kUpdBlock = 256;

__device__ inline F3 pos3(const float* p, uint32_t i) { return F3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

__device__ inline F3 face_normal(const TriShade* sh, int32_t f) {
  if (f < 0) return F3{0.0f, 0.0f, 0.0f};
  return F3{sh[f].n[0], sh[f].n[1], sh[f].n[2]};
}

// std::min / std::max as the host builder uses them: min(a, b) = b < a ? b : a.
__device__ inline float mn(float a, float b) { return b < a ? b : a; }
__device__ inline float mx(float a, float b) { return a < b ? b : a; }

struct DBox {
  float lo[3], hi[3];
  __device__ void clear() {
    lo[0] = lo[1] = lo[2] = FLT_MAX;
    hi[0] = hi[1] = hi[2] = -FLT_MAX;
  }
  __device__ void grow(const float* p) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = mn(lo[k], p[k]);
      hi[k] = mx(hi[k], p[k]);
    }
  }
  __device__ void grow(const DBox& b) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = mn(lo[k], b.lo[k]);
      hi[k] = mx(hi[k], b.hi[k]);
    }
  }
};

// Outward rounding of the stored child boxes (as the host builder's).
__device__ inline float pad_lo(float x) { return x - (fabsf(x) * 9.5367431640625e-7f + 1e-7f); }
__device__ inline float pad_hi(float x) { return x + (fabsf(x) * 9.5367431640625e-7f + 1e-7f); }

struct SecPred {
  const float* pos;
  const TriShade* shade;
  __device__ bool operator()(const EdgeDev& e) const {
    if (e.f1 < 0) return true;
    const F3 n0 = face_normal(shade, e.f0), n1 = face_normal(shade, e.f1);
    const F3 p0 = pos3(pos, e.v0), p2 = pos3(pos, e.v2);
    const bool noncoplanar = dot(n0, n1) < 1.0f - 1e-6f;
    const bool convex = dot(n1, normalize(p2 - p0)) < -kEdgeEpsilon;
    return noncoplanar && convex;
  }
};

struct PrimPred {
  const float* pos;
  const TriShade* shade;
  F3 eye;
  __device__ bool operator()(const EdgeDev& e) const {
    const F3 d = eye - pos3(pos, e.v0);
    const bool ff0 = dot(d, face_normal(shade, e.f0)) > 0.0f;
    const bool ff1 = dot(d, face_normal(shade, e.f1)) > 0.0f;
    return e.f1 < 0 || ff0 != ff1;
  }
};

__global__ void k_face_normals(const float* pos, const uint32_t* idx, TriShade* shade, uint32_t T) {
  const uint32_t t = blockIdx.x * kUpdBlock + threadIdx.x;
  if (t >= T) return;
  const F3 p0 = pos3(pos, idx[3 * t]), p1 = pos3(pos, idx[3 * t + 1]), p2 = pos3(pos, idx[3 * t + 2]);
  const F3 c = cross(p1 - p0, p2 - p0);
  const float l = length(c);
  const F3 n = c / l;
  shade[t].n[0] = n.x;
  shade[t].n[1] = n.y;
  shade[t].n[2] = n.z;
  shade[t].area = l * 0.5f;
}

// faces of vertex v: list[off[v] .. off[v+1]) in ascending triangle order
__global__ void k_vertex_normals(const TriShade* shade, const uint32_t* off, const uint32_t* list, float* vnrm,
                                 uint32_t V) {
  const uint32_t v = blockIdx.x * kUpdBlock + threadIdx.x;
  if (v >= V) return;
  const uint32_t a = off[v], b = off[v + 1];
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  for (uint32_t i = a; i < b; ++i) {
    const TriShade& s = shade[list[i]];
    sx += s.n[0];
    sy += s.n[1];
    sz += s.n[2];
  }
  if (b == a) {
    vnrm[3 * v] = vnrm[3 * v + 1] = vnrm[3 * v + 2] = 0.0f;
    return;
  }
  const float c = (float)(b - a);
  const F3 n = normalize(F3{sx / c, sy / c, sz / c});
  vnrm[3 * v] = n.x;
  vnrm[3 * v + 1] = n.y;
  vnrm[3 * v + 2] = n.z;
}

// One lane: the emitter table is a handful of triangles.
__global__ void k_light_table(const TriShade* shade, const int32_t* light_tri, const float* emit_rad, float* light_f,
                              float* light_cdf, int32_t NL, UpdateScalars* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  float integral = 0.0f;
  if (NL > 0) {
    for (int32_t i = 0; i < NL; ++i) {
      const TriShade& s = shade[light_tri[i]];
      const float* I = emit_rad + 3 * s.emitter;
      const float lum = 0.212671f * I[0] + 0.715160f * I[1] + 0.072169f * I[2];
      light_f[i] = lum * s.area;
    }
    const float inv_n = 1.0f / (float)NL;
    float run = 0.0f;
    for (int32_t i = 0; i < NL; ++i) {
      run += light_f[i];
      light_cdf[i] = run * inv_n;
    }
    integral = light_cdf[NL - 1];
    if (integral > 0.0f)
      for (int32_t i = 0; i < NL; ++i) light_cdf[i] = light_cdf[i] / integral;
  }
  out->light_integral = integral;
}

// Edge selection in tiles of kEdgeTile edges per block: count, scan of the
// block counts, scatter in edge order.
template <class Pred>
__global__ void k_edge_count(const EdgeDev* edges, uint32_t n, Pred pred, uint32_t* block_cnt) {
  __shared__ uint32_t s[kUpdBlock];
  const uint32_t base = blockIdx.x * kEdgeTile;
  uint32_t c = 0;
  for (int k = 0; k < kEdgeTile / kUpdBlock; ++k) {
    const uint32_t i = base + k * kUpdBlock + threadIdx.x;
    if (i < n && pred(edges[i])) ++c;
  }
  s[threadIdx.x] = c;
  __syncthreads();
  for (int w = kUpdBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) block_cnt[blockIdx.x] = s[0];
}

// Exclusive scan of nb block counts in place (one block of 1024 threads, each
// summing a contiguous run), total into *count.
__global__ void __launch_bounds__(1024) k_block_scan(uint32_t* cnt, uint32_t nb, EdgeCount* count) {
  __shared__ uint32_t s[1024];
  const uint32_t per = (nb + 1023) / 1024;
  const uint32_t a = threadIdx.x * per, b = min(nb, a + per);
  uint32_t sum = 0;
  for (uint32_t i = a; i < b; ++i) sum += cnt[i];
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int w = 1; w < 1024; w <<= 1) {
    const uint32_t v = threadIdx.x >= (uint32_t)w ? s[threadIdx.x - w] : 0u;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  uint32_t run = threadIdx.x > 0 ? s[threadIdx.x - 1] : 0u;
  for (uint32_t i = a; i < b; ++i) {
    const uint32_t c = cnt[i];
    cnt[i] = run;
    run += c;
  }
  if (threadIdx.x == 1023) count->n = (int32_t)s[1023];
}

template <class Pred>
__global__ void k_edge_scatter(const EdgeDev* edges, uint32_t n, Pred pred, const uint32_t* block_off, EdgeDev* out) {
  __shared__ uint32_t wave_cnt[kUpdBlock / 64];
  const uint32_t base = blockIdx.x * kEdgeTile;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t off = block_off[blockIdx.x];
  for (int k = 0; k < kEdgeTile / kUpdBlock; ++k) {
    const uint32_t i = base + k * kUpdBlock + threadIdx.x;
    const bool keep = i < n && pred(edges[i]);
    const unsigned long long m = __ballot(keep);
    const uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_cnt[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (uint32_t w = 0; w < kUpdBlock / 64; ++w) {
      if (w < wave) woff += wave_cnt[w];
      tot += wave_cnt[w];
    }
    if (keep) out[off + woff + below] = edges[i];
    off += tot;
    __syncthreads();
  }
}

__global__ void k_edge_len(const EdgeDev* e, const float* pos, float* len, uint32_t cap, const EdgeCount* count) {
  const uint32_t i = blockIdx.x * kUpdBlock + threadIdx.x;
  if (i >= cap || (int32_t)i >= count->n) return;
  len[i] = length(pos3(pos, e[i].v1) - pos3(pos, e[i].v0));
}

// Inclusive prefix of the lengths in the blocked order of the host code:
// k_cdf_local sums each chunk of kCdfChunk lengths sequentially (one lane
// per chunk) into cdf and the chunk sum into part; k_cdf_offsets turns the
// chunk sums into sequential offsets; k_cdf_final adds the offset and
// scales by 1/n; k_cdf_integral reads the integral (cdf[n-1]).
__global__ void k_cdf_local(const float* __restrict__ len, float* __restrict__ cdf, float* __restrict__ part,
                            uint32_t cap, const EdgeCount* __restrict__ count) {
  const uint32_t c = blockIdx.x * kUpdBlock + threadIdx.x;
  const uint32_t n = (uint32_t)max(count->n, 0);
  const uint32_t a = c * (uint32_t)kCdfChunk;
  if (c * (uint64_t)kCdfChunk >= cap || a >= n) return;
  const uint32_t b = min(n, a + (uint32_t)kCdfChunk);
  float local = 0.0f;
  uint32_t i = a;
  for (; i + 8 <= b; i += 8) {  // loads ahead of the dependent additions
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = len[i + k];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      local = local + v[k];
      cdf[i + k] = local;
    }
  }
  for (; i < b; ++i) {
    local = local + len[i];
    cdf[i] = local;
  }
  part[c] = local;
}

// One lane: the chunk sums (n / kCdfChunk of them) become exclusive offsets.
__global__ void k_cdf_offsets(float* __restrict__ part, const EdgeCount* __restrict__ count) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const uint32_t n = (uint32_t)max(count->n, 0);
  const uint32_t nc = (n + kCdfChunk - 1) / kCdfChunk;
  float off = 0.0f;
  uint32_t c = 0;
  for (; c + 16 <= nc; c += 16) {
    float v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = part[c + k];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      part[c + k] = off;
      off = off + v[k];
    }
  }
  for (; c < nc; ++c) {
    const float acc = part[c];
    part[c] = off;
    off = off + acc;
  }
}

__global__ void k_cdf_final(float* cdf, const float* part, uint32_t cap, const EdgeCount* count) {
  const uint32_t i = blockIdx.x * kUpdBlock + threadIdx.x;
  const int32_t n = count->n;
  if (i >= cap || (int32_t)i >= n) return;
  const float inv = 1.0f / (float)n;
  cdf[i] = (part[i / kCdfChunk] + cdf[i]) * inv;
}

__global__ void k_cdf_integral(const float* cdf, EdgeCount* count) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int32_t n = count->n;
  count->integral = n > 0 ? cdf[n - 1] : 0.0f;
}

__global__ void k_edge_cdf_norm(float* cdf, uint32_t cap, const EdgeCount* count) {
  const uint32_t i = blockIdx.x * kUpdBlock + threadIdx.x;
  const float integral = count->integral;
  if (i >= cap || (int32_t)i >= count->n || !(integral > 0.0f)) return;
  cdf[i] = cdf[i] / integral;
}

// Triangle records in leaf order follow the new positions (the triangle id
// stays in v0[3]).
__global__ void k_tri_records(const float* pos, const uint32_t* idx, TriRecord* tris, uint32_t T) {
  const uint32_t i = blockIdx.x * kUpdBlock + threadIdx.x;
  if (i >= T) return;
  TriRecord& r = tris[i];
  const uint32_t t = __float_as_uint(r.v0[3]);
  for (int k = 0; k < 3; ++k) {
    r.v0[k] = pos[3 * (size_t)idx[3 * t] + k];
    r.v1[k] = pos[3 * (size_t)idx[3 * t + 1] + k];
    r.v2[k] = pos[3 * (size_t)idx[3 * t + 2] + k];
  }
}

// One level of the refit: every node's child boxes from its leaves'
// triangles or its children's (unpadded) boxes of the level below; nbox
// keeps each node's unpadded box for its parent.
__global__ void k_refit_level(Bvh4Node* nodes, float* nbox, const TriRecord* tris, const uint32_t* level_nodes,
                              uint32_t count) {
  const uint32_t j = blockIdx.x * kUpdBlock + threadIdx.x;
  if (j >= count) return;
  const uint32_t ni = level_nodes[j];
  Bvh4Node& n = nodes[ni];
  DBox all;
  all.clear();
  for (int s = 0; s < 4; ++s) {
    const uint32_t c = n.child[s];
    if (c == kEmpty) continue;
    DBox cb;
    cb.clear();
    if (child_is_leaf(c)) {
      for (uint32_t i = leaf_first(c); i < leaf_first(c) + leaf_count(c); ++i) {
        cb.grow(tris[i].v0);
        cb.grow(tris[i].v1);
        cb.grow(tris[i].v2);
      }
    } else {
      const float* b = nbox + 6 * (size_t)c;
      for (int k = 0; k < 3; ++k) {
        cb.lo[k] = b[k];
        cb.hi[k] = b[3 + k];
      }
    }
    n.lox[s] = pad_lo(cb.lo[0]);
    n.loy[s] = pad_lo(cb.lo[1]);
    n.loz[s] = pad_lo(cb.lo[2]);
    n.hix[s] = pad_hi(cb.hi[0]);
    n.hiy[s] = pad_hi(cb.hi[1]);
    n.hiz[s] = pad_hi(cb.hi[2]);
    all.grow(cb);
  }
  float* o = nbox + 6 * (size_t)ni;
  for (int k = 0; k < 3; ++k) {
    o[k] = all.lo[k];
    o[3 + k] = all.hi[k];
  }
}

// SAH cost of the 4-wide tree relative to the root box (the build's
// diagnostic): sum over child slots of area(child) / area(root) times the
// slot's triangle count x kSahIntersect (leaves) or kSahTraverse (inner
// nodes).  Order of the double
// additions is not fixed; the value only drives the rebuild decision.
__global__ void k_sah(const Bvh4Node* nodes, uint32_t N, const float* nbox, UpdateScalars* out) {
  const uint32_t i = blockIdx.x * kUpdBlock + threadIdx.x;
  __shared__ double part[kUpdBlock];
  double cost = 0.0;
  if (i < N) {
    const float* r = nbox;
    const double rx = (double)r[3] - r[0], ry = (double)r[4] - r[1], rz = (double)r[5] - r[2];
    double ra = 2.0 * (rx * ry + ry * rz + rz * rx);
    ra = ra > 1e-30 ? ra : 1e-30;
    const Bvh4Node& n = nodes[i];
    for (int s = 0; s < 4; ++s) {
      const uint32_t c = n.child[s];
      if (c == kEmpty) continue;
      const double dx = (double)n.hix[s] - n.lox[s], dy = (double)n.hiy[s] - n.loy[s],
                   dz = (double)n.hiz[s] - n.loz[s];
      const double a = 2.0 * (dx * dy + dy * dz + dz * dx) / ra;
      cost += a * (child_is_leaf(c) ? (double)leaf_count(c) * kSahIntersect : kSahTraverse);
    }
  }
  part[threadIdx.x] = cost;
  __syncthreads();
  for (int w = kUpdBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&out->sah, part[0]);
}

__global__ void k_zero_sah(UpdateScalars* out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) out->sah = 0.0;
}

inline unsigned blocks_for(uint64_t n) { return (unsigned)((n + kUpdBlock - 1) / kUpdBlock); }

template <class Pred>
hipError_t select_edges(const EdgeDev* all, uint32_t n_all, Pred pred, uint32_t* scratch, float* part,
                        EdgeDev* out_e, float* out_len, float* out_cdf, EdgeCount* count, const float* pos,
                        hipStream_t st) {
  const uint32_t nb = n_all == 0 ? 1 : (n_all + kEdgeTile - 1) / kEdgeTile;
  hipLaunchKernelGGL(k_edge_count<Pred>, dim3(nb), dim3(kUpdBlock), 0, st, all, n_all, pred, scratch);
  hipLaunchKernelGGL(k_block_scan, dim3(1), dim3(1024), 0, st, scratch, nb, count);
  hipLaunchKernelGGL(k_edge_scatter<Pred>, dim3(nb), dim3(kUpdBlock), 0, st, all, n_all, pred,
                     (const uint32_t*)scratch, out_e);
  const uint32_t cap = n_all > 0 ? n_all : 1;
  hipLaunchKernelGGL(k_edge_len, dim3(blocks_for(cap)), dim3(kUpdBlock), 0, st, out_e, pos, out_len, cap,
                     (const EdgeCount*)count);
  const uint32_t nchunks = (cap + kCdfChunk - 1) / kCdfChunk;
  hipLaunchKernelGGL(k_cdf_local, dim3(blocks_for(nchunks)), dim3(kUpdBlock), 0, st, (const float*)out_len, out_cdf,
                     part, cap, (const EdgeCount*)count);
  hipLaunchKernelGGL(k_cdf_offsets, dim3(1), dim3(1), 0, st, part, (const EdgeCount*)count);
  hipLaunchKernelGGL(k_cdf_final, dim3(blocks_for(cap)), dim3(kUpdBlock), 0, st, out_cdf, (const float*)part, cap,
                     (const EdgeCount*)count);
  hipLaunchKernelGGL(k_cdf_integral, dim3(1), dim3(1), 0, st, (const float*)out_cdf, count);
  hipLaunchKernelGGL(k_edge_cdf_norm, dim3(blocks_for(cap)), dim3(kUpdBlock), 0, st, out_cdf, cap,
                     (const EdgeCount*)count);
  return hipGetLastError();
}

}  // namespace

uint32_t edge_scan_words(uint32_t n_edges) { return n_edges == 0 ? 1u : (n_edges + kEdgeTile - 1) / kEdgeTile; }

hipError_t launch_scene_update(const SceneDev& S, const SceneUpdateDev& U, hipStream_t st) {
  TriShade* shade = const_cast<TriShade*>(S.shade);
  const float* pos = S.pos;
  const uint32_t T = (uint32_t)S.n_tris, V = (uint32_t)S.n_verts;
  if (T == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_face_normals, dim3(blocks_for(T)), dim3(kUpdBlock), 0, st, pos, S.idx, shade, T);
  if (U.vs_off)
    hipLaunchKernelGGL(k_vertex_normals, dim3(blocks_for(V)), dim3(kUpdBlock), 0, st, (const TriShade*)shade,
                       U.vs_off, U.vs_list, const_cast<float*>(S.vnrm), V);
  hipLaunchKernelGGL(k_light_table, dim3(1), dim3(1), 0, st, (const TriShade*)shade, S.light_tri, S.emit_rad,
                     const_cast<float*>(S.light_f), const_cast<float*>(S.light_cdf), S.n_light, U.scal);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (U.n_edges > 0) {
    e = select_edges(U.all_edges, U.n_edges, SecPred{pos, shade}, U.scan, U.part, const_cast<EdgeDev*>(S.sec.e),
                     const_cast<float*>(S.sec.len), const_cast<float*>(S.sec.cdf), &U.scal->sec, pos, st);
    if (e != hipSuccess) return e;
  }
  TriRecord* tris = const_cast<TriRecord*>(S.tris);
  Bvh4Node* nodes = const_cast<Bvh4Node*>(S.nodes);
  hipLaunchKernelGGL(k_tri_records, dim3(blocks_for(T)), dim3(kUpdBlock), 0, st, pos, S.idx, tris, T);
  for (size_t l = 0; l + 1 < U.level_off.size(); ++l) {
    const uint32_t a = U.level_off[l], n = U.level_off[l + 1] - a;
    if (n == 0) continue;
    hipLaunchKernelGGL(k_refit_level, dim3(blocks_for(n)), dim3(kUpdBlock), 0, st, nodes, U.nbox,
                       (const TriRecord*)tris, U.refit_nodes + a, n);
  }
  hipLaunchKernelGGL(k_zero_sah, dim3(1), dim3(1), 0, st, U.scal);
  hipLaunchKernelGGL(k_sah, dim3(blocks_for(U.n_nodes)), dim3(kUpdBlock), 0, st, (const Bvh4Node*)nodes, U.n_nodes,
                     (const float*)U.nbox, U.scal);
  return hipGetLastError();
}

hipError_t launch_prim_edges(const SceneDev& S, const SceneUpdateDev& U, const float eye[3], EdgeDev* out_e,
                             float* out_len, float* out_cdf, EdgeCount* count, hipStream_t st) {
  return select_edges(U.all_edges, U.n_edges, PrimPred{S.pos, S.shade, F3{eye[0], eye[1], eye[2]}}, U.scan, U.part,
                      out_e, out_len, out_cdf, count, S.pos, st);
}

}  // namespace trc
