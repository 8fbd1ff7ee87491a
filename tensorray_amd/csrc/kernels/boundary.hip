// Boundary (edge) terms of the differentiable integrator, one sample per
// thread iteration.  Each term has zero value and the gradient
//
//   d/dtheta [ detach(K) * xDotN ],   xDotN = n . u2,
//   u2 = (1-u-v) P0d + u P1d + v P2d, (u, v) = RayIntersectAD(dir, org, P0, P1-P0, P2-P0)
//
// where K carries the detached radiance / importance / baseVal / pdf of the
// sample and (dir, org, P) may all depend on vertex positions.  The weight
// w = <dL/dI[pixel], K> / spp is folded in before the VJP, so each sample
// scatters a handful of vertex-gradient atomics.
//
// Reference: BoundaryPrimary.cpp:36-181, BoundaryPixel.cpp:36-159,
// BoundaryDirect.cpp:38-143, BoundaryIndirect.cpp:38-192, Edge.cpp:90-409,
// Camera.cpp:345-467, EvalImportance (Integrator.cpp:75-176).
#include <hip/hip_runtime.h>

#include "../trc_internal.h"
#include "adjoint.h"

namespace trc {

constexpr uint32_t kSensorBlock0 = 64u;

__device__ inline TraceCtx make_trace_ctx_b(uint32_t* lds, const LaunchCfg& L) {
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

__device__ inline F3 cam_eye(const CameraDev& c) { return F3{c.eye[0], c.eye[1], c.eye[2]}; }

// Calls f(i) for the items i of [first, end).  With a counter (zeroed before
// the launch) a wave takes 64 consecutive items at a time as it finishes the
// last ones, so no wave is left with a long tail of items dealt to it in
// advance; without one, items are dealt grid-stride.  f returns to skip the
// rest of an item.  All lanes of the wave must call it together.
template <class F>
__device__ inline void for_each_item(uint64_t first, uint64_t end, uint32_t* ctr, const LaunchCfg& L, F&& f) {
  if (ctr) {
    const uint32_t lane = threadIdx.x & 63u;
    for (;;) {
      uint32_t c = 0u;
      if (lane == 0u) c = atomicAdd(ctr, 1u);
      c = (uint32_t)__shfl((int)c, 0);
      const uint64_t i0 = first + (uint64_t)c * 64u;
      if (i0 >= end) break;
      if (i0 + lane < end) f(i0 + lane);
    }
  } else {
    const uint64_t stride = (uint64_t)L.grid * kBlock;
    for (uint64_t i = first + (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < end; i += stride) f(i);
  }
}

// Edge sampling with the length distribution; returns the edge index and
// the pdf per unit length.
__device__ inline int32_t sample_edge(const EdgeSet& E, float u, float* pdf) {
  const int32_t n = E.dyn ? E.dyn->n : E.n;
  const float integral = E.dyn ? E.dyn->integral : E.integral;
  int32_t i = cdf_search(E.cdf, n, u);
  float pmf = E.len[i] / (integral * (float)n);
  *pdf = pmf / E.len[i];
  return i;
}

struct EdgeGeo {
  F3 p0, e1, n0, n1, p2;
  bool boundary;
};
__device__ inline EdgeGeo edge_geo(const SceneDev& S, const EdgeDev& e) {
  EdgeGeo g;
  g.p0 = load3(S.pos, e.v0);
  g.e1 = load3(S.pos, e.v1) - g.p0;
  g.p2 = load3(S.pos, e.v2);
  const TriShade& s0 = S.shade[e.f0];
  g.n0 = F3{s0.n[0], s0.n[1], s0.n[2]};
  g.boundary = e.f1 < 0;
  if (g.boundary) {
    g.n1 = F3{0.0f, 0.0f, 0.0f};
  } else {
    const TriShade& s1 = S.shade[e.f1];
    g.n1 = F3{s1.n[0], s1.n[1], s1.n[2]};
  }
  return g;
}

__device__ inline F3 nan0(F3 v) { return finite3(v) ? v : F3{0.0f, 0.0f, 0.0f}; }

// Radiance along a camera boundary ray: EvalRadianceEmitted + bounces with
// throughput = sensor importance (GenerateBoundaryRays), all detached.
__device__ F3 boundary_radiance(const SceneDev& S, TraceCtx& tc, const Its& its, F3 raydir, float sensor_val,
                                uint32_t seed, uint32_t domain, uint64_t index, int B) {
  F3 beta{sensor_val, sensor_val, sensor_val};
  F3 rad{0.0f, 0.0f, 0.0f};
  if (its.emitter >= 0) {
    F3 e = beta * emitted(S, its.emitter, its.ng, -raydir);
    if (finite3(e)) rad = rad + e;
  }
  path_bounces<false, false>(S, tc, its, beta, raydir, seed, domain, index, 1u, B, rad, nullptr, nullptr);
  return rad;
}

// BSDF::Eval with the shading-normal correction (BSDF.cpp:172-195).
__device__ inline F3 bsdf_eval_corrected(const SceneDev& S, const Its& s, F3 wo, F3 wi) {
  F3 f = bsdf_eval(S, s, wo, wi);
  F3 lo = to_local(s, wo), li = to_local(s, wi);
  float corr = fabsf(lo.z * dot(s.ng, wi) / (li.z * dot(s.ng, wo)));
  return f * corr;
}

// One vertex of EvalImportance (Integrator.cpp:75-176): the contribution to
// the pixel seen through the camera and the next (BSDF-sampled) vertex.
struct ImpStep {
  bool contrib;
  int32_t pixel;
  F3 value;
  bool next;
};
__device__ ImpStep importance_step(const SceneDev& S, const CameraDev& C, TraceCtx& tc, Its& its, F3& rdir, F3& thr,
                                   Rnd4 rnd) {
  ImpStep r;
  r.contrib = false;
  r.next = false;
  r.pixel = 0;
  r.value = F3{0.0f, 0.0f, 0.0f};
  const F3 eye = cam_eye(C);
  F3 wo = -rdir;
  SensorSample sd = sensor_sample(C, its.p);
  if (sd.valid) {
    F3 cd = camera_dir(C, sd.qx * (float)C.width, sd.qy * (float)C.height);
    float tmax = length(its.p - eye) - kShadowEpsilon;
    bool vis = trace_ray<true>(S, tc, eye, cd, kShadowEpsilon, tmax).tri < 0;
    if (vis) {
      F3 t = sd.value * thr;
      F3 bv{0.0f, 0.0f, 0.0f};
      if (bsdf_dir_valid(S, its, wo)) {
        F3 wi = normalize(eye - its.p);
        float wiok = bsdf_dir_valid(S, its, wi) ? 1.0f : 0.0f;
        bv = bsdf_eval_corrected(S, its, wo, wi) * fabsf(dot(its.ns, wi)) * wiok;
      }
      F3 v = t * bv;
      if (finite3(v)) {
        r.contrib = true;
        r.pixel = sd.pixel;
        r.value = v;
      }
    }
  }
  if (bsdf_dir_valid(S, its, wo)) {
    float pdf;
    F3 wi = bsdf_sample(S, its, wo, rnd.a, rnd.b, &pdf);
    F3 bv = bsdf_eval_corrected(S, its, wo, wi) * fabsf(dot(its.ns, wi));
    Hit h = trace_ray<false>(S, tc, its.p, wi, kShadowEpsilon, 1e32f);
    if (h.tri >= 0 && pdf > 0.0f && bsdf_dir_valid(S, its, wi)) {
      r.next = true;
      thr = thr * bv / pdf;
      its = post_intersect(S, h.tri, h.u, h.v);
      rdir = wi;
    }
  }
  return r;
}

// baseVal of the camera-side terms (BoundaryPrimary.cpp:117-129,
// BoundaryPixel.cpp:95-107; the pixel term reverts the sign).
__device__ inline float base_val_camera(F3 itp, F3 ng, F3 raydir, F3 p_edge, F3 eye, F3 edge, F3 edge2, F3* n_out,
                                        bool revert) {
  float dist = length(itp - eye);
  float dist1 = length(p_edge - eye);
  float cos2 = fabsf(dot(ng, -raydir));
  F3 e = cross(edge, raydir);
  float sinphi = length(e);
  F3 proj = normalize(cross(e, ng));
  float sinphi2 = length(cross(raydir, proj));
  F3 n = normalize(cross(ng, proj));
  bool sign0 = dot(e, edge2) > 0.0f, sign1 = dot(e, n) > 0.0f;
  float bv = (dist / dist1) * (sinphi / sinphi2) * cos2;
  bv = bv * (sinphi > kEpsilon ? 1.0f : 0.0f) * (sinphi2 > kEpsilon ? 1.0f : 0.0f);
  float sg = (sign0 == sign1) ? 1.0f : -1.0f;
  *n_out = n;
  return bv * (revert ? -sg : sg);
}

// baseVal of the secondary terms (Edge.cpp:245-257, BoundaryIndirect.cpp:100-112).
__device__ inline float base_val_secondary(F3 its_p, F3 far_p, F3 far_n, F3 dir_to_near, F3 p_edge, F3 edge, F3 edge2,
                                           F3* n_out) {
  float dist = length(far_p - its_p);
  float cos2 = fabsf(dot(far_n, dir_to_near));
  F3 e = cross(edge, -dir_to_near);
  float sinphi = length(e);
  F3 proj = normalize(cross(e, far_n));
  float sinphi2 = length(cross(-dir_to_near, proj));
  float itsT = length(its_p - p_edge);
  F3 n = normalize(cross(far_n, proj));
  bool sign0 = dot(e, edge2) > 0.0f, sign1 = dot(e, n) > 0.0f;
  float bv = (itsT / dist) * (sinphi / sinphi2) * cos2;
  bv = bv * (sinphi > kEpsilon ? 1.0f : 0.0f) * (sinphi2 > kEpsilon ? 1.0f : 0.0f);
  *n_out = n;
  return bv * ((sign0 == sign1) ? 1.0f : -1.0f);
}

// VJP of w * xDotN with respect to the triangle vertices P0..P2 (of the
// RayIntersectAD), the ray origin and the point the ray passes through.
// xDotN = n . ((1-u-v) P0d + u P1d + v P2d).
struct XdnAdj {
  F3 P0, P1, P2, org, through;
};
__device__ inline XdnAdj x_dot_n_vjp(F3 P0, F3 P1, F3 P2, F3 org, F3 through, F3 n, float w) {
  XdnAdj a;
  F3 dr = through - org;
  float dl = length(dr);
  F3 dir = dr / dl;
  F3 e1 = P1 - P0, e2 = P2 - P0;
  F3 h = cross(dir, e2);
  float aa = dot(e1, h);
  float f = 1.0f / aa;
  F3 s0 = org - P0;
  float su = dot(s0, h);
  F3 q = cross(s0, e1);
  float sv = dot(dir, q);
  float adj_u = w * dot(n, P1 - P0);
  float adj_v = w * dot(n, P2 - P0);
  float adj_f = adj_u * su + adj_v * sv;
  float adj_su = adj_u * f, adj_sv = adj_v * f;
  float adj_a = -adj_f * f * f;
  F3 adj_s0 = adj_su * h;
  F3 adj_h = adj_su * s0 + adj_a * e1;
  F3 adj_dir = adj_sv * q;
  F3 adj_q = adj_sv * dir;
  adj_s0 += cross(e1, adj_q);
  F3 adj_e1 = cross(adj_q, s0) + adj_a * h;
  adj_dir += cross(e2, adj_h);
  F3 adj_e2 = cross(adj_h, dir);
  a.org = adj_s0;
  a.P0 = -(adj_s0 + adj_e1 + adj_e2);
  a.P1 = adj_e1;
  a.P2 = adj_e2;
  F3 adj_dr = normalize_vjp(dir, dl, adj_dir);
  a.through = adj_dr;
  a.org -= adj_dr;
  return a;
}

template <class Sink>
__device__ inline void vtx_add(const SceneDev& S, uint32_t v, int32_t tri, F3 g, Sink& sk) {
  if (!(S.shade[tri].flags & kTriDiff)) return;
  if (!finite3(g)) return;
  sk.pos(v, g);
}

__device__ inline F3 dLdI_at(const float* dLdI, uint32_t WH, int32_t pix) {
  return F3{dLdI[pix], dLdI[WH + pix], dLdI[2 * WH + pix]};
}

// Scatter of one sample: triangle of the RayIntersectAD (tri), edge point
// (through = q0 + t (q1 - q0) when the edge is given) and material-form
// origin (otri, ou, ov) when given.  GradSink: gradients of w * xDotN;
// TangentSink with w = 1: the derivative of xDotN along the tangents.
template <class Sink>
__device__ void scatter_xdn(const SceneDev& S, int32_t tri, F3 org, F3 through, F3 n, float w, const EdgeDev* edge,
                            float et, int32_t ef0, int32_t otri, float ou, float ov, Sink& sk) {
  F3 P0, P1, P2;
  tri_verts(S, tri, P0, P1, P2);
  XdnAdj a = x_dot_n_vjp(P0, P1, P2, org, through, n, w);
  vtx_add(S, S.idx[3 * tri], tri, a.P0, sk);
  vtx_add(S, S.idx[3 * tri + 1], tri, a.P1, sk);
  vtx_add(S, S.idx[3 * tri + 2], tri, a.P2, sk);
  if (edge) {
    vtx_add(S, edge->v0, ef0, (1.0f - et) * a.through, sk);
    vtx_add(S, edge->v1, ef0, et * a.through, sk);
  }
  if (otri >= 0) {
    float ow = 1.0f - ou - ov;
    vtx_add(S, S.idx[3 * otri], otri, ow * a.org, sk);
    vtx_add(S, S.idx[3 * otri + 1], otri, ou * a.org, sk);
    vtx_add(S, S.idx[3 * otri + 2], otri, ov * a.org, sk);
  }
}

// Derivative-image contribution v at a pixel (non-finite values dropped).
__device__ inline void add_deriv(float* dimg, uint32_t WH, int32_t pix, F3 v) {
  if (!finite3(v)) return;
  if (v.x != 0.0f) atomicAdd(&dimg[pix], v.x);
  if (v.y != 0.0f) atomicAdd(&dimg[WH + pix], v.y);
  if (v.z != 0.0f) atomicAdd(&dimg[2 * WH + pix], v.z);
}

struct BoundaryParams {
  uint32_t seed;
  int32_t max_bounces;
  int32_t spp;
  uint32_t pad;
};

// ---------------------------------------------------------------------------
// Survivor queues.  The primary, direct and indirect terms run in two
// kernels per chunk of samples: a select kernel draws the samples, traces
// their first rays and keeps the few that can contribute; an eval kernel
// traces the long radiance / importance paths of those only, so its waves
// start full instead of carrying the lanes whose samples died at the first
// test.  The queue is structure-of-arrays (word k of slot s at
// w[k * cap + s]) and filled by wave-level ballot compaction: one atomic
// per wave and consecutive slots for the surviving lanes.  Chunks are large
// (4M samples, one chunk for a 512^2 x 16 request): every chunk ends in two
// kernel tails, and at 1M samples per chunk those cost more than the
// compaction saved.  Survivors of the 512^2 metric scene: 10 % of the
// primary, 0.8 % of the direct and 13 % of the indirect samples.
// ---------------------------------------------------------------------------
struct BndQueue {
  uint32_t* count;  // survivors of this chunk
  uint32_t* w;      // kBndWords x cap words
  uint32_t cap;
};
constexpr int kBndWords = 20;
constexpr uint64_t kBndChunk = 1ull << 22;  // 335 MB of queue at most

// Slot of the calling lane in the queue when `keep`, else -1.  Every active
// lane of the wave calls it at the same point.
__device__ inline int64_t wave_append(bool keep, uint32_t* count) {
  const uint64_t act = __ballot(1);
  const uint64_t m = __ballot(keep);
  if (m == 0ull) return -1;
  const int lane = (int)(threadIdx.x & 63u);
  const int leader = __ffsll((unsigned long long)act) - 1;
  uint32_t base = 0u;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = (uint32_t)__shfl((int)base, leader);
  if (!keep) return -1;
  return (int64_t)base + (int64_t)__popcll(m & ((1ull << lane) - 1ull));
}

struct QWriter {
  const BndQueue& q;
  uint32_t s;
  int k = 0;
  __device__ void f(float x) { q.w[(size_t)(k++) * q.cap + s] = __float_as_uint(x); }
  __device__ void u(uint32_t x) { q.w[(size_t)(k++) * q.cap + s] = x; }
  __device__ void i(int32_t x) { u((uint32_t)x); }
  __device__ void f3(F3 v) { f(v.x); f(v.y); f(v.z); }
  __device__ void u64(uint64_t x) { u((uint32_t)x); u((uint32_t)(x >> 32)); }
};
struct QReader {
  const BndQueue& q;
  uint32_t s;
  int k = 0;
  __device__ float f() { return __uint_as_float(q.w[(size_t)(k++) * q.cap + s]); }
  __device__ uint32_t u() { return q.w[(size_t)(k++) * q.cap + s]; }
  __device__ int32_t i() { return (int32_t)u(); }
  __device__ F3 f3() { F3 v; v.x = f(); v.y = f(); v.z = f(); return v; }
  __device__ uint64_t u64() { uint64_t lo = u(); uint64_t hi = u(); return lo | (hi << 32); }
};

// ---------------------------------------------------------------------------
// Primary edges (BoundaryPrimary.cpp:36-181)
// ---------------------------------------------------------------------------
// select: edge point, the ray away from the camera, its projection onto the
// sensor, the edge point's visibility from the camera and baseVal.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kBndWaves))) k_bnd_primary_select(SceneDev S, CameraDev C, BoundaryParams P, LaunchCfg L,
                                                               EdgeSet E, uint64_t first, uint64_t end, BndQueue Q) {
  __shared__ uint32_t lds_stack[kTraceLdsWords * kBlock];
  TraceCtx tc = make_trace_ctx_b(lds_stack, L);
  const F3 eye = cam_eye(C);
  // no silhouette or boundary edge seen from this camera: no samples
  if ((E.dyn ? E.dyn->n : E.n) <= 0) return;
  for_each_item(first, end, L.tiles, L, [&](uint64_t i) {
    bool keep = false;
    Hit h;
    SensorSample sd;
    F3 p, rd, n;
    float pdf, bv, rb;
    int32_t ei;
    do {
      Rnd4 rnd = rng4(P.seed, kRngPrimaryEdge, i, 0u);
      ei = sample_edge(E, rnd.a, &pdf);
      const EdgeDev ed = E.e[ei];
      EdgeGeo g = edge_geo(S, ed);
      rb = rnd.b;
      p = g.p0 + rnd.b * g.e1;
      F3 edge = normalize(g.e1);
      F3 edge2 = normalize(g.p2 - g.p0);
      F3 dir = normalize(p - eye);
      h = trace_ray<false>(S, tc, p, dir, kShadowEpsilon, 1e32f);
      if (h.tri < 0) break;
      F3 hitP = p + h.t * dir;
      sd = sensor_sample(C, hitP);
      if (!sd.valid) break;
      rd = camera_dir(C, sd.qx * (float)C.width, sd.qy * (float)C.height);
      float tmax = length(p - eye) - kShadowEpsilon;
      if (trace_ray<true>(S, tc, eye, rd, kShadowEpsilon, tmax).tri >= 0) break;
      Its its = post_intersect(S, h.tri, h.u, h.v);
      bv = base_val_camera(its.p, its.ng, rd, p, eye, edge, edge2, &n, false);
      // K = rad * bv / pdf vanishes or is not finite for every radiance
      keep = bv != 0.0f && isfinite(bv);
    } while (false);
    const int64_t s = wave_append(keep, Q.count);
    if (s >= 0) {
      QWriter w{Q, (uint32_t)s};
      w.u64(i); w.i(h.tri); w.f(h.u); w.f(h.v); w.f3(rd); w.f(sd.value); w.i(sd.pixel);
      w.f(bv); w.f(pdf); w.f3(n); w.i(ei); w.f(rb); w.f3(p);
    }
  });
}

// eval: the radiance seen through the sensor point and the scatter.
// kTangent: the derivative image along tpos goes to dimg instead.
template <bool kTangent>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kBndWaves))) k_bnd_primary_eval(SceneDev S, CameraDev C, BoundaryParams P, LaunchCfg L,
                                                             EdgeSet E, BndQueue Q, const float* dLdI, float* dpos,
                                                             const float* tpos, float* dimg) {
  __shared__ uint32_t lds_stack[kTraceLdsWords * kBlock];
  TraceCtx tc = make_trace_ctx_b(lds_stack, L);
  const uint32_t WH = (uint32_t)C.width * (uint32_t)C.height;
  const F3 eye = cam_eye(C);
  const uint32_t n_q = *Q.count;
  const uint32_t stride = (uint32_t)L.grid * kBlock;
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < n_q; s += stride) {
    QReader r{Q, s};
    const uint64_t i = r.u64();
    const int32_t tri = r.i();
    const float u = r.f(), v = r.f();
    const F3 rd = r.f3();
    const float value = r.f();
    const int32_t pixel = r.i();
    const float bv = r.f(), pdf = r.f();
    const F3 n = r.f3();
    const int32_t ei = r.i();
    const float rb = r.f();
    const F3 p = r.f3();
    Its its = post_intersect(S, tri, u, v);
    F3 rad = boundary_radiance(S, tc, its, rd, value, P.seed, kRngPrimaryEdge, i, P.max_bounces);
    F3 K = rad * bv / pdf;
    if (!(length(K) > 0.0f) || !finite3(K)) continue;
    const EdgeDev ed = E.e[ei];
    if constexpr (kTangent) {
      TangentSink sk{tpos, nullptr, nullptr, 0.0f};
      scatter_xdn(S, tri, eye, p, n, 1.0f, &ed, rb, ed.f0, -1, 0.0f, 0.0f, sk);
      add_deriv(dimg, WH, pixel, K * (sk.acc / (float)P.spp));
    } else {
      float w = dot(dLdI_at(dLdI, WH, pixel), K) / (float)P.spp;
      if (w == 0.0f) continue;
      GradSink sk{dpos, nullptr, nullptr, nullptr};
      scatter_xdn(S, tri, eye, p, n, w, &ed, rb, ed.f0, -1, 0.0f, 0.0f, sk);
    }
  }
}

// ---------------------------------------------------------------------------
// Pixel boundary of the box filter (BoundaryPixel.cpp:36-159, Edge.cpp:377-409,
// Camera.cpp:433-467).  4 antithetic positions per random number.
// ---------------------------------------------------------------------------
template <bool kTangent>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kBndWaves))) k_boundary_pixel(SceneDev S, CameraDev C, BoundaryParams P, LaunchCfg L,
                                                           uint64_t N, const float* dLdI, float* dpos,
                                                           const float* tpos, float* dimg) {
  __shared__ uint32_t lds_stack[kTraceLdsWords * kBlock];
  TraceCtx tc = make_trace_ctx_b(lds_stack, L);
  const uint32_t WH = (uint32_t)C.width * (uint32_t)C.height;
  const F3 eye = cam_eye(C);
  const F3 R{C.right[0], C.right[1], C.right[2]}, U{C.up[0], C.up[1], C.up[2]}, D{C.dir[0], C.dir[1], C.dir[2]};
  const float plx = 2.0f * C.ratio * C.tan_half, ply = 2.0f * C.tan_half;
  const F3 vdir = normalize(C.ay * U), hdir = normalize(C.ax * R);
  const uint64_t per_anti = N / 4;
  for_each_item(0, N, L.tiles, L, [&](uint64_t i) {
    const uint64_t j = i % per_anti, anti = i / per_anti;
    float rr = rng4(P.seed, kRngPixelEdge, j, 0u).a;
    float rv;
    if (anti == 0) rv = rr;
    else if (anti == 1) rv = rr < 0.5f ? rr + 0.5f : rr - 0.5f;
    else if (anti == 2) rv = rr < 0.75f ? 0.75f - rr : 1.75f - rr;
    else rv = rr < 0.25f ? 0.25f - rr : 1.25f - rr;
    float r4 = rv * 4.0f;
    int side = (int)floorf(r4);
    float ox, oy;
    if (side == 0) { ox = 0.0f; oy = r4; }
    else if (side == 1) { ox = r4 - 1.0f; oy = 1.0f; }
    else if (side == 2) { ox = 1.0f; oy = 3.0f - r4; }
    else { ox = 4.0f - r4; oy = 0.0f; }
    const uint32_t pix = (uint32_t)(i % WH);
    float rx = (float)(pix % (uint32_t)C.width) + ox, ry = (float)(pix / (uint32_t)C.width) + oy;
    F3 p0 = ((eye + (C.ax * rx + C.bx) * R) + (C.ay * ry + C.by) * U) + D;
    F3 edge = (side == 0 || side == 2) ? vdir : hdir;
    F3 edge2 = side == 0 ? hdir : (side == 1 ? -vdir : (side == 2 ? -hdir : vdir));
    float pdf = (side == 0 || side == 2) ? 0.25f / (ply * (float)C.width) : 0.25f / (plx * (float)C.height);
    F3 dir = normalize(p0 - eye);
    Hit h = trace_ray<false>(S, tc, p0, dir, kShadowEpsilon, 1e32f);
    if (h.tri < 0) return;
    F3 hitP = p0 + h.t * dir;
    SensorSample sd = sensor_sample(C, hitP);
    F3 rd = camera_dir(C, sd.qx * (float)C.width, sd.qy * (float)C.height);
    Its its = post_intersect(S, h.tri, h.u, h.v);
    F3 n;
    float bv = base_val_camera(its.p, its.ng, rd, p0, eye, edge, edge2, &n, true);
    F3 rad = boundary_radiance(S, tc, its, rd, sd.value, P.seed, kRngPixelEdge, j, P.max_bounces);
    F3 K = rad * bv / pdf;
    if (!finite3(K)) return;
    if constexpr (kTangent) {
      TangentSink sk{tpos, nullptr, nullptr, 0.0f};
      scatter_xdn(S, h.tri, eye, p0, n, 1.0f, nullptr, 0.0f, -1, -1, 0.0f, 0.0f, sk);
      add_deriv(dimg, WH, (int32_t)pix, K * (sk.acc / (float)P.spp));
    } else {
      float w = dot(dLdI_at(dLdI, WH, (int32_t)pix), K) / (float)P.spp;
      if (w == 0.0f) return;
      GradSink sk{dpos, nullptr, nullptr, nullptr};
      scatter_xdn(S, h.tri, eye, p0, n, w, nullptr, 0.0f, -1, -1, 0.0f, 0.0f, sk);
    }
  });
}

// ---------------------------------------------------------------------------
// Secondary edges, direct illumination (BoundaryDirect.cpp:38-143,
// Edge.cpp:108-144 and 201-309)
// ---------------------------------------------------------------------------
// Guide evaluation (AdaptiveQuadratureDistribution::setup and
// GuideDirectBoundary_Helper, AQ_distrb.h:63-76 and 100-175): sample k of
// leaf i draws x uniformly in the leaf's CDF interval and (y, z) in [0,1]^2
// and records |baseVal| / spp of the guiding-mode evaluation, where only
// the two hit tests and the same-point test of Edge.cpp:201-262 apply.
// Waves take samples 64 at a time from the launch's counter (2M-triangle
// sphere, 16 samples per leaf: 26.7 ms dealt grid-stride, 21.6 ms so).
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kBndWaves))) k_guide_eval(SceneDev S, LaunchCfg L, uint32_t seed, GuideDev G) {
  __shared__ uint32_t lds_stack[kTraceLdsWords * kBlock];
  TraceCtx tc = make_trace_ctx_b(lds_stack, L);
  const EdgeSet& E = S.sec;
  const uint64_t N = (uint64_t)G.n * (uint64_t)G.spp;
  for_each_item(0, N, L.tiles, L, [&](uint64_t j) {
    const int32_t leaf = (int32_t)(j / (uint64_t)G.spp);
    Rnd4 r = rng4(seed, kRngGuide, j, 0u);
    const float x0 = leaf > 0 ? E.cdf[leaf - 1] : 0.0f, x1 = E.cdf[leaf];
    const float x = x0 + r.a * (x1 - x0);
    float pdf_e;
    int32_t ei = sample_edge(E, x, &pdf_e);
    const EdgeDev ed = E.e[ei];
    EdgeGeo g = edge_geo(S, ed);
    float t = cdf_reuse(E.cdf, ei, x);
    F3 p = g.p0 + t * g.e1;
    F3 edge = normalize(g.e1);
    F3 edge2 = normalize(g.p2 - g.p0);
    LightSample ls = light_sample(S, r.b, r.c);
    F3 rdir = normalize(p - ls.p);
    float v = 0.0f;
    Hit h0 = trace_ray<false>(S, tc, p, -rdir, kShadowEpsilon, 1e32f);
    if (h0.tri >= 0) {
      Hit h = trace_ray<false>(S, tc, p, rdir, kShadowEpsilon, 1e32f);
      F3 hitP = p + h0.t * (-rdir);
      if (h.tri >= 0 && length(hitP - ls.p) < kShadowEpsilon) {
        Its its = post_intersect(S, h.tri, h.u, h.v);
        F3 n;
        v = fabsf(base_val_secondary(its.p, ls.p, ls.n, rdir, p, edge, edge2, &n));
        if (!isfinite(v)) v = 0.0f;
      }
    }
    G.vals[j] = v / (float)G.spp;
  });
}

// Guide normalization (AQ_distrb.h:166-180): poly_i = sum_k vals in sample
// order, w = poly / sum(poly), poly = eps + (1 - eps) w, f = poly * volume,
// then the Distribution1D CDF of f (Distribution.cpp:45-60).  The sums over
// leaves use the blocked order of kSumChunk, so the result does not depend on
// the launch shape and every pass is parallel.
__global__ void k_guide_poly(GuideDev G) {
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < G.n; i += stride) {
    float acc = 0.0f;
    const float* v = G.vals + (size_t)i * (size_t)G.spp;
    for (int k = 0; k < G.spp; ++k) acc = acc + v[k];
    G.poly[i] = acc;
  }
}

// part[c] = sequential sum of chunk c of v[0, n).
__global__ void k_chunk_sums(const float* v, int n, float* part) {
  const int nc = (n + kSumChunk - 1) / kSumChunk;
  const int stride = gridDim.x * blockDim.x;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += stride) {
    const int lo = c * kSumChunk, hi = min(n, lo + kSumChunk);
    float acc = 0.0f;
    for (int i = lo; i < hi; ++i) acc = acc + v[i];
    part[c] = acc;
  }
}

// One thread: part[c] becomes the exclusive prefix of the chunk sums,
// part[nc] their total.
__global__ void k_chunk_scan(float* part, int nc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  float run = 0.0f;
  for (int c = 0; c < nc; ++c) {
    const float s = part[c];
    part[c] = run;
    run = run + s;
  }
  part[nc] = run;
}

// f_i = (eps + (1 - eps) poly_i / total) * (leaf's x extent), stored in cdf.
__global__ void k_guide_weights(const float* sec_cdf, GuideDev G, const float* total) {
  const float sum = *total;
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < G.n; i += stride) {
    const float w = sum > 0.0f ? G.poly[i] / sum : 0.0f;
    const float poly = G.eps + (1.0f - G.eps) * w;
    const float x0 = i > 0 ? sec_cdf[i - 1] : 0.0f;
    G.poly[i] = poly;
    G.cdf[i] = poly * ((sec_cdf[i] - x0) * 1.0f * 1.0f);
  }
}

// cdf_i = (blocked inclusive prefix of f) / n, from the chunk offsets.
__global__ void k_chunk_prefix(float* v, int n, const float* part, float inv) {
  const int nc = (n + kSumChunk - 1) / kSumChunk;
  const int stride = gridDim.x * blockDim.x;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += stride) {
    const int lo = c * kSumChunk, hi = min(n, lo + kSumChunk);
    float local = 0.0f;
    const float off = part[c];
    for (int i = lo; i < hi; ++i) {
      local = local + v[i];
      v[i] = (off + local) * inv;
    }
  }
}

// fsum = total of f; cdf_i /= cdf_{n-1} (the integral) when it is positive.
__global__ void k_guide_normalize(GuideDev G, const float* total) {
  const float integral = G.cdf[G.n - 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) G.fsum[0] = *total;
  if (!(integral > 0.0f)) return;
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < G.n - 1; i += stride) G.cdf[i] = G.cdf[i] / integral;
}

// The last element is divided after every other thread has read it.
__global__ void k_guide_normalize_last(GuideDev G) {
  const float integral = G.cdf[G.n - 1];
  if (integral > 0.0f) G.cdf[G.n - 1] = integral / integral;
}

// select: edge point and light point, the two rays from the edge point,
// the emitted radiance and baseVal.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kBndWaves))) k_bnd_direct_select(SceneDev S, CameraDev C, BoundaryParams P, LaunchCfg L,
                                                              GuideDev G, uint64_t first, uint64_t end, BndQueue Q) {
  __shared__ uint32_t lds_stack[kTraceLdsWords * kBlock];
  TraceCtx tc = make_trace_ctx_b(lds_stack, L);
  const EdgeSet& E = S.sec;
  for_each_item(first, end, L.tiles, L, [&](uint64_t i) {
    bool keep = false;
    Hit h0, h;
    F3 K0, rdir, p, n;
    float t;
    int32_t ei;
    do {
      Rnd4 rnd = rng4(P.seed, kRngDirectEdge, i, 0u);
      float pdf_b = 1.0f;
      if (G.enabled) {
        // AdaptiveQuadratureDistribution::sample (AQ_distrb.h:182-197)
        const int32_t leaf = cdf_search(G.cdf, G.n, rnd.d);
        const float x0 = leaf > 0 ? E.cdf[leaf - 1] : 0.0f, x1 = E.cdf[leaf];
        rnd.a = x0 + rnd.a * (x1 - x0);
        pdf_b = G.poly[leaf] / G.fsum[0];
      }
      float pdf_e;
      ei = sample_edge(E, rnd.a, &pdf_e);
      const EdgeDev ed = E.e[ei];
      EdgeGeo g = edge_geo(S, ed);
      t = cdf_reuse(E.cdf, ei, rnd.a);
      p = g.p0 + t * g.e1;
      F3 edge = normalize(g.e1);
      F3 edge2 = normalize(g.p2 - g.p0);
      LightSample ls = light_sample(S, rnd.b, rnd.c);
      F3 ev = ls.p - p;
      float d2 = len2(ev), dist = sqrtf(d2);
      F3 en = ev / dist;
      float cosT = dot(ls.n, -en);
      float c0 = dot(en, g.n0), c1 = dot(en, g.n1);
      bool valid0 = fabsf(c0) > kEdgeEpsilon;
      bool valid1 = (c0 > kEdgeEpsilon && c1 < -kEdgeEpsilon) || (c0 < -kEdgeEpsilon && c1 > kEdgeEpsilon);
      bool right = (g.boundary && valid0) || (!g.boundary && valid1);
      if (!(right && cosT > kEpsilon && dist > kShadowEpsilon)) break;
      float pdf = pdf_e * ls.pdf * d2 / cosT * pdf_b;
      rdir = normalize(p - ls.p);
      h0 = trace_ray<false>(S, tc, p, -rdir, kShadowEpsilon, 1e32f);
      if (h0.tri < 0) break;
      h = trace_ray<false>(S, tc, p, rdir, kShadowEpsilon, 1e32f);
      if (h.tri < 0) break;
      F3 hitP = p + h0.t * (-rdir);
      if (!(length(hitP - ls.p) < kShadowEpsilon)) break;
      F3 Le{0.0f, 0.0f, 0.0f};
      const TriShade& sh0 = S.shade[h0.tri];
      if (sh0.emitter >= 0) Le = emitted(S, sh0.emitter, F3{sh0.n[0], sh0.n[1], sh0.n[2]}, rdir);
      Its its = post_intersect(S, h.tri, h.u, h.v);
      float bv = base_val_secondary(its.p, ls.p, ls.n, rdir, p, edge, edge2, &n);
      K0 = Le * bv / pdf;
      // every contribution K0 * importance then vanishes or is not finite
      keep = finite3(K0) && (K0.x != 0.0f || K0.y != 0.0f || K0.z != 0.0f);
    } while (false);
    const int64_t s = wave_append(keep, Q.count);
    if (s >= 0) {
      QWriter w{Q, (uint32_t)s};
      w.u64(i); w.f3(K0); w.i(h.tri); w.f(h.u); w.f(h.v); w.f3(rdir); w.i(h0.tri); w.f3(p); w.f3(n); w.i(ei);
      w.f(t);
    }
  });
}

// eval: EvalImportance from the sensor-side hit for B bounces and the scatter.
// kTangent: the derivative image along tpos goes to dimg instead.
template <bool kTangent>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kBndWaves))) k_bnd_direct_eval(SceneDev S, CameraDev C, BoundaryParams P, LaunchCfg L,
                                                            BndQueue Q, const float* dLdI, float* dpos,
                                                            const float* tpos, float* dimg) {
  __shared__ uint32_t lds_stack[kTraceLdsWords * kBlock];
  TraceCtx tc = make_trace_ctx_b(lds_stack, L);
  const uint32_t WH = (uint32_t)C.width * (uint32_t)C.height;
  const EdgeSet& E = S.sec;
  const uint32_t n_q = *Q.count;
  const uint32_t stride = (uint32_t)L.grid * kBlock;
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < n_q; s += stride) {
    QReader r{Q, s};
    const uint64_t i = r.u64();
    const F3 K0 = r.f3();
    const int32_t tri = r.i();
    const float hu = r.f(), hv = r.f();
    const F3 rdir = r.f3();
    const int32_t tri0 = r.i();
    const F3 p = r.f3();
    const F3 n = r.f3();
    const int32_t ei = r.i();
    const float t = r.f();
    Its its = post_intersect(S, tri, hu, hv);
    const EdgeDev ed = E.e[ei];
    float ds = 0.0f;  // tangent mode: derivative of xDotN
    if constexpr (kTangent) {
      TangentSink sk{tpos, nullptr, nullptr, 0.0f};
      scatter_xdn(S, tri0, its.p, p, n, 1.0f, &ed, t, ed.f0, tri, hu, hv, sk);
      ds = sk.acc / (float)P.spp;
      if (ds == 0.0f || !isfinite(ds)) continue;
    }
    // sensor side: EvalImportance for B bounces, throughput 1
    float w = 0.0f;
    Its cur = its;
    F3 cdir = rdir, thr{1.0f, 1.0f, 1.0f};
    for (int b = 0; b < P.max_bounces; ++b) {
      ImpStep st = importance_step(S, C, tc, cur, cdir, thr, rng4(P.seed, kRngDirectEdge, i, kSensorBlock0 + (uint32_t)b));
      if (st.contrib) {
        F3 v = K0 * st.value;
        if (length(v) > 0.0f) {
          if constexpr (kTangent) add_deriv(dimg, WH, st.pixel, v * ds);
          else w += dot(dLdI_at(dLdI, WH, st.pixel), v);
        }
      }
      if (!st.next) break;
    }
    if constexpr (!kTangent) {
      w = w / (float)P.spp;
      if (w == 0.0f || !isfinite(w)) continue;
      GradSink sk{dpos, nullptr, nullptr, nullptr};
      scatter_xdn(S, tri0, its.p, p, n, w, &ed, t, ed.f0, tri, hu, hv, sk);
    }
  }
}

// ---------------------------------------------------------------------------
// Secondary edges, indirect illumination (BoundaryIndirect.cpp:38-192,
// Edge.cpp:146-199)
// ---------------------------------------------------------------------------
// select: edge point and direction, the two rays from the edge point and
// baseVal / pdf.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kBndWaves))) k_bnd_indirect_select(SceneDev S, CameraDev C, BoundaryParams P, LaunchCfg L,
                                                                uint64_t first, uint64_t end, BndQueue Q) {
  __shared__ uint32_t lds_stack[kTraceLdsWords * kBlock];
  TraceCtx tc = make_trace_ctx_b(lds_stack, L);
  const EdgeSet& E = S.sec;
  for_each_item(first, end, L.tiles, L, [&](uint64_t i) {
    bool keep = false;
    Hit he, hs;
    F3 dir, p, n;
    float v0, t;
    int32_t ei;
    do {
      Rnd4 rnd = rng4(P.seed, kRngIndirectEdge, i, 0u);
      float pdf_e;
      ei = sample_edge(E, rnd.a, &pdf_e);
      const EdgeDev ed = E.e[ei];
      EdgeGeo g = edge_geo(S, ed);
      t = cdf_reuse(E.cdf, ei, rnd.a);
      p = g.p0 + t * g.e1;
      F3 edge = normalize(g.e1);
      F3 edge2 = normalize(g.p2 - g.p0);
      float pdf_d;
      if (g.boundary) {
        F3 nrm = cross(g.n0, edge);
        F3 ft, fb;
        coordinate_system(nrm, &ft, &fb);
        float z = minf(1.0f - 2.0f * rnd.b, 1.0f - kEdgeEpsilon);
        float rr = sqrtf(maxf(0.0f, 1.0f - z * z));
        float sp, cp;
        tr_sincos(2.0f * kPi * rnd.c, &sp, &cp);
        dir = ((rr * cp) * ft + (rr * sp) * fb) + z * nrm;
        pdf_d = 0.25f / kPi;
      } else {
        float phi0 = tr_acos(dot(g.n0, g.n1));
        pdf_d = 0.25f / phi0;
        F3 z = normalize(g.n0 + g.n1);
        F3 y = normalize(cross(g.n0, z));
        F3 x = cross(y, z);
        float phi = (rnd.b - 0.5f) * phi0;
        phi = minf(maxf(phi, -0.5f * phi0 + kEdgeEpsilon), 0.5f * phi0 - kEdgeEpsilon);
        float sp, cp;
        tr_sincos(phi, &sp, &cp);
        F3 x1 = x * cp + z * sp;
        bool m0 = rnd.c > 0.5f;
        float bb = 4.0f * rnd.c - (m0 ? 3.0f : 1.0f);
        float aa = sqrtf(1.0f - bb * bb) * (m0 ? -1.0f : 1.0f);
        dir = x1 * aa + y * bb;
      }
      float pdf = pdf_e * pdf_d;
      he = trace_ray<false>(S, tc, p, -dir, kShadowEpsilon, 1e32f);
      if (he.tri < 0) break;
      hs = trace_ray<false>(S, tc, p, dir, kShadowEpsilon, 1e32f);
      if (hs.tri < 0) break;
      Its ite = post_intersect(S, he.tri, he.u, he.v);
      Its its = post_intersect(S, hs.tri, hs.u, hs.v);
      float bv = base_val_secondary(its.p, ite.p, ite.ng, dir, p, edge, edge2, &n);
      v0 = bv / pdf;
      // with max_bounces <= 1 no path length fits; a zero or non-finite v0
      // makes every contribution vanish or not finite
      keep = P.max_bounces > 1 && v0 != 0.0f && isfinite(v0);
    } while (false);
    const int64_t s = wave_append(keep, Q.count);
    if (s >= 0) {
      QWriter w{Q, (uint32_t)s};
      w.u64(i); w.i(he.tri); w.f(he.u); w.f(he.v); w.i(hs.tri); w.f(hs.u); w.f(hs.v); w.f3(dir); w.f(v0);
      w.f3(p); w.f3(n); w.i(ei); w.f(t);
    }
  });
}

// eval: the emitter-side radiance after 1..B-1 bounces, the sensor-side
// importance, their combination by total length and the scatter.
// kTangent: the derivative image along tpos goes to dimg instead.
template <bool kTangent>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kBndWaves))) k_bnd_indirect_eval(SceneDev S, CameraDev C, BoundaryParams P, LaunchCfg L,
                                                              BndQueue Q, const float* dLdI, float* dpos,
                                                              const float* tpos, float* dimg) {
  __shared__ uint32_t lds_stack[kTraceLdsWords * kBlock];
  TraceCtx tc = make_trace_ctx_b(lds_stack, L);
  const uint32_t WH = (uint32_t)C.width * (uint32_t)C.height;
  const EdgeSet& E = S.sec;
  const int B = P.max_bounces;
  const uint32_t n_q = *Q.count;
  const uint32_t stride = (uint32_t)L.grid * kBlock;
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < n_q; s += stride) {
    QReader r{Q, s};
    const uint64_t i = r.u64();
    const int32_t tri_e = r.i();
    const float ue = r.f(), ve = r.f();
    const int32_t tri_s = r.i();
    const float us = r.f(), vs = r.f();
    const F3 dir = r.f3();
    const float v0 = r.f();
    const F3 p = r.f3();
    const F3 n = r.f3();
    const int32_t ei = r.i();
    const float t = r.f();
    Its ite = post_intersect(S, tri_e, ue, ve);
    Its its = post_intersect(S, tri_s, us, vs);
    const EdgeDev ed = E.e[ei];
    float ds = 0.0f;  // tangent mode: derivative of xDotN
    if constexpr (kTangent) {
      TangentSink sk{tpos, nullptr, nullptr, 0.0f};
      scatter_xdn(S, tri_e, its.p, p, n, 1.0f, &ed, t, ed.f0, tri_s, us, vs, sk);
      ds = sk.acc / (float)P.spp;
      if (ds == 0.0f || !isfinite(ds)) continue;
    }
    // emitter side: cumulative radiance after 1..B-1 bounces
    F3 cum[kMaxBounces];
    F3 per[kMaxBounces];
    for (int k = 0; k < kMaxBounces; ++k) per[k] = F3{0.0f, 0.0f, 0.0f};
    F3 acc{0.0f, 0.0f, 0.0f};
    if (B > 1)
      path_bounces<false, true>(S, tc, ite, F3{1.0f, 1.0f, 1.0f}, -dir, P.seed, kRngIndirectEdge, i, 1u, B - 1, acc,
                                nullptr, per);
    F3 run{0.0f, 0.0f, 0.0f};
    for (int k = 0; k < B - 1; ++k) {
      run = run + per[k];
      cum[k] = run;
    }
    // sensor side
    float w = 0.0f;
    Its cur = its;
    F3 cdir = dir, thr{1.0f, 1.0f, 1.0f};
    for (int b = 0; b < B - 1; ++b) {
      ImpStep st = importance_step(S, C, tc, cur, cdir, thr, rng4(P.seed, kRngIndirectEdge, i, kSensorBlock0 + (uint32_t)b));
      int ri = B - b - 2;
      if (st.contrib && ri >= 0) {
        F3 v = (v0 * cum[ri]) * st.value;
        if (length(v) > 0.0f) {
          if constexpr (kTangent) add_deriv(dimg, WH, st.pixel, v * ds);
          else w += dot(dLdI_at(dLdI, WH, st.pixel), v);
        }
      }
      if (!st.next) break;
    }
    if constexpr (!kTangent) {
      w = w / (float)P.spp;
      if (w == 0.0f || !isfinite(w)) continue;
      GradSink sk{dpos, nullptr, nullptr, nullptr};
      scatter_xdn(S, tri_e, its.p, p, n, w, &ed, t, ed.f0, tri_s, us, vs, sk);
    }
  }
}

hipError_t launch_guide_build(const SceneDev& S, const LaunchCfg& L, uint32_t seed, const GuideDev& g,
                              hipStream_t st) {
  const LaunchCfg Lg = fit_grid(k_guide_eval, L);
  if (Lg.tiles) {
    const hipError_t z = launch_fill_zero(reinterpret_cast<float*>(Lg.tiles), 1, st);
    if (z != hipSuccess) return z;
  }
  hipLaunchKernelGGL(k_guide_eval, dim3(Lg.grid), dim3(kBlock), 0, st, S, Lg, seed, g);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int n = g.n, nc = (n + kSumChunk - 1) / kSumChunk;
  const int gl = (n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096;  // grid over leaves
  const int gc = (nc + 63) / 64;                                     // grid over chunks
  float* total = g.part + nc;
  hipLaunchKernelGGL(k_guide_poly, dim3(gl), dim3(256), 0, st, g);
  hipLaunchKernelGGL(k_chunk_sums, dim3(gc), dim3(64), 0, st, g.poly, n, g.part);
  hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(64), 0, st, g.part, nc);
  hipLaunchKernelGGL(k_guide_weights, dim3(gl), dim3(256), 0, st, S.sec.cdf, g, total);
  hipLaunchKernelGGL(k_chunk_sums, dim3(gc), dim3(64), 0, st, g.cdf, n, g.part);
  hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(64), 0, st, g.part, nc);
  hipLaunchKernelGGL(k_chunk_prefix, dim3(gc), dim3(64), 0, st, g.cdf, n, g.part, 1.0f / (float)n);
  hipLaunchKernelGGL(k_guide_normalize, dim3(gl), dim3(256), 0, st, g, total);
  hipLaunchKernelGGL(k_guide_normalize_last, dim3(1), dim3(64), 0, st, g);
  return hipGetLastError();
}

uint64_t boundary_queue_bytes(int kind, uint64_t n_samples) {
  if (kind == 1 || n_samples == 0) return 0;
  const uint64_t nchunk = (n_samples + kBndChunk - 1) / kBndChunk;
  const uint64_t cap = n_samples < kBndChunk ? n_samples : kBndChunk;
  return ((uint64_t)kBndWords * cap + nchunk) * sizeof(uint32_t);
}

template <bool kTangent>
static hipError_t launch_boundary_t(int kind, const SceneDev& S, const CameraDev& C, const BoundaryParams& P,
                                    uint64_t N, const LaunchCfg& L, const EdgeSet& prim, const GuideDev& guide,
                                    const float* dLdI, float* dpos, const float* tpos, float* dimg, void* queue,
                                    hipStream_t st) {
  if (kind == 1) {
    const LaunchCfg Lk = fit_grid(k_boundary_pixel<kTangent>, L);
    if (Lk.tiles) {
      const hipError_t e = launch_fill_zero(reinterpret_cast<float*>(Lk.tiles), 1, st);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_boundary_pixel<kTangent>, dim3(Lk.grid), dim3(kBlock), 0, st, S, C, P, Lk, N, dLdI, dpos,
                       tpos, dimg);
    return hipGetLastError();
  }
  // select / eval per chunk of kBndChunk samples over one survivor queue;
  // the chunks' counters are cleared by a kernel (no memset node when captured)
  if (!queue) return hipErrorInvalidValue;
  const uint64_t nchunk = (N + kBndChunk - 1) / kBndChunk;
  const uint32_t cap = (uint32_t)(N < kBndChunk ? N : kBndChunk);
  uint32_t* buf = (uint32_t*)queue;
  uint32_t* counts = buf + (size_t)kBndWords * cap;
  hipError_t e = launch_fill_zero((float*)counts, nchunk, st);
  for (uint64_t c = 0; c < nchunk && e == hipSuccess; ++c) {
    const uint64_t first = c * kBndChunk, end = first + kBndChunk < N ? first + kBndChunk : N;
    const BndQueue Q{counts + c, buf, cap};
    if (kind == 0) {
      const LaunchCfg Ls = fit_grid(k_bnd_primary_select, L), Le = fit_grid(k_bnd_primary_eval<kTangent>, L);
      if (Ls.tiles && (e = launch_fill_zero(reinterpret_cast<float*>(Ls.tiles), 1, st)) != hipSuccess) break;
      hipLaunchKernelGGL(k_bnd_primary_select, dim3(Ls.grid), dim3(kBlock), 0, st, S, C, P, Ls, prim, first, end, Q);
      hipLaunchKernelGGL(k_bnd_primary_eval<kTangent>, dim3(Le.grid), dim3(kBlock), 0, st, S, C, P, Le, prim, Q,
                         dLdI, dpos, tpos, dimg);
    } else if (kind == 2) {
      const LaunchCfg Ls = fit_grid(k_bnd_direct_select, L), Le = fit_grid(k_bnd_direct_eval<kTangent>, L);
      if (Ls.tiles && (e = launch_fill_zero(reinterpret_cast<float*>(Ls.tiles), 1, st)) != hipSuccess) break;
      hipLaunchKernelGGL(k_bnd_direct_select, dim3(Ls.grid), dim3(kBlock), 0, st, S, C, P, Ls, guide, first, end, Q);
      hipLaunchKernelGGL(k_bnd_direct_eval<kTangent>, dim3(Le.grid), dim3(kBlock), 0, st, S, C, P, Le, Q, dLdI, dpos,
                         tpos, dimg);
    } else {
      const LaunchCfg Ls = fit_grid(k_bnd_indirect_select, L), Le = fit_grid(k_bnd_indirect_eval<kTangent>, L);
      if (Ls.tiles && (e = launch_fill_zero(reinterpret_cast<float*>(Ls.tiles), 1, st)) != hipSuccess) break;
      hipLaunchKernelGGL(k_bnd_indirect_select, dim3(Ls.grid), dim3(kBlock), 0, st, S, C, P, Ls, first, end, Q);
      hipLaunchKernelGGL(k_bnd_indirect_eval<kTangent>, dim3(Le.grid), dim3(kBlock), 0, st, S, C, P, Le, Q, dLdI,
                         dpos, tpos, dimg);
    }
    e = hipGetLastError();
  }
  return e;
}

hipError_t launch_boundary(int kind, const SceneDev& S, const CameraDev& C, uint32_t seed, int max_bounces, int spp,
                           const LaunchCfg& L, const EdgeSet& prim, const GuideDev& guide, const float* dLdI,
                           float* dpos, const float* tpos, float* dimg, void* queue, hipStream_t st) {
  const BoundaryParams P{seed, max_bounces, spp, 0u};
  const uint64_t N = (uint64_t)spp * (uint64_t)C.width * (uint64_t)C.height;
  if (N == 0) return hipSuccess;
  if (dimg)
    return launch_boundary_t<true>(kind, S, C, P, N, L, prim, guide, nullptr, nullptr, tpos, dimg, queue, st);
  return launch_boundary_t<false>(kind, S, C, P, N, L, prim, guide, dLdI, dpos, nullptr, nullptr, queue, st);
}

}  // namespace trc
