// One camera path of the interior integrator with next-event estimation and
// MIS, evaluated per thread.  The same routine drives the forward kernel
// (kRecord = false) and the first sweep of the adjoint kernel (kRecord =
// true), which stores the discrete decisions and throughputs of the path in
// a small per-thread record instead of an expression tape.
//
// Reference: PathTracer::Radiance (LibTensorRay/Renderer/PathTracer.cpp:37-76),
// EvalRadianceEmitted (Integrator.cpp:55-73), EvalRadianceDirect
// (Integrator.cpp:178-303), Camera::GenerateAntitheticRays (Camera.cpp:224-277)
// and Camera::EvalFilter (Camera.cpp:279-303).
#pragma once
#include "../common/tr_bsdf.h"
#include "trace.h"

namespace trc {

struct PathParams {
  uint32_t seed;
  int32_t max_bounces;   // B: NEE + BSDF steps after the primary hit
  int32_t spp;
  uint32_t path_offset;  // index of the first path of this launch (spp passes)
};

// Per-path record for replay: hit vertices x_0..x_K, throughputs, light
// samples and the constant (detached) factors of every contribution.
struct PathRec {
  int32_t K;                     // index of the last hit vertex, -1 = camera ray missed
  int32_t tri[kMaxBounces + 1];
  float bu[kMaxBounces + 1], bv[kMaxBounces + 1];
  F3 beta[kMaxBounces + 1];
  F3 e0c;                        // emission at x_0 (constant factor)
  int32_t ltri[kMaxBounces];     // light sample of bounce i
  float lbu[kMaxBounces], lbv[kMaxBounces];
  F3 neec[kMaxBounces];          // valid * Le * mis / pdf_L, zero when unused
  float ppdf[kMaxBounces];       // pdf_bsdf * detach(G) of segment i -> i+1
  F3 emc[kMaxBounces];           // Le * mis at x_{i+1}, zero when not an emitter
  uint32_t tflag;                // bit i: throughput of segment i is nonzero (G > 0)
};

TR_HD F3 nan_to_zero(F3 v) { return finite3(v) ? v : F3{0.0f, 0.0f, 0.0f}; }

// Path checkpoint: the outcome of every ray query of a path, written by the
// forward pass so that the adjoint re-evaluates the path without tracing.
// Per path: meta = (K + 1) | (NEE visibility of bounce b) << (8 + b), where
// K is the last hit vertex (-1: the camera ray missed), and for each vertex
// i <= K the hit triangle and barycentrics.  Fields are [vertex][path]
// planes, so lanes with consecutive paths touch consecutive words.
struct CkptDev {
  uint32_t* meta;   // [npaths]
  int32_t* tri;     // [nv][npaths]
  float* u;         // [nv][npaths]
  float* v;         // [nv][npaths]
  uint64_t npaths;
  int32_t nv;       // max_bounces + 1
};
// 4 bytes of meta + 12 bytes per vertex.
TR_HD uint64_t ckpt_bytes(uint64_t npaths, int max_bounces) {
  return npaths * (4ull + 12ull * (uint64_t)(max_bounces + 1));
}
TR_HD CkptDev ckpt_view(void* base, uint64_t npaths, int max_bounces) {
  CkptDev c;
  const uint64_t nv = (uint64_t)(max_bounces + 1);
  c.meta = (uint32_t*)base;
  c.tri = (int32_t*)(c.meta + npaths);
  c.u = (float*)(c.tri + nv * npaths);
  c.v = c.u + nv * npaths;
  c.npaths = npaths;
  c.nv = (int32_t)nv;
  return c;
}
__device__ inline void ckpt_put_vertex(const CkptDev& c, uint64_t path, int i, int32_t tri, float u, float v) {
  const uint64_t k = (uint64_t)i * c.npaths + path;
  c.tri[k] = tri;
  c.u[k] = u;
  c.v[k] = v;
}

// Where a path's ray queries come from: the BVH (TraceCtx) or a checkpoint.
__device__ inline Hit src_camera(const SceneDev& S, TraceCtx& tc, F3 o, F3 d) {
  return trace_ray<false>(S, tc, o, d, kShadowEpsilon, 1e32f);
}
__device__ inline bool src_visible(const SceneDev& S, TraceCtx& tc, int b, F3 o, F3 wi, float dist) {
  return trace_ray<true>(S, tc, o, wi, kShadowEpsilon, dist - kShadowEpsilon).tri < 0;
}
__device__ inline Hit src_bsdf(const SceneDev& S, TraceCtx& tc, int b, F3 o, F3 d) {
  return trace_ray<false>(S, tc, o, d, kShadowEpsilon, 1e32f);
}

struct CkptSrc {
  const CkptDev* c;
  uint64_t path;
  uint32_t meta;
  __device__ Hit vertex(int i) const {
    Hit h;
    const int K = (int)(meta & 0xFFu) - 1;
    if (i > K) {
      h.tri = -1; h.u = 0.0f; h.v = 0.0f; h.t = -1.0f;
      return h;
    }
    const uint64_t k = (uint64_t)i * c->npaths + path;
    h.tri = c->tri[k];
    h.u = c->u[k];
    h.v = c->v[k];
    h.t = 0.0f;
    return h;
  }
};
__device__ inline Hit src_camera(const SceneDev&, CkptSrc& s, F3, F3) { return s.vertex(0); }
__device__ inline bool src_visible(const SceneDev&, CkptSrc& s, int b, F3, F3, float) {
  return ((s.meta >> (8 + b)) & 1u) != 0u;
}
__device__ inline Hit src_bsdf(const SceneDev&, CkptSrc& s, int b, F3, F3) { return s.vertex(b + 1); }

// Bounce loop (PathTracer.cpp:55-73, EvalRadianceDirect Integrator.cpp:178-303)
// from surface point `its` reached along `raydir` with throughput `beta`.
// Bounce b draws blocks block0 + 2b (light) and block0 + 2b + 1 (BSDF) of
// (seed, domain, index).  kRecord fills the replay record (interior
// adjoint); kPerBounce stores each executed bounce's contribution.
template <bool kRecord, bool kPerBounce, class Src>
__device__ void path_bounces(const SceneDev& S, Src& tc, Its its, F3 beta, F3 raydir, uint32_t seed,
                             uint32_t domain, uint64_t index, uint32_t block0, int B, F3& contrib, PathRec* R,
                             F3* per_bounce) {
  const F3 zero{0.0f, 0.0f, 0.0f};
  for (int b = 0; b < B; ++b) {
    Rnd4 rl = rng4(seed, domain, index, block0 + 2u * (uint32_t)b);
    Rnd4 rb = rng4(seed, domain, index, block0 + 2u * (uint32_t)b + 1u);
    F3 ret = zero;
    F3 wo = -raydir;
    // ---- next-event estimation ----
    LightSample ls = light_sample(S, rl.a, rl.b);
    F3 neec = zero;
    {
      F3 wi = ls.p - its.p;
      float d2 = len2(wi);
      float dist = sqrtf(d2);
      wi = wi / dist;
      float dsh = dot(its.ns, wi), dgeo = dot(its.ng, wi);
      bool stable = dist > kShadowEpsilon && fabsf(dsh) > kEdgeEpsilon;
      bool noleak = (dsh * dgeo) > 0.0f;
      bool visible = false;
      if (stable && noleak) visible = src_visible(S, tc, b, its.p, wi, dist);
      float valid = (stable && noleak && visible) ? 1.0f : 0.0f;
      float Gl = fabsf(dot(ls.n, -wi)) / d2;
      float pbs;
      F3 f = bsdf_eval_pdf(S, its, wo, wi, &pbs) * fabsf(dsh);
      F3 Le = emitted(S, ls.emitter, ls.n, -wi);
      float JL = S.shade[ls.tri].area / S.shade[ls.tri].area;
      F3 val = valid * beta * f * JL * Gl * Le / ls.pdf;
      float pb = pbs * Gl;
      float mis = (ls.pdf * ls.pdf) / (ls.pdf * ls.pdf + pb * pb);
      val = val * mis;
      ret = ret + val;
      if (kRecord) neec = valid * Le / ls.pdf * mis;
    }
    if (kRecord) {
      R->ltri[b] = ls.tri; R->lbu[b] = ls.bu; R->lbv[b] = ls.bv;
      R->neec[b] = neec;
      R->emc[b] = zero;
    }
    // ---- BSDF sampling ----
    float pdfb;
    F3 wib = bsdf_sample(S, its, wo, rb.a, rb.b, &pdfb);
    Hit h2 = src_bsdf(S, tc, b, its.p, wib);
    bool ok = h2.tri >= 0 && pdfb > 0.0f && bsdf_dir_valid(S, its, wib);
    if (!ok) {
      if (finite3(ret)) {
        contrib = contrib + ret;
        if (kPerBounce) per_bounce[b] = ret;
      } else if (kRecord) {
        R->neec[b] = zero;
      }
      break;
    }
    Its its2 = post_intersect(S, h2.tri, h2.u, h2.v);
    F3 dd = its2.p - its.p;
    float dd2 = len2(dd);
    F3 nd = dd / sqrtf(dd2);
    float G2 = fabsf(dot(its2.ng, -nd)) / dd2;
    F3 f2 = bsdf_eval(S, its, wo, nd) * fabsf(dot(its.ns, nd));
    float ppdf = pdfb * G2;
    float J2 = S.shade[its2.tri].area / S.shade[its2.tri].area;
    F3 beta2 = beta * f2 * G2 * J2 / ppdf;
    bool gpos = G2 > 0.0f;
    if (!gpos) beta2 = zero;
    F3 emc = zero;
    if (its2.emitter >= 0) {
      F3 Le2 = emitted(S, its2.emitter, its2.ng, -nd);
      F3 val2 = beta2 * Le2;
      float pn = light_pdf_tri(S, its2.tri);
      float mis2 = (ppdf * ppdf) / (ppdf * ppdf + pn * pn);
      val2 = val2 * mis2;
      ret = ret + val2;
      emc = Le2 * mis2;
    }
    bool fin = finite3(ret);
    if (fin) {
      contrib = contrib + ret;
      if (kPerBounce) per_bounce[b] = ret;
    }
    if (kRecord) {
      if (!fin) { R->neec[b] = zero; emc = zero; }
      R->emc[b] = emc;
      R->ppdf[b] = ppdf;
      if (gpos) R->tflag |= 1u << b;
      R->K = b + 1;
      R->tri[b + 1] = its2.tri; R->bu[b + 1] = h2.u; R->bv[b + 1] = h2.v;
      R->beta[b + 1] = beta2;
    }
    its = its2;
    beta = beta2;
    raydir = nd;
  }
}

template <bool kRecord, class Src>
__device__ F3 path_forward(const SceneDev& S, const CameraDev& C, const PathParams& P, uint64_t path,
                           Src& tc, PathRec* R) {
  const F3 zero{0.0f, 0.0f, 0.0f};
  const uint32_t WH = (uint32_t)C.width * (uint32_t)C.height;
  const uint32_t pix = (uint32_t)(path % WH);
  const uint32_t px = pix % (uint32_t)C.width, py = pix / (uint32_t)C.width;
  if (kRecord) { R->K = -1; R->tflag = 0u; }
  // --- primary ray (box filter: the sample stays inside its pixel) ---
  Rnd4 rc = rng4(P.seed, kRngInterior, path, 0u);
  const F3 eye{C.eye[0], C.eye[1], C.eye[2]};
  const F3 cdir{C.dir[0], C.dir[1], C.dir[2]};
  F3 d0 = camera_dir(C, (float)px + rc.a, (float)py + rc.b);
  Hit h = src_camera(S, tc, eye, d0);
  if (h.tri < 0) return zero;
  Its its = post_intersect(S, h.tri, h.u, h.v);
  // EvalFilter: G / detach(G) with G = |cos_x| / (cos_y^3 d^2); times J
  float beta0;
  {
    F3 cd = its.p - eye;
    float dist = length(cd);
    cd = cd / dist;
    float cosy = dot(cd, cdir);
    float cosx = dot(-cd, its.ng);
    float G = fabsf(cosx) / (cosy * cosy * cosy * (dist * dist));
    float area = S.shade[its.tri].area;
    beta0 = (G / G) * (area / area);
  }
  F3 beta{beta0, beta0, beta0};
  F3 raydir = normalize(its.p - eye);
  F3 contrib = zero;
  F3 e0c = zero;
  if (its.emitter >= 0) {
    F3 Le = emitted(S, its.emitter, its.ng, -raydir);
    F3 e0 = beta * Le;
    if (finite3(e0)) { contrib = contrib + e0; e0c = Le; }
  }
  if (kRecord) {
    R->K = 0;
    R->tri[0] = its.tri; R->bu[0] = h.u; R->bv[0] = h.v;
    R->beta[0] = beta;
    R->e0c = e0c;
  }
  path_bounces<kRecord, false>(S, tc, its, beta, raydir, P.seed, kRngInterior, path, 1u, P.max_bounces, contrib,
                               R, nullptr);
  return contrib;
}

// ---------------------------------------------------------------------------
// Forward paths as a per-lane state machine with path regeneration.
//
// path_forward runs one path to completion per lane, so a wave waits for its
// longest path and every traversal is entered by whichever lanes still live.
// Here each loop iteration advances one query per lane -- the camera ray, or
// the two rays of a bounce (the NEE shadow ray, then the BSDF-sampled ray,
// as one chained query from the vertex) -- and shades the lanes whose query
// has finished; a lane whose path ends takes the next path of its wave's
// chunk (ballot + popcount hand-out, no global atomics).  A query that is
// still running when most of the wave's have finished keeps its traversal
// state and continues in the next iteration.  Each path evaluates the same
// expressions in the same order as path_forward<false>, so the per-path
// radiance is identical.
// ---------------------------------------------------------------------------
enum PathStage : uint32_t { kStageIdle = 0, kStageCamera = 1, kStageBounce = 2 };

// Fraction (of 64) of a wave's queries still unfinished at which the wave
// stops traversing and shades the finished ones.
constexpr int kTraceKeep64 = 16;

struct PathLane {
  uint64_t path;
  // stage (bits 0-1), the bounce's BSDF direction has a nonzero pdf (bit 2),
  // bounce index b (bits 3-31): one register
  uint32_t sb;
  __device__ uint32_t stage() const { return sb & 3u; }
  __device__ uint32_t b() const { return sb >> 3; }
  __device__ void set_stage(uint32_t s) { sb = (sb & ~3u) | s; }
  int32_t tri;           // current vertex x_b
  float u, v;
  F3 beta, raydir, contrib;
  float la, lb;          // light-sample uniforms of bounce b
  float pdfb;            // pdf of the BSDF direction d2 of bounce b
  // query of this iteration from o (tmin = kShadowEpsilon): the camera ray
  // along d, or the bounce's shadow ray along d (any hit) chained with its
  // BSDF-sampled ray along d2; the traversal state holds the far bound
  F3 o, d, d2;
  TraceState ts;
};

// NEE geometry of bounce b at `its` (Integrator.cpp:196-213).
struct NeeGeo {
  LightSample ls;
  F3 wi;
  float d2, dist, dsh;
  bool ray;  // stable && noleak: a shadow ray is traced
};
__device__ inline NeeGeo nee_geo(const SceneDev& S, const Its& its, float la, float lb) {
  NeeGeo g;
  g.ls = light_sample(S, la, lb);
  F3 wi = g.ls.p - its.p;
  g.d2 = len2(wi);
  g.dist = sqrtf(g.d2);
  g.wi = wi / g.dist;
  g.dsh = dot(its.ns, g.wi);
  float dgeo = dot(its.ng, g.wi);
  bool stable = g.dist > kShadowEpsilon && fabsf(g.dsh) > kEdgeEpsilon;
  bool noleak = (g.dsh * dgeo) > 0.0f;
  g.ray = stable && noleak;
  return g;
}

// The path ranges a wave renders: tiles t, t + step, ... (below tile `tend`)
// of `tile` paths each, tile k covering [base + k*tile, ...) clipped to
// `limit`.  The wave moves to its next tile as soon as the current one is
// handed out, without waiting for its lanes to drain.  All fields are
// wave-uniform.
struct TileSupply {
  uint64_t next, end;            // paths not yet handed out of the current tile
  uint64_t t, tend, step;        // tile cursor
  uint64_t base, limit, tile, t0;
  // tile counter: when set, the wave's tiles after its first (tile t, one
  // per wave: tiles 0 .. step-1 of the launch are dealt so) are step + the
  // counter's next value, taken as the wave needs one; otherwise t,
  // t + step, ...  Tile numbers count over the whole launch.
  uint32_t* ctr;
  __device__ void load() {
    next = base + (t - t0) * tile;
    end = next + tile < limit ? next + tile : limit;
  }
  // Starts at tile t; tiles t0.. map to paths base..limit.
  __device__ static TileSupply make(uint64_t t, uint64_t tend, uint64_t step, uint64_t t0, uint64_t base,
                                    uint64_t limit, uint64_t tile) {
    TileSupply s;
    s.t = t;
    s.tend = tend;
    s.step = step;
    s.t0 = t0;
    s.base = base;
    s.limit = limit;
    s.tile = tile;
    s.ctr = nullptr;
    if (t < tend) s.load();
    else s.next = s.end = 0;
    return s;
  }
  // Moves to the wave's next tile; false (and an empty range) when none is
  // left, in which case t stays at the first tile past tend.
  __device__ bool more() {
    if (t >= tend) {
      next = end;
      return false;
    }
    if (ctr) {
      uint32_t v = 0u;
      if ((threadIdx.x & 63u) == 0u) v = atomicAdd(ctr, 1u);
      t = step + (uint64_t)(uint32_t)__shfl((int)v, 0);
    } else {
      t += step;
    }
    if (t >= tend) {
      next = end;
      return false;
    }
    load();
    return true;
  }
};

// The BSDF-sampled ray of bounce b from `its` came back with hit `h`:
// MIS-weighted emission, throughput update and the move to the next vertex,
// or the end of the path (Integrator.cpp:232-293).  ret: the bounce's NEE
// contribution.  Returns true, with `its` replaced by the new vertex
// x_{b+1}, when the path continues.
template <bool kCkpt, class Finish>
__device__ inline bool bsdf_hit(const SceneDev& S, const PathParams& P, const CkptDev& ck, Its& its, F3 wo,
                                const Hit& h, F3 ret, PathLane& L, uint32_t& meta, Finish& finish) {
  const F3 zero{0.0f, 0.0f, 0.0f};
  bool ok = h.tri >= 0 && L.pdfb > 0.0f && bsdf_dir_valid(S, its, L.d2);
  if (!ok) {
    if (finite3(ret)) L.contrib = L.contrib + ret;
    if (kCkpt) ck.meta[L.path] = meta;
    finish(L.path, L.contrib);
    L.set_stage(kStageIdle);
    return false;
  }
  Its its2 = post_intersect(S, h.tri, h.u, h.v);
  F3 dd = its2.p - its.p;
  float dd2 = len2(dd);
  F3 nd = dd / sqrtf(dd2);
  float G2 = fabsf(dot(its2.ng, -nd)) / dd2;
  F3 f2 = bsdf_eval(S, its, wo, nd) * fabsf(dot(its.ns, nd));
  float ppdf = L.pdfb * G2;
  float J2 = S.shade[its2.tri].area / S.shade[its2.tri].area;
  F3 beta2 = L.beta * f2 * G2 * J2 / ppdf;
  if (!(G2 > 0.0f)) beta2 = zero;
  if (its2.emitter >= 0) {
    F3 Le2 = emitted(S, its2.emitter, its2.ng, -nd);
    F3 val2 = beta2 * Le2;
    float pn = light_pdf_tri(S, its2.tri);
    float mis2 = (ppdf * ppdf) / (ppdf * ppdf + pn * pn);
    val2 = val2 * mis2;
    ret = ret + val2;
  }
  if (finite3(ret)) L.contrib = L.contrib + ret;
  L.tri = h.tri; L.u = h.u; L.v = h.v;
  L.beta = beta2;
  L.raydir = nd;
  L.sb += 8u;  // b + 1
  if (kCkpt) {
    ckpt_put_vertex(ck, L.path, (int)L.b(), h.tri, h.u, h.v);
    meta += 1u;
  }
  if ((int)L.b() >= P.max_bounces) {
    if (kCkpt) ck.meta[L.path] = meta;
    finish(L.path, L.contrib);
    L.set_stage(kStageIdle);
    return false;
  }
  its = its2;
  return true;
}

// The wave renders the paths of its tile supply.  A bounce is two ray
// queries, the shadow ray of next-event estimation and the BSDF-sampled
// ray, traced as one chained query and shaded together.
template <bool kCkpt, class Finish>
__device__ void render_paths_regen(const SceneDev& S, const CameraDev& C, const PathParams& P, TraceCtx& tc,
                                   TileSupply& sup, const CkptDev& ck, Finish finish) {
  const F3 zero{0.0f, 0.0f, 0.0f};
  const uint32_t WH = (uint32_t)C.width * (uint32_t)C.height;
  const F3 eye{C.eye[0], C.eye[1], C.eye[2]};
  const F3 cdir{C.dir[0], C.dir[1], C.dir[2]};
  const int lane = (int)(threadIdx.x & 63u);
  PathLane L;
  L.sb = kStageIdle;
  trace_none(L.ts);
  uint32_t meta = 0u;  // checkpoint meta word of the current path
  for (;;) {
    // ---- regenerate: idle lanes take the next paths of the wave's tiles ----
    if (sup.next >= sup.end) sup.more();
    const uint64_t next = sup.next, end = sup.end;
    const bool idle = L.stage() == kStageIdle;
    const uint64_t idle_mask = __ballot(idle);
    if (idle) {
      const uint64_t rank = (uint64_t)__popcll(idle_mask & ((1ull << lane) - 1ull));
      if (next + rank < end) {
        L.path = next + rank;
        const uint32_t pix = (uint32_t)(L.path % WH);
        const uint32_t px = pix % (uint32_t)C.width, py = pix / (uint32_t)C.width;
        Rnd4 rc = rng4(P.seed, kRngInterior, L.path, 0u);
        L.sb = kStageCamera;  // b = 0
        L.o = eye;
        L.d = camera_dir(C, (float)px + rc.a, (float)py + rc.b);
        trace_begin(L.ts, 1e32f);
      }
    }
    {
      const uint64_t n_idle = (uint64_t)__popcll(idle_mask);
      sup.next = next + n_idle < end ? next + n_idle : end;
    }
    if (__ballot(L.stage() != kStageIdle) == 0ull) break;
    // ---- one query per lane ----
    // The wave leaves the traversal loop once at most kTraceKeep64 / 64 of
    // its queries are unfinished; those lanes skip shading and resume their
    // query in the next iteration, beside the new rays of the others.
    // (a BSDF direction of pdf 0 ends the path whatever it hits: its query
    // is given an empty interval)
    trace_run<true>(S, tc, L.o, L.d, L.d2, (L.sb & 4u) ? 1e32f : -1.0f, kShadowEpsilon,
                    L.stage() == kStageBounce, L.ts, kTraceKeep64);
    if (!trace_finished(L.ts)) continue;
    const Hit h = trace_hit(L.ts);
    // ---- shade ----
    // Both stages start from one surface interaction -- the first vertex for
    // a camera ray, the current vertex x_b for a bounce -- and end with the
    // next bounce's set-up (light sample, shadow ray, BSDF direction), each
    // evaluated once for the whole wave.
    const uint32_t stage = L.stage();
    if (stage == kStageIdle) continue;  // no path left for this lane
    const bool cam = stage == kStageCamera;
    if (cam && h.tri < 0) {
      if (kCkpt) ck.meta[L.path] = 0u;
      finish(L.path, zero);
      L.set_stage(kStageIdle);
      continue;
    }
    Its its = post_intersect(S, cam ? h.tri : L.tri, cam ? h.u : L.u, cam ? h.v : L.v);
    if (cam) {
      if (kCkpt) {
        ckpt_put_vertex(ck, L.path, 0, h.tri, h.u, h.v);
        meta = 1u;
      }
      float beta0;
      {
        F3 cd = its.p - eye;
        float dist = length(cd);
        cd = cd / dist;
        float cosy = dot(cd, cdir);
        float cosx = dot(-cd, its.ng);
        float G = fabsf(cosx) / (cosy * cosy * cosy * (dist * dist));
        float area = S.shade[its.tri].area;
        beta0 = (G / G) * (area / area);
      }
      L.beta = F3{beta0, beta0, beta0};
      L.raydir = normalize(its.p - eye);
      L.contrib = zero;
      if (its.emitter >= 0) {
        F3 Le = emitted(S, its.emitter, its.ng, -L.raydir);
        F3 e0 = L.beta * Le;
        if (finite3(e0)) L.contrib = L.contrib + e0;
      }
      L.tri = h.tri; L.u = h.u; L.v = h.v;
      if (P.max_bounces <= 0) {
        if (kCkpt) ck.meta[L.path] = meta;
        finish(L.path, L.contrib);
        L.set_stage(kStageIdle);
        continue;
      }
    } else {
      // next-event estimation of bounce b (Integrator.cpp:196-230) with the
      // traced shadow ray's outcome, then the BSDF-sampled ray's hit
      const F3 wo = -L.raydir;
      const NeeGeo g = nee_geo(S, its, L.la, L.lb);
      bool visible = g.ray && !trace_first_hit(L.ts);
      if (kCkpt && visible) meta |= 1u << (8u + L.b());
      float valid = visible ? 1.0f : 0.0f;
      float Gl = fabsf(dot(g.ls.n, -g.wi)) / g.d2;
      float pbs;
      F3 f = bsdf_eval_pdf(S, its, wo, g.wi, &pbs) * fabsf(g.dsh);
      F3 Le = emitted(S, g.ls.emitter, g.ls.n, -g.wi);
      float JL = S.shade[g.ls.tri].area / S.shade[g.ls.tri].area;
      F3 val = valid * L.beta * f * JL * Gl * Le / g.ls.pdf;
      float pb = pbs * Gl;
      float mis = (g.ls.pdf * g.ls.pdf) / (g.ls.pdf * g.ls.pdf + pb * pb);
      val = val * mis;
      F3 ret = zero;
      ret = ret + val;
      if (!bsdf_hit<kCkpt>(S, P, ck, its, wo, h, ret, L, meta, finish)) continue;
    }
    // ---- bounce b at `its`: light sample, shadow ray and BSDF direction ----
    // (Integrator.cpp:196-213 and 232-245; the draws of the reference's
    // loop in its order: light sample 1 + 2b, BSDF sample 2 + 2b)
    const Rnd4 rl = rng4(P.seed, kRngInterior, L.path, 1u + 2u * L.b());
    L.la = rl.a;
    L.lb = rl.b;
    const NeeGeo g2 = nee_geo(S, its, rl.a, rl.b);
    const Rnd4 rb = rng4(P.seed, kRngInterior, L.path, 1u + 2u * L.b() + 1u);
    L.d2 = bsdf_sample(S, its, -L.raydir, rb.a, rb.b, &L.pdfb);
    L.sb = (L.sb & ~7u) | kStageBounce | (L.pdfb > 0.0f ? 4u : 0u);
    L.o = its.p;
    L.d = g2.wi;
    trace_begin_chain(L.ts, g2.ray ? g2.dist - kShadowEpsilon : -1.0f);
  }
}

}  // namespace trc
