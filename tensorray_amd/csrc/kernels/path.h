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

// Bounce loop (PathTracer.cpp:55-73, EvalRadianceDirect Integrator.cpp:178-303)
// from surface point `its` reached along `raydir` with throughput `beta`.
// Bounce b draws blocks block0 + 2b (light) and block0 + 2b + 1 (BSDF) of
// (seed, domain, index).  kRecord fills the replay record (interior
// adjoint); kPerBounce stores each executed bounce's contribution.
template <bool kRecord, bool kPerBounce>
__device__ void path_bounces(const SceneDev& S, TraceCtx& tc, Its its, F3 beta, F3 raydir, uint32_t seed,
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
      if (stable && noleak)
        visible = trace_ray<true>(S, tc, its.p, wi, kShadowEpsilon, dist - kShadowEpsilon).tri < 0;
      float valid = (stable && noleak && visible) ? 1.0f : 0.0f;
      float Gl = fabsf(dot(ls.n, -wi)) / d2;
      F3 f = bsdf_eval(S, its, wo, wi) * fabsf(dsh);
      F3 Le = emitted(S, ls.emitter, ls.n, -wi);
      float JL = S.shade[ls.tri].area / S.shade[ls.tri].area;
      F3 val = valid * beta * f * JL * Gl * Le / ls.pdf;
      float pb = bsdf_pdf(S, its, wo, wi) * Gl;
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
    Hit h2 = trace_ray<false>(S, tc, its.p, wib, kShadowEpsilon, 1e32f);
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

template <bool kRecord>
__device__ F3 path_forward(const SceneDev& S, const CameraDev& C, const PathParams& P, uint64_t path,
                           TraceCtx& tc, PathRec* R) {
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
  Hit h = trace_ray<false>(S, tc, eye, d0, kShadowEpsilon, 1e32f);
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

}  // namespace trc
