// Hand-written reverse-mode adjoint of one interior path.
//
// The reference differentiates its whole per-pass expression graph
// (PathTracer::Integrate -> Tensor::Backward, PathTracer.cpp:107-111,
// Tensor.cpp:601-635).  Here each path is replayed from its counter-based
// random numbers (path_forward<true>), which records the hit triangles,
// barycentrics, light samples and throughputs, and then swept backwards:
//
//   C = sum_i beta_i * (e_i + a_i),   beta_{i+1} = beta_i * t_i
//   Q_i = g * (e_i + a_i) + t_i * Q_{i+1}            (adjoint of beta_i)
//
// where a_i is the NEE term at vertex i, e_i the (constant) emission and MIS
// factor, t_i the area-measure throughput of segment i -> i+1 and g = dL/dI
// of the path's pixel.  Local vector-Jacobian products of a_i and t_i move
// the adjoint into the hit positions (material form, fixed barycentrics),
// face normals, areas, the shading frame, the incoming direction and the
// BSDF parameters; triangle adjoints are flushed with float atomics once no
// later step can touch them.
#pragma once
#include "path.h"

namespace trc {

// Material parameter gradient slots: [roughness, albedo rgb, spec rgb, pad].
constexpr int kMatGradStride = 8;
constexpr int kMaxMatGrad = 32;

struct TriAdj {
  F3 a0, a1, a2;
};
TR_HD void tri_adj_zero(TriAdj& t) { t.a0 = t.a1 = t.a2 = F3{0.0f, 0.0f, 0.0f}; }

TR_HD F3 sgn3(float s, F3 v) { return s >= 0.0f ? v : -v; }
TR_HD float sgnf(float s) { return s > 0.0f ? 1.0f : (s < 0.0f ? -1.0f : 0.0f); }

// y = x / |x|  ->  adj_x = (adj_y - y (y . adj_y)) / |x|
TR_HD F3 normalize_vjp(F3 y, float len, F3 adj_y) { return (adj_y - y * dot(y, adj_y)) / len; }

// Face normal n = c/|c| and area |c|/2 with c = (p1-p0) x (p2-p0).
struct FaceGeom {
  F3 p0, p1, p2, e0, e1, c, n;
  float len, area;
};
TR_HD FaceGeom face_geom(F3 p0, F3 p1, F3 p2) {
  FaceGeom g;
  g.p0 = p0; g.p1 = p1; g.p2 = p2;
  g.e0 = p1 - p0; g.e1 = p2 - p0;
  g.c = cross(g.e0, g.e1);
  g.len = length(g.c);
  g.n = g.c / g.len;
  g.area = g.len * 0.5f;
  return g;
}
TR_HD void face_vjp(const FaceGeom& g, F3 adj_n, float adj_area, TriAdj& out) {
  F3 adj_c = normalize_vjp(g.n, g.len, adj_n) + g.n * (0.5f * adj_area);
  F3 ae0 = cross(g.e1, adj_c);
  F3 ae1 = cross(adj_c, g.e0);
  out.a1 += ae0;
  out.a2 += ae1;
  out.a0 -= ae0 + ae1;
}

// CoordinateSystem(n) -> (t, b): adjoint of n from adj_t, adj_b.
TR_HD F3 coordsys_vjp(F3 n, F3 adj_t, F3 adj_b) {
  F3 t, b;
  coordinate_system(n, &t, &b);
  F3 adj_n = cross(t, adj_b);
  adj_t += cross(adj_b, n);
  bool cond = fabsf(n.x) > fabsf(n.y);
  if (cond) {
    float q = n.x * n.x + n.z * n.z;
    float il = 1.0f / sqrtf(maxf(q, 1e-3f));
    adj_n.z += -adj_t.x * il;
    adj_n.x += adj_t.z * il;
    float adj_il = -adj_t.x * n.z + adj_t.z * n.x;
    if (q > 1e-3f) {
      float adj_q = adj_il * (-0.5f) * il * il * il;
      adj_n.x += adj_q * 2.0f * n.x;
      adj_n.z += adj_q * 2.0f * n.z;
    }
  } else {
    float q = n.y * n.y + n.z * n.z;
    float il = 1.0f / sqrtf(maxf(q, 1e-3f));
    adj_n.z += adj_t.y * il;
    adj_n.y += -adj_t.z * il;
    float adj_il = adj_t.y * n.z - adj_t.z * n.y;
    if (q > 1e-3f) {
      float adj_q = adj_il * (-0.5f) * il * il * il;
      adj_n.y += adj_q * 2.0f * n.y;
      adj_n.z += adj_q * 2.0f * n.z;
    }
  }
  return adj_n;
}

// Microfacet::EvalInner VJP (forward in microfacet_inner).  Only called
// when both cosines are positive (otherwise the value and its derivative
// are zero).
TR_HD void microfacet_inner_vjp(F3 wo, F3 wi, float r, F3 kd, F3 F0, F3 adj_f, F3* adj_wo, F3* adj_wi,
                                float* adj_r, F3* adj_kd, F3* adj_F0) {
  const float c1 = -5.55473f, c2 = 6.8316f;
  float nv = wo.z, nl = wi.z;
  F3 hr = wo + wi;
  float hl = length(hr);
  F3 wh = hr / hl;
  float nh = wh.z, vh = dot(wo, wh);
  float alpha = r * r;
  float k = (r + 1.0f) * (r + 1.0f) * 0.125f;
  float a2m1 = alpha * alpha - 1.0f;
  float dt = nh * nh * a2m1 + 1.0f;
  float tmp = alpha / dt;
  float ggx = tmp * tmp * kInvPi;
  float coeff = vh * (vh * c1 - c2);
  float e = tr_exp2(coeff);
  F3 one{1.0f, 1.0f, 1.0f};
  F3 fres = F0 + (one - F0) * e;
  float D1 = nv * (1.0f - k) + k, D2 = nl * (1.0f - k) + k;
  float g1 = nv / D1, g2 = nl / D2, G = g1 * g2;
  float den = 4.0f * fabsf(nv) * fabsf(nl) + 1e-6f;
  F3 spec = (ggx * G) * fres / den;

  *adj_kd = adj_f * kInvPi;
  F3 adj_num = adj_f / den;
  float adj_den = -dot(adj_f, spec) / den;
  float s = dot(adj_num, fres);
  F3 adj_fres = adj_num * (ggx * G);
  float adj_ggx = s * G, adj_G = s * ggx;
  *adj_F0 = adj_fres * (one - F3{e, e, e});
  float adj_e = dot(adj_fres, one - F0);
  float adj_coeff = adj_e * e * 0.693147180559945309f;
  float adj_vh = adj_coeff * (2.0f * c1 * vh - c2);
  float adj_g1 = adj_G * g2, adj_g2 = adj_G * g1;
  float adj_nv = adj_g1 * k / (D1 * D1);
  float adj_k = adj_g1 * (-nv * (1.0f - nv)) / (D1 * D1);
  float adj_nl = adj_g2 * k / (D2 * D2);
  adj_k += adj_g2 * (-nl * (1.0f - nl)) / (D2 * D2);
  float adj_tmp = adj_ggx * 2.0f * tmp * kInvPi;
  float adj_alpha = adj_tmp / dt;
  float adj_dt = -adj_tmp * tmp / dt;
  float adj_nh = adj_dt * 2.0f * nh * a2m1;
  float adj_a2m1 = adj_dt * nh * nh;
  adj_alpha += adj_a2m1 * 2.0f * alpha;
  *adj_r = adj_alpha * 2.0f * r + adj_k * 0.25f * (r + 1.0f);
  adj_nv += adj_den * 4.0f * sgnf(nv) * fabsf(nl);
  adj_nl += adj_den * 4.0f * fabsf(nv) * sgnf(nl);
  F3 adj_wh = F3{0.0f, 0.0f, adj_nh} + adj_vh * wo;
  F3 ahr = normalize_vjp(wh, hl, adj_wh);
  *adj_wo = F3{0.0f, 0.0f, adj_nv} + adj_vh * wh + ahr;
  *adj_wi = F3{0.0f, 0.0f, adj_nl} + ahr;
}

// Adjoint of a BSDF evaluation f(wo_world, wi_world) at shading point `s`
// (frame t, b, ns).  Accumulates into adj_wo / adj_wi / frame adjoints and
// the material slot.
struct FrameAdj {
  F3 t, b, ns;
};

__device__ inline void bsdf_eval_vjp(const SceneDev& S, const Its& s, F3 wo_w, F3 wi_w, F3 adj_f,
                                     F3& adj_wo_w, F3& adj_wi_w, FrameAdj& fa, float* mat_grad) {
  const Material& m = S.mats[s.mat];
  F3 wo = to_local(s, wo_w), wi = to_local(s, wi_w);
  F3 awo{0.0f, 0.0f, 0.0f}, awi{0.0f, 0.0f, 0.0f};
  if (m.type == kMatDiffuse) {
    float inner = diffuse_inner(wo, wi);
    if (inner != 0.0f && m.tex < 0 && s.mat < kMaxMatGrad) {
      F3 ga = adj_f * inner;
      atomicAdd(&mat_grad[s.mat * kMatGradStride + 1], ga.x);
      atomicAdd(&mat_grad[s.mat * kMatGradStride + 2], ga.y);
      atomicAdd(&mat_grad[s.mat * kMatGradStride + 3], ga.z);
    }
    return;  // no directional dependence
  }
  if (!(wo.z > 0.0f && wi.z > 0.0f)) return;
  F3 kd = mat_albedo(S, m, s);
  F3 F0{m.spec[0], m.spec[1], m.spec[2]};
  float ar;
  F3 akd, aF0;
  microfacet_inner_vjp(wo, wi, m.roughness, kd, F0, adj_f, &awo, &awi, &ar, &akd, &aF0);
  if (s.mat < kMaxMatGrad) {
    float* g = &mat_grad[s.mat * kMatGradStride];
    atomicAdd(&g[0], ar);
    if (m.tex < 0) {
      atomicAdd(&g[1], akd.x); atomicAdd(&g[2], akd.y); atomicAdd(&g[3], akd.z);
    }
    atomicAdd(&g[4], aF0.x); atomicAdd(&g[5], aF0.y); atomicAdd(&g[6], aF0.z);
  }
  // local = (v.t, v.b, v.ns)
  adj_wo_w += awo.x * s.t + awo.y * s.b + awo.z * s.ns;
  adj_wi_w += awi.x * s.t + awi.y * s.b + awi.z * s.ns;
  fa.t += awo.x * wo_w + awi.x * wi_w;
  fa.b += awo.y * wo_w + awi.y * wi_w;
  fa.ns += awo.z * wo_w + awi.z * wi_w;
}

// Forward value of the BSDF (same as bsdf_eval) -- kept next to its VJP.
__device__ inline F3 bsdf_value(const SceneDev& S, const Its& s, F3 wo_w, F3 wi_w) {
  return bsdf_eval(S, s, wo_w, wi_w);
}

__device__ inline void tri_flush(const SceneDev& S, int32_t tri, const TriAdj& a, float* dpos) {
  if (tri < 0) return;
  if (!(S.shade[tri].flags & kTriDiff)) return;
  uint32_t v[3] = {S.idx[3 * tri], S.idx[3 * tri + 1], S.idx[3 * tri + 2]};
  F3 aa[3] = {a.a0, a.a1, a.a2};
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (aa[j].x != 0.0f) atomicAdd(&dpos[3 * v[j]], aa[j].x);
    if (aa[j].y != 0.0f) atomicAdd(&dpos[3 * v[j] + 1], aa[j].y);
    if (aa[j].z != 0.0f) atomicAdd(&dpos[3 * v[j] + 2], aa[j].z);
  }
}

__device__ inline void tri_verts(const SceneDev& S, int32_t tri, F3& p0, F3& p1, F3& p2) {
  p0 = load3(S.pos, S.idx[3 * tri]);
  p1 = load3(S.pos, S.idx[3 * tri + 1]);
  p2 = load3(S.pos, S.idx[3 * tri + 2]);
}

// Backward sweep over a recorded path.  g = dL/dI of the pixel times 1/spp.
__device__ void path_backward(const SceneDev& S, const CameraDev& C, const PathParams& P, const PathRec& R,
                              F3 g, float* dpos, float* mat_grad) {
  const F3 zero{0.0f, 0.0f, 0.0f};
  const int K = R.K;
  if (K < 0) return;
  const F3 eye{C.eye[0], C.eye[1], C.eye[2]};
  F3 Q = zero;                 // adjoint of beta_{i+1}
  TriAdj A_next, A_cur, A_prev;
  tri_adj_zero(A_next);
  tri_adj_zero(A_cur);
  int32_t tri_next = -1;
  for (int i = K; i >= 0; --i) {
    tri_adj_zero(A_prev);
    const int32_t ti = R.tri[i];
    F3 p0, p1, p2;
    tri_verts(S, ti, p0, p1, p2);
    FaceGeom fg = face_geom(p0, p1, p2);
    const float u = R.bu[i], v = R.bv[i], w = 1.0f - u - v;
    Its its = post_intersect(S, ti, u, v);
    const F3 xi = its.p;
    F3 xprev;
    if (i == 0) {
      xprev = eye;
    } else {
      F3 q0, q1, q2;
      tri_verts(S, R.tri[i - 1], q0, q1, q2);
      float uu = R.bu[i - 1], vv = R.bv[i - 1];
      xprev = (1.0f - uu - vv) * q0 + uu * q1 + vv * q2;
    }
    F3 dinr = xi - xprev;
    float dinl = length(dinr);
    F3 din = dinr / dinl;
    F3 wo = -din;
    const F3 beta = R.beta[i];
    F3 adj_x = zero, adj_wo = zero, adj_ng = zero, adj_xprev = zero;
    float adj_area = 0.0f;
    FrameAdj fa{zero, zero, zero};
    F3 a_val = zero, t_val = zero;
    // ---- NEE term a_i ----
    if (i < P.max_bounces) {
      F3 nc = R.neec[i];
      if (nc.x != 0.0f || nc.y != 0.0f || nc.z != 0.0f) {
        const int32_t lt = R.ltri[i];
        F3 q0, q1, q2;
        tri_verts(S, lt, q0, q1, q2);
        FaceGeom lg = face_geom(q0, q1, q2);
        float lbu = R.lbu[i], lbv = R.lbv[i], lbw = 1.0f - lbu - lbv;
        F3 pL = lbw * q0 + lbu * q1 + lbv * q2;
        F3 wr = pL - xi;
        float d2 = len2(wr), dist = sqrtf(d2);
        F3 wl = wr / dist;
        float dsh = dot(its.ns, wl);
        float absSh = fabsf(dsh);
        float q = dot(lg.n, -wl);
        float GL = fabsf(q) / d2;
        F3 f = bsdf_value(S, its, wo, wl);
        a_val = f * absSh * GL * nc;
        F3 wn = (g * beta) * nc;
        F3 adj_f = wn * (absSh * GL);
        float sv = dot(wn, f);
        float adj_absSh = sv * GL, adj_GL = sv * absSh, adj_JL = sv * absSh * GL;
        F3 adj_wl = zero;
        bsdf_eval_vjp(S, its, wo, wl, adj_f, adj_wo, adj_wl, fa, mat_grad);
        float adj_dsh = adj_absSh * sgnf(dsh);
        fa.ns += adj_dsh * wl;
        adj_wl += adj_dsh * its.ns;
        float adj_q = adj_GL * sgnf(q) / d2;
        float adj_d2 = -adj_GL * GL / d2;
        F3 adj_nL = adj_q * (-wl);
        adj_wl += adj_q * (-lg.n);
        F3 adj_wr = normalize_vjp(wl, dist, adj_wl) + (2.0f * adj_d2) * wr;
        adj_x -= adj_wr;
        if (S.shade[lt].flags & kTriDiff) {
          TriAdj la;
          tri_adj_zero(la);
          la.a0 += lbw * adj_wr; la.a1 += lbu * adj_wr; la.a2 += lbv * adj_wr;
          face_vjp(lg, adj_nL, adj_JL / lg.area, la);
          tri_flush(S, lt, la, dpos);
        }
      }
    }
    // ---- transfer t_i of segment i -> i+1 ----
    if (i < K && (R.tflag >> i & 1u)) {
      const int32_t tn = R.tri[i + 1];
      F3 r0, r1, r2;
      tri_verts(S, tn, r0, r1, r2);
      FaceGeom ng = face_geom(r0, r1, r2);
      float nu = R.bu[i + 1], nvv = R.bv[i + 1], nw = 1.0f - nu - nvv;
      F3 xn = nw * r0 + nu * r1 + nvv * r2;
      F3 dr = xn - xi;
      float dd2 = len2(dr), dl = sqrtf(dd2);
      F3 d = dr / dl;
      float qn = dot(ng.n, -d);
      float G = fabsf(qn) / dd2;
      F3 f = bsdf_value(S, its, wo, d);
      float dc = dot(its.ns, d);
      float absc = fabsf(dc);
      float inv_pp = 1.0f / R.ppdf[i];
      t_val = f * (absc * G * inv_pp);
      F3 gT = beta * Q;
      F3 adj_f = gT * (absc * G * inv_pp);
      float sv = dot(gT, f) * inv_pp;
      float adj_absc = sv * G, adj_G = sv * absc, adj_J = sv * absc * G;
      F3 adj_d = zero;
      bsdf_eval_vjp(S, its, wo, d, adj_f, adj_wo, adj_d, fa, mat_grad);
      float adj_dc = adj_absc * sgnf(dc);
      fa.ns += adj_dc * d;
      adj_d += adj_dc * its.ns;
      float adj_qn = adj_G * sgnf(qn) / dd2;
      float adj_dd2 = -adj_G * G / dd2;
      F3 adj_nn = adj_qn * (-d);
      adj_d += adj_qn * (-ng.n);
      F3 adj_dr = normalize_vjp(d, dl, adj_d) + (2.0f * adj_dd2) * dr;
      adj_x -= adj_dr;
      A_next.a0 += nw * adj_dr; A_next.a1 += nu * adj_dr; A_next.a2 += nvv * adj_dr;
      face_vjp(ng, adj_nn, adj_J / ng.area, A_next);
    }
    // ---- adjoint of beta_i ----
    F3 e_i = (i == 0) ? R.e0c : R.emc[i - 1];
    F3 Qi = g * (e_i + a_val) + t_val * Q;
    if (i == 0) {
      // beta_0 = G_cam / detach(G_cam) * area / detach(area)
      float ab = Qi.x + Qi.y + Qi.z;
      F3 cdr = xi - eye;
      float dist = length(cdr);
      F3 cd = cdr / dist;
      const F3 cdir{C.dir[0], C.dir[1], C.dir[2]};
      float cosy = dot(cd, cdir);
      float cosx = dot(-cd, its.ng);
      float den = cosy * cosy * cosy * (dist * dist);
      float Gc = fabsf(cosx) / den;
      float adj_Gc = ab / Gc;
      adj_area += ab / fg.area;
      float adj_cosx = adj_Gc * sgnf(cosx) / den;
      float adj_cosy = adj_Gc * (-3.0f) * Gc / cosy;
      float adj_dist = adj_Gc * (-2.0f) * Gc / dist;
      F3 adj_cd = -adj_cosx * its.ng + adj_cosy * cdir;
      adj_ng += -adj_cosx * cd;
      adj_x += normalize_vjp(cd, dist, adj_cd) + adj_dist * cd;
    }
    Q = Qi;
    // ---- incoming direction wo = -normalize(x_i - x_prev) ----
    {
      F3 adj_din = -adj_wo;
      F3 adj_dr = normalize_vjp(din, dinl, adj_din);
      adj_x += adj_dr;
      adj_xprev -= adj_dr;
    }
    // ---- shading frame -> normal -> triangle ----
    F3 adj_ns = coordsys_vjp(its.ns, fa.t, fa.b) + fa.ns;
    adj_ng += adj_ns;  // flat shading: ns = ng
    A_cur.a0 += w * adj_x; A_cur.a1 += u * adj_x; A_cur.a2 += v * adj_x;
    face_vjp(fg, adj_ng, adj_area, A_cur);
    if (i > 0) {
      float uu = R.bu[i - 1], vv = R.bv[i - 1], ww = 1.0f - uu - vv;
      A_prev.a0 += ww * adj_xprev; A_prev.a1 += uu * adj_xprev; A_prev.a2 += vv * adj_xprev;
    }
    // triangle i+1 is complete
    tri_flush(S, tri_next, A_next, dpos);
    tri_next = ti;
    A_next = A_cur;
    A_cur = A_prev;
  }
  tri_flush(S, tri_next, A_next, dpos);
}

}  // namespace trc
