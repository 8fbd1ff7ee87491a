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

// Material parameter gradient slots: [roughness, albedo rgb, spec rgb, pad];
// for a rough conductor [roughness, eta rgb, k rgb, pad].
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
// Returns the value microfacet_inner(wo, wi, r, kd, F0) of an active pair
// (wo.z > 0, wi.z > 0), computed from the same intermediates.
TR_HD F3 microfacet_inner_vjp(F3 wo, F3 wi, float r, F3 kd, F3 F0, F3 adj_f, F3* adj_wo, F3* adj_wi,
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
  return kd * kInvPi + spec;
}

// d tan^2(theta) / d z for tan^2 = max(1 - z^2, 0) / z^2.
TR_HD float tan2_dz(float z) {
  float c2 = z * z;
  return (1.0f - c2) > 0.0f ? -2.0f * z / (c2 * c2) : 0.0f;
}

// VJP of ggx_d(wh, a) (value D) into wh.z and a.
TR_HD void ggx_d_vjp(float z, float a, float D, float gD, float* g_z, float* g_a) {
  float c2 = z * z;
  float om = 1.0f - c2;
  float t2 = (om > 0.0f ? om : 0.0f) / c2;
  float e = 1.0f / (a * a) * t2;
  float root = a * c2 * (1.0f + e);
  float g_root = gD * (-2.0f * D / root);
  float g_e = g_root * a * c2;
  *g_a += g_root * c2 * (1.0f + e) + g_e * (-2.0f * t2 / (a * a * a));
  float g_t2 = g_e / (a * a);
  *g_z += g_root * a * (1.0f + e) * 2.0f * z + g_t2 * tan2_dz(z);
}

// VJP of smith_g(v, wh, a) into wh.z and a (v enters only through masks).
TR_HD void smith_g_vjp(F3 v, F3 wh, float a, float gG, float* g_z, float* g_a) {
  float t2 = tan_theta2(wh);
  if (!(t2 < 3.402823466e38f) || t2 == 0.0f || (dot(v, wh) * v.z) <= 0.0f) return;
  float root = (a * a) * t2;
  float sq = sqrtf(1.0f + root * root);
  float g_root = gG * (-2.0f / ((1.0f + sq) * (1.0f + sq))) * root / sq;
  *g_a += g_root * 2.0f * a * t2;
  *g_z += g_root * (a * a) * tan2_dz(wh.z);
}

// VJP of fresnel_conductor(ci, eta, k) (BSDF.cpp:304-323).
TR_HD void fresnel_conductor_vjp(float ci, float eta, float k, float gF, float* g_ci, float* g_eta, float* g_k) {
  float ci2 = ci * ci;
  float si2 = 1.0f - ci2;
  float si4 = si2 * si2;
  float eta2 = eta * eta, k2 = k * k;
  float temp1 = eta2 - k2 - si2;
  float q = temp1 * temp1 + k2 * eta2 * 4.0f;
  float a2pb2 = safe_sqrt(q);
  float p = (a2pb2 + temp1) * 0.5f;
  float a = safe_sqrt(p);
  float term1 = a2pb2 + ci2;
  float term2 = a * (2.0f * ci);
  float U = term1 - term2, W = term1 + term2;
  float Rs2 = U / W;
  float term3 = a2pb2 * ci2 + si4, term4 = term2 * si2;
  float X = term3 - term4, Y = term3 + term4;
  float Rp2 = Rs2 * X / Y;
  float g_Rp2 = 0.5f * gF, g_Rs2 = 0.5f * gF;
  g_Rs2 += g_Rp2 * X / Y;
  float g_X = g_Rp2 * Rs2 / Y, g_Y = -g_Rp2 * Rp2 / Y;
  float g_term3 = g_X + g_Y, g_term4 = g_Y - g_X;
  float g_U = g_Rs2 / W, g_W = -g_Rs2 * Rs2 / W;
  float g_term1 = g_U + g_W, g_term2 = g_W - g_U;
  g_term2 += g_term4 * si2;
  float g_si2 = g_term4 * term2;
  float g_a2pb2 = g_term3 * ci2 + g_term1;
  float g_ci2 = g_term3 * a2pb2 + g_term1;
  float g_si4 = g_term3;
  float g_a = g_term2 * 2.0f * ci;
  float gci = g_term2 * 2.0f * a;
  float g_p = p > 0.0f ? g_a * 0.5f / a : 0.0f;
  g_a2pb2 += 0.5f * g_p;
  float g_temp1 = 0.5f * g_p;
  float g_q = q > 0.0f ? g_a2pb2 * 0.5f / a2pb2 : 0.0f;
  g_temp1 += g_q * 2.0f * temp1;
  float g_k2 = g_q * 4.0f * eta2, g_eta2 = g_q * 4.0f * k2;
  g_eta2 += g_temp1;
  g_k2 -= g_temp1;
  g_si2 -= g_temp1;
  g_si2 += g_si4 * 2.0f * si2;
  g_ci2 -= g_si2;
  gci += g_ci2 * 2.0f * ci;
  *g_ci += gci;
  *g_eta = g_eta2 * 2.0f * eta;
  *g_k = g_k2 * 2.0f * k;
}

// VJP of roughconductor_inner.
TR_HD void roughconductor_inner_vjp(F3 wo, F3 wi, float r, F3 eta, F3 kk, F3 adj_f, F3* adj_wo, F3* adj_wi,
                                    float* adj_r, F3* adj_eta, F3* adj_k) {
  const F3 zero{0.0f, 0.0f, 0.0f};
  *adj_wo = *adj_wi = *adj_eta = *adj_k = zero;
  *adj_r = 0.0f;
  if (!(wo.z > 0.0f && wi.z > 0.0f)) return;
  F3 hr = wo + wi;
  float hl = length(hr);
  F3 wh = hr / hl;
  float D = ggx_d(wh, r);
  if (D == 0.0f) return;
  float Go = smith_g(wo, wh, r), Gi = smith_g(wi, wh, r), G = Go * Gi;
  float coso = dot(wo, wh);
  float den = 4.0f * fabsf(wi.z) * fabsf(wo.z);
  F3 Fr{fresnel_conductor(coso, eta.x, kk.x), fresnel_conductor(coso, eta.y, kk.y),
        fresnel_conductor(coso, eta.z, kk.z)};
  F3 gFr = adj_f * (D * G / den);
  float sFr = dot(adj_f, Fr);
  float gD = sFr * G / den, gG = sFr * D / den;
  float gden = -sFr * D * G / (den * den);
  float g_coso = 0.0f;
  fresnel_conductor_vjp(coso, eta.x, kk.x, gFr.x, &g_coso, &adj_eta->x, &adj_k->x);
  fresnel_conductor_vjp(coso, eta.y, kk.y, gFr.y, &g_coso, &adj_eta->y, &adj_k->y);
  fresnel_conductor_vjp(coso, eta.z, kk.z, gFr.z, &g_coso, &adj_eta->z, &adj_k->z);
  float g_whz = 0.0f, g_r = 0.0f;
  ggx_d_vjp(wh.z, r, D, gD, &g_whz, &g_r);
  smith_g_vjp(wo, wh, r, gG * Gi, &g_whz, &g_r);
  smith_g_vjp(wi, wh, r, gG * Go, &g_whz, &g_r);
  *adj_r = g_r;
  F3 g_wh = F3{0.0f, 0.0f, g_whz} + g_coso * wo;
  F3 g_hr = normalize_vjp(wh, hl, g_wh);
  *adj_wo = F3{0.0f, 0.0f, gden * 4.0f * wi.z} + g_coso * wh + g_hr;
  *adj_wi = F3{0.0f, 0.0f, gden * 4.0f * wo.z} + g_hr;
}

// Adjoint of a BSDF evaluation f(wo_world, wi_world) at shading point `s`
// (frame t, b, ns).  Accumulates into adj_wo / adj_wi / frame adjoints and
// the material slot.
struct FrameAdj {
  F3 t, b, ns;
};

// Where the backward sweep sends the adjoints of the scene parameters.
// GradSink: float atomics into the gradient buffers -- vertex positions
// dpos [V][3], the block's material accumulators, texels (BSDF::DiffTexture,
// BSDF.cpp:350-356: the nearest-texel lookup scatters into one texel) and
// the vertex-normal adjoints of smooth-shaded meshes.
struct GradSink {
  float* dpos;
  float* mat_grad;
  float* adj_vn;
  float* dtex;
  __device__ bool has_vn() const { return adj_vn != nullptr; }
  __device__ void pos(uint32_t v, F3 a) {
    if (a.x != 0.0f) atomicAdd(&dpos[3 * v], a.x);
    if (a.y != 0.0f) atomicAdd(&dpos[3 * v + 1], a.y);
    if (a.z != 0.0f) atomicAdd(&dpos[3 * v + 2], a.z);
  }
  __device__ void mat(int i, float g) { atomicAdd(&mat_grad[i], g); }
  __device__ void tex(const SceneDev& S, const Material& m, const Its& s, F3 g) {
    if (!dtex) return;
    const TexDesc& t = S.texs[m.tex];
    float* d = dtex + t.grad_off + 3 * tex_index(t, s.tu, s.tv);
    if (g.x != 0.0f) atomicAdd(&d[0], g.x);
    if (g.y != 0.0f) atomicAdd(&d[1], g.y);
    if (g.z != 0.0f) atomicAdd(&d[2], g.z);
  }
  __device__ void vn(uint32_t v, F3 c) {
    if (c.x != 0.0f) atomicAdd(&adj_vn[3 * v], c.x);
    if (c.y != 0.0f) atomicAdd(&adj_vn[3 * v + 1], c.y);
    if (c.z != 0.0f) atomicAdd(&adj_vn[3 * v + 2], c.z);
  }
};

// TangentSink: forward-mode derivative of one path's contribution along a
// parameter tangent -- tpos [V][3] = d position / d theta, tmat [M][8] =
// d material slot / d theta (or null), tvn [V][3] = d vertex normal / d theta
// (or null): every adjoint is dotted with its tangent and summed into acc.
// Texels carry no tangent.
struct TangentSink {
  const float* tpos;
  const float* tmat;
  const float* tvn;
  float acc;
  __device__ bool has_vn() const { return tvn != nullptr; }
  __device__ void pos(uint32_t v, F3 a) { acc += dot(a, load3(tpos, v)); }
  __device__ void mat(int i, float g) {
    if (tmat) acc += g * tmat[i];
  }
  __device__ void tex(const SceneDev&, const Material&, const Its&, F3) {}
  __device__ void vn(uint32_t v, F3 c) { acc += dot(c, load3(tvn, v)); }
};

// VJP of bsdf_eval(S, s, wo_w, wi_w) for the adjoint adj_f of its value;
// returns the value (the same as bsdf_eval's).
template <class Sink>
__device__ inline F3 bsdf_eval_vjp(const SceneDev& S, const Its& s, F3 wo_w, F3 wi_w, F3 adj_f,
                                   F3& adj_wo_w, F3& adj_wi_w, FrameAdj& fa, Sink& sk) {
  const Material& m = S.mats[s.mat];
  F3 wo = to_local(s, wo_w), wi = to_local(s, wi_w);
  F3 awo{0.0f, 0.0f, 0.0f}, awi{0.0f, 0.0f, 0.0f};
  F3 f;
  if (m.type == kMatDiffuse) {
    float inner = diffuse_inner(wo, wi);
    f = mat_albedo(S, m, s) * inner;
    if (inner == 0.0f) return f;
    F3 ga = adj_f * inner;
    if (m.tex >= 0) {
      sk.tex(S, m, s, ga);
    } else if (s.mat < kMaxMatGrad) {
      sk.mat(s.mat * kMatGradStride + 1, ga.x);
      sk.mat(s.mat * kMatGradStride + 2, ga.y);
      sk.mat(s.mat * kMatGradStride + 3, ga.z);
    }
    return f;  // no directional dependence
  }
  if (!(wo.z > 0.0f && wi.z > 0.0f)) return F3{0.0f, 0.0f, 0.0f};
  if (m.type == kMatRoughConductor) {
    f = roughconductor_inner(wo, wi, m.roughness, F3{m.eta[0], m.eta[1], m.eta[2]}, F3{m.k[0], m.k[1], m.k[2]});
    float ar;
    F3 aeta, ak;
    roughconductor_inner_vjp(wo, wi, m.roughness, F3{m.eta[0], m.eta[1], m.eta[2]}, F3{m.k[0], m.k[1], m.k[2]},
                             adj_f, &awo, &awi, &ar, &aeta, &ak);
    if (s.mat < kMaxMatGrad) {
      const int gb = s.mat * kMatGradStride;
      sk.mat(gb, ar);
      sk.mat(gb + 1, aeta.x); sk.mat(gb + 2, aeta.y); sk.mat(gb + 3, aeta.z);
      sk.mat(gb + 4, ak.x); sk.mat(gb + 5, ak.y); sk.mat(gb + 6, ak.z);
    }
  } else {
    F3 kd = mat_albedo(S, m, s);
    F3 F0{m.spec[0], m.spec[1], m.spec[2]};
    float ar;
    F3 akd, aF0;
    f = microfacet_inner_vjp(wo, wi, m.roughness, kd, F0, adj_f, &awo, &awi, &ar, &akd, &aF0);
    if (m.tex >= 0) sk.tex(S, m, s, akd);
    if (s.mat < kMaxMatGrad) {
      const int gb = s.mat * kMatGradStride;
      sk.mat(gb, ar);
      if (m.tex < 0) {
        sk.mat(gb + 1, akd.x); sk.mat(gb + 2, akd.y); sk.mat(gb + 3, akd.z);
      }
      sk.mat(gb + 4, aF0.x); sk.mat(gb + 5, aF0.y); sk.mat(gb + 6, aF0.z);
    }
  }
  // local = (v.t, v.b, v.ns)
  adj_wo_w += awo.x * s.t + awo.y * s.b + awo.z * s.ns;
  adj_wi_w += awi.x * s.t + awi.y * s.b + awi.z * s.ns;
  fa.t += awo.x * wo_w + awi.x * wi_w;
  fa.b += awo.y * wo_w + awi.y * wi_w;
  fa.ns += awo.z * wo_w + awi.z * wi_w;
  return f;
}


template <class Sink>
__device__ inline void tri_flush(const SceneDev& S, int32_t tri, const TriAdj& a, Sink& sk) {
  if (tri < 0) return;
  if (!(S.shade[tri].flags & kTriDiff)) return;
  uint32_t v[3] = {S.idx[3 * tri], S.idx[3 * tri + 1], S.idx[3 * tri + 2]};
  F3 aa[3] = {a.a0, a.a1, a.a2};
#pragma unroll
  for (int j = 0; j < 3; ++j) sk.pos(v[j], aa[j]);
}

__device__ inline void tri_verts(const SceneDev& S, int32_t tri, F3& p0, F3& p1, F3& p2) {
  p0 = load3(S.pos, S.idx[3 * tri]);
  p1 = load3(S.pos, S.idx[3 * tri + 1]);
  p2 = load3(S.pos, S.idx[3 * tri + 2]);
}

// Backward sweep over a recorded path.  g = dL/dI of the pixel times 1/spp;
// the parameter adjoints go to `sk` (GradSink or TangentSink).
template <class Sink>
__device__ void path_backward(const SceneDev& S, const CameraDev& C, const PathParams& P, const PathRec& R,
                              F3 g, Sink& sk) {
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
        F3 wn = (g * beta) * nc;
        F3 adj_f = wn * (absSh * GL);
        F3 adj_wl = zero;
        const F3 f = bsdf_eval_vjp(S, its, wo, wl, adj_f, adj_wo, adj_wl, fa, sk);
        a_val = f * absSh * GL * nc;
        float sv = dot(wn, f);
        float adj_absSh = sv * GL, adj_GL = sv * absSh, adj_JL = sv * absSh * GL;
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
          tri_flush(S, lt, la, sk);
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
      float dc = dot(its.ns, d);
      float absc = fabsf(dc);
      float inv_pp = 1.0f / R.ppdf[i];
      F3 gT = beta * Q;
      F3 adj_f = gT * (absc * G * inv_pp);
      F3 adj_d = zero;
      const F3 f = bsdf_eval_vjp(S, its, wo, d, adj_f, adj_wo, adj_d, fa, sk);
      t_val = f * (absc * G * inv_pp);
      float sv = dot(gT, f) * inv_pp;
      float adj_absc = sv * G, adj_G = sv * absc, adj_J = sv * absc * G;
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
    const uint32_t tflags = S.shade[ti].flags;
    if (!(tflags & kTriSmooth)) {
      adj_ng += adj_ns;  // flat shading: ns = ng
    } else if (sk.has_vn() && (tflags & kTriDiff)) {
      // ns = w n0 + u n1 + v n2 (vertex normals, not renormalized)
      const uint32_t k0 = S.idx[3 * ti], k1 = S.idx[3 * ti + 1], k2 = S.idx[3 * ti + 2];
      const F3 c[3] = {w * adj_ns, u * adj_ns, v * adj_ns};
      const uint32_t ks[3] = {k0, k1, k2};
#pragma unroll
      for (int j = 0; j < 3; ++j) sk.vn(ks[j], c[j]);
    }
    A_cur.a0 += w * adj_x; A_cur.a1 += u * adj_x; A_cur.a2 += v * adj_x;
    face_vjp(fg, adj_ng, adj_area, A_cur);
    if (i > 0) {
      float uu = R.bu[i - 1], vv = R.bv[i - 1], ww = 1.0f - uu - vv;
      A_prev.a0 += ww * adj_xprev; A_prev.a1 += uu * adj_xprev; A_prev.a2 += vv * adj_xprev;
    }
    // triangle i+1 is complete
    tri_flush(S, tri_next, A_next, sk);
    tri_next = ti;
    A_next = A_cur;
    A_cur = A_prev;
  }
  tri_flush(S, tri_next, A_next, sk);
}

// Adjoint of the smooth-shading vertex normals (TriangleMesh::
// ComputeVertexNormal, Primitive.cpp:323-340): n_v = normalize(a / c),
// a = sum of the c incident face normals.  One thread per vertex folds
// adj_vn[v] into the face normals' vertices and clears adj_vn[v].
__device__ inline void vertex_normal_backward(const SceneDev& S, uint32_t vtx, float* adj_vn, float* dpos) {
  const uint32_t b = S.vf_off[vtx], e = S.vf_off[vtx + 1];
  const F3 g{adj_vn[3 * vtx], adj_vn[3 * vtx + 1], adj_vn[3 * vtx + 2]};
  if (b == e || (g.x == 0.0f && g.y == 0.0f && g.z == 0.0f)) return;
  adj_vn[3 * vtx] = adj_vn[3 * vtx + 1] = adj_vn[3 * vtx + 2] = 0.0f;
  F3 a{0.0f, 0.0f, 0.0f};
  for (uint32_t j = b; j < e; ++j) {
    const TriShade& sh = S.shade[S.vf_list[j]];
    a = a + F3{sh.n[0], sh.n[1], sh.n[2]};
  }
  const float c = (float)(e - b);
  const F3 q = a / c;
  const float l = length(q);
  const F3 adj_a = normalize_vjp(q / l, l, g) / c;
  for (uint32_t j = b; j < e; ++j) {
    const int32_t t = (int32_t)S.vf_list[j];
    F3 p0, p1, p2;
    tri_verts(S, t, p0, p1, p2);
    TriAdj ta;
    tri_adj_zero(ta);
    face_vjp(face_geom(p0, p1, p2), adj_a, 0.0f, ta);
    GradSink sk{dpos, nullptr, nullptr, nullptr};
    tri_flush(S, t, ta, sk);
  }
}

// Tangent of the smooth-shading vertex normal of `vtx` along tpos (forward
// mode of n_v = normalize(a / c), a = sum of the unit normals of the c
// incident faces, each n_f = normalize(cross(p1 - p0, p2 - p0))).
__device__ inline F3 vertex_normal_tangent(const SceneDev& S, uint32_t vtx, const float* tpos) {
  const uint32_t b = S.vf_off[vtx], e = S.vf_off[vtx + 1];
  const F3 zero{0.0f, 0.0f, 0.0f};
  if (b == e) return zero;
  F3 a = zero, da = zero;
  for (uint32_t j = b; j < e; ++j) {
    const int32_t t = (int32_t)S.vf_list[j];
    F3 p0, p1, p2;
    tri_verts(S, t, p0, p1, p2);
    const F3 t0 = load3(tpos, S.idx[3 * t]), t1 = load3(tpos, S.idx[3 * t + 1]), t2 = load3(tpos, S.idx[3 * t + 2]);
    const FaceGeom g = face_geom(p0, p1, p2);
    const F3 dc = cross(t1 - t0, g.e1) + cross(g.e0, t2 - t0);
    const TriShade& sh = S.shade[t];
    a = a + F3{sh.n[0], sh.n[1], sh.n[2]};
    da = da + (dc - g.n * dot(g.n, dc)) / g.len;
  }
  const float c = (float)(e - b);
  const F3 q = a / c, dq = da / c;
  const float l = length(q);
  const F3 nv = q / l;
  return (dq - nv * dot(nv, dq)) / l;
}

}  // namespace trc
