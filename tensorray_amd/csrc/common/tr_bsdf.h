// Shading-point reconstruction, BSDFs, emitter and discrete sampling, camera
// rays -- the per-sample math of the interior integrator, written once for
// the HIP kernels (and host unit tests).  Operation order follows the
// reference expressions cited at each function.
#pragma once
#include "tr_layout.h"

namespace trc {

// Intersection record after Scene::PostIntersect (Scene.cpp:668-730).
struct Its {
  F3 p;      // w*p0 + u*p1 + v*p2
  F3 ng;     // face normal
  F3 ns;     // shading normal (flat: ng; smooth: barycentric vertex normals, not renormalized)
  F3 t, b;   // CoordinateSystem(ns)
  float tu, tv;  // texture coordinate (0 when the scene has none)
  int32_t tri, mat, emitter;
};

// CoordinateSystem (LibTensorRay/Renderer/Utils.cpp:118-140).
TR_HD void coordinate_system(F3 n, F3* t, F3* b) {
  bool cond = fabsf(n.x) > fabsf(n.y);
  float il = cond ? 1.0f / sqrtf(maxf(n.x * n.x + n.z * n.z, 1e-3f))
                  : 1.0f / sqrtf(maxf(n.y * n.y + n.z * n.z, 1e-3f));
  *t = cond ? F3{-n.z * il, 0.0f, n.x * il} : F3{0.0f, n.z * il, -n.y * il};
  *b = cross(n, *t);
}

TR_HD F3 to_local(const Its& s, F3 v) { return F3{dot(v, s.t), dot(v, s.b), dot(v, s.ns)}; }
TR_HD F3 to_world(const Its& s, F3 l) { return l.x * s.t + l.y * s.b + l.z * s.ns; }

TR_HD F3 load3(const float* p, uint32_t i) { return F3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// Scene::PostIntersect (Scene.cpp:668-730) for triangle `tri` and
// barycentrics (u, v).
TR_HD Its post_intersect(const SceneDev& S, int32_t tri, float u, float v) {
  Its s;
  uint32_t i0 = S.idx[3 * tri], i1 = S.idx[3 * tri + 1], i2 = S.idx[3 * tri + 2];
  F3 p0 = load3(S.pos, i0), p1 = load3(S.pos, i1), p2 = load3(S.pos, i2);
  float w = 1.0f - u - v;
  s.p = w * p0 + u * p1 + v * p2;
  const TriShade& sh = S.shade[tri];
  s.ng = F3{sh.n[0], sh.n[1], sh.n[2]};
  if (sh.flags & kTriSmooth) {
    F3 n0 = load3(S.vnrm, i0), n1 = load3(S.vnrm, i1), n2 = load3(S.vnrm, i2);
    s.ns = w * n0 + u * n1 + v * n2;
  } else {
    s.ns = s.ng;
  }
  if (S.uv) {
    uint32_t t0 = S.uv_idx[3 * tri], t1 = S.uv_idx[3 * tri + 1], t2 = S.uv_idx[3 * tri + 2];
    s.tu = w * S.uv[2 * t0] + u * S.uv[2 * t1] + v * S.uv[2 * t2];
    s.tv = w * S.uv[2 * t0 + 1] + u * S.uv[2 * t1 + 1] + v * S.uv[2 * t2 + 1];
  } else {
    s.tu = 0.0f;
    s.tv = 0.0f;
  }
  coordinate_system(s.ns, &s.t, &s.b);
  s.tri = tri;
  s.mat = sh.material;
  s.emitter = sh.emitter;
  return s;
}

// ImageTexture::Eval (BSDF.cpp:72-98): nearest texel at floor(frac(u) * (w-1)).
TR_HD size_t tex_index(const TexDesc& t, float u, float v) {
  float fu = u - floorf(u), fv = v - floorf(v);
  int x = (int)(fu * (float)(t.w - 1)), y = (int)(fv * (float)(t.h - 1));
  return (size_t)y * t.w + x;
}
TR_HD F3 tex_eval(const TexDesc& t, float u, float v) {
  const float* p = t.texels + 3 * tex_index(t, u, v);
  return F3{p[0], p[1], p[2]};
}

TR_HD F3 mat_albedo(const SceneDev& S, const Material& m, const Its& s) {
  if (m.tex >= 0) return tex_eval(S.texs[m.tex], s.tu, s.tv);
  return F3{m.albedo[0], m.albedo[1], m.albedo[2]};
}

// ---------------------------------------------------------------------------
// Sampling primitives (Utils.cpp:70-115)
// ---------------------------------------------------------------------------
TR_HD F3 cosine_sample_hemisphere(float s0, float s1) {
  float r1 = 2.0f * s0 - 1.0f, r2 = 2.0f * s1 - 1.0f;
  bool cond = (r1 * r1) > (r2 * r2);
  float r = cond ? r1 : r2;
  bool zero = (r1 == 0.0f) && (r2 == 0.0f);
  float phi = 0.0f;
  if (!zero) phi = cond ? kPiOver4 * (r2 / r1) : kPiOver2 - (r1 / r2) * kPiOver4;
  float r_ = zero ? 0.0f : r;
  float sp, cp;
  tr_sincos(phi, &sp, &cp);
  float x = r_ * cp, y = r_ * sp;
  float z = sqrtf(maxf(1e-4f, 1.0f - x * x - y * y));
  return F3{x, y, z};
}

TR_HD void square_to_disk_concentric(float s0, float s1, float* ox, float* oy) {
  float x = 2.0f * s0 - 1.0f, y = 2.0f * s1 - 1.0f;
  bool zero = (x == 0.0f && y == 0.0f);
  bool q13 = fabsf(x) < fabsf(y);
  float r = q13 ? y : x;
  float rp = q13 ? x : y;
  float phi = 0.0f;
  if (!zero) {
    float phi_ = (0.25f * kPi) * rp / r;
    phi = q13 ? 0.5f * kPi - phi_ : phi_;
  }
  float sp, cp;
  tr_sincos(phi, &sp, &cp);
  *ox = r * cp;
  *oy = r * sp;
}

// ---------------------------------------------------------------------------
// Lambertian (Diffuse.cpp:37-85)
// ---------------------------------------------------------------------------
TR_HD float diffuse_inner(F3 wo, F3 wi) {
  bool refl = wo.z > 0.0f && (wo.z * wi.z) > 0.0f;
  return kInvPi * (refl ? 1.0f : 0.0f);
}
TR_HD float diffuse_pdf(F3 wo, F3 wi) {
  bool refl = wo.z > 0.0f && (wo.z * wi.z) > 0.0f;
  return fabsf(wi.z) / kPi * (refl ? 1.0f : 0.0f);
}
TR_HD bool diffuse_dir_valid(const Its& s, F3 w) {
  return dot(s.ns, w) > kEdgeEpsilon && dot(s.ng, w) > kEdgeEpsilon;
}

// ---------------------------------------------------------------------------
// GGX helpers (BSDF.cpp:100-302) and the "Microfacet" BRDF (Microfacet.cpp)
// ---------------------------------------------------------------------------
TR_HD float tan_theta2(F3 v) {
  float c2 = v.z * v.z;
  return maxf(1.0f - c2, 0.0f) / c2;
}
TR_HD float sin_theta(F3 v) { return sqrtf(maxf(1.0f - v.z * v.z, 0.0f)); }
TR_HD float cos_phi(F3 v) {
  float st = sin_theta(v);
  return st == 0.0f ? 1.0f : clampf(v.x / st, -1.0f, 1.0f);
}
TR_HD float sin_phi(F3 v) {
  float st = sin_theta(v);
  return st == 0.0f ? 0.0f : clampf(v.y / st, -1.0f, 1.0f);
}

// BSDF::GGX_D (BSDF.cpp:232-241)
TR_HD float ggx_d(F3 wh, float alpha) {
  float t2 = tan_theta2(wh);
  bool ok = t2 < 3.402823466e38f;
  float c2 = wh.z * wh.z;
  float e = 1.0f / (alpha * alpha) * t2;
  float root = alpha * c2 * (1.0f + e);
  float D = kInvPi / (root * root);
  return ok ? D : 0.0f;
}

// BSDF::Smith_G (BSDF.cpp:243-256), restated as written (the tangent is the
// half vector's, the root squared).
TR_HD float smith_g(F3 v, F3 wh, float alpha) {
  float t2 = tan_theta2(wh);
  bool ok = t2 < 3.402823466e38f;
  float root = (alpha * alpha) * t2;
  float G = 2.0f / (1.0f + sqrtf(1.0f + root * root));
  G = ok ? G : 0.0f;
  float voh = dot(v, wh);
  G = (t2 == 0.0f) ? 1.0f : G;
  G = (voh * v.z) <= 0.0f ? 0.0f : G;
  return G;
}

// BSDF::GGX_VNDF_Pdf (BSDF.cpp:263-270)
TR_HD float ggx_vndf_pdf(F3 v, F3 wh, float alpha) {
  float D = ggx_d(wh, alpha);
  float G = smith_g(v, wh, alpha);
  float voh = fabsf(dot(v, wh));
  return D * G * voh / fabsf(v.z);
}

// BSDF::GGX_VNDF_SampleVisible11 (BSDF.cpp:290-302)
TR_HD void ggx_sample_visible11(float cos_t, float s0, float s1, float* sx, float* sy) {
  float px, py;
  square_to_disk_concentric(s0, s1, &px, &py);
  float s = 0.5f * (1.0f + cos_t);
  float x = px;
  float a = safe_sqrt(1.0f - x * x);
  float y = (1.0f - s) * a + s * py;
  float z = safe_sqrt(1.0f - (px * px + py * py));
  float sin_t = safe_sqrt(1.0f - cos_t * cos_t);
  float norm = 1.0f / (sin_t * y + cos_t * z);
  *sx = (cos_t * y - sin_t * z) * norm;
  *sy = x * norm;
}

// BSDF::GGX_VNDF_Sample (BSDF.cpp:272-288)
TR_HD F3 ggx_vndf_sample(F3 v, float alpha, float s0, float s1) {
  F3 V = normalize(F3{v.x * alpha, v.y * alpha, v.z * 1.0f});
  float sp = sin_phi(V), cp = cos_phi(V), ct = V.z;
  float slx, sly;
  ggx_sample_visible11(ct, s0, s1, &slx, &sly);
  float sx = (cp * slx - sp * sly) * alpha;
  float sy = (sp * slx + cp * sly) * alpha;
  return normalize(F3{-sx, -sy, 1.0f});
}

// Microfacet::EvalInner (Microfacet.cpp:37-73): value per channel, given
// the half vector wh = normalize(wo + wi).
TR_HD F3 microfacet_inner_wh(F3 wo, F3 wi, F3 wh, float r, F3 kd, F3 F0) {
  float nv = wo.z, nl = wi.z;
  bool active = nv > 0.0f && nl > 0.0f;
  F3 diff = kd * kInvPi;
  float nh = wh.z, vh = dot(wo, wh);
  float alpha = r * r;
  float k = (r + 1.0f) * (r + 1.0f) * 0.125f;
  float tmp = alpha / (nh * nh * (alpha * alpha - 1.0f) + 1.0f);
  float ggx = tmp * tmp * kInvPi;
  float coeff = vh * (vh * -5.55473f - 6.8316f);
  float e = tr_exp2(coeff);
  F3 fres = F0 + (F3{1.0f, 1.0f, 1.0f} - F0) * e;
  float g1 = nv / (nv * (1.0f - k) + k);
  float g2 = nl / (nl * (1.0f - k) + k);
  float G = g1 * g2;
  F3 num = ggx * G * fres;
  float den = 4.0f * fabsf(nv) * fabsf(nl);
  F3 spec = num / (den + 1e-6f);
  F3 val = diff + spec;
  return active ? val : F3{0.0f, 0.0f, 0.0f};
}
TR_HD F3 microfacet_inner(F3 wo, F3 wi, float r, F3 kd, F3 F0) {
  return microfacet_inner_wh(wo, wi, normalize(wo + wi), r, kd, F0);
}

// BSDF::FresnelConductor (BSDF.cpp:304-323), one channel.
TR_HD float fresnel_conductor(float ci, float eta, float k) {
  float ci2 = ci * ci;
  float si2 = 1.0f - ci2;
  float si4 = si2 * si2;
  float eta2 = eta * eta, k2 = k * k;
  float temp1 = eta2 - k2 - si2;
  float a2pb2 = safe_sqrt(temp1 * temp1 + k2 * eta2 * 4.0f);
  float a = safe_sqrt((a2pb2 + temp1) * 0.5f);
  float term1 = a2pb2 + ci2;
  float term2 = a * (2.0f * ci);
  float Rs2 = (term1 - term2) / (term1 + term2);
  float term3 = a2pb2 * ci2 + si4, term4 = term2 * si2;
  float Rp2 = Rs2 * (term3 - term4) / (term3 + term4);
  return 0.5f * (Rp2 + Rs2);
}

// RoughConductor::EvalInner (Roughconductor.cpp:37-57): GGX with alpha =
// roughness (no squaring, unlike Microfacet), conductor Fresnel.
TR_HD F3 roughconductor_inner_wh(F3 wo, F3 wi, F3 wh, float r, F3 eta, F3 k) {
  const F3 zero{0.0f, 0.0f, 0.0f};
  if (!(wo.z > 0.0f && wi.z > 0.0f)) return zero;
  float D = ggx_d(wh, r);
  if (D == 0.0f) return zero;
  float G = smith_g(wo, wh, r) * smith_g(wi, wh, r);
  float coso = dot(wo, wh);
  F3 Fr{fresnel_conductor(coso, eta.x, k.x), fresnel_conductor(coso, eta.y, k.y),
        fresnel_conductor(coso, eta.z, k.z)};
  float den = 4.0f * fabsf(wi.z) * fabsf(wo.z);
  return Fr * D * G / den;
}
TR_HD F3 roughconductor_inner(F3 wo, F3 wi, float r, F3 eta, F3 k) {
  return roughconductor_inner_wh(wo, wi, normalize(wo + wi), r, eta, k);
}

// Microfacet::PdfInner (Microfacet.cpp:152-165); RoughConductor::PdfInner
// (Roughconductor.cpp:107-121) is the same expression.
TR_HD float microfacet_pdf_wh(F3 wo, F3 wi, F3 wh, float r) {
  float mp = ggx_vndf_pdf(wo, wh, r);
  float dwh = 1.0f / (4.0f * dot(wi, wh));
  float pdf = fabsf(mp * dwh);
  bool refl = wo.z > 0.0f && mp != 0.0f && wi.z > 0.0f;
  return refl ? pdf : 0.0f;
}
TR_HD float microfacet_pdf(F3 wo, F3 wi, float r) { return microfacet_pdf_wh(wo, wi, normalize(wo + wi), r); }

// BSDF::OutDirValid (BSDF.h:133) for non-Lambertian materials.
TR_HD bool base_dir_valid(const Its& s, F3 w) { return dot(s.ns, w) * dot(s.ng, w) > 0.0f; }

// ---------------------------------------------------------------------------
// Material dispatch (per-material loop of Integrator.cpp:185-296)
// ---------------------------------------------------------------------------
// BSDF::Eval (BSDF.cpp:172-195) without the shading-normal correction.
TR_HD F3 bsdf_eval(const SceneDev& S, const Its& s, F3 wo_w, F3 wi_w) {
  const Material& m = S.mats[s.mat];
  F3 wo = to_local(s, wo_w), wi = to_local(s, wi_w);
  if (m.type == kMatDiffuse) {
    return mat_albedo(S, m, s) * diffuse_inner(wo, wi);
  }
  if (m.type == kMatRoughConductor)
    return roughconductor_inner(wo, wi, m.roughness, F3{m.eta[0], m.eta[1], m.eta[2]}, F3{m.k[0], m.k[1], m.k[2]});
  return microfacet_inner(wo, wi, m.roughness, mat_albedo(S, m, s), F3{m.spec[0], m.spec[1], m.spec[2]});
}

TR_HD float bsdf_pdf(const SceneDev& S, const Its& s, F3 wo_w, F3 wi_w) {
  const Material& m = S.mats[s.mat];
  F3 wo = to_local(s, wo_w), wi = to_local(s, wi_w);
  if (m.type == kMatDiffuse) return diffuse_pdf(wo, wi);
  return microfacet_pdf(wo, wi, m.roughness);
}

// bsdf_eval and bsdf_pdf of one pair of directions (the same values as
// the two calls), sharing the local frame and the half vector.
TR_HD F3 bsdf_eval_pdf(const SceneDev& S, const Its& s, F3 wo_w, F3 wi_w, float* pdf) {
  const Material& m = S.mats[s.mat];
  F3 wo = to_local(s, wo_w), wi = to_local(s, wi_w);
  if (m.type == kMatDiffuse) {
    *pdf = diffuse_pdf(wo, wi);
    return mat_albedo(S, m, s) * diffuse_inner(wo, wi);
  }
  const F3 wh = normalize(wo + wi);
  *pdf = microfacet_pdf_wh(wo, wi, wh, m.roughness);
  if (m.type == kMatRoughConductor)
    return roughconductor_inner_wh(wo, wi, wh, m.roughness, F3{m.eta[0], m.eta[1], m.eta[2]},
                                   F3{m.k[0], m.k[1], m.k[2]});
  return microfacet_inner_wh(wo, wi, wh, m.roughness, mat_albedo(S, m, s), F3{m.spec[0], m.spec[1], m.spec[2]});
}

TR_HD bool bsdf_dir_valid(const SceneDev& S, const Its& s, F3 w) {
  return S.mats[s.mat].type == kMatDiffuse ? diffuse_dir_valid(s, w) : base_dir_valid(s, w);
}

// SampleOnly (Diffuse.cpp:66-78, Microfacet.cpp:127-150,
// Roughconductor.cpp:83-105; the last two are the same VNDF draw): world
// direction and its solid-angle pdf.
TR_HD F3 bsdf_sample(const SceneDev& S, const Its& s, F3 wo_w, float s0, float s1, float* pdf) {
  const Material& m = S.mats[s.mat];
  F3 wo = to_local(s, wo_w);
  if (m.type == kMatDiffuse) {
    F3 wi = cosine_sample_hemisphere(s0, s1);
    *pdf = diffuse_pdf(wo, wi);
    return to_world(s, wi);
  }
  bool refl = wo.z > 0.0f;
  F3 wh = ggx_vndf_sample(wo, m.roughness, s0, s1);
  float mp = ggx_vndf_pdf(wo, wh, m.roughness);
  F3 wi = reflect(-wo, wh);
  refl = refl && mp != 0.0f && wi.z > 0.0f;
  float dwh = 1.0f / (4.0f * dot(wi, wh));
  float p = fabsf(mp * dwh);
  *pdf = refl ? p : 0.0f;
  return to_world(s, wi);
}

// ---------------------------------------------------------------------------
// Discrete distribution (Distribution.cpp:37-58, Distribution.cu:37-48)
// ---------------------------------------------------------------------------
// First index with cdf[i] >= u, clamped to [0, n-1].
TR_HD int32_t cdf_search(const float* cdf, int32_t n, float u) {
  int32_t first = 0, count = n;
  while (count > 0) {
    int32_t step = count / 2;
    int32_t mid = first + step;
    if (cdf[mid] < u) {
      first = mid + 1;
      count -= step + 1;
    } else {
      count = step;
    }
  }
  return first < n - 1 ? first : n - 1;
}
TR_HD float cdf_reuse(const float* cdf, int32_t i, float u) {
  float lo = i > 0 ? cdf[i - 1] : 0.0f;
  return (u - lo) / (cdf[i] - lo);
}

// AreaLight::Sample (Light.cpp:146-175)
struct LightSample {
  F3 p, n;
  float pdf;       // pmf / area (area measure)
  int32_t tri;     // global triangle id
  int32_t emitter;
  float bu, bv;    // barycentrics of the sampled point
};

TR_HD LightSample light_sample(const SceneDev& S, float u0, float u1) {
  LightSample ls;
  int32_t i = cdf_search(S.light_cdf, S.n_light, u0);
  float pmf = S.light_f[i] / (S.light_integral * (float)S.n_light);
  float reuse = cdf_reuse(S.light_cdf, i, u0);
  float t = sqrtf(reuse);
  float bu = 1.0f - t;
  float bv = t * u1;
  float bw = 1.0f - bu - bv;
  int32_t tri = S.light_tri[i];
  F3 q0 = load3(S.pos, S.idx[3 * tri]), q1 = load3(S.pos, S.idx[3 * tri + 1]), q2 = load3(S.pos, S.idx[3 * tri + 2]);
  ls.p = bw * q0 + bu * q1 + bv * q2;
  const TriShade& sh = S.shade[tri];
  ls.n = F3{sh.n[0], sh.n[1], sh.n[2]};
  ls.pdf = pmf / sh.area;
  ls.tri = tri;
  ls.emitter = sh.emitter;
  ls.bu = bu;
  ls.bv = bv;
  return ls;
}

// AreaLight::Pdf (Light.cpp:177-184) for a BSDF-sampled emitter hit.
TR_HD float light_pdf_tri(const SceneDev& S, int32_t tri) {
  const TriShade& sh = S.shade[tri];
  float pmf = S.light_f[sh.emit_tri] / (S.light_integral * (float)S.n_light);
  return pmf / sh.area;
}

// AreaLight::Eval (Light.cpp:134-144): one-sided emission.
TR_HD F3 emitted(const SceneDev& S, int32_t emitter, F3 ng, F3 w) {
  float on = dot(w, ng) > 0.0f ? 1.0f : 0.0f;
  const float* I = S.emit_rad + 3 * emitter;
  return F3{I[0], I[1], I[2]} * on;
}

// ---------------------------------------------------------------------------
// Camera (Camera.cpp:224-303, 367-384)
// ---------------------------------------------------------------------------
TR_HD F3 camera_dir(const CameraDev& c, float rx, float ry) {
  F3 v = normalize(F3{c.ax * rx + c.bx, c.ay * ry + c.by, 1.0f});
  F3 R = F3{c.right[0], c.right[1], c.right[2]};
  F3 U = F3{c.up[0], c.up[1], c.up[2]};
  F3 D = F3{c.dir[0], c.dir[1], c.dir[2]};
  return v.x * R + v.y * U + v.z * D;
}

// Camera-space coordinates of a world point.
TR_HD F3 camera_local(const CameraDev& c, F3 p) {
  F3 d = p - F3{c.eye[0], c.eye[1], c.eye[2]};
  return F3{dot(d, F3{c.right[0], c.right[1], c.right[2]}), dot(d, F3{c.up[0], c.up[1], c.up[2]}),
            dot(d, F3{c.dir[0], c.dir[1], c.dir[2]})};
}

// Camera::sampleDirect: image-plane coordinate q in [0,1)^2 of a world
// point, validity (inside the image and in front of the camera) and the
// sensor importance 1 / (d^2 cos^3) * inv_area.
struct SensorSample {
  float qx, qy;
  bool valid;
  int32_t pixel;
  float value;
};
TR_HD SensorSample sensor_sample(const CameraDev& c, F3 p) {
  SensorSample s;
  F3 l = camera_local(c, p);
  // x_cam / z = ratio * tan * (1 - 2 qx)  =>  qx = (1 - x/(z ratio tan)) / 2
  float sx = l.x / (l.z * c.ratio * c.tan_half);
  float sy = l.y / (l.z * c.tan_half);
  s.qx = (1.0f - sx) * 0.5f;
  s.qy = (1.0f - sy) * 0.5f;
  s.valid = l.z > 0.0f && s.qx >= 0.0f && s.qx < 1.0f && s.qy >= 0.0f && s.qy < 1.0f;
  int px = (int)floorf(s.qx * (float)c.width), py = (int)floorf(s.qy * (float)c.height);
  s.pixel = py * c.width + px;
  F3 d = p - F3{c.eye[0], c.eye[1], c.eye[2]};
  float d2 = len2(d);
  F3 dn = d / sqrtf(d2);
  float ct = dot(F3{c.dir[0], c.dir[1], c.dir[2]}, dn);
  float ic = 1.0f / ct;
  s.value = 1.0f / d2 * ic * ic * ic * c.inv_area;
  return s;
}

}  // namespace trc
