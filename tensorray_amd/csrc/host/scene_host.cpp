// Host-side scene flattening, derived data, generators and the camera
// reduction.  See scene_host.h for the reference counterparts.
#include "scene_host.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <thread>

namespace trc {

namespace {

inline F3 vget(const std::vector<float>& p, uint32_t i) { return F3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

}  // namespace

std::vector<EdgeHost> build_edge_list(const std::vector<uint32_t>& idx) {
  // ObjMesh::processEdges (EDXUtil/Graphics/ObjMesh.cpp:584-632): one entry
  // per undirected edge, keyed (min v, max v) in lexicographic order; the
  // value list starts with the opposite vertex of the first incident face,
  // followed by the incident faces.
  std::map<std::pair<uint32_t, uint32_t>, std::vector<int64_t>> m;
  const size_t T = idx.size() / 3;
  for (size_t i = 0; i < T; ++i) {
    for (int j = 0; j < 3; ++j) {
      uint32_t a = idx[3 * i + j], b = idx[3 * i + (j + 1) % 3], c = idx[3 * i + (j + 2) % 3];
      auto key = a < b ? std::make_pair(a, b) : std::make_pair(b, a);
      auto it = m.find(key);
      if (it == m.end()) it = m.emplace(key, std::vector<int64_t>{(int64_t)c}).first;
      it->second.push_back((int64_t)i);
    }
  }
  std::vector<EdgeHost> out;
  out.reserve(m.size());
  for (auto& kv : m) {
    EdgeHost e;
    e.v0 = kv.first.first;
    e.v1 = kv.first.second;
    e.v2 = (uint32_t)kv.second[0];
    e.f0 = (int32_t)kv.second[1];
    e.f1 = kv.second.size() == 3 ? (int32_t)kv.second[2] : -1;
    out.push_back(e);
  }
  return out;
}

void SceneHost::flatten() {
  if (meshes.empty()) throw std::runtime_error("scene has no shapes");
  if (materials.empty()) throw std::runtime_error("scene has no BSDFs");
  pos.clear(); idx.clear(); uv.clear(); uv_idx.clear();
  mesh_vert_offset.clear(); mesh_tri_offset.clear(); mesh_edge_offset.clear();
  edges.clear();
  emit_rad.clear();
  bool any_uv = false;
  for (auto& m : meshes) any_uv = any_uv || !m.uv.empty();
  uint32_t voff = 0, toff = 0, uvoff = 0;
  int32_t emitter_id = 0;
  std::vector<int32_t> tri_mesh;
  for (size_t mi = 0; mi < meshes.size(); ++mi) {
    MeshHost& m = meshes[mi];
    if (m.pos.size() % 3 || m.idx.size() % 3 || m.idx.empty())
      throw std::runtime_error("mesh " + std::to_string(mi) + ": malformed buffers");
    const uint32_t V = (uint32_t)(m.pos.size() / 3), T = (uint32_t)(m.idx.size() / 3);
    for (uint32_t i : m.idx)
      if (i >= V) throw std::runtime_error("mesh " + std::to_string(mi) + ": index out of range");
    if (m.material < 0 || m.material >= (int32_t)materials.size())
      throw std::runtime_error("mesh " + std::to_string(mi) + ": material out of range");
    if (m.is_emitter) {
      m.emitter = emitter_id++;
      emit_rad.insert(emit_rad.end(), m.radiance, m.radiance + 3);
    } else {
      m.emitter = -1;
    }
    mesh_vert_offset.push_back(voff);
    mesh_tri_offset.push_back(toff);
    pos.insert(pos.end(), m.pos.begin(), m.pos.end());
    for (uint32_t i : m.idx) idx.push_back(i + voff);
    if (any_uv) {
      if (!m.uv.empty()) {
        if (m.uv_idx.size() != m.idx.size()) throw std::runtime_error("uv index size mismatch");
        uv.insert(uv.end(), m.uv.begin(), m.uv.end());
        for (uint32_t i : m.uv_idx) uv_idx.push_back(i + uvoff);
        uvoff += (uint32_t)(m.uv.size() / 2);
      } else {
        // meshes without texture coordinates read uv (0, 0)
        uv.push_back(0.0f); uv.push_back(0.0f);
        for (uint32_t t = 0; t < 3 * T; ++t) uv_idx.push_back(uvoff);
        uvoff += 1;
      }
    }
    for (uint32_t t = 0; t < T; ++t) tri_mesh.push_back((int32_t)mi);
    // edges of this mesh, global ids
    mesh_edge_offset.push_back((uint32_t)edges.size());
    for (auto e : build_edge_list(m.idx)) {
      e.v0 += voff; e.v1 += voff; e.v2 += voff;
      e.f0 += (int32_t)toff;
      if (e.f1 >= 0) e.f1 += (int32_t)toff;
      edges.push_back(e);
    }
    voff += V;
    toff += T;
  }
  shade.assign(toff, TriShade{});
  for (uint32_t t = 0; t < toff; ++t) {
    const MeshHost& m = meshes[tri_mesh[t]];
    shade[t].material = m.material;
    shade[t].emitter = m.emitter;
    shade[t].emit_tri = -1;
    shade[t].flags = (m.smooth ? kTriSmooth : 0u) | (m.diff ? kTriDiff : 0u);
  }
  // emitter triangle table (Scene.cpp:63-98 mTriIdToEmitTriIdBuffer)
  light_tri.clear();
  for (uint32_t t = 0; t < toff; ++t)
    if (shade[t].emitter >= 0) {
      shade[t].emit_tri = (int32_t)light_tri.size();
      light_tri.push_back((int32_t)t);
    }
  update_positions(nullptr);
}

void SceneHost::update_positions(const float* new_pos) {
  if (new_pos) std::copy(new_pos, new_pos + pos.size(), pos.begin());
  const uint32_t T = (uint32_t)(idx.size() / 3), V = (uint32_t)(pos.size() / 3);
  // TriangleMesh::ComputeFaceNormalAndDistrib (Primitive.cpp:299-321)
  for (uint32_t t = 0; t < T; ++t) {
    F3 p0 = vget(pos, idx[3 * t]), p1 = vget(pos, idx[3 * t + 1]), p2 = vget(pos, idx[3 * t + 2]);
    F3 c = cross(p1 - p0, p2 - p0);
    float l = length(c);
    F3 n = c / l;
    shade[t].n[0] = n.x; shade[t].n[1] = n.y; shade[t].n[2] = n.z;
    shade[t].area = l * 0.5f;
  }
  // TriangleMesh::ComputeVertexNormal (Primitive.cpp:323-340): normalized
  // mean of the incident face normals.
  vnrm.assign(3 * (size_t)V, 0.0f);
  std::vector<float> cnt(V, 0.0f);
  for (uint32_t t = 0; t < T; ++t) {
    if (!(shade[t].flags & kTriSmooth)) continue;
    for (int j = 0; j < 3; ++j) {
      uint32_t v = idx[3 * t + j];
      vnrm[3 * v] += shade[t].n[0];
      vnrm[3 * v + 1] += shade[t].n[1];
      vnrm[3 * v + 2] += shade[t].n[2];
      cnt[v] += 1.0f;
    }
  }
  for (uint32_t v = 0; v < V; ++v) {
    if (cnt[v] == 0.0f) continue;
    F3 a = F3{vnrm[3 * v] / cnt[v], vnrm[3 * v + 1] / cnt[v], vnrm[3 * v + 2] / cnt[v]};
    F3 n = normalize(a);
    vnrm[3 * v] = n.x; vnrm[3 * v + 1] = n.y; vnrm[3 * v + 2] = n.z;
  }
  // Scene::ConfigureLight (Scene.cpp:471-520): f = luminance * area per
  // emitter triangle; Distribution1D::SetFunction (Distribution.cpp:45-58):
  // cdf = inclusive_scan(f) / N, integral = cdf[N-1], cdf /= integral.
  const int32_t NL = (int32_t)light_tri.size();
  light_f.assign(NL, 0.0f);
  light_cdf.assign(NL, 0.0f);
  light_integral = 0.0f;
  if (NL > 0) {
    for (int32_t i = 0; i < NL; ++i) {
      const TriShade& s = shade[light_tri[i]];
      const float* I = &emit_rad[3 * s.emitter];
      float lum = 0.212671f * I[0] + 0.715160f * I[1] + 0.072169f * I[2];
      light_f[i] = lum * s.area;
    }
    const float inv_n = 1.0f / (float)NL;
    float run = 0.0f;
    for (int32_t i = 0; i < NL; ++i) {
      run += light_f[i];
      light_cdf[i] = run * inv_n;
    }
    light_integral = light_cdf[NL - 1];
    if (light_integral > 0.0f)
      for (int32_t i = 0; i < NL; ++i) light_cdf[i] = light_cdf[i] / light_integral;
  }
  // secondary edges depend on the face normals
  select_sec_edges(*this, sec_edges);
  sec_integral = edge_distribution(*this, sec_edges, sec_len, sec_cdf);
}

namespace {
inline F3 face_n(const SceneHost& s, int32_t f) {
  if (f < 0) return F3{0.0f, 0.0f, 0.0f};
  return F3{s.shade[f].n[0], s.shade[f].n[1], s.shade[f].n[2]};
}

// The edges e with keep(e), in edge order.  Large lists are filtered in
// contiguous chunks on up to 16 threads and concatenated in chunk order, so
// the result is the same as the sequential filter.
template <class Keep>
void filter_edges(const std::vector<EdgeHost>& edges, std::vector<EdgeDev>& out, Keep keep) {
  out.clear();
  const size_t n = edges.size();
  unsigned nt = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (n < (size_t)1 << 16) nt = 1;
  std::vector<std::vector<EdgeDev>> part(nt);
  auto run = [&](unsigned t) {
    const size_t lo = n * t / nt, hi = n * (t + 1) / nt;
    for (size_t i = lo; i < hi; ++i) {
      const EdgeHost& e = edges[i];
      if (keep(e)) part[t].push_back(EdgeDev{e.v0, e.v1, e.v2, e.f0, e.f1, {0, 0, 0}});
    }
  };
  if (nt == 1) {
    run(0);
  } else {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) th.emplace_back(run, t);
    for (auto& x : th) x.join();
  }
  size_t total = 0;
  for (auto& p : part) total += p.size();
  out.reserve(total);
  for (auto& p : part) out.insert(out.end(), p.begin(), p.end());
}
}  // namespace

void select_sec_edges(const SceneHost& s, std::vector<EdgeDev>& out) {
  filter_edges(s.edges, out, [&](const EdgeHost& e) {
    if (e.f1 < 0) return true;
    F3 n0 = face_n(s, e.f0), n1 = face_n(s, e.f1);
    F3 p0 = vget(s.pos, e.v0), p2 = vget(s.pos, e.v2);
    bool noncoplanar = dot(n0, n1) < 1.0f - 1e-6f;
    bool convex = dot(n1, normalize(p2 - p0)) < -kEdgeEpsilon;
    return noncoplanar && convex;
  });
}

float edge_distribution(const SceneHost& s, const std::vector<EdgeDev>& e, std::vector<float>& len,
                        std::vector<float>& cdf) {
  const size_t n = e.size();
  len.resize(n);
  cdf.resize(n);
  if (n == 0) return 0.0f;
  for (size_t i = 0; i < n; ++i) len[i] = length(vget(s.pos, e[i].v1) - vget(s.pos, e[i].v0));
  // inclusive prefix in the blocked order: sequential within chunks of
  // kCdfChunk edges, chunk sums sequential, chunk offset + local prefix
  const float inv = 1.0f / (float)n;
  const size_t nc = (n + kCdfChunk - 1) / kCdfChunk;
  float off = 0.0f;
  for (size_t c = 0; c < nc; ++c) {
    float local = 0.0f;
    const size_t end = std::min(n, (c + 1) * (size_t)kCdfChunk);
    for (size_t i = c * kCdfChunk; i < end; ++i) {
      local = local + len[i];
      cdf[i] = (off + local) * inv;
    }
    off = off + local;
  }
  float integral = cdf[n - 1];
  if (integral > 0.0f)
    for (size_t i = 0; i < n; ++i) cdf[i] = cdf[i] / integral;
  return integral;
}

void SceneHost::build_bvh() { bvh_build(pos, idx, nodes, tris, tri_order, bvh_stats, &refit_plan); }

void SceneHost::refit_bvh() { bvh_refit(pos, idx, tri_order, nodes, tris); }

CameraDev make_camera(const float pos[3], const float target[3], const float up[3], int w, int h,
                      float fov_deg) {
  // Matrix::LookAt (EDXUtil/Math/Matrix.cpp:181-209): dir = normalize(target
  // - eye), right = normalize(cross(normalize(up), dir)), up' = cross(dir,
  // right).  Matrix::Perspective (Matrix.cpp:211-220) with the vertical fov
  // and BaseCamera::Resize's screen-to-raster map (BaseCamera.cpp:73-88)
  // give, on the near plane z = 1:
  //   x_cam = ratio * tan(fov/2) * (1 - 2 rx / W),  y_cam = tan(fov/2) * (1 - 2 ry / H).
  // Computed in double, stored as float.
  double e[3] = {pos[0], pos[1], pos[2]};
  double d[3] = {(double)target[0] - e[0], (double)target[1] - e[1], (double)target[2] - e[2]};
  double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (double& x : d) x /= dl;
  double u[3] = {up[0], up[1], up[2]};
  double ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (double& x : u) x /= ul;
  double r[3] = {u[1] * d[2] - u[2] * d[1], u[2] * d[0] - u[0] * d[2], u[0] * d[1] - u[1] * d[0]};
  double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  for (double& x : r) x /= rl;
  double nu[3] = {d[1] * r[2] - d[2] * r[1], d[2] * r[0] - d[0] * r[2], d[0] * r[1] - d[1] * r[0]};
  double tanh = std::tan(fov_deg * 3.14159265358979323846 / 180.0 / 2.0);
  double ratio = (double)w / (double)h;
  CameraDev c{};
  for (int k = 0; k < 3; ++k) {
    c.eye[k] = (float)e[k];
    c.right[k] = (float)r[k];
    c.up[k] = (float)nu[k];
    c.dir[k] = (float)d[k];
  }
  c.ax = (float)(-2.0 * ratio * tanh / w);
  c.bx = (float)(ratio * tanh);
  c.ay = (float)(-2.0 * tanh / h);
  c.by = (float)tanh;
  c.tan_half = (float)tanh;
  c.ratio = (float)ratio;
  // Camera::Resize (Camera.cpp:120-132): 1 / (plane width * plane height) *
  // |centre|^2, with the centre of the z = 1 plane at distance 1.
  c.inv_area = (float)(1.0 / ((2.0 * ratio * tanh) * (2.0 * tanh)));
  c.width = w;
  c.height = h;
  return c;
}

}  // namespace trc
