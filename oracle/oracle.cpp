// oracle.cpp -- CPU restatement of libumbreon's per-ray path.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker. The MI355X product never calls
// it. Built with -ffp-contract=off so every float operation rounds exactly as
// written, in the order the reference writes it.
//
// Each function cites the reference file:line it restates (paths relative to
// the CueMol/umbreon source tree). The template parameter T selects the
// precision: float follows the reference's float-op order, double is the FP64
// oracle of the same integrator.
//
// Pinning (which check fixes which function) is listed at the top of
// tests/test_oracle.py. Functions whose exact arithmetic only a live Embree
// could fix (curve t/u/Ng, triangle edge watertightness, interpolation
// rounding, tie order among equal-t hits) are "parity unpinned": their
// oracle here is checked against closed forms and FP64 brute force only.

#include "oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <type_traits>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// Vector math (scene.hpp:20-76). Same operation order as the reference.
template <class T>
struct V3 {
  T x, y, z;
};
template <class T>
inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T>
inline V3<T> operator*(T s, V3<T> a) { return a * s; }
template <class T>
inline V3<T> neg(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T>
inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T>
inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
// scene.hpp:35-38
template <class T>
inline V3<T> normalize(V3<T> a) {
  T l = length(a);
  return l > T(0) ? a * (T(1) / l) : a;
}
// scene.hpp:44-49
template <class T>
inline V3<T> safeNormalize(V3<T> a) {
  T s = T(1) / std::sqrt(std::fmax(dot(a, a), T(0x1.0p-126)));
  return a * s;
}
// scene.hpp:53-57
template <class T>
inline V3<T> safeNormalize(V3<T> a, V3<T> fallback) {
  T d = dot(a, a);
  if (d < T(1.0e-18)) return fallback;
  return a * (T(1) / std::sqrt(d));
}
template <class T>
struct Frame {
  V3<T> t, b, n;
};
// scene.hpp:69-76
template <class T>
inline Frame<T> frameFromNormal(V3<T> n) {
  const V3<T> seed0{T(0), n.z, -n.y};
  const V3<T> seed1{-n.z, T(0), n.x};
  const V3<T> t = safeNormalize(std::fabs(n.x) < std::fabs(n.y) ? seed0 : seed1,
                                V3<T>{T(1), T(0), T(0)});
  const V3<T> b = cross(n, t);
  return {t, b, n};
}
template <class T>
inline T radiansT(T deg) { return deg * T(0.01745329252f); }

template <class T>
inline V3<T> v3(const float* p) { return {T(p[0]), T(p[1]), T(p[2])}; }

// shading.hpp:27-30
template <class T>
inline V3<T> faceForward(V3<T> n, V3<T> rayDir) {
  return (dot(n, rayDir) > T(0)) ? neg(n) : n;
}
// shading.hpp:34-41
template <class T>
inline T blinnExp(T roughness) {
  T r = roughness;
  if (r < T(1e-6f)) r = T(1e-6f);
  T e = T(1) / r;
  if (e < T(1)) e = T(1);
  if (e > T(1.0e6f)) e = T(1.0e6f);
  return e;
}
// secondary_rays.hpp:38-46
template <class T>
inline T selfIntersectEps(V3<T> P, V3<T> dir, T t) {
  const T dirMax = std::fmax(std::fabs(dir.x), std::fmax(std::fabs(dir.y), std::fabs(dir.z)));
  const T epsScale = std::fmax(std::fmax(std::fabs(P.x), std::fabs(P.y)),
                               std::fmax(std::fabs(P.z), dirMax * t));
  return epsScale * T(0x1.0fp-21f);
}
// secondary_rays.hpp:51-58 (TEA, 8 rounds)
inline void tea2(uint32_t& v0, uint32_t& v1) {
  uint32_t sum = 0;
  for (int i = 0; i < 8; ++i) {
    sum += 0x9E3779B9u;
    v0 += ((v1 << 4) + 0xA341316Cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xC8013EA4u);
    v1 += ((v0 << 4) + 0xAD90777Du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7E95761Eu);
  }
}
// secondary_rays.hpp:61
template <class T>
inline T u32ToUnorm(uint32_t u) { return T(u >> 8) * T(0x1.0p-24f); }
// secondary_rays.hpp:66-73 (Malley)
template <class T>
inline V3<T> cosineSampleHemisphere(T u1, T u2, const Frame<T>& f) {
  const T phi = T(6.2831853072f) * u1;
  const T r = std::sqrt(u2);
  const T z = std::sqrt(std::fmax(T(0), T(1) - u2));
  const T x = std::cos(phi) * r;
  const T y = std::sin(phi) * r;
  return f.t * x + f.b * y + f.n * z;
}

// ---------------------------------------------------------------------------
// Scene assembly (scene_build.cpp:16-170, curve_build.cpp:41-267).

struct Mat {
  float ambient, diffuse, specular, roughness, brilliance, phong, phongSize;
  bool metallic;
  float reflection, emission;
  int model;
  // Material::Pbr (scene.hpp:150-166)
  float pbrMetallic, pbrRoughness, pbrSpecular, pbrAnisotropy, pbrRotation;
};
inline Mat toMat(const ubg_material& m) {
  return Mat{m.ambient,        m.diffuse,       m.specular,      m.roughness,       m.brilliance,
             m.phong,          m.phong_size,    m.metallic != 0, m.reflection,      m.emission,
             m.model,          m.pbr_metallic,  m.pbr_roughness, m.pbr_specular,    m.pbr_anisotropy,
             m.pbr_anisotropy_rotation};
}

struct PrimSide {  // BuiltScene side-table entry for one primitive
  float color[4];
  int32_t mat;
  uint16_t group;
  uint8_t fromEdge;
  float opacity1;
  // cylinders: unit p0 -> p1 axis (BuiltScene::cylAxis / cylCapAxis,
  // curve_build.cpp:163-167 baked, :243-245 unbaked), the principled
  // anisotropy tangent
  float axis[3] = {0.0f, 0.0f, 0.0f};
};

struct Geometry {
  // Mesh (baked per instance): positions/normals/colors per vertex, index per
  // corner (scene_build.cpp:37-68).
  int64_t baseV = 0, baseT = 0, copies = 1;
  std::vector<float> pos, nrm, col;
  std::vector<uint32_t> idx;
  std::vector<Mat> meshMats;
  std::vector<uint8_t> triMat;
  std::vector<uint16_t> triGroup;
  // Spheres (scene_build.cpp:102-123).
  std::vector<float> sph;  // x y z r
  std::vector<PrimSide> sphSide;
  // Round curves: vertex buffer (x y z r), segment start vertex, flags.
  std::vector<float> rv;
  std::vector<uint32_t> rseg;
  std::vector<uint8_t> rflag;
  std::vector<int32_t> rsrc;
  std::vector<PrimSide> rSide;
  // Cone curves: 8 floats per segment (p0 r p1 r).
  std::vector<float> cone;
  std::vector<PrimSide> cSide;
  std::vector<Mat> primMats;
};

static inline uint32_t cornerVertex(const ubg_scene_desc* d, int64_t corner) {
  return d->index ? d->index[corner] : static_cast<uint32_t>(corner);
}

// scene.hpp:44-49 safeNormalize (no fallback) of p1 - p0, into ps.axis
static void setAxis(PrimSide& ps, const float p0[3], const float p1[3]) {
  const float x = p1[0] - p0[0], y = p1[1] - p0[1], z = p1[2] - p0[2];
  const float sc = 1.0f / std::sqrt(std::fmax(x * x + y * y + z * z, 0x1.0p-126f));
  ps.axis[0] = x * sc;
  ps.axis[1] = y * sc;
  ps.axis[2] = z * sc;
}

// curve_build.cpp:88-199 -- endpoint-hash chaining of `open` cylinders.
static void chainOpenCylinders(const ubg_scene_desc* d, const std::vector<int>& openIdx,
                               const float off[3], Geometry& g) {
  const std::size_t no = openIdx.size();
  const float eps = 1.0e-4f;
  auto keyOf = [&](const float p[3]) {
    return std::array<long long, 3>{std::llround(p[0] / eps), std::llround(p[1] / eps),
                                    std::llround(p[2] / eps)};
  };
  std::vector<std::array<float, 3>> a0(no), a1(no);
  for (std::size_t i = 0; i < no; ++i) {
    const int c = openIdx[i];
    for (int k = 0; k < 3; ++k) {
      a0[i][k] = d->cyl_p0[3 * c + k] + off[k];
      a1[i][k] = d->cyl_p1[3 * c + k] + off[k];
    }
  }
  std::map<std::array<long long, 3>, std::vector<std::pair<int, int>>> inc;
  for (int i = 0; i < static_cast<int>(no); ++i) {
    inc[keyOf(a0[i].data())].push_back({i, 0});
    inc[keyOf(a1[i].data())].push_back({i, 1});
  }
  auto grp = [&](int s) -> uint16_t { return d->cyl_group ? d->cyl_group[openIdx[s]] : 0; };
  auto op1 = [&](int s) -> float { return d->cyl_opacity1 ? d->cyl_opacity1[openIdx[s]] : -1.0f; };
  auto compatible = [&](int s, int t) {
    return grp(s) == grp(t) && (op1(s) < 0.0f) == (op1(t) < 0.0f);
  };
  std::vector<int> next(no, -1), prev(no, -1);
  for (int s = 0; s < static_cast<int>(no); ++s) {
    const auto& ends = inc[keyOf(a1[s].data())];
    if (ends.size() != 2) continue;
    int t = -1;
    for (const auto& e : ends)
      if (e.first != s) t = (e.second == 0) ? e.first : -1;
    if (t >= 0 && t != s && compatible(s, t)) next[s] = t;
  }
  for (int s = 0; s < static_cast<int>(no); ++s)
    if (next[s] >= 0) prev[next[s]] = s;
  std::vector<char> visited(no, 0);
  auto emitChain = [&](int start) {
    std::vector<int> chain;
    for (int s = start; s >= 0 && !visited[s]; s = next[s]) {
      visited[s] = 1;
      chain.push_back(s);
    }
    const uint32_t base = static_cast<uint32_t>(g.rv.size() / 4);
    auto pushV = [&](const std::array<float, 3>& p, float r) {
      g.rv.push_back(p[0]);
      g.rv.push_back(p[1]);
      g.rv.push_back(p[2]);
      g.rv.push_back(r);
    };
    pushV(a0[chain[0]], d->cyl_radius[openIdx[chain[0]]]);
    for (int cs : chain) pushV(a1[cs], d->cyl_radius[openIdx[cs]]);
    const int n = static_cast<int>(chain.size());
    for (int j = 0; j < n; ++j) {
      g.rseg.push_back(base + static_cast<uint32_t>(j));
      uint8_t fl = 0;
      if (j > 0) fl |= 1;
      if (j < n - 1) fl |= 2;
      g.rflag.push_back(fl);
      const int c = openIdx[chain[j]];
      g.rsrc.push_back(c);
      PrimSide ps;
      for (int k = 0; k < 4; ++k) ps.color[k] = d->cyl_color[4 * c + k];
      ps.mat = d->cyl_material ? d->cyl_material[c] : 0;
      ps.group = d->cyl_group ? d->cyl_group[c] : 0;
      ps.fromEdge = d->cyl_from_edge ? d->cyl_from_edge[c] : 0;
      ps.opacity1 = d->cyl_opacity1 ? d->cyl_opacity1[c] : -1.0f;
      setAxis(ps, a0[chain[j]].data(), a1[chain[j]].data());
      g.rSide.push_back(ps);
    }
  };
  for (int s = 0; s < static_cast<int>(no); ++s)
    if (prev[s] < 0 && !visited[s]) emitChain(s);
  for (int s = 0; s < static_cast<int>(no); ++s)
    if (!visited[s]) emitChain(s);
}

static void buildGeometry(const ubg_scene_desc* d, Geometry& g) {
  std::vector<std::array<float, 3>> offs;
  if (d->n_instances > 0 && d->instance_offsets) {
    for (int64_t i = 0; i < d->n_instances; ++i)
      offs.push_back({d->instance_offsets[3 * i], d->instance_offsets[3 * i + 1],
                      d->instance_offsets[3 * i + 2]});
  } else {
    offs.push_back({0.0f, 0.0f, 0.0f});
  }
  g.copies = static_cast<int64_t>(offs.size());
  // Mesh (scene_build.cpp:16-88): skipped unless >= 3 vertices and >= 1 tri.
  if (d->n_vertices >= 3 && d->n_triangles > 0) {
    g.baseV = d->n_vertices;
    g.baseT = d->n_triangles;
    for (const auto& off : offs) {
      for (int64_t v = 0; v < g.baseV; ++v) {
        for (int k = 0; k < 3; ++k) g.pos.push_back(d->positions[3 * v + k] + off[k]);
        for (int k = 0; k < 3; ++k) g.nrm.push_back(d->normals ? d->normals[3 * v + k] : 0.0f);
        for (int k = 0; k < 4; ++k)
          g.col.push_back(d->colors ? d->colors[4 * v + k] : (k == 3 ? 1.0f : 0.0f));
      }
    }
    for (int64_t c = 0; c < g.copies; ++c)
      for (int64_t t = 0; t < g.baseT; ++t)
        for (int k = 0; k < 3; ++k)
          g.idx.push_back(static_cast<uint32_t>(c * g.baseV + cornerVertex(d, t * 3 + k)));
    for (int32_t i = 0; i < d->n_mesh_materials; ++i) g.meshMats.push_back(toMat(d->mesh_materials[i]));
    if (d->tri_material) g.triMat.assign(d->tri_material, d->tri_material + g.baseT);
    if (d->tri_group) g.triGroup.assign(d->tri_group, d->tri_group + g.baseT);
  }
  for (int32_t i = 0; i < d->n_prim_materials; ++i) g.primMats.push_back(toMat(d->prim_materials[i]));
  // Spheres (scene_build.cpp:91-129): instance-major order.
  for (const auto& off : offs) {
    for (int64_t s = 0; s < d->n_spheres; ++s) {
      for (int k = 0; k < 3; ++k) g.sph.push_back(d->sphere_center_radius[4 * s + k] + off[k]);
      g.sph.push_back(d->sphere_center_radius[4 * s + 3]);
      PrimSide ps;
      for (int k = 0; k < 4; ++k) ps.color[k] = d->sphere_color[4 * s + k];
      ps.mat = d->sphere_material ? d->sphere_material[s] : 0;
      ps.group = d->sphere_group ? d->sphere_group[s] : 0;
      ps.fromEdge = d->sphere_from_edge ? d->sphere_from_edge[s] : 0;
      ps.opacity1 = -1.0f;
      g.sphSide.push_back(ps);
    }
  }
  // Cylinders (curve_build.cpp:41-267): open -> chained round, else cone.
  std::vector<int> openIdx, capIdx;
  for (int64_t i = 0; i < d->n_cylinders; ++i) {
    if (d->cyl_open && d->cyl_open[i]) openIdx.push_back(static_cast<int>(i));
    else capIdx.push_back(static_cast<int>(i));
  }
  if (!openIdx.empty())
    for (const auto& off : offs) chainOpenCylinders(d, openIdx, off.data(), g);
  for (const auto& off : offs) {
    for (int c : capIdx) {
      for (int k = 0; k < 3; ++k) g.cone.push_back(d->cyl_p0[3 * c + k] + off[k]);
      g.cone.push_back(d->cyl_radius[c]);
      for (int k = 0; k < 3; ++k) g.cone.push_back(d->cyl_p1[3 * c + k] + off[k]);
      g.cone.push_back(d->cyl_radius[c]);
      PrimSide ps;
      for (int k = 0; k < 4; ++k) ps.color[k] = d->cyl_color[4 * c + k];
      ps.mat = d->cyl_material ? d->cyl_material[c] : 0;
      ps.group = d->cyl_group ? d->cyl_group[c] : 0;
      ps.fromEdge = d->cyl_from_edge ? d->cyl_from_edge[c] : 0;
      ps.opacity1 = d->cyl_opacity1 ? d->cyl_opacity1[c] : -1.0f;
      setAxis(ps, &d->cyl_p0[3 * c], &d->cyl_p1[3 * c]);
      g.cSide.push_back(ps);
    }
  }
}

// ---------------------------------------------------------------------------
// Intersectors (Embree 4 semantics, restated). A hit is accepted when
// tnear <= t <= tfar. Curves and spheres solve in a frame whose origin is moved
// along the ray to the closest approach of the primitive's center, so far
// cameras do not cancel the discriminant.

template <class T>
struct Hit {
  T t;
  T u, v;
  int kind;  // -1 = miss
  uint32_t id;
  V3<T> ng;
};

template <class T>
inline T maxAbsT(V3<T> v) {
  return std::fmax(std::fmax(std::fabs(v.x), std::fabs(v.y)), std::fabs(v.z));
}

// Closest-approach origin shift: returns ts with o + ts*d nearest to c.
template <class T>
inline T shiftTo(V3<T> o, V3<T> d, V3<T> c, T A) {
  return dot(c - o, d) / A;
}

// FP64 polish of grazing roots, the FP32 path only (the FP64 path solves in
// double throughout). In the shifted frame a grazing ray's discriminant is a
// small difference of squares whose FP32 terms carry the rounding of the
// shifted origin (an ulp of the coordinates' magnitude m). When |disc| lies
// within that budget the quadratic is set up again in double from the exact
// inputs -- shifted origin (o - p) + d ts -- solved, and each root takes one
// Newton step (when the step is at most half the roots' separation). This is
// the library's own rule (its intersectors restate it bit for bit); Embree's
// ROBUST mode (scene_build.cpp:138) is what it stands in for, and no
// reference test pins either.
namespace graze {
struct D3 {
  double x, y, z;
};
inline D3 toD(V3<float> v) { return D3{v.x, v.y, v.z}; }
inline double dotD(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 subD(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 shifted(V3<float> o, V3<float> p, V3<float> d, float ts) {
  const double t = ts;
  return D3{(double(o.x) - double(p.x)) + double(d.x) * t, (double(o.y) - double(p.y)) + double(d.y) * t,
            (double(o.z) - double(p.z)) + double(d.z) * t};
}
inline float maxAbs(V3<float> v) { return std::fmax(std::fmax(std::fabs(v.x), std::fabs(v.y)), std::fabs(v.z)); }
inline float slack(float scale, float size, float m) { return scale * 0x1.0p-19f * size * m; }
inline bool quad(double A, double B, double C, double& x0, double& x1) {
  const double disc = B * B - A * C;
  if (disc < 0.0) return false;
  const double q = std::sqrt(disc);
  x0 = (-B - q) / A;
  x1 = (-B + q) / A;
  const double half = 0.5 * std::fabs(x1 - x0);
  for (double* x : {&x0, &x1}) {
    const double g = (A * *x + 2.0 * B) * *x + C;
    const double gp = 2.0 * (A * *x + B);
    if (gp != 0.0) {
      const double st = g / gp;
      if (std::fabs(st) <= half) *x = *x - st;
    }
  }
  return true;
}
}  // namespace graze

// Sphere (RTC_GEOMETRY_TYPE_SPHERE_POINT): near root if in range, else far.
template <class T>
inline bool isectSphere(V3<T> o, V3<T> d, T tnear, T tfar, V3<T> c, T r, Hit<T>& h) {
  const T A = dot(d, d);
  const T ts = shiftTo(o, d, c, A);
  const V3<T> f = (o + d * ts) - c;
  const T B = dot(f, d);
  const T C = dot(f, f) - r * r;
  const T disc = B * B - A * C;
  T t0, t1;
  bool polished = false;
  if constexpr (std::is_same_v<T, float>) {
    if (std::fabs(disc) <= graze::slack(A, r, std::fmax(graze::maxAbs(o), graze::maxAbs(c)))) {
      const graze::D3 fd = graze::shifted(o, c, d, ts), dd = graze::toD(d);
      const double rd = r;
      double x0, x1;
      if (!graze::quad(graze::dotD(dd, dd), graze::dotD(fd, dd), graze::dotD(fd, fd) - rd * rd, x0, x1))
        return false;
      t0 = float(double(ts) + x0);
      t1 = float(double(ts) + x1);
      polished = true;
    }
  }
  if (!polished) {
    if (disc < T(0)) return false;
    const T q = std::sqrt(disc);
    t0 = ts + (-B - q) / A;
    t1 = ts + (-B + q) / A;
  }
  T t;
  if (tnear <= t0 && t0 <= tfar) t = t0;
  else if (tnear <= t1 && t1 <= tfar) t = t1;
  else return false;
  h.t = t;
  h.u = T(0);
  h.v = T(0);
  h.ng = (o + d * t) - c;
  return true;
}

// Infinite-cone body roots in a shifted frame. Radius r(hax) = r0 + k*hax
// along the unit axis w from a. Returns number of roots (0 or 2) in dt[].
// (FP32 path: oa = (o + d ts) - a rounded; a grazing discriminant is solved
// again from (o - a) + d ts in double, with the axis and slope from ax =
// b - a and the end radii; m the magnitude of the rounded coordinates, ext
// the segment length.)
template <class T>
inline int coneBody(V3<T> oa, V3<T> d, V3<T> w, T r0, T k, T dt[2], V3<T> o, V3<T> a, V3<T> ax, T r1, T ts, T m,
                    T ext) {
  const T ho = dot(oa, w);
  const T hd = dot(d, w);
  const T rr = r0 + k * ho;
  const T A = dot(d, d) - hd * hd - k * k * hd * hd;
  const T B = dot(oa, d) - ho * hd - k * hd * rr;
  const T C = dot(oa, oa) - ho * ho - rr * rr;
  if (A == T(0)) return 0;
  const T disc = B * B - A * C;
  if constexpr (std::is_same_v<T, float>) {
    if (std::fabs(disc) <= graze::slack(std::fabs(A) + dot(d, d), std::fabs(rr) + std::fabs(r0) + ext, m)) {
      const graze::D3 od = graze::shifted(o, a, d, ts), dd = graze::toD(d), axd = graze::toD(ax);
      const double Ld = std::sqrt(graze::dotD(axd, axd));
      const graze::D3 wd{axd.x / Ld, axd.y / Ld, axd.z / Ld};
      const double kd = (double(r1) - double(r0)) / Ld;
      const double hoD = graze::dotD(od, wd), hdD = graze::dotD(dd, wd), rrD = double(r0) + kd * hoD;
      const double AD = graze::dotD(dd, dd) - hdD * hdD - kd * kd * hdD * hdD;
      const double BD = graze::dotD(od, dd) - hoD * hdD - kd * hdD * rrD;
      const double CD = graze::dotD(od, od) - hoD * hoD - rrD * rrD;
      double x0, x1;
      if (AD == 0.0 || !graze::quad(AD, BD, CD, x0, x1)) return 0;
      dt[0] = float(x0);
      dt[1] = float(x1);
      return 2;
    }
  }
  if (disc < T(0)) return 0;
  const T q = std::sqrt(disc);
  dt[0] = (-B - q) / A;
  dt[1] = (-B + q) / A;
  return 2;
}

// CONE_LINEAR_CURVE: cone body between the end planes plus flat disk caps.
// u = axial fraction, Ng = outward surface normal (caps: -w at p0, +w at p1).
template <class T>
inline bool isectCone(V3<T> o, V3<T> d, T tnear, T tfar, V3<T> a, T ra, V3<T> b, T rb,
                      Hit<T>& h) {
  const V3<T> ax = b - a;
  const T L = length(ax);
  if (!(L > T(0))) return false;
  const V3<T> w = ax * (T(1) / L);
  const T k = (rb - ra) / L;
  const T A = dot(d, d);
  const V3<T> mid = (a + b) * T(0.5f);
  const T ts = shiftTo(o, d, mid, A);
  const V3<T> oa = (o + d * ts) - a;
  const T ho = dot(oa, w);
  const T hd = dot(d, w);
  bool found = false;
  T best = tfar;
  T bu = T(0);
  V3<T> bn{T(0), T(0), T(0)};
  T dt[2];
  const T m = std::fmax(maxAbsT(o), maxAbsT(mid));
  if (coneBody(oa, d, w, ra, k, dt, o, a, ax, rb, ts, m, L) == 2) {
    for (int i = 0; i < 2; ++i) {
      const T t = ts + dt[i];
      if (!(tnear <= t && t <= best)) continue;
      const T hh = ho + dt[i] * hd;
      if (hh < T(0) || hh > L) continue;
      if (ra + k * hh < T(0)) continue;
      const V3<T> p = oa + d * dt[i];
      const V3<T> radial = p - w * hh;
      // cone normal: radial direction tilted by the slope
      const T rl = length(radial);
      V3<T> n = radial;
      if (k != T(0) && rl > T(0)) n = radial * (T(1) / rl) - w * k;
      best = t;
      bu = hh / L;
      bn = n;
      found = true;
      break;
    }
  }
  if (hd != T(0)) {
    // cap at a: plane hax = 0
    {
      const T dtc = -ho / hd;
      const T t = ts + dtc;
      if (tnear <= t && t <= best) {
        const V3<T> p = oa + d * dtc;
        if (dot(p, p) <= ra * ra) {
          best = t;
          bu = T(0);
          bn = neg(w);
          found = true;
        }
      }
    }
    {
      const T dtc = (L - ho) / hd;
      const T t = ts + dtc;
      if (tnear <= t && t <= best) {
        const V3<T> p = (oa + d * dtc) - w * L;
        if (dot(p, p) <= rb * rb) {
          best = t;
          bu = T(1);
          bn = w;
          found = true;
        }
      }
    }
  }
  if (!found) return false;
  h.t = best;
  h.u = bu;
  h.v = T(0);
  h.ng = bn;
  return true;
}

// ROUND_LINEAR_CURVE segment with neighbor flags: swept sphere (cone body +
// end spheres) restricted to the segment's cell between the bisector planes of
// its chained joints, so a chain has no internal surfaces (the dropped caps of
// RTC_CURVE_FLAG_NEIGHBOR_LEFT/RIGHT).
template <class T>
inline bool isectRound(V3<T> o, V3<T> d, T tnear, T tfar, V3<T> a, T ra, V3<T> b, T rb,
                       bool hasL, V3<T> pl, bool hasR, V3<T> pr, Hit<T>& h) {
  const V3<T> ax = b - a;
  const T L = length(ax);
  if (!(L > T(0))) return false;
  const V3<T> w = ax * (T(1) / L);
  const T k = (rb - ra) / L;
  V3<T> nL = w, nR = w;
  if (hasL) nL = safeNormalize(normalize(a - pl) + w, w);
  if (hasR) nR = safeNormalize(w + normalize(pr - b), w);
  const T A = dot(d, d);
  const V3<T> mid = (a + b) * T(0.5f);
  const T ts = shiftTo(o, d, mid, A);
  const V3<T> os = o + d * ts;
  const V3<T> oa = os - a;
  const T ho = dot(oa, w);
  const T hd = dot(d, w);
  bool found = false;
  T best = tfar;
  T bu = T(0);
  V3<T> bn{T(0), T(0), T(0)};
  auto inCell = [&](V3<T> p /* relative to a */) {
    if (hasL && dot(p, nL) < T(0)) return false;
    if (hasR && dot(p - ax, nR) > T(0)) return false;
    return true;
  };
  auto consider = [&](T dtv, V3<T> n, T hh) {
    const T t = ts + dtv;
    if (!(tnear <= t && t <= best)) return;
    const V3<T> p = oa + d * dtv;
    if (!inCell(p)) return;
    best = t;
    T uu = hh / L;
    if (uu < T(0)) uu = T(0);
    if (uu > T(1)) uu = T(1);
    bu = uu;
    bn = n;
    found = true;
  };
  T dt[2];
  const T m = std::fmax(maxAbsT(o), maxAbsT(mid));
  if (coneBody(oa, d, w, ra, k, dt, o, a, ax, rb, ts, m, L) == 2) {
    for (int i = 0; i < 2; ++i) {
      const T hh = ho + dt[i] * hd;
      if (hh < T(0) || hh > L) continue;
      const V3<T> p = oa + d * dt[i];
      consider(dt[i], p - w * hh, hh);
    }
  }
  // end spheres: the part outside the body slab
  for (int e = 0; e < 2; ++e) {
    const V3<T> c = e == 0 ? V3<T>{T(0), T(0), T(0)} : ax;
    const T r = e == 0 ? ra : rb;
    const V3<T> f = oa - c;
    const T B = dot(f, d);
    const T C = dot(f, f) - r * r;
    const T disc = B * B - A * C;
    T r0, r1;
    bool polished = false;
    if constexpr (std::is_same_v<T, float>) {
      if (std::fabs(disc) <= graze::slack(A, r + L, m)) {
        const graze::D3 fd = graze::subD(graze::shifted(o, a, d, ts), graze::toD(c)), dd = graze::toD(d);
        const double rd = r;
        double x0, x1;
        if (!graze::quad(graze::dotD(dd, dd), graze::dotD(fd, dd), graze::dotD(fd, fd) - rd * rd, x0, x1)) continue;
        r0 = float(x0);
        r1 = float(x1);
        polished = true;
      }
    }
    if (!polished) {
      if (disc < T(0)) continue;
      const T q = std::sqrt(disc);
      r0 = (-B - q) / A;
      r1 = (-B + q) / A;
    }
    for (T dtv : {r0, r1}) {
      const T hh = ho + dtv * hd;
      if (e == 0 && hh >= T(0)) continue;
      if (e == 1 && hh <= L) continue;
      const V3<T> p = oa + d * dtv;
      consider(dtv, p - c, hh);
    }
  }
  if (!found) return false;
  h.t = best;
  h.u = bu;
  h.v = T(0);
  h.ng = bn;
  return true;
}

// TRIANGLE: Moller-Trumbore; P = (1-u-v) v0 + u v1 + v v2; Ng = (v1-v0)x(v2-v0).
template <class T>
inline bool isectTri(V3<T> o, V3<T> d, T tnear, T tfar, V3<T> v0, V3<T> v1, V3<T> v2,
                     Hit<T>& h) {
  const V3<T> e1 = v1 - v0;
  const V3<T> e2 = v2 - v0;
  const V3<T> pv = cross(d, e2);
  const T det = dot(e1, pv);
  if (det == T(0)) return false;
  const T inv = T(1) / det;
  const V3<T> tv = o - v0;
  const T u = dot(tv, pv) * inv;
  if (u < T(0) || u > T(1)) return false;
  const V3<T> qv = cross(tv, e1);
  const T v = dot(d, qv) * inv;
  if (v < T(0) || u + v > T(1)) return false;
  const T t = dot(e2, qv) * inv;
  if (!(tnear <= t && t <= tfar)) return false;
  h.t = t;
  h.u = u;
  h.v = v;
  h.ng = cross(e1, e2);
  return true;
}

// ---------------------------------------------------------------------------
// BVH over all primitives (binned SAH, 16 bins). Only speed depends on it.

struct Box {
  double lo[3], hi[3];
};
struct PRef {
  int kind;
  uint32_t id;
};
struct BNode {
  Box b;
  int32_t left = -1, right = -1;
  int32_t first = 0, count = 0;
};

class Bvh {
 public:
  std::vector<BNode> nodes;
  std::vector<PRef> refs;
  void build(std::vector<PRef> prefs, std::vector<Box> boxes) {
    refs = std::move(prefs);
    boxes_ = std::move(boxes);
    order_.resize(refs.size());
    for (std::size_t i = 0; i < order_.size(); ++i) order_[i] = static_cast<int32_t>(i);
    cent_.resize(refs.size());
    for (std::size_t i = 0; i < refs.size(); ++i)
      for (int k = 0; k < 3; ++k) cent_[i][k] = 0.5 * (boxes_[i].lo[k] + boxes_[i].hi[k]);
    nodes.clear();
    if (refs.empty()) return;
    nodes.reserve(2 * refs.size());
    nodes.push_back(BNode{});
    buildRec(0, 0, static_cast<int32_t>(refs.size()), 0);
    std::vector<PRef> r2(refs.size());
    for (std::size_t i = 0; i < refs.size(); ++i) r2[i] = refs[order_[i]];
    refs.swap(r2);
  }

 private:
  std::vector<Box> boxes_;
  std::vector<int32_t> order_;
  std::vector<std::array<double, 3>> cent_;
  static double area(const Box& b) {
    const double dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    if (dx < 0) return 0;
    return 2 * (dx * dy + dy * dz + dz * dx);
  }
  static void grow(Box& a, const Box& b) {
    for (int k = 0; k < 3; ++k) {
      a.lo[k] = std::min(a.lo[k], b.lo[k]);
      a.hi[k] = std::max(a.hi[k], b.hi[k]);
    }
  }
  static Box empty() {
    Box b;
    for (int k = 0; k < 3; ++k) {
      b.lo[k] = 1e300;
      b.hi[k] = -1e300;
    }
    return b;
  }
  void buildRec(int32_t ni, int32_t first, int32_t count, int depth) {
    Box b = empty(), cb = empty();
    for (int32_t i = first; i < first + count; ++i) {
      grow(b, boxes_[order_[i]]);
      for (int k = 0; k < 3; ++k) {
        cb.lo[k] = std::min(cb.lo[k], cent_[order_[i]][k]);
        cb.hi[k] = std::max(cb.hi[k], cent_[order_[i]][k]);
      }
    }
    nodes[ni].b = b;
    if (count <= 2 || depth > 60) {
      nodes[ni].first = first;
      nodes[ni].count = count;
      return;
    }
    constexpr int NB = 16;
    double bestCost = 1e300;
    int bestAxis = -1, bestSplit = -1;
    for (int ax = 0; ax < 3; ++ax) {
      const double ext = cb.hi[ax] - cb.lo[ax];
      if (!(ext > 0)) continue;
      Box bb[NB];
      int cnt[NB] = {0};
      for (int i = 0; i < NB; ++i) bb[i] = empty();
      for (int32_t i = first; i < first + count; ++i) {
        int bi = static_cast<int>(NB * (cent_[order_[i]][ax] - cb.lo[ax]) / ext);
        bi = std::min(NB - 1, std::max(0, bi));
        cnt[bi]++;
        grow(bb[bi], boxes_[order_[i]]);
      }
      double la[NB], ra[NB];
      int lc[NB], rc[NB];
      Box acc = empty();
      int c = 0;
      for (int i = 0; i < NB; ++i) {
        grow(acc, bb[i]);
        c += cnt[i];
        la[i] = area(acc);
        lc[i] = c;
      }
      acc = empty();
      c = 0;
      for (int i = NB - 1; i >= 0; --i) {
        grow(acc, bb[i]);
        c += cnt[i];
        ra[i] = area(acc);
        rc[i] = c;
      }
      for (int i = 0; i < NB - 1; ++i) {
        if (lc[i] == 0 || rc[i + 1] == 0) continue;
        const double cost = la[i] * lc[i] + ra[i + 1] * rc[i + 1];
        if (cost < bestCost) {
          bestCost = cost;
          bestAxis = ax;
          bestSplit = i;
        }
      }
    }
    const double leafCost = area(b) * count;
    if (bestAxis < 0 || (count <= 4 && bestCost >= leafCost)) {
      nodes[ni].first = first;
      nodes[ni].count = count;
      return;
    }
    const double ext = cb.hi[bestAxis] - cb.lo[bestAxis];
    auto mid = std::partition(order_.begin() + first, order_.begin() + first + count, [&](int32_t p) {
      int bi = static_cast<int>(NB * (cent_[p][bestAxis] - cb.lo[bestAxis]) / ext);
      bi = std::min(NB - 1, std::max(0, bi));
      return bi <= bestSplit;
    });
    int32_t nl = static_cast<int32_t>(mid - (order_.begin() + first));
    if (nl == 0 || nl == count) nl = count / 2;
    const int32_t l = static_cast<int32_t>(nodes.size());
    nodes.push_back(BNode{});
    nodes.push_back(BNode{});
    nodes[ni].left = l;
    nodes[ni].right = l + 1;
    buildRec(l, first, nl, depth + 1);
    buildRec(l + 1, first + nl, count - nl, depth + 1);
  }
};

// ---------------------------------------------------------------------------
// Scene with acceleration structure.

struct OScene {
  Geometry g;
  Bvh bvh;
  bool brute = false;
  std::vector<PRef> all;  // geomID order: mesh, spheres, round, cone
};

static Box primBox(const Geometry& g, const PRef& p) {
  Box b;
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = 1e300;
    b.hi[k] = -1e300;
  }
  auto ext = [&](double x, double y, double z, double r) {
    const double c[3] = {x, y, z};
    for (int k = 0; k < 3; ++k) {
      b.lo[k] = std::min(b.lo[k], c[k] - r);
      b.hi[k] = std::max(b.hi[k], c[k] + r);
    }
  };
  if (p.kind == UBG_KIND_MESH) {
    for (int c = 0; c < 3; ++c) {
      const uint32_t vi = g.idx[3 * p.id + c];
      ext(g.pos[3 * vi], g.pos[3 * vi + 1], g.pos[3 * vi + 2], 0.0);
    }
  } else if (p.kind == UBG_KIND_SPHERE) {
    const float* s = &g.sph[4 * p.id];
    ext(s[0], s[1], s[2], s[3]);
  } else if (p.kind == UBG_KIND_ROUND_CURVE) {
    const uint32_t v = g.rseg[p.id];
    const float* a = &g.rv[4 * v];
    ext(a[0], a[1], a[2], a[3]);
    ext(a[4], a[5], a[6], a[7]);
  } else {
    const float* c = &g.cone[8 * p.id];
    ext(c[0], c[1], c[2], c[3]);
    ext(c[4], c[5], c[6], c[7]);
  }
  // conservative inflation (covers float rounding in the intersectors)
  for (int k = 0; k < 3; ++k) {
    const double m = std::max(std::fabs(b.lo[k]), std::fabs(b.hi[k]));
    b.lo[k] -= 1e-5 * m + 1e-6;
    b.hi[k] += 1e-5 * m + 1e-6;
  }
  return b;
}

static void buildScene(const ubg_scene_desc* d, bool brute, OScene& s) {
  buildGeometry(d, s.g);
  const Geometry& g = s.g;
  for (uint32_t i = 0; i < g.idx.size() / 3; ++i) s.all.push_back({UBG_KIND_MESH, i});
  for (uint32_t i = 0; i < g.sph.size() / 4; ++i) s.all.push_back({UBG_KIND_SPHERE, i});
  for (uint32_t i = 0; i < g.rseg.size(); ++i) s.all.push_back({UBG_KIND_ROUND_CURVE, i});
  for (uint32_t i = 0; i < g.cone.size() / 8; ++i) s.all.push_back({UBG_KIND_CONE_CURVE, i});
  s.brute = brute;
  if (!brute) {
    std::vector<Box> boxes;
    boxes.reserve(s.all.size());
    for (const auto& p : s.all) boxes.push_back(primBox(g, p));
    s.bvh.build(s.all, std::move(boxes));
  }
}

// Reuse of the last built scene across render calls (orc_set_scene_cache;
// off by default). The key hashes every array the description points to, so
// a changed scene always rebuilds; only the timing of repeated renders of
// one scene changes.
namespace cache {
std::mutex mu;
bool enabled = false;
uint64_t key = 0;
std::shared_ptr<OScene> scene;

uint64_t mix(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    std::memcpy(&w, b + i, 8);
    h = (h ^ (w * 0x9E3779B97F4A7C15ull)) * 0xBF58476D1CE4E5B9ull;
    h ^= h >> 31;
  }
  for (; i < n; ++i) h = (h ^ b[i]) * 0x100000001B3ull;
  return (h ^ n) * 0x94D049BB133111EBull;
}

uint64_t descKey(const ubg_scene_desc* d, bool brute) {
  uint64_t h = brute ? 1 : 2;
  auto arr = [&](const void* p, int64_t count, size_t elem) {
    const int64_t n = p ? count : -1;
    h = mix(h, &n, sizeof(n));
    if (p && count > 0) h = mix(h, p, static_cast<size_t>(count) * elem);
  };
  const int64_t nv = d->n_vertices, nt = d->n_triangles, ns = d->n_spheres, nc = d->n_cylinders;
  arr(d->positions, 3 * nv, 4);
  arr(d->normals, 3 * nv, 4);
  arr(d->colors, 4 * nv, 4);
  arr(d->index, 3 * nt, 4);
  arr(d->mesh_materials, d->n_mesh_materials, sizeof(ubg_material));
  arr(d->tri_material, nt, 1);
  arr(d->tri_group, nt, 2);
  arr(d->sphere_center_radius, 4 * ns, 4);
  arr(d->sphere_color, 4 * ns, 4);
  arr(d->sphere_material, ns, 4);
  arr(d->sphere_group, ns, 2);
  arr(d->sphere_from_edge, ns, 1);
  arr(d->cyl_p0, 3 * nc, 4);
  arr(d->cyl_p1, 3 * nc, 4);
  arr(d->cyl_radius, nc, 4);
  arr(d->cyl_color, 4 * nc, 4);
  arr(d->cyl_opacity1, nc, 4);
  arr(d->cyl_material, nc, 4);
  arr(d->cyl_group, nc, 2);
  arr(d->cyl_open, nc, 1);
  arr(d->cyl_from_edge, nc, 1);
  arr(d->prim_materials, d->n_prim_materials, sizeof(ubg_material));
  arr(d->instance_offsets, 3 * d->n_instances, 4);
  return h;
}
}  // namespace cache

// The scene of a render call: built here, or the cached one when reuse is on
// and the description is unchanged.
static std::shared_ptr<const OScene> sceneFor(const ubg_scene_desc* d, bool brute) {
  std::lock_guard<std::mutex> lk(cache::mu);
  if (!cache::enabled) {
    auto s = std::make_shared<OScene>();
    buildScene(d, brute, *s);
    return s;
  }
  const uint64_t k = cache::descKey(d, brute);
  if (!cache::scene || cache::key != k) {
    cache::scene.reset();
    auto s = std::make_shared<OScene>();
    buildScene(d, brute, *s);
    cache::scene = s;
    cache::key = k;
  }
  return cache::scene;
}

template <class T>
static bool isectPrim(const Geometry& g, const PRef& p, V3<T> o, V3<T> d, T tnear, T tfar,
                      Hit<T>& h) {
  if (p.kind == UBG_KIND_MESH) {
    const uint32_t* ix = &g.idx[3 * p.id];
    return isectTri(o, d, tnear, tfar, v3<T>(&g.pos[3 * ix[0]]), v3<T>(&g.pos[3 * ix[1]]),
                    v3<T>(&g.pos[3 * ix[2]]), h);
  }
  if (p.kind == UBG_KIND_SPHERE) {
    const float* s = &g.sph[4 * p.id];
    return isectSphere(o, d, tnear, tfar, v3<T>(s), T(s[3]), h);
  }
  if (p.kind == UBG_KIND_ROUND_CURVE) {
    const uint32_t v = g.rseg[p.id];
    const uint8_t fl = g.rflag[p.id];
    const float* a = &g.rv[4 * v];
    const V3<T> pl = (fl & 1) ? v3<T>(&g.rv[4 * (v - 1)]) : V3<T>{T(0), T(0), T(0)};
    const V3<T> pr = (fl & 2) ? v3<T>(&g.rv[4 * (v + 2)]) : V3<T>{T(0), T(0), T(0)};
    return isectRound(o, d, tnear, tfar, v3<T>(a), T(a[3]), v3<T>(a + 4), T(a[7]), (fl & 1) != 0,
                      pl, (fl & 2) != 0, pr, h);
  }
  const float* c = &g.cone[8 * p.id];
  return isectCone(o, d, tnear, tfar, v3<T>(c), T(c[3]), v3<T>(c + 4), T(c[7]), h);
}

static inline bool boxHit(const Box& b, const double o[3], const double inv[3], double t0,
                          double t1) {
  for (int k = 0; k < 3; ++k) {
    double a = (b.lo[k] - o[k]) * inv[k];
    double c = (b.hi[k] - o[k]) * inv[k];
    if (std::isnan(a)) a = -1e300;
    if (std::isnan(c)) c = 1e300;
    if (a > c) std::swap(a, c);
    t0 = std::max(t0, a);
    t1 = std::min(t1, c);
  }
  return t0 <= t1 * (1.0 + 1e-9) + 1e-12;
}

// rtcIntersect1 analogue: closest hit in [tnear, tfar]; of primitives hit at
// exactly the same t the one with the smaller (kind << 30 | id) wins, so the
// answer does not depend on the BVH (Embree's own tie order is its traversal
// order, which no test pins).
template <class T>
static Hit<T> intersect(const OScene& s, V3<T> o, V3<T> d, T tnear, T tfar) {
  Hit<T> best;
  best.kind = -1;
  best.t = tfar;
  Hit<T> h;
  auto key = [](int kind, uint32_t id) { return static_cast<uint32_t>(kind) << 30 | id; };
  auto test = [&](const PRef& p) {
    if (isectPrim(s.g, p, o, d, tnear, best.t, h) &&
        (best.kind < 0 || h.t < best.t || (h.t == best.t && key(p.kind, p.id) < key(best.kind, best.id)))) {
      best = h;
      best.kind = p.kind;
      best.id = p.id;
    }
  };
  if (s.brute) {
    for (const auto& p : s.all) test(p);
    return best;
  }
  if (s.bvh.nodes.empty()) return best;
  const double od[3] = {double(o.x), double(o.y), double(o.z)};
  const double inv[3] = {1.0 / double(d.x), 1.0 / double(d.y), 1.0 / double(d.z)};
  int32_t stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const BNode& n = s.bvh.nodes[stack[--sp]];
    const double tb = best.kind < 0 ? double(tfar) : double(best.t);
    if (!boxHit(n.b, od, inv, double(tnear), tb)) continue;
    if (n.count) {
      for (int32_t i = n.first; i < n.first + n.count; ++i) test(s.bvh.refs[i]);
    } else {
      stack[sp++] = n.right;
      stack[sp++] = n.left;
    }
  }
  return best;
}

// rtcOccluded1 analogue.
template <class T>
static bool occluded(const OScene& s, V3<T> o, V3<T> d, T tnear, T tfar) {
  Hit<T> h;
  if (s.brute) {
    for (const auto& p : s.all)
      if (isectPrim(s.g, p, o, d, tnear, tfar, h)) return true;
    return false;
  }
  if (s.bvh.nodes.empty()) return false;
  const double od[3] = {double(o.x), double(o.y), double(o.z)};
  const double inv[3] = {1.0 / double(d.x), 1.0 / double(d.y), 1.0 / double(d.z)};
  int32_t stack[128];
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    const BNode& n = s.bvh.nodes[stack[--sp]];
    if (!boxHit(n.b, od, inv, double(tnear), double(tfar))) continue;
    if (n.count) {
      for (int32_t i = n.first; i < n.first + n.count; ++i)
        if (isectPrim(s.g, s.bvh.refs[i], o, d, tnear, tfar, h)) return true;
    } else {
      stack[sp++] = n.right;
      stack[sp++] = n.left;
    }
  }
  return false;
}

// ---------------------------------------------------------------------------
// Shading (hit_shader.hpp:147-496, shading.hpp:76-192, secondary_rays.hpp,
// ambient_occlusion.hpp:33-57, ao_shade.hpp:128-142).

template <class T>
struct Light {
  V3<T> L, color;
  bool highlight;
  T radius;
};

struct Counters {
  uint64_t primary = 0, shaded = 0, aoTraced = 0, aoSampled = 0, shadow = 0, capped = 0;
  uint64_t giRays = 0, giNee = 0;
};

// scene.hpp:171-196 -- Material::diffuseWeight (principled: the metal cut
// diffuse * (1 - metallic); POV and toon: diffuse) and toonLike
inline float diffuseWeight(const Mat& m) { return m.model == 1 ? m.diffuse * (1.0f - m.pbrMetallic) : m.diffuse; }
inline bool toonLike(const Mat& m) {
  if (m.model == 2) return true;
  if (m.model != 0) return false;
  if (m.specular > 1.0f || m.phong > 1.0f || m.brilliance == 0.0f) return true;
  return m.diffuse <= 0.0f && m.specular <= 0.0f && m.phong <= 0.0f && m.reflection <= 0.0f;
}

template <class T>
struct Ctx {
  const OScene* s;
  const ubg_render_params* p;
  std::vector<Light<T>> lights;
  V3<T> ambLight, bg, camDir;
  V3<T> aoUp;  // resolved bent-normal gradient axis (embree_renderer.cpp:1320-1323)
  int wHi, hHi;
  const struct CoarseAo* grid = nullptr;  // coarse-grid AO (aoResDiv > 1), else null
  // hit_shader.hpp:101-107 -- seed lattice width (the hi-res width, or W*k for
  // the adaptive-AA fine samples) and the replicated-centre sample multipliers
  int seedW = 0;
  int aoMul = 1, shMul = 1;
  bool gi = false;  // pt1 GI: receivers drop their constant ambient
};

// ao_coarse.hpp:54-65 -- the coarse AO grid: gathered openness (and bent
// normal for the enhanced estimator) plus the lookup guides
struct CoarseAo {
  int w = 0, h = 0;
  std::vector<double> openness, bent, normal, depth;
  std::vector<uint8_t> hit;
};

// ao_coarse.hpp:138-205 -- per-hit bilateral lookup over the 2x2 surrounding
// cells (bilinear x normal^32 x depth weights); false below the 0.3 weight sum
template <class T>
inline bool sampleCoarseAo(const CoarseAo& g, int latW, int latH, uint32_t px, uint32_t py, V3<T> N, T z,
                           T& openness, V3<T>& bent) {
  if (g.w <= 0 || g.h <= 0) return false;
  const T sx = (T(static_cast<float>(px)) + T(0.5f)) * T(static_cast<float>(g.w)) / T(static_cast<float>(latW)) -
               T(0.5f);
  const T sy = (T(static_cast<float>(py)) + T(0.5f)) * T(static_cast<float>(g.h)) / T(static_cast<float>(latH)) -
               T(0.5f);
  const int bx = static_cast<int>(std::floor(sx));
  const int by = static_cast<int>(std::floor(sy));
  const T fx = sx - T(static_cast<float>(bx));
  const T fy = sy - T(static_cast<float>(by));
  const bool quality = !g.bent.empty();
  T wSum = T(0), opSum = T(0);
  V3<T> bentSum{T(0), T(0), T(0)};
  for (int dy = 0; dy < 2; ++dy) {
    const int cy = by + dy;
    if (cy < 0 || cy >= g.h) continue;
    for (int dx = 0; dx < 2; ++dx) {
      const int cx = bx + dx;
      if (cx < 0 || cx >= g.w) continue;
      const std::size_t cp = static_cast<std::size_t>(cy) * g.w + cx;
      if (!g.hit[cp]) continue;
      const T wb = (dx ? fx : T(1) - fx) * (dy ? fy : T(1) - fy);
      const V3<T> Nc{T(g.normal[3 * cp]), T(g.normal[3 * cp + 1]), T(g.normal[3 * cp + 2])};
      const T nd = dot(N, Nc);
      if (nd <= T(0)) continue;
      const T wn = std::pow(nd, T(32.0f));
      const T zc = T(g.depth[cp]);
      const T wz = std::exp(-std::fabs(z - zc) / (T(0.02f) * z + T(1e-6f)));
      const T wgt = wb * wn * wz;
      if (wgt <= T(0)) continue;
      wSum += wgt;
      opSum += wgt * T(g.openness[cp]);
      if (quality) {
        bentSum.x += wgt * T(g.bent[3 * cp]);
        bentSum.y += wgt * T(g.bent[3 * cp + 1]);
        bentSum.z += wgt * T(g.bent[3 * cp + 2]);
      }
    }
  }
  if (wSum < T(0.3f)) return false;
  const T inv = T(1) / wSum;
  openness = opSum * inv;
  if (quality) bent = safeNormalize(V3<T>{bentSum.x * inv, bentSum.y * inv, bentSum.z * inv}, N);
  return true;
}

// secondary_rays.hpp:102-109
template <class T>
inline V3<T> sampleLightDir(const Light<T>& l, T u1, T u2) {
  if (l.radius <= T(0)) return l.L;
  const Frame<T> f = frameFromNormal(l.L);
  const T phi = T(6.2831853072f) * u1;
  const T rr = std::tan(l.radius) * std::sqrt(u2);
  const V3<T> dd = f.n + f.t * (rr * std::cos(phi)) + f.b * (rr * std::sin(phi));
  return safeNormalize(dd, l.L);
}

// secondary_rays.hpp:118-134
template <class T>
inline T computeShadow(const Ctx<T>& c, V3<T> P, V3<T> Ng, V3<T> N, T eps, const Light<T>& l,
                       int shadowSamples, uint32_t& s0, uint32_t& s1, Counters& cnt) {
  V3<T> ng = (dot(Ng, N) < T(0)) ? neg(Ng) : Ng;
  ng = safeNormalize(ng, N);
  const V3<T> O = P + ng * eps;
  const T far = std::numeric_limits<T>::infinity();
  if (shadowSamples <= 1 || l.radius <= T(0)) {
    cnt.shadow++;
    return occluded(*c.s, O, l.L, eps, far) ? T(0) : T(1);
  }
  int hits = 0;
  for (int i = 0; i < shadowSamples; ++i) {
    tea2(s0, s1);
    const V3<T> dir = sampleLightDir(l, u32ToUnorm<T>(s0), u32ToUnorm<T>(s1));
    cnt.shadow++;
    if (occluded(*c.s, O, dir, eps, far)) ++hits;
  }
  return T(1) - T(hits) / T(shadowSamples);
}

// ambient_occlusion.hpp:33-57
template <class T>
inline T computeAO(const Ctx<T>& c, V3<T> P, V3<T> Ng, V3<T> Ns, T eps, int nSamples, T aoRadius,
                   uint32_t px, uint32_t py, int wSeed, Counters& cnt) {
  if (nSamples <= 0) return T(1);
  const Frame<T> f = frameFromNormal(Ns);
  V3<T> ng = (dot(Ng, Ns) < T(0)) ? neg(Ng) : Ng;
  ng = safeNormalize(ng, Ns);
  const V3<T> O = P + ng * eps;
  const uint32_t base = px + py * static_cast<uint32_t>(wSeed);
  int hits = 0;
  for (int i = 0; i < nSamples; ++i) {
    uint32_t s0 = base;
    uint32_t s1 = static_cast<uint32_t>(i);
    tea2(s0, s1);
    const V3<T> dir = cosineSampleHemisphere(u32ToUnorm<T>(s0), u32ToUnorm<T>(s1), f);
    if (dot(dir, Ns) < T(0.01f)) {
      ++hits;
      cnt.aoSampled++;
      continue;
    }
    cnt.aoTraced++;
    if (occluded(*c.s, O, dir, eps, aoRadius)) ++hits;
  }
  return T(1) - T(hits) / T(nSamples);
}

// ambient_occlusion.hpp:96-104 -- base-2 radical inverse (bit reversal)
inline uint32_t reverseBits32(uint32_t b) {
  b = (b << 16) | (b >> 16);
  b = ((b & 0x55555555u) << 1) | ((b & 0xAAAAAAAAu) >> 1);
  b = ((b & 0x33333333u) << 2) | ((b & 0xCCCCCCCCu) >> 2);
  b = ((b & 0x0F0F0F0Fu) << 4) | ((b & 0xF0F0F0F0u) >> 4);
  b = ((b & 0x00FF00FFu) << 8) | ((b & 0xFF00FF00u) >> 8);
  return b;
}

// ambient_occlusion.hpp:111-125 -- the i-th AO sample: the tea2 stream, or
// the Hammersley point (i/N, radicalInverse2(i)) shifted by (cpx, cpy) mod 1
template <class T>
inline void aoSample2d(bool lowDiscrepancy, uint32_t base, int i, int n, T cpx, T cpy, T& u1, T& u2) {
  if (lowDiscrepancy) {
    u1 = T(static_cast<float>(i)) / T(static_cast<float>(n)) + cpx;
    u2 = T(static_cast<float>(reverseBits32(static_cast<uint32_t>(i)))) * T(0x1.0p-32f) + cpy;
    if (u1 >= T(1)) u1 -= T(1);
    if (u2 >= T(1)) u2 -= T(1);
  } else {
    uint32_t s0 = base, s1 = static_cast<uint32_t>(i);
    tea2(s0, s1);
    u1 = u32ToUnorm<T>(s0);
    u2 = u32ToUnorm<T>(s1);
  }
}

// ambient_occlusion.hpp:133-138 -- GTAO multibounce lift of one channel
template <class T>
inline T aoMultibounce(T ao, T albedo) {
  const T a = T(2.0404f) * albedo - T(0.3324f);
  const T b = T(-4.7951f) * albedo + T(0.6417f);
  const T cc = T(2.7552f) * albedo + T(0.6903f);
  return std::fmax(ao, ((ao * a + b) * ao + cc) * ao);
}

template <class T>
struct AOQuality {
  T contact = T(1), shape = T(1), openness = T(1);
  V3<T> bent{T(0), T(0), T(0)};
  T avgHitDist = T(0);
};

template <class T>
inline bool aoEnhanced(const ubg_render_params* p) {
  return p->ao_falloff_power > 0.0f || p->ao_multi_scale || p->ao_bent_normal || p->ao_multibounce ||
         p->ao_low_discrepancy || p->ao_diffuse_factor > 0.0f;
}

// ambient_occlusion.hpp:171-264 -- nearest-hit multi-scale / falloff / bent
// normal estimator from one cosine ray set
template <class T>
inline AOQuality<T> computeAOQuality(const Ctx<T>& c, V3<T> P, V3<T> Ng, V3<T> Ns, T eps, int nSamples,
                                     uint32_t px, uint32_t py, int wSeed, Counters& cnt) {
  const ubg_render_params* p = c.p;
  AOQuality<T> r;
  if (nSamples <= 0) {
    r.bent = Ns;
    return r;
  }
  const Frame<T> f = frameFromNormal(Ns);
  V3<T> ng = (dot(Ng, Ns) < T(0)) ? neg(Ng) : Ng;
  ng = safeNormalize(ng, Ns);
  const V3<T> O = P + ng * eps;
  const uint32_t base = px + py * static_cast<uint32_t>(wSeed);
  const T R = T(p->ao_distance);
  T radius[3], weight[3];
  int K;
  if (p->ao_multi_scale) {
    K = 3;
    radius[0] = T(0.08f) * R;
    radius[1] = T(0.30f) * R;
    radius[2] = T(1.00f) * R;
    weight[0] = T(0.55f);
    weight[1] = T(0.35f);
    weight[2] = T(0.10f);
  } else {
    K = 1;
    radius[0] = radius[1] = radius[2] = R;
    weight[0] = T(1);
    weight[1] = weight[2] = T(0);
  }
  uint32_t c0 = px + 1u, c1 = py + 1u;
  tea2(c0, c1);
  const T cpx = u32ToUnorm<T>(c0), cpy = u32ToUnorm<T>(c1);
  T obsScale[3] = {T(0), T(0), T(0)};
  T weightedObs = T(0);
  V3<T> bentAccum{T(0), T(0), T(0)};
  T hitDistSum = T(0);
  int nHit = 0;
  const T power = T(p->ao_falloff_power);
  for (int i = 0; i < nSamples; ++i) {
    T u1, u2;
    aoSample2d(p->ao_low_discrepancy != 0, base, i, nSamples, cpx, cpy, u1, u2);
    const V3<T> dir = cosineSampleHemisphere(u1, u2, f);
    if (dot(dir, Ns) < T(0.01f)) {
      cnt.aoSampled++;
      obsScale[0] += T(1);
      obsScale[1] += T(1);
      obsScale[2] += T(1);
      weightedObs += T(1);
      continue;
    }
    cnt.aoTraced++;
    const Hit<T> h = intersect(*c.s, O, dir, eps, R);
    const T t = h.kind >= 0 ? h.t : std::numeric_limits<T>::infinity();
    T perSample = T(0);
    for (int k = 0; k < K; ++k) {
      if (t < radius[k]) {
        T fk = T(1);
        if (power > T(0)) {
          T a = T(1) - t / radius[k];
          if (a < T(0)) a = T(0);
          fk = std::pow(a, power);
        }
        obsScale[k] += fk;
        perSample += weight[k] * fk;
      }
    }
    weightedObs += perSample;
    bentAccum = bentAccum + dir * (T(1) - perSample);
    if (t < R) {
      hitDistSum += t;
      ++nHit;
    }
  }
  const T invN = T(1) / T(static_cast<float>(nSamples));
  r.openness = T(1) - weightedObs * invN;
  r.contact = T(1) - obsScale[0] * invN;
  if (p->ao_multi_scale) {
    const T wMid = weight[1] + weight[2];
    r.shape = T(1) - (weight[1] * obsScale[1] + weight[2] * obsScale[2]) / (wMid * T(static_cast<float>(nSamples)));
  } else {
    r.shape = r.openness;
  }
  r.bent = safeNormalize(bentAccum, Ns);
  r.avgHitDist = nHit > 0 ? hitDistSum / T(static_cast<float>(nHit)) : R;
  return r;
}

// ---- principled.hpp:33-305 -- the GGX principled subset, direct pass ----
template <class T>
inline T ggxD(T a2, T ndh) {  // principled.hpp:63-68
  if (ndh <= T(0)) return T(0);
  const T t = ndh * ndh * (a2 - T(1)) + T(1);
  return a2 / (T(3.14159265358979323846f) * t * t);
}
template <class T>
inline T ggxLambda(T a2, T cosT) {  // principled.hpp:72-76
  const T c2 = std::fmax(cosT * cosT, T(1.0e-12f));
  const T t2 = (T(1) - c2) / c2;
  return T(0.5f) * (std::sqrt(T(1) + a2 * t2) - T(1));
}
template <class T>
inline T ggxDAniso(T ax, T ay, T hx, T hy, T hz) {  // principled.hpp:87-92
  if (hz <= T(0)) return T(0);
  const T sx = hx / ax, sy = hy / ay;
  const T t = sx * sx + sy * sy + hz * hz;
  return T(1) / (T(3.14159265358979323846f) * ax * ay * t * t);
}
template <class T>
inline T ggxLambdaAniso(T ax, T ay, T vx, T vy, T vz) {  // principled.hpp:95-99
  const T z2 = std::fmax(vz * vz, T(1.0e-12f));
  const T t2 = (ax * ax * vx * vx + ay * ay * vy * vy) / z2;
  return T(0.5f) * (std::sqrt(T(1) + t2) - T(1));
}
template <class T>
inline V3<T> schlickF(V3<T> F0, T u) {  // principled.hpp:52-59
  T c = T(1) - u;
  if (c < T(0)) c = T(0);
  if (c > T(1)) c = T(1);
  const T c5 = c * c * c * c * c;
  return V3<T>{F0.x + (T(1) - F0.x) * c5, F0.y + (T(1) - F0.y) * c5, F0.z + (T(1) - F0.z) * c5};
}
template <class T>
inline T principledAspect(T anisotropy) {  // principled.hpp:79-83
  T a = std::fabs(anisotropy);
  if (a > T(1)) a = T(1);
  return std::sqrt(T(1) - T(0.9f) * a);
}

// principled.hpp:139-305 -- shadePrincipled (the traced-reflection flag is
// always off here: configurations that would set it are refused)
template <class T>
inline V3<T> shadePrincipled(const Ctx<T>& c, const Mat& mat, V3<T> C, V3<T> N, V3<T> V, V3<T> ambLight,
                             V3<T> aoFactor, T diffuseAo, V3<T> P, V3<T> Ng, T eps, bool shadowsOn,
                             int shadowSamples, uint32_t px, uint32_t py, const V3<T>* tangent,
                             Counters& cnt) {
  const T em = T(mat.emission), amb = T(mat.ambient);
  V3<T> out{em * C.x + aoFactor.x * amb * C.x * ambLight.x, em * C.y + aoFactor.y * amb * C.y * ambLight.y,
            em * C.z + aoFactor.z * amb * C.z * ambLight.z};
  const T kd = T(mat.diffuse) * (T(1) - T(mat.pbrMetallic));  // principledDiffuseWeight
  const T dd = T(0.08f) * T(mat.pbrSpecular), tm = T(mat.pbrMetallic);  // principledF0
  const V3<T> F0{dd + (C.x - dd) * tm, dd + (C.y - dd) * tm, dd + (C.z - dd) * tm};
  const T f0max = std::fmax(F0.x, std::fmax(F0.y, F0.z));
  const T kAlphaMin = T(1.0e-3f);  // kGgxDirectAlphaMin
  T alpha = T(mat.pbrRoughness) * T(mat.pbrRoughness);
  if (alpha < kAlphaMin) alpha = kAlphaMin;
  const T a2 = alpha * alpha;
  const T ndv = std::fmax(dot(N, V), T(1.0e-4f));
  const T lamV = ggxLambda(a2, ndv);
  const T kPi = T(3.14159265358979323846f);
  // principledAnisoCtx (principled.hpp:113-128)
  const bool an = tangent != nullptr && mat.pbrAnisotropy != 0.0f;
  V3<T> t1{T(0), T(0), T(0)}, t2{T(0), T(0), T(0)};
  T ax = T(0), ay = T(0);
  if (an) {
    t1 = *tangent;
    t2 = cross(N, t1);
    const T aspect = principledAspect(T(mat.pbrAnisotropy));
    ax = alpha / aspect;
    if (ax > T(1)) ax = T(1);
    ay = alpha * aspect;
    if (ax < kAlphaMin) ax = kAlphaMin;
    if (ay < kAlphaMin) ay = kAlphaMin;
  }
  const T lamVA = an ? ggxLambdaAniso(ax, ay, dot(V, t1), dot(V, t2), ndv) : T(0);
  auto specWeight = [&](V3<T> w, T ndlw, V3<T> H, T ndh) -> T {
    if (!an) {
      const T G2 = T(1) / (T(1) + lamV + ggxLambda(a2, ndlw));
      return kPi * ggxD(a2, ndh) * G2 / (T(4) * ndv);
    }
    const T G2 = T(1) / (T(1) + lamVA + ggxLambdaAniso(ax, ay, dot(w, t1), dot(w, t2), ndlw));
    return kPi * ggxDAniso(ax, ay, dot(H, t1), dot(H, t2), ndh) * G2 / (T(4) * ndv);
  };
  uint32_t s0 = px, s1 = py;
  uint32_t q0 = px ^ 0x50425244u, q1 = py;  // the area path's 'PBRD'-tagged stream
  for (const Light<T>& l : c.lights) {
    const T ndl = dot(N, l.L);
    if (ndl <= T(0)) continue;
    const bool areaPath = shadowsOn && shadowSamples > 1 && l.radius > T(0);
    if (!areaPath) {
      const T sf = shadowsOn ? computeShadow(c, P, Ng, N, eps, l, shadowSamples, s0, s1, cnt) : T(1);
      const V3<T> Lc{l.color.x * sf, l.color.y * sf, l.color.z * sf};
      const T dk = kd * ndl * diffuseAo;
      out.x += dk * C.x * Lc.x;
      out.y += dk * C.y * Lc.y;
      out.z += dk * C.z * Lc.z;
      if (!l.highlight) continue;
      if (f0max > T(0)) {
        const V3<T> H = normalize(V3<T>{l.L.x + V.x, l.L.y + V.y, l.L.z + V.z});
        const T ndh = dot(N, H);
        if (ndh > T(0)) {
          const T sW = specWeight(l.L, ndl, H, ndh);
          const V3<T> F = schlickF(F0, dot(V, H));
          out.x += sW * F.x * Lc.x;
          out.y += sW * F.y * Lc.y;
          out.z += sW * F.z * Lc.z;
        }
      }
    } else {
      V3<T> ng = (dot(Ng, N) < T(0)) ? neg(Ng) : Ng;
      ng = safeNormalize(ng, N);
      const V3<T> O = P + ng * eps;
      const T far = std::numeric_limits<T>::infinity();
      T accD = T(0);
      V3<T> accS{T(0), T(0), T(0)};
      for (int i = 0; i < shadowSamples; ++i) {
        tea2(q0, q1);
        const V3<T> w = sampleLightDir(l, u32ToUnorm<T>(q0), u32ToUnorm<T>(q1));
        const T ndlS = dot(N, w);
        if (ndlS <= T(0)) continue;
        cnt.shadow++;
        if (occluded(*c.s, O, w, eps, far)) continue;
        accD += ndlS;
        if (l.highlight && f0max > T(0)) {
          const V3<T> H = normalize(V3<T>{w.x + V.x, w.y + V.y, w.z + V.z});
          const T ndh = dot(N, H);
          if (ndh > T(0)) {
            const T sW = specWeight(w, ndlS, H, ndh);
            const V3<T> F = schlickF(F0, dot(V, H));
            accS.x += sW * F.x;
            accS.y += sW * F.y;
            accS.z += sW * F.z;
          }
        }
      }
      const T inv = T(1) / T(shadowSamples);
      const T dk = kd * diffuseAo * inv * accD;
      out.x += dk * C.x * l.color.x;
      out.y += dk * C.y * l.color.y;
      out.z += dk * C.z * l.color.z;
      out.x += inv * accS.x * l.color.x;
      out.y += inv * accS.y * l.color.y;
      out.z += inv * accS.z * l.color.z;
    }
  }
  // the fake environment specular: a converted POV reflection scalar, else
  // the constant F0 x background
  if (mat.reflection > 0.0f) {
    const T r = T(mat.reflection);
    out.x += r * c.bg.x;
    out.y += r * c.bg.y;
    out.z += r * c.bg.z;
  } else if (f0max > T(0)) {
    out.x += F0.x * c.bg.x;
    out.y += F0.y * c.bg.y;
    out.z += F0.z * c.bg.z;
  }
  return out;
}

// shading.hpp:76-192 (POV model; principled materials go to their own
// evaluator, shading.hpp:88-91)
template <class T>
inline V3<T> shadeLocal(const Ctx<T>& c, const Mat& mat, V3<T> C, V3<T> N, V3<T> V,
                        V3<T> ambLight, V3<T> aoFactor, T diffuseAo, V3<T> P, V3<T> Ng, T eps,
                        bool shadowsOn, int shadowSamples, uint32_t px, uint32_t py,
                        Counters& cnt, const V3<T>* tangent = nullptr) {
  if (mat.model == 1)
    return shadePrincipled(c, mat, C, N, V, ambLight, aoFactor, diffuseAo, P, Ng, eps, shadowsOn, shadowSamples,
                           px, py, tangent, cnt);
  const T em = T(mat.emission), amb = T(mat.ambient);
  V3<T> out{em * C.x + aoFactor.x * amb * C.x * ambLight.x,
            em * C.y + aoFactor.y * amb * C.y * ambLight.y,
            em * C.z + aoFactor.z * amb * C.z * ambLight.z};
  const T ex = blinnExp(T(mat.roughness));
  uint32_t s0 = px, s1 = py;
  for (const Light<T>& l : c.lights) {
    const T ndl = dot(N, l.L);
    if (ndl <= T(0)) continue;
    const T shadowFactor =
        shadowsOn ? computeShadow(c, P, Ng, N, eps, l, shadowSamples, s0, s1, cnt) : T(1);
    const V3<T> Lc{l.color.x * shadowFactor, l.color.y * shadowFactor, l.color.z * shadowFactor};
    T dd;
    if (T(mat.brilliance) == T(1)) dd = ndl;
    else if (T(mat.brilliance) == T(0)) dd = T(1);
    else dd = std::pow(ndl, T(mat.brilliance));
    const T dk = T(mat.diffuse) * dd * diffuseAo;
    out.x += dk * C.x * Lc.x;
    out.y += dk * C.y * Lc.y;
    out.z += dk * C.z * Lc.z;
    if (!l.highlight) continue;
    V3<T> hl;
    if (mat.metallic) {
      T cc = ndl;
      if (cc > T(1)) cc = T(1);
      const T x = std::acos(cc) * T(0.63661977f);
      T f = T(0.014567225f) / ((x - T(1.12f)) * (x - T(1.12f))) - T(0.011612903f);
      if (f < T(0)) f = T(0);
      if (f > T(1)) f = T(1);
      hl = V3<T>{Lc.x * (f + (T(1) - f) * C.x), Lc.y * (f + (T(1) - f) * C.y),
                 Lc.z * (f + (T(1) - f) * C.z)};
    } else {
      hl = Lc;
    }
    T specW = T(0);
    if (mat.specular > 0.0f) {
      const V3<T> H = normalize(V3<T>{l.L.x + V.x, l.L.y + V.y, l.L.z + V.z});
      const T nh = dot(N, H);
      if (nh > T(0)) specW += T(mat.specular) * std::pow(nh, ex);
    }
    if (mat.phong > 0.0f) {
      const V3<T> Rr = (T(2) * ndl) * N - l.L;
      const T rv = dot(Rr, V);
      if (rv > T(0) && (T(mat.phongSize) < T(60) || rv > T(0.0008f)))
        specW += T(mat.phong) * std::pow(rv, T(mat.phongSize));
    }
    if (specW > T(0)) {
      const T sc = specW * T(c.p->specular_scale);
      out.x += sc * hl.x;
      out.y += sc * hl.y;
      out.z += sc * hl.z;
    }
  }
  if (mat.reflection > 0.0f) {
    const T r = T(mat.reflection);
    out.x += r * c.bg.x;
    out.y += r * c.bg.y;
    out.z += r * c.bg.z;
  }
  return out;
}

template <class T>
struct HitShade {
  V3<T> color;
  T opacity;
  T aoOpen;
  V3<T> normal;   // face-forwarded shading normal (the adaptive-AA mask's)
  uint32_t group; // section group (BuiltScene *Group tables, Mesh::groupForTri)
  V3<T> giRefl{T(0), T(0), T(0)};  // diffuse reflectance of a GI receiver
  bool giElig = false;
  // the normal AOV under GI (hit_shader.hpp:214-220 mesh, :364 receivers)
  V3<T> giGuideN{T(0), T(0), T(0)};
};

template <class T>
struct AoShadeT {
  V3<T> aoFactor{T(1), T(1), T(1)};
  V3<T> ambLight;
  T diffuseAo = T(1);
  T openness = T(1);
};

// ao_shade.hpp:44-141 -- gather (legacy binary or enhanced), then the shading
// factors: bent-normal sky/ground ambient, per-channel (multibounce) aoFactor,
// direct-diffuse darkening
template <class T>
inline AoShadeT<T> aoShade(const Ctx<T>& c, V3<T> P, V3<T> Ng, V3<T> N, V3<T> C, T eps, T hitT, uint32_t px,
                           uint32_t py, Counters& cnt) {
  const ubg_render_params* p = c.p;
  AoShadeT<T> r;
  r.ambLight = c.ambLight;
  if (p->ao_samples <= 0) return r;
  const bool enh = aoEnhanced<T>(p);
  T openness = T(1);
  V3<T> bent = N;
  // hit_shader.hpp:124-140: the coarse grid's bilateral lookup, else the
  // exact inline gather with the hi-res seeds
  const int latW = c.seedW;
  const int latH = static_cast<int>(static_cast<int64_t>(latW) * c.hHi / c.wHi);
  const bool coarse = c.grid != nullptr && sampleCoarseAo(*c.grid, latW, latH, px, py, N, hitT, openness, bent);
  if (coarse) {
  } else if (enh) {
    const AOQuality<T> q = computeAOQuality(c, P, Ng, N, eps, p->ao_samples * c.aoMul, px, py, c.seedW, cnt);
    openness = q.openness;
    bent = q.bent;
  } else {
    openness = computeAO(c, P, Ng, N, eps, p->ao_samples * c.aoMul, T(p->ao_distance), px, py, c.seedW, cnt);
  }
  r.openness = openness;
  if (enh && p->ao_bent_normal) {
    const T w = T(0.5f) * (dot(bent, c.aoUp) + T(1));
    const T gx = T(p->ao_ground_color[0]), sx = T(p->ao_sky_color[0]);
    const T gy = T(p->ao_ground_color[1]), sy = T(p->ao_sky_color[1]);
    const T gz = T(p->ao_ground_color[2]), sz = T(p->ao_sky_color[2]);
    r.ambLight = V3<T>{c.ambLight.x * (gx + (sx - gx) * w), c.ambLight.y * (gy + (sy - gy) * w),
                       c.ambLight.z * (gz + (sz - gz) * w)};
  }
  const T I = T(p->ao_intensity);
  if (p->ao_multibounce) {
    const T ox = aoMultibounce(openness, C.x), oy = aoMultibounce(openness, C.y),
            oz = aoMultibounce(openness, C.z);
    r.aoFactor = V3<T>{T(1) - I * (T(1) - ox), T(1) - I * (T(1) - oy), T(1) - I * (T(1) - oz)};
  } else {
    const T s = T(1) - I * (T(1) - openness);
    r.aoFactor = V3<T>{s, s, s};
  }
  if (p->ao_diffuse_factor > 0.0f) r.diffuseAo = T(1) - T(p->ao_diffuse_factor) * (T(1) - openness);
  return r;
}

// hit_shader.hpp:147-496
template <class T>
inline HitShade<T> shadeHit(const Ctx<T>& c, const Hit<T>& h, V3<T> rd, V3<T> org, uint32_t px,
                            uint32_t py, Counters& cnt) {
  const Geometry& g = c.s->g;
  HitShade<T> hs;
  const bool shadowsActive = c.p->shadows != 0;
  const V3<T> V = normalize(neg(rd));
  if (h.kind == UBG_KIND_MESH) {
    const uint32_t* ix = &g.idx[3 * h.id];
    const T w0 = T(1) - h.u - h.v;
    V3<T> nb{T(0), T(0), T(0)};
    T cb[4];
    for (int k = 0; k < 3; ++k) {
      const T a0 = T(g.nrm[3 * ix[0] + k]), a1 = T(g.nrm[3 * ix[1] + k]), a2 = T(g.nrm[3 * ix[2] + k]);
      const T val = w0 * a0 + h.u * a1 + h.v * a2;
      if (k == 0) nb.x = val;
      else if (k == 1) nb.y = val;
      else nb.z = val;
    }
    for (int k = 0; k < 4; ++k) {
      const T a0 = T(g.col[4 * ix[0] + k]), a1 = T(g.col[4 * ix[1] + k]), a2 = T(g.col[4 * ix[2] + k]);
      cb[k] = w0 * a0 + h.u * a1 + h.v * a2;
    }
    const V3<T> Ng = h.ng;
    const V3<T> NgF = faceForward(Ng, rd);
    V3<T> N = normalize(nb);
    if (dot(N, NgF) < T(0)) N = neg(N);
    const V3<T> C{cb[0], cb[1], cb[2]};
    const int64_t baseTri = static_cast<int64_t>(h.id % static_cast<uint32_t>(g.baseT));
    const Mat& mat = g.triMat.empty() ? g.meshMats[0] : g.meshMats[g.triMat[baseTri]];
    const V3<T> P{org.x + rd.x * h.t, org.y + rd.y * h.t, org.z + rd.z * h.t};
    const T secEps = selfIntersectEps(P, rd, h.t);
    AoShadeT<T> ao = aoShade(c, P, Ng, N, C, secEps, h.t, px, py, cnt);
    // hit_shader.hpp:229-243: under GI a mesh hit gathers indirect instead of
    // its constant ambient (toon finishes: unit ambient, no indirect)
    if (c.gi) hs.giGuideN = N;
    if (c.gi && toonLike(mat)) {
      ao.ambLight = V3<T>{T(1), T(1), T(1)};
    } else if (c.gi) {
      const T kd = T(diffuseWeight(mat));
      hs.giRefl = V3<T>{kd * C.x, kd * C.y, kd * C.z};
      hs.giElig = true;
      ao.ambLight = V3<T>{T(0), T(0), T(0)};
    }
    hs.aoOpen = ao.openness;
    hs.color = shadeLocal(c, mat, C, N, V, ao.ambLight, ao.aoFactor, ao.diffuseAo, P, Ng, secEps,
                          shadowsActive, c.p->shadow_samples * c.shMul, px, py, cnt);
    hs.opacity = cb[3];
    hs.normal = N;
    // Mesh::groupForTri by base triangle (an instance copy's primID would
    // index past the per-triangle table)
    hs.group = g.triGroup.empty() ? 0u : g.triGroup[baseTri];
    return hs;
  }
  const PrimSide* side = h.kind == UBG_KIND_SPHERE        ? &g.sphSide[h.id]
                         : h.kind == UBG_KIND_CONE_CURVE ? &g.cSide[h.id]
                                                         : &g.rSide[h.id];
  const Mat& pm = g.primMats[side->mat];
  const V3<T> N = faceForward(normalize(h.ng), rd);
  const V3<T> C{T(side->color[0]), T(side->color[1]), T(side->color[2])};
  const V3<T> P{org.x + rd.x * h.t, org.y + rd.y * h.t, org.z + rd.z * h.t};
  const V3<T> Ng = h.ng;
  const T secEps = selfIntersectEps(P, rd, h.t);
  const bool fromEdge = side->fromEdge != 0;
  AoShadeT<T> ao;
  ao.ambLight = c.ambLight;
  if (!fromEdge) {
    ao = aoShade(c, P, Ng, N, C, secEps, h.t, px, py, cnt);
    // hit_shader.hpp:349-368: a real primitive is a GI receiver like the mesh
    if (c.gi && toonLike(pm)) {
      ao.ambLight = V3<T>{T(1), T(1), T(1)};
    } else if (c.gi) {
      const T kd = T(diffuseWeight(pm));
      hs.giRefl = V3<T>{kd * C.x, kd * C.y, kd * C.z};
      hs.giElig = true;
      hs.giGuideN = N;
      ao.ambLight = V3<T>{T(0), T(0), T(0)};
    }
  }
  hs.aoOpen = ao.openness;
  const bool primShadows = !fromEdge && shadowsActive;
  // hit_shader.hpp:373-425: the anisotropy tangent of a principled sphere
  // (meridian of the world +Y pole frame) or cylinder (its axis projected on
  // the tangent plane); poles, caps and degenerate projections stay isotropic
  V3<T> anisoT{T(0), T(0), T(0)};
  bool tanValid = false;
  if (pm.model == 1 && pm.pbrAnisotropy != 0.0f && !fromEdge) {
    V3<T> t1{T(0), T(0), T(0)};
    if (h.kind == UBG_KIND_SPHERE) {
      if (std::fabs(N.y) < T(0.999f)) {
        const V3<T> tPhi = safeNormalize(cross(V3<T>{T(0), T(1), T(0)}, N));
        t1 = cross(N, tPhi);
        tanValid = true;
      }
    } else {
      const V3<T> A{T(side->axis[0]), T(side->axis[1]), T(side->axis[2])};
      const V3<T> ngu = safeNormalize(Ng, N);
      if (std::fabs(dot(ngu, A)) < T(0.99f)) {
        const T andn = dot(A, N);
        const V3<T> proj{A.x - N.x * andn, A.y - N.y * andn, A.z - N.z * andn};
        if (dot(proj, proj) >= T(1.0e-12f)) {
          t1 = safeNormalize(proj);
          tanValid = true;
        }
      }
    }
    if (tanValid) {
      if (pm.pbrRotation != 0.0f) {
        const T ang = T(pm.pbrRotation) * T(6.2831853072f);
        const V3<T> t2 = cross(N, t1);
        const T ca = std::cos(ang), sa = std::sin(ang);
        t1 = V3<T>{ca * t1.x + sa * t2.x, ca * t1.y + sa * t2.y, ca * t1.z + sa * t2.z};
      }
      anisoT = t1;
    }
  }
  hs.color = shadeLocal(c, pm, C, N, V, ao.ambLight, ao.aoFactor, ao.diffuseAo, P, Ng, secEps, primShadows,
                        c.p->shadow_samples * c.shMul, px, py, cnt, tanValid ? &anisoT : nullptr);
  hs.opacity = T(side->color[3]);
  hs.normal = N;
  hs.group = side->group;
  if (h.kind != UBG_KIND_SPHERE && side->opacity1 >= 0.0f) {
    T u = h.u;
    if (u < T(0)) u = T(0);
    if (u > T(1)) u = T(1);
    hs.opacity = T(side->color[3]) * (T(1) - u) + T(side->opacity1) * u;
  }
  return hs;
}

template <class T>
struct PixelOut {
  T r, g, b, a, depth, viewZ, aoOpen;
  uint32_t hitId;
  // adaptive-AA mask inputs (transparency.hpp:402-413): first-hit shading
  // normal (0 on a miss) and (kind << 16 | group), all ones on a miss
  V3<T> normal{T(0), T(0), T(0)};
  uint32_t gid = 0xFFFFFFFFu;
  // GI seed of the first hit (transparency.hpp:405-417)
  V3<T> worldPos{T(0), T(0), T(0)};
  V3<T> giRefl{T(0), T(0), T(0)};
  bool giElig = false;
  V3<T> giGuideN{T(0), T(0), T(0)};
};

// transparency.hpp:311-522 (grid path; clip t-range per :118-129)
template <class T>
inline PixelOut<T> integratePixel(const Ctx<T>& c, V3<T> org, V3<T> rd, uint32_t px, uint32_t py,
                                  Counters& cnt) {
  const ubg_render_params* p = c.p;
  const T inf = std::numeric_limits<T>::infinity();
  T clipNear = T(0), clipFar = inf;
  {
    const bool off = !(p->clip_near > -std::numeric_limits<float>::infinity()) &&
                     !(p->clip_far < std::numeric_limits<float>::infinity());
    if (!off) {
      const T cosr = dot(rd, c.camDir);
      if (cosr > T(1.0e-6f)) {
        if (p->clip_near > 0.0f) clipNear = T(p->clip_near) / cosr;
        if (p->clip_far < std::numeric_limits<float>::infinity()) clipFar = T(p->clip_far) / cosr;
      }
    }
  }
  V3<T> Cf{T(0), T(0), T(0)};
  T Af = T(0);
  T nearDepth = T(0);
  T firstViewZ = T(0);
  T firstAo = T(1);
  uint32_t firstId = 0xFFFFFFFFu;
  V3<T> firstN{T(0), T(0), T(0)};
  uint32_t firstGid = 0xFFFFFFFFu;
  V3<T> firstPos{T(0), T(0), T(0)}, firstRefl{T(0), T(0), T(0)}, firstGuideN{T(0), T(0), T(0)};
  bool firstElig = false;
  V3<T> base = c.bg;
  T baseCov = p->transparent_background ? T(0) : T(1);
  const T kOpaque = T(1.0f - 1e-4f);
  T tnear = clipNear;
  const int maxIters = p->transparency ? (p->max_transparent_layers + 1) : 1;
  for (int iter = 0; iter < maxIters; ++iter) {
    cnt.primary++;
    const Hit<T> h = intersect(*c.s, org, rd, tnear, clipFar);
    if (h.kind < 0) break;
    cnt.shaded++;
    const HitShade<T> hs = shadeHit(c, h, rd, org, px, py, cnt);
    if (nearDepth == T(0)) {
      nearDepth = h.t;
      firstViewZ = h.t * dot(rd, c.camDir);
      firstAo = hs.aoOpen;
      firstId = (static_cast<uint32_t>(h.kind) << 30) | h.id;
      firstN = hs.normal;
      firstGid = (static_cast<uint32_t>(h.kind) << 16) | hs.group;
      firstPos = V3<T>{org.x + rd.x * h.t, org.y + rd.y * h.t, org.z + rd.z * h.t};
      firstRefl = hs.giRefl;
      firstElig = hs.giElig;
      firstGuideN = hs.giGuideN;
    }
    if (!p->transparency || hs.opacity >= kOpaque) {
      base = hs.color;
      baseCov = T(1);
      break;
    }
    const T w = (T(1) - Af) * hs.opacity;
    Cf.x += w * hs.color.x;
    Cf.y += w * hs.color.y;
    Cf.z += w * hs.color.z;
    Af += w;
    if (Af >= kOpaque) break;
    const V3<T> hitP{org.x + rd.x * h.t, org.y + rd.y * h.t, org.z + rd.z * h.t};
    const T dirMax = std::fmax(std::fabs(rd.x), std::fmax(std::fabs(rd.y), std::fabs(rd.z)));
    const T epsScale = std::fmax(std::fmax(std::fabs(hitP.x), std::fabs(hitP.y)),
                                 std::fmax(std::fabs(hitP.z), dirMax * h.t));
    tnear = h.t + epsScale * T(0x1.0fp-21f);
    if (iter == maxIters - 1) cnt.capped++;
  }
  const T floorW = (T(1) - Af) * baseCov;
  T outA = Af + floorW;
  outA = std::fmin(T(1), std::fmax(T(0), outA));
  PixelOut<T> o;
  o.r = Cf.x + floorW * base.x;
  o.g = Cf.y + floorW * base.y;
  o.b = Cf.z + floorW * base.z;
  o.a = outA;
  o.depth = nearDepth;
  o.viewZ = firstViewZ;
  o.aoOpen = firstAo;
  o.hitId = firstId;
  o.normal = firstN;
  o.gid = firstGid;
  o.worldPos = firstPos;
  o.giRefl = firstRefl;
  o.giElig = firstElig;
  o.giGuideN = firstGuideN;
  return o;
}

// fog.cpp:7-15
inline float fogFactor(const ubg_fog& fog, float z) {
  if (!fog.enabled) return 1.0f;
  const float denom = fog.end - fog.start;
  if (denom <= 0.0f) return z < fog.end ? 1.0f : 0.0f;
  float f = (fog.end - z) / denom;
  if (f < 0.0f) return 0.0f;
  if (f > 1.0f) return 1.0f;
  return f;
}

// image_ops.cpp:9-35
void boxDownsample(const float* src, int w, int h, int channels, int ss, float* out) {
  const int ow = w / ss, oh = h / ss;
  const float inv = 1.0f / static_cast<float>(ss * ss);
  for (int y = 0; y < oh; ++y)
    for (int x = 0; x < ow; ++x)
      for (int c = 0; c < channels; ++c) {
        float acc = 0.0f;
        for (int dy = 0; dy < ss; ++dy)
          for (int dx = 0; dx < ss; ++dx) {
            const std::size_t si =
                (static_cast<std::size_t>(y * ss + dy) * w + (x * ss + dx)) * channels + c;
            const float v = src[si];
            acc += std::isfinite(v) ? v : 0.0f;
          }
        out[(static_cast<std::size_t>(y) * ow + x) * channels + c] = acc * inv;
      }
}

// ---------------------------------------------------------------------------
// pt1 GI gather (integrator/pt1/pt1_gather.hpp), restated.

// irradiance_cache.hpp:176-184
inline uint32_t hashU32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7FEB352Du;
  x ^= x >> 15;
  x *= 0x846CA68Bu;
  x ^= x >> 16;
  return x;
}

// The pt1 gather's fixed inputs (IrradianceCacheParams as runPt1GiPass fills
// them, embree_renderer.cpp:581-611).
template <class T>
struct GiParams {
  V3<T> sky, ground, up;
  float envIntensity = 1.0f;
  int bounces = 1;
  T maxDist;
  // pt2 (giIntegrator 2): Sobol-Burley first bounce with this pixel
  // arrangement, the path's stream for area-light NEE, emission at vertices
  bool pt2 = false;
  int pattern = 1;
  bool emissive = true;
};

// pt2_sampler.hpp -- hash-based Owen-scrambled Sobol (Burley, "Practical
// Hash-based Owen Scrambling", JCGT 9(4) 2020): the Laine-Karras style hash
// on reversed bits, its base-4 variant for the Morton pixel shuffle, and the
// second Sobol dimension's direction numbers (rows of Pascal's triangle mod 2
// in the reversed-bit convention).
inline uint32_t owenHashRev(uint32_t n, uint32_t seed) {
  n ^= n * 0x3d20adeau;
  n += seed;
  n *= (seed >> 16) | 1u;
  n ^= n * 0x05526c56u;
  n ^= n * 0x53a22864u;
  return n;
}
inline uint32_t owenHashRevBase4(uint32_t n, uint32_t seed) {
  n ^= n * 0x3d20adeau;
  n ^= (n >> 1) & (n << 1) & 0x55555555u;
  n += seed;
  n *= (seed >> 16) | 1u;
  n ^= (n >> 1) & (n << 1) & 0x55555555u;
  n ^= n * 0x05526c56u;
  n ^= n * 0x53a22864u;
  return n;
}
inline uint32_t interleaveBits16(uint32_t x) {
  x &= 0x0000ffffu;
  x = (x ^ (x << 8)) & 0x00ff00ffu;
  x = (x ^ (x << 4)) & 0x0f0f0f0fu;
  x = (x ^ (x << 2)) & 0x33333333u;
  x = (x ^ (x << 1)) & 0x55555555u;
  return x;
}
inline float sobolToUnit(uint32_t n) { return static_cast<float>(n) * (1.0f / 4294967808.0f); }
inline float sobolBurleyDim(uint32_t revIdx, int dim, uint32_t seed) {
  uint32_t r;
  if (dim == 0) {
    r = reverseBits32(revIdx);
  } else {
    r = 0;
    uint32_t row = 1;  // Pascal's triangle mod 2, one row per bit from the top
    for (int b = 31; b >= 0; --b) {
      if ((revIdx >> b) & 1u) r ^= row;
      row ^= row << 1;
    }
  }
  return sobolToUnit(reverseBits32(owenHashRev(r, seed)));
}
inline void sobolBurley2D(uint32_t index, uint32_t seed, float& u1, float& u2) {
  index = owenHashRev(reverseBits32(index), seed ^ 0xf8ade99au);
  u1 = sobolBurleyDim(index, 0, seed ^ 0xe0aaaf76u);
  u2 = sobolBurleyDim(index, 1, seed ^ 0x94964d4eu);
}
// pt2MakePixelSampler: blue noise = one global sequence, pixel sections of
// sppPow2 samples ordered along the base-4 scrambled Morton curve; sobol = an
// independent sequence per pixel
inline void pt2PixelSampler(int pattern, uint32_t px, uint32_t py, uint32_t frameSeed, uint32_t pixelHash,
                            uint32_t sppPow2, uint32_t& indexBase, uint32_t& seed) {
  if (pattern == 1) {
    const uint32_t m = (interleaveBits16(px) << 1) | interleaveBits16(py);
    indexBase = reverseBits32(owenHashRevBase4(reverseBits32(m), frameSeed)) * sppPow2;
    seed = frameSeed;
  } else {
    indexBase = 0;
    seed = pixelHash;
  }
}
inline uint32_t nextPow2(uint32_t n) {
  uint32_t q = 1;
  while (q < n) q <<= 1;
  return q;
}

template <class T>
struct Pt1Vertex {
  V3<T> radiance, emitted, albedo, P, N;
  T tfar;
};

// pt1_gather.hpp:117-212 (pt1: no emission, the hard NEE ray per light)
template <class T>
static bool pt1EvalVertex(const Ctx<T>& c, const Hit<T>& h, V3<T> O, V3<T> wi, Pt1Vertex<T>& v, Counters& cnt,
                          bool addEmission = false, uint32_t* softRng0 = nullptr, uint32_t* softRng1 = nullptr) {
  const Geometry& g = c.s->g;
  V3<T> Ny, Cy, NgS;
  T kd;
  float em = 0.0f;
  if (h.kind == UBG_KIND_MESH) {
    const uint32_t* ix = &g.idx[3 * h.id];
    const T w0 = T(1) - h.u - h.v;
    T nb[3], cb[3];
    for (int k = 0; k < 3; ++k) {
      nb[k] = w0 * T(g.nrm[3 * ix[0] + k]) + h.u * T(g.nrm[3 * ix[1] + k]) + h.v * T(g.nrm[3 * ix[2] + k]);
      cb[k] = w0 * T(g.col[4 * ix[0] + k]) + h.u * T(g.col[4 * ix[1] + k]) + h.v * T(g.col[4 * ix[2] + k]);
    }
    Ny = normalize(V3<T>{nb[0], nb[1], nb[2]});
    Cy = V3<T>{cb[0], cb[1], cb[2]};
    const int64_t baseTri = static_cast<int64_t>(h.id % static_cast<uint32_t>(g.baseT));
    const Mat& mat = g.triMat.empty() ? g.meshMats[0] : g.meshMats[g.triMat[baseTri]];
    kd = T(diffuseWeight(mat));
    if (addEmission) em = mat.emission;
    NgS = h.ng;
  } else {
    const PrimSide* side = h.kind == UBG_KIND_SPHERE        ? &g.sphSide[h.id]
                           : h.kind == UBG_KIND_CONE_CURVE ? &g.cSide[h.id]
                                                           : &g.rSide[h.id];
    if (side->fromEdge) return false;
    Ny = safeNormalize(h.ng);
    Cy = V3<T>{T(side->color[0]), T(side->color[1]), T(side->color[2])};
    kd = T(diffuseWeight(g.primMats[side->mat]));
    if (addEmission) em = g.primMats[side->mat].emission;
    NgS = Ny;
  }
  if (dot(Ny, wi) > T(0)) Ny = neg(Ny);
  const V3<T> Py{O.x + wi.x * h.t, O.y + wi.y * h.t, O.z + wi.z * h.t};
  const T eps = selfIntersectEps(Py, wi, h.t);
  V3<T> E{T(0), T(0), T(0)};
  for (const Light<T>& l : c.lights) {
    const T ndl = dot(Ny, l.L);
    if (ndl <= T(0)) continue;
    // computeShadow's one hard ray (secondary_rays.hpp:118-124), or (pt2, an
    // area light) one cone-jittered ray drawn from the path's stream
    V3<T> ng = (dot(NgS, Ny) < T(0)) ? neg(NgS) : NgS;
    ng = safeNormalize(ng, Ny);
    const V3<T> So = Py + ng * eps;
    V3<T> ldir = l.L;
    if (softRng0 && softRng1 && l.radius > T(0)) {
      tea2(*softRng0, *softRng1);
      ldir = sampleLightDir(l, u32ToUnorm<T>(*softRng0), u32ToUnorm<T>(*softRng1));
    }
    cnt.giNee++;
    const T sh = occluded(*c.s, So, ldir, eps, std::numeric_limits<T>::infinity()) ? T(0) : T(1);
    E.x += ndl * l.color.x * sh;
    E.y += ndl * l.color.y * sh;
    E.z += ndl * l.color.z * sh;
  }
  v.radiance = V3<T>{kd * Cy.x * E.x, kd * Cy.y * E.y, kd * Cy.z * E.z};
  v.emitted = em > 0.0f ? V3<T>{T(em) * Cy.x, T(em) * Cy.y, T(em) * Cy.z} : V3<T>{T(0), T(0), T(0)};
  v.albedo = V3<T>{kd * Cy.x, kd * Cy.y, kd * Cy.z};
  v.P = Py;
  v.N = Ny;
  v.tfar = h.t;
  return true;
}

// pt1_gather.hpp:228-485 (pt1: no Sobol, reservoir, emissive NEE or halves)
template <class T>
static V3<T> pt1GatherPoint(const Ctx<T>& c, const GiParams<T>& gp, V3<T> P, V3<T> N, V3<T> Ng, int spp,
                            uint32_t seed, T epsT, bool ld, float clampLum, T* outOcc, Counters& cnt,
                            uint32_t sobolBase = 0, uint32_t sobolSeed = 0) {
  const Frame<T> f = frameFromNormal(N);
  const T eps = selfIntersectEps(P, N, epsT);
  const V3<T> O = P + N * eps;
  const int maxB = gp.bounces < 1 ? 1 : gp.bounces;
  T cpx = T(0), cpy = T(0);
  if (ld) {
    uint32_t c0 = seed, c1 = 0x9E3779B9u;
    tea2(c0, c1);
    cpx = u32ToUnorm<T>(c0);
    cpy = u32ToUnorm<T>(c1);
  }
  V3<T> Esum{T(0), T(0), T(0)};
  int nOcc = 0;
  for (int si = 0; si < spp; ++si) {
    uint32_t s0 = seed, s1 = static_cast<uint32_t>(si);
    tea2(s0, s1);
    T u1, u2;
    if (gp.pt2) {
      float a, b2;
      sobolBurley2D(sobolBase + static_cast<uint32_t>(si), sobolSeed, a, b2);
      u1 = T(a);
      u2 = T(b2);
    } else if (ld) {
      u1 = T(static_cast<float>(si)) / T(static_cast<float>(spp)) + cpx;
      u2 = T(static_cast<float>(reverseBits32(static_cast<uint32_t>(si)))) * T(0x1.0p-32f) + cpy;
      if (u1 >= T(1)) u1 -= T(1);
      if (u2 >= T(1)) u2 -= T(1);
    } else {
      u1 = u32ToUnorm<T>(s0);
      u2 = u32ToUnorm<T>(s1);
    }
    V3<T> wi = cosineSampleHemisphere(u1, u2, f);
    const bool horizonOk = dot(wi, Ng) > T(0);
    if (!horizonOk) ++nOcc;
    V3<T> org = O;
    T tnear = eps;
    V3<T> tp{T(1), T(1), T(1)};
    V3<T> L{T(0), T(0), T(0)};
    for (int b = 1; horizonOk && b <= maxB; ++b) {
      cnt.giRays++;
      const Hit<T> h = intersect(*c.s, org, wi, tnear, gp.maxDist);
      if (h.kind < 0) {
        // irradiance_cache.hpp:188-194
        const T w = T(0.5f) * (dot(wi, gp.up) + T(1));
        const T e = T(gp.envIntensity);
        L.x += tp.x * (e * c.ambLight.x * (gp.ground.x + (gp.sky.x - gp.ground.x) * w));
        L.y += tp.y * (e * c.ambLight.y * (gp.ground.y + (gp.sky.y - gp.ground.y) * w));
        L.z += tp.z * (e * c.ambLight.z * (gp.ground.z + (gp.sky.z - gp.ground.z) * w));
        break;
      }
      if (b == 1) ++nOcc;
      Pt1Vertex<T> v;
      if (!pt1EvalVertex(c, h, org, wi, v, cnt, gp.pt2 && gp.emissive, gp.pt2 ? &s0 : nullptr,
                         gp.pt2 ? &s1 : nullptr))
        break;
      L.x += tp.x * (v.radiance.x + v.emitted.x);
      L.y += tp.y * (v.radiance.y + v.emitted.y);
      L.z += tp.z * (v.radiance.z + v.emitted.z);
      if (b == maxB) break;
      tp.x *= v.albedo.x;
      tp.y *= v.albedo.y;
      tp.z *= v.albedo.z;
      if (b >= 2) {
        tea2(s0, s1);
        T pc = std::fmax(tp.x, std::fmax(tp.y, tp.z));
        pc = std::fmin(T(0.95f), std::fmax(T(0.05f), pc));
        if (u32ToUnorm<T>(s0) >= pc) break;
        const T inv = T(1) / pc;
        tp.x *= inv;
        tp.y *= inv;
        tp.z *= inv;
      }
      tea2(s0, s1);
      const Frame<T> fv = frameFromNormal(v.N);
      wi = cosineSampleHemisphere(u32ToUnorm<T>(s0), u32ToUnorm<T>(s1), fv);
      const T epsV = selfIntersectEps(v.P, wi, v.tfar);
      org = V3<T>{v.P.x + v.N.x * epsV, v.P.y + v.N.y * epsV, v.P.z + v.N.z * epsV};
      tnear = epsV;
    }
    if (clampLum > 0.0f) {
      const T lum = T(0.2126f) * L.x + T(0.7152f) * L.y + T(0.0722f) * L.z;
      if (lum > T(clampLum)) {
        const T k = T(clampLum) / lum;
        L.x *= k;
        L.y *= k;
        L.z *= k;
      }
    }
    Esum.x += L.x;
    Esum.y += L.y;
    Esum.z += L.z;
  }
  const T invN = T(1) / T(static_cast<float>(spp > 0 ? spp : 1));
  if (outOcc) *outOcc = T(static_cast<float>(nOcc)) * invN;
  return V3<T>{Esum.x * invN, Esum.y * invN, Esum.z * invN};
}

// ---------------------------------------------------------------------------
// Adaptive antialiasing mask (render/adaptive_aa.hpp), restated.

// What the mask reads of one output pixel's centre result (PixelResult's
// r, g, b, a, viewZ, worldNormal, firstGeomID, firstGroup). The geometry id
// is the hit's kind: the reference attaches one Embree geometry per kind
// (scene_build.cpp:158-161), so equal kinds <=> equal geomIDs.
struct AaCenter {
  float r, g, b, a, viewZ;
  float n[3];
  uint32_t geom;   // 0xFFFFFFFF = background
  int32_t group;   // -1 = background
};

template <class T>
static AaCenter aaCenterOf(const PixelOut<T>& po) {
  AaCenter c;
  c.r = static_cast<float>(po.r);
  c.g = static_cast<float>(po.g);
  c.b = static_cast<float>(po.b);
  c.a = static_cast<float>(po.a);
  c.viewZ = static_cast<float>(po.viewZ);
  c.n[0] = static_cast<float>(po.normal.x);
  c.n[1] = static_cast<float>(po.normal.y);
  c.n[2] = static_cast<float>(po.normal.z);
  c.geom = po.gid == 0xFFFFFFFFu ? 0xFFFFFFFFu : (po.gid >> 16);
  c.group = po.gid == 0xFFFFFFFFu ? -1 : static_cast<int32_t>(po.gid & 0xFFFFu);
  return c;
}

// adaptive_aa.hpp:20-29 defaults (normal delta 0.3, depth k 3, 12 pixel
// worlds) and :66-115 (aaPairNeedsRefine)
static bool aaPairNeedsRefine(const AaCenter& a, const AaCenter& b, float za0, float zb1, float pws, float threshold,
                              bool perspective) {
  if (a.geom != b.geom || a.group != b.group) return true;
  if (std::fabs(a.r - b.r) > threshold || std::fabs(a.g - b.g) > threshold || std::fabs(a.b - b.b) > threshold ||
      std::fabs(a.a - b.a) > threshold)
    return true;
  const bool aHit = a.geom != 0xFFFFFFFFu, bHit = b.geom != 0xFFFFFFFFu;
  if (!aHit || !bHit) return false;
  const float la = a.n[0] * a.n[0] + a.n[1] * a.n[1] + a.n[2] * a.n[2];
  const float lb = b.n[0] * b.n[0] + b.n[1] * b.n[1] + b.n[2] * b.n[2];
  if (la > 0.0f && lb > 0.0f) {
    const float d = a.n[0] * b.n[0] + a.n[1] * b.n[1] + a.n[2] * b.n[2];
    if (1.0f - d > 0.3f) return true;
  }
  const float step = std::fabs(b.viewZ - a.viewZ);
  const float sa = std::fabs(a.viewZ - za0);
  const float sb = std::fabs(zb1 - b.viewZ);
  const float slope = sa < sb ? sa : sb;
  float absTerm = 12.0f * pws;
  if (perspective) absTerm *= std::fabs(a.viewZ) < std::fabs(b.viewZ) ? std::fabs(a.viewZ) : std::fabs(b.viewZ);
  const float thresh = 3.0f * slope > absTerm ? 3.0f * slope : absTerm;
  return step > thresh;
}

// adaptive_aa.hpp:128-194 (pairs along both axes with clamped outer slopes)
// and :36-57 (3x3 dilation); the edges-mode hi-res test is not part of it
static std::vector<uint8_t> buildAaMask(int Wf, int Hf, const AaCenter* c, float pws, float threshold,
                                        bool perspective) {
  std::vector<uint8_t> raw(static_cast<std::size_t>(Wf) * Hf, 0);
  auto at = [&](int x, int y) -> const AaCenter& { return c[static_cast<std::size_t>(y) * Wf + x]; };
  for (int y = 0; y < Hf; ++y)
    for (int x = 0; x + 1 < Wf; ++x) {
      const AaCenter& a = at(x, y);
      const AaCenter& b = at(x + 1, y);
      const float za0 = x > 0 ? at(x - 1, y).viewZ : 2.0f * a.viewZ - b.viewZ;
      const float zb1 = x + 2 < Wf ? at(x + 2, y).viewZ : 2.0f * b.viewZ - a.viewZ;
      if (aaPairNeedsRefine(a, b, za0, zb1, pws, threshold, perspective)) {
        raw[static_cast<std::size_t>(y) * Wf + x] = 1;
        raw[static_cast<std::size_t>(y) * Wf + x + 1] = 1;
      }
    }
  for (int y = 0; y + 1 < Hf; ++y)
    for (int x = 0; x < Wf; ++x) {
      const AaCenter& a = at(x, y);
      const AaCenter& b = at(x, y + 1);
      const float za0 = y > 0 ? at(x, y - 1).viewZ : 2.0f * a.viewZ - b.viewZ;
      const float zb1 = y + 2 < Hf ? at(x, y + 2).viewZ : 2.0f * b.viewZ - a.viewZ;
      if (aaPairNeedsRefine(a, b, za0, zb1, pws, threshold, perspective)) {
        raw[static_cast<std::size_t>(y) * Wf + x] = 1;
        raw[static_cast<std::size_t>(y + 1) * Wf + x] = 1;
      }
    }
  std::vector<uint8_t> out(raw.size(), 0);
  for (int y = 0; y < Hf; ++y)
    for (int x = 0; x < Wf; ++x) {
      uint8_t v = 0;
      for (int dy = -1; dy <= 1 && !v; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= Hf) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          const int xx = x + dx;
          if (xx >= 0 && xx < Wf && raw[static_cast<std::size_t>(yy) * Wf + xx]) {
            v = 1;
            break;
          }
        }
      }
      out[static_cast<std::size_t>(y) * Wf + x] = v;
    }
  return out;
}

// embree_renderer.cpp:436-458 (lights), 1320-1323 (AO / environment up axis)
template <class T>
static void initLights(Ctx<T>& c, const ubg_render_params* p, V3<T> trueUp) {
  for (int i = 0; i < p->n_lights; ++i) {
    const ubg_light& dl = p->lights[i];
    Light<T> l;
    l.L = normalize(V3<T>{-T(dl.direction[0]), -T(dl.direction[1]), -T(dl.direction[2])});
    l.color = V3<T>{T(dl.color[0]) * T(dl.intensity), T(dl.color[1]) * T(dl.intensity),
                    T(dl.color[2]) * T(dl.intensity)};
    l.highlight = dl.casts_highlight != 0;
    l.radius = (p->gi_integrator == 2 && dl.angular_radius > 0.0f) ? T(dl.angular_radius)
                                                                   : radiansT(T(p->light_radius_deg));
    c.lights.push_back(l);
  }
  c.ambLight = v3<T>(p->ambient_color);
  c.aoUp = p->ao_use_camera_up ? trueUp : safeNormalize(v3<T>(p->ao_up), trueUp);
  c.bg = v3<T>(p->background);
}

// denoise_atrous.cpp (denoiseAtrous) on a 3-channel buffer without albedo
// demodulation, as denoisePt1E runs it: edge-avoiding a-trous (B3 spline,
// 2^i holes) with normal^sigmaN, exp(-|dp| / (sigmaZ grad step)) and
// exp(-|dl| / (sigmaL sqrt(var))) edge stops; pixels whose guide normal is
// shorter than 0.5 neither change nor contribute.
static void denoiseAtrous(int W, int H, std::vector<float>& E, const std::vector<float>& nrm,
                          const std::vector<float>& pos, int itersIn, float sigmaZIn, float sigmaN, float sigmaLIn) {
  const std::size_t N = static_cast<std::size_t>(W) * H;
  const int iters = std::max(0, itersIn);
  if (W <= 0 || H <= 0 || iters == 0) return;
  const float kEps = 1.0e-3f;
  const float sigmaZ = std::max(1.0e-4f, sigmaZIn);
  const float sigmaL = std::max(1.0e-4f, sigmaLIn);
  auto lumOf = [](const float* v) { return 0.2126f * v[0] + 0.7152f * v[1] + 0.0722f * v[2]; };
  auto dist = [&](std::size_t a, std::size_t b) {
    const float dx = pos[3 * a] - pos[3 * b], dy = pos[3 * a + 1] - pos[3 * b + 1], dz = pos[3 * a + 2] - pos[3 * b + 2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
  };
  std::vector<uint8_t> valid(N);
  for (std::size_t p = 0; p < N; ++p) {
    const float nx = nrm[3 * p], ny = nrm[3 * p + 1], nz = nrm[3 * p + 2];
    valid[p] = (nx * nx + ny * ny + nz * nz) > 0.25f;
  }
  std::vector<float> irr = E;
  std::vector<float> grad(N, kEps), var(N, 0.0f);
  for (int y = 0; y < H; ++y)
    for (int x = 0; x < W; ++x) {
      const std::size_t p = static_cast<std::size_t>(y) * W + x;
      if (!valid[p]) continue;
      float g = kEps;
      const int nb[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
      for (auto& d : nb) {
        const int qx = x + d[0], qy = y + d[1];
        if (qx < 0 || qx >= W || qy < 0 || qy >= H) continue;
        const std::size_t q = static_cast<std::size_t>(qy) * W + qx;
        if (valid[q]) g = std::max(g, dist(p, q));
      }
      grad[p] = g;
      float s1 = 0.0f, s2 = 0.0f;
      int n = 0;
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int qx = x + dx, qy = y + dy;
          if (qx < 0 || qx >= W || qy < 0 || qy >= H) continue;
          const std::size_t q = static_cast<std::size_t>(qy) * W + qx;
          if (!valid[q]) continue;
          const float l = lumOf(&irr[3 * q]);
          s1 += l;
          s2 += l * l;
          ++n;
        }
      if (n > 0) {
        const float mean = s1 / n;
        var[p] = std::max(0.0f, s2 / n - mean * mean);
      }
    }
  const float h5[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
  std::vector<float> next(N * 3);
  for (int it = 0; it < iters; ++it) {
    const int step = 1 << it;
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) {
        const std::size_t p = static_cast<std::size_t>(y) * W + x;
        if (!valid[p]) {
          for (int k = 0; k < 3; ++k) next[3 * p + k] = irr[3 * p + k];
          continue;
        }
        const float lp = lumOf(&irr[3 * p]);
        const float lDenom = sigmaL * std::sqrt(var[p]) + kEps;
        const float zDenom = sigmaZ * grad[p] * static_cast<float>(step) + kEps;
        float sum[3] = {0.0f, 0.0f, 0.0f}, wsum = 0.0f;
        for (int dy = -2; dy <= 2; ++dy) {
          const int qy = y + dy * step;
          if (qy < 0 || qy >= H) continue;
          for (int dx = -2; dx <= 2; ++dx) {
            const int qx = x + dx * step;
            if (qx < 0 || qx >= W) continue;
            const std::size_t q = static_cast<std::size_t>(qy) * W + qx;
            if (!valid[q]) continue;
            float w = h5[dx + 2] * h5[dy + 2];
            const float nd = std::max(0.0f, nrm[3 * p] * nrm[3 * q] + nrm[3 * p + 1] * nrm[3 * q + 1] +
                                                nrm[3 * p + 2] * nrm[3 * q + 2]);
            w *= std::pow(nd, sigmaN);
            w *= std::exp(-dist(p, q) / zDenom);
            w *= std::exp(-std::fabs(lp - lumOf(&irr[3 * q])) / lDenom);
            for (int k = 0; k < 3; ++k) sum[k] += w * irr[3 * q + k];
            wsum += w;
          }
        }
        for (int k = 0; k < 3; ++k) next[3 * p + k] = wsum > 0.0f ? sum[k] / wsum : irr[3 * p + k];
      }
    irr.swap(next);
  }
  for (std::size_t p = 0; p < N; ++p)
    if (valid[p])
      for (int k = 0; k < 3; ++k) E[3 * p + k] = std::max(0.0f, irr[3 * p + k]);
}

// GI integrators the restatement covers: pt1, and pt2 without ReSTIR,
// adaptive spp, emissive NEE sources (pt2_emissive.hpp:111-126) or
// reflective receivers (hit_shader.hpp:251-254, :417-420)
static bool giSupported(const ubg_scene_desc* d, const ubg_render_params* p) {
  if (!p->gi) return true;
  if (p->gi_integrator == 1) return true;
  if (p->gi_integrator != 2 || p->pt2_rounds > 0 || p->pt2_adaptive) return false;
  bool emit = false, refl = false;
  // hit_shader.hpp:250-256: a principled material owns a traced reflection
  // whenever its specular lobe exists, a POV one when reflection > 0
  auto owns = [](const ubg_material& m) {
    return m.model == 1 ? (m.pbr_metallic > 1e-4f || m.pbr_specular > 1e-4f) : m.reflection > 0.0f;
  };
  if (d->n_triangles > 0 && d->n_vertices >= 3)
    for (int32_t m = 0; m < d->n_mesh_materials; ++m) {
      emit = emit || d->mesh_materials[m].emission > 0.0f;
      refl = refl || owns(d->mesh_materials[m]);
    }
  for (int64_t i = 0; i < d->n_spheres; ++i) {
    if (d->sphere_from_edge && d->sphere_from_edge[i]) continue;
    const ubg_material& m = d->prim_materials[d->sphere_material ? d->sphere_material[i] : 0];
    emit = emit || (m.emission > 0.0f && d->sphere_center_radius[4 * i + 3] > 0.0f);
    refl = refl || owns(m);
  }
  for (int64_t i = 0; i < d->n_cylinders; ++i) {
    if (d->cyl_from_edge && d->cyl_from_edge[i]) continue;
    const ubg_material& m = d->prim_materials[d->cyl_material ? d->cyl_material[i] : 0];
    emit = emit || (!(d->cyl_open && d->cyl_open[i]) && m.emission > 0.0f && d->cyl_radius[i] > 0.0f);
    refl = refl || owns(m);
  }
  if (p->pt2_emissive && p->pt2_emissive_nee && emit) return false;
  if (p->pt2_reflect && refl) return false;
  return true;
}

template <class T>
static int renderT(const ubg_scene_desc* desc, const ubg_render_params* p, ubg_frame* fr,
                   const orc_options* opt) {
  const std::shared_ptr<const OScene> scene = sceneFor(desc, opt->brute_force != 0);
  const OScene& s = *scene;
  const int ss = std::max(1, p->supersample);
  const int W = p->width * ss, H = p->height * ss;
  Ctx<T> c;
  c.s = &s;
  c.p = p;
  c.wHi = W;
  c.hHi = H;
  c.seedW = W;
  const bool gi = p->gi != 0;
  c.gi = gi;
  // embree_renderer.cpp:375-386 (camera basis)
  const ubg_camera& cam = p->camera;
  const T aspect = T(static_cast<float>(W)) / T(static_cast<float>(H));
  const V3<T> dir = normalize(v3<T>(cam.direction));
  const V3<T> right = normalize(cross(dir, v3<T>(cam.up)));
  const V3<T> trueUp = normalize(cross(right, dir));
  const T halfH = T(cam.height) * T(0.5f);
  const T halfW = halfH * aspect;
  const T persHalfH = std::tan(radiansT(T(cam.fovy)) * T(0.5f));
  const T persHalfW = persHalfH * aspect;
  c.camDir = dir;
  initLights(c, p, trueUp);
  const V3<T> pos = v3<T>(cam.position);

  // embree_renderer.cpp:527-553 + pipeline.cpp:68: coarse-grid AO pre-pass
  // (aoResDiv -1 = the supersample factor) on a first-hit G-buffer
  // (pt1_gbuffer.hpp:86-200), gathered with the coarse lattice seeds
  // (ao_coarse.hpp:73-129)
  CoarseAo grid;
  Counters coarseCnt;
  {
    // (pipeline.cpp:72-80: coarse AO is off under GI)
    const int div = gi ? 0 : (p->ao_res_div < 0 ? ss : p->ao_res_div);
    const Geometry& g = s.g;
    bool real = !g.idx.empty();
    for (const auto& sd : g.sphSide) real = real || sd.fromEdge == 0;
    for (const auto& sd : g.cSide) real = real || sd.fromEdge == 0;
    for (const auto& sd : g.rSide) real = real || sd.fromEdge == 0;
    if (div > 1 && p->ao_samples > 0 && real) {
      const int cw = (W + div - 1) / div, ch = (H + div - 1) / div;
      const bool quality = aoEnhanced<T>(p);
      grid.w = cw;
      grid.h = ch;
      const std::size_t nc = static_cast<std::size_t>(cw) * ch;
      grid.openness.assign(nc, 1.0);
      if (quality) grid.bent.assign(3 * nc, 0.0);
      grid.normal.assign(3 * nc, 0.0);
      grid.depth.assign(nc, 0.0);
      grid.hit.assign(nc, 0);
      const V3<T> posC = v3<T>(cam.position);
      for (int cy = 0; cy < ch; ++cy) {
        const T v = T(1) - T(2) * (T(static_cast<float>(cy)) + T(0.5f)) / T(static_cast<float>(ch));
        for (int cx = 0; cx < cw; ++cx) {
          const T u = T(2) * (T(static_cast<float>(cx)) + T(0.5f)) / T(static_cast<float>(cw)) - T(1);
          V3<T> org, rd;
          if (cam.orthographic) {
            org = posC + right * (u * halfW) + trueUp * (v * halfH);
            rd = dir;
          } else {
            org = posC;
            rd = normalize(dir + right * (u * persHalfW) + trueUp * (v * persHalfH));
          }
          coarseCnt.primary++;
          const Hit<T> h = intersect(s, org, rd, T(0), std::numeric_limits<T>::infinity());
          if (h.kind < 0) continue;
          V3<T> N, Ng;
          if (h.kind == UBG_KIND_MESH) {
            const uint32_t* ix = &g.idx[3 * h.id];
            const T w0 = T(1) - h.u - h.v;
            V3<T> nb{T(0), T(0), T(0)};
            for (int k = 0; k < 3; ++k) {
              const T a0 = T(g.nrm[3 * ix[0] + k]), a1 = T(g.nrm[3 * ix[1] + k]), a2 = T(g.nrm[3 * ix[2] + k]);
              const T val = w0 * a0 + h.u * a1 + h.v * a2;
              if (k == 0) nb.x = val;
              else if (k == 1) nb.y = val;
              else nb.z = val;
            }
            N = safeNormalize(nb);
            Ng = safeNormalize(h.ng);
          } else {
            const PrimSide* sd = h.kind == UBG_KIND_SPHERE        ? &g.sphSide[h.id]
                                 : h.kind == UBG_KIND_CONE_CURVE ? &g.cSide[h.id]
                                                                 : &g.rSide[h.id];
            if (sd->fromEdge) continue;
            N = safeNormalize(h.ng);
            Ng = N;
          }
          if (dot(N, rd) > T(0)) N = neg(N);
          if (dot(Ng, rd) > T(0)) Ng = neg(Ng);
          const V3<T> P{org.x + rd.x * h.t, org.y + rd.y * h.t, org.z + rd.z * h.t};
          const std::size_t cp = static_cast<std::size_t>(cy) * cw + cx;
          grid.hit[cp] = 1;
          grid.normal[3 * cp] = double(N.x);
          grid.normal[3 * cp + 1] = double(N.y);
          grid.normal[3 * cp + 2] = double(N.z);
          grid.depth[cp] = double(h.t);
          const T secEps = selfIntersectEps(P, N, h.t);
          if (quality) {
            const AOQuality<T> q = computeAOQuality(c, P, Ng, N, secEps, p->ao_samples, static_cast<uint32_t>(cx),
                                                    static_cast<uint32_t>(cy), cw, coarseCnt);
            grid.openness[cp] = double(q.openness);
            grid.bent[3 * cp] = double(q.bent.x);
            grid.bent[3 * cp + 1] = double(q.bent.y);
            grid.bent[3 * cp + 2] = double(q.bent.z);
          } else {
            grid.openness[cp] = double(computeAO(c, P, Ng, N, secEps, p->ao_samples, T(p->ao_distance),
                                                 static_cast<uint32_t>(cx), static_cast<uint32_t>(cy), cw,
                                                 coarseCnt));
          }
        }
      }
      c.grid = &grid;
    }
  }

  const std::size_t npx = static_cast<std::size_t>(W) * H;
  std::vector<float> hiColor(npx * 4, 0.0f), hiDepth(npx, 0.0f), hiViewZ(npx, 0.0f),
      hiAo(npx, 1.0f);
  std::vector<uint32_t> hiId(npx, 0xFFFFFFFFu);
  int nth = opt->threads > 0 ? opt->threads : static_cast<int>(std::thread::hardware_concurrency());
  if (nth < 1) nth = 1;
  std::vector<Counters> cnts(nth);
  // rows [0, nRows) of some grid over the worker pool, dynamically scheduled
  auto parallelRows = [&](const std::vector<int>& rowList, const std::function<void(int, Counters&)>& fn) {
    std::atomic<std::size_t> nextRow{0};
    auto worker = [&](int tid) {
      for (;;) {
        const std::size_t ri = nextRow.fetch_add(1);
        if (ri >= rowList.size()) break;
        fn(rowList[ri], cnts[tid]);
      }
    };
    std::vector<std::thread> pool;
    for (int i = 1; i < nth; ++i) pool.emplace_back(worker, i);
    worker(0);
    for (auto& t : pool) t.join();
  };
  // embree_renderer.cpp:1400-1414 / :1444-1455 -- the primary ray through
  // (lx, ly) of a latW x latH lattice of the image plane, integrated
  auto shadeAt = [&](const Ctx<T>& cx, int lx, int ly, int latW, int latH, Counters& cnt) {
    const T v = T(1) - T(2) * (T(static_cast<float>(ly)) + T(0.5f)) / T(static_cast<float>(latH));
    const T u = T(2) * (T(static_cast<float>(lx)) + T(0.5f)) / T(static_cast<float>(latW)) - T(1);
    V3<T> org, rd;
    if (cam.orthographic) {
      org = pos + right * (u * halfW) + trueUp * (v * halfH);
      rd = dir;
    } else {
      org = pos;
      rd = normalize(dir + right * (u * persHalfW) + trueUp * (v * persHalfH));
    }
    return integratePixel(cx, org, rd, static_cast<uint32_t>(lx), static_cast<uint32_t>(ly), cnt);
  };
  // GI first-hit G-buffer (embree_renderer.cpp:184-194)
  std::vector<float> giPos(gi ? npx * 3 : 0), giNrm(gi ? npx * 3 : 0), giRefl(gi ? npx * 3 : 0);
  std::vector<uint8_t> giElig(gi ? npx : 0);
  auto store = [&](std::size_t pix, const PixelOut<T>& po) {
    if (gi) {
      giPos[3 * pix] = static_cast<float>(po.worldPos.x);
      giPos[3 * pix + 1] = static_cast<float>(po.worldPos.y);
      giPos[3 * pix + 2] = static_cast<float>(po.worldPos.z);
      giNrm[3 * pix] = static_cast<float>(po.giGuideN.x);
      giNrm[3 * pix + 1] = static_cast<float>(po.giGuideN.y);
      giNrm[3 * pix + 2] = static_cast<float>(po.giGuideN.z);
      giRefl[3 * pix] = static_cast<float>(po.giRefl.x);
      giRefl[3 * pix + 1] = static_cast<float>(po.giRefl.y);
      giRefl[3 * pix + 2] = static_cast<float>(po.giRefl.z);
      giElig[pix] = po.giElig ? 1 : 0;
    }
    hiColor[4 * pix + 0] = static_cast<float>(po.r);
    hiColor[4 * pix + 1] = static_cast<float>(po.g);
    hiColor[4 * pix + 2] = static_cast<float>(po.b);
    hiColor[4 * pix + 3] = static_cast<float>(po.a);
    hiDepth[pix] = static_cast<float>(po.depth);
    hiViewZ[pix] = static_cast<float>(po.viewZ);
    hiAo[pix] = static_cast<float>(po.aoOpen);
    hiId[pix] = po.hitId;
  };
  // (pipeline.cpp:60-65: adaptive AA falls back to the grid under GI)
  const bool adaptive = p->aa_mode == 1 && (ss > 1 || p->aa_depth > 1) && !gi;
  const auto t0 = std::chrono::steady_clock::now();
  int64_t aaFlagged = 0;
  if (!adaptive) {
    const int rowB = std::max(0, opt->row_begin);
    const int rowE = opt->row_end > 0 ? std::min(H, opt->row_end) : H;
    const int rowStep = opt->row_step > 0 ? opt->row_step : 1;
    std::vector<int> rows;
    for (int r = rowB; r < rowE; r += rowStep) rows.push_back(r);
    parallelRows(rows, [&](int py, Counters& cnt) {
      for (int px = 0; px < W; ++px) store(static_cast<std::size_t>(py) * W + px, shadeAt(c, px, py, W, H, cnt));
    });
  } else {
    // embree_renderer.cpp:1428-1674 -- adaptive AA, three deterministic phases
    const int Wf = p->width, Hf = p->height, cSub = ss / 2;
    const int kFine = std::max(1, (p->aa_depth + ss - 1) / ss);
    std::vector<int> outRows(Hf);
    for (int y = 0; y < Hf; ++y) outRows[y] = y;
    // phase 1: the centre subpixel of every output pixel, plain sample counts
    std::vector<PixelOut<T>> centers(static_cast<std::size_t>(Wf) * Hf);
    parallelRows(outRows, [&](int Y, Counters& cnt) {
      for (int X = 0; X < Wf; ++X)
        centers[static_cast<std::size_t>(Y) * Wf + X] = shadeAt(c, X * ss + cSub, Y * ss + cSub, W, H, cnt);
    });
    // phase 2: the refinement mask (adaptive_aa.hpp), a function of the centres
    std::vector<AaCenter> ac(centers.size());
    for (std::size_t i = 0; i < centers.size(); ++i) ac[i] = aaCenterOf(centers[i]);
    const float pws = cam.orthographic ? (cam.height / static_cast<float>(Hf))
                                       : (2.0f * static_cast<float>(persHalfH) / static_cast<float>(Hf));
    const std::vector<uint8_t> mask = buildAaMask(Wf, Hf, ac.data(), pws, p->aa_threshold, !cam.orthographic);
    for (uint8_t m : mask) aaFlagged += m ? 1 : 0;
    if (fr->aa_mask)
      for (std::size_t i = 0; i < mask.size(); ++i) fr->aa_mask[i] = mask[i] ? 1.0f : 0.0f;
    // the fine lattice (aaDepth > ss) and the replicated-centre boost
    Ctx<T> cFine = c;
    cFine.seedW = W * kFine;
    const bool coarseOn = c.grid != nullptr;
    const bool softShadows = p->shadow_samples > 1 && p->light_radius_deg > 0.0f;
    const bool needBoost = ss > 1 && ((p->ao_samples > 0 && !coarseOn) || softShadows);
    Ctx<T> cBoost = c;
    cBoost.aoMul = coarseOn ? 1 : ss * ss;
    cBoost.shMul = softShadows ? ss * ss : 1;
    auto shadeCell = [&](int px, int py, Counters& cnt) {
      if (kFine == 1) return shadeAt(c, px, py, W, H, cnt);
      PixelOut<T> acc{};
      for (int bb = 0; bb < kFine; ++bb)
        for (int a = 0; a < kFine; ++a) {
          const PixelOut<T> pr = shadeAt(cFine, px * kFine + a, py * kFine + bb, W * kFine, H * kFine, cnt);
          if (a == 0 && bb == 0) {
            acc = pr;
          } else {
            acc.r += pr.r;
            acc.g += pr.g;
            acc.b += pr.b;
            acc.a += pr.a;
            acc.depth += pr.depth;
            acc.viewZ += pr.viewZ;
          }
        }
      const T inv = T(1.0f / static_cast<float>(kFine * kFine));
      acc.r *= inv;
      acc.g *= inv;
      acc.b *= inv;
      acc.a *= inv;
      acc.depth *= inv;
      acc.viewZ *= inv;
      return acc;
    };
    // phase 3: flagged pixels shade every subpixel, the rest replicate
    parallelRows(outRows, [&](int Y, Counters& cnt) {
      for (int X = 0; X < Wf; ++X) {
        const std::size_t XY = static_cast<std::size_t>(Y) * Wf + X;
        if (!mask[XY]) {
          const PixelOut<T> rep = needBoost ? shadeAt(cBoost, X * ss + cSub, Y * ss + cSub, W, H, cnt) : centers[XY];
          for (int j = 0; j < ss; ++j)
            for (int i = 0; i < ss; ++i) store(static_cast<std::size_t>(Y * ss + j) * W + X * ss + i, rep);
          continue;
        }
        for (int j = 0; j < ss; ++j)
          for (int i = 0; i < ss; ++i) {
            const int px = X * ss + i, py = Y * ss + j;
            const PixelOut<T> pr = (kFine == 1 && i == cSub && j == cSub) ? centers[XY] : shadeCell(px, py, cnt);
            store(static_cast<std::size_t>(py) * W + px, pr);
          }
      }
    });
  }
  // embree_renderer.cpp:1708-1716 -> runPt1GiPass full-res branch
  // (:959-1049): gather E at every receiver, then color += giIntensity *
  // reflectance * E, before the fog
  std::vector<float> giInd(gi ? npx * 3 : 0, 0.0f), giOcc(gi ? npx : 0, 0.0f);
  uint64_t gbufferRays = 0;
  if (gi) {
    GiParams<T> gp;
    gp.sky = v3<T>(p->pt1_sky_radiance);
    gp.ground = p->pt1_sky_mode == 1 ? v3<T>(p->ao_ground_color) : gp.sky;
    gp.up = c.aoUp;
    gp.envIntensity = p->gi_env_intensity;
    gp.bounces = std::max(1, p->gi_bounces);
    gp.maxDist = p->gi_max_distance > 0.0f ? T(p->gi_max_distance) : std::numeric_limits<T>::infinity();
    gp.pt2 = p->gi_integrator == 2;
    gp.pattern = p->pt2_pattern;
    gp.emissive = p->pt2_emissive != 0;
    const int spp = std::max(1, p->pt1_spp);
    const uint32_t seedMix = hashU32(p->pt1_seed);
    // gatherPt1Grid (pt1_gather.hpp:525-612) over a w x h grid: every
    // pixel with elig[pix] gathers at its position / normal (horizon test
    // against geo, or the normal itself), seeded by its index in the grid
    auto gatherGrid = [&](int w, int h, const std::vector<float>& pos, const std::vector<float>& nrm,
                          const std::vector<float>* geo, const std::vector<uint8_t>& elig, const std::vector<float>& dep,
                          int nspp, std::vector<float>& Eo, std::vector<float>& occo) {
      std::vector<int> rws(h);
      for (int y = 0; y < h; ++y) rws[y] = y;
      parallelRows(rws, [&](int py, Counters& cnt) {
        for (int px = 0; px < w; ++px) {
          const std::size_t pix = static_cast<std::size_t>(py) * w + px;
          if (!elig[pix]) continue;
          const V3<T> P = v3<T>(&pos[3 * pix]);
          const V3<T> N = v3<T>(&nrm[3 * pix]);
          const V3<T> Ng = geo ? v3<T>(&(*geo)[3 * pix]) : N;
          const uint32_t seed = hashU32(static_cast<uint32_t>(pix) ^ seedMix);
          uint32_t sBase = 0, sSeed = 0;
          if (gp.pt2)
            pt2PixelSampler(gp.pattern, static_cast<uint32_t>(px), static_cast<uint32_t>(py), seedMix, seed,
                            nextPow2(static_cast<uint32_t>(nspp)), sBase, sSeed);
          T occ = T(0);
          const V3<T> E = pt1GatherPoint(c, gp, P, N, Ng, nspp, seed, T(dep[pix]), p->pt1_ld != 0, p->pt1_clamp,
                                         &occ, cnt, sBase, sSeed);
          Eo[3 * pix] = static_cast<float>(E.x);
          Eo[3 * pix + 1] = static_cast<float>(E.y);
          Eo[3 * pix + 2] = static_cast<float>(E.z);
          occo[pix] = static_cast<float>(occ);
        }
      });
    };
    // pt1_denoise.cpp:20-81: NaN scrub, then the a-trous filter on E with the
    // normal / position guides (no albedo demodulation)
    auto denoiseE = [&](int w, int h, std::vector<float>& E, const std::vector<float>& nrm,
                        const std::vector<float>& pos) {
      if (!p->pt1_denoise) return;
      for (float& e : E)
        if (!std::isfinite(e)) e = 0.0f;
      denoiseAtrous(w, h, E, nrm, pos, p->denoise_iters, p->denoise_sigma_z, p->denoise_sigma_n,
                    p->denoise_sigma_l);
    };
    int div = p->pt1_gather_div < 0 ? ss : p->pt1_gather_div;
    if (div == 0) div = p->pt1_half_res ? 2 : 1;
    if (div < 1) div = 1;
    std::vector<float> Ebuf(npx * 3, 0.0f);
    if (div <= 1) {
      gatherGrid(W, H, giPos, giNrm, nullptr, giElig, hiDepth, spp, Ebuf, giOcc);
      if (fr->gi_features) {
        // the learned denoiser's features of the gathered E: log(1 + E), the
        // receiver's albedo, the normal AOV, distance / its frame maximum
        float dmax = 0.0f;
        for (std::size_t q = 0; q < npx; ++q) dmax = std::max(dmax, hiDepth[q]);
        for (std::size_t q = 0; q < npx; ++q) {
          float* f = fr->gi_features + 10 * q;
          for (int k = 0; k < 3; ++k) {
            f[k] = std::log1p(std::max(Ebuf[3 * q + k], 0.0f));
            f[3 + k] = giElig[q] ? giRefl[3 * q + k] : 0.0f;
            f[6 + k] = giNrm[3 * q + k];
          }
          f[9] = dmax > 0.0f ? hiDepth[q] / dmax : 0.0f;
        }
      }
      denoiseE(W, H, Ebuf, giNrm, giPos);
    } else {
      // embree_renderer.cpp:719-958
      const int hw = (W + div - 1) / div, hh = (H + div - 1) / div;
      const std::size_t ln = static_cast<std::size_t>(hw) * hh;
      std::vector<float> lPos(3 * ln, 0.0f), lNrm(3 * ln, 0.0f), lGeo(3 * ln, 0.0f), lDep(ln, 0.0f);
      std::vector<uint8_t> lHit(ln, 0);
      gbufferRays += ln;
      // tracePt1GBuffer (pt1_gbuffer.hpp:86-200)
      for (int py = 0; py < hh; ++py) {
        const T v = T(1) - T(2) * (T(static_cast<float>(py)) + T(0.5f)) / T(static_cast<float>(hh));
        for (int px = 0; px < hw; ++px) {
          const T u = T(2) * (T(static_cast<float>(px)) + T(0.5f)) / T(static_cast<float>(hw)) - T(1);
          V3<T> org, rd;
          if (cam.orthographic) {
            org = pos + right * (u * halfW) + trueUp * (v * halfH);
            rd = dir;
          } else {
            org = pos;
            rd = normalize(dir + right * (u * persHalfW) + trueUp * (v * persHalfH));
          }
          const Hit<T> h = intersect(s, org, rd, T(0), std::numeric_limits<T>::infinity());
          if (h.kind < 0) continue;
          const Geometry& g = s.g;
          V3<T> N, Ng;
          if (h.kind == UBG_KIND_MESH) {
            const uint32_t* ix = &g.idx[3 * h.id];
            const T w0 = T(1) - h.u - h.v;
            T nb[3];
            for (int k = 0; k < 3; ++k)
              nb[k] = w0 * T(g.nrm[3 * ix[0] + k]) + h.u * T(g.nrm[3 * ix[1] + k]) + h.v * T(g.nrm[3 * ix[2] + k]);
            N = safeNormalize(V3<T>{nb[0], nb[1], nb[2]});
            Ng = safeNormalize(h.ng);
          } else {
            const PrimSide* sd = h.kind == UBG_KIND_SPHERE        ? &g.sphSide[h.id]
                                 : h.kind == UBG_KIND_CONE_CURVE ? &g.cSide[h.id]
                                                                 : &g.rSide[h.id];
            if (sd->fromEdge) continue;
            N = safeNormalize(h.ng);
            Ng = N;
          }
          if (dot(N, rd) > T(0)) N = neg(N);
          if (dot(Ng, rd) > T(0)) Ng = neg(Ng);
          const std::size_t lp = static_cast<std::size_t>(py) * hw + px;
          const V3<T> P{org.x + rd.x * h.t, org.y + rd.y * h.t, org.z + rd.z * h.t};
          lPos[3 * lp] = static_cast<float>(P.x);
          lPos[3 * lp + 1] = static_cast<float>(P.y);
          lPos[3 * lp + 2] = static_cast<float>(P.z);
          lNrm[3 * lp] = static_cast<float>(N.x);
          lNrm[3 * lp + 1] = static_cast<float>(N.y);
          lNrm[3 * lp + 2] = static_cast<float>(N.z);
          lGeo[3 * lp] = static_cast<float>(Ng.x);
          lGeo[3 * lp + 1] = static_cast<float>(Ng.y);
          lGeo[3 * lp + 2] = static_cast<float>(Ng.z);
          lDep[lp] = static_cast<float>(h.t);
          lHit[lp] = 1;
        }
      }
      std::vector<float> Eh(3 * ln, 0.0f), occh(ln, 0.0f);
      gatherGrid(hw, hh, lPos, lNrm, &lGeo, lHit, lDep, spp, Eh, occh);
      denoiseE(hw, hh, Eh, lNrm, lPos);
      // upsampleJointBilateral (pt1_upsample.hpp:35-176)
      const float nPow = p->pt1_upsample_normal_pow, dScale = p->pt1_upsample_depth_scale;
      const int fbR = div <= 2 ? 0 : div;
      std::vector<uint8_t> patch(npx, 0);
      for (int py = 0; py < H; ++py) {
        const float sy = (static_cast<float>(py) + 0.5f) * static_cast<float>(hh) / static_cast<float>(H) - 0.5f;
        const int by = static_cast<int>(std::floor(sy));
        const float fy = sy - static_cast<float>(by);
        for (int px = 0; px < W; ++px) {
          const std::size_t pix = static_cast<std::size_t>(py) * W + px;
          if (!giElig[pix]) continue;
          const float sx = (static_cast<float>(px) + 0.5f) * static_cast<float>(hw) / static_cast<float>(W) - 0.5f;
          const int bx = static_cast<int>(std::floor(sx));
          const float fx = sx - static_cast<float>(bx);
          const float* nf = &giNrm[3 * pix];
          const float zf = hiDepth[pix];
          float wSum = 0.0f, e[3] = {0.0f, 0.0f, 0.0f}, oSum = 0.0f, bestBilin = -1.0f;
          long best = -1;
          for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
              const int hx = std::min(std::max(bx + dx, 0), hw - 1), hy = std::min(std::max(by + dy, 0), hh - 1);
              const std::size_t hp = static_cast<std::size_t>(hy) * hw + hx;
              if (!lHit[hp]) continue;
              const float wb = (dx ? fx : 1.0f - fx) * (dy ? fy : 1.0f - fy);
              if (wb > bestBilin) {
                bestBilin = wb;
                best = static_cast<long>(hp);
              }
              const float nd = nf[0] * lNrm[3 * hp] + nf[1] * lNrm[3 * hp + 1] + nf[2] * lNrm[3 * hp + 2];
              const float wn = std::pow(std::fmax(0.0f, nd), nPow);
              const float wz = std::exp(-std::fabs(zf - lDep[hp]) / (dScale * zf + 1e-6f));
              const float wgt = wb * wn * wz;
              wSum += wgt;
              for (int k = 0; k < 3; ++k) e[k] += wgt * Eh[3 * hp + k];
              oSum += wgt * occh[hp];
            }
          if (wSum >= 1e-6f) {
            if (p->pt1_edge_patch && wSum < p->pt1_edge_patch_thresh) patch[pix] = 1;
            const float inv = 1.0f / wSum;
            for (int k = 0; k < 3; ++k) Ebuf[3 * pix + k] = e[k] * inv;
            giOcc[pix] = oSum * inv;
            continue;
          }
          if (p->pt1_edge_patch) patch[pix] = 1;
          if (best < 0 && fbR > 0) {
            float bestD2 = 1e30f;
            for (int wy = by - fbR + 1; wy <= by + fbR; ++wy) {
              if (wy < 0 || wy >= hh) continue;
              for (int wx = bx - fbR + 1; wx <= bx + fbR; ++wx) {
                if (wx < 0 || wx >= hw) continue;
                const std::size_t hp = static_cast<std::size_t>(wy) * hw + wx;
                if (!lHit[hp]) continue;
                const float ddx = static_cast<float>(wx) - sx, ddy = static_cast<float>(wy) - sy;
                const float d2 = ddx * ddx + ddy * ddy;
                if (d2 < bestD2) {
                  bestD2 = d2;
                  best = static_cast<long>(hp);
                }
              }
            }
          }
          if (best >= 0) {
            for (int k = 0; k < 3; ++k) Ebuf[3 * pix + k] = Eh[3 * best + k];
            giOcc[pix] = occh[best];
          }
        }
      }
      // the rim patch (embree_renderer.cpp:800-850) and smoothPatchedPixels
      // (pt1_upsample.hpp:188-245, radius 2)
      bool anyPatch = false;
      for (uint8_t m : patch) anyPatch = anyPatch || m;
      if (anyPatch) {
        std::vector<float> Ep(npx * 3, 0.0f), occp(npx, 0.0f);
        gatherGrid(W, H, giPos, giNrm, nullptr, patch, hiDepth, spp * std::max(1, p->pt1_edge_patch_spp_mul), Ep,
                   occp);
        for (std::size_t i = 0; i < npx; ++i) {
          if (!patch[i]) continue;
          for (int k = 0; k < 3; ++k) Ebuf[3 * i + k] = Ep[3 * i + k];
          giOcc[i] = occp[i];
        }
        const std::vector<float> Esrc = Ebuf, oSrc = giOcc;
        for (int py = 0; py < H; ++py)
          for (int px = 0; px < W; ++px) {
            const std::size_t pix = static_cast<std::size_t>(py) * W + px;
            if (!patch[pix]) continue;
            const float* nf = &giNrm[3 * pix];
            const float zf = hiDepth[pix];
            float wSum = 1.0f, e[3] = {Esrc[3 * pix], Esrc[3 * pix + 1], Esrc[3 * pix + 2]}, oSum = oSrc[pix];
            for (int dy = -2; dy <= 2; ++dy) {
              const int ny = py + dy;
              if (ny < 0 || ny >= H) continue;
              for (int dx = -2; dx <= 2; ++dx) {
                if (dx == 0 && dy == 0) continue;
                const int nx = px + dx;
                if (nx < 0 || nx >= W) continue;
                const std::size_t q = static_cast<std::size_t>(ny) * W + nx;
                if (!giElig[q]) continue;
                const float nd = nf[0] * giNrm[3 * q] + nf[1] * giNrm[3 * q + 1] + nf[2] * giNrm[3 * q + 2];
                const float wn = std::pow(std::fmax(0.0f, nd), nPow);
                const float wz = std::exp(-std::fabs(zf - hiDepth[q]) / (dScale * zf + 1e-6f));
                const float wgt = wn * wz;
                if (wgt < 1e-4f) continue;
                wSum += wgt;
                for (int k = 0; k < 3; ++k) e[k] += wgt * Esrc[3 * q + k];
                oSum += wgt * oSrc[q];
              }
            }
            const float inv = 1.0f / wSum;
            for (int k = 0; k < 3; ++k) Ebuf[3 * pix + k] = e[k] * inv;
            giOcc[pix] = oSum * inv;
          }
      }
    }
    const float gI = p->gi_intensity;
    for (std::size_t pix = 0; pix < npx; ++pix) {
      if (!giElig[pix]) continue;
      for (int k = 0; k < 3; ++k) {
        const float ind = gI * giRefl[3 * pix + k] * Ebuf[3 * pix + k];
        hiColor[4 * pix + k] += ind;
        giInd[3 * pix + k] = ind;
      }
    }
  }
  const auto t1 = std::chrono::steady_clock::now();
  // pipeline.cpp:126-129 -> fog.cpp:17-36 (hi-res, before the downsample)
  if (p->fog.enabled) {
    for (std::size_t i = 0; i < npx; ++i) {
      const float z = hiViewZ[i];
      if (z <= 0.0f) continue;
      const float f = fogFactor(p->fog, z);
      float* col = &hiColor[4 * i];
      if (p->transparent_background) {
        col[3] *= f;
      } else {
        col[0] = p->fog.color[0] * (1.0f - f) + col[0] * f;
        col[1] = p->fog.color[1] * (1.0f - f) + col[1] * f;
        col[2] = p->fog.color[2] * (1.0f - f) + col[2] * f;
      }
    }
  }
  const auto t2 = std::chrono::steady_clock::now();
  Counters tot = coarseCnt;
  for (const auto& k : cnts) {
    tot.primary += k.primary;
    tot.shaded += k.shaded;
    tot.aoTraced += k.aoTraced;
    tot.aoSampled += k.aoSampled;
    tot.shadow += k.shadow;
    tot.capped += k.capped;
    tot.giRays += k.giRays;
    tot.giNee += k.giNee;
  }
  fr->hi_w = W;
  fr->hi_h = H;
  fr->aa_flagged = aaFlagged;
  fr->stats.primary_rays = tot.primary;
  fr->stats.shaded_hits = tot.shaded;
  fr->stats.ao_rays_traced = tot.aoTraced;
  fr->stats.ao_rays_sampled = tot.aoSampled;
  fr->stats.shadow_rays = tot.shadow;
  fr->stats.capped_rays = tot.capped;
  fr->stats.gi_rays = tot.giRays;
  fr->stats.gi_nee_rays = tot.giNee;
  fr->stats.gi_gbuffer_rays = gbufferRays;
  if (gi && fr->gi_indirect) std::memcpy(fr->gi_indirect, giInd.data(), npx * 3 * sizeof(float));
  if (gi && fr->gi_occlusion) std::memcpy(fr->gi_occlusion, giOcc.data(), npx * sizeof(float));
  fr->render_seconds = std::chrono::duration<double>(t2 - t0).count();
  (void)t1;
  fr->cancelled = 0;
  if (fr->depth) std::memcpy(fr->depth, hiDepth.data(), npx * sizeof(float));
  if (fr->view_z) std::memcpy(fr->view_z, hiViewZ.data(), npx * sizeof(float));
  if (fr->hit_id) std::memcpy(fr->hit_id, hiId.data(), npx * sizeof(uint32_t));
  if (fr->ao_open) std::memcpy(fr->ao_open, hiAo.data(), npx * sizeof(float));
  if (p->flags & UBG_FLAG_HIRES_ONLY) {
    fr->out_w = W;
    fr->out_h = H;
    if (fr->color) std::memcpy(fr->color, hiColor.data(), npx * 4 * sizeof(float));
    fr->post_seconds = 0.0;
    return UBG_OK;
  }
  // pipeline.cpp:182-223 (box downsample, ss > 1) and :254 (assumed gamma)
  const int ow = p->width, oh = p->height;
  std::vector<float> out;
  if (ss > 1) {
    out.assign(static_cast<std::size_t>(ow) * oh * 4, 0.0f);
    boxDownsample(hiColor.data(), W, H, 4, ss, out.data());
  } else {
    out.swap(hiColor);
  }
  // image_ops.cpp:37-46
  const float gm = p->assumed_gamma;
  if (!(std::fabs(gm - 1.0f) <= 1.0e-4f)) {
    const std::size_t n = static_cast<std::size_t>(ow) * oh;
    for (std::size_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        const float vv = out[4 * i + k];
        out[4 * i + k] = (vv > 0.0f) ? std::pow(vv, gm) : 0.0f;
      }
  }
  const auto t3 = std::chrono::steady_clock::now();
  fr->post_seconds = std::chrono::duration<double>(t3 - t2).count();
  fr->out_w = ow;
  fr->out_h = oh;
  if (fr->color) std::memcpy(fr->color, out.data(), out.size() * sizeof(float));
  return UBG_OK;
}

template <class T>
static int intersectT(const ubg_scene_desc* d, int64_t n, const ubg_ray* rays, ubg_hit* hits,
                      const orc_options* opt) {
  OScene s;
  buildScene(d, opt->brute_force != 0, s);
  for (int64_t i = 0; i < n; ++i) {
    const ubg_ray& r = rays[i];
    const Hit<T> h = intersect(s, v3<T>(r.org), v3<T>(r.dir), T(r.tnear), T(r.tfar));
    ubg_hit& o = hits[i];
    if (h.kind < 0) {
      o.t = std::numeric_limits<float>::infinity();
      o.u = o.v = 0.0f;
      o.kind = 0xFFFFFFFFu;
      o.prim_id = 0xFFFFFFFFu;
      o.ng[0] = o.ng[1] = o.ng[2] = 0.0f;
    } else {
      o.t = static_cast<float>(h.t);
      o.u = static_cast<float>(h.u);
      o.v = static_cast<float>(h.v);
      o.kind = static_cast<uint32_t>(h.kind);
      o.prim_id = h.id;
      o.ng[0] = static_cast<float>(h.ng.x);
      o.ng[1] = static_cast<float>(h.ng.y);
      o.ng[2] = static_cast<float>(h.ng.z);
    }
  }
  return UBG_OK;
}

template <class T>
static int occludedT(const ubg_scene_desc* d, int64_t n, const ubg_ray* rays, uint8_t* out,
                     const orc_options* opt) {
  OScene s;
  buildScene(d, opt->brute_force != 0, s);
  for (int64_t i = 0; i < n; ++i) {
    const ubg_ray& r = rays[i];
    out[i] = occluded(s, v3<T>(r.org), v3<T>(r.dir), T(r.tnear), T(r.tfar)) ? 1 : 0;
  }
  return UBG_OK;
}

// embree_renderer.cpp:294-357 with the argument filter of :71-131, FP32,
// every primitive tested (no BVH); a rejected root moves the search just
// past it along the same primitive.
static int occludedSegments(const ubg_scene_desc* d, int64_t n, const ubg_segment* segs, const int32_t* excl,
                            int64_t nExcl, float eps, float graze, float coplanar, uint8_t* out) {
  OScene s;
  buildScene(d, true, s);
  const Geometry& g = s.g;
  const bool hasMesh = !g.idx.empty();
  for (int64_t i = 0; i < n; ++i) {
    const ubg_segment& sg = segs[i];
    out[i] = 0;
    if (sg.exclude_first < 0 || sg.exclude_count < 0 ||
        static_cast<int64_t>(sg.exclude_first) + sg.exclude_count > nExcl)
      return UBG_ERR_INVALID_ARGUMENT;
    const V3<float> p = v3<float>(sg.p), q = v3<float>(sg.q);
    const V3<float> dd{q.x - p.x, q.y - p.y, q.z - p.z};
    const float len = length(dd);
    if (!(len > 0.0f)) continue;
    const V3<float> dir{dd.x / len, dd.y / len, dd.z / len};
    const float tnear = std::max(eps * len, eps);
    const float tfar = len * (1.0f - eps);
    if (tnear >= tfar) continue;
    const int nEx = hasMesh ? sg.exclude_count : 0;
    const int32_t* ex = nEx > 0 ? excl + sg.exclude_first : nullptr;
    const bool plain = nEx == 0 && graze <= 0.0f && coplanar <= 0.0f;
    auto accept = [&](const Hit<float>& h, int kind, uint32_t id) {
      if (nEx > 0 && kind == UBG_KIND_MESH) {
        const int64_t base = g.baseT > 0 ? static_cast<int64_t>(id) % g.baseT : static_cast<int64_t>(id);
        for (int k = 0; k < nEx; ++k)
          if (static_cast<int64_t>(ex[k]) == base) return false;
      }
      if (graze > 0.0f || coplanar > 0.0f) {
        const float nx = h.ng.x, ny = h.ng.y, nz = h.ng.z;
        const float nlen = std::sqrt(nx * nx + ny * ny + nz * nz);
        const float dlen = std::sqrt(dir.x * dir.x + dir.y * dir.y + dir.z * dir.z);
        if (nlen > 0.0f && dlen > 0.0f) {
          const float dDotN = nx * dir.x + ny * dir.y + nz * dir.z;
          const float c = std::fabs(dDotN) / (nlen * dlen);
          if (graze > 0.0f && c <= graze) return false;
          if (coplanar > 0.0f && h.t * std::fabs(dDotN) / nlen <= coplanar) return false;
        }
      }
      return true;
    };
    Hit<float> h;
    for (const PRef& pr : s.all) {
      float lo = tnear;
      bool hit = false;
      for (int k = 0; k < 4; ++k) {
        if (!isectPrim(g, pr, p, dir, lo, tfar, h)) break;
        if (plain || accept(h, pr.kind, pr.id)) {
          hit = true;
          break;
        }
        lo = std::nextafter(h.t, std::numeric_limits<float>::infinity());
      }
      if (hit) {
        out[i] = 1;
        break;
      }
    }
  }
  return UBG_OK;
}

}  // namespace orc

namespace orc {
// ---------------------------------------------------------------------------
// Object-space silhouette edges of the analytic primitives
// (edges/object_space_edges.cpp:172-229, edges/analytic_silhouette.cpp:31-223)
// with the union-boundary clip tested against every solid (the reference's
// own brute force). Mesh feature edges are not restated here.

struct EdgeAcc {
  std::vector<float> p0, p1, color, opacity1;
  std::vector<uint16_t> group;
  void add(V3<float> a, V3<float> b, const ubg_edge_options& o, uint16_t g) {
    p0.insert(p0.end(), {a.x, a.y, a.z});
    p1.insert(p1.end(), {b.x, b.y, b.z});
    color.insert(color.end(), {o.color[0], o.color[1], o.color[2], 1.0f});
    opacity1.push_back(-1.0f);  // both ends opaque
    group.push_back(g);
  }
};

struct SilLoop {
  std::vector<V3<float>> pts;
  bool ring = false;
  uint16_t group = 0;
  int sph = -1, cyl = -1;
};

static void silhouetteLoops(const ubg_scene_desc* d, const ubg_camera& c, int N, float raise,
                            std::vector<SilLoop>& loops) {
  using F = V3<float>;
  const F eye{c.position[0], c.position[1], c.position[2]};
  const F view{c.direction[0], c.direction[1], c.direction[2]};
  const bool ortho = c.orthographic != 0;
  const float pi = 3.14159265358979323846f;
  auto ringPoints = [&](F center, Frame<float> fr, float rad, int n, std::vector<F>& pts) {
    pts.clear();
    for (int i = 0; i < n; ++i) {
      const float th = 2.0f * pi * static_cast<float>(i) / static_cast<float>(n);
      pts.push_back(center + (fr.t * std::cos(th) + fr.b * std::sin(th)) * rad);
    }
  };
  for (int64_t i = 0; i < d->n_spheres; ++i) {
    if (d->sphere_from_edge && d->sphere_from_edge[i]) continue;
    const F O{d->sphere_center_radius[4 * i], d->sphere_center_radius[4 * i + 1], d->sphere_center_radius[4 * i + 2]};
    const float r = d->sphere_center_radius[4 * i + 3];
    if (r <= 0.0f) continue;
    F axis, Q;
    float rho;
    if (ortho) {
      axis = normalize(view);
      Q = O;
      rho = r;
    } else {
      const F dd = O - eye;
      const float L = length(dd);
      if (L <= r * 1.0001f) continue;
      axis = dd * (1.0f / L);
      Q = O - axis * (r * r / L);
      const float disc = L * L - r * r;
      rho = r * std::sqrt(disc > 0.0f ? disc : 0.0f) / L;
    }
    if (rho <= 0.0f) continue;
    SilLoop l;
    ringPoints(Q, frameFromNormal(axis), rho, N, l.pts);
    for (F& P : l.pts) P = P + normalize(P - O) * raise;
    l.ring = true;
    l.group = d->sphere_group ? d->sphere_group[i] : 0;
    l.sph = static_cast<int>(i);
    loops.push_back(l);
  }
  // the two grazing contacts on the cross-section circle at P
  auto contacts = [&](F P, F u, float r, F& dp, F& dm) {
    if (ortho) {
      F eT = cross(u, normalize(view) * -1.0f);
      const float l = length(eT);
      if (l < 1.0e-4f) return false;
      eT = eT * (1.0f / l);
      dp = eT;
      dm = eT * -1.0f;
      return true;
    }
    const F cp = eye - P;
    const F perp = cp - u * dot(u, cp);
    const float Lp = length(perp);
    if (Lp <= r * 1.0001f) return false;
    const F eR = perp * (1.0f / Lp);
    F eT = cross(u, eR);
    const float lt = length(eT);
    if (lt < 1.0e-4f) return false;
    eT = eT * (1.0f / lt);
    const float ca = r / Lp, sa = std::sqrt(std::fmax(0.0f, 1.0f - ca * ca));
    dp = eR * ca + eT * sa;
    dm = eR * ca - eT * sa;
    return true;
  };
  for (int64_t i = 0; i < d->n_cylinders; ++i) {
    if (d->cyl_from_edge && d->cyl_from_edge[i]) continue;
    const F A{d->cyl_p0[3 * i], d->cyl_p0[3 * i + 1], d->cyl_p0[3 * i + 2]};
    const F B{d->cyl_p1[3 * i], d->cyl_p1[3 * i + 1], d->cyl_p1[3 * i + 2]};
    const float r = d->cyl_radius[i];
    if (r <= 0.0f) continue;
    const float axLen = length(B - A);
    if (axLen <= 0.0f) continue;
    const F u = (B - A) * (1.0f / axLen);
    const uint16_t g = d->cyl_group ? d->cyl_group[i] : 0;
    F ap, am, bp, bm;
    if (contacts(A, u, r, ap, am) && contacts(B, u, r, bp, bm)) {
      const float off = r + raise;
      SilLoop l1, l2;
      l1.pts = {A + ap * off, B + bp * off};
      l2.pts = {A + am * off, B + bm * off};
      l1.group = l2.group = g;
      l1.cyl = l2.cyl = static_cast<int>(i);
      loops.push_back(l1);
      loops.push_back(l2);
    } else {
      // end-on: the rim of the nearer cap
      const F dir = normalize(view);
      const F cap = ortho ? (dot(A, dir) < dot(B, dir) ? A : B) : (length(eye - A) < length(eye - B) ? A : B);
      SilLoop l;
      ringPoints(cap, frameFromNormal(normalize(u)), r + raise, N, l.pts);
      l.ring = true;
      l.group = g;
      l.cyl = static_cast<int>(i);
      loops.push_back(l);
    }
  }
}

// strictly inside a solid other than the chord's own source
static bool insideOtherSolid(const ubg_scene_desc* d, V3<float> X, int skipSph, int skipCyl, float ringRs) {
  using F = V3<float>;
  for (int64_t i = 0; i < d->n_spheres; ++i) {
    if (static_cast<int>(i) == skipSph) continue;
    const float* s = &d->sphere_center_radius[4 * i];
    if (length(X - F{s[0], s[1], s[2]}) < s[3]) return true;
  }
  for (int64_t j = 0; j < d->n_cylinders; ++j) {
    if (static_cast<int>(j) == skipCyl) continue;
    const float r = d->cyl_radius[j];
    if (ringRs > 0.0f && r < ringRs - 1.0e-4f) continue;  // a thinner bond keeps the ring
    const F p0{d->cyl_p0[3 * j], d->cyl_p0[3 * j + 1], d->cyl_p0[3 * j + 2]};
    const F ab = F{d->cyl_p1[3 * j], d->cyl_p1[3 * j + 1], d->cyl_p1[3 * j + 2]} - p0;
    const float L = length(ab);
    if (L <= 0.0f) continue;
    const F u = ab * (1.0f / L);
    const float t = dot(X - p0, u);
    if (t < 0.0f || t > L) continue;
    if (length((X - p0) - u * t) < r) return true;
  }
  return false;
}

static void objectSpaceEdges(const ubg_scene_desc* d, const ubg_camera& cam, const ubg_edge_options& o, EdgeAcc& out) {
  std::vector<SilLoop> loops;
  silhouetteLoops(d, cam, o.segments < 3 ? 3 : o.segments, o.raise, loops);
  for (const SilLoop& l : loops) {
    const size_t n = l.pts.size();
    if (n < 2) continue;
    for (size_t i = 0; i < (l.ring ? n : n - 1); ++i) {
      const V3<float> a = l.pts[i], b = l.pts[(i + 1) % n];
      if (!o.clip) {
        out.add(a, b, o, l.group);
        continue;
      }
      const V3<float> dd = b - a;
      const float len = length(dd);
      if (len <= 0.0f) continue;
      const float step = o.width > 0.0f ? o.width : len;
      const int K = std::max(2, static_cast<int>(std::ceil(len / step)));
      auto at = [&](int k) { return a + dd * (static_cast<float>(k) / static_cast<float>(K)); };
      const float ringRs = l.sph >= 0 ? d->sphere_center_radius[4 * l.sph + 3] : -1.0f;
      int start = -1;
      for (int k = 0; k <= K; ++k) {
        const bool in = insideOtherSolid(d, at(k), l.sph, l.cyl, ringRs);
        if (!in && start < 0) start = k;
        if (in && start >= 0) {
          if (k - 1 > start) out.add(at(start), at(k - 1), o, l.group);
          start = -1;
        }
      }
      if (start >= 0 && K > start) out.add(at(start), at(K), o, l.group);
    }
  }
}
}  // namespace orc

extern "C" {

int orc_occluded_segments(const ubg_scene_desc* desc, int64_t n, const ubg_segment* segs, const int32_t* excl,
                          int64_t n_excl, float eps, float graze_cos_eps, float coplanar_eps, uint8_t* occ) {
  if (!desc || n < 0 || (n > 0 && (!segs || !occ)) || (n_excl > 0 && !excl)) return UBG_ERR_INVALID_ARGUMENT;
  return orc::occludedSegments(desc, n, segs, excl, n_excl, eps, graze_cos_eps, coplanar_eps, occ);
}

static const orc_options kDefaultOpt = {32, 0, 0, 0, 1, 0};

int orc_render(const ubg_scene_desc* desc, const ubg_render_params* params, ubg_frame* frame,
               const orc_options* opt) {
  if (!desc || !params || !frame) return UBG_ERR_INVALID_ARGUMENT;
  if (params->width <= 0 || params->height <= 0) return UBG_ERR_INVALID_ARGUMENT;
  const orc_options* o = opt ? opt : &kDefaultOpt;
  if (!orc::giSupported(desc, params)) return UBG_ERR_UNSUPPORTED;
  // the learned denoiser is checked through its network, not restated here
  if (params->gi && params->pt1_denoise == 2) return UBG_ERR_UNSUPPORTED;
  return o->precision == 64 ? orc::renderT<double>(desc, params, frame, o)
                            : orc::renderT<float>(desc, params, frame, o);
}

int orc_intersect(const ubg_scene_desc* desc, int64_t n, const ubg_ray* rays, ubg_hit* hits,
                  const orc_options* opt) {
  const orc_options* o = opt ? opt : &kDefaultOpt;
  return o->precision == 64 ? orc::intersectT<double>(desc, n, rays, hits, o)
                            : orc::intersectT<float>(desc, n, rays, hits, o);
}

int orc_occluded(const ubg_scene_desc* desc, int64_t n, const ubg_ray* rays, uint8_t* occ,
                 const orc_options* opt) {
  const orc_options* o = opt ? opt : &kDefaultOpt;
  return o->precision == 64 ? orc::occludedT<double>(desc, n, rays, occ, o)
                            : orc::occludedT<float>(desc, n, rays, occ, o);
}

int64_t orc_chain_curves(const ubg_scene_desc* desc, int64_t cap, int32_t* src, uint8_t* flags) {
  orc::Geometry g;
  orc::buildGeometry(desc, g);
  const int64_t n = static_cast<int64_t>(g.rseg.size());
  if (n > cap) return -1;
  for (int64_t i = 0; i < n; ++i) {
    src[i] = g.rsrc[i];
    flags[i] = g.rflag[i];
  }
  return n;
}

void orc_tea2(uint32_t* v0, uint32_t* v1) { orc::tea2(*v0, *v1); }

void orc_set_scene_cache(int32_t enable) {
  std::lock_guard<std::mutex> lk(orc::cache::mu);
  orc::cache::enabled = enable != 0;
  if (!orc::cache::enabled) orc::cache::scene.reset();
}

int orc_object_space_edges(const ubg_scene_desc* desc, const ubg_camera* cam, const ubg_edge_options* opt,
                           ubg_edges* out) {
  if (!desc || !cam || !opt || !out) return UBG_ERR_INVALID_ARGUMENT;
  orc::EdgeAcc acc;
  orc::objectSpaceEdges(desc, *cam, *opt, acc);
  const int64_t n = static_cast<int64_t>(acc.group.size());
  *out = ubg_edges{};
  out->n = n;
  out->radius = opt->width > 0.0f ? opt->width : 0.0f;
  if (n == 0) return UBG_OK;
  auto dup = [](const auto& v) {
    using T = typename std::decay_t<decltype(v)>::value_type;
    T* p = static_cast<T*>(std::malloc(v.size() * sizeof(T)));
    std::memcpy(p, v.data(), v.size() * sizeof(T));
    return p;
  };
  out->p0 = dup(acc.p0);
  out->p1 = dup(acc.p1);
  out->color = dup(acc.color);
  out->opacity1 = dup(acc.opacity1);
  out->group = dup(acc.group);
  return UBG_OK;
}

void orc_edges_free(ubg_edges* e) {
  if (!e) return;
  std::free(e->p0);
  std::free(e->p1);
  std::free(e->color);
  std::free(e->opacity1);
  std::free(e->group);
  *e = ubg_edges{};
}

void orc_pt1_denoise(int32_t w, int32_t h, float* E, const float* normal, const float* position, int32_t iters,
                     float sigma_z, float sigma_n, float sigma_l) {
  const std::size_t n = static_cast<std::size_t>(w) * h;
  std::vector<float> e(E, E + 3 * n), nrm(normal, normal + 3 * n), pos(position, position + 3 * n);
  for (float& v : e)
    if (!std::isfinite(v)) v = 0.0f;
  orc::denoiseAtrous(w, h, e, nrm, pos, iters, sigma_z, sigma_n, sigma_l);
  std::memcpy(E, e.data(), 3 * n * sizeof(float));
}

void orc_pt2_sobol_2d(uint32_t index, uint32_t seed, float* u1, float* u2) { orc::sobolBurley2D(index, seed, *u1, *u2); }

void orc_pt2_pixel_sampler(int32_t pattern, uint32_t px, uint32_t py, uint32_t frame_seed, uint32_t pixel_hash,
                           uint32_t spp, uint32_t* index_base, uint32_t* seed) {
  orc::pt2PixelSampler(pattern, px, py, frame_seed, pixel_hash, orc::nextPow2(spp < 1 ? 1 : spp), *index_base, *seed);
}

int orc_pt1_gather_point(const ubg_scene_desc* desc, const ubg_render_params* params, const float* P,
                         const float* N, const float* Ng, int32_t spp, uint32_t seed, float eps_t, float* E,
                         float* occlusion, uint64_t* gi_rays) {
  if (!desc || !params || !P || !N || !Ng || !E) return UBG_ERR_INVALID_ARGUMENT;
  orc::OScene s;
  orc::buildScene(desc, false, s);
  orc::Ctx<float> c;
  c.s = &s;
  c.p = params;
  const ubg_camera& cam = params->camera;
  const orc::V3<float> dir = orc::normalize(orc::v3<float>(cam.direction));
  const orc::V3<float> right = orc::normalize(orc::cross(dir, orc::v3<float>(cam.up)));
  orc::initLights(c, params, orc::normalize(orc::cross(right, dir)));
  orc::GiParams<float> gp;
  gp.sky = orc::v3<float>(params->pt1_sky_radiance);
  gp.ground = params->pt1_sky_mode == 1 ? orc::v3<float>(params->ao_ground_color) : gp.sky;
  gp.up = c.aoUp;
  gp.envIntensity = params->gi_env_intensity;
  gp.bounces = std::max(1, params->gi_bounces);
  gp.maxDist = params->gi_max_distance > 0.0f ? params->gi_max_distance : std::numeric_limits<float>::infinity();
  gp.pt2 = params->gi_integrator == 2;
  gp.pattern = params->pt2_pattern;
  gp.emissive = params->pt2_emissive != 0;
  // (a pixel at the origin of the gather grid: the per-pixel Sobol
  // arrangement is seeded by the point's seed, the blue-noise one by the
  // frame seed of pt1_seed)
  uint32_t sBase = 0, sSeed = 0;
  if (gp.pt2)
    orc::pt2PixelSampler(gp.pattern, 0u, 0u, orc::hashU32(params->pt1_seed), seed,
                         orc::nextPow2(static_cast<uint32_t>(std::max(1, spp))), sBase, sSeed);
  orc::Counters cnt;
  float occ = 0.0f;
  const orc::V3<float> e = orc::pt1GatherPoint(c, gp, orc::v3<float>(P), orc::v3<float>(N), orc::v3<float>(Ng),
                                               spp, seed, eps_t, params->pt1_ld != 0, params->pt1_clamp, &occ, cnt,
                                               sBase, sSeed);
  E[0] = e.x;
  E[1] = e.y;
  E[2] = e.z;
  if (occlusion) *occlusion = occ;
  if (gi_rays) *gi_rays = cnt.giRays;
  return UBG_OK;
}

void orc_aa_mask(int32_t wf, int32_t hf, const float* rgba, const float* view_z, const float* normal,
                 const uint32_t* geom, const int32_t* group, float pixel_world_size, float threshold,
                 int32_t perspective, uint8_t* mask) {
  std::vector<orc::AaCenter> c(static_cast<std::size_t>(wf) * hf);
  for (std::size_t i = 0; i < c.size(); ++i) {
    c[i].r = rgba[4 * i];
    c[i].g = rgba[4 * i + 1];
    c[i].b = rgba[4 * i + 2];
    c[i].a = rgba[4 * i + 3];
    c[i].viewZ = view_z[i];
    for (int k = 0; k < 3; ++k) c[i].n[k] = normal[3 * i + k];
    c[i].geom = geom[i];
    c[i].group = group[i];
  }
  const std::vector<uint8_t> m =
      orc::buildAaMask(wf, hf, c.data(), pixel_world_size, threshold, perspective != 0);
  std::memcpy(mask, m.data(), m.size());
}

float orc_fog_factor(const ubg_fog* fog, float z) { return orc::fogFactor(*fog, z); }

void orc_box_downsample(const float* src, int w, int h, int channels, int ss, float* dst) {
  orc::boxDownsample(src, w, h, channels, ss, dst);
}

// image_ops.cpp:48-58
float orc_srgb_encode(float v) {
  v = std::min(1.0f, std::max(0.0f, v));
  return (v <= 0.0031308f) ? 12.92f * v : 1.055f * std::pow(v, 1.0f / 2.4f) - 0.055f;
}
float orc_srgb_decode(float s) {
  s = std::min(1.0f, std::max(0.0f, s));
  return (s <= 0.04045f) ? s / 12.92f : std::pow((s + 0.055f) / 1.055f, 2.4f);
}

}  // extern "C"

// ===========================================================================
// Physically based path tracer of ubg_pathtrace (no reference counterpart:
// the reference's integrators are diffuse-only and docs/api/libumbreon.md:33
// puts refraction out of its scope). This restates the MI355X kernels'
// algorithm step for step -- the same counter-based random numbers and
// dimension layout, the same float-op order for T = float -- so the FP32
// oracle follows the device per sample, and T = double is the FP64 oracle of
// the same estimator. Its pins are closed forms: the white furnace of a
// smooth dielectric / diffuse / mirror sphere (radiance exactly the
// environment's), the irradiance of one spherical light on a plane, the
// BSDF sampling weights against quadrature of the evaluated BSDF, the
// (generalised) reciprocity of the evaluated BSDF, and ReSTIR's estimate
// against the all-lights next-event estimator (tests/test_pt_oracle.py).
namespace orc {
namespace pt {

template <class T>
struct Onb {
  V3<T> t, b, n;
};
template <class T>
inline T tpi() { return T(3.14159265358979323846); }

inline uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7FEB352Du;
  x ^= x >> 15;
  x *= 0x846CA68Bu;
  x ^= x >> 16;
  return x;
}
inline float rndf(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim) {
  const uint32_t h = hash32(seed ^ hash32(pixel ^ hash32((sample * 0x9E3779B9u) ^ hash32(dim + 0x632BE5ABu))));
  return static_cast<float>(h >> 8) * 0x1.0p-24f;
}
template <class T>
inline T rnd(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim) {
  return T(rndf(seed, pixel, sample, dim));
}

template <class T>
inline V3<T> mkv(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T>
inline V3<T> mul(V3<T> a, V3<T> b) { return V3<T>{a.x * b.x, a.y * b.y, a.z * b.z}; }

// Duff et al. 2017
template <class T>
inline Onb<T> onb(V3<T> n) {
  const T sgn = n.z >= T(0) ? T(1) : T(-1);
  const T a = T(-1) / (sgn + n.z);
  const T b = n.x * n.y * a;
  return Onb<T>{mkv<T>(T(1) + sgn * n.x * n.x * a, sgn * b, -sgn * n.x), mkv<T>(b, sgn + n.y * n.y * a, -n.y), n};
}
template <class T>
inline V3<T> toLocal(const Onb<T>& f, V3<T> v) { return mkv<T>(dot(v, f.t), dot(v, f.b), dot(v, f.n)); }
template <class T>
inline V3<T> toWorld(const Onb<T>& f, V3<T> v) {
  return mkv<T>(f.t.x * v.x + f.b.x * v.y + f.n.x * v.z, f.t.y * v.x + f.b.y * v.y + f.n.y * v.z,
                f.t.z * v.x + f.b.z * v.y + f.n.z * v.z);
}

// GGX in the local frame (n = +z): D, Smith Lambda / G1 / height-correlated G2
template <class T>
inline T ggxD(V3<T> m, T a) {
  const T c2 = m.z * m.z;
  if (c2 <= T(0)) return T(0);
  const T a2 = a * a;
  const T k = c2 * (a2 - T(1)) + T(1);
  return a2 / (tpi<T>() * k * k);
}
template <class T>
inline T ggxLambda(V3<T> w, T a) {
  const T c2 = w.z * w.z;
  if (c2 <= T(0)) return T(0);
  const T t2 = std::fmax(T(0), T(1) - c2) / c2;
  return T(0.5f) * (std::sqrt(T(1) + a * a * t2) - T(1));
}
template <class T>
inline T ggxG1(V3<T> w, T a) { return T(1) / (T(1) + ggxLambda(w, a)); }
template <class T>
inline T ggxG2(V3<T> wo, V3<T> wi, T a) { return T(1) / (T(1) + ggxLambda(wo, a) + ggxLambda(wi, a)); }
// visible-normal sampling (Heitz 2018)
template <class T>
inline V3<T> ggxSampleVisible(V3<T> wo, T a, T u1, T u2) {
  const V3<T> vh = normalize(mkv<T>(a * wo.x, a * wo.y, wo.z));
  const T l2 = vh.x * vh.x + vh.y * vh.y;
  const V3<T> T1 = l2 > T(0) ? mkv<T>(-vh.y, vh.x, T(0)) * (T(1) / std::sqrt(l2)) : mkv<T>(T(1), T(0), T(0));
  const V3<T> T2 = cross(vh, T1);
  const T r = std::sqrt(u1);
  const T phi = T(2) * tpi<T>() * u2;
  const T t1 = r * std::cos(phi);
  T t2 = r * std::sin(phi);
  const T s = T(0.5f) * (T(1) + vh.z);
  t2 = (T(1) - s) * std::sqrt(std::fmax(T(0), T(1) - t1 * t1)) + s * t2;
  const T tz = std::sqrt(std::fmax(T(0), T(1) - t1 * t1 - t2 * t2));
  const V3<T> nh = mkv<T>(T1.x * t1 + T2.x * t2 + vh.x * tz, T1.y * t1 + T2.y * t2 + vh.y * tz,
                          T1.z * t1 + T2.z * t2 + vh.z * tz);
  return normalize(mkv<T>(a * nh.x, a * nh.y, std::fmax(T(1.0e-6f), nh.z)));
}
// unpolarised dielectric Fresnel; eta = n_t / n_i
template <class T>
inline T fresnelDielectric(T cosI, T eta) {
  const T s2 = std::fmax(T(0), T(1) - cosI * cosI) / (eta * eta);
  if (s2 >= T(1)) return T(1);
  const T cosT = std::sqrt(T(1) - s2);
  const T rs = (cosI - eta * cosT) / (cosI + eta * cosT);
  const T rp = (eta * cosI - cosT) / (eta * cosI + cosT);
  return T(0.5f) * (rs * rs + rp * rp);
}
template <class T>
inline V3<T> schlick(V3<T> f0, T c) {
  const T m = std::fmin(T(1), std::fmax(T(0), T(1) - c));
  const T m5 = (m * m) * (m * m) * m;
  return mkv<T>(f0.x + (T(1) - f0.x) * m5, f0.y + (T(1) - f0.y) * m5, f0.z + (T(1) - f0.z) * m5);
}

template <class T>
struct Surf {
  Onb<T> f;
  V3<T> C;
  int type;
  T alpha, ior;
};
constexpr int kDiffuse = 0, kConductor = 1, kDielectric = 2;
constexpr float kSmoothAlpha = 4.0e-4f;

template <class T>
inline V3<T> bsdfCos(const Surf<T>& s, V3<T> wo, V3<T> wi) {
  const V3<T> zero{T(0), T(0), T(0)};
  if (s.type == kDiffuse) {
    if (wo.z <= T(0) || wi.z <= T(0)) return zero;
    return s.C * (wi.z / tpi<T>());
  }
  if (s.alpha < T(kSmoothAlpha)) return zero;
  if (s.type == kConductor) {
    if (wo.z <= T(0) || wi.z <= T(0)) return zero;
    const V3<T> m = normalize(wo + wi);
    const V3<T> F = schlick(s.C, dot(wi, m));
    const T k = ggxD(m, s.alpha) * ggxG2(wo, wi, s.alpha) / (T(4) * wo.z);
    return F * k;
  }
  const T cosO = wo.z, cosI = wi.z;
  if (cosO == T(0) || cosI == T(0)) return zero;
  const bool refl = cosO * cosI > T(0);
  const T eta = cosO > T(0) ? s.ior : T(1) / s.ior;
  V3<T> m = refl ? (wo + wi) : (wi * eta + wo);
  if (dot(m, m) == T(0)) return zero;
  m = normalize(m);
  if (m.z < T(0)) m = neg(m);
  if (dot(m, wi) * cosI < T(0) || dot(m, wo) * cosO < T(0)) return zero;
  const T F = fresnelDielectric(std::fabs(dot(wo, m)), cosO > T(0) ? s.ior : T(1) / s.ior);
  const T D = ggxD(m, s.alpha);
  const T G = ggxG2(mkv<T>(wo.x, wo.y, std::fabs(wo.z)), mkv<T>(wi.x, wi.y, std::fabs(wi.z)), s.alpha);
  if (refl) {
    const T k = D * G * F / (T(4) * std::fabs(cosO));
    return mkv<T>(k, k, k);
  }
  const T dn = dot(wi, m) + dot(wo, m) / eta;
  const T k = D * (T(1) - F) * G * std::fabs(dot(wi, m) * dot(wo, m) / (dn * dn * cosO)) / (eta * eta);
  return s.C * k;
}

template <class T>
inline bool bsdfSample(const Surf<T>& s, V3<T> wo, T u1, T u2, T u3, V3<T>& wi, V3<T>& weight) {
  if (s.type == kDiffuse) {
    if (wo.z <= T(0)) return false;
    const T r = std::sqrt(u1), phi = T(2) * tpi<T>() * u2;
    wi = mkv<T>(r * std::cos(phi), r * std::sin(phi), std::sqrt(std::fmax(T(0), T(1) - u1)));
    weight = s.C;
    return true;
  }
  if (s.type == kConductor) {
    if (wo.z <= T(0)) return false;
    const V3<T> m = s.alpha < T(kSmoothAlpha) ? mkv<T>(T(0), T(0), T(1)) : ggxSampleVisible(wo, s.alpha, u1, u2);
    const T om = dot(wo, m);
    wi = m * (T(2) * om) - wo;
    if (wi.z <= T(0)) return false;
    const V3<T> F = schlick(s.C, dot(wi, m));
    weight = s.alpha < T(kSmoothAlpha) ? F : F * (ggxG2(wo, wi, s.alpha) / ggxG1(wo, s.alpha));
    return true;
  }
  const bool outside = wo.z > T(0);
  const V3<T> w = outside ? wo : mkv<T>(wo.x, wo.y, -wo.z);
  if (w.z == T(0)) return false;
  const T eta = outside ? s.ior : T(1) / s.ior;
  const V3<T> m = s.alpha < T(kSmoothAlpha) ? mkv<T>(T(0), T(0), T(1)) : ggxSampleVisible(w, s.alpha, u1, u2);
  const T c = dot(w, m);
  const T F = fresnelDielectric(c, eta);
  V3<T> wl;
  bool refracted = false;
  if (u3 < F) {
    wl = m * (T(2) * c) - w;
    if (wl.z <= T(0)) return false;
  } else {
    const T s2 = std::fmax(T(0), T(1) - c * c) / (eta * eta);
    if (s2 >= T(1)) return false;
    const T ct = std::sqrt(T(1) - s2);
    wl = m * (c / eta - ct) - w * (T(1) / eta);
    if (wl.z >= T(0)) return false;
    refracted = true;
  }
  const T g = s.alpha < T(kSmoothAlpha) ? T(1) : ggxG2(w, mkv<T>(wl.x, wl.y, std::fabs(wl.z)), s.alpha) / ggxG1(w, s.alpha);
  wi = outside ? wl : mkv<T>(wl.x, wl.y, -wl.z);
  weight = refracted ? s.C * (g / (eta * eta)) : mkv<T>(g, g, g);
  return true;
}

// Density (solid angle) with which bsdfSample draws wi for wo
// (ub_pathtrace.hip `bsdfPdf`); smooth lobes are handled by the callers.
template <class T>
inline T bsdfPdf(const Surf<T>& s, V3<T> wo, V3<T> wi) {
  if (s.type == kDiffuse) return (wo.z > T(0) && wi.z > T(0)) ? wi.z / tpi<T>() : T(0);
  if (s.type == kConductor) {
    if (wo.z <= T(0) || wi.z <= T(0)) return T(0);
    const V3<T> m = normalize(wo + wi);
    return ggxG1(wo, s.alpha) * ggxD(m, s.alpha) / (T(4) * wo.z);
  }
  const bool outside = wo.z > T(0);
  const V3<T> w = outside ? wo : mkv<T>(wo.x, wo.y, -wo.z);
  const V3<T> wl = outside ? wi : mkv<T>(wi.x, wi.y, -wi.z);
  if (w.z == T(0) || wl.z == T(0)) return T(0);
  const T eta = outside ? s.ior : T(1) / s.ior;
  const T g1 = ggxG1(w, s.alpha);
  if (wl.z > T(0)) {
    const V3<T> m = normalize(w + wl);
    return fresnelDielectric(dot(w, m), eta) * g1 * ggxD(m, s.alpha) / (T(4) * w.z);
  }
  V3<T> m = w + wl * eta;
  if (dot(m, m) == T(0)) return T(0);
  m = normalize(m);
  if (m.z < T(0)) m = neg(m);
  const T cw = dot(w, m), cl = dot(wl, m);
  if (cw <= T(0) || cl >= T(0)) return T(0);
  const T dn = cl + cw / eta;
  return (T(1) - fresnelDielectric(cw, eta)) * g1 * cw * ggxD(m, s.alpha) / w.z * std::fabs(cl) / (dn * dn);
}

template <class T>
struct SLight {
  V3<T> c;
  T r;
  V3<T> Le;
};
template <class T>
inline bool lightSample(const SLight<T>& L, V3<T> P, T u1, T u2, int nLights, V3<T>& Q, T& pdfA) {
  const V3<T> dcv = L.c - P;
  const T d2 = dot(dcv, dcv);
  const T r2 = L.r * L.r;
  if (d2 <= r2) return false;
  const T d = std::sqrt(d2);
  const V3<T> axis = dcv * (T(1) / d);
  const T cosMax = std::sqrt(std::fmax(T(0), T(1) - r2 / d2));
  const T cosT = T(1) - u1 * (T(1) - cosMax);
  const T sinT = std::sqrt(std::fmax(T(0), T(1) - cosT * cosT));
  const T phi = T(2) * tpi<T>() * u2;
  const Onb<T> f = onb(axis);
  const V3<T> dir = toWorld(f, mkv<T>(sinT * std::cos(phi), sinT * std::sin(phi), cosT));
  const T b = dot(dcv, dir);
  const T disc = b * b - (d2 - r2);
  const T t = b - std::sqrt(std::fmax(T(0), disc));
  Q = P + dir * t;
  const V3<T> nl = (Q - L.c) * (T(1) / L.r);
  const T cosL = std::fabs(dot(nl, dir));
  const T pdfW = T(1) / (T(2) * tpi<T>() * (T(1) - cosMax));
  pdfA = pdfW * cosL / (t * t) / T(static_cast<float>(nLights));
  return pdfA > T(0);
}
template <class T>
inline T lum(V3<T> c) { return T(0.2126f) * c.x + T(0.7152f) * c.y + T(0.0722f) * c.z; }

struct PMat {
  int type;
  float alpha, ior;
};

template <class T>
struct PCtx {
  const OScene* s;
  int W, H;
  bool ortho;
  V3<T> pos, dir, right, up;
  T halfW, halfH, persHalfW, persHalfH;
  uint32_t seed;
  int bounces, nLights, M, K;
  float radius;
  bool neeAll;
  V3<T> env;
  std::vector<SLight<T>> lights;
  std::vector<PMat> primMats, meshMats;
};

template <class T>
inline V3<T> lightContrib(const PCtx<T>& c, const Surf<T>& s, V3<T> P, V3<T> wo, V3<T> Q, int li, T& dist) {
  const SLight<T>& L = c.lights[li];
  const V3<T> dv = Q - P;
  const T d2 = dot(dv, dv);
  dist = std::sqrt(d2);
  const V3<T> zero{T(0), T(0), T(0)};
  if (!(d2 > T(0))) return zero;
  const V3<T> wi = dv * (T(1) / dist);
  const V3<T> nl = (Q - L.c) * (T(1) / L.r);
  const T cosL = -dot(nl, wi);
  if (cosL <= T(0)) return zero;
  const V3<T> fc = bsdfCos(s, toLocal(s.f, wo), toLocal(s.f, wi));
  const T g = cosL / d2;
  return mkv<T>(fc.x * L.Le.x * g, fc.y * L.Le.y * g, fc.z * L.Le.z * g);
}

// Multiple importance sampling of the lights between next-event estimation
// and BSDF rays (ub_pathtrace.hip `misVertex` / `neeDensity` / `misNee` /
// `misBsdf`): balance heuristic at conductor and dielectric vertices, the
// next-event density of a direction towards light l = (1 / n, or 1 with
// neeAll) / (2 pi (1 - cos theta_max)); a BSDF ray collects every light
// sphere whose near side it crosses before the scene hit.
template <class T>
inline bool misVertex(const Surf<T>& s) { return s.type != kDiffuse; }
template <class T>
inline T neeDensity(const PCtx<T>& c, const SLight<T>& L, V3<T> P) {
  const V3<T> dcv = L.c - P;
  const T d2 = dot(dcv, dcv), r2 = L.r * L.r;
  if (d2 <= r2) return T(0);
  const T cosMax = std::sqrt(std::fmax(T(0), T(1) - r2 / d2));
  const T sel = c.neeAll ? T(1) : T(1) / T(static_cast<float>(c.nLights));
  return sel / (T(2) * tpi<T>() * (T(1) - cosMax));
}
template <class T>
inline T misNee(const PCtx<T>& c, const Surf<T>& sf, V3<T> P, V3<T> wo, V3<T> wi, int li) {
  if (!misVertex(sf)) return T(1);
  const T pb = sf.alpha < T(kSmoothAlpha) ? T(0) : bsdfPdf(sf, toLocal(sf.f, wo), toLocal(sf.f, wi));
  const T pn = neeDensity(c, c.lights[li], P);
  return pn > T(0) ? pn / (pn + pb) : T(0);
}
template <class T>
inline V3<T> misBsdf(const PCtx<T>& c, V3<T> P, V3<T> wi, T tHit, T pb) {
  V3<T> sum{T(0), T(0), T(0)};
  for (int li = 0; li < c.nLights; ++li) {
    const SLight<T>& L = c.lights[li];
    const V3<T> dcv = L.c - P;
    const T b = dot(dcv, wi);
    if (b <= T(0)) continue;
    const T d2 = dot(dcv, dcv), r2 = L.r * L.r;
    const T disc = b * b - (d2 - r2);
    if (disc < T(0) || d2 <= r2) continue;
    const T t = b - std::sqrt(disc);
    if (!(t < tHit)) continue;
    const T pn = neeDensity(c, L, P);
    const T w = std::isinf(pb) ? T(1) : pb / (pb + pn);
    sum = sum + L.Le * w;
  }
  return sum;
}

template <class T>
inline Surf<T> surfaceAt(const PCtx<T>& c, const Hit<T>& h, V3<T> rd, V3<T>& Ng) {
  const Geometry& g = c.s->g;
  V3<T> N, C;
  uint32_t mat;
  bool mesh = false;
  if (h.kind == UBG_KIND_MESH) {
    const uint32_t* ix = &g.idx[3 * static_cast<size_t>(h.id)];
    const T w0 = T(1) - h.u - h.v;
    T nb[3], cb[3];
    for (int k = 0; k < 3; ++k) {
      nb[k] = w0 * T(g.nrm[3 * ix[0] + k]) + h.u * T(g.nrm[3 * ix[1] + k]) + h.v * T(g.nrm[3 * ix[2] + k]);
      cb[k] = w0 * T(g.col[4 * ix[0] + k]) + h.u * T(g.col[4 * ix[1] + k]) + h.v * T(g.col[4 * ix[2] + k]);
    }
    N = normalize(mkv<T>(nb[0], nb[1], nb[2]));
    C = mkv<T>(cb[0], cb[1], cb[2]);
    Ng = normalize(h.ng);
    if (dot(Ng, N) < T(0)) Ng = neg(Ng);
    const int64_t baseTri = static_cast<int64_t>(h.id % static_cast<uint32_t>(g.baseT));
    mat = g.triMat.empty() ? 0u : g.triMat[baseTri];
    mesh = true;
  } else {
    const PrimSide* side = h.kind == UBG_KIND_SPHERE        ? &g.sphSide[h.id]
                           : h.kind == UBG_KIND_CONE_CURVE ? &g.cSide[h.id]
                                                           : &g.rSide[h.id];
    N = normalize(h.ng);
    Ng = N;
    C = mkv<T>(T(side->color[0]), T(side->color[1]), T(side->color[2]));
    mat = static_cast<uint32_t>(side->mat);
  }
  const std::vector<PMat>& tab = mesh ? c.meshMats : c.primMats;
  const PMat m = tab[mat < tab.size() ? mat : 0u];
  if (m.type != kDielectric && dot(N, rd) > T(0)) {
    N = neg(N);
    Ng = neg(Ng);
  }
  Surf<T> sf;
  sf.f = onb(N);
  sf.C = C;
  sf.type = m.type;
  sf.alpha = T(m.alpha);
  sf.ior = T(m.ior);
  return sf;
}

template <class T>
inline V3<T> offsetOrigin(V3<T> P, V3<T> Ng, V3<T> d, T t) {
  const T e = std::fmax(std::fmax(std::fabs(P.x), std::fabs(P.y)), std::fmax(std::fabs(P.z), t)) * T(0x1.0p-19f) +
              T(1.0e-6f);
  return P + Ng * (dot(d, Ng) >= T(0) ? e : -e);
}

template <class T>
inline void cameraRay(const PCtx<T>& c, T fx, T fy, V3<T>& o, V3<T>& d) {
  const T sx = T(2) * fx / T(static_cast<float>(c.W)) - T(1);
  const T sy = T(1) - T(2) * fy / T(static_cast<float>(c.H));
  if (c.ortho) {
    o = c.pos + c.right * (sx * c.halfW) + c.up * (sy * c.halfH);
    d = c.dir;
  } else {
    o = c.pos;
    d = normalize(c.dir + c.right * (sx * c.persHalfW) + c.up * (sy * c.persHalfH));
  }
}

constexpr uint32_t kDimCam = 0;
inline uint32_t dimVertex(int b) { return 16u + 64u * static_cast<uint32_t>(b); }
constexpr uint32_t kDimSpatial = 0x10000u;
inline uint32_t dimAllLights(int b) { return 0x100000u + 0x2000u * static_cast<uint32_t>(b); }

template <class T>
struct Pick {
  V3<T> Q;
  int li;
  T wsum, target;
};

template <class T>
inline bool traceClosest(const PCtx<T>& c, V3<T> o, V3<T> d, Hit<T>& h) {
  h = intersect<T>(*c.s, o, d, T(0), std::numeric_limits<T>::infinity());
  return h.kind >= 0;
}
template <class T>
inline bool traceAny(const PCtx<T>& c, V3<T> o, V3<T> d, T tf) { return occluded<T>(*c.s, o, d, T(0), tf); }

template <class T>
inline Pick<T> risLights(const PCtx<T>& c, const Surf<T>& sf, V3<T> P, V3<T> wo, uint32_t pix, uint32_t smp,
                         uint32_t dim) {
  Pick<T> k{mkv<T>(T(0), T(0), T(0)), -1, T(0), T(0)};
  for (int ci = 0; ci < c.M; ++ci) {
    const uint32_t d = dim + 4u * static_cast<uint32_t>(ci);
    const int li = std::min(c.nLights - 1, static_cast<int>(rndf(c.seed, pix, smp, d) * static_cast<float>(c.nLights)));
    V3<T> Q;
    T pdfA;
    if (!lightSample(c.lights[li], P, rnd<T>(c.seed, pix, smp, d + 1), rnd<T>(c.seed, pix, smp, d + 2), c.nLights, Q,
                     pdfA))
      continue;
    T dist;
    const T t = lum(lightContrib(c, sf, P, wo, Q, li, dist));
    const T w = t / pdfA;
    if (!(w > T(0))) continue;
    k.wsum += w;
    if (rnd<T>(c.seed, pix, smp, d + 3) * k.wsum < w) {
      k.Q = Q;
      k.li = li;
      k.target = t;
    }
  }
  return k;
}

template <class T>
inline V3<T> directLight(const PCtx<T>& c, const Surf<T>& sf, V3<T> P, V3<T> Ng, V3<T> wo, T tIn, uint32_t pix,
                         uint32_t smp, int b, uint64_t& nr) {
  V3<T> sum{T(0), T(0), T(0)};
  if (c.nLights <= 0) return sum;
  if (c.neeAll) {
    for (int li = 0; li < c.nLights; ++li) {
      const uint32_t d = dimAllLights(b) + 2u * static_cast<uint32_t>(li);
      V3<T> Q;
      T pdfA;
      if (!lightSample(c.lights[li], P, rnd<T>(c.seed, pix, smp, d), rnd<T>(c.seed, pix, smp, d + 1), 1, Q, pdfA))
        continue;
      T dist;
      const V3<T> ct = lightContrib(c, sf, P, wo, Q, li, dist);
      if (!(lum(ct) > T(0))) continue;
      const V3<T> wi = (Q - P) * (T(1) / dist);
      ++nr;
      if (traceAny(c, offsetOrigin(P, Ng, wi, tIn), wi, dist * T(0.9999f))) continue;
      sum = sum + ct * (misNee(c, sf, P, wo, wi, li) / pdfA);
    }
    return sum;
  }
  const Pick<T> k = risLights(c, sf, P, wo, pix, smp, dimVertex(b));
  if (k.li < 0 || !(k.target > T(0))) return sum;
  T dist;
  const V3<T> ct = lightContrib(c, sf, P, wo, k.Q, k.li, dist);
  const V3<T> wi = (k.Q - P) * (T(1) / dist);
  ++nr;
  if (traceAny(c, offsetOrigin(P, Ng, wi, tIn), wi, dist * T(0.9999f))) return sum;
  const T Wt = k.wsum / (T(static_cast<float>(c.M)) * k.target);
  return ct * (Wt * misNee(c, sf, P, wo, wi, k.li));
}

// Denoiser guides of a first hit (ub_pathtrace.hip `guides`): a dielectric
// first hit follows the ideal refraction about the interface normal through
// up to 4 dielectric interfaces to the first other surface; albedo = F * white
// + (1 - F) * tint * its colour (F of the first interface, tint = the product
// of the crossed interfaces' colours), normal = its normal; a chain that
// misses, is totally reflected or runs out of hops keeps F * white and the
// first hit's normal. The normal is returned in camera space.
template <class T>
static void guides(const PCtx<T>& c, const Surf<T>& sf0, V3<T> P0, V3<T> Ng0, V3<T> d0, T t0, V3<T>& alb,
                   V3<T>& nrm) {
  V3<T> n = sf0.f.n;
  if (sf0.type != kDielectric) {
    alb = sf0.C;
  } else {
    Surf<T> sf = sf0;
    V3<T> P = P0, Ng = Ng0, d = d0;
    T t = t0, F0 = T(0);
    V3<T> tint = mkv<T>(T(1), T(1), T(1)), behind = mkv<T>(T(0), T(0), T(0));
    for (int k = 0; k < 4; ++k) {
      const T cosO = -dot(d, sf.f.n);
      const bool outside = cosO > T(0);
      const T cc = std::fabs(cosO);
      const T eta = outside ? sf.ior : T(1) / sf.ior;
      const T F = fresnelDielectric(cc, eta);
      if (k == 0) F0 = F;
      const T s2 = std::fmax(T(0), T(1) - cc * cc) / (eta * eta);
      if (s2 >= T(1)) break;
      const T ct = std::sqrt(T(1) - s2);
      const V3<T> ns = outside ? sf.f.n : neg(sf.f.n);
      d = normalize(d * (T(1) / eta) + ns * (cc / eta - ct));
      tint = mkv<T>(tint.x * sf.C.x, tint.y * sf.C.y, tint.z * sf.C.z);
      const V3<T> o = offsetOrigin(P, Ng, d, t);
      Hit<T> h;
      if (!traceClosest(c, o, d, h)) break;
      V3<T> ng;
      sf = surfaceAt(c, h, d, ng);
      P = o + d * h.t;
      Ng = ng;
      t = h.t;
      if (sf.type != kDielectric) {
        behind = mkv<T>(tint.x * sf.C.x, tint.y * sf.C.y, tint.z * sf.C.z);
        n = sf.f.n;
        break;
      }
    }
    const T Tr = T(1) - F0;
    alb = mkv<T>(F0 + Tr * behind.x, F0 + Tr * behind.y, F0 + Tr * behind.z);
  }
  nrm = mkv<T>(dot(n, c.right), dot(n, c.up), -dot(n, c.dir));
}

template <class T>
struct GP {  // first-hit record of one pixel and sample
  bool hit = false;
  V3<T> P, N, Ng, C, wo;
  T t;
  int type;
  T alpha, ior;
  // reservoir
  V3<T> Q;
  int li = -1;
  T wsum = T(0), M = T(0), W = T(0), target = T(0);
};
template <class T>
inline Surf<T> surfOf(const GP<T>& g) {
  Surf<T> sf;
  sf.f = onb(g.N);
  sf.C = g.C;
  sf.type = g.type;
  sf.alpha = g.alpha;
  sf.ior = g.ior;
  return sf;
}

template <class T>
static void primaryPixel(const PCtx<T>& c, int pix, uint32_t smp, GP<T>& g, V3<T>& acc, float* first) {
  const uint32_t upix = static_cast<uint32_t>(pix);
  const int px = pix % c.W, py = pix / c.W;
  V3<T> o, d;
  cameraRay(c, T(static_cast<float>(px)) + rnd<T>(c.seed, upix, smp, kDimCam),
            T(static_cast<float>(py)) + rnd<T>(c.seed, upix, smp, kDimCam + 1), o, d);
  Hit<T> h;
  g = GP<T>();
  if (!traceClosest(c, o, d, h)) {
    acc = acc + c.env;
    if (first && smp == 0)
      for (int k = 0; k < 8; ++k) first[8 * static_cast<size_t>(pix) + k] = 0.0f;
    return;
  }
  V3<T> Ng;
  const Surf<T> sf = surfaceAt(c, h, d, Ng);
  const V3<T> P = o + d * h.t;
  const V3<T> wo = neg(d);
  g.hit = true;
  g.P = P;
  g.t = h.t;
  g.N = sf.f.n;
  g.Ng = Ng;
  g.C = sf.C;
  g.wo = wo;
  g.type = sf.type;
  g.alpha = sf.alpha;
  g.ior = sf.ior;
  if (first && smp == 0) {
    float* f = first + 8 * static_cast<size_t>(pix);
    V3<T> alb, nrm;
    guides(c, sf, P, Ng, d, h.t, alb, nrm);
    f[0] = float(alb.x), f[1] = float(alb.y), f[2] = float(alb.z), f[3] = float(h.t);
    f[4] = float(nrm.x), f[5] = float(nrm.y), f[6] = float(nrm.z), f[7] = 1.0f;
  }
  if (c.nLights > 0 && !c.neeAll) {
    const Pick<T> k = risLights(c, sf, P, wo, upix, smp, dimVertex(0));
    g.Q = k.Q;
    g.li = k.li;
    g.wsum = k.wsum;
    g.M = T(static_cast<float>(c.M));
    g.target = k.target;
    g.W = (k.li >= 0 && k.target > T(0)) ? k.wsum / (g.M * k.target) : T(0);
  }
}

template <class T>
static V3<T> shadePixel(const PCtx<T>& c, int pix, uint32_t smp, const std::vector<GP<T>>& gb, uint64_t& nr) {
  const GP<T>& g = gb[pix];
  const uint32_t upix = static_cast<uint32_t>(pix);
  const Surf<T> sf0 = surfOf(g);
  const V3<T> P0 = g.P, Ng0 = g.Ng, wo0 = g.wo;
  V3<T> L{T(0), T(0), T(0)};
  if (c.nLights > 0) {
    if (c.neeAll) {
      L = directLight(c, sf0, P0, Ng0, wo0, g.t, upix, smp, 0, nr);
    } else {
      const int px = pix % c.W, py = pix / c.W;
      int src[17];
      int nsrc = 0;
      src[nsrc++] = pix;
      for (int j = 0; j < c.K; ++j) {
        const uint32_t d = kDimSpatial + 3u * static_cast<uint32_t>(j);
        const int qx = std::min(c.W - 1, std::max(0, px + static_cast<int>(std::floor(
                                                             (2.0f * rndf(c.seed, upix, smp, d) - 1.0f) * c.radius + 0.5f))));
        const int qy = std::min(c.H - 1, std::max(0, py + static_cast<int>(std::floor(
                                                             (2.0f * rndf(c.seed, upix, smp, d + 1) - 1.0f) * c.radius + 0.5f))));
        const int q = qy * c.W + qx;
        if (q == pix || !gb[q].hit) continue;
        src[nsrc++] = q;
      }
      // balance-heuristic resampling weights m_i = M_i p_i / sum_j M_j p_j
      V3<T> Q{T(0), T(0), T(0)};
      int li = -1;
      T wsum = T(0), tsel = T(0);
      for (int i = 0; i < nsrc; ++i) {
        const GP<T>& r = gb[src[i]];
        if (r.li < 0 || !(r.wsum > T(0))) continue;
        T dist;
        const T t = lum(lightContrib(c, sf0, P0, wo0, r.Q, r.li, dist));
        if (!(t > T(0))) continue;
        T den = T(static_cast<float>(c.M)) * t;
        for (int j = 1; j < nsrc; ++j) {
          const GP<T>& q = gb[src[j]];
          const T tj = j == i ? r.target : lum(lightContrib(c, surfOf(q), q.P, q.wo, r.Q, r.li, dist));
          den += q.M * tj;
        }
        const T w = t * r.wsum / den;
        if (!(w > T(0))) continue;
        wsum += w;
        if (rnd<T>(c.seed, upix, smp, kDimSpatial + 3u * static_cast<uint32_t>(i) + 2u) * wsum < w) {
          Q = r.Q;
          li = r.li;
          tsel = t;
        }
      }
      if (li >= 0 && tsel > T(0)) {
        T dist;
        const V3<T> ct = lightContrib(c, sf0, P0, wo0, Q, li, dist);
        const V3<T> wi = (Q - P0) * (T(1) / dist);
        ++nr;
        if (!traceAny(c, offsetOrigin(P0, Ng0, wi, g.t), wi, dist * T(0.9999f)))
          L = ct * ((wsum / tsel) * misNee(c, sf0, P0, wo0, wi, li));
      }
    }
  }
  V3<T> thr{T(1), T(1), T(1)};
  Surf<T> sf = sf0;
  V3<T> P = P0, Ng = Ng0, wo = wo0;
  T tPrev = g.t;
  for (int b = 1; b <= c.bounces; ++b) {
    const uint32_t dv = dimVertex(b - 1) + 56u;
    V3<T> wiL, wgt;
    if (!bsdfSample(sf, toLocal(sf.f, wo), rnd<T>(c.seed, upix, smp, dv), rnd<T>(c.seed, upix, smp, dv + 1),
                    rnd<T>(c.seed, upix, smp, dv + 2), wiL, wgt))
      break;
    thr = mul(thr, wgt);
    if (!(lum(thr) > T(0))) break;
    const V3<T> wi = toWorld(sf.f, wiL);
    const V3<T> o = offsetOrigin(P, Ng, wi, tPrev);
    Hit<T> h;
    ++nr;
    const bool hit = traceClosest(c, o, wi, h);
    if (c.nLights > 0 && misVertex(sf)) {
      const T pb = sf.alpha < T(kSmoothAlpha) ? std::numeric_limits<T>::infinity()
                                              : bsdfPdf(sf, toLocal(sf.f, wo), wiL);
      L = L + mul(thr, misBsdf(c, P, wi, hit ? h.t : std::numeric_limits<T>::infinity(), pb));
    }
    if (!hit) {
      L = L + mul(thr, c.env);
      break;
    }
    if (b == c.bounces) break;
    V3<T> ng;
    sf = surfaceAt(c, h, wi, ng);
    P = o + wi * h.t;
    Ng = ng;
    wo = neg(wi);
    tPrev = h.t;
    const V3<T> Ld = directLight(c, sf, P, Ng, wo, tPrev, upix, smp, b, nr);
    L = L + mul(thr, Ld);
  }
  return L;
}

template <class T>
static int pathtraceT(const ubg_scene_desc* desc, const ubg_pt_params* p, float* color, float* first,
                      uint64_t* rays, const orc_options* opt) {
  const std::shared_ptr<const OScene> scene = sceneFor(desc, opt->brute_force != 0);
  PCtx<T> c;
  c.s = scene.get();
  c.W = p->width;
  c.H = p->height;
  const ubg_camera& cam = p->camera;
  const T aspect = T(static_cast<float>(c.W)) / T(static_cast<float>(c.H));
  c.dir = normalize(v3<T>(cam.direction));
  c.right = normalize(cross(c.dir, v3<T>(cam.up)));
  c.up = normalize(cross(c.right, c.dir));
  c.pos = v3<T>(cam.position);
  c.ortho = cam.orthographic != 0;
  c.halfH = T(cam.height) * T(0.5f);
  c.halfW = c.halfH * aspect;
  c.persHalfH = std::tan(radiansT(T(cam.fovy)) * T(0.5f));
  c.persHalfW = c.persHalfH * aspect;
  c.seed = p->seed;
  c.bounces = p->bounces;
  c.nLights = p->n_lights;
  c.M = p->restir_candidates > 0 ? p->restir_candidates : 8;
  c.K = p->restir_spatial;
  c.radius = p->restir_radius > 0.0f ? p->restir_radius : 16.0f;
  c.neeAll = p->nee_all_lights != 0;
  c.env = v3<T>(p->env_radiance);
  for (int i = 0; i < p->n_lights; ++i) {
    const ubg_sphere_light& l = p->lights[i];
    c.lights.push_back(SLight<T>{v3<T>(l.center), T(l.radius), v3<T>(l.radiance)});
  }
  auto mats = [](const ubg_pt_material* m, int n) {
    std::vector<PMat> v;
    for (int i = 0; i < n; ++i) v.push_back(PMat{m[i].type, m[i].roughness * m[i].roughness, m[i].ior});
    if (v.empty()) v.push_back(PMat{kDiffuse, 0.0f, 1.5f});
    return v;
  };
  c.primMats = mats(p->prim_materials, p->n_prim_materials);
  c.meshMats = mats(p->mesh_materials, p->n_mesh_materials);
  const int n = c.W * c.H;
  std::vector<GP<T>> gb(static_cast<size_t>(n));
  std::vector<V3<T>> acc(static_cast<size_t>(n), V3<T>{T(0), T(0), T(0)});
  int nth = opt->threads > 0 ? opt->threads : static_cast<int>(std::thread::hardware_concurrency());
  nth = std::max(1, std::min(nth, c.H));
  std::atomic<uint64_t> total{0};
  auto rowsParallel = [&](const std::function<void(int)>& fn) {
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nth; ++t)
      pool.emplace_back([&]() {
        for (int y; (y = next.fetch_add(1)) < c.H;) fn(y);
      });
    for (auto& t : pool) t.join();
  };
  for (int smp = 0; smp < p->spp; ++smp) {
    const uint32_t us = static_cast<uint32_t>(smp);
    rowsParallel([&](int y) {
      for (int x = 0; x < c.W; ++x) {
        const int pix = y * c.W + x;
        primaryPixel(c, pix, us, gb[pix], acc[pix], first);
      }
      total.fetch_add(static_cast<uint64_t>(c.W));
    });
    rowsParallel([&](int y) {
      uint64_t nr = 0;
      for (int x = 0; x < c.W; ++x) {
        const int pix = y * c.W + x;
        if (!gb[pix].hit) continue;
        acc[pix] = acc[pix] + shadePixel(c, pix, us, gb, nr);
      }
      total.fetch_add(nr);
    });
  }
  const T inv = T(1) / T(static_cast<float>(p->spp));
  for (int i = 0; i < n; ++i) {
    color[4 * i] = float(acc[i].x * inv);
    color[4 * i + 1] = float(acc[i].y * inv);
    color[4 * i + 2] = float(acc[i].z * inv);
    color[4 * i + 3] = 1.0f;
  }
  if (rays) *rays = total.load();
  return UBG_OK;
}

template <class T>
static Surf<T> surfFromArgs(int type, float roughness, float ior, const float* C) {
  Surf<T> s;
  s.f = onb(mkv<T>(T(0), T(0), T(1)));
  s.C = v3<T>(C);
  s.type = type;
  s.alpha = T(roughness * roughness);
  s.ior = T(ior);
  return s;
}

}  // namespace pt
}  // namespace orc

extern "C" {

int orc_pathtrace(const ubg_scene_desc* desc, const ubg_pt_params* params, float* color, float* first,
                  uint64_t* rays, const orc_options* opt) {
  if (!desc || !params || !color) return UBG_ERR_INVALID_ARGUMENT;
  if (params->width <= 0 || params->height <= 0 || params->spp < 1 || params->bounces < 1)
    return UBG_ERR_INVALID_ARGUMENT;
  const orc_options* o = opt ? opt : &kDefaultOpt;
  return o->precision == 64 ? orc::pt::pathtraceT<double>(desc, params, color, first, rays, o)
                            : orc::pt::pathtraceT<float>(desc, params, color, first, rays, o);
}

// bsdfCos / bsdfSample in the local frame (n = +z), double precision, over
// n direction pairs / uniform triples
void orc_pt_bsdf_eval(int32_t type, float roughness, float ior, const float* C, int64_t n, const double* wo,
                      const double* wi, double* out) {
  using namespace orc::pt;
  const Surf<double> s = surfFromArgs<double>(type, roughness, ior, C);
  for (int64_t i = 0; i < n; ++i) {
    const double* a = wo + 3 * i;
    const double* b = wi + 3 * i;
    const orc::V3<double> f = bsdfCos(s, orc::V3<double>{a[0], a[1], a[2]}, orc::V3<double>{b[0], b[1], b[2]});
    out[3 * i] = f.x, out[3 * i + 1] = f.y, out[3 * i + 2] = f.z;
  }
}
void orc_pt_bsdf_sample(int32_t type, float roughness, float ior, const float* C, int64_t n, const double* wo,
                        const double* u, double* wi, double* weight, uint8_t* ok) {
  using namespace orc::pt;
  const Surf<double> s = surfFromArgs<double>(type, roughness, ior, C);
  for (int64_t i = 0; i < n; ++i) {
    const double* a = wo + 3 * i;
    orc::V3<double> w{0.0, 0.0, 0.0}, g{0.0, 0.0, 0.0};
    ok[i] = bsdfSample(s, orc::V3<double>{a[0], a[1], a[2]}, u[3 * i], u[3 * i + 1], u[3 * i + 2], w, g) ? 1 : 0;
    wi[3 * i] = w.x, wi[3 * i + 1] = w.y, wi[3 * i + 2] = w.z;
    weight[3 * i] = g.x, weight[3 * i + 1] = g.y, weight[3 * i + 2] = g.z;
  }
}
void orc_pt_bsdf_pdf(int32_t type, float roughness, float ior, const float* C, int64_t n, const double* wo,
                     const double* wi, double* out) {
  using namespace orc::pt;
  const Surf<double> s = surfFromArgs<double>(type, roughness, ior, C);
  for (int64_t i = 0; i < n; ++i) {
    const double* a = wo + 3 * i;
    const double* b = wi + 3 * i;
    out[i] = bsdfPdf(s, orc::V3<double>{a[0], a[1], a[2]}, orc::V3<double>{b[0], b[1], b[2]});
  }
}
float orc_pt_rnd(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim) {
  return orc::pt::rndf(seed, pixel, sample, dim);
}

}  // extern "C"
