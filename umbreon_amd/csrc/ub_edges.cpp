// ub_edges.cpp -- host side of the object-space edge generator. The float
// expressions follow the reference's (edges/analytic_silhouette.cpp,
// edges/mesh_feature_edges.cpp) operation by operation: the edges' endpoints
// are geometry, so they must come out bit for bit.
#include "ub_edges.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <unordered_map>

namespace ubg {
namespace {

constexpr float kPiF = 3.14159265358979323846f;

struct P3 {
  float x, y, z;
};
inline P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline P3 operator*(P3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline P3 cross(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(P3 a) { return std::sqrt(dot(a, a)); }
// scene.hpp:35-38 / :53-57
inline P3 normalize(P3 a) {
  const float l = length(a);
  return l > 0.0f ? a * (1.0f / l) : a;
}
inline P3 safeNormalize(P3 a, P3 fb) {
  const float d = dot(a, a);
  if (d < 1.0e-18f) return fb;
  return a * (1.0f / std::sqrt(d));
}
inline P3 ld(const float* p) { return {p[0], p[1], p[2]}; }
inline void st(P3 v, float* p) {
  p[0] = v.x;
  p[1] = v.y;
  p[2] = v.z;
}

struct Fr {
  P3 t, b, n;
};
// scene.hpp:69-76
inline Fr frameFromNormal(P3 n) {
  const P3 s0{0.0f, n.z, -n.y};
  const P3 s1{-n.z, 0.0f, n.x};
  const P3 t = safeNormalize(std::fabs(n.x) < std::fabs(n.y) ? s0 : s1, P3{1.0f, 0.0f, 0.0f});
  return {t, cross(n, t), n};
}

struct Cam {
  P3 pos, dir;
  bool ortho;
};

// One analytic loop: ordered contour points, closed (ring) or open.
struct Loop {
  std::vector<P3> pts;
  bool closed = false;
  uint16_t group = 0;
  int32_t srcSphere = -1, srcCyl = -1;
};

// The horizon ring of a sphere (analytic_silhouette.cpp:31-79): the great
// circle across the view (orthographic) or the tangent circle seen from the
// eye (perspective), raised outward along the surface normal.
bool sphereRing(P3 O, float r, const Cam& cam, int N, float raise, Loop& out) {
  if (r <= 0.0f) return false;
  P3 view, ringC;  // the ring lies in the plane through ringC normal to view
  float ringR;
  if (cam.ortho) {
    view = normalize(cam.dir);
    ringC = O;
    ringR = r;
  } else {
    const P3 d = O - cam.pos;
    const float L = length(d);
    if (L <= r * 1.0001f) return false;  // the eye on or inside the sphere
    view = d * (1.0f / L);
    ringC = O - view * (r * r / L);
    const float disc = L * L - r * r;
    ringR = r * std::sqrt(disc > 0.0f ? disc : 0.0f) / L;
  }
  if (ringR <= 0.0f) return false;
  const Fr fr = frameFromNormal(view);
  out.pts.resize(static_cast<size_t>(N));
  for (int i = 0; i < N; ++i) {
    const float theta = 2.0f * kPiF * static_cast<float>(i) / static_cast<float>(N);
    const P3 dir = fr.t * std::cos(theta) + fr.b * std::sin(theta);
    P3 P = ringC + dir * ringR;
    const P3 nrm = normalize(P - O);
    P = P + nrm * raise;
    out.pts[static_cast<size_t>(i)] = P;
  }
  out.closed = true;
  return true;
}

// The grazing contact directions on the cross-section circle at axis point P
// (analytic_silhouette.cpp:118-150); false end-on or with the eye within the
// radius of the axis.
bool contactDirs(P3 P, P3 u, float r, const Cam& cam, P3& dPlus, P3& dMinus) {
  constexpr float kEps = 1.0e-4f;
  if (cam.ortho) {
    const P3 toEye = normalize(cam.dir) * -1.0f;
    P3 side = cross(u, toEye);  // across the axis, normal to the view
    const float l = length(side);
    if (l < kEps) return false;
    side = side * (1.0f / l);
    dPlus = side;
    dMinus = side * -1.0f;
    return true;
  }
  const P3 toEye = cam.pos - P;
  const P3 radial = toEye - u * dot(u, toEye);  // the eye's offset from the axis
  const float dist = length(radial);
  if (dist <= r * 1.0001f) return false;
  const P3 inward = radial * (1.0f / dist);
  P3 side = cross(u, inward);
  const float ls = length(side);
  if (ls < kEps) return false;
  side = side * (1.0f / ls);
  // the tangent from the eye touches the circle at angle acos(r / dist) off
  // the direction to the eye
  const float c = r / dist;
  const float s = std::sqrt(std::fmax(0.0f, 1.0f - c * c));
  dPlus = inward * c + side * s;
  dMinus = inward * c - side * s;
  return true;
}

// A cylinder's silhouette (analytic_silhouette.cpp:152-197): two side lines
// through the contacts at both ends, or end-on the rim of the nearer cap.
void cylinderLoops(P3 A, P3 B, float r, uint16_t group, int32_t src, const Cam& cam, int N, float raise,
                   std::vector<Loop>& out) {
  if (r <= 0.0f) return;
  const P3 ab = B - A;
  const float axisLen = length(ab);
  if (axisLen <= 0.0f) return;
  const P3 u = ab * (1.0f / axisLen);
  P3 dAp, dAm, dBp, dBm;
  if (!contactDirs(A, u, r, cam, dAp, dAm) || !contactDirs(B, u, r, cam, dBp, dBm)) {
    const P3 dir = normalize(cam.dir);
    P3 nearCap;
    if (cam.ortho) nearCap = dot(A, dir) < dot(B, dir) ? A : B;
    else nearCap = length(cam.pos - A) < length(cam.pos - B) ? A : B;
    Loop cap;
    const Fr fr = frameFromNormal(normalize(u));
    const float rr = r + raise;
    cap.pts.resize(static_cast<size_t>(N));
    for (int i = 0; i < N; ++i) {
      const float theta = 2.0f * kPiF * static_cast<float>(i) / static_cast<float>(N);
      const P3 dir2 = fr.t * std::cos(theta) + fr.b * std::sin(theta);
      cap.pts[static_cast<size_t>(i)] = nearCap + dir2 * rr;
    }
    cap.closed = true;
    cap.group = group;
    cap.srcCyl = src;
    out.push_back(std::move(cap));
    return;
  }
  const float off = r + raise;
  Loop l1;
  l1.pts = {A + dAp * off, B + dBp * off};
  l1.group = group;
  l1.srcCyl = src;
  out.push_back(std::move(l1));
  Loop l2;
  l2.pts = {A + dAm * off, B + dBm * off};
  l2.group = group;
  l2.srcCyl = src;
  out.push_back(std::move(l2));
}

Cam camOf(const ubg_camera& c) { return Cam{ld(c.position), ld(c.direction), c.orthographic != 0}; }

}  // namespace

void EdgeList::push(const float a[3], const float b[3], const ubg_edge_options& o, uint16_t grp, float op0,
                    float op1) {
  for (int k = 0; k < 3; ++k) p0.push_back(a[k]);
  for (int k = 0; k < 3; ++k) p1.push_back(b[k]);
  for (int k = 0; k < 3; ++k) color.push_back(o.color[k]);
  color.push_back(op0);
  // both ends opaque: uniform (-1), which keeps the edges chainable
  opacity1.push_back(op0 >= 1.0f && op1 >= 1.0f ? -1.0f : op1);
  group.push_back(grp);
}

void analyticChords(const ubg_scene_desc* d, const ubg_camera& camera, int segments, float raise,
                    std::vector<EdgeChord>& out) {
  const Cam cam = camOf(camera);
  const int N = segments < 3 ? 3 : segments;
  std::vector<Loop> loops;
  for (int64_t i = 0; i < d->n_spheres; ++i) {
    if (d->sphere_from_edge && d->sphere_from_edge[i]) continue;
    Loop l;
    const float* s = &d->sphere_center_radius[4 * i];
    if (sphereRing(ld(s), s[3], cam, N, raise, l)) {
      l.group = d->sphere_group ? d->sphere_group[i] : 0;
      l.srcSphere = static_cast<int32_t>(i);
      loops.push_back(std::move(l));
    }
  }
  for (int64_t i = 0; i < d->n_cylinders; ++i) {
    if (d->cyl_from_edge && d->cyl_from_edge[i]) continue;
    cylinderLoops(ld(&d->cyl_p0[3 * i]), ld(&d->cyl_p1[3 * i]), d->cyl_radius[i],
                  d->cyl_group ? d->cyl_group[i] : 0, static_cast<int32_t>(i), cam, N, raise, loops);
  }
  // flattenLoop (object_space_edges.cpp:71-80): N chords of a ring, N - 1 of
  // an open polyline
  for (const Loop& l : loops) {
    const size_t n = l.pts.size();
    if (n < 2) continue;
    const size_t last = l.closed ? n : n - 1;
    for (size_t i = 0; i < last; ++i) {
      EdgeChord c;
      st(l.pts[i], c.a);
      st(l.pts[(i + 1) % n], c.b);
      c.group = l.group;
      c.srcSphere = l.srcSphere;
      c.srcCyl = l.srcCyl;
      out.push_back(c);
    }
  }
}

int chordSamples(const EdgeChord& c, float width) {
  const P3 dd = ld(c.b) - ld(c.a);
  const float len = length(dd);
  if (len <= 0.0f) return 0;
  const float step = width > 0.0f ? width : len;
  return std::max(2, static_cast<int>(std::ceil(len / step)));
}

void chordPoint(const EdgeChord& c, int K, int i, float p[3]) {
  const P3 a = ld(c.a);
  const P3 dd = ld(c.b) - a;
  st(a + dd * (static_cast<float>(i) / static_cast<float>(K)), p);
}

void buildSolidGrid(const ubg_scene_desc* d, SolidGrid& g) {
  const int64_t nS = d->n_spheres, nC = d->n_cylinders;
  g.spheres.assign(d->sphere_center_radius, d->sphere_center_radius + 4 * nS);
  g.cyls.resize(static_cast<size_t>(8 * nC));
  for (int64_t j = 0; j < nC; ++j) {
    for (int k = 0; k < 3; ++k) {
      g.cyls[8 * j + k] = d->cyl_p0[3 * j + k];
      g.cyls[8 * j + 4 + k] = d->cyl_p1[3 * j + k];
    }
    g.cyls[8 * j + 3] = d->cyl_radius[j];
    g.cyls[8 * j + 7] = 0.0f;
  }
  // conservative boxes: a point the float tests call inside lies within
  // its solid's box widened by a relative margin
  std::vector<std::array<float, 6>> box;
  for (int64_t i = 0; i < nS; ++i) {
    const float* s = &g.spheres[4 * i];
    const float r = std::fabs(s[3]);
    const float m = 1e-5f * (r + std::fabs(s[0]) + std::fabs(s[1]) + std::fabs(s[2])) + 1e-30f;
    box.push_back({s[0] - r - m, s[1] - r - m, s[2] - r - m, s[0] + r + m, s[1] + r + m, s[2] + r + m});
  }
  for (int64_t j = 0; j < nC; ++j) {
    const float* c = &g.cyls[8 * j];
    const float r = std::fabs(c[3]);
    std::array<float, 6> b;
    float mag = r;
    for (int k = 0; k < 3; ++k) {
      b[k] = std::fmin(c[k], c[4 + k]) - r;
      b[3 + k] = std::fmax(c[k], c[4 + k]) + r;
      mag += std::fabs(c[k]) + std::fabs(c[4 + k]);
    }
    const float m = 1e-5f * mag + 1e-30f;
    for (int k = 0; k < 3; ++k) {
      b[k] -= m;
      b[3 + k] += m;
    }
    box.push_back(b);
  }
  const size_t n = box.size();
  g.cellStart.assign(1, 0u);
  g.cellItems.clear();
  if (n == 0) {
    g.dim[0] = g.dim[1] = g.dim[2] = 0;
    return;
  }
  float lo[3] = {box[0][0], box[0][1], box[0][2]}, hi[3] = {box[0][3], box[0][4], box[0][5]};
  for (const auto& b : box)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::fmin(lo[k], b[k]);
      hi[k] = std::fmax(hi[k], b[3 + k]);
    }
  // about one solid per cell, at most 2^22 cells
  double ext[3], vol = 1.0;
  for (int k = 0; k < 3; ++k) {
    ext[k] = std::max(1e-6, static_cast<double>(hi[k]) - lo[k]);
    vol *= ext[k];
  }
  double h = std::cbrt(vol / static_cast<double>(n));
  for (;;) {
    double cells = 1.0;
    for (int k = 0; k < 3; ++k) cells *= std::floor(ext[k] / h) + 1.0;
    if (cells <= static_cast<double>(1u << 22)) break;
    h *= 1.25;
  }
  g.inv = static_cast<float>(1.0 / h);
  for (int k = 0; k < 3; ++k) {
    g.lo[k] = lo[k];
    g.dim[k] = static_cast<int>(std::floor(ext[k] / h)) + 1;
  }
  auto cellOf = [&](float v, int k) {
    const float f = std::floor((v - g.lo[k]) * g.inv);
    return std::min(std::max(static_cast<int>(f), 0), g.dim[k] - 1);
  };
  const size_t nCells = static_cast<size_t>(g.dim[0]) * g.dim[1] * g.dim[2];
  std::vector<uint32_t> count(nCells + 1, 0u);
  for (int pass = 0; pass < 2; ++pass) {
    for (size_t i = 0; i < n; ++i) {
      const auto& b = box[i];
      const int x0 = cellOf(b[0], 0), x1 = cellOf(b[3], 0);
      const int y0 = cellOf(b[1], 1), y1 = cellOf(b[4], 1);
      const int z0 = cellOf(b[2], 2), z1 = cellOf(b[5], 2);
      const uint32_t item = i < static_cast<size_t>(nS) ? static_cast<uint32_t>(i)
                                                        : (static_cast<uint32_t>(i - nS) | 0x80000000u);
      for (int z = z0; z <= z1; ++z)
        for (int y = y0; y <= y1; ++y)
          for (int x = x0; x <= x1; ++x) {
            const size_t c = (static_cast<size_t>(z) * g.dim[1] + y) * g.dim[0] + x;
            if (pass == 0) ++count[c + 1];
            else g.cellItems[count[c]++] = item;
          }
    }
    if (pass == 0) {
      for (size_t c = 0; c < nCells; ++c) count[c + 1] += count[c];
      g.cellStart = count;
      g.cellItems.resize(count[nCells]);
    }
  }
}

// ---- mesh feature edges (mesh_feature_edges.cpp:83-704) ----
namespace {

// the undirected welded edge (a, b) as one key
inline uint64_t edgeKey(int a, int b) {
  const uint32_t lo = static_cast<uint32_t>(a < b ? a : b), hi = static_cast<uint32_t>(a < b ? b : a);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

struct FaceAdj {
  int f1 = -1, f2 = -1;
};

// the weld grid of mesh_weld.hpp: positions quantized to 1e-4 world units
struct WeldCell {
  int x, y, z;
  bool operator==(const WeldCell& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct WeldCellHash {
  size_t operator()(const WeldCell& k) const {
    return (static_cast<size_t>(k.x) * 73856093u) ^ (static_cast<size_t>(k.y) * 19349663u) ^
           (static_cast<size_t>(k.z) * 83492791u);
  }
};
inline WeldCell weldCell(P3 p) {
  return WeldCell{static_cast<int>(std::lround(p.x * 1.0e4f)), static_cast<int>(std::lround(p.y * 1.0e4f)),
                  static_cast<int>(std::lround(p.z * 1.0e4f))};
}

// toward the viewer at P: the reversed view direction (orthographic) or the
// direction to the eye
inline P3 viewerDir(P3 P, const Cam& cam) {
  const P3 toCam = normalize(cam.dir * -1.0f);
  if (cam.ortho) return toCam;
  const P3 dd = cam.pos - P;
  const float l = length(dd);
  return l > 0.0f ? dd * (1.0f / l) : toCam;
}

}  // namespace

void meshFeatureEdges(const ubg_scene_desc* d, const ubg_camera& camera, const ubg_edge_options& o, EdgeList& out) {
  if (!(d->n_vertices >= 3 && d->n_triangles > 0)) return;
  if (!o.mesh_silhouette && !o.mesh_crease && !o.mesh_border) return;
  const Cam cam = camOf(camera);
  const size_t nTri = static_cast<size_t>(d->n_triangles), nCorner = 3 * nTri;
  const bool haveNrm = d->normals != nullptr;
  auto cornerV = [&](size_t c) -> size_t { return d->index ? d->index[c] : c; };
  auto cornerPos = [&](size_t c) { return ld(&d->positions[3 * cornerV(c)]); };
  auto cornerNrm = [&](size_t c) { return ld(&d->normals[3 * cornerV(c)]); };
  auto groupOf = [&](size_t t) -> uint16_t { return d->tri_group ? d->tri_group[t] : 0; };
  const float hardCos = std::cos(o.mesh_hard_edge_deg * kPiF / 180.0f);

  // 1) positional weld in first-seen corner order; summed corner normals
  std::vector<P3> vpos, vacc;
  std::vector<int> corner2v(nCorner);
  {
    std::unordered_map<WeldCell, int, WeldCellHash> vmap;
    vmap.reserve(nCorner);
    for (size_t c = 0; c < nCorner; ++c) {
      const P3 p = cornerPos(c);
      auto it = vmap.find(weldCell(p));
      if (it == vmap.end()) {
        const int vi = static_cast<int>(vpos.size());
        vmap.emplace(weldCell(p), vi);
        vpos.push_back(p);
        vacc.push_back(haveNrm ? cornerNrm(c) : P3{0.0f, 0.0f, 0.0f});
        corner2v[c] = vi;
      } else {
        corner2v[c] = it->second;
        if (haveNrm) vacc[static_cast<size_t>(it->second)] = vacc[static_cast<size_t>(it->second)] + cornerNrm(c);
      }
    }
  }
  const size_t nV = vpos.size();

  // 2) welded faces, geometric normals, averaged vertex normals, adjacency
  std::vector<int> fa(nTri), fb(nTri), fc(nTri);
  std::vector<P3> fNg(nTri);
  for (size_t f = 0; f < nTri; ++f) {
    fa[f] = corner2v[3 * f];
    fb[f] = corner2v[3 * f + 1];
    fc[f] = corner2v[3 * f + 2];
    fNg[f] = normalize(cross(vpos[static_cast<size_t>(fb[f])] - vpos[static_cast<size_t>(fa[f])],
                             vpos[static_cast<size_t>(fc[f])] - vpos[static_cast<size_t>(fa[f])]));
  }
  std::vector<P3> vN(nV, P3{0.0f, 0.0f, 0.0f});
  for (size_t v = 0; v < nV; ++v) {
    const float l = length(vacc[v]);
    if (haveNrm && l > 1.0e-6f) vN[v] = vacc[v] * (1.0f / l);
  }
  for (size_t f = 0; f < nTri; ++f)
    for (int e = 0; e < 3; ++e) {
      const size_t v = static_cast<size_t>(e == 0 ? fa[f] : e == 1 ? fb[f] : fc[f]);
      if (length(vN[v]) < 0.5f) vN[v] = fNg[f];
    }
  // The edge map is iterated below, and its iteration order is the order the
  // hard-edge silhouettes, creases and borders come out in: the same
  // container, reservation and insertion sequence keep that order.
  std::unordered_map<uint64_t, FaceAdj> emap;
  emap.reserve(nTri * 3);
  for (size_t f = 0; f < nTri; ++f) {
    const int v[3] = {fa[f], fb[f], fc[f]};
    for (int e = 0; e < 3; ++e) {
      FaceAdj& adj = emap[edgeKey(v[e], v[(e + 1) % 3])];
      if (adj.f1 < 0) adj.f1 = static_cast<int>(f);
      else if (adj.f2 < 0) adj.f2 = static_cast<int>(f);
    }
  }
  // per-corner normal averaged within its smoothing cluster at the welded
  // vertex (corner normals within the hard-edge angle of a cluster's first)
  std::vector<P3> cnrm(nCorner);
  {
    std::vector<std::vector<int>> posCorners(nV);
    for (size_t c = 0; c < nCorner; ++c) posCorners[static_cast<size_t>(corner2v[c])].push_back(static_cast<int>(c));
    for (size_t v = 0; v < nV; ++v) {
      const std::vector<int>& cs = posCorners[v];
      std::vector<P3> rep, sum;
      std::vector<int> cl(cs.size());
      for (size_t k = 0; k < cs.size(); ++k) {
        const size_t c = static_cast<size_t>(cs[k]);
        P3 n = haveNrm ? cornerNrm(c) : fNg[c / 3];
        const float ln = length(n);
        n = ln > 1.0e-12f ? n * (1.0f / ln) : fNg[c / 3];
        int found = -1;
        for (size_t t = 0; t < rep.size(); ++t)
          if (dot(rep[t], n) >= hardCos) {
            found = static_cast<int>(t);
            break;
          }
        if (found < 0) {
          rep.push_back(n);
          sum.push_back(n);
          found = static_cast<int>(rep.size()) - 1;
        } else {
          sum[static_cast<size_t>(found)] = sum[static_cast<size_t>(found)] + n;
        }
        cl[k] = found;
      }
      std::vector<P3> avg(rep.size());
      for (size_t t = 0; t < rep.size(); ++t) {
        const float l = length(sum[t]);
        avg[t] = l > 1.0e-12f ? sum[t] * (1.0f / l) : rep[t];
      }
      for (size_t k = 0; k < cs.size(); ++k) cnrm[static_cast<size_t>(cs[k])] = avg[static_cast<size_t>(cl[k])];
    }
  }
  auto emit = [&](P3 A, P3 B, uint16_t grp) {
    float a[3], b[3];
    st(A, a);
    st(B, b);
    out.push(a, b, o, grp);
  };

  // 3) silhouettes
  if (o.mesh_silhouette) {
    const float w = o.width > 0.0f ? o.width : 0.0f;
    const float silOff = o.raise + w;
    double elsum = 0.0;
    for (size_t f = 0; f < nTri; ++f) {
      const size_t a = static_cast<size_t>(fa[f]), b = static_cast<size_t>(fb[f]), c = static_cast<size_t>(fc[f]);
      elsum += length(vpos[b] - vpos[a]) + length(vpos[c] - vpos[b]) + length(vpos[a] - vpos[c]);
    }
    const float meanEdge = nTri ? static_cast<float>(elsum / (3.0 * nTri)) : 0.0f;
    // eye-ward lift off the grazing tangent shell
    const float camBias = std::fmax(0.5f * w, 0.15f * meanEdge);
    // 3a) the smooth n.v == 0 contour: per face, the segment between the two
    // sign changes of the interpolated n.v
    for (size_t f = 0; f < nTri; ++f) {
      P3 P[3], NN[3];
      float dv[3];
      for (int k = 0; k < 3; ++k) {
        P[k] = vpos[static_cast<size_t>(corner2v[3 * f + k])];
        NN[k] = cnrm[3 * f + k];
        dv[k] = dot(NN[k], viewerDir(P[k], cam));
      }
      P3 cp[2], cn[2];
      int nc = 0;
      auto addCross = [&](int i, int j) {
        const float t = dv[i] / (dv[i] - dv[j]);
        cp[nc] = P[i] + (P[j] - P[i]) * t;
        cn[nc] = normalize(NN[i] + (NN[j] - NN[i]) * t);
        ++nc;
      };
      auto addVert = [&](int i) {
        cp[nc] = P[i];
        cn[nc] = NN[i];
        ++nc;
      };
      int zero[3] = {-1, -1, -1}, nz = 0;
      for (int k = 0; k < 3; ++k)
        if (dv[k] == 0.0f) zero[nz++] = k;
      if (nz == 0) {
        for (int e = 0; e < 3 && nc < 2; ++e)
          if (dv[e] * dv[(e + 1) % 3] < 0.0f) addCross(e, (e + 1) % 3);
      } else if (nz == 1) {
        const int o0 = (zero[0] + 1) % 3, o1 = (zero[0] + 2) % 3;
        if (dv[o0] * dv[o1] < 0.0f) {
          addVert(zero[0]);
          addCross(o0, o1);
        }
      } else if (nz == 2) {
        addVert(zero[0]);
        addVert(zero[1]);
      }
      if (nc == 2) {
        const P3 v0 = viewerDir(cp[0], cam) * camBias, v1 = viewerDir(cp[1], cam) * camBias;
        emit(cp[0] + cn[0] * silOff + v0, cp[1] + cn[1] * silOff + v1, groupOf(f));
      }
    }
    // 3b) hard edges whose two faces straddle the view at an endpoint
    for (const auto& kv : emap) {
      const FaceAdj& adj = kv.second;
      if (adj.f2 < 0) continue;
      const size_t f1 = static_cast<size_t>(adj.f1), f2 = static_cast<size_t>(adj.f2);
      if (dot(fNg[f1], fNg[f2]) >= hardCos) continue;
      const size_t a = static_cast<size_t>(kv.first & 0xffffffffu), b = static_cast<size_t>(kv.first >> 32);
      const P3 da = viewerDir(vpos[a], cam), db = viewerDir(vpos[b], cam);
      const bool sA = dot(da, fNg[f1]) * dot(da, fNg[f2]) < 0.0f;
      const bool sB = dot(db, fNg[f1]) * dot(db, fNg[f2]) < 0.0f;
      if (!sA && !sB) continue;
      const P3 outw = normalize(fNg[f1] + fNg[f2]);
      emit(vpos[a] + outw * silOff + da * camBias, vpos[b] + outw * silOff + db * camBias, groupOf(f1));
    }
  }

  // 4) creases (dihedral beyond the angle, with the optional gates) and
  // borders (one incident face, with the optional continuation veto)
  if (o.mesh_crease || o.mesh_border) {
    const float creaseCos = std::cos(o.crease_angle_deg * kPiF / 180.0f);
    const bool smoothVeto = o.mesh_crease_smooth_veto_deg > 0.0f;
    const float smoothCos = std::cos(o.mesh_crease_smooth_veto_deg * kPiF / 180.0f);
    const bool borderVeto = o.mesh_border_coplanar_veto_deg > 0.0f;
    const float colinCos = std::cos(o.mesh_border_coplanar_veto_deg * kPiF / 180.0f);
    // the raw shading normal face f interpolates at welded vertex v
    auto cornerNrmOf = [&](int f, size_t v) -> P3 {
      const size_t c0 = 3 * static_cast<size_t>(f);
      size_t c = c0;
      if (static_cast<size_t>(corner2v[c0 + 1]) == v) c = c0 + 1;
      else if (static_cast<size_t>(corner2v[c0 + 2]) == v) c = c0 + 2;
      if (!haveNrm) return fNg[static_cast<size_t>(f)];
      const P3 n = cornerNrm(c);
      const float l = length(n);
      return l > 1.0e-12f ? n * (1.0f / l) : fNg[static_cast<size_t>(f)];
    };
    std::unordered_map<int, std::vector<int>> borderAdj;
    if (o.mesh_border && borderVeto)
      for (const auto& kv : emap) {
        if (kv.second.f2 >= 0) continue;
        const int a = static_cast<int>(kv.first & 0xffffffffu), b = static_cast<int>(kv.first >> 32);
        borderAdj[a].push_back(b);
        borderAdj[b].push_back(a);
      }
    // the border chain continues through p nearly straight
    auto continues = [&](size_t p, size_t other) {
      const auto it = borderAdj.find(static_cast<int>(p));
      if (it == borderAdj.end()) return false;
      const P3 in = normalize(vpos[p] - vpos[other]);
      for (int q : it->second) {
        if (static_cast<size_t>(q) == other) continue;
        if (dot(in, normalize(vpos[static_cast<size_t>(q)] - vpos[p])) >= colinCos) return true;
      }
      return false;
    };
    // a valley: the mean face normal points toward the apexes
    auto concave = [&](int f1, int f2, size_t a, size_t b) {
      auto apex = [&](size_t f) -> P3 {
        const size_t x = static_cast<size_t>(fa[f]), y = static_cast<size_t>(fb[f]), z = static_cast<size_t>(fc[f]);
        if (x != a && x != b) return vpos[x];
        if (y != a && y != b) return vpos[y];
        return vpos[z];
      };
      const P3 nAvg = normalize(fNg[static_cast<size_t>(f1)] + fNg[static_cast<size_t>(f2)]);
      const P3 edgeMid = (vpos[a] + vpos[b]) * 0.5f;
      const P3 apexMid = (apex(static_cast<size_t>(f1)) + apex(static_cast<size_t>(f2))) * 0.5f;
      return dot(nAvg, apexMid - edgeMid) > 0.0f;
    };
    auto lifted = [&](size_t v) { return vpos[v] + vN[v] * o.raise; };
    struct Cand {
      size_t a, b;
      uint16_t grp;
    };
    std::vector<Cand> cand;
    std::vector<int> deg;
    const int maxDeg = o.mesh_crease_max_degree;
    if (maxDeg > 0) deg.assign(nV, 0);
    for (const auto& kv : emap) {
      const size_t a = static_cast<size_t>(kv.first & 0xffffffffu), b = static_cast<size_t>(kv.first >> 32);
      const FaceAdj& adj = kv.second;
      if (adj.f2 < 0) {
        if (!o.mesh_border) continue;
        if (borderVeto && continues(a, b) && continues(b, a)) continue;
        emit(lifted(a), lifted(b), groupOf(static_cast<size_t>(adj.f1)));
        continue;
      }
      if (!o.mesh_crease) continue;
      if (dot(fNg[static_cast<size_t>(adj.f1)], fNg[static_cast<size_t>(adj.f2)]) >= creaseCos) continue;
      if (smoothVeto) {
        const float agree = std::fmin(dot(cornerNrmOf(adj.f1, a), cornerNrmOf(adj.f2, a)),
                                      dot(cornerNrmOf(adj.f1, b), cornerNrmOf(adj.f2, b)));
        if (agree >= smoothCos) continue;
      }
      if (o.mesh_crease_convex_only && concave(adj.f1, adj.f2, a, b)) continue;
      const uint16_t grp = groupOf(static_cast<size_t>(adj.f1));
      if (maxDeg > 0) {
        ++deg[a];
        ++deg[b];
        cand.push_back({a, b, grp});
      } else {
        emit(lifted(a), lifted(b), grp);
      }
    }
    for (const Cand& c : cand)
      if (deg[c.a] <= maxDeg && deg[c.b] <= maxDeg) emit(lifted(c.a), lifted(c.b), c.grp);
  }
}

}  // namespace ubg
