// ub_device.h -- math, primitive intersectors and BVH traversal of the
// per-ray path. Every function is __host__ __device__: the gfx950 kernels run
// them, and the library's host test hook (ubgx_host_query) runs the very same
// code on the CPU so the traversal can be checked without a GPU.
//
// The whole library is compiled with -ffp-contract=off and correctly rounded
// fp32 divide/sqrt, so every expression below rounds exactly as written. The
// intersectors and the shading chain use the same operation order as the
// reference's float code (scene.hpp vector helpers, shading/*.hpp), which is
// what makes first hits bit-identical to the CPU restatement. Only the
// traversal box tests use explicit FMAs: they merely decide which nodes are
// visited and every box is inflated by the builder, so they can never cull a
// primitive the exact test would hit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ub_layout.h"

#define UBG_HD __host__ __device__ __forceinline__

namespace ubg {

UBG_HD uint32_t f2u(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return u;
}
UBG_HD float u2f(uint32_t u) {
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

struct V3 {
  float x, y, z;
};
UBG_HD V3 mk(float x, float y, float z) { return V3{x, y, z}; }
UBG_HD V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
UBG_HD V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
UBG_HD V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
UBG_HD V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
UBG_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
UBG_HD V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
UBG_HD float len(V3 a) { return sqrtf(dot(a, a)); }
// scene.hpp:35-38
UBG_HD V3 normalize(V3 a) {
  const float l = len(a);
  return l > 0.0f ? a * (1.0f / l) : a;
}
// scene.hpp:44-49
UBG_HD V3 safeNormalize(V3 a) {
  const float s = 1.0f / sqrtf(fmaxf(dot(a, a), 0x1.0p-126f));
  return a * s;
}
// scene.hpp:53-57
UBG_HD V3 safeNormalize(V3 a, V3 fb) {
  const float d = dot(a, a);
  if (d < 1.0e-18f) return fb;
  return a * (1.0f / sqrtf(d));
}
// shading.hpp:27-30
UBG_HD V3 faceForward(V3 n, V3 rd) { return dot(n, rd) > 0.0f ? neg(n) : n; }

struct Frame3 {
  V3 t, b, n;
};
// scene.hpp:69-76
UBG_HD Frame3 frameFromNormal(V3 n) {
  const V3 s0 = mk(0.0f, n.z, -n.y);
  const V3 s1 = mk(-n.z, 0.0f, n.x);
  const V3 t = safeNormalize(fabsf(n.x) < fabsf(n.y) ? s0 : s1, mk(1.0f, 0.0f, 0.0f));
  return Frame3{t, cross(n, t), n};
}
// secondary_rays.hpp:38-46
UBG_HD float selfIntersectEps(V3 P, V3 d, float t) {
  const float dirMax = fmaxf(fabsf(d.x), fmaxf(fabsf(d.y), fabsf(d.z)));
  const float s = fmaxf(fmaxf(fabsf(P.x), fabsf(P.y)), fmaxf(fabsf(P.z), dirMax * t));
  return s * 0x1.0fp-21f;
}
// secondary_rays.hpp:51-58 (TEA-8)
UBG_HD void tea2(uint32_t& v0, uint32_t& v1) {
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    sum += 0x9E3779B9u;
    v0 += ((v1 << 4) + 0xA341316Cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xC8013EA4u);
    v1 += ((v0 << 4) + 0xAD90777Du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7E95761Eu);
  }
}
UBG_HD float u32ToUnorm(uint32_t u) { return (float)(u >> 8) * 0x1.0p-24f; }
// secondary_rays.hpp:66-73 (Malley cosine hemisphere)
UBG_HD V3 cosineSample(float u1, float u2, const Frame3& f) {
  const float phi = 6.2831853072f * u1;
  const float r = sqrtf(u2);
  const float z = sqrtf(fmaxf(0.0f, 1.0f - u2));
  const float x = cosf(phi) * r;
  const float y = sinf(phi) * r;
  return f.t * x + f.b * y + f.n * z;
}
// secondary_rays.hpp:102-109
UBG_HD V3 sampleLightDir(V3 L, float tanR, float u1, float u2) {
  const Frame3 f = frameFromNormal(L);
  const float phi = 6.2831853072f * u1;
  const float rr = tanR * sqrtf(u2);
  const V3 d = f.n + f.t * (rr * cosf(phi)) + f.b * (rr * sinf(phi));
  return safeNormalize(d, L);
}

UBG_HD V3 ld3(const float4& a) { return mk(a.x, a.y, a.z); }

// ambient_occlusion.hpp:96-104 (base-2 radical inverse = bit reversal)
UBG_HD uint32_t reverseBits32(uint32_t b) {
  b = (b << 16) | (b >> 16);
  b = ((b & 0x55555555u) << 1) | ((b & 0xAAAAAAAAu) >> 1);
  b = ((b & 0x33333333u) << 2) | ((b & 0xCCCCCCCCu) >> 2);
  b = ((b & 0x0F0F0F0Fu) << 4) | ((b & 0xF0F0F0F0u) >> 4);
  b = ((b & 0x00FF00FFu) << 8) | ((b & 0xFF00FF00u) >> 8);
  return b;
}
// ambient_occlusion.hpp:111-125: the legacy tea2 stream or the Hammersley
// point shifted per pixel by (cpx, cpy) mod 1
UBG_HD void aoSample2d(bool lowDisc, uint32_t base, int i, int n, float cpx, float cpy, float& u1, float& u2) {
  if (lowDisc) {
    u1 = static_cast<float>(i) / static_cast<float>(n) + cpx;
    u2 = static_cast<float>(reverseBits32(static_cast<uint32_t>(i))) * 0x1.0p-32f + cpy;
    if (u1 >= 1.0f) u1 -= 1.0f;
    if (u2 >= 1.0f) u2 -= 1.0f;
  } else {
    uint32_t s0 = base, s1 = static_cast<uint32_t>(i);
    tea2(s0, s1);
    u1 = u32ToUnorm(s0);
    u2 = u32ToUnorm(s1);
  }
}
// ambient_occlusion.hpp:133-138 (GTAO multibounce)
UBG_HD float aoMultibounce(float ao, float albedo) {
  const float a = 2.0404f * albedo - 0.3324f;
  const float b = -4.7951f * albedo + 0.6417f;
  const float c = 2.7552f * albedo + 0.6903f;
  return fmaxf(ao, ((ao * a + b) * ao + c) * ao);
}

struct Hit {
  float t, u, v;
  uint32_t bits;
  V3 ng;
};

// Closest-hit order: the nearer t first; at equal t the smaller
// (kind << 30 | primID) wins, so which of two primitives hit at exactly the
// same distance is reported does not depend on the traversal order (BVH
// layout, builder, leaf size).
UBG_HD bool closer(const Hit& h, const Hit& best) { return h.t < best.t || (h.t == best.t && h.bits < best.bits); }

// ---------------------------------------------------------------------------
// Intersectors. A hit is accepted when tnear <= t <= tfar. Spheres and curves
// solve in a frame moved along the ray to the closest approach of the
// primitive's center so far cameras do not cancel the discriminant.

// ---------------------------------------------------------------------------
// FP64 polish of grazing roots. In the shifted frame a sphere's or a cone
// body's discriminant is a small difference of squares for a ray passing
// near the silhouette, and its FP32 terms carry the rounding of the shifted
// origin: an ulp of the coordinates' magnitude m. At large coordinates
// (m ~ 1e4) a ray grazing within ~m 2^-24 of the silhouette is then dropped
// or hit at a wrong t. When |disc| lies within that error budget
// (grazeSlack) the quadratic is set up again in FP64 from the exact inputs
// -- the shifted origin rebuilt as (o - p) + d ts, exact float inputs
// subtracted in double -- solved there, and each root takes one Newton step
// on the FP64 quadratic. Away from the silhouette the FP32 path is untouched.
struct V3d {
  double x, y, z;
};
UBG_HD V3d toD(V3 v) { return V3d{v.x, v.y, v.z}; }
UBG_HD double dotD(V3d a, V3d b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
UBG_HD V3d subD(V3d a, V3d b) { return V3d{a.x - b.x, a.y - b.y, a.z - b.z}; }
// (o - p) + d ts in FP64
UBG_HD V3d shiftedD(V3 o, V3 p, V3 d, float ts) {
  const double t = ts;
  return V3d{(static_cast<double>(o.x) - static_cast<double>(p.x)) + static_cast<double>(d.x) * t,
             (static_cast<double>(o.y) - static_cast<double>(p.y)) + static_cast<double>(d.y) * t,
             (static_cast<double>(o.z) - static_cast<double>(p.z)) + static_cast<double>(d.z) * t};
}
UBG_HD float maxAbs(V3 v) { return fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fabsf(v.z)); }
// |disc| below which the FP32 discriminant is not trusted: scale x size x
// magnitude of the rounded coordinates, with a 2^5 margin over 2^-24.
UBG_HD float grazeSlack(float scale, float size, float m) { return scale * 0x1.0p-19f * size * m; }
// One Newton step on A x^2 + 2 B x + C at x, taken when it moves x by at most
// `half`.
UBG_HD double newtonStep64(double A, double B, double C, double x, double half) {
  const double g = (A * x + 2.0 * B) * x + C;
  const double gp = 2.0 * (A * x + B);
  if (gp != 0.0) {
    const double st = g / gp;
    if (fabs(st) <= half) return x - st;
  }
  return x;
}
// Roots of A x^2 + 2 B x + C in FP64, each with one Newton step (taken when
// it moves the root by at most half the roots' separation).
UBG_HD bool quadPolish64(double A, double B, double C, double& x0, double& x1) {
  const double disc = B * B - A * C;
  if (disc < 0.0) return false;
  const double q = sqrt(disc);
  x0 = (-B - q) / A;
  x1 = (-B + q) / A;
  const double half = 0.5 * fabs(x1 - x0);
  x0 = newtonStep64(A, B, C, x0, half);
  x1 = newtonStep64(A, B, C, x1, half);
  return true;
}

UBG_HD bool isectSphere(V3 o, V3 d, float tnear, float tfar, V3 c, float r, Hit& h) {
  const float A = dot(d, d);
  const float ts = dot(c - o, d) / A;
  const V3 f = (o + d * ts) - c;
  const float B = dot(f, d);
  const float C = dot(f, f) - r * r;
  const float disc = B * B - A * C;
  float t0, t1;
  if (fabsf(disc) <= grazeSlack(A, r, fmaxf(maxAbs(o), maxAbs(c)))) {
    const V3d fd = shiftedD(o, c, d, ts), dd = toD(d);
    const double rd = r;
    double x0, x1;
    if (!quadPolish64(dotD(dd, dd), dotD(fd, dd), dotD(fd, fd) - rd * rd, x0, x1)) return false;
    t0 = static_cast<float>(static_cast<double>(ts) + x0);
    t1 = static_cast<float>(static_cast<double>(ts) + x1);
  } else {
    if (disc < 0.0f) return false;
    const float q = sqrtf(disc);
    t0 = ts + (-B - q) / A;
    t1 = ts + (-B + q) / A;
  }
  float t;
  if (tnear <= t0 && t0 <= tfar) t = t0;
  else if (tnear <= t1 && t1 <= tfar) t = t1;
  else return false;
  h.t = t;
  h.u = 0.0f;
  h.v = 0.0f;
  h.ng = (o + d * t) - c;
  return true;
}

// Roots (in the shifted frame) of the infinite cone r(h) = r0 + k*h around
// the unit axis w through the origin of `oa`. Returns false when none.
// oa = (o + d ts) - a rounded in FP32; a grazing discriminant is solved again
// from (o - a) + d ts in FP64, with the axis and slope from the segment ax =
// b - a and the end radii (m: magnitude of the rounded coordinates, ext: the
// segment length, bounding |oa|).
UBG_HD bool coneBody(V3 oa, V3 d, V3 w, float r0, float k, float& dt0, float& dt1, V3 o, V3 a, V3 ax, float r1,
                     float ts, float m, float ext) {
  const float ho = dot(oa, w);
  const float hd = dot(d, w);
  const float rr = r0 + k * ho;
  const float A = dot(d, d) - hd * hd - k * k * hd * hd;
  const float B = dot(oa, d) - ho * hd - k * hd * rr;
  const float C = dot(oa, oa) - ho * ho - rr * rr;
  if (A == 0.0f) return false;
  const float disc = B * B - A * C;
  if (fabsf(disc) <= grazeSlack(fabsf(A) + dot(d, d), fabsf(rr) + fabsf(r0) + ext, m)) {
    // the axis, its length and the slope again in FP64 from the segment
    const V3d od = shiftedD(o, a, d, ts), dd = toD(d), axd = toD(ax);
    const double Ld = sqrt(dotD(axd, axd));
    const V3d wd{axd.x / Ld, axd.y / Ld, axd.z / Ld};
    const double kd = (static_cast<double>(r1) - static_cast<double>(r0)) / Ld;
    const double hoD = dotD(od, wd), hdD = dotD(dd, wd), rrD = static_cast<double>(r0) + kd * hoD;
    const double AD = dotD(dd, dd) - hdD * hdD - kd * kd * hdD * hdD;
    const double BD = dotD(od, dd) - hoD * hdD - kd * hdD * rrD;
    const double CD = dotD(od, od) - hoD * hoD - rrD * rrD;
    double x0, x1;
    if (AD == 0.0 || !quadPolish64(AD, BD, CD, x0, x1)) return false;
    dt0 = static_cast<float>(x0);
    dt1 = static_cast<float>(x1);
    return true;
  }
  if (disc < 0.0f) return false;
  const float q = sqrtf(disc);
  dt0 = (-B - q) / A;
  dt1 = (-B + q) / A;
  return true;
}

// CONE_LINEAR_CURVE: body between the end planes plus flat disk caps. L =
// |b - a|, invL = 1 / L and k = (rb - ra) / L come precomputed with the record
// (the same float operations, done once by the builder).
UBG_HD bool isectCone(V3 o, V3 d, float tnear, float tfar, V3 a, float ra, V3 b, float rb, float L, float invL,
                      float k, Hit& h) {
  const V3 ax = b - a;
  if (!(L > 0.0f)) return false;
  const V3 w = ax * invL;
  const float A = dot(d, d);
  const V3 mid = (a + b) * 0.5f;
  const float ts = dot(mid - o, d) / A;
  const V3 oa = (o + d * ts) - a;
  const float ho = dot(oa, w);
  const float hd = dot(d, w);
  bool found = false;
  float best = tfar, bu = 0.0f;
  V3 bn = mk(0.0f, 0.0f, 0.0f);
  float r0, r1;
  if (coneBody(oa, d, w, ra, k, r0, r1, o, a, ax, rb, ts, fmaxf(maxAbs(o), maxAbs(mid)), L)) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const float dt = i == 0 ? r0 : r1;
      const float t = ts + dt;
      if (!(tnear <= t && t <= best)) continue;
      const float hh = ho + dt * hd;
      if (hh < 0.0f || hh > L) continue;
      if (ra + k * hh < 0.0f) continue;
      const V3 p = oa + d * dt;
      const V3 radial = p - w * hh;
      const float rl = len(radial);
      V3 n = radial;
      if (k != 0.0f && rl > 0.0f) n = radial * (1.0f / rl) - w * k;
      best = t;
      bu = hh / L;
      bn = n;
      found = true;
    }
  }
  if (hd != 0.0f) {
    {
      const float dtc = -ho / hd;
      const float t = ts + dtc;
      if (tnear <= t && t <= best) {
        const V3 p = oa + d * dtc;
        if (dot(p, p) <= ra * ra) {
          best = t;
          bu = 0.0f;
          bn = neg(w);
          found = true;
        }
      }
    }
    {
      const float dtc = (L - ho) / hd;
      const float t = ts + dtc;
      if (tnear <= t && t <= best) {
        const V3 p = (oa + d * dtc) - w * L;
        if (dot(p, p) <= rb * rb) {
          best = t;
          bu = 1.0f;
          bn = w;
          found = true;
        }
      }
    }
  }
  if (!found) return false;
  h.t = best;
  h.u = bu;
  h.v = 0.0f;
  h.ng = bn;
  return true;
}

// ROUND_LINEAR_CURVE segment restricted to its cell between the joint planes
// (nL at a when hasL, nR at b when hasR): a chained polyline has no internal
// caps, which is what RTC_CURVE_FLAG_NEIGHBOR_LEFT/RIGHT request.
UBG_HD bool isectRound(V3 o, V3 d, float tnear, float tfar, V3 a, float ra, V3 b, float rb, float4 pl,
                       float4 pr, Hit& h) {
  const V3 ax = b - a;
  const float L = len(ax);
  if (!(L > 0.0f)) return false;
  const V3 w = ax * (1.0f / L);
  const float k = (rb - ra) / L;
  const bool hasL = pl.w != 0.0f, hasR = pr.w != 0.0f;
  const V3 nL = ld3(pl), nR = ld3(pr);
  const float A = dot(d, d);
  const V3 mid = (a + b) * 0.5f;
  const float ts = dot(mid - o, d) / A;
  const V3 oa = (o + d * ts) - a;
  const float ho = dot(oa, w);
  const float hd = dot(d, w);
  const float m = fmaxf(maxAbs(o), maxAbs(mid));
  bool found = false;
  float best = tfar, bu = 0.0f;
  V3 bn = mk(0.0f, 0.0f, 0.0f);
  // candidate roots: 2 on the body, 2 on each end sphere
  float cdt[6];
  int ckind[6];  // 0 body, 1 sphere at a, 2 sphere at b, -1 none
  {
    float r0 = 0.0f, r1 = 0.0f;
    const bool ok = coneBody(oa, d, w, ra, k, r0, r1, o, a, ax, rb, ts, m, L);
    cdt[0] = r0;
    cdt[1] = r1;
    ckind[0] = ok ? 0 : -1;
    ckind[1] = ok ? 0 : -1;
  }
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const V3 c = e == 0 ? mk(0.0f, 0.0f, 0.0f) : ax;
    const float r = e == 0 ? ra : rb;
    const V3 f = oa - c;
    const float B = dot(f, d);
    const float C = dot(f, f) - r * r;
    const float disc = B * B - A * C;
    bool ok;
    if (fabsf(disc) <= grazeSlack(A, r + L, m)) {
      const V3d fd = subD(shiftedD(o, a, d, ts), toD(c)), dd = toD(d);
      const double rd = r;
      double x0 = 0.0, x1 = 0.0;
      ok = quadPolish64(dotD(dd, dd), dotD(fd, dd), dotD(fd, fd) - rd * rd, x0, x1);
      cdt[2 + 2 * e] = static_cast<float>(x0);
      cdt[3 + 2 * e] = static_cast<float>(x1);
    } else {
      ok = disc >= 0.0f;
      const float q = ok ? sqrtf(disc) : 0.0f;
      cdt[2 + 2 * e] = (-B - q) / A;
      cdt[3 + 2 * e] = (-B + q) / A;
    }
    ckind[2 + 2 * e] = ok ? 1 + e : -1;
    ckind[3 + 2 * e] = ok ? 1 + e : -1;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    if (ckind[i] < 0) continue;
    const float dtv = cdt[i];
    const float hh = ho + dtv * hd;
    if (ckind[i] == 0 && (hh < 0.0f || hh > L)) continue;
    if (ckind[i] == 1 && hh >= 0.0f) continue;
    if (ckind[i] == 2 && hh <= L) continue;
    const float t = ts + dtv;
    if (!(tnear <= t && t <= best)) continue;
    const V3 p = oa + d * dtv;
    if (hasL && dot(p, nL) < 0.0f) continue;
    if (hasR && dot(p - ax, nR) > 0.0f) continue;
    const V3 n = ckind[i] == 0 ? p - w * hh : (ckind[i] == 1 ? p : p - ax);
    best = t;
    float uu = hh / L;
    if (uu < 0.0f) uu = 0.0f;
    if (uu > 1.0f) uu = 1.0f;
    bu = uu;
    bn = n;
    found = true;
  }
  if (!found) return false;
  h.t = best;
  h.u = bu;
  h.v = 0.0f;
  h.ng = bn;
  return true;
}

// TRIANGLE (Moller-Trumbore): P = (1-u-v) v0 + u v1 + v v2, Ng = (v1-v0)x(v2-v0).
UBG_HD bool isectTri(V3 o, V3 d, float tnear, float tfar, V3 v0, V3 v1, V3 v2, Hit& h) {
  const V3 e1 = v1 - v0;
  const V3 e2 = v2 - v0;
  const V3 pv = cross(d, e2);
  const float det = dot(e1, pv);
  if (det == 0.0f) return false;
  const float inv = 1.0f / det;
  const V3 tv = o - v0;
  const float u = dot(tv, pv) * inv;
  if (u < 0.0f || u > 1.0f) return false;
  const V3 qv = cross(tv, e1);
  const float v = dot(d, qv) * inv;
  if (v < 0.0f || u + v > 1.0f) return false;
  const float t = dot(e2, qv) * inv;
  if (!(tnear <= t && t <= tfar)) return false;
  h.t = t;
  h.u = u;
  h.v = v;
  h.ng = cross(e1, e2);
  return true;
}

UBG_HD bool isectPrim(const SceneView& s, uint32_t i, V3 o, V3 d, float tnear, float tfar, Hit& h) {
  const float4* p = reinterpret_cast<const float4*>(s.prims) + 3 * static_cast<size_t>(i);
  const float4 p0 = p[0];
  const float4 p1 = p[1];
  const float4 p2 = p[2];
  const uint32_t bits = f2u(p2.w);
  const uint32_t kind = bits >> kKindShift;
  bool r;
  if (kind == 1u) {
    r = isectSphere(o, d, tnear, tfar, ld3(p0), p0.w, h);
  } else if (kind == 3u) {
    r = isectCone(o, d, tnear, tfar, ld3(p0), p0.w, ld3(p1), p1.w, p2.x, p2.y, p2.z, h);
  } else if (kind == 0u) {
    r = isectTri(o, d, tnear, tfar, ld3(p0), mk(p0.w, p1.x, p1.y), mk(p1.z, p1.w, p2.x), h);
  } else {
    const float4* pl = reinterpret_cast<const float4*>(s.rndPlanes) + 2 * (bits & kIdMask);
    r = isectRound(o, d, tnear, tfar, ld3(p0), p0.w, ld3(p1), p1.w, pl[0], pl[1], h);
  }
  h.bits = bits;
  return r;
}

// Hit filters (the Embree argument-filter stage): a rejected hit does not
// count and the query goes on. AcceptAll is the plain query.
struct AcceptAll {
  UBG_HD bool operator()(const Hit&) const { return true; }
};

// isectPrim with a filter: when the filter rejects the first root in
// [tnear, tfar], the primitive's next root beyond it is tried (a primitive
// has at most four roots along a line: two body roots and two caps).
template <class F>
UBG_HD bool isectPrimF(const SceneView& s, uint32_t i, V3 o, V3 d, float tnear, float tfar, Hit& h, const F& f) {
  float lo = tnear;
#pragma unroll 1
  for (int k = 0; k < 4; ++k) {
    if (!isectPrim(s, i, o, d, lo, tfar, h)) return false;
    if (f(h)) return true;
    lo = h.t > 0.0f ? u2f(f2u(h.t) + 1u) : (h.t < 0.0f ? u2f(f2u(h.t) - 1u) : 0x1.0p-149f);
  }
  return false;
}

// Reciprocal direction with zero components replaced by a signed tiny value
// (keeps every slab product finite).
UBG_HD V3 safeInv(V3 d) {
  return mk(1.0f / (d.x != 0.0f ? d.x : copysignf(1e-30f, d.x)),
            1.0f / (d.y != 0.0f ? d.y : copysignf(1e-30f, d.y)),
            1.0f / (d.z != 0.0f ? d.z : copysignf(1e-30f, d.z)));
}

// ---------------------------------------------------------------------------
// BVH2 traversal (full-precision nodes). Stack slot i of this lane lives at
// stk[i * stride] (LDS on the GPU: stride = block size, so the 64 lanes of a
// wave use 64 distinct banks). kAny returns at the first hit in
// [tnear, tfar]; otherwise the closest hit in the order of closer().
template <bool kAny, bool kCount, class F = AcceptAll>
UBG_HD bool traverse2(const SceneView& s, V3 o, V3 d, float tnear, float tfar, Hit& best, uint32_t* stk,
                      int stride, uint32_t* nodeVisits, uint32_t* primTests, const F& filt = F()) {
  if (s.nNodes == 0) return false;
  const V3 iv = safeInv(d);
  const float oix = o.x * iv.x, oiy = o.y * iv.y, oiz = o.z * iv.z;
  const float4* nodes = reinterpret_cast<const float4*>(s.nodes);
  bool found = false;
  float tmax = tfar;
  int sp = 0;
  int node = 0;
  Hit h;
  for (;;) {
    const float4* n = nodes + 4 * node;
    const float4 a = n[0];
    const float4 b = n[1];
    const float4 c = n[2];
    const float4 cf = n[3];
    const int c0 = static_cast<int>(f2u(cf.x)), c1 = static_cast<int>(f2u(cf.y));
    if (kCount) ++*nodeVisits;
    const float tmx = tmax * 1.00000012f;
    const float l0x = __builtin_fmaf(a.x, iv.x, -oix), h0x = __builtin_fmaf(a.y, iv.x, -oix);
    const float l0y = __builtin_fmaf(a.z, iv.y, -oiy), h0y = __builtin_fmaf(a.w, iv.y, -oiy);
    const float l0z = __builtin_fmaf(c.x, iv.z, -oiz), h0z = __builtin_fmaf(c.y, iv.z, -oiz);
    const float l1x = __builtin_fmaf(b.x, iv.x, -oix), h1x = __builtin_fmaf(b.y, iv.x, -oix);
    const float l1y = __builtin_fmaf(b.z, iv.y, -oiy), h1y = __builtin_fmaf(b.w, iv.y, -oiy);
    const float l1z = __builtin_fmaf(c.z, iv.z, -oiz), h1z = __builtin_fmaf(c.w, iv.z, -oiz);
    const float e0 = fmaxf(fmaxf(fminf(l0x, h0x), fminf(l0y, h0y)), fmaxf(fminf(l0z, h0z), tnear));
    const float x0 = fminf(fminf(fmaxf(l0x, h0x), fmaxf(l0y, h0y)), fminf(fmaxf(l0z, h0z), tmx));
    const float e1 = fmaxf(fmaxf(fminf(l1x, h1x), fminf(l1y, h1y)), fmaxf(fminf(l1z, h1z), tnear));
    const float x1 = fminf(fminf(fmaxf(l1x, h1x), fmaxf(l1y, h1y)), fminf(fmaxf(l1z, h1z), tmx));
    bool hit0 = e0 <= x0;
    bool hit1 = e1 <= x1;
    const bool near1 = e1 < e0;
    // Leaves are intersected on the spot, nearer one first.
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const bool isChild0 = (k == 0) != near1;
      const int cidx = isChild0 ? c0 : c1;
      const bool hc = isChild0 ? hit0 : hit1;
      if (!hc || cidx >= 0) continue;
      const uint32_t code = ~static_cast<uint32_t>(cidx);
      const uint32_t beg = code >> 3;
      const uint32_t cnt = (code & 7u) + 1u;
      for (uint32_t i = beg; i < beg + cnt; ++i) {
        if (kCount) ++*primTests;
        if (isectPrimF(s, i, o, d, tnear, tmax, h, filt)) {
          if (kAny) return true;
          if (!found || closer(h, best)) {
            best = h;
            tmax = h.t;
            found = true;
          }
        }
      }
      if (isChild0) hit0 = false;
      else hit1 = false;
    }
    if (hit0 && hit1) {
      stk[sp * stride] = static_cast<uint32_t>(near1 ? c0 : c1);
      ++sp;
      node = near1 ? c1 : c0;
    } else if (hit0) {
      node = c0;
    } else if (hit1) {
      node = c1;
    } else {
      if (sp == 0) break;
      --sp;
      node = static_cast<int>(stk[sp * stride]);
    }
  }
  return found;
}

// ---------------------------------------------------------------------------
// Compressed 8-wide BVH (80 B nodes, 8-bit child boxes).
//
// Node = 5 x uint4:
//   [0] px, py, pz (box origin, f32), w = ex | ey << 8 | ez << 16 | imask << 24
//       with per-axis scale 2^(e - 127) and imask bit i = slot i is an inner node
//   [1] childBase (first inner child), primBase (first leaf primitive),
//       meta[0..3], meta[4..7]: 0 = empty slot; leaf = (unary prim count) << 5
//       | offset from primBase; inner = 1 << 5 | (24 + slot)
//   [2..4] qlo.x[8], qlo.y[8], qlo.z[8], qhi.x[8], qhi.y[8], qhi.z[8] (u8)
// The builder places each child in the slot whose octant (bit k = child centre
// on the + side of the node centre along axis k) matches its position, so
// visiting inner children in increasing (slot ^ rayOctant) order is roughly
// front to back without sorting. A traversal stack entry is a node group
// (childBase, pending inner-hit bits 24..31 | imask 0..7): one entry per
// level, i.e. the stack never holds more entries than the tree is deep.

UBG_HD uint32_t byteOf(uint32_t lo, uint32_t hi, int i) { return ((i < 4 ? lo : hi) >> (8 * (i & 3))) & 0xFFu; }

// Box tests of one node's 8 children: returns pending inner-child bits
// (24 + (slot ^ oct)) | imask, and the leaf-primitive bits via `leafBits`.
template <bool kCount>
UBG_HD uint32_t intersectNode8(const uint4* nodes, uint32_t idx, V3 iv, V3 o, float tnear, float tmax, uint32_t oct,
                               uint32_t& childBase, uint32_t& primBase, uint32_t& leafBits) {
  const uint4* n = nodes + 5 * static_cast<size_t>(idx);
  const uint4 w0 = n[0];
  const uint4 w1 = n[1];
  const uint4 w2 = n[2];
  const uint4 w3 = n[3];
  const uint4 w4 = n[4];
  const float sx = u2f((w0.w & 0xFFu) << 23) * iv.x;
  const float sy = u2f(((w0.w >> 8) & 0xFFu) << 23) * iv.y;
  const float sz = u2f(((w0.w >> 16) & 0xFFu) << 23) * iv.z;
  const float ox = (u2f(w0.x) - o.x) * iv.x;
  const float oy = (u2f(w0.y) - o.y) * iv.y;
  const float oz = (u2f(w0.z) - o.z) * iv.z;
  const uint32_t imask = w0.w >> 24;
  const float tmx = tmax * 1.00000012f;
  // Per axis, the plane the ray enters through is the low one for a positive
  // direction and the high one for a negative direction: pick the byte arrays
  // once per node instead of min/max per child.
  const bool px = iv.x >= 0.0f, py = iv.y >= 0.0f, pz = iv.z >= 0.0f;
  const uint32_t nx0 = px ? w2.x : w3.z, nx1 = px ? w2.y : w3.w, fx0 = px ? w3.z : w2.x, fx1 = px ? w3.w : w2.y;
  const uint32_t ny0 = py ? w2.z : w4.x, ny1 = py ? w2.w : w4.y, fy0 = py ? w4.x : w2.z, fy1 = py ? w4.y : w2.w;
  const uint32_t nz0 = pz ? w3.x : w4.z, nz1 = pz ? w3.y : w4.w, fz0 = pz ? w4.z : w3.x, fz1 = pz ? w4.w : w3.y;
  // Branch-free box tests collect the slots whose box is hit; they are then
  // split into inner children (pending bits, reordered by the ray octant) and
  // leaf primitives. An empty slot needs no test of its own: its meta is 0,
  // so it adds no primitive bits, and imask has no bit for it.
  uint32_t hitSlots = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const float nxv = __builtin_fmaf(static_cast<float>(byteOf(nx0, nx1, i)), sx, ox);
    const float fxv = __builtin_fmaf(static_cast<float>(byteOf(fx0, fx1, i)), sx, ox);
    const float nyv = __builtin_fmaf(static_cast<float>(byteOf(ny0, ny1, i)), sy, oy);
    const float fyv = __builtin_fmaf(static_cast<float>(byteOf(fy0, fy1, i)), sy, oy);
    const float nzv = __builtin_fmaf(static_cast<float>(byteOf(nz0, nz1, i)), sz, oz);
    const float fzv = __builtin_fmaf(static_cast<float>(byteOf(fz0, fz1, i)), sz, oz);
    const float en = fmaxf(fmaxf(nxv, nyv), fmaxf(nzv, tnear));
    const float ex = fminf(fminf(fxv, fyv), fminf(fzv, tmx));
    hitSlots |= en <= ex ? (1u << i) : 0u;
  }
  // leaf slots hit: meta = (unary count) << 5 | offset of their primitives
  uint32_t leaf = 0;
  for (uint32_t ls = hitSlots & ~imask; ls != 0u; ls &= ls - 1u) {
    const uint32_t meta = byteOf(w1.z, w1.w, __builtin_ctz(ls));
    leaf |= (meta >> 5) << (meta & 31u);
  }
  uint32_t inner = hitSlots & imask;
  // slot i -> key i ^ oct, as three conditional swaps of bit groups
  if (oct & 1u) inner = ((inner & 0x55u) << 1) | ((inner >> 1) & 0x55u);
  if (oct & 2u) inner = ((inner & 0x33u) << 2) | ((inner >> 2) & 0x33u);
  if (oct & 4u) inner = ((inner & 0x0Fu) << 4) | ((inner >> 4) & 0x0Fu);
  childBase = w1.x;
  primBase = w1.y;
  leafBits = leaf;
  return (inner << 24) | imask;
}

template <bool kAny, bool kCount, class F = AcceptAll>
UBG_HD bool traverse8(const SceneView& s, V3 o, V3 d, float tnear, float tfar, Hit& best, uint32_t* stkBase,
                      uint32_t* stkHits, int stride, uint32_t* nodeVisits, uint32_t* primTests,
                      const F& filt = F()) {
  if (s.nNodes == 0) return false;
  const V3 iv = safeInv(d);
  const uint32_t oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
  const uint4* nodes = reinterpret_cast<const uint4*>(s.nodes);
  uint32_t gBase = 0, gHits = (1u << (24 + oct)) | 1u;  // the root as slot 0 of a virtual parent
  uint32_t tBase = 0, tHits = 0;
  int sp = 0;
  bool found = false;
  float tmax = tfar;
  Hit h;
  for (;;) {
    while (tHits != 0u) {
      const uint32_t b = static_cast<uint32_t>(__builtin_ctz(tHits));
      tHits &= tHits - 1u;
      if (kCount) ++*primTests;
      if (isectPrimF(s, tBase + b, o, d, tnear, tmax, h, filt)) {
        if (kAny) return true;
        if (!found || closer(h, best)) {
          best = h;
          tmax = h.t;
          found = true;
        }
      }
    }
    if ((gHits & 0xFF000000u) == 0u) {
      if (sp == 0) break;
      --sp;
      gBase = stkBase[sp * stride];
      gHits = stkHits[sp * stride];
    }
    const uint32_t key = static_cast<uint32_t>(__builtin_ctz(gHits >> 24));
    gHits &= ~(1u << (24 + key));
    const uint32_t slot = key ^ oct;
    const uint32_t idx = gBase + static_cast<uint32_t>(__builtin_popcount(gHits & 0xFFu & ((1u << slot) - 1u)));
    if (gHits & 0xFF000000u) {
      stkBase[sp * stride] = gBase;
      stkHits[sp * stride] = gHits;
      ++sp;
    }
    if (kCount) ++*nodeVisits;
    gHits = intersectNode8<kCount>(nodes, idx, iv, o, tnear, tmax, oct, gBase, tBase, tHits);
  }
  return found;
}

// Any-hit BVH8 traversal as a resumable state machine: one call does one unit
// of work (one primitive test or one node test), so a lane that finishes its
// ray can take a new one while the rest of its wave is still traversing.
struct Trav8 {
  V3 o, d, iv;
  float tnear, tmax;
  uint32_t oct, gBase, gHits, tBase, tHits;
  uint32_t kLo, kHi;  // kind bit planes of the pending primitives (trav8Node<.., true>)
  int sp;
};

UBG_HD void trav8Init(Trav8& t, V3 o, V3 d, float tnear, float tfar) {
  t.o = o;
  t.d = d;
  t.iv = safeInv(d);
  t.tnear = tnear;
  t.tmax = tfar;
  t.oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
  t.gBase = 0;
  t.gHits = (1u << (24 + t.oct)) | 1u;
  t.tBase = 0;
  t.tHits = 0;
  t.kLo = 0;
  t.kHi = 0;
  t.sp = 0;
}

// Node work is pending: inner children of the current group or a stacked group.
UBG_HD bool trav8HasNodes(const Trav8& t) { return (t.gHits & 0xFF000000u) != 0u || t.sp != 0; }

// One node test (requires no pending primitives and trav8HasNodes). Returns
// -1 when the traversal is exhausted, else 0. kKinds also fetches the kinds
// of the node's leaf primitives (for trav8PrimKind).
template <bool kCount, bool kKinds = false>
UBG_HD int trav8Node(const SceneView& s, Trav8& t, uint32_t* stkBase, uint32_t* stkHits, int stride,
                     uint32_t* nodeVisits) {
  if ((t.gHits & 0xFF000000u) == 0u) {
    --t.sp;
    t.gBase = stkBase[t.sp * stride];
    t.gHits = stkHits[t.sp * stride];
  }
  const uint32_t key = static_cast<uint32_t>(__builtin_ctz(t.gHits >> 24));
  t.gHits &= ~(1u << (24 + key));
  const uint32_t slot = key ^ t.oct;
  const uint32_t idx = t.gBase + static_cast<uint32_t>(__builtin_popcount(t.gHits & 0xFFu & ((1u << slot) - 1u)));
  if (t.gHits & 0xFF000000u) {
    stkBase[t.sp * stride] = t.gBase;
    stkHits[t.sp * stride] = t.gHits;
    ++t.sp;
  }
  if (kCount) ++*nodeVisits;
  t.gHits = intersectNode8<kCount>(reinterpret_cast<const uint4*>(s.nodes), idx, t.iv, t.o, t.tnear, t.tmax, t.oct,
                                   t.gBase, t.tBase, t.tHits);
  if (kKinds && t.tHits != 0u) {
    const uint32_t* kw = reinterpret_cast<const uint32_t*>(s.nodeKinds) + 2 * static_cast<size_t>(idx);
    t.kLo = kw[0];
    t.kHi = kw[1];
  }
  return (t.tHits == 0u && !trav8HasNodes(t)) ? -1 : 0;
}

// Pending primitives of kind `kind` (GeomKind numbering: mesh 0, sphere 1,
// round curve 2, cone curve 3).
UBG_HD uint32_t trav8KindMask(const Trav8& t, uint32_t kind) {
  const uint32_t lo = (kind & 1u) ? t.kLo : ~t.kLo;
  const uint32_t hi = (kind & 2u) ? t.kHi : ~t.kHi;
  return t.tHits & lo & hi;
}

// One any-hit test of a pending primitive of kind kKind (requires one; any
// order of primitive tests gives the same any-hit answer). Returns 1 when it
// occludes the ray, -1 when nothing is left to visit, else 0.
template <int kKind, bool kCount>
UBG_HD int trav8PrimKind(const SceneView& s, Trav8& t, uint32_t* primTests) {
  const uint32_t b = static_cast<uint32_t>(__builtin_ctz(trav8KindMask(t, kKind)));
  t.tHits &= ~(1u << b);
  if (kCount) ++*primTests;
  const float4* p = reinterpret_cast<const float4*>(s.prims) + 3 * static_cast<size_t>(t.tBase + b);
  const float4 p0 = p[0];
  Hit h;
  bool r;
  if (kKind == 1) {
    r = isectSphere(t.o, t.d, t.tnear, t.tmax, ld3(p0), p0.w, h);
  } else if (kKind == 3) {
    const float4 p1 = p[1], p2 = p[2];
    r = isectCone(t.o, t.d, t.tnear, t.tmax, ld3(p0), p0.w, ld3(p1), p1.w, p2.x, p2.y, p2.z, h);
  } else if (kKind == 0) {
    const float4 p1 = p[1], p2 = p[2];
    r = isectTri(t.o, t.d, t.tnear, t.tmax, ld3(p0), mk(p0.w, p1.x, p1.y), mk(p1.z, p1.w, p2.x), h);
  } else {
    const float4 p1 = p[1], p2 = p[2];
    const float4* pl = reinterpret_cast<const float4*>(s.rndPlanes) + 2 * (f2u(p2.w) & kIdMask);
    r = isectRound(t.o, t.d, t.tnear, t.tmax, ld3(p0), p0.w, ld3(p1), p1.w, pl[0], pl[1], h);
  }
  if (r) return 1;
  return (t.tHits == 0u && !trav8HasNodes(t)) ? -1 : 0;
}

// One primitive test (requires pending primitives). Returns 1 when it
// occludes the ray, -1 when nothing is left to visit, else 0.
template <bool kCount>
UBG_HD int trav8Prim(const SceneView& s, Trav8& t, uint32_t* primTests) {
  const uint32_t b = static_cast<uint32_t>(__builtin_ctz(t.tHits));
  t.tHits &= t.tHits - 1u;
  if (kCount) ++*primTests;
  Hit h;
  if (isectPrim(s, t.tBase + b, t.o, t.d, t.tnear, t.tmax, h)) return 1;
  return (t.tHits == 0u && !trav8HasNodes(t)) ? -1 : 0;
}

// One primitive test of a closest-hit traversal: a strictly nearer hit
// replaces `best` and shortens the ray. Returns -1 when nothing is left to
// visit, else 0.
template <bool kCount>
UBG_HD int trav8PrimClosest(const SceneView& s, Trav8& t, Hit& best, bool& found, uint32_t* primTests) {
  const uint32_t b = static_cast<uint32_t>(__builtin_ctz(t.tHits));
  t.tHits &= t.tHits - 1u;
  if (kCount) ++*primTests;
  Hit h;
  if (isectPrim(s, t.tBase + b, t.o, t.d, t.tnear, t.tmax, h)) {
    if (!found || closer(h, best)) {
      best = h;
      t.tmax = h.t;
      found = true;
    }
  }
  return (t.tHits == 0u && !trav8HasNodes(t)) ? -1 : 0;
}

// One unit of any-hit work: a primitive test if one is pending, else a node
// test. Returns 1 when a primitive occludes the ray, -1 when the traversal is
// exhausted without a hit, 0 to continue.
template <bool kCount>
UBG_HD int trav8StepAny(const SceneView& s, Trav8& t, uint32_t* stkBase, uint32_t* stkHits, int stride,
                        uint32_t* nodeVisits, uint32_t* primTests) {
  if (t.tHits != 0u) return trav8Prim<kCount>(s, t, primTests);
  return trav8Node<kCount>(s, t, stkBase, stkHits, stride, nodeVisits);
}

}  // namespace ubg
