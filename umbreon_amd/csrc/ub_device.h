// ub_device.h -- gfx950 device math, primitive intersectors and BVH
// traversal. Included by ub_kernels.hip only.
//
// The whole library is compiled with -ffp-contract=off and correctly rounded
// fp32 divide/sqrt, so every expression below rounds exactly as written. The
// intersectors and the shading chain use the same operation order as the
// reference's float code (scene.hpp vector helpers, shading/*.hpp), which is
// what makes first hits bit-identical to the CPU restatement. Only the
// traversal box test uses explicit FMAs: it merely decides which nodes are
// visited and every box is inflated by the builder, so it can never cull a
// primitive the exact test would hit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ub_layout.h"

namespace ubg {

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 mk(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len(V3 a) { return sqrtf(dot(a, a)); }
// scene.hpp:35-38
__device__ __forceinline__ V3 normalize(V3 a) {
  const float l = len(a);
  return l > 0.0f ? a * (1.0f / l) : a;
}
// scene.hpp:44-49
__device__ __forceinline__ V3 safeNormalize(V3 a) {
  const float s = 1.0f / sqrtf(fmaxf(dot(a, a), 0x1.0p-126f));
  return a * s;
}
// scene.hpp:53-57
__device__ __forceinline__ V3 safeNormalize(V3 a, V3 fb) {
  const float d = dot(a, a);
  if (d < 1.0e-18f) return fb;
  return a * (1.0f / sqrtf(d));
}
// shading.hpp:27-30
__device__ __forceinline__ V3 faceForward(V3 n, V3 rd) { return dot(n, rd) > 0.0f ? neg(n) : n; }

struct Frame3 {
  V3 t, b, n;
};
// scene.hpp:69-76
__device__ __forceinline__ Frame3 frameFromNormal(V3 n) {
  const V3 s0 = mk(0.0f, n.z, -n.y);
  const V3 s1 = mk(-n.z, 0.0f, n.x);
  const V3 t = safeNormalize(fabsf(n.x) < fabsf(n.y) ? s0 : s1, mk(1.0f, 0.0f, 0.0f));
  return Frame3{t, cross(n, t), n};
}
// secondary_rays.hpp:38-46
__device__ __forceinline__ float selfIntersectEps(V3 P, V3 d, float t) {
  const float dirMax = fmaxf(fabsf(d.x), fmaxf(fabsf(d.y), fabsf(d.z)));
  const float s = fmaxf(fmaxf(fabsf(P.x), fabsf(P.y)), fmaxf(fabsf(P.z), dirMax * t));
  return s * 0x1.0fp-21f;
}
// secondary_rays.hpp:51-58 (TEA-8)
__device__ __forceinline__ void tea2(uint32_t& v0, uint32_t& v1) {
  uint32_t sum = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    sum += 0x9E3779B9u;
    v0 += ((v1 << 4) + 0xA341316Cu) ^ (v1 + sum) ^ ((v1 >> 5) + 0xC8013EA4u);
    v1 += ((v0 << 4) + 0xAD90777Du) ^ (v0 + sum) ^ ((v0 >> 5) + 0x7E95761Eu);
  }
}
__device__ __forceinline__ float u32ToUnorm(uint32_t u) { return (float)(u >> 8) * 0x1.0p-24f; }
// secondary_rays.hpp:66-73 (Malley cosine hemisphere)
__device__ __forceinline__ V3 cosineSample(float u1, float u2, const Frame3& f) {
  const float phi = 6.2831853072f * u1;
  const float r = sqrtf(u2);
  const float z = sqrtf(fmaxf(0.0f, 1.0f - u2));
  const float x = cosf(phi) * r;
  const float y = sinf(phi) * r;
  return f.t * x + f.b * y + f.n * z;
}
// secondary_rays.hpp:102-109
__device__ __forceinline__ V3 sampleLightDir(V3 L, float tanR, float u1, float u2) {
  const Frame3 f = frameFromNormal(L);
  const float phi = 6.2831853072f * u1;
  const float rr = tanR * sqrtf(u2);
  const V3 d = f.n + f.t * (rr * cosf(phi)) + f.b * (rr * sinf(phi));
  return safeNormalize(d, L);
}

__device__ __forceinline__ V3 ld3(const float4& a) { return mk(a.x, a.y, a.z); }

struct Hit {
  float t, u, v;
  uint32_t bits;
  V3 ng;
};

// ---------------------------------------------------------------------------
// Intersectors. A hit is accepted when tnear <= t <= tfar. Spheres and curves
// solve in a frame moved along the ray to the closest approach of the
// primitive's center so far cameras do not cancel the discriminant.

__device__ __forceinline__ bool isectSphere(V3 o, V3 d, float tnear, float tfar, V3 c, float r,
                                            Hit& h) {
  const float A = dot(d, d);
  const float ts = dot(c - o, d) / A;
  const V3 f = (o + d * ts) - c;
  const float B = dot(f, d);
  const float C = dot(f, f) - r * r;
  const float disc = B * B - A * C;
  if (disc < 0.0f) return false;
  const float q = sqrtf(disc);
  const float t0 = ts + (-B - q) / A;
  const float t1 = ts + (-B + q) / A;
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

__device__ __forceinline__ int coneBody(V3 oa, V3 d, V3 w, float r0, float k, float dt[2]) {
  const float ho = dot(oa, w);
  const float hd = dot(d, w);
  const float rr = r0 + k * ho;
  const float A = dot(d, d) - hd * hd - k * k * hd * hd;
  const float B = dot(oa, d) - ho * hd - k * hd * rr;
  const float C = dot(oa, oa) - ho * ho - rr * rr;
  if (A == 0.0f) return 0;
  const float disc = B * B - A * C;
  if (disc < 0.0f) return 0;
  const float q = sqrtf(disc);
  dt[0] = (-B - q) / A;
  dt[1] = (-B + q) / A;
  return 2;
}

// CONE_LINEAR_CURVE: body between the end planes plus flat disk caps.
__device__ __forceinline__ bool isectCone(V3 o, V3 d, float tnear, float tfar, V3 a, float ra, V3 b,
                                          float rb, Hit& h) {
  const V3 ax = b - a;
  const float L = len(ax);
  if (!(L > 0.0f)) return false;
  const V3 w = ax * (1.0f / L);
  const float k = (rb - ra) / L;
  const float A = dot(d, d);
  const V3 mid = (a + b) * 0.5f;
  const float ts = dot(mid - o, d) / A;
  const V3 oa = (o + d * ts) - a;
  const float ho = dot(oa, w);
  const float hd = dot(d, w);
  bool found = false;
  float best = tfar, bu = 0.0f;
  V3 bn = mk(0.0f, 0.0f, 0.0f);
  float dt[2];
  if (coneBody(oa, d, w, ra, k, dt) == 2) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const float t = ts + dt[i];
      if (!(tnear <= t && t <= best)) continue;
      const float hh = ho + dt[i] * hd;
      if (hh < 0.0f || hh > L) continue;
      if (ra + k * hh < 0.0f) continue;
      const V3 p = oa + d * dt[i];
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
__device__ __noinline__ bool isectRound(V3 o, V3 d, float tnear, float tfar, V3 a, float ra, V3 b,
                                        float rb, float4 pl, float4 pr, Hit& h) {
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
  bool found = false;
  float best = tfar, bu = 0.0f;
  V3 bn = mk(0.0f, 0.0f, 0.0f);
  auto consider = [&](float dtv, V3 n, float hh) {
    const float t = ts + dtv;
    if (!(tnear <= t && t <= best)) return;
    const V3 p = oa + d * dtv;
    if (hasL && dot(p, nL) < 0.0f) return;
    if (hasR && dot(p - ax, nR) > 0.0f) return;
    best = t;
    float uu = hh / L;
    if (uu < 0.0f) uu = 0.0f;
    if (uu > 1.0f) uu = 1.0f;
    bu = uu;
    bn = n;
    found = true;
  };
  float dt[2];
  if (coneBody(oa, d, w, ra, k, dt) == 2) {
    for (int i = 0; i < 2; ++i) {
      const float hh = ho + dt[i] * hd;
      if (hh < 0.0f || hh > L) continue;
      const V3 p = oa + d * dt[i];
      consider(dt[i], p - w * hh, hh);
    }
  }
  for (int e = 0; e < 2; ++e) {
    const V3 c = e == 0 ? mk(0.0f, 0.0f, 0.0f) : ax;
    const float r = e == 0 ? ra : rb;
    const V3 f = oa - c;
    const float B = dot(f, d);
    const float C = dot(f, f) - r * r;
    const float disc = B * B - A * C;
    if (disc < 0.0f) continue;
    const float q = sqrtf(disc);
    const float roots[2] = {(-B - q) / A, (-B + q) / A};
    for (int i = 0; i < 2; ++i) {
      const float dtv = roots[i];
      const float hh = ho + dtv * hd;
      if (e == 0 && hh >= 0.0f) continue;
      if (e == 1 && hh <= L) continue;
      const V3 p = oa + d * dtv;
      consider(dtv, p - c, hh);
    }
  }
  if (!found) return false;
  h.t = best;
  h.u = bu;
  h.v = 0.0f;
  h.ng = bn;
  return true;
}

// TRIANGLE (Moller-Trumbore): P = (1-u-v) v0 + u v1 + v v2, Ng = (v1-v0)x(v2-v0).
__device__ __forceinline__ bool isectTri(V3 o, V3 d, float tnear, float tfar, V3 v0, V3 v1, V3 v2,
                                         Hit& h) {
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

__device__ __forceinline__ bool isectPrim(const SceneView& s, int i, V3 o, V3 d, float tnear,
                                          float tfar, Hit& h) {
  const float4* p = reinterpret_cast<const float4*>(s.prims) + 3 * i;
  const float4 p0 = p[0];
  const float4 p1 = p[1];
  const float4 p2 = p[2];
  const uint32_t bits = __float_as_uint(p2.w);
  const uint32_t kind = bits >> kKindShift;
  bool r;
  if (kind == 1u) {
    r = isectSphere(o, d, tnear, tfar, ld3(p0), p0.w, h);
  } else if (kind == 3u) {
    r = isectCone(o, d, tnear, tfar, ld3(p0), p0.w, ld3(p1), p1.w, h);
  } else if (kind == 0u) {
    r = isectTri(o, d, tnear, tfar, ld3(p0), mk(p0.w, p1.x, p1.y), mk(p1.z, p1.w, p2.x), h);
  } else {
    const float4* pl = reinterpret_cast<const float4*>(s.rndPlanes) + 2 * (bits & kIdMask);
    r = isectRound(o, d, tnear, tfar, ld3(p0), p0.w, ld3(p1), p1.w, pl[0], pl[1], h);
  }
  h.bits = bits;
  return r;
}

// ---------------------------------------------------------------------------
// BVH2 traversal with a per-lane stack in LDS (stack slot i of lane `lane`
// lives at stk[i * stride + lane], so the 64 lanes of a wave hit 64 distinct
// banks). kAny: return at the first hit in [tnear, tfar]. Otherwise the
// closest hit (strictly nearer hits replace the current best).
template <bool kAny, bool kCount>
__device__ __forceinline__ bool traverse(const SceneView& s, V3 o, V3 d, float tnear, float tfar,
                                         Hit& best, int* stk, int stride, uint32_t* nodeVisits,
                                         uint32_t* primTests) {
  if (s.nNodes == 0) return false;
  const float ix = 1.0f / (d.x != 0.0f ? d.x : copysignf(1e-30f, d.x));
  const float iy = 1.0f / (d.y != 0.0f ? d.y : copysignf(1e-30f, d.y));
  const float iz = 1.0f / (d.z != 0.0f ? d.z : copysignf(1e-30f, d.z));
  const float oix = o.x * ix, oiy = o.y * iy, oiz = o.z * iz;
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
    const int4 ch = *reinterpret_cast<const int4*>(n + 3);
    if (kCount) ++*nodeVisits;
    const float tmx = tmax * 1.00000012f;
    const float l0x = __builtin_fmaf(a.x, ix, -oix), h0x = __builtin_fmaf(a.y, ix, -oix);
    const float l0y = __builtin_fmaf(a.z, iy, -oiy), h0y = __builtin_fmaf(a.w, iy, -oiy);
    const float l0z = __builtin_fmaf(c.x, iz, -oiz), h0z = __builtin_fmaf(c.y, iz, -oiz);
    const float l1x = __builtin_fmaf(b.x, ix, -oix), h1x = __builtin_fmaf(b.y, ix, -oix);
    const float l1y = __builtin_fmaf(b.z, iy, -oiy), h1y = __builtin_fmaf(b.w, iy, -oiy);
    const float l1z = __builtin_fmaf(c.z, iz, -oiz), h1z = __builtin_fmaf(c.w, iz, -oiz);
    const float e0 = fmaxf(fmaxf(fminf(l0x, h0x), fminf(l0y, h0y)), fmaxf(fminf(l0z, h0z), tnear));
    const float x0 = fminf(fminf(fmaxf(l0x, h0x), fmaxf(l0y, h0y)), fminf(fmaxf(l0z, h0z), tmx));
    const float e1 = fmaxf(fmaxf(fminf(l1x, h1x), fminf(l1y, h1y)), fmaxf(fminf(l1z, h1z), tnear));
    const float x1 = fminf(fminf(fmaxf(l1x, h1x), fmaxf(l1y, h1y)), fminf(fmaxf(l1z, h1z), tmx));
    bool hit0 = e0 <= x0;
    bool hit1 = e1 <= x1;
    // Leaves are intersected on the spot, nearer one first.
    const bool swapOrder = e1 < e0;
    for (int k = 0; k < 2; ++k) {
      const bool first = (k == 0) != swapOrder;
      const int cidx = first ? ch.x : ch.y;
      const bool hc = first ? hit0 : hit1;
      if (!hc || cidx >= 0) continue;
      const uint32_t code = ~static_cast<uint32_t>(cidx);
      const int beg = static_cast<int>(code >> 3);
      const int cnt = static_cast<int>(code & 7u) + 1;
      for (int i = beg; i < beg + cnt; ++i) {
        if (kCount) ++*primTests;
        if (isectPrim(s, i, o, d, tnear, tmax, h)) {
          if (kAny) return true;
          if (!found || h.t < best.t) {
            best = h;
            tmax = h.t;
            found = true;
          }
        }
      }
      if (first) hit0 = false;
      else hit1 = false;
    }
    if (hit0 && hit1) {
      const bool near1 = e1 < e0;
      stk[sp * stride] = near1 ? ch.x : ch.y;
      ++sp;
      node = near1 ? ch.y : ch.x;
    } else if (hit0) {
      node = ch.x;
    } else if (hit1) {
      node = ch.y;
    } else {
      if (sp == 0) break;
      --sp;
      node = stk[sp * stride];
    }
  }
  return found;
}

}  // namespace ubg
