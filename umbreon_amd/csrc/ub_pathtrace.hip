// ub_pathtrace.hip -- the physically based path tracer of ubg_pathtrace:
// GGX conductors and dielectrics (Walter et al. 2007 microfacet BSDF with
// visible-normal sampling, Heitz 2018), Lambertian surfaces, spherical area
// lights sampled by resampled importance sampling with ReSTIR spatial reuse
// at the first vertex (Bitterli et al. 2020; unbiased, with balance-heuristic
// resampling weights over the combined reservoirs),
// a constant environment, a fixed number of path vertices.
//
// No reference counterpart: the reference's integrators are diffuse-only
// (pt1 / pt2); docs/api/libumbreon.md:33 puts refraction out of its scope.
//
// One sample of every pixel per pass: k_pt_primary traces the camera rays,
// stores the first-hit G-buffer and draws M light candidates per pixel into
// a reservoir; k_pt_shade combines each pixel's reservoir with K neighbours'
// (each candidate's target re-evaluated at every contributing pixel for its
// resampling weight), traces the shadow ray, then continues the path through the
// remaining vertices with per-vertex RIS next-event estimation; the
// radiance adds into a per-pixel accumulator in sample order. At glossy
// vertices the lights are also reached by the BSDF rays, combined with the
// next-event estimates by multiple importance sampling (balance heuristic);
// camera rays do not see the lights, and rays that leave the scene return
// the environment radiance.
//
// Random numbers are a pure function of (seed, pixel, sample, dimension)
// (hash32 below), so an image does not depend on launch shapes.

#include <hip/hip_runtime.h>

#include "ub_device.h"
#include "ub_pathtrace.h"

namespace ubg {
namespace pt {

extern __shared__ uint32_t ptLds[];

constexpr float kPi = 3.14159265358979f;
constexpr int kBlock = 128;
// Waves per SIMD the sample-pass kernels are compiled for (their register
// budget: 512 / waves VGPRs). Traversal is latency-bound, so occupancy pays
// for itself up to the point where the budget forces spills.
#ifndef UBG_PT_PRIMARY_WAVES
#define UBG_PT_PRIMARY_WAVES 4
#endif
#ifndef UBG_PT_SHADE_WAVES
#define UBG_PT_SHADE_WAVES 3
#endif
constexpr int kPrimaryWaves = UBG_PT_PRIMARY_WAVES;
constexpr int kShadeWaves = UBG_PT_SHADE_WAVES;

// ---------------------------------------------------------------- random
__host__ __device__ inline uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7FEB352Du;
  x ^= x >> 15;
  x *= 0x846CA68Bu;
  x ^= x >> 16;
  return x;
}
__device__ __forceinline__ float rnd(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim) {
  const uint32_t h = hash32(seed ^ hash32(pixel ^ hash32((sample * 0x9E3779B9u) ^ hash32(dim + 0x632BE5ABu))));
  return static_cast<float>(h >> 8) * 0x1.0p-24f;
}

// ---------------------------------------------------------------- frames
struct Onb {
  V3 t, b, n;
};
// Orthonormal basis around a unit n (Duff et al. 2017, "Building an
// Orthonormal Basis, Revisited").
__device__ __forceinline__ Onb onb(V3 n) {
  const float sgn = n.z >= 0.0f ? 1.0f : -1.0f;
  const float a = -1.0f / (sgn + n.z);
  const float b = n.x * n.y * a;
  return Onb{mk(1.0f + sgn * n.x * n.x * a, sgn * b, -sgn * n.x), mk(b, sgn + n.y * n.y * a, -n.y), n};
}
__device__ __forceinline__ V3 toLocal(const Onb& f, V3 v) { return mk(dot(v, f.t), dot(v, f.b), dot(v, f.n)); }
__device__ __forceinline__ V3 toWorld(const Onb& f, V3 v) {
  return mk(f.t.x * v.x + f.b.x * v.y + f.n.x * v.z, f.t.y * v.x + f.b.y * v.y + f.n.y * v.z,
            f.t.z * v.x + f.b.z * v.y + f.n.z * v.z);
}

// ---------------------------------------------------------------- GGX (local frame, n = +z)
__device__ __forceinline__ float ggxD(V3 m, float a) {
  const float c2 = m.z * m.z;
  if (c2 <= 0.0f) return 0.0f;
  const float a2 = a * a;
  const float k = c2 * (a2 - 1.0f) + 1.0f;
  return a2 / (kPi * k * k);
}
__device__ __forceinline__ float ggxLambda(V3 w, float a) {
  const float c2 = w.z * w.z;
  if (c2 <= 0.0f) return 0.0f;
  const float t2 = fmaxf(0.0f, 1.0f - c2) / c2;
  return 0.5f * (sqrtf(1.0f + a * a * t2) - 1.0f);
}
__device__ __forceinline__ float ggxG1(V3 w, float a) { return 1.0f / (1.0f + ggxLambda(w, a)); }
__device__ __forceinline__ float ggxG2(V3 wo, V3 wi, float a) { return 1.0f / (1.0f + ggxLambda(wo, a) + ggxLambda(wi, a)); }
// Visible normal of wo (wo.z > 0) from two uniforms (Heitz 2018).
__device__ __forceinline__ V3 ggxSampleVisible(V3 wo, float a, float u1, float u2) {
  const V3 vh = normalize(mk(a * wo.x, a * wo.y, wo.z));
  const float l2 = vh.x * vh.x + vh.y * vh.y;
  const V3 T1 = l2 > 0.0f ? mk(-vh.y, vh.x, 0.0f) * (1.0f / sqrtf(l2)) : mk(1.0f, 0.0f, 0.0f);
  const V3 T2 = cross(vh, T1);
  const float r = sqrtf(u1);
  const float phi = 2.0f * kPi * u2;
  const float t1 = r * cosf(phi);
  float t2 = r * sinf(phi);
  const float s = 0.5f * (1.0f + vh.z);
  t2 = (1.0f - s) * sqrtf(fmaxf(0.0f, 1.0f - t1 * t1)) + s * t2;
  const float tz = sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2));
  const V3 nh = mk(T1.x * t1 + T2.x * t2 + vh.x * tz, T1.y * t1 + T2.y * t2 + vh.y * tz,
                   T1.z * t1 + T2.z * t2 + vh.z * tz);
  return normalize(mk(a * nh.x, a * nh.y, fmaxf(1.0e-6f, nh.z)));
}
// Density of that visible normal: G1(wo) max(0, wo.m) D(m) / wo.z.
__device__ __forceinline__ float ggxVisiblePdf(V3 wo, V3 m, float a) {
  return ggxG1(wo, a) * fmaxf(0.0f, dot(wo, m)) * ggxD(m, a) / wo.z;
}
// Unpolarized Fresnel reflectance of a dielectric; cosI = cos of the
// incident angle to the normal on the incident side, eta = n_t / n_i.
__device__ __forceinline__ float fresnelDielectric(float cosI, float eta) {
  const float s2 = fmaxf(0.0f, 1.0f - cosI * cosI) / (eta * eta);
  if (s2 >= 1.0f) return 1.0f;
  const float cosT = sqrtf(1.0f - s2);
  const float rs = (cosI - eta * cosT) / (cosI + eta * cosT);
  const float rp = (eta * cosI - cosT) / (eta * cosI + cosT);
  return 0.5f * (rs * rs + rp * rp);
}
__device__ __forceinline__ V3 schlick(V3 f0, float c) {
  const float m = fminf(1.0f, fmaxf(0.0f, 1.0f - c));
  const float m5 = (m * m) * (m * m) * m;
  return mk(f0.x + (1.0f - f0.x) * m5, f0.y + (1.0f - f0.y) * m5, f0.z + (1.0f - f0.z) * m5);
}

// The surface seen at a vertex: local frame around the outward normal n
// (geometric for primitives, interpolated for meshes), base colour and
// material.
struct Surf {
  Onb f;
  V3 C;
  int type;
  float alpha;   // GGX alpha (roughness squared)
  float ior;
};
constexpr float kSmoothAlpha = 4.0e-4f;  // below: a smooth (delta) interface

// f(wo, wi) * |cos wi| in the local frame. Conductors and diffuse surfaces
// are one-sided (both directions on the outward side); dielectrics reflect
// and transmit by the microfacet model (no delta lobe is evaluated).
__device__ __forceinline__ V3 bsdfCos(const Surf& s, V3 wo, V3 wi) {
  if (s.type == kMatDiffuse) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return mk(0.0f, 0.0f, 0.0f);
    return s.C * (wi.z / kPi);
  }
  if (s.alpha < kSmoothAlpha) return mk(0.0f, 0.0f, 0.0f);
  if (s.type == kMatConductor) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return mk(0.0f, 0.0f, 0.0f);
    const V3 m = normalize(wo + wi);
    const V3 F = schlick(s.C, dot(wi, m));
    const float k = ggxD(m, s.alpha) * ggxG2(wo, wi, s.alpha) / (4.0f * wo.z);
    return F * k;
  }
  // dielectric: eta = n_inside / n_outside seen from the side of wo
  const float cosO = wo.z, cosI = wi.z;
  if (cosO == 0.0f || cosI == 0.0f) return mk(0.0f, 0.0f, 0.0f);
  const bool refl = cosO * cosI > 0.0f;
  const float eta = cosO > 0.0f ? s.ior : 1.0f / s.ior;  // n_t / n_i for light leaving along wo
  V3 m = refl ? (wo + wi) : (wi * eta + wo);
  if (dot(m, m) == 0.0f) return mk(0.0f, 0.0f, 0.0f);
  m = normalize(m);
  if (m.z < 0.0f) m = neg(m);
  if (dot(m, wi) * cosI < 0.0f || dot(m, wo) * cosO < 0.0f) return mk(0.0f, 0.0f, 0.0f);
  const float F = fresnelDielectric(fabsf(dot(wo, m)), cosO > 0.0f ? s.ior : 1.0f / s.ior);
  const float D = ggxD(m, s.alpha);
  const float G = ggxG2(mk(wo.x, wo.y, fabsf(wo.z)), mk(wi.x, wi.y, fabsf(wi.z)), s.alpha);
  if (refl) {
    const float k = D * G * F / (4.0f * fabsf(cosO));
    return mk(k, k, k);
  }
  const float dn = dot(wi, m) + dot(wo, m) / eta;
  const float k = D * (1.0f - F) * G * fabsf(dot(wi, m) * dot(wo, m) / (dn * dn * cosO)) / (eta * eta);
  return s.C * k;
}

// Samples wi for wo: returns false when the path ends; weight = f |cos| / pdf.
__device__ __forceinline__ bool bsdfSample(const Surf& s, V3 wo, float u1, float u2, float u3, V3& wi, V3& weight) {
  if (s.type == kMatDiffuse) {
    if (wo.z <= 0.0f) return false;
    const float r = sqrtf(u1), phi = 2.0f * kPi * u2;
    wi = mk(r * cosf(phi), r * sinf(phi), sqrtf(fmaxf(0.0f, 1.0f - u1)));
    weight = s.C;
    return true;
  }
  if (s.type == kMatConductor) {
    if (wo.z <= 0.0f) return false;
    V3 m;
    if (s.alpha < kSmoothAlpha) {
      m = mk(0.0f, 0.0f, 1.0f);
    } else {
      m = ggxSampleVisible(wo, s.alpha, u1, u2);
    }
    const float om = dot(wo, m);
    wi = m * (2.0f * om) - wo;
    if (wi.z <= 0.0f) return false;
    const V3 F = schlick(s.C, dot(wi, m));
    weight = s.alpha < kSmoothAlpha ? F : F * (ggxG2(wo, wi, s.alpha) / ggxG1(wo, s.alpha));
    return true;
  }
  // dielectric, in the frame flipped to the side of wo
  const bool outside = wo.z > 0.0f;
  const V3 w = outside ? wo : mk(wo.x, wo.y, -wo.z);
  if (w.z == 0.0f) return false;
  const float eta = outside ? s.ior : 1.0f / s.ior;  // n_t / n_i
  const V3 m = s.alpha < kSmoothAlpha ? mk(0.0f, 0.0f, 1.0f) : ggxSampleVisible(w, s.alpha, u1, u2);
  const float c = dot(w, m);
  const float F = fresnelDielectric(c, eta);
  V3 wl;
  bool refracted = false;
  if (u3 < F) {
    wl = m * (2.0f * c) - w;
    if (wl.z <= 0.0f) return false;
  } else {
    const float s2 = fmaxf(0.0f, 1.0f - c * c) / (eta * eta);
    if (s2 >= 1.0f) return false;  // (F is 1 under total internal reflection)
    const float ct = sqrtf(1.0f - s2);
    wl = m * (c / eta - ct) - w * (1.0f / eta);
    if (wl.z >= 0.0f) return false;
    refracted = true;
  }
  const float g = s.alpha < kSmoothAlpha ? 1.0f
                                         : ggxG2(w, mk(wl.x, wl.y, fabsf(wl.z)), s.alpha) / ggxG1(w, s.alpha);
  wi = outside ? wl : mk(wl.x, wl.y, -wl.z);
  // radiance crossing the interface scales by (n_i / n_t)^2 = 1 / eta^2
  weight = refracted ? s.C * (g / (eta * eta)) : mk(g, g, g);
  return true;
}

// Density (solid angle, local frame) with which bsdfSample draws wi for wo;
// 0 where it never does. Smooth (delta) lobes are not evaluated here: a
// light reached through one takes the whole weight (see misBsdf).
__device__ __forceinline__ float bsdfPdf(const Surf& s, V3 wo, V3 wi) {
  if (s.type == kMatDiffuse) return (wo.z > 0.0f && wi.z > 0.0f) ? wi.z / kPi : 0.0f;
  if (s.type == kMatConductor) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
    const V3 m = normalize(wo + wi);
    return ggxG1(wo, s.alpha) * ggxD(m, s.alpha) / (4.0f * wo.z);
  }
  // dielectric, in the frame flipped to the side of wo (as bsdfSample)
  const bool outside = wo.z > 0.0f;
  const V3 w = outside ? wo : mk(wo.x, wo.y, -wo.z);
  const V3 wl = outside ? wi : mk(wi.x, wi.y, -wi.z);
  if (w.z == 0.0f || wl.z == 0.0f) return 0.0f;
  const float eta = outside ? s.ior : 1.0f / s.ior;  // n_t / n_i
  const float g1 = ggxG1(w, s.alpha);
  if (wl.z > 0.0f) {  // reflection: the visible-normal density / (4 w.m), times F
    const V3 m = normalize(w + wl);
    return fresnelDielectric(dot(w, m), eta) * g1 * ggxD(m, s.alpha) / (4.0f * w.z);
  }
  // refraction: m along -(w + eta wl); Jacobian |wl.m| / (wl.m + w.m / eta)^2
  V3 m = w + wl * eta;
  if (dot(m, m) == 0.0f) return 0.0f;
  m = normalize(m);
  if (m.z < 0.0f) m = neg(m);
  const float c = dot(w, m), cl = dot(wl, m);
  if (c <= 0.0f || cl >= 0.0f) return 0.0f;
  const float dn = cl + c / eta;
  return (1.0f - fresnelDielectric(c, eta)) * g1 * c * ggxD(m, s.alpha) / w.z * fabsf(cl) / (dn * dn);
}

// ---------------------------------------------------------------- lights
// Uniform direction in the cone the light subtends from P; returns the hit
// point on the sphere and the area-measure density (light choice included).
__device__ __forceinline__ bool lightSample(const Light& L, V3 P, float u1, float u2, int nLights, V3& Q, float& pdfA) {
  const V3 c = mk(L.c[0], L.c[1], L.c[2]);
  const V3 dcv = c - P;
  const float d2 = dot(dcv, dcv);
  const float r2 = L.r * L.r;
  if (d2 <= r2) return false;
  const float d = sqrtf(d2);
  const V3 axis = dcv * (1.0f / d);
  const float cosMax = sqrtf(fmaxf(0.0f, 1.0f - r2 / d2));
  const float cosT = 1.0f - u1 * (1.0f - cosMax);
  const float sinT = sqrtf(fmaxf(0.0f, 1.0f - cosT * cosT));
  const float phi = 2.0f * kPi * u2;
  const Onb f = onb(axis);
  const V3 dir = toWorld(f, mk(sinT * cosf(phi), sinT * sinf(phi), cosT));
  // nearest intersection of P + t dir with the sphere
  const float b = dot(dcv, dir);
  const float disc = b * b - (d2 - r2);
  const float t = b - sqrtf(fmaxf(0.0f, disc));
  Q = P + dir * t;
  const V3 nl = (Q - c) * (1.0f / L.r);
  const float cosL = fabsf(dot(nl, dir));
  const float pdfW = 1.0f / (2.0f * kPi * (1.0f - cosMax));
  pdfA = pdfW * cosL / (t * t) / static_cast<float>(nLights);
  return pdfA > 0.0f;
}
// Unshadowed contribution of light point Q (on light li) at the surface:
// Multiple importance sampling of the lights (balance heuristic, Veach
// 1997) between next-event estimation and BSDF rays, at glossy vertices
// (conductors and dielectrics; a diffuse vertex keeps next-event estimation
// alone, a smooth one BSDF rays alone, since next-event estimation cannot
// reach a delta lobe). Lights are mutually transparent emitters that camera
// rays do not see; a BSDF ray collects every light sphere whose near side it
// crosses before the scene hit. The next-event density of a direction
// towards light l is that of its candidates -- a uniform light choice
// (1 / n, or 1 with neeAll) times the uniform cone density 1 / (2 pi (1 -
// cos theta_max)) -- whatever RIS / ReSTIR then selects: the weights sum to
// one for every path, so the estimate stays unbiased.
__device__ __forceinline__ bool misVertex(const Surf& s) { return s.type != kMatDiffuse; }
__device__ __forceinline__ float neeDensity(const Params& p, const Light& L, V3 P) {
  const V3 dcv = mk(L.c[0], L.c[1], L.c[2]) - P;
  const float d2 = dot(dcv, dcv), r2 = L.r * L.r;
  if (d2 <= r2) return 0.0f;
  const float cosMax = sqrtf(fmaxf(0.0f, 1.0f - r2 / d2));
  const float sel = p.neeAll ? 1.0f : 1.0f / static_cast<float>(p.nLights);
  return sel / (2.0f * kPi * (1.0f - cosMax));
}
// Weight of a next-event sample towards light li in direction wi (world).
__device__ __forceinline__ float misNee(const Params& p, const Surf& sf, V3 P, V3 wo, V3 wi, int li) {
  if (!misVertex(sf)) return 1.0f;
  const float pb = sf.alpha < kSmoothAlpha ? 0.0f : bsdfPdf(sf, toLocal(sf.f, wo), toLocal(sf.f, wi));
  const float pn = neeDensity(p, p.lights[li], P);
  return pn > 0.0f ? pn / (pn + pb) : 0.0f;
}
// Light a BSDF ray from vertex P in direction wi (world) collects before the
// scene hit at tHit, with its weights; pb = the ray's BSDF density (+inf for
// a delta lobe).
// The lights are tested in groups of kLightGroup behind their bounding sphere:
// a ray that misses the group's sphere misses each of its lights, which
// would add nothing, so the sum (over the lights hit, in index order) is the
// one of testing every light.
__device__ __forceinline__ V3 misBsdf(const Params& p, V3 P, V3 wi, float tHit, float pb) {
  V3 sum = mk(0.0f, 0.0f, 0.0f);
  for (int g0 = 0; g0 < p.nLights; g0 += kLightGroup) {
    const float4 G = p.lightGroups[g0 / kLightGroup];
    const V3 dg = mk(G.x, G.y, G.z) - P;
    const float bg = dot(dg, wi), dg2 = dot(dg, dg), rg2 = G.w * G.w;
    // the forward ray meets the group's sphere: origin inside, or ahead and
    // within the radius
    if (dg2 > rg2 && (bg <= 0.0f || bg * bg - (dg2 - rg2) < 0.0f)) continue;
    const int g1 = min(p.nLights, g0 + kLightGroup);
    for (int li = g0; li < g1; ++li) {
      const Light& L = p.lights[li];
      const V3 dcv = mk(L.c[0], L.c[1], L.c[2]) - P;
      const float b = dot(dcv, wi);
      if (b <= 0.0f) continue;
      const float d2 = dot(dcv, dcv), r2 = L.r * L.r;
      const float disc = b * b - (d2 - r2);
      if (disc < 0.0f || d2 <= r2) continue;
      const float t = b - sqrtf(disc);
      if (!(t < tHit)) continue;
      const float pn = neeDensity(p, L, P);
      const float w = pb == INFINITY ? 1.0f : pb / (pb + pn);
      sum = sum + mk(L.Le[0] * w, L.Le[1] * w, L.Le[2] * w);
    }
  }
  return sum;
}

// f |cos| Le |cos_l| / dist^2 (area measure); target = its luminance.
__device__ __forceinline__ V3 lightContrib(const Params& p, const Surf& s, V3 P, V3 wo, V3 Q, int li, float& dist) {
  const Light& L = p.lights[li];
  const V3 dv = Q - P;
  const float d2 = dot(dv, dv);
  dist = sqrtf(d2);
  if (!(d2 > 0.0f)) return mk(0.0f, 0.0f, 0.0f);
  const V3 wi = dv * (1.0f / dist);
  const V3 nl = (Q - mk(L.c[0], L.c[1], L.c[2])) * (1.0f / L.r);
  const float cosL = -dot(nl, wi);
  if (cosL <= 0.0f) return mk(0.0f, 0.0f, 0.0f);
  const V3 fc = bsdfCos(s, toLocal(s.f, wo), toLocal(s.f, wi));
  const float g = cosL / d2;
  return mk(fc.x * L.Le[0] * g, fc.y * L.Le[1] * g, fc.z * L.Le[2] * g);
}
__device__ __forceinline__ float lum(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// Adds the wave's traced rays to the frame counter (one atomic per wave).
__device__ __forceinline__ void countRays(unsigned long long* c, uint32_t n) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
  if (__lane_id() == 0 && n) atomicAdd(c, static_cast<unsigned long long>(n));
}

// ---------------------------------------------------------------- scene access
__device__ __forceinline__ bool traceClosest(const SceneView& s, V3 o, V3 d, float tf, Hit& h) {
  uint32_t* base = ptLds + threadIdx.x;
  return traverse8<false, false>(s, o, d, 0.0f, tf, h, base, base + s.stackDepth * blockDim.x, blockDim.x, nullptr,
                                 nullptr);
}
__device__ __forceinline__ bool traceAny(const SceneView& s, V3 o, V3 d, float tf) {
  uint32_t* base = ptLds + threadIdx.x;
  Hit h;
  return traverse8<true, false>(s, o, d, 0.0f, tf, h, base, base + s.stackDepth * blockDim.x, blockDim.x, nullptr,
                                nullptr);
}

// Surface at a hit: outward normal, base colour, material.
__device__ __forceinline__ Surf surfaceAt(const Params& p, const SceneView& s, const Hit& h, V3 rd, V3& Ng) {
  const uint32_t kind = h.bits >> kKindShift, id = h.bits & kIdMask;
  V3 N, C;
  uint32_t mat;
  bool mesh = false;
  if (kind == 0u) {
    const uint4 tri = reinterpret_cast<const uint4*>(s.triIdx)[id];
    const float4* vn = reinterpret_cast<const float4*>(s.vtxNrm);
    const float4* vc = reinterpret_cast<const float4*>(s.vtxCol);
    const float4 n0 = vn[tri.x], n1 = vn[tri.y], n2 = vn[tri.z];
    const float4 c0 = vc[tri.x], c1 = vc[tri.y], c2 = vc[tri.z];
    const float w0 = 1.0f - h.u - h.v;
    N = normalize(mk(w0 * n0.x + h.u * n1.x + h.v * n2.x, w0 * n0.y + h.u * n1.y + h.v * n2.y,
                     w0 * n0.z + h.u * n1.z + h.v * n2.z));
    C = mk(w0 * c0.x + h.u * c1.x + h.v * c2.x, w0 * c0.y + h.u * c1.y + h.v * c2.y,
           w0 * c0.z + h.u * c1.z + h.v * c2.z);
    Ng = normalize(h.ng);
    if (dot(Ng, N) < 0.0f) Ng = neg(Ng);  // the mesh's outward side is its normals'
    mat = tri.w & kInfoMatMask;
    mesh = true;
  } else {
    float4 col;
    uint32_t info;
    if (kind == 1u) {
      col = reinterpret_cast<const float4*>(s.sphColor)[id];
      info = s.sphInfo[id];
    } else if (kind == 3u) {
      col = reinterpret_cast<const float4*>(s.coneColor)[id];
      info = s.coneInfo[id];
    } else {
      col = reinterpret_cast<const float4*>(s.rndColor)[id];
      info = s.rndInfo[id];
    }
    N = normalize(h.ng);  // primitive normals point outward
    Ng = N;
    C = mk(col.x, col.y, col.z);
    mat = info & kInfoMatMask;
  }
  // the mesh table follows the primitive table (Params::meshMats ==
  // primMats + nPrimMats): an offset, not a choice between two pointers,
  // which the compiler would keep in private memory
  const Mat* tab = p.primMats + (mesh ? p.nPrimMats : 0);
  const int n = mesh ? p.nMeshMats : p.nPrimMats;
  const Mat m = n > 0 ? tab[mat < static_cast<uint32_t>(n) ? mat : 0u] : Mat{kMatDiffuse, 0.0f, 1.5f, 0};
  Surf sf;
  // opaque surfaces shade the side the ray arrives on
  if (m.type != kMatDielectric && dot(N, rd) > 0.0f) {
    N = neg(N);
    Ng = neg(Ng);
  }
  sf.f = onb(N);
  sf.C = C;
  sf.type = m.type;
  sf.alpha = m.alpha;
  sf.ior = m.ior;
  return sf;
}

// Offset origin for a ray leaving the surface at P along d.
__device__ __forceinline__ V3 offsetOrigin(V3 P, V3 Ng, V3 d, float t) {
  const float e = fmaxf(fmaxf(fabsf(P.x), fabsf(P.y)), fmaxf(fabsf(P.z), t)) * 0x1.0p-19f + 1.0e-6f;
  return P + Ng * (dot(d, Ng) >= 0.0f ? e : -e);
}

__device__ __forceinline__ void cameraRay(const Params& p, float fx, float fy, V3& o, V3& d) {
  const float sx = 2.0f * fx / static_cast<float>(p.W) - 1.0f;
  const float sy = 1.0f - 2.0f * fy / static_cast<float>(p.H);
  const V3 dir = mk(p.dir[0], p.dir[1], p.dir[2]);
  const V3 right = mk(p.right[0], p.right[1], p.right[2]);
  const V3 up = mk(p.up[0], p.up[1], p.up[2]);
  const V3 pos = mk(p.pos[0], p.pos[1], p.pos[2]);
  if (p.ortho) {
    o = pos + right * (sx * p.halfW) + up * (sy * p.halfH);
    d = dir;
  } else {
    o = pos;
    d = normalize(dir + right * (sx * p.persHalfW) + up * (sy * p.persHalfH));
  }
}

// Dimension layout of one sample's random numbers: camera jitter 0..1;
// vertex b: RIS candidates at 4c .. 4c + 3 (c < 12), the BSDF sample at
// 56 .. 58; the spatial reuse's neighbour offsets and selections; the
// all-lights estimator's cone samples, 2 per light.
constexpr uint32_t kDimCam = 0;
__device__ __forceinline__ uint32_t dimVertex(int b) { return 16u + 64u * static_cast<uint32_t>(b); }
constexpr uint32_t kDimSpatial = 0x10000u;
__device__ __forceinline__ uint32_t dimAllLights(int b) { return 0x100000u + 0x2000u * static_cast<uint32_t>(b); }

// RIS over M light candidates at a surface: the selected point, its light,
// the reservoir weight sum and the selected target.
struct Pick {
  V3 Q;
  int li;
  float wsum, target;
};
__device__ __forceinline__ Pick risLights(const Params& p, const Surf& sf, V3 P, V3 wo, uint32_t pix, uint32_t smp,
                                          uint32_t dim) {
  Pick k{mk(0.0f, 0.0f, 0.0f), -1, 0.0f, 0.0f};
  for (int c = 0; c < p.restirM; ++c) {
    const uint32_t d = dim + 4u * static_cast<uint32_t>(c);
    const int li = min(p.nLights - 1, static_cast<int>(rnd(p.seed, pix, smp, d) * static_cast<float>(p.nLights)));
    V3 Q;
    float pdfA;
    if (!lightSample(p.lights[li], P, rnd(p.seed, pix, smp, d + 1), rnd(p.seed, pix, smp, d + 2), p.nLights, Q,
                     pdfA))
      continue;
    float dist;
    const float t = lum(lightContrib(p, sf, P, wo, Q, li, dist));
    const float w = t / pdfA;
    if (!(w > 0.0f)) continue;
    k.wsum += w;
    if (rnd(p.seed, pix, smp, d + 3) * k.wsum < w) {
      k.Q = Q;
      k.li = li;
      k.target = t;
    }
  }
  return k;
}

// Direct light at a vertex by RIS NEE (or, with neeAll, one cone sample of
// every light), shadow rays included; tIn = the distance along the ray that
// found the vertex (it scales the shadow rays' origin offset, like every
// other ray's).
__device__ __forceinline__ V3 directLight(const Params& p, const SceneView& s, const Surf& sf, V3 P, V3 Ng, V3 wo,
                                          float tIn, uint32_t pix, uint32_t smp, int b, uint32_t& nr) {
  V3 sum = mk(0.0f, 0.0f, 0.0f);
  if (p.nLights <= 0) return sum;
  if (p.neeAll) {
    for (int li = 0; li < p.nLights; ++li) {
      const uint32_t d = dimAllLights(b) + 2u * static_cast<uint32_t>(li);
      V3 Q;
      float pdfA;
      if (!lightSample(p.lights[li], P, rnd(p.seed, pix, smp, d), rnd(p.seed, pix, smp, d + 1), 1, Q, pdfA)) continue;
      float dist;
      const V3 c = lightContrib(p, sf, P, wo, Q, li, dist);
      if (!(lum(c) > 0.0f)) continue;
      const V3 wi = (Q - P) * (1.0f / dist);
      ++nr;
      if (traceAny(s, offsetOrigin(P, Ng, wi, tIn), wi, dist * 0.9999f)) continue;
      sum = sum + c * (misNee(p, sf, P, wo, wi, li) / pdfA);
    }
    return sum;
  }
  const Pick k = risLights(p, sf, P, wo, pix, smp, dimVertex(b));
  if (k.li < 0 || !(k.target > 0.0f)) return sum;
  float dist;
  const V3 c = lightContrib(p, sf, P, wo, k.Q, k.li, dist);
  const V3 wi = (k.Q - P) * (1.0f / dist);
  ++nr;
  if (traceAny(s, offsetOrigin(P, Ng, wi, tIn), wi, dist * 0.9999f)) return sum;
  const float W = k.wsum / (static_cast<float>(p.restirM) * k.target);
  return c * (W * misNee(p, sf, P, wo, wi, k.li));
}

// ---------------------------------------------------------------- denoiser guides
// Albedo and normal guides of a first hit. A dielectric is mostly seen
// through, so from one the guides follow the ideal refraction (about the
// interface's own normal, no microfacet sampling) through up to kGuideHops
// dielectric interfaces to the first other surface: albedo = F * white +
// (1 - F) * (the colours of the interfaces crossed) * that surface's colour,
// F the first interface's Fresnel reflectance, and the normal is that
// surface's. A chain that leaves the scene, meets total internal reflection
// or runs out of hops keeps F * white and the first hit's normal. Normals
// are returned in camera space (right, up, towards the eye). Guide rays are
// not counted as traced rays: they feed the network, not the estimate.
constexpr int kGuideHops = 4;
__device__ __forceinline__ void guides(const Params& p, const SceneView& s, const Surf& sf0, V3 P0, V3 Ng0, V3 d0,
                                    float t0, V3& alb, V3& nrm) {
  V3 n = sf0.f.n;
  if (sf0.type != kMatDielectric) {
    alb = sf0.C;
  } else {
    Surf sf = sf0;
    V3 P = P0, Ng = Ng0, d = d0;
    float t = t0, F0 = 0.0f;
    V3 tint = mk(1.0f, 1.0f, 1.0f), behind = mk(0.0f, 0.0f, 0.0f);
    for (int k = 0; k < kGuideHops; ++k) {
      const float cosO = -dot(d, sf.f.n);
      const bool outside = cosO > 0.0f;
      const float c = fabsf(cosO);
      const float eta = outside ? sf.ior : 1.0f / sf.ior;  // n_t / n_i
      const float F = fresnelDielectric(c, eta);
      if (k == 0) F0 = F;
      const float s2 = fmaxf(0.0f, 1.0f - c * c) / (eta * eta);
      if (s2 >= 1.0f) break;  // total internal reflection
      const float ct = sqrtf(1.0f - s2);
      const V3 ns = outside ? sf.f.n : neg(sf.f.n);  // normal on the incident side
      d = normalize(d * (1.0f / eta) + ns * (c / eta - ct));
      tint = mk(tint.x * sf.C.x, tint.y * sf.C.y, tint.z * sf.C.z);
      const V3 o = offsetOrigin(P, Ng, d, t);
      Hit h;
      if (!traceClosest(s, o, d, INFINITY, h)) break;
      V3 ng;
      sf = surfaceAt(p, s, h, d, ng);
      P = o + d * h.t;
      Ng = ng;
      t = h.t;
      if (sf.type != kMatDielectric) {
        behind = mk(tint.x * sf.C.x, tint.y * sf.C.y, tint.z * sf.C.z);
        n = sf.f.n;
        break;
      }
    }
    const float T = 1.0f - F0;
    alb = mk(F0 + T * behind.x, F0 + T * behind.y, F0 + T * behind.z);
  }
  nrm = mk(n.x * p.right[0] + n.y * p.right[1] + n.z * p.right[2], n.x * p.up[0] + n.y * p.up[1] + n.z * p.up[2],
           -(n.x * p.dir[0] + n.y * p.dir[1] + n.z * p.dir[2]));
}

// ---------------------------------------------------------------- kernels
// One camera ray; returns the rays traced (1).
__device__ __forceinline__ uint32_t primaryPixel(const Params& p, const SceneView& s, int pix, GPix* __restrict__ gb,
                                                 Reservoir* __restrict__ res, float4* __restrict__ acc) {
  const uint32_t upix = static_cast<uint32_t>(pix), smp = static_cast<uint32_t>(p.sample);
  const int px = pix % p.W, py = pix / p.W;
  V3 o, d;
  cameraRay(p, px + rnd(p.seed, upix, smp, kDimCam), py + rnd(p.seed, upix, smp, kDimCam + 1), o, d);
  Hit h;
  GPix g;
  Reservoir r;
  r.Q = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
  r.wsum = 0.0f;
  r.M = 0.0f;
  r.W = 0.0f;
  r.target = 0.0f;
  if (!traceClosest(s, o, d, INFINITY, h)) {
    g.P = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // w = 0: no surface
    gb[pix] = g;
    res[pix] = r;
    const float l = (p.env[0] + p.env[1] + p.env[2]) * (1.0f / 3.0f);
    float4 a = acc[pix];
    acc[pix] = make_float4(a.x + p.env[0], a.y + p.env[1], a.z + p.env[2], a.w + l * l);
    return 1u;
  }
  V3 Ng;
  const Surf sf = surfaceAt(p, s, h, d, Ng);
  const V3 P = o + d * h.t;
  const V3 wo = neg(d);
  g.P = make_float4(P.x, P.y, P.z, h.t);
  g.N = make_float4(sf.f.n.x, sf.f.n.y, sf.f.n.z, __int_as_float(sf.type));
  g.Ng = make_float4(Ng.x, Ng.y, Ng.z, sf.alpha);
  g.C = make_float4(sf.C.x, sf.C.y, sf.C.z, sf.ior);
  g.wo = make_float4(wo.x, wo.y, wo.z, 0.0f);
  gb[pix] = g;
  if (p.nLights > 0 && !p.neeAll) {
    const Pick k = risLights(p, sf, P, wo, upix, smp, dimVertex(0));
    r.Q = make_float4(k.Q.x, k.Q.y, k.Q.z, __int_as_float(k.li));
    r.wsum = k.wsum;
    r.M = static_cast<float>(p.restirM);
    r.target = k.target;
    r.W = (k.li >= 0 && k.target > 0.0f) ? k.wsum / (r.M * k.target) : 0.0f;
  }
  res[pix] = r;
  return 1u;
}
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kPrimaryWaves))) void k_pt_primary(
    Params p, SceneView s, GPix* __restrict__ gb, Reservoir* __restrict__ res, float4* __restrict__ acc) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nr = pix < p.W * p.H ? primaryPixel(p, s, pix, gb, res, acc) : 0u;
  countRays(p.rays, nr);
}

__device__ __forceinline__ Surf surfOfG(const GPix& g) {
  Surf sf;
  sf.f = onb(mk(g.N.x, g.N.y, g.N.z));
  sf.C = mk(g.C.x, g.C.y, g.C.z);
  sf.type = __float_as_int(g.N.w);
  sf.alpha = g.Ng.w;
  sf.ior = g.C.w;
  return sf;
}

// The denoiser's first-hit record of sample 0 (guides + distance), from the
// G-buffer k_pt_primary wrote for that sample, and the frame's maximum
// first-hit distance (>= 0: float bits order like the floats; one atomic
// per wave). A kernel of its own, launched once per frame with features, so
// the guide chain's registers do not lower the sample passes' occupancy.
__global__ __launch_bounds__(kBlock) void k_pt_guides(Params p, SceneView s, const GPix* __restrict__ gb,
                                                     float4* __restrict__ first) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  float tFirst = 0.0f;
  if (pix < p.W * p.H) {
    const GPix g = gb[pix];
    if (g.P.w == 0.0f) {
      first[2 * pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      first[2 * pix + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else {
      V3 alb, nrm;
      guides(p, s, surfOfG(g), mk(g.P.x, g.P.y, g.P.z), mk(g.Ng.x, g.Ng.y, g.Ng.z), mk(-g.wo.x, -g.wo.y, -g.wo.z),
             g.P.w, alb, nrm);
      first[2 * pix] = make_float4(alb.x, alb.y, alb.z, g.P.w);
      first[2 * pix + 1] = make_float4(nrm.x, nrm.y, nrm.z, 1.0f);
      tFirst = g.P.w;
    }
  }
  int v = static_cast<int>(__float_as_uint(fmaxf(0.0f, tFirst)));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  if (__lane_id() == 0 && v > 0) atomicMax(p.depthMax, static_cast<unsigned int>(v));
}

// The rest of one path; returns the rays traced.
__device__ __forceinline__ uint32_t shadePixel(const Params& p, const SceneView& s, int pix,
                                               const GPix* __restrict__ gb, const Reservoir* __restrict__ res,
                                               float4* __restrict__ acc) {
  const GPix g = gb[pix];
  if (g.P.w == 0.0f) return 0u;  // the environment was added by the primary pass
  uint32_t nr = 0;
  const uint32_t upix = static_cast<uint32_t>(pix), smp = static_cast<uint32_t>(p.sample);
  const Surf sf0 = surfOfG(g);
  const V3 P0 = mk(g.P.x, g.P.y, g.P.z), Ng0 = mk(g.Ng.x, g.Ng.y, g.Ng.z), wo0 = mk(g.wo.x, g.wo.y, g.wo.z);
  V3 L = mk(0.0f, 0.0f, 0.0f);
  // ---- first vertex: ReSTIR over the own and K neighbours' reservoirs
  if (p.nLights > 0) {
    if (p.neeAll) {
      L = directLight(p, s, sf0, P0, Ng0, wo0, g.P.w, upix, smp, 0, nr);
    } else {
      const int px = pix % p.W, py = pix / p.W;
      int src[1 + kMaxSpatial];
      int nsrc = 0;
      src[nsrc++] = pix;
      for (int j = 0; j < p.restirK; ++j) {
        const uint32_t d = kDimSpatial + 3u * static_cast<uint32_t>(j);
        const int qx = min(p.W - 1, max(0, px + static_cast<int>(floorf((2.0f * rnd(p.seed, upix, smp, d) - 1.0f) *
                                                                           p.restirRadius + 0.5f))));
        const int qy = min(p.H - 1, max(0, py + static_cast<int>(floorf((2.0f * rnd(p.seed, upix, smp, d + 1) - 1.0f) *
                                                                           p.restirRadius + 0.5f))));
        const int q = qy * p.W + qx;
        if (q == pix || gb[q].P.w == 0.0f) continue;
        src[nsrc++] = q;
      }
      // combine with balance-heuristic resampling weights: candidate y_i of
      // reservoir i (weight sum wsum_i, M_i candidates, target p_i) enters
      // with m_i(y_i) p_pix(y_i) W_i = p_pix(y_i) wsum_i / sum_j M_j p_j(y_i),
      // m_i = M_i p_i / sum_j M_j p_j summing to one wherever p_pix > 0
      V3 Q = mk(0.0f, 0.0f, 0.0f);
      int li = -1;
      float wsum = 0.0f, tsel = 0.0f;
      for (int i = 0; i < nsrc; ++i) {
        const Reservoir r = res[src[i]];
        const int l = __float_as_int(r.Q.w);
        if (l < 0 || !(r.wsum > 0.0f)) continue;
        const V3 Qi = mk(r.Q.x, r.Q.y, r.Q.z);
        float dist;
        const float t = lum(lightContrib(p, sf0, P0, wo0, Qi, l, dist));
        if (!(t > 0.0f)) continue;
        float den = static_cast<float>(p.restirM) * t;
        for (int j = 1; j < nsrc; ++j) {
          const GPix gq = gb[src[j]];
          const float tj =
              j == i ? r.target
                     : lum(lightContrib(p, surfOfG(gq), mk(gq.P.x, gq.P.y, gq.P.z), mk(gq.wo.x, gq.wo.y, gq.wo.z), Qi,
                                        l, dist));
          den += res[src[j]].M * tj;
        }
        const float w = t * r.wsum / den;
        if (!(w > 0.0f)) continue;
        wsum += w;
        if (rnd(p.seed, upix, smp, kDimSpatial + 3u * static_cast<uint32_t>(i) + 2u) * wsum < w) {
          Q = Qi;
          li = l;
          tsel = t;
        }
      }
      if (li >= 0 && tsel > 0.0f) {
        float dist;
        const V3 c = lightContrib(p, sf0, P0, wo0, Q, li, dist);
        const V3 wi = (Q - P0) * (1.0f / dist);
        ++nr;
        if (!traceAny(s, offsetOrigin(P0, Ng0, wi, g.P.w), wi, dist * 0.9999f))
          L = c * ((wsum / tsel) * misNee(p, sf0, P0, wo0, wi, li));
      }
    }
  }
  // ---- the scatterings: a BSDF sample at vertex b - 1, then the
  // environment on an escape, or vertex b with its own next-event estimate
  // while b < bounces (vertices 0 .. bounces - 1 are shaded)
  V3 thr = mk(1.0f, 1.0f, 1.0f);
  Surf sf = sf0;
  V3 P = P0, Ng = Ng0, wo = wo0;
  float tPrev = g.P.w;
  for (int b = 1; b <= p.bounces; ++b) {
    const uint32_t dv = dimVertex(b - 1) + 56u;
    V3 wiL, wgt;
    if (!bsdfSample(sf, toLocal(sf.f, wo), rnd(p.seed, upix, smp, dv), rnd(p.seed, upix, smp, dv + 1),
                    rnd(p.seed, upix, smp, dv + 2), wiL, wgt))
      break;
    thr = mk(thr.x * wgt.x, thr.y * wgt.y, thr.z * wgt.z);
    if (!(lum(thr) > 0.0f)) break;
    const V3 wi = toWorld(sf.f, wiL);
    const V3 o = offsetOrigin(P, Ng, wi, tPrev);
    Hit h;
    ++nr;
    const bool hit = traceClosest(s, o, wi, INFINITY, h);
    if (p.nLights > 0 && misVertex(sf)) {
      const float pb = sf.alpha < kSmoothAlpha ? INFINITY : bsdfPdf(sf, toLocal(sf.f, wo), wiL);
      const V3 Le = misBsdf(p, P, wi, hit ? h.t : INFINITY, pb);
      L = L + mk(thr.x * Le.x, thr.y * Le.y, thr.z * Le.z);
    }
    if (!hit) {
      L = L + mk(thr.x * p.env[0], thr.y * p.env[1], thr.z * p.env[2]);
      break;
    }
    if (b == p.bounces) break;  // the vertex budget is spent
    V3 ng;
    sf = surfaceAt(p, s, h, wi, ng);
    P = o + wi * h.t;
    Ng = ng;
    wo = neg(wi);
    tPrev = h.t;
    const V3 Ld = directLight(p, s, sf, P, Ng, wo, tPrev, upix, smp, b, nr);
    L = L + mk(thr.x * Ld.x, thr.y * Ld.y, thr.z * Ld.z);
  }
  const float l = (L.x + L.y + L.z) * (1.0f / 3.0f);
  float4 a = acc[pix];
  acc[pix] = make_float4(a.x + L.x, a.y + L.y, a.z + L.z, a.w + l * l);
  return nr;
}
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kShadeWaves))) void k_pt_shade(
    Params p, SceneView s, const GPix* __restrict__ gb, const Reservoir* __restrict__ res, float4* __restrict__ acc) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nr = pix < p.W * p.H ? shadePixel(p, s, pix, gb, res, acc) : 0u;
  countRays(p.rays, nr);
}

// Mean over the samples, and the denoiser features: the encoded colour
// u = log(1 + 32 c) / 4, albedo and normal guides, first-hit distance / its
// frame maximum (maxBits: that maximum as float bits, from k_pt_guides), and
// two noise descriptors: 1 / sqrt(n) for the n samples, and the standard error
// of the mean of the samples' (r + g + b) / 3, s = sqrt(var / n) with the
// unbiased var over the n samples (0 for n = 1), carried into u through its
// slope at the mean m: 8 s / (1 + 32 m), at most 4.
__global__ void k_pt_resolve(int W, int H, int Wp, int Hp, int spp, const float4* __restrict__ acc,
                             const float4* __restrict__ first, const unsigned int* __restrict__ maxBits,
                             float4* __restrict__ color, float* __restrict__ feat) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (feat) {
    if (i >= Wp * Hp) return;
    const int x = min(i % Wp, W - 1), y = min(i / Wp, H - 1);  // edge-replicated padding
    const int j = y * W + x;
    const float inv = 1.0f / static_cast<float>(spp);
    const float4 a = acc[j];
    const float4 f0 = first[2 * j], f1 = first[2 * j + 1];
    const float mx = __uint_as_float(*maxBits);
    float* o = feat + static_cast<size_t>(i) * kPtFeatures;
    o[0] = kColourScale * log1pf(kColourGain * fmaxf(0.0f, a.x * inv));
    o[1] = kColourScale * log1pf(kColourGain * fmaxf(0.0f, a.y * inv));
    o[2] = kColourScale * log1pf(kColourGain * fmaxf(0.0f, a.z * inv));
    o[3] = f0.x;
    o[4] = f0.y;
    o[5] = f0.z;
    o[6] = f1.x;
    o[7] = f1.y;
    o[8] = f1.z;
    o[9] = mx > 0.0f ? f0.w / mx : 0.0f;
    const float m = (a.x + a.y + a.z) * (1.0f / 3.0f) * inv;
    const float var = spp > 1 ? fmaxf(0.0f, a.w * inv - m * m) * static_cast<float>(spp) / static_cast<float>(spp - 1)
                              : 0.0f;
    o[10] = sqrtf(inv);
    o[11] = fminf(4.0f, kColourScale * kColourGain * sqrtf(var * inv) / (1.0f + kColourGain * fmaxf(0.0f, m)));
    if (color && i < W * H) {
      const float4 c = acc[i];
      color[i] = make_float4(c.x * inv, c.y * inv, c.z * inv, 1.0f);
    }
    return;
  }
  if (i >= W * H) return;
  const float inv = 1.0f / static_cast<float>(spp);
  const float4 c = acc[i];
  color[i] = make_float4(c.x * inv, c.y * inv, c.z * inv, 1.0f);
}
// The denoised colour: the network's prediction decoded, c = expm1(4 u) / 32.
__global__ void k_pt_apply_net(int W, int H, int Wp, const float* __restrict__ pred, float4* __restrict__ color) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W * H) return;
  const int x = i % W, y = i / W;
  const float* q = pred + (static_cast<size_t>(y) * Wp + x) * 3;
  const float k = 1.0f / kColourScale, g = 1.0f / kColourGain;
  color[i] = make_float4(fmaxf(0.0f, expm1f(q[0] * k) * g), fmaxf(0.0f, expm1f(q[1] * k) * g),
                         fmaxf(0.0f, expm1f(q[2] * k) * g), 1.0f);
}
// The self-ensemble of the denoiser. Turn k (0..7) of the Wp x Hp feature
// image: a transpose if bit 2 is set, then the columns mirrored if bit 0 is,
// then the rows if bit 1 is. The camera-space normal (right, up) turns with
// the pixels: a transpose maps (x, y) to (-y, -x), a column mirror negates x,
// a row mirror y. The turned image is W1 x H1 (Hp x Wp after a transpose).
__global__ void k_pt_feat_turn(int k, int Wp, int Hp, const float* __restrict__ src, float* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Wp * Hp) return;
  const bool tr = (k & 4) != 0;
  const int W1 = tr ? Hp : Wp, H1 = tr ? Wp : Hp;
  const int x1 = i % W1, y1 = i / W1;
  const int xa = (k & 1) ? W1 - 1 - x1 : x1, ya = (k & 2) ? H1 - 1 - y1 : y1;  // before the mirrors
  const int x = tr ? ya : xa, y = tr ? xa : ya;                                // before the transpose
  const float* a = src + (static_cast<size_t>(y) * Wp + x) * kPtFeatures;
  float* d = dst + static_cast<size_t>(i) * kPtFeatures;
#pragma unroll
  for (int c = 0; c < kPtFeatures; ++c) d[c] = a[c];
  float nx = a[6], ny = a[7];
  if (tr) {
    const float t = nx;
    nx = -ny;
    ny = -t;
  }
  d[6] = (k & 1) ? -nx : nx;
  d[7] = (k & 2) ? -ny : ny;
}
// The network's prediction on turn k (W1 x H1) turned back onto the frame's
// W x H pixels (stride Wp) and weighted by wgt: turn 0 writes, the others add.
__global__ void k_pt_pred_unturn(int k, int W, int H, int Wp, int Hp, float wgt, const float* __restrict__ pk,
                                 float* __restrict__ sum) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W * H) return;
  const int x = i % W, y = i / W;
  const bool tr = (k & 4) != 0;
  const int W1 = tr ? Hp : Wp, H1 = tr ? Wp : Hp;
  const int xa = tr ? y : x, ya = tr ? x : y;
  const int x1 = (k & 1) ? W1 - 1 - xa : xa, y1 = (k & 2) ? H1 - 1 - ya : ya;
  const float* q = pk + (static_cast<size_t>(y1) * W1 + x1) * 3;
  float* o = sum + (static_cast<size_t>(y) * Wp + x) * 3;
  if (k == 0) {
    o[0] = wgt * q[0];
    o[1] = wgt * q[1];
    o[2] = wgt * q[2];
  } else {
    o[0] += wgt * q[0];
    o[1] += wgt * q[1];
    o[2] += wgt * q[2];
  }
}
__global__ void k_pt_zero(float4* __restrict__ a, int n, unsigned int* __restrict__ maxBits,
                          unsigned long long* __restrict__ rays) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (i == 0) {
    *maxBits = 0u;
    *rays = 0ull;
  }
}

}  // namespace pt

// ---------------------------------------------------------------- launchers
namespace {
inline unsigned gridOf(int64_t n, int b) { return static_cast<unsigned>((n + b - 1) / b); }
}  // namespace

void ptZero(float4* acc, int n, unsigned int* maxBits, unsigned long long* rays, hipStream_t st) {
  hipLaunchKernelGGL(pt::k_pt_zero, dim3(gridOf(n, 256)), dim3(256), 0, st, acc, n, maxBits, rays);
}
void ptSample(const pt::Params& p, const SceneView& s, pt::GPix* gb, pt::Reservoir* res, float4* acc, float4* first,
              hipStream_t st) {
  const int n = p.W * p.H;
  const size_t lds = static_cast<size_t>(2) * static_cast<size_t>(s.stackDepth > 0 ? s.stackDepth : 1) * pt::kBlock * 4u;
  hipLaunchKernelGGL(pt::k_pt_primary, dim3(gridOf(n, pt::kBlock)), dim3(pt::kBlock), lds, st, p, s, gb, res, acc);
  if (first && p.sample == 0)
    hipLaunchKernelGGL(pt::k_pt_guides, dim3(gridOf(n, pt::kBlock)), dim3(pt::kBlock), lds, st, p, s, gb, first);
  hipLaunchKernelGGL(pt::k_pt_shade, dim3(gridOf(n, pt::kBlock)), dim3(pt::kBlock), lds, st, p, s, gb, res, acc);
}
void ptResolve(int W, int H, int Wp, int Hp, int spp, const float4* acc, const float4* first,
               const unsigned int* maxBits, float4* color, float* feat, hipStream_t st) {
  if (feat) {
    hipLaunchKernelGGL(pt::k_pt_resolve, dim3(gridOf(static_cast<int64_t>(Wp) * Hp, 256)), dim3(256), 0, st, W, H, Wp,
                       Hp, spp, acc, first, maxBits, color, feat);
    return;
  }
  hipLaunchKernelGGL(pt::k_pt_resolve, dim3(gridOf(W * H, 256)), dim3(256), 0, st, W, H, Wp, Hp, spp, acc, first,
                     maxBits, color, nullptr);
}
void ptFeatTurn(int k, int Wp, int Hp, const float* feat, float* turned, hipStream_t st) {
  hipLaunchKernelGGL(pt::k_pt_feat_turn, dim3(gridOf(static_cast<int64_t>(Wp) * Hp, 256)), dim3(256), 0, st, k, Wp, Hp,
                     feat, turned);
}
void ptPredUnturn(int k, int W, int H, int Wp, int Hp, float wgt, const float* predK, float* sum, hipStream_t st) {
  hipLaunchKernelGGL(pt::k_pt_pred_unturn, dim3(gridOf(static_cast<int64_t>(W) * H, 256)), dim3(256), 0, st, k, W, H,
                     Wp, Hp, wgt, predK, sum);
}
void ptApplyNet(int W, int H, int Wp, const float* pred, float4* color, hipStream_t st) {
  hipLaunchKernelGGL(pt::k_pt_apply_net, dim3(gridOf(W * H, 256)), dim3(256), 0, st, W, H, Wp, pred, color);
}

}  // namespace ubg
