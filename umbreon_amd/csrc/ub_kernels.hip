// ub_kernels.hip -- wavefront per-ray kernels for gfx950 (MI355X).
//
// One frame runs, per transparency layer:
//   k_primary : camera (or continuation) ray -> closest hit -> hit setup
//               (interpolation / side tables, offset origin), compacted into
//               a structure-of-arrays hit queue with one atomic per wave;
//               misses finalize their pixel (composite + fused fog).
//   k_ao      : aoLanes lanes per queued hit, each tracing its share of the
//               hit's AO rays (TEA-seeded, cosine-weighted), reduced with
//               cross-lane shuffles to one occluded count per hit.
//   k_resolve : per hit, the shadow rays, POV local shading, "over"
//               compositing and either the final store (+ fog) or a push to
//               the compacted continuation queue for the next layer.
//   k_post    : supersample box resolve + assumed gamma.
// Reference: integratePixel (shading/transparency.hpp:311-522), shadeHit
// (shading/hit_shader.hpp:147-496), shadeLocal (shading/shading.hpp:76-192),
// computeShadow / computeAO (shading/secondary_rays.hpp:118-134,
// ao/ambient_occlusion.hpp:33-57), applyFog (postprocess/fog.cpp:17-36),
// boxDownsample / applyAssumedGamma (postprocess/image_ops.cpp:9-46).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "ub_device.h"
#include "ub_kernels.h"

namespace ubg {

constexpr int kTraceBlock = 128;
// Waves per SIMD the traversal kernels are register-limited to: k_primary to
// 6 (<= 80 VGPRs), the persistent AO kernel to kAoWavesPerSimd = 5 (<= 96
// VGPRs). The FP64
// polish of grazing roots (ub_device.h) needs ~30 more VGPRs than the FP32
// tests; under the tighter budgets of 7 / 6 waves the whole traversal loop
// spilled to scratch around it (molecular-surface frame 7.38 ms; 6.75 ms
// here; 6.19 ms with the polish compiled out).
constexpr int kTraceWavesPerSimd = 6;
static_assert(kAoWavesPerSimd == 5, "the AO register budget (ub_kernels.h)");

// Dynamic LDS: the per-lane traversal stacks, sized at launch from the scene's
// tree depth (BVH2: one u32 node index per entry; BVH8: a u32 base array and
// a u32 hit-mask array).
extern __shared__ uint32_t ubgLds[];

__host__ size_t stackLdsBytes(const SceneView& s, int block) {
  const int words = s.layout == kLayoutBvh8 ? 2 : 1;
  return static_cast<size_t>(words) * static_cast<size_t>(s.stackDepth > 0 ? s.stackDepth : 1) * block * 4u;
}

// Closest hit over the BVH8 for every lane of a wave at once: each
// iteration the wave runs either primitive tests or node tests (whichever
// more lanes are waiting for, primitives preferred at half the node lanes),
// so the two code paths are never both executed by one iteration. Each lane
// visits nodes and primitives in the same order as traverse8, so the result
// (ties included) is the same.
template <bool kCount>
__device__ __forceinline__ bool traceClosestWave8(const SceneView& s, bool live, V3 o, V3 d, float tn, float tf,
                                                  Hit& best, uint32_t* nv, uint32_t* nt) {
  uint32_t* sb = ubgLds + threadIdx.x;
  uint32_t* sh = sb + s.stackDepth * blockDim.x;
  Trav8 t;
  trav8Init(t, o, d, tn, tf);
  bool found = false;
  live = live && s.nNodes > 0;
  while (__ballot(live) != 0ull) {
    const bool wantPrim = live && t.tHits != 0u;
    const int nPrim = __popcll(__ballot(wantPrim));
    const int nNode = __popcll(__ballot(live && !wantPrim));
    int r = 0;
    if (2 * nPrim >= nNode) {
      if (wantPrim) r = trav8PrimClosest<kCount>(s, t, best, found, nt);
    } else {
      if (live && !wantPrim) r = trav8Node<kCount>(s, t, sb, sh, blockDim.x, nv);
    }
    if (r != 0) live = false;
  }
  return found;
}

// Any hit over the BVH8 for every lane of a wave at once, voted like
// traceClosestWave8; each lane's own sequence of node and primitive tests is
// traverse8's any-hit one, so the answer and the counts are the same.
template <bool kCount>
__device__ __forceinline__ bool traceAnyWave8(const SceneView& s, bool live, V3 o, V3 d, float tn, float tf,
                                              uint32_t* nv, uint32_t* nt) {
  uint32_t* sb = ubgLds + threadIdx.x;
  uint32_t* sh = sb + s.stackDepth * blockDim.x;
  Trav8 t;
  trav8Init(t, o, d, tn, tf);
  bool occluded = false;
  live = live && s.nNodes > 0;
  while (__ballot(live) != 0ull) {
    const bool wantPrim = live && t.tHits != 0u;
    const int nPrim = __popcll(__ballot(wantPrim));
    const int nNode = __popcll(__ballot(live && !wantPrim));
    int r = 0;
    if (2 * nPrim >= nNode) {
      if (wantPrim) r = trav8Prim<kCount>(s, t, nt);
    } else {
      if (live && !wantPrim) r = trav8Node<kCount>(s, t, sb, sh, blockDim.x, nv);
    }
    if (r != 0) {
      live = false;
      occluded = r > 0;
    }
  }
  return occluded;
}

template <int kLayout, bool kAny, bool kCount>
__device__ __forceinline__ bool trace(const SceneView& s, V3 o, V3 d, float tn, float tf, Hit& h, uint32_t* nv,
                                      uint32_t* nt) {
  uint32_t* base = ubgLds + threadIdx.x;
  if (kLayout == kLayoutBvh8)
    return traverse8<kAny, kCount>(s, o, d, tn, tf, h, base, base + s.stackDepth * blockDim.x, blockDim.x, nv, nt);
  return traverse2<kAny, kCount>(s, o, d, tn, tf, h, base, blockDim.x, nv, nt);
}

// hitN.w flag bits
constexpr uint32_t kFlagMesh = 1u << 24;
constexpr uint32_t kFlagAo = 1u << 25;
constexpr uint32_t kFlagShadow = 1u << 26;
constexpr uint32_t kFlagCoarse = 1u << 27;  // AO openness taken from the coarse grid
constexpr uint32_t kFlagGiRecv = 1u << 28;  // GI receiver: constant ambient dropped
constexpr uint32_t kFlagGiToon = 1u << 29;  // toon finish under GI: unit ambient

// This block's statistics group of the counter buffer.
__device__ __forceinline__ unsigned long long* statGroup(unsigned long long* cnt) {
  return cnt + kNumCounters * static_cast<int>(blockIdx.x % kStatSpread);
}

// Fill of partition p of the queue whose fill counters start at `fillBase`.
__device__ __forceinline__ uint32_t queueFill(const unsigned long long* cnt, int fillBase, int p) {
  return static_cast<uint32_t>(cnt[fillBase + kLineWords * p]);
}

// Appends the lanes with `pred` to partition (block % kQueueParts) of a queue
// (one atomic per wave); returns the lane's slot, meaningful where pred.
__device__ __forceinline__ uint32_t queueAppend(unsigned long long* cnt, int fillBase, int partCap, bool pred) {
  const unsigned long long m = __ballot(pred);
  const uint32_t p = blockIdx.x % kQueueParts;
  uint32_t base = 0;
  if (__lane_id() == 0 && m != 0ull)
    base = static_cast<uint32_t>(atomicAdd(&cnt[fillBase + kLineWords * p], (unsigned long long)__popcll(m)));
  base = __shfl(base, 0, 64);
  return p * static_cast<uint32_t>(partCap) + base +
         static_cast<uint32_t>(__popcll(m & ((1ull << __lane_id()) - 1ull)));
}

// Adds the number of lanes with `pred` to statistic `slot`.
__device__ __forceinline__ void waveCount(unsigned long long* cnt, int slot, bool pred) {
  const unsigned long long m = __ballot(pred);
  if (__lane_id() == 0 && m != 0ull) atomicAdd(&statGroup(cnt)[slot], (unsigned long long)__popcll(m));
}

__device__ __forceinline__ uint32_t waveSum(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Visit-counting launches: add the wave's node fetches / primitive tests to
// the counters of ray class `cls` (0 primary, 1 AO, 2 shadow).
__device__ __forceinline__ void addVisits(unsigned long long* cnt, int cls, uint32_t nv, uint32_t nt) {
  const uint32_t sv = waveSum(nv), st = waveSum(nt);
  if (__lane_id() == 0) {
    atomicAdd(&statGroup(cnt)[kCntNodes + cls], (unsigned long long)sv);
    atomicAdd(&statGroup(cnt)[kCntPrimTests + cls], (unsigned long long)st);
  }
}

// embree_renderer.cpp:1400-1414 (top-left origin, pixel centers) on a w x h
// lattice of the camera plane
__device__ __forceinline__ void genRayWH(const FrameConst& fc, uint32_t px, uint32_t py, int w, int h, V3& org,
                                         V3& rd) {
  const float v = 1.0f - 2.0f * ((float)py + 0.5f) / (float)h;
  const float u = 2.0f * ((float)px + 0.5f) / (float)w - 1.0f;
  const V3 pos = mk(fc.pos[0], fc.pos[1], fc.pos[2]);
  const V3 dir = mk(fc.dir[0], fc.dir[1], fc.dir[2]);
  const V3 right = mk(fc.right[0], fc.right[1], fc.right[2]);
  const V3 up = mk(fc.up[0], fc.up[1], fc.up[2]);
  if (fc.ortho) {
    org = pos + right * (u * fc.halfW) + up * (v * fc.halfH);
    rd = dir;
  } else {
    org = pos;
    rd = normalize(dir + right * (u * fc.persHalfW) + up * (v * fc.persHalfH));
  }
}
// Primary ray through (px, py) of the pass's ray lattice (latW x latH).
__device__ __forceinline__ void genRay(const FrameConst& fc, uint32_t px, uint32_t py, V3& org, V3& rd) {
  genRayWH(fc, px, py, fc.latW, fc.latH, org, rd);
}

// Output index of the current pass -> its ray-lattice position (the sample
// maps kSm* of ub_layout.h; embree_renderer.cpp:1512-1518 centres,
// :1654-1660 flagged subpixels, :1551-1567 the aaDepth fine lattice).
__device__ __forceinline__ void sampleLattice(const FrameConst& fc, const FrameBufs& b, uint32_t out, uint32_t& lx,
                                              uint32_t& ly) {
  if (fc.smMode == kSmGrid) {
    lx = out % static_cast<uint32_t>(fc.W);
    ly = out / static_cast<uint32_t>(fc.W);
    return;
  }
  const uint32_t sIdx = fc.smBase + out;
  const uint32_t ss = static_cast<uint32_t>(fc.smSs), wf = static_cast<uint32_t>(fc.smWf);
  if (fc.smMode == kSmCenter) {
    const uint32_t xy = b.smList ? b.smList[sIdx] : sIdx;
    lx = (xy % wf) * ss + static_cast<uint32_t>(fc.smCSub);
    ly = (xy / wf) * ss + static_cast<uint32_t>(fc.smCSub);
    return;
  }
  const uint32_t per = static_cast<uint32_t>(fc.smPer);
  const uint32_t xy = b.smList[sIdx / per];
  const uint32_t r = sIdx % per;
  const uint32_t X = xy % wf, Y = xy / wf;
  if (fc.smMode == kSmCells) {
    const uint32_t centre = static_cast<uint32_t>(fc.smCSub) * ss + static_cast<uint32_t>(fc.smCSub);
    const uint32_t cell = r + (r >= centre ? 1u : 0u);
    lx = X * ss + cell % ss;
    ly = Y * ss + cell / ss;
    return;
  }
  const uint32_t k = static_cast<uint32_t>(fc.smK), kk = k * k;
  const uint32_t cell = r / kk, f = r % kk;
  lx = (X * ss + cell % ss) * k + f % k;
  ly = (Y * ss + cell / ss) * k + f / k;
}

// transparency.hpp:118-129 (view-z clip planes -> ray t range)
__device__ __forceinline__ void clipRange(const FrameConst& fc, V3 rd, float& tn, float& tf) {
  tn = 0.0f;
  tf = __builtin_inff();
  if (!fc.clipOn) return;
  const float cosr = dot(rd, mk(fc.dir[0], fc.dir[1], fc.dir[2]));
  if (cosr <= 1.0e-6f) return;
  if (fc.clipNear > 0.0f) tn = fc.clipNear / cosr;
  if (fc.clipFar < __builtin_inff()) tf = fc.clipFar / cosr;
}

// fog.cpp:7-15
__device__ __forceinline__ float fogFactor(const FrameConst& fc, float z) {
  const float denom = fc.fogEnd - fc.fogStart;
  if (denom <= 0.0f) return z < fc.fogEnd ? 1.0f : 0.0f;
  float f = (fc.fogEnd - z) / denom;
  if (f < 0.0f) return 0.0f;
  if (f > 1.0f) return 1.0f;
  return f;
}

// transparency.hpp:489-521 (floor over background) + fog.cpp:17-36 fused.
__device__ __forceinline__ void finalizePixel(const FrameConst& fc, const FrameBufs& b, uint32_t pix,
                                              float4 acc, V3 base, float baseCov) {
  const float floorW = (1.0f - acc.w) * baseCov;
  float outA = acc.w + floorW;
  outA = fminf(1.0f, fmaxf(0.0f, outA));
  float r = acc.x + floorW * base.x;
  float g = acc.y + floorW * base.y;
  float bl = acc.z + floorW * base.z;
  if (fc.fogOn) {
    const float z = b.viewZ[pix];
    if (z > 0.0f) {
      const float f = fogFactor(fc, z);
      if (fc.transparentBg) {
        outA *= f;
      } else {
        r = fc.fogColor[0] * (1.0f - f) + r * f;
        g = fc.fogColor[1] * (1.0f - f) + g * f;
        bl = fc.fogColor[2] * (1.0f - f) + bl * f;
      }
    }
  }
  b.hiColor[pix] = make_float4(r, g, bl, outA);
}

// ao_coarse.hpp:138-205: bilateral lookup of the coarse AO grid for a hit at
// ray-lattice position (px, py) (hit_shader.hpp:129-133: the lattice is the
// seed lattice) with shading normal N at ray distance z: the 2x2
// surrounding cells weighted bilinear x dot(N, Nc)^32 x exp(-|z - zc| /
// (0.02 z + 1e-6)); false when the weight sum is below 0.3 (the caller then
// gathers inline).
__device__ bool sampleCoarseAo(const FrameConst& fc, const FrameBufs& b, uint32_t px, uint32_t py, V3 N, float z,
                               float& openness, V3& bent) {
  const float sx = ((float)px + 0.5f) * (float)fc.cW / (float)fc.latW - 0.5f;
  const float sy = ((float)py + 0.5f) * (float)fc.cH / (float)fc.latH - 0.5f;
  const int bx = static_cast<int>(floorf(sx));
  const int by = static_cast<int>(floorf(sy));
  const float fx = sx - (float)bx;
  const float fy = sy - (float)by;
  float wSum = 0.0f, opSum = 0.0f;
  V3 bentSum = mk(0.0f, 0.0f, 0.0f);
  for (int dy = 0; dy < 2; ++dy) {
    const int cy = by + dy;
    if (cy < 0 || cy >= fc.cH) continue;
    for (int dx = 0; dx < 2; ++dx) {
      const int cx = bx + dx;
      if (cx < 0 || cx >= fc.cW) continue;
      const int cp = cy * fc.cW + cx;
      const float4 g = b.cGuide[cp];
      if (!(g.w > 0.0f)) continue;  // no AO cell (miss / outline)
      const float wb = (dx ? fx : 1.0f - fx) * (dy ? fy : 1.0f - fy);
      const float nd = dot(N, ld3(g));
      if (nd <= 0.0f) continue;
      const float wn = powf(nd, 32.0f);
      const float wz = expf(-fabsf(z - g.w) / (0.02f * z + 1e-6f));
      const float wgt = wb * wn * wz;
      if (wgt <= 0.0f) continue;
      wSum += wgt;
      opSum += wgt * b.cOpen[cp];
      if (fc.aoEnh) bentSum = bentSum + ld3(b.cBent[cp]) * wgt;
    }
  }
  if (wSum < 0.3f) return false;
  const float inv = 1.0f / wSum;
  openness = opSum * inv;
  if (fc.aoEnh) bent = safeNormalize(mk(bentSum.x * inv, bentSum.y * inv, bentSum.z * inv), N);
  return true;
}

// One block of the primary pass; blk = the block's index within its frame's
// launch (the block itself for a single frame, its offset into the frame's
// block range for a batch).
template <bool kCount, int kLayout>
__device__ __forceinline__ void primaryBlock(const FrameConst& fc, const SceneView& s, const FrameBufs& b, int iter,
                                             const PartFill& fill, int partBlocks, uint32_t blk) {
  const int tid = static_cast<int>(blk) * blockDim.x + threadIdx.x;
  bool valid;
  uint32_t pix = 0;  // output index of the pass
  if (iter == 0 && (fc.smMode == kSmGrid || (fc.smMode == kSmCenter && b.smList == nullptr))) {
    // first layer over a grid (the hi-res pixels, or the output pixels'
    // centres): each wave takes an 8x8 tile (coherent rays)
    const uint32_t gw = static_cast<uint32_t>(fc.smMode == kSmGrid ? fc.W : fc.smWf);
    const uint32_t gh = static_cast<uint32_t>(fc.smMode == kSmGrid ? fc.H : fc.smHf);
    const uint32_t tilesX = (gw + 7u) >> 3;
    const uint32_t wave = static_cast<uint32_t>(tid) >> 6, lane = static_cast<uint32_t>(tid) & 63u;
    const uint32_t px = (wave % tilesX) * 8u + (lane & 7u), py = (wave / tilesX) * 8u + (lane >> 3);
    valid = px < gw && py < gh;
    pix = py * gw + px;
  } else if (iter == 0) {
    // first layer of a list pass: one lane per sample
    valid = static_cast<uint32_t>(tid) < fc.nList;
    pix = static_cast<uint32_t>(tid);
  } else {
    // later layers: partBlocks blocks per continuation-queue partition
    const uint32_t p = blk / static_cast<uint32_t>(partBlocks);
    const uint32_t local = (blk % static_cast<uint32_t>(partBlocks)) * blockDim.x + threadIdx.x;
    valid = local < fill.n[p];
    if (valid) pix = b.activeIn[p * static_cast<uint32_t>(b.partCap) + local];
  }
  bool hit = false;
  Hit h;
  V3 org = mk(0, 0, 0), rd = mk(0, 0, 1);
  float tn = 0.0f, tf = 0.0f;
  uint32_t nv = 0, nt = 0;
  uint32_t lx = 0, ly = 0;  // ray-lattice position
  if (valid) {
    sampleLattice(fc, b, pix, lx, ly);
    genRay(fc, lx, ly, org, rd);
    clipRange(fc, rd, tn, tf);
    if (iter > 0) tn = b.tnear[pix];
  }
  if (kLayout == kLayoutBvh8) {
    hit = traceClosestWave8<kCount>(s, valid, org, rd, tn, tf, h, &nv, &nt);
  } else if (valid) {
    hit = trace<kLayout, false, kCount>(s, org, rd, tn, tf, h, &nv, &nt);
  }
  if (valid) {
    if (iter == 0) {
      const float vz = hit ? h.t * dot(rd, mk(fc.dir[0], fc.dir[1], fc.dir[2])) : 0.0f;
      b.depth[pix] = hit ? h.t : 0.0f;
      b.viewZ[pix] = vz;
      if (b.hitId) b.hitId[pix] = hit ? h.bits : 0xFFFFFFFFu;
      if (b.aoOpen && !hit) b.aoOpen[pix] = 1.0f;
      if (b.firstN && !hit) b.firstN[pix] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0xFFFFFFFFu));
      if (fc.gi && !hit) {
        b.giNrm[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        b.giPos[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
    }
    if (!hit) {
      const float4 acc = iter == 0 ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : b.acc[pix];
      finalizePixel(fc, b, pix, acc, mk(fc.bg[0], fc.bg[1], fc.bg[2]),
                    fc.transparentBg ? 0.0f : 1.0f);
    }
  }
  waveCount(b.counters, kCntPrimary, valid);
  if (kCount) addVisits(b.counters, 0, nv, nt);
  // Compaction into the hit queue: one atomic per wave reserves its slots.
  const uint32_t slot = queueAppend(b.counters, kHitFill, b.partCap, hit);
  if (!hit) return;

  // ---- hit setup (hit_shader.hpp:147-316) ----
  const uint32_t kind = h.bits >> kKindShift;
  const uint32_t id = h.bits & kIdMask;
  V3 N, C;
  float opacity;
  uint32_t flags;
  bool fromEdge = false;
  if (kind == 0u) {
    const uint4 tri = reinterpret_cast<const uint4*>(s.triIdx)[id];
    const float4* vn = reinterpret_cast<const float4*>(s.vtxNrm);
    const float4* vc = reinterpret_cast<const float4*>(s.vtxCol);
    const float4 n0 = vn[tri.x], n1 = vn[tri.y], n2 = vn[tri.z];
    const float4 c0 = vc[tri.x], c1 = vc[tri.y], c2 = vc[tri.z];
    const float w0 = 1.0f - h.u - h.v;
    const V3 nb = mk(w0 * n0.x + h.u * n1.x + h.v * n2.x, w0 * n0.y + h.u * n1.y + h.v * n2.y,
                     w0 * n0.z + h.u * n1.z + h.v * n2.z);
    C = mk(w0 * c0.x + h.u * c1.x + h.v * c2.x, w0 * c0.y + h.u * c1.y + h.v * c2.y,
           w0 * c0.z + h.u * c1.z + h.v * c2.z);
    opacity = w0 * c0.w + h.u * c1.w + h.v * c2.w;
    const V3 ngF = faceForward(h.ng, rd);
    N = normalize(nb);
    if (dot(N, ngF) < 0.0f) N = neg(N);
    flags = (tri.w & kInfoMatMask) | kFlagMesh;
  } else {
    float4 col;
    uint32_t info;
    float op1 = -1.0f;
    if (kind == 1u) {
      col = reinterpret_cast<const float4*>(s.sphColor)[id];
      info = s.sphInfo[id];
    } else if (kind == 3u) {
      col = reinterpret_cast<const float4*>(s.coneColor)[id];
      info = s.coneInfo[id];
      op1 = s.coneOp1[id];
    } else {
      col = reinterpret_cast<const float4*>(s.rndColor)[id];
      info = s.rndInfo[id];
      op1 = s.rndOp1[id];
    }
    N = faceForward(normalize(h.ng), rd);
    C = mk(col.x, col.y, col.z);
    opacity = col.w;
    if (op1 >= 0.0f) {
      float u = h.u;
      if (u < 0.0f) u = 0.0f;
      if (u > 1.0f) u = 1.0f;
      opacity = col.w * (1.0f - u) + op1 * u;
    }
    fromEdge = (info & kInfoFromEdge) != 0u;
    flags = info & kInfoMatMask;
  }
  if (iter == 0 && b.firstN) {
    // adaptive-AA first-hit record (transparency.hpp:402-413, hit_shader.hpp:
    // 214-220 / :321): shading normal, geometry kind and section group
    uint32_t grp;
    if (kind == 0u) grp = s.triGroup[static_cast<int64_t>(id) % s.meshBaseTris];
    else if (kind == 1u) grp = s.sphGroup[id];
    else if (kind == 3u) grp = s.coneGroup[id];
    else grp = s.rndGroup[id];
    b.firstN[pix] = make_float4(N.x, N.y, N.z, __uint_as_float((kind << 16) | grp));
  }
  if (!fromEdge && fc.aoSamples > 0) flags |= kFlagAo;
  if (!fromEdge && fc.shadows) flags |= kFlagShadow;
  const V3 P = mk(org.x + rd.x * h.t, org.y + rd.y * h.t, org.z + rd.z * h.t);
  const float eps = selfIntersectEps(P, rd, h.t);
  // ambient_occlusion.hpp:40-43 / secondary_rays.hpp:121-123: the same offset
  // origin serves the AO and the shadow rays.
  V3 ng = dot(h.ng, N) < 0.0f ? neg(h.ng) : h.ng;
  ng = safeNormalize(ng, N);
  const V3 O = P + ng * eps;
  if (fc.coarse && (flags & kFlagAo)) {
    // hit_shader.hpp:124-140: a hit the coarse grid covers needs no AO rays
    float open;
    V3 bent = N;
    if (sampleCoarseAo(fc, b, lx, ly, N, h.t, open, bent)) {
      flags = (flags & ~kFlagAo) | kFlagCoarse;
      b.hitOpen[slot] = open;
      if (fc.aoEnh) b.hitBent[slot] = make_float4(bent.x, bent.y, bent.z, 0.0f);
    }
  }
  if (fc.gi) {
    // hit_shader.hpp:229-243 / :355-368: a mesh or real-primitive hit gathers
    // indirect instead of its constant ambient; toon finishes keep unit ambient
    bool recv = false;
    float kd = 0.0f;
    if (!fromEdge) {
      const DevMaterial& m = (flags & kFlagMesh) ? s.meshMats[flags & kInfoMatMask] : s.primMats[flags & kInfoMatMask];
      recv = m.toon == 0u;
      kd = diffuseWeight(m);
      flags |= recv ? kFlagGiRecv : kFlagGiToon;
    }
    if (iter == 0) {
      // transparency.hpp:402-413 first-hit G-buffer: position, the normal AOV
      // (mesh hits, hit_shader.hpp:214-220; receivers, :364), reflectance
      const bool nOn = recv || (flags & kFlagMesh) != 0u;
      b.giPos[pix] = make_float4(P.x, P.y, P.z, 0.0f);
      b.giNrm[pix] = nOn ? make_float4(N.x, N.y, N.z, recv ? 1.0f : 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      b.giRefl[pix] = make_float4(kd * C.x, kd * C.y, kd * C.z, 0.0f);
    }
  }
  b.hitPix[slot] = lx + ly * static_cast<uint32_t>(fc.seedW);
  if (b.hitOut) b.hitOut[slot] = pix;
  b.hitT[slot] = h.t;
  b.hitO[slot] = make_float4(O.x, O.y, O.z, eps);
  b.hitN[slot] = make_float4(N.x, N.y, N.z, __uint_as_float(flags));
  b.hitC[slot] = make_float4(C.x, C.y, C.z, opacity);
  b.hitAo[slot] = 0u;
  if (b.hitPrim) b.hitPrim[slot] = h.bits;
}

template <bool kCount, int kLayout>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(kTraceWavesPerSimd))) void k_primary(FrameConst fc, SceneView s, FrameBufs b,
                                                         int iter, PartFill fill, int partBlocks) {
  primaryBlock<kCount, kLayout>(fc, s, b, iter, fill, partBlocks, blockIdx.x);
}

template <bool kCount, int kLayout>
__global__ __launch_bounds__(kTraceBlock) void k_ao(FrameConst fc, SceneView s, FrameBufs b,
                                                    int unused) {
  (void)unused;
  const int lanes = fc.aoLanes;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int hi = tid / lanes;  // hit-queue slot
  const int j = tid % lanes;
  const int part = hi / b.partCap;
  const bool valid = part < kQueueParts && static_cast<uint32_t>(hi % b.partCap) < queueFill(b.counters, kHitFill, part);
  float4 O4 = make_float4(0, 0, 0, 0), N4 = make_float4(0, 0, 1, 0);
  uint32_t pix = 0;
  bool on = false;
  if (valid) {
    O4 = b.hitO[hi];
    N4 = b.hitN[hi];
    pix = b.hitPix[hi];
    on = (__float_as_uint(N4.w) & kFlagAo) != 0u;
  }
  const V3 N = ld3(N4);
  const V3 O = ld3(O4);
  const Frame3 f = frameFromNormal(N);
  const uint32_t seed = pix;  // px + py * seedW
  const int n = fc.aoSamples;
  const int rounds = (n + lanes - 1) / lanes;
  uint32_t occ = 0, nv = 0, nt = 0;
  for (int r = 0; r < rounds; ++r) {
    const int si = j + r * lanes;
    const bool act = valid && on && si < n;
    bool traced = false, sampled = false;
    if (act) {
      uint32_t s0 = seed, s1 = static_cast<uint32_t>(si);
      tea2(s0, s1);
      const V3 dir = cosineSample(u32ToUnorm(s0), u32ToUnorm(s1), f);
      if (dot(dir, N) < 0.01f) {
        sampled = true;
        ++occ;
      } else {
        traced = true;
        Hit h;
        if (trace<kLayout, true, kCount>(s, O, dir, O4.w, fc.aoDist, h, &nv, &nt))
          ++occ;
      }
    }
    waveCount(b.counters, kCntAoTraced, traced);
    waveCount(b.counters, kCntAoSampled, sampled);
  }
  if (kCount) addVisits(b.counters, 1, nv, nt);
  for (int off = lanes >> 1; off > 0; off >>= 1) occ += __shfl_xor(occ, off, lanes);
  if (valid && j == 0) b.hitAo[hi] = occ;
}

// Persistent AO over the compressed BVH8. The AO rays of hit-queue
// partition p are (hit, sample) pairs 0 .. fill(p) * aoSamples - 1; a wave
// takes them in chunks of kAoChunk from partition p's cursor (one atomic per
// chunk), starting at partition (wave % kQueueParts) and moving on to the
// next partition when one is used up, and hands them to its lanes from a
// wave-uniform cursor. Each lane advances its own traversal one node or one
// primitive per iteration and takes the next ray once its own has finished,
// so lanes do not idle behind the longest ray of their wave. Refills wait
// until at least kAoRefill lanes are idle (or the wave would otherwise
// stall), and each iteration runs one kind of step for the whole wave (node
// tests or primitive tests, by lane-count vote). Occluded rays add 1 to their
// hit's count: integer atomics, so the result does not depend on the order.
// The chunk is 256 rays while a partition holds that many for every wave the
// grid gives it, and shrinks (to 128, then 64) when it holds fewer, so the
// small late transparency layers spread over more waves instead of running
// four rays after one another on each lane of a few waves. Chunks of
// different sizes still split a partition into disjoint ranges (each wave
// takes [base, base + its chunk) from the atomic cursor).
constexpr int kAoChunk = 256;
constexpr int kAoChunkMin = 64;
constexpr int kAoRefill = 16;

// The persistent AO loop of one frame: returns once every partition of its
// hit queue is used up and every lane of the wave is idle.
template <bool kCount>
__device__ __forceinline__ void aoPersistentFrame(const FrameConst& fc, const SceneView& s, const FrameBufs& b) {
  const uint32_t n = static_cast<uint32_t>(fc.aoSamples);
  uint32_t* sb = ubgLds + threadIdx.x;
  uint32_t* sh = sb + s.stackDepth * blockDim.x;
  const int stride = blockDim.x;
  const uint64_t laneBit = 1ull << __lane_id();
  // wave-uniform: current partition and cursor into the current chunk
  uint32_t part = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) % kQueueParts;
  uint32_t chunkPart = part;
  const uint64_t wavesPerPart = max(1u, (gridDim.x * (blockDim.x >> 6)) / static_cast<uint32_t>(kQueueParts));
  int tried = 0;
  uint64_t next = 0, end = 0;
  bool drained = false;  // every partition is used up
  bool active = false;
  uint32_t hitIdx = 0;
  uint32_t nTraced = 0, nSampled = 0, nv = 0, nt = 0;
  Trav8 t;
  trav8Init(t, mk(0, 0, 0), mk(0, 0, 1), 0.0f, 0.0f);
  for (;;) {
    const unsigned long long act = __ballot(active);
    const unsigned long long idle = ~act;
    const int nIdle = __popcll(idle);
    if (!drained && (nIdle >= kAoRefill || act == 0ull)) {
      while (next == end) {
        const uint64_t len = static_cast<uint64_t>(queueFill(b.counters, kHitFill, part)) * n;
        unsigned long long* cur = &b.counters[kAoCursor + kLineWords * part];
        uint64_t chunk = kAoChunk;
        while (chunk > kAoChunkMin && len < chunk * wavesPerPart) chunk >>= 1;
        uint64_t base = len;
        if (__lane_id() == 0) {
          // a plain read first: a used-up partition costs no atomic
          const uint64_t seen = __hip_atomic_load(cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (seen < len) base = atomicAdd(cur, (unsigned long long)chunk);
        }
        base = __shfl(base, 0, 64);
        if (base < len) {
          next = base;
          end = min(base + chunk, len);
          chunkPart = part;
        } else {
          if (++tried == kQueueParts) {
            drained = true;
            break;
          }
          part = (part + 1u) % kQueueParts;
        }
      }
      if (next < end) {
        const uint64_t rank = static_cast<uint64_t>(__popcll(idle & (laneBit - 1ull)));
        const uint64_t idx = next + rank;
        const uint64_t taken = min(static_cast<uint64_t>(nIdle), end - next);
        next += taken;
        if (!active && rank < taken) {
          const uint32_t local = static_cast<uint32_t>(idx / n);
          const uint32_t si = static_cast<uint32_t>(idx - static_cast<uint64_t>(local) * n);
          hitIdx = chunkPart * static_cast<uint32_t>(b.partCap) + local;
          const float4 N4 = b.hitN[hitIdx];
          if (__float_as_uint(N4.w) & kFlagAo) {
            const float4 O4 = b.hitO[hitIdx];
            const V3 N = ld3(N4);
            const Frame3 f = frameFromNormal(N);
            uint32_t s0 = b.hitPix[hitIdx], s1 = si;  // seed (px + py*W, i)
            tea2(s0, s1);
            const V3 dir = cosineSample(u32ToUnorm(s0), u32ToUnorm(s1), f);
            if (dot(dir, N) < 0.01f) {
              ++nSampled;
              atomicAdd(&b.hitAo[hitIdx], 1u);
            } else {
              ++nTraced;
              trav8Init(t, ld3(O4), dir, O4.w, fc.aoDist);
              active = true;
            }
          }
        }
      }
    }
    if (act == 0ull && drained) break;
    // Each iteration runs one kind of step for the whole wave: node tests, or
    // the tests of one primitive kind (the kind most lanes have pending; any
    // order of primitive tests gives the same any-hit answer). Primitives go
    // first once their lanes number at least half the node lanes.
    const bool wantPrim = active && t.tHits != 0u;
    const int nPrim = __popcll(__ballot(wantPrim));
    const int nNode = __popcll(__ballot(active && !wantPrim));
    int r = 0;
    if (2 * nPrim >= nNode) {
      const unsigned long long ms = __ballot(wantPrim && trav8KindMask(t, 1u) != 0u);
      const unsigned long long mc = __ballot(wantPrim && trav8KindMask(t, 3u) != 0u);
      const unsigned long long mt = __ballot(wantPrim && trav8KindMask(t, 0u) != 0u);
      const unsigned long long mr = __ballot(wantPrim && trav8KindMask(t, 2u) != 0u);
      const int cs = __popcll(ms), cc = __popcll(mc), ct = __popcll(mt), cr = __popcll(mr);
      const uint64_t me = 1ull << __lane_id();
      if (cs >= cc && cs >= ct && cs >= cr) {
        if (ms & me) r = trav8PrimKind<1, kCount>(s, t, &nt);
      } else if (cc >= ct && cc >= cr) {
        if (mc & me) r = trav8PrimKind<3, kCount>(s, t, &nt);
      } else if (ct >= cr) {
        if (mt & me) r = trav8PrimKind<0, kCount>(s, t, &nt);
      } else {
        if (mr & me) r = trav8PrimKind<2, kCount>(s, t, &nt);
      }
    } else {
      if (active && !wantPrim) r = trav8Node<kCount, true>(s, t, sb, sh, stride, &nv);
    }
    if (r != 0) {
      active = false;
      if (r > 0) atomicAdd(&b.hitAo[hitIdx], 1u);
    }
  }
  const uint32_t tr = waveSum(nTraced), sa = waveSum(nSampled);
  if (__lane_id() == 0) {
    if (tr) atomicAdd(&statGroup(b.counters)[kCntAoTraced], (unsigned long long)tr);
    if (sa) atomicAdd(&statGroup(b.counters)[kCntAoSampled], (unsigned long long)sa);
  }
  if (kCount) addVisits(b.counters, 1, nv, nt);
}

template <bool kCount>
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(kAoWavesPerSimd))) void k_ao_persistent(
    FrameConst fc, SceneView s, FrameBufs b) {
  aoPersistentFrame<kCount>(fc, s, b);
}

// Enhanced AO rays (ambient_occlusion.hpp:171-246): one lane per (hit-queue
// slot, sample) records the nearest-hit distance within the AO radius (+inf
// on a miss, -1 for a below-horizon sample that counts as full contact);
// k_resolve reduces a hit's samples in order, so the float sums are the same
// on every run.
template <bool kCount, int kLayout>
__global__ __launch_bounds__(kTraceBlock) void k_ao_quality(FrameConst fc, SceneView s, FrameBufs b) {
  const int n = fc.aoSamples;
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t slot = gid / n;
  const int si = static_cast<int>(gid - slot * n);
  const int part = static_cast<int>(slot / b.partCap);
  bool traced = false, sampled = false;
  uint32_t nv = 0, nt = 0;
  if (part < kQueueParts && static_cast<uint32_t>(slot % b.partCap) < queueFill(b.counters, kHitFill, part)) {
    const float4 N4 = b.hitN[slot];
    if (__float_as_uint(N4.w) & kFlagAo) {
      const float4 O4 = b.hitO[slot];
      const uint32_t pix = b.hitPix[slot];  // seed base px + py * seedW
      const uint32_t px = pix % static_cast<uint32_t>(fc.seedW), py = pix / static_cast<uint32_t>(fc.seedW);
      const V3 N = ld3(N4);
      const Frame3 f = frameFromNormal(N);
      uint32_t c0 = px + 1u, c1 = py + 1u;
      tea2(c0, c1);
      float u1, u2;
      aoSample2d(fc.aoLowDisc != 0, pix, si, n, u32ToUnorm(c0), u32ToUnorm(c1), u1, u2);
      const V3 dir = cosineSample(u1, u2, f);
      float t;
      if (dot(dir, N) < 0.01f) {
        sampled = true;
        t = -1.0f;
      } else {
        traced = true;
        Hit h;
        t = trace<kLayout, false, kCount>(s, ld3(O4), dir, O4.w, fc.aoDist, h, &nv, &nt) ? h.t
                                                                                          : __builtin_inff();
      }
      b.aoT[slot * n + si] = t;
    }
  }
  waveCount(b.counters, kCntAoTraced, traced);
  waveCount(b.counters, kCntAoSampled, sampled);
  if (kCount) addVisits(b.counters, 1, nv, nt);
}

// computeAOQuality's reduction (ambient_occlusion.hpp:211-264) over the
// recorded distances of one hit, in sample order: openness and bent normal.
// (px, py) is the hit's cell on the seed lattice of width fc.seedW.
__device__ void aoQualityGather(const FrameConst& fc, const FrameBufs& b, int64_t slot, uint32_t px, uint32_t py,
                                V3 N, float& openness, V3& bentOut) {
  const int n = fc.aoSamples;
  const float R = fc.aoDist;
  float radius[3], weight[3];
  int K;
  if (fc.aoMultiScale) {
    K = 3;
    radius[0] = 0.08f * R;
    radius[1] = 0.30f * R;
    radius[2] = 1.00f * R;
    weight[0] = 0.55f;
    weight[1] = 0.35f;
    weight[2] = 0.10f;
  } else {
    K = 1;
    radius[0] = radius[1] = radius[2] = R;
    weight[0] = 1.0f;
    weight[1] = weight[2] = 0.0f;
  }
  const Frame3 f = frameFromNormal(N);
  uint32_t c0 = px + 1u, c1 = py + 1u;
  tea2(c0, c1);
  const float cpx = u32ToUnorm(c0), cpy = u32ToUnorm(c1);
  const uint32_t base = px + py * static_cast<uint32_t>(fc.seedW);
  // (the per-scale sums of the reference only feed the contact / shape AOVs,
  // which this path does not write; the colour needs the weighted sum)
  float weightedObs = 0.0f;
  V3 bentAccum = mk(0.0f, 0.0f, 0.0f);
  const float power = fc.aoFalloff;
  for (int i = 0; i < n; ++i) {
    const float t = b.aoT[slot * n + i];
    if (t == -1.0f) {
      weightedObs += 1.0f;
      continue;
    }
    float perSample = 0.0f;
    for (int k = 0; k < K; ++k) {
      if (t < radius[k]) {
        float fk = 1.0f;
        if (power > 0.0f) {
          float a = 1.0f - t / radius[k];
          if (a < 0.0f) a = 0.0f;
          fk = powf(a, power);
        }
        perSample += weight[k] * fk;
      }
    }
    weightedObs += perSample;
    if (fc.aoBent) {
      float u1, u2;
      aoSample2d(fc.aoLowDisc != 0, base, i, n, cpx, cpy, u1, u2);
      const V3 dir = cosineSample(u1, u2, f);
      bentAccum = bentAccum + dir * (1.0f - perSample);
    }
  }
  const float invN = 1.0f / static_cast<float>(n);
  openness = 1.0f - weightedObs * invN;
  bentOut = fc.aoBent ? safeNormalize(bentAccum, N) : N;
}

// aoApplyFactors (ao_shade.hpp:83-120): per-channel AO factor (scalar for the
// legacy estimator, multibounce when asked), ambient light (bent-normal
// sky/ground gradient) and the direct-diffuse factor.
__device__ void aoApply(const FrameConst& fc, float openness, V3 bent, V3 C, V3& aoF, V3& ambL, float& diffAo) {
  ambL = mk(fc.amb[0], fc.amb[1], fc.amb[2]);
  if (!fc.aoEnh) {
    const float sv = 1.0f - fc.aoIntensity * (1.0f - openness);
    aoF = mk(sv, sv, sv);
    diffAo = 1.0f;
    return;
  }
  if (fc.aoBent) {
    const float w = 0.5f * (dot(bent, mk(fc.aoUpAxis[0], fc.aoUpAxis[1], fc.aoUpAxis[2])) + 1.0f);
    ambL = mk(fc.amb[0] * (fc.aoGround[0] + (fc.aoSky[0] - fc.aoGround[0]) * w),
              fc.amb[1] * (fc.aoGround[1] + (fc.aoSky[1] - fc.aoGround[1]) * w),
              fc.amb[2] * (fc.aoGround[2] + (fc.aoSky[2] - fc.aoGround[2]) * w));
  }
  if (fc.aoMultibounce) {
    const float ox = aoMultibounce(openness, C.x), oy = aoMultibounce(openness, C.y),
                oz = aoMultibounce(openness, C.z);
    aoF = mk(1.0f - fc.aoIntensity * (1.0f - ox), 1.0f - fc.aoIntensity * (1.0f - oy),
             1.0f - fc.aoIntensity * (1.0f - oz));
  } else {
    const float sv = 1.0f - fc.aoIntensity * (1.0f - openness);
    aoF = mk(sv, sv, sv);
  }
  diffAo = fc.aoDiffuse > 0.0f ? 1.0f - fc.aoDiffuse * (1.0f - openness) : 1.0f;
}

// ---- principled direct shading (shading/principled.hpp:139-305) ----
constexpr float kGgxDirectAlphaMin = 1.0e-3f;
constexpr float kPiF = 3.14159265358979323846f;

// isotropic GGX NDF, 0 below the horizon
__device__ __forceinline__ float ggxD(float a2, float ndh) {
  if (ndh <= 0.0f) return 0.0f;
  const float t = ndh * ndh * (a2 - 1.0f) + 1.0f;
  return a2 / (kPiF * t * t);
}
// Smith Lambda of the isotropic GGX (cosT = |cos| to the normal)
__device__ __forceinline__ float ggxLambda(float a2, float cosT) {
  const float c2 = fmaxf(cosT * cosT, 1.0e-12f);
  const float t2 = (1.0f - c2) / c2;
  return 0.5f * (sqrtf(1.0f + a2 * t2) - 1.0f);
}
// anisotropic NDF and Lambda, components in the (t1, t2, N) frame
__device__ __forceinline__ float ggxDAniso(float ax, float ay, float hx, float hy, float hz) {
  if (hz <= 0.0f) return 0.0f;
  const float sx = hx / ax, sy = hy / ay;
  const float t = sx * sx + sy * sy + hz * hz;
  return 1.0f / (kPiF * ax * ay * t * t);
}
__device__ __forceinline__ float ggxLambdaAniso(float ax, float ay, float vx, float vy, float vz) {
  const float z2 = fmaxf(vz * vz, 1.0e-12f);
  const float t2 = (ax * ax * vx * vx + ay * ay * vy * vy) / z2;
  return 0.5f * (sqrtf(1.0f + t2) - 1.0f);
}
// Schlick Fresnel, F90 = 1
__device__ __forceinline__ V3 schlickF(V3 F0, float u) {
  float c = 1.0f - u;
  if (c < 0.0f) c = 0.0f;
  if (c > 1.0f) c = 1.0f;
  const float c5 = c * c * c * c * c;
  return mk(F0.x + (1.0f - F0.x) * c5, F0.y + (1.0f - F0.y) * c5, F0.z + (1.0f - F0.z) * c5);
}
// Disney aspect of the anisotropy stretch
__device__ __forceinline__ float principledAspect(float anisotropy) {
  float a = fabsf(anisotropy);
  if (a > 1.0f) a = 1.0f;
  return sqrtf(1.0f - 0.9f * a);
}

// hit_shader.hpp:382-425: the anisotropy tangent of a principled sphere
// (meridian of the world +Y pole frame) or cylinder (axis projected on the
// tangent plane, caps excluded), rotated by the material's rotation. Returns
// false where the frame degenerates (poles, caps): the hit shades isotropic.
__device__ __forceinline__ bool principledTangent(const SceneView& s, const DevMaterial& m, uint32_t bits, V3 N,
                                                  V3& t1) {
  const uint32_t kind = bits >> kKindShift, id = bits & kIdMask;
  bool ok = false;
  if (kind == 1u) {
    if (fabsf(N.y) < 0.999f) {
      const V3 tPhi = safeNormalize(cross(mk(0.0f, 1.0f, 0.0f), N));
      t1 = cross(N, tPhi);
      ok = true;
    }
  } else if (kind == 2u || kind == 3u) {
    const float4* axes = reinterpret_cast<const float4*>(kind == 3u ? s.coneAxis : s.rndAxis);
    if (axes) {
      const V3 A = ld3(axes[id]);
      // the normalized geometric normal: N up to its sign
      if (fabsf(dot(N, A)) < 0.99f) {
        const float andn = dot(A, N);
        const V3 proj = mk(A.x - N.x * andn, A.y - N.y * andn, A.z - N.z * andn);
        if (dot(proj, proj) >= 1.0e-12f) {
          t1 = safeNormalize(proj);
          ok = true;
        }
      }
    }
  }
  if (ok && m.pbrRotation != 0.0f) {
    const V3 t2 = cross(N, t1);
    t1 = mk(m.rotCos * t1.x + m.rotSin * t2.x, m.rotCos * t1.y + m.rotSin * t2.y,
            m.rotCos * t1.z + m.rotSin * t2.z);
  }
  return ok;
}

// shadePrincipled for one hit: the POV model's ambient / emission start and
// light gates, then per light a diffuse lobe (kd = diffuse * (1 - metallic))
// and a GGX highlight with Schlick Fresnel, pi-scaled to the POV light units;
// under an area light with soft shadows, visibility and BRDF per cone sample
// (the 'PBRD'-tagged per-pixel stream); the Fresnel x background stand-in
// for the environment. `O`, `eps`: the shadow rays' offset origin.
template <bool kCount, int kLayout>
__device__ V3 shadePrincipled(const FrameConst& fc, const SceneView& s, const DevMaterial& m, V3 C, V3 N, V3 V,
                              V3 ambL, V3 aoF, float diffAo, V3 O, float eps, bool shOn, uint32_t px, uint32_t py,
                              bool aniso, V3 t1, uint32_t& nShadow, uint32_t* nv, uint32_t* nt) {
  V3 out = mk(m.emission * C.x + aoF.x * m.ambient * C.x * ambL.x,
              m.emission * C.y + aoF.y * m.ambient * C.y * ambL.y,
              m.emission * C.z + aoF.z * m.ambient * C.z * ambL.z);
  const float kd = m.diffuse * (1.0f - m.pbrMetallic);
  const float d = 0.08f * m.pbrSpecular, mt = m.pbrMetallic;
  const V3 F0 = mk(d + (C.x - d) * mt, d + (C.y - d) * mt, d + (C.z - d) * mt);
  const float f0max = fmaxf(F0.x, fmaxf(F0.y, F0.z));
  float alpha = m.pbrRoughness * m.pbrRoughness;
  if (alpha < kGgxDirectAlphaMin) alpha = kGgxDirectAlphaMin;
  const float a2 = alpha * alpha;
  const float ndv = fmaxf(dot(N, V), 1.0e-4f);
  const float lamV = ggxLambda(a2, ndv);
  // anisotropic frame (principledAnisoCtx): only with a tangent and a
  // non-zero anisotropy; otherwise every expression stays isotropic
  const bool an = aniso && m.pbrAnisotropy != 0.0f;
  V3 t2 = mk(0.0f, 0.0f, 0.0f);
  float ax = 0.0f, ay = 0.0f, lamVA = 0.0f;
  if (an) {
    t2 = cross(N, t1);
    const float aspect = principledAspect(m.pbrAnisotropy);
    ax = alpha / aspect;
    if (ax > 1.0f) ax = 1.0f;
    ay = alpha * aspect;
    if (ax < kGgxDirectAlphaMin) ax = kGgxDirectAlphaMin;
    if (ay < kGgxDirectAlphaMin) ay = kGgxDirectAlphaMin;
    lamVA = ggxLambdaAniso(ax, ay, dot(V, t1), dot(V, t2), ndv);
  }
  // pi * f * (N.L) = pi * D * G2 * F / (4 ndv) for light direction w
  auto specW = [&](V3 w, float ndlw, V3 H, float ndh) {
    if (!an) {
      const float G2 = 1.0f / (1.0f + lamV + ggxLambda(a2, ndlw));
      return kPiF * ggxD(a2, ndh) * G2 / (4.0f * ndv);
    }
    const float G2 = 1.0f / (1.0f + lamVA + ggxLambdaAniso(ax, ay, dot(w, t1), dot(w, t2), ndlw));
    return kPiF * ggxDAniso(ax, ay, dot(H, t1), dot(H, t2), ndh) * G2 / (4.0f * ndv);
  };
  uint32_t q0 = px ^ 0x50425244u, q1 = py;  // 'PBRD'
  for (int li = 0; li < fc.nLights; ++li) {
    const DevLight& l = fc.lights[li];
    const V3 L = mk(l.L[0], l.L[1], l.L[2]);
    const float ndl = dot(N, L);
    if (ndl <= 0.0f) continue;
    Hit h;
    if (!(shOn && fc.shadowSamples > 1 && l.radius > 0.0f)) {
      // one direction (hard shadow, shadows off, or one sample)
      float sf = 1.0f;
      if (shOn) {
        ++nShadow;
        sf = trace<kLayout, true, kCount>(s, O, L, eps, __builtin_inff(), h, nv, nt) ? 0.0f : 1.0f;
      }
      const V3 Lc = mk(l.color[0] * sf, l.color[1] * sf, l.color[2] * sf);
      const float dk = kd * ndl * diffAo;
      out.x += dk * C.x * Lc.x;
      out.y += dk * C.y * Lc.y;
      out.z += dk * C.z * Lc.z;
      if (!l.highlight) continue;  // fill lights: diffuse only
      if (f0max > 0.0f) {
        const V3 H = normalize(mk(L.x + V.x, L.y + V.y, L.z + V.z));
        const float ndh = dot(N, H);
        if (ndh > 0.0f) {
          const float sW = specW(L, ndl, H, ndh);
          const V3 F = schlickF(F0, dot(V, H));
          out.x += sW * F.x * Lc.x;
          out.y += sW * F.y * Lc.y;
          out.z += sW * F.z * Lc.z;
        }
      }
    } else {
      // area light: visibility x BRDF per cone sample; below-horizon samples
      // add nothing but stay in the denominator
      float accD = 0.0f;
      V3 accS = mk(0.0f, 0.0f, 0.0f);
      for (int i = 0; i < fc.shadowSamples; ++i) {
        tea2(q0, q1);
        const V3 w = sampleLightDir(L, l.tanRadius, u32ToUnorm(q0), u32ToUnorm(q1));
        const float ndlS = dot(N, w);
        if (ndlS <= 0.0f) continue;
        ++nShadow;
        if (trace<kLayout, true, kCount>(s, O, w, eps, __builtin_inff(), h, nv, nt)) continue;
        accD += ndlS;
        if (l.highlight && f0max > 0.0f) {
          const V3 H = normalize(mk(w.x + V.x, w.y + V.y, w.z + V.z));
          const float ndh = dot(N, H);
          if (ndh > 0.0f) {
            const float sW = specW(w, ndlS, H, ndh);
            const V3 F = schlickF(F0, dot(V, H));
            accS.x += sW * F.x;
            accS.y += sW * F.y;
            accS.z += sW * F.z;
          }
        }
      }
      const float inv = 1.0f / static_cast<float>(fc.shadowSamples);
      const float dk = kd * diffAo * inv * accD;
      out.x += dk * C.x * l.color[0];
      out.y += dk * C.y * l.color[1];
      out.z += dk * C.z * l.color[2];
      out.x += inv * accS.x * l.color[0];
      out.y += inv * accS.y * l.color[1];
      out.z += inv * accS.z * l.color[2];
    }
  }
  // the environment stand-in outside traced reflections: a converted POV
  // reflection scalar, else F0 x background
  if (m.reflection > 0.0f) {
    out.x += m.reflection * fc.bg[0];
    out.y += m.reflection * fc.bg[1];
    out.z += m.reflection * fc.bg[2];
  } else if (f0max > 0.0f) {
    out.x += F0.x * fc.bg[0];
    out.y += F0.y * fc.bg[1];
    out.z += F0.z * fc.bg[2];
  }
  return out;
}

template <bool kCount, int kLayout, bool kPbr>
__device__ __forceinline__ void resolveHits(const FrameConst& fc, const SceneView& s, const FrameBufs& b, int iter,
                                            uint32_t blk) {
  const int tid = static_cast<int>(blk) * blockDim.x + threadIdx.x;  // hit-queue slot
  const int part = tid / b.partCap;
  const bool valid = part < kQueueParts && static_cast<uint32_t>(tid % b.partCap) < queueFill(b.counters, kHitFill, part);
  uint32_t nShadow = 0, nv = 0, nt = 0;
  bool capped = false, cont = false;
  uint32_t pix = 0;  // output index
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float tnext = 0.0f;
  if (valid) {
    const uint32_t seedPix = b.hitPix[tid];
    pix = b.hitOut ? b.hitOut[tid] : seedPix;
    const float t = b.hitT[tid];
    const float4 O4 = b.hitO[tid];
    const float4 N4 = b.hitN[tid];
    const float4 C4 = b.hitC[tid];
    const uint32_t flags = __float_as_uint(N4.w);
    const uint32_t px = seedPix % static_cast<uint32_t>(fc.seedW), py = seedPix / static_cast<uint32_t>(fc.seedW);
    V3 org, rd;
    genRay(fc, px, py, org, rd);
    const V3 V = normalize(neg(rd));
    const V3 N = ld3(N4);
    const V3 C = ld3(C4);
    const V3 O = ld3(O4);
    const float eps = O4.w;
    const DevMaterial m = (flags & kFlagMesh) ? s.meshMats[flags & kInfoMatMask]
                                              : s.primMats[flags & kInfoMatMask];
    // ao_shade.hpp:83-120 (legacy: scalar aoFactor, scene ambient)
    V3 aoF = mk(1.0f, 1.0f, 1.0f), ambL = mk(fc.amb[0], fc.amb[1], fc.amb[2]);
    float open = 1.0f, diffAo = 1.0f;
    if (flags & (kFlagAo | kFlagCoarse)) {
      V3 bent = N;
      if (flags & kFlagCoarse) {
        open = b.hitOpen[tid];
        if (fc.aoEnh) bent = ld3(b.hitBent[tid]);
      } else if (fc.aoEnh) {
        aoQualityGather(fc, b, tid, px, py, N, open, bent);
      } else {
        open = 1.0f - (float)b.hitAo[tid] / (float)fc.aoSamples;
      }
      aoApply(fc, open, bent, C, aoF, ambL, diffAo);
    }
    if (flags & kFlagGiRecv) ambL = mk(0.0f, 0.0f, 0.0f);
    else if (flags & kFlagGiToon) ambL = mk(1.0f, 1.0f, 1.0f);
    if (iter == 0 && b.aoOpen) b.aoOpen[pix] = open;
    const bool shOn = (flags & kFlagShadow) != 0u;
    V3 out;
    if (kPbr && m.principled) {
      // shading.hpp:88-91: principled materials take their own evaluator
      V3 t1 = mk(0.0f, 0.0f, 0.0f);
      bool aniso = false;
      if (b.hitPrim && m.pbrAnisotropy != 0.0f) {
        const uint32_t bits = b.hitPrim[tid], kind = bits >> kKindShift, id = bits & kIdMask;
        const uint32_t info = kind == 0u ? 0u : kind == 1u ? s.sphInfo[id] : kind == 3u ? s.coneInfo[id] : s.rndInfo[id];
        if (kind != 0u && !(info & kInfoFromEdge)) aniso = principledTangent(s, m, bits, N, t1);
      }
      out = shadePrincipled<kCount, kLayout>(fc, s, m, C, N, V, ambL, aoF, diffAo, O, eps, shOn, px, py, aniso, t1,
                                             nShadow, &nv, &nt);
    } else {
    // shading.hpp:92-190
    out = mk(m.emission * C.x + aoF.x * m.ambient * C.x * ambL.x,
             m.emission * C.y + aoF.y * m.ambient * C.y * ambL.y,
             m.emission * C.z + aoF.z * m.ambient * C.z * ambL.z);
    uint32_t s0 = px, s1 = py;
    for (int li = 0; li < fc.nLights; ++li) {
      const DevLight& l = fc.lights[li];
      const V3 L = mk(l.L[0], l.L[1], l.L[2]);
      const float ndl = dot(N, L);
      if (ndl <= 0.0f) continue;
      float sf = 1.0f;
      if (shOn) {
        Hit h;
        if (!l.soft) {
          ++nShadow;
          sf = trace<kLayout, true, kCount>(s, O, L, eps, __builtin_inff(), h, &nv, &nt)
                   ? 0.0f
                   : 1.0f;
        } else {
          int hits = 0;
          for (int i = 0; i < fc.shadowSamples; ++i) {
            tea2(s0, s1);
            const V3 dir = sampleLightDir(L, l.tanRadius, u32ToUnorm(s0), u32ToUnorm(s1));
            ++nShadow;
            if (trace<kLayout, true, kCount>(s, O, dir, eps, __builtin_inff(), h, &nv, &nt))
              ++hits;
          }
          sf = 1.0f - (float)hits / (float)fc.shadowSamples;
        }
      }
      const V3 Lc = mk(l.color[0] * sf, l.color[1] * sf, l.color[2] * sf);
      float dd;
      if (m.brilliance == 1.0f) dd = ndl;
      else if (m.brilliance == 0.0f) dd = 1.0f;
      else dd = powf(ndl, m.brilliance);
      const float dk = m.diffuse * dd * diffAo;
      out.x += dk * C.x * Lc.x;
      out.y += dk * C.y * Lc.y;
      out.z += dk * C.z * Lc.z;
      if (!l.highlight) continue;
      V3 hl;
      if (m.metallic) {
        float cc = ndl;
        if (cc > 1.0f) cc = 1.0f;
        const float x = acosf(cc) * 0.63661977f;
        float ff = 0.014567225f / ((x - 1.12f) * (x - 1.12f)) - 0.011612903f;
        if (ff < 0.0f) ff = 0.0f;
        if (ff > 1.0f) ff = 1.0f;
        hl = mk(Lc.x * (ff + (1.0f - ff) * C.x), Lc.y * (ff + (1.0f - ff) * C.y),
                Lc.z * (ff + (1.0f - ff) * C.z));
      } else {
        hl = Lc;
      }
      float specW = 0.0f;
      if (m.specular > 0.0f) {
        const V3 H = normalize(mk(L.x + V.x, L.y + V.y, L.z + V.z));
        const float nh = dot(N, H);
        if (nh > 0.0f) specW += m.specular * powf(nh, m.blinnExp);
      }
      if (m.phong > 0.0f) {
        const V3 Rr = N * (2.0f * ndl) - L;
        const float rv = dot(Rr, V);
        if (rv > 0.0f && (m.phongSize < 60.0f || rv > 0.0008f))
          specW += m.phong * powf(rv, m.phongSize);
      }
      if (specW > 0.0f) {
        const float sc = specW * fc.specScale;
        out.x += sc * hl.x;
        out.y += sc * hl.y;
        out.z += sc * hl.z;
      }
    }
    if (m.reflection > 0.0f) {
      out.x += m.reflection * fc.bg[0];
      out.y += m.reflection * fc.bg[1];
      out.z += m.reflection * fc.bg[2];
    }
    }
    // transparency.hpp:431-467 (opaque floor / over / step past the surface)
    if (iter > 0) acc = b.acc[pix];
    const float opacity = C4.w;
    const float kOpaque = 1.0f - 1e-4f;
    if (!fc.transparency || opacity >= kOpaque) {
      finalizePixel(fc, b, pix, acc, out, 1.0f);
    } else {
      const float w = (1.0f - acc.w) * opacity;
      acc.x += w * out.x;
      acc.y += w * out.y;
      acc.z += w * out.z;
      acc.w += w;
      if (acc.w >= kOpaque) {
        finalizePixel(fc, b, pix, acc, mk(fc.bg[0], fc.bg[1], fc.bg[2]),
                      fc.transparentBg ? 0.0f : 1.0f);
      } else {
        const V3 hp = mk(org.x + rd.x * t, org.y + rd.y * t, org.z + rd.z * t);
        const float dirMax = fmaxf(fabsf(rd.x), fmaxf(fabsf(rd.y), fabsf(rd.z)));
        const float es = fmaxf(fmaxf(fabsf(hp.x), fabsf(hp.y)), fmaxf(fabsf(hp.z), dirMax * t));
        tnext = t + es * 0x1.0fp-21f;
        if (iter == fc.maxIters - 1) {
          capped = true;
          finalizePixel(fc, b, pix, acc, mk(fc.bg[0], fc.bg[1], fc.bg[2]),
                        fc.transparentBg ? 0.0f : 1.0f);
        } else {
          cont = true;
        }
      }
    }
  }
  waveCount(b.counters, kCntShaded, valid);
  waveCount(b.counters, kCntCapped, capped);
  {
    const uint32_t sh = waveSum(nShadow);
    if (__lane_id() == 0 && sh) atomicAdd(&statGroup(b.counters)[kCntShadow], (unsigned long long)sh);
  }
  if (kCount) addVisits(b.counters, 2, nv, nt);
  const uint32_t slot = queueAppend(b.counters, kActiveFill, b.partCap, cont);
  if (!cont) return;
  b.activeOut[slot] = pix;
  b.acc[pix] = acc;
  b.tnear[pix] = tnext;
}

template <bool kCount, int kLayout>
__global__ __launch_bounds__(kTraceBlock) void k_resolve(FrameConst fc, SceneView s, FrameBufs b, int iter) {
  resolveHits<kCount, kLayout, false>(fc, s, b, iter, blockIdx.x);
}
template <bool kCount, int kLayout>
__global__ __launch_bounds__(kTraceBlock) void k_resolve_pbr(FrameConst fc, SceneView s, FrameBufs b, int iter) {
  resolveHits<kCount, kLayout, true>(fc, s, b, iter, blockIdx.x);
}

// image_ops.cpp:9-46 -- box resolve (non-finite samples contribute 0) and
// assumed gamma. ss == 1 keeps values as they are (the reference skips the
// downsample entirely).
__device__ __forceinline__ void postPixel(const float4* __restrict__ hi, int W, int H, int ss, float gamma,
                                          float4* __restrict__ out, int i) {
  const int ow = W / ss, oh = H / ss;
  if (i >= ow * oh) return;
  const int x = i % ow, y = i / ow;
  float4 v;
  if (ss == 1) {
    v = hi[i];
  } else {
    const float inv = 1.0f / (float)(ss * ss);
    float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
    for (int dy = 0; dy < ss; ++dy)
      for (int dx = 0; dx < ss; ++dx) {
        const float4 sv = hi[(size_t)(y * ss + dy) * W + (x * ss + dx)];
        a0 += isfinite(sv.x) ? sv.x : 0.0f;
        a1 += isfinite(sv.y) ? sv.y : 0.0f;
        a2 += isfinite(sv.z) ? sv.z : 0.0f;
        a3 += isfinite(sv.w) ? sv.w : 0.0f;
      }
    v = make_float4(a0 * inv, a1 * inv, a2 * inv, a3 * inv);
  }
  if (!(fabsf(gamma - 1.0f) <= 1.0e-4f)) {
    v.x = v.x > 0.0f ? powf(v.x, gamma) : 0.0f;
    v.y = v.y > 0.0f ? powf(v.y, gamma) : 0.0f;
    v.z = v.z > 0.0f ? powf(v.z, gamma) : 0.0f;
  }
  out[i] = v;
}
__global__ void k_post(const float4* __restrict__ hi, int W, int H, int ss, float gamma,
                       float4* __restrict__ out) {
  postPixel(hi, W, H, ss, gamma, out, blockIdx.x * blockDim.x + threadIdx.x);
}


// ---------------------------------------------------------------------------
// Batched frames (ubg_render_batch): the frames of several requests share
// each launch. A BatchItem (device memory) holds one request's frame
// constants, scene and buffers; a BatchLaunch lists the requests a launch
// covers with their block ranges, so a block finds its request by a binary
// search over the block prefix (block-uniform, scalar loads) and runs the
// single-frame block code on that request's data.

// Entry of the launch list that block `blk` belongs to.
__device__ __forceinline__ int batchEntry(const BatchLaunch& L, uint32_t blk) {
  int lo = 0, hi = L.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (blk < L.blockEnd[mid]) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(kTraceWavesPerSimd))) void k_primary_batch(
    const BatchItem* __restrict__ items, const BatchLaunch* __restrict__ L, int iter) {
  const int e = batchEntry(*L, blockIdx.x);
  const uint32_t first = e == 0 ? 0u : L->blockEnd[e - 1];
  const BatchItem& it = items[L->item[e]];
  primaryBlock<false, kLayoutBvh8>(it.fc, it.s, it.b, iter, L->fill[e], L->partBlocks[e], blockIdx.x - first);
}

__global__ __launch_bounds__(kTraceBlock) void k_resolve_batch(const BatchItem* __restrict__ items,
                                                               const BatchLaunch* __restrict__ L, int iter) {
  const int e = batchEntry(*L, blockIdx.x);
  const uint32_t first = e == 0 ? 0u : L->blockEnd[e - 1];
  const BatchItem& it = items[L->item[e]];
  resolveHits<false, kLayoutBvh8, false>(it.fc, it.s, it.b, iter, blockIdx.x - first);
}

// One persistent grid walks the listed requests in order: every wave drains
// a request's hit-queue partitions (its lanes all idle) before it moves to
// the next, so a wave's lanes always trace one scene.
__global__ __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(kAoWavesPerSimd))) void k_ao_batch(
    const BatchItem* __restrict__ items, const BatchLaunch* __restrict__ L) {
  for (int e = 0; e < L->n; ++e) {
    const BatchItem& it = items[L->item[e]];
    if (it.fc.aoSamples > 0) aoPersistentFrame<false>(it.fc, it.s, it.b);
  }
}

__global__ void k_post_batch(const BatchItem* __restrict__ items, const BatchLaunch* __restrict__ L) {
  const int e = batchEntry(*L, blockIdx.x);
  const uint32_t first = e == 0 ? 0u : L->blockEnd[e - 1];
  const BatchItem& it = items[L->item[e]];
  postPixel(it.b.hiColor, it.fc.W, it.fc.H, it.ss, it.gamma, it.out,
            static_cast<int>((blockIdx.x - first) * blockDim.x + threadIdx.x));
}

// Counters of every request to zero (block y = request), or only the queue
// fills and AO cursors of the listed ones.
__global__ void k_zero_batch(const BatchItem* __restrict__ items, int words) {
  unsigned long long* c = items[blockIdx.y].b.counters;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x) c[i] = 0ull;
}
__global__ void k_reset_batch(const BatchItem* __restrict__ items, const BatchLaunch* __restrict__ L) {
  if (static_cast<int>(blockIdx.x) >= L->n) return;
  unsigned long long* c = items[L->item[blockIdx.x]].b.counters;
  if (threadIdx.x < 3 * kQueueParts) c[kHitFill + kLineWords * threadIdx.x] = 0ull;
}
// Continuation-queue fills (request-major, kQueueParts each) and the summed
// statistics (request-major, kNumCounters each) of every request.
__global__ void k_fills_batch(const BatchItem* __restrict__ items, uint32_t* __restrict__ fills) {
  const unsigned long long* c = items[blockIdx.x].b.counters;
  if (threadIdx.x < kQueueParts)
    fills[blockIdx.x * kQueueParts + threadIdx.x] = static_cast<uint32_t>(c[kActiveFill + kLineWords * threadIdx.x]);
}
__global__ void k_stats_batch(const BatchItem* __restrict__ items, unsigned long long* __restrict__ stats) {
  const unsigned long long* c = items[blockIdx.x].b.counters;
  if (threadIdx.x < kNumCounters) {
    unsigned long long v = 0;
    for (int g = 0; g < kStatSpread; ++g) v += c[g * kNumCounters + threadIdx.x];
    stats[blockIdx.x * kNumCounters + threadIdx.x] = v;
  }
}

struct RayIn {
  float org[3];
  float tnear;
  float dir[3];
  float tfar;
};
struct HitOut {
  float t, u, v;
  uint32_t kind, prim;
  float ng[3];
};

template <int kLayout>
__global__ __launch_bounds__(kTraceBlock) void k_query(SceneView s, const RayIn* rays, int64_t n,
                                                       HitOut* hits, uint8_t* occ, int any) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const RayIn r = rays[i];
  const V3 o = mk(r.org[0], r.org[1], r.org[2]);
  const V3 d = mk(r.dir[0], r.dir[1], r.dir[2]);
  Hit h;
  uint32_t nv = 0, nt = 0;
  if (any) {
    occ[i] = trace<kLayout, true, false>(s, o, d, r.tnear, r.tfar, h, &nv,
                                   &nt)
                 ? 1
                 : 0;
    return;
  }
  HitOut out;
  if (trace<kLayout, false, false>(s, o, d, r.tnear, r.tfar, h, &nv,
                             &nt)) {
    out.t = h.t;
    out.u = h.u;
    out.v = h.v;
    out.kind = h.bits >> kKindShift;
    out.prim = h.bits & kIdMask;
    out.ng[0] = h.ng.x;
    out.ng[1] = h.ng.y;
    out.ng[2] = h.ng.z;
  } else {
    out.t = __builtin_inff();
    out.u = out.v = 0.0f;
    out.kind = 0xFFFFFFFFu;
    out.prim = 0xFFFFFFFFu;
    out.ng[0] = out.ng[1] = out.ng[2] = 0.0f;
  }
  hits[i] = out;
}

// Coarse-grid AO, first half (pt1_gbuffer.hpp:86-200 and the setup of
// ao_coarse.hpp:73-129): one thread per coarse cell traces the cell's camera
// ray over [0, inf) and takes the first hit's shading and geometric normals
// (mesh: interpolated vertex normal; real primitive: the hit normal; outline
// primitives and misses are no AO cell), both faced to the viewer. It writes
// the lookup guide (N, t) and appends every AO cell to the hit queue (offset
// origin, normal, seed base cx + cy * cW) for the AO kernels.
template <bool kCount, int kLayout>
__global__ __launch_bounds__(kTraceBlock) void k_coarse_gbuf(FrameConst fc, SceneView s, FrameBufs b) {
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = cell < fc.cW * fc.cH;
  bool ao = false;
  V3 N = mk(0.0f, 0.0f, 0.0f), Ng = N, P = N;
  float t = 0.0f;
  uint32_t nv = 0, nt = 0;
  if (valid) {
    const uint32_t cx = static_cast<uint32_t>(cell % fc.cW), cy = static_cast<uint32_t>(cell / fc.cW);
    V3 org, rd;
    genRayWH(fc, cx, cy, fc.cW, fc.cH, org, rd);
    Hit h;
    if (trace<kLayout, false, kCount>(s, org, rd, 0.0f, __builtin_inff(), h, &nv, &nt)) {
      const uint32_t kind = h.bits >> kKindShift, id = h.bits & kIdMask;
      ao = true;
      if (kind == 0u) {
        const uint4 tri = reinterpret_cast<const uint4*>(s.triIdx)[id];
        const float4* vn = reinterpret_cast<const float4*>(s.vtxNrm);
        const float4 n0 = vn[tri.x], n1 = vn[tri.y], n2 = vn[tri.z];
        const float w0 = 1.0f - h.u - h.v;
        N = safeNormalize(mk(w0 * n0.x + h.u * n1.x + h.v * n2.x, w0 * n0.y + h.u * n1.y + h.v * n2.y,
                             w0 * n0.z + h.u * n1.z + h.v * n2.z));
        Ng = safeNormalize(h.ng);
      } else {
        const uint32_t info = kind == 1u ? s.sphInfo[id] : kind == 3u ? s.coneInfo[id] : s.rndInfo[id];
        if (info & kInfoFromEdge) ao = false;
        N = safeNormalize(h.ng);
        Ng = N;
      }
      if (ao) {
        if (dot(N, rd) > 0.0f) N = neg(N);
        if (dot(Ng, rd) > 0.0f) Ng = neg(Ng);
        t = h.t;
        P = mk(org.x + rd.x * t, org.y + rd.y * t, org.z + rd.z * t);
      }
    }
    b.cGuide[cell] = ao ? make_float4(N.x, N.y, N.z, t) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  waveCount(b.counters, kCntPrimary, valid);
  if (kCount) addVisits(b.counters, 0, nv, nt);
  const uint32_t slot = queueAppend(b.counters, kHitFill, b.partCap, ao);
  if (!ao) return;
  // ambient_occlusion.hpp:40-43 with the G-buffer's eps rule (ao_coarse.hpp:106-110)
  const float eps = selfIntersectEps(P, N, t);
  V3 ng = dot(Ng, N) < 0.0f ? neg(Ng) : Ng;
  ng = safeNormalize(ng, N);
  const V3 O = P + ng * eps;
  b.hitPix[slot] = static_cast<uint32_t>(cell);
  b.hitO[slot] = make_float4(O.x, O.y, O.z, eps);
  b.hitN[slot] = make_float4(N.x, N.y, N.z, __uint_as_float(kFlagAo));
  b.hitAo[slot] = 0u;
}

// Coarse-grid AO, second half: each queued cell's openness (and bent normal)
// from its AO rays, stored in the grid. fc.seedW = cW.
__global__ __launch_bounds__(kTraceBlock) void k_coarse_reduce(FrameConst fc, FrameBufs b) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int part = tid / b.partCap;
  if (part >= kQueueParts || static_cast<uint32_t>(tid % b.partCap) >= queueFill(b.counters, kHitFill, part))
    return;
  const uint32_t cell = b.hitPix[tid];
  const V3 N = ld3(b.hitN[tid]);
  float open;
  if (fc.aoEnh) {
    V3 bent;
    aoQualityGather(fc, b, tid, cell % static_cast<uint32_t>(fc.cW), cell / static_cast<uint32_t>(fc.cW), N, open,
                    bent);
    b.cBent[cell] = make_float4(bent.x, bent.y, bent.z, 0.0f);
  } else {
    open = 1.0f - (float)b.hitAo[tid] / (float)fc.aoSamples;
  }
  b.cOpen[cell] = open;
}

struct SegIn {
  float p[3];
  float q[3];
  int32_t first, count;
};

// The edge pass's argument filter (embree_renderer.cpp:71-131).
struct SegFilter {
  const int32_t* excl;
  int32_t nExcl;
  int64_t baseTris;
  float graze, coplanar;
  V3 dir;
  __device__ bool operator()(const Hit& h) const {
    if (nExcl > 0 && (h.bits >> kKindShift) == 0u) {
      int64_t id = static_cast<int64_t>(h.bits & kIdMask);
      if (baseTris > 0) id %= baseTris;  // instance copies map to their base triangle
      for (int k = 0; k < nExcl; ++k)
        if (static_cast<int64_t>(excl[k]) == id) return false;
    }
    if (graze > 0.0f || coplanar > 0.0f) {
      const float nx = h.ng.x, ny = h.ng.y, nz = h.ng.z;
      const float nlen = sqrtf(nx * nx + ny * ny + nz * nz);
      const float dlen = sqrtf(dir.x * dir.x + dir.y * dir.y + dir.z * dir.z);
      if (nlen > 0.0f && dlen > 0.0f) {
        const float dDotN = nx * dir.x + ny * dir.y + nz * dir.z;
        const float c = fabsf(dDotN) / (nlen * dlen);
        if (graze > 0.0f && c <= graze) return false;
        if (coplanar > 0.0f && h.t * fabsf(dDotN) / nlen <= coplanar) return false;
      }
    }
    return true;
  }
};

// embree_renderer.cpp:294-357, one thread per segment.
template <int kLayout>
__global__ __launch_bounds__(kTraceBlock) void k_segments(SceneView s, const SegIn* segs, int64_t n,
                                                          const int32_t* excl, int64_t nExcl, float eps,
                                                          float graze, float coplanar, uint8_t* occ) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SegIn sg = segs[i];
  const V3 p = mk(sg.p[0], sg.p[1], sg.p[2]);
  const V3 d = mk(sg.q[0] - sg.p[0], sg.q[1] - sg.p[1], sg.q[2] - sg.p[2]);
  const float L = len(d);
  uint8_t r = 0;
  if (L > 0.0f) {
    const V3 dir = mk(d.x / L, d.y / L, d.z / L);
    const float tnear = fmaxf(eps * L, eps);
    const float tfar = L * (1.0f - eps);
    if (tnear < tfar) {
      SegFilter f;
      const bool rangeOk = sg.first >= 0 && sg.count >= 0 && static_cast<int64_t>(sg.first) + sg.count <= nExcl;
      f.excl = rangeOk && sg.count > 0 ? excl + sg.first : nullptr;
      f.nExcl = (rangeOk && s.meshBaseTris > 0) ? sg.count : 0;
      f.baseTris = s.meshBaseTris;
      f.graze = graze;
      f.coplanar = coplanar;
      f.dir = dir;
      uint32_t* base = ubgLds + threadIdx.x;
      uint32_t nv = 0, nt = 0;
      Hit h;
      bool o;
      if (f.nExcl == 0 && graze <= 0.0f && coplanar <= 0.0f) {
        o = trace<kLayout, true, false>(s, p, dir, tnear, tfar, h, &nv, &nt);
      } else if (kLayout == kLayoutBvh8) {
        o = traverse8<true, false>(s, p, dir, tnear, tfar, h, base, base + s.stackDepth * blockDim.x, blockDim.x,
                                   &nv, &nt, f);
      } else {
        o = traverse2<true, false>(s, p, dir, tnear, tfar, h, base, blockDim.x, &nv, &nt, f);
      }
      r = o ? 1 : 0;
    }
  }
  occ[i] = r;
}

// ---- pt1 GI (integrator/pt1/pt1_gather.hpp, embree_renderer.cpp:560-1049) ----

// irradiance_cache.hpp:176-184 (32-bit integer hash)
__device__ __forceinline__ uint32_t hashU32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7FEB352Du;
  x ^= x >> 15;
  x *= 0x846CA68Bu;
  x ^= x >> 16;
  return x;
}

// pt2 first-bounce sampler (pt2_sampler.hpp): hash-based Owen-scrambled
// Sobol (Burley, JCGT 9(4) 2020) -- the Laine-Karras style hash on reversed
// bits, its base-4 variant for the Morton pixel shuffle of the blue-noise
// arrangement, and the second Sobol dimension's direction numbers (rows of
// Pascal's triangle mod 2 in the reversed-bit convention).
__device__ __forceinline__ uint32_t owenHashRev(uint32_t n, uint32_t seed) {
  n ^= n * 0x3d20adeau;
  n += seed;
  n *= (seed >> 16) | 1u;
  n ^= n * 0x05526c56u;
  n ^= n * 0x53a22864u;
  return n;
}
__device__ __forceinline__ uint32_t owenHashRevBase4(uint32_t n, uint32_t seed) {
  n ^= n * 0x3d20adeau;
  n ^= (n >> 1) & (n << 1) & 0x55555555u;
  n += seed;
  n *= (seed >> 16) | 1u;
  n ^= (n >> 1) & (n << 1) & 0x55555555u;
  n ^= n * 0x05526c56u;
  n ^= n * 0x53a22864u;
  return n;
}
__device__ __forceinline__ uint32_t interleaveBits16(uint32_t x) {
  x &= 0x0000ffffu;
  x = (x ^ (x << 8)) & 0x00ff00ffu;
  x = (x ^ (x << 4)) & 0x0f0f0f0fu;
  x = (x ^ (x << 2)) & 0x33333333u;
  x = (x ^ (x << 1)) & 0x55555555u;
  return x;
}
__device__ __forceinline__ float sobolToUnit(uint32_t n) { return static_cast<float>(n) * (1.0f / 4294967808.0f); }
__device__ __forceinline__ void sobolBurley2D(uint32_t index, uint32_t seed, float& u1, float& u2) {
  index = owenHashRev(reverseBits32(index), seed ^ 0xf8ade99au);
  // dimension 0: the van der Corput sequence (the reversed index itself)
  u1 = sobolToUnit(reverseBits32(owenHashRev(reverseBits32(index), seed ^ 0xe0aaaf76u)));
  // dimension 1: XOR of the direction numbers of the index's set bits
  uint32_t r = 0, row = 1;
  for (int bit = 31; bit >= 0; --bit) {
    if ((index >> bit) & 1u) r ^= row;
    row ^= row << 1;
  }
  u2 = sobolToUnit(reverseBits32(owenHashRev(r, seed ^ 0x94964d4eu)));
}

// The pixels of a gather grid to gather (f.nrm.w != 0, or mask != 0) into
// its partitioned list (one atomic per wave, partition = block % kQueueParts);
// with zeroOthers every other pixel gets E = 0 (gatherPt1Grid's zero fill).
__global__ void k_gi_list(GiField f, const uint8_t* __restrict__ mask, int zeroOthers, FrameBufs b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = i < f.w * f.h;
  const bool on = in && (mask ? mask[i] != 0 : f.nrm[i].w != 0.0f);
  const uint32_t slot = queueAppend(b.counters, f.fillBase, b.partCap, on);
  if (on) f.list[slot] = static_cast<uint32_t>(i);
  else if (in && zeroOthers) f.E[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// tracePt1GBuffer (pt1_gbuffer.hpp:86-200): one camera ray per pixel centre
// of the reduced grid (the render camera's plane, no clip range), mesh hits
// and real primitives gather-eligible (toon finishes included), normals
// face-forwarded against the ray, albedo = diffuse weight * pigment.
template <int kLayout>
__global__ __launch_bounds__(kTraceBlock) void k_gi_gbuf(FrameConst fc, SceneView s, GiField f, float4* albedo,
                                                         unsigned long long* counters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < f.w * f.h;
  Hit h;
  V3 org = mk(0, 0, 0), rd = mk(0, 0, 1);
  bool hit = false;
  if (valid) {
    genRayWH(fc, static_cast<uint32_t>(i % f.w), static_cast<uint32_t>(i / f.w), f.w, f.h, org, rd);
    uint32_t nv = 0, nt = 0;
    hit = trace<kLayout, false, false>(s, org, rd, 0.0f, __builtin_inff(), h, &nv, &nt);
  }
  waveCount(counters, kCntGiGbuf, valid);
  if (!valid) return;
  float4* pos = const_cast<float4*>(f.pos);
  float4* nrm = const_cast<float4*>(f.nrm);
  float4* geo = const_cast<float4*>(f.geoN);
  float* dep = const_cast<float*>(f.depth);
  const uint32_t kind = h.bits >> kKindShift, id = h.bits & kIdMask;
  bool ok = hit;
  V3 N = mk(0, 0, 0), Ng = mk(0, 0, 0), A = mk(0, 0, 0);
  if (hit && kind == 0u) {
    const uint4 tri = reinterpret_cast<const uint4*>(s.triIdx)[id];
    const float4* vn = reinterpret_cast<const float4*>(s.vtxNrm);
    const float4* vc = reinterpret_cast<const float4*>(s.vtxCol);
    const float4 n0 = vn[tri.x], n1 = vn[tri.y], n2 = vn[tri.z];
    const float4 c0 = vc[tri.x], c1 = vc[tri.y], c2 = vc[tri.z];
    const float w0 = 1.0f - h.u - h.v;
    N = safeNormalize(mk(w0 * n0.x + h.u * n1.x + h.v * n2.x, w0 * n0.y + h.u * n1.y + h.v * n2.y,
                         w0 * n0.z + h.u * n1.z + h.v * n2.z));
    Ng = safeNormalize(h.ng);
    const float kd = diffuseWeight(s.meshMats[tri.w & kInfoMatMask]);
    A = mk(kd * (w0 * c0.x + h.u * c1.x + h.v * c2.x), kd * (w0 * c0.y + h.u * c1.y + h.v * c2.y),
           kd * (w0 * c0.z + h.u * c1.z + h.v * c2.z));
  } else if (hit) {
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
    ok = (info & kInfoFromEdge) == 0u;
    N = safeNormalize(h.ng);
    Ng = N;
    const float kd = diffuseWeight(s.primMats[info & kInfoMatMask]);
    A = mk(kd * col.x, kd * col.y, kd * col.z);
  }
  if (!ok) {
    pos[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    nrm[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    geo[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    albedo[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    dep[i] = 0.0f;
    return;
  }
  if (dot(N, rd) > 0.0f) N = neg(N);
  if (dot(Ng, rd) > 0.0f) Ng = neg(Ng);
  pos[i] = make_float4(org.x + rd.x * h.t, org.y + rd.y * h.t, org.z + rd.z * h.t, 0.0f);
  nrm[i] = make_float4(N.x, N.y, N.z, 1.0f);
  geo[i] = make_float4(Ng.x, Ng.y, Ng.z, 0.0f);
  albedo[i] = make_float4(A.x, A.y, A.z, 0.0f);
  dep[i] = h.t;
}

// ---- pt1 E denoise: denoise_atrous.cpp as denoisePt1E runs it (no albedo
// demodulation; guides = the normal and the first-hit position of the grid) ----
constexpr float kAtrousEps = 1.0e-3f;

__device__ __forceinline__ float lum709(float4 v) { return 0.2126f * v.x + 0.7152f * v.y + 0.0722f * v.z; }
__device__ __forceinline__ bool atrousValid(const float4* nrm, int p) {
  const float4 n = nrm[p];
  return (n.x * n.x + n.y * n.y + n.z * n.z) > 0.25f;
}
__device__ __forceinline__ float posDist(const float4 a, const float4 c) {
  const float dx = a.x - c.x, dy = a.y - c.y, dz = a.z - c.z;
  return sqrtf(dx * dx + dy * dy + dz * dz);
}

// pt1_denoise.cpp:26-29 NaN/Inf scrub into the working buffer
__global__ void k_atrous_scrub(int n, const float4* __restrict__ E, float4* __restrict__ irr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 e = E[i];
  irr[i] = make_float4(isfinite(e.x) ? e.x : 0.0f, isfinite(e.y) ? e.y : 0.0f, isfinite(e.z) ? e.z : 0.0f, 0.0f);
}

// denoise_atrous.cpp:79-128: per valid pixel, the largest world-space step to a
// valid 4-neighbour (floor 1e-3) and the 3x3 luminance variance
__global__ void k_atrous_stats(int W, int H, const float4* __restrict__ nrm, const float4* __restrict__ pos,
                               const float4* __restrict__ irr, float* __restrict__ grad, float* __restrict__ var) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= W * H || !atrousValid(nrm, p)) return;
  const int x = p % W, y = p / W;
  const float4 pp = pos[p];
  float g = kAtrousEps;
  const int nb[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};
  for (int k = 0; k < 4; ++k) {
    const int qx = x + nb[k][0], qy = y + nb[k][1];
    if (qx < 0 || qx >= W || qy < 0 || qy >= H) continue;
    const int q = qy * W + qx;
    if (atrousValid(nrm, q)) g = fmaxf(g, posDist(pp, pos[q]));
  }
  grad[p] = g;
  float s1 = 0.0f, s2 = 0.0f;
  int n = 0;
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) {
      const int qx = x + dx, qy = y + dy;
      if (qx < 0 || qx >= W || qy < 0 || qy >= H) continue;
      const int q = qy * W + qx;
      if (!atrousValid(nrm, q)) continue;
      const float l = lum709(irr[q]);
      s1 += l;
      s2 += l * l;
      ++n;
    }
  float v = 0.0f;
  if (n > 0) {
    const float mean = s1 / static_cast<float>(n);
    v = fmaxf(0.0f, s2 / static_cast<float>(n) - mean * mean);
  }
  var[p] = v;
}

// denoise_atrous.cpp:134-197: one a-trous iteration with holes of `step`
__global__ void k_atrous_iter(int W, int H, int step, float sigmaZ, float sigmaN, float sigmaL,
                              const float4* __restrict__ nrm, const float4* __restrict__ pos,
                              const float4* __restrict__ in, const float* __restrict__ grad,
                              const float* __restrict__ var, float4* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= W * H) return;
  const float4 ip = in[p];
  if (!atrousValid(nrm, p)) {
    out[p] = ip;
    return;
  }
  const int x = p % W, y = p / W;
  const float h5[5] = {0.0625f, 0.25f, 0.375f, 0.25f, 0.0625f};
  const float4 np = nrm[p];
  const float4 pp = pos[p];
  const float lp = lum709(ip);
  const float lDenom = sigmaL * sqrtf(var[p]) + kAtrousEps;
  const float zDenom = sigmaZ * grad[p] * static_cast<float>(step) + kAtrousEps;
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, wsum = 0.0f;
  for (int dy = -2; dy <= 2; ++dy) {
    const int qy = y + dy * step;
    if (qy < 0 || qy >= H) continue;
    for (int dx = -2; dx <= 2; ++dx) {
      const int qx = x + dx * step;
      if (qx < 0 || qx >= W) continue;
      const int q = qy * W + qx;
      if (!atrousValid(nrm, q)) continue;
      const float4 nq = nrm[q];
      const float4 iq = in[q];
      float w = h5[dx + 2] * h5[dy + 2];
      const float nd = fmaxf(0.0f, np.x * nq.x + np.y * nq.y + np.z * nq.z);
      w *= powf(nd, sigmaN);
      w *= expf(-posDist(pp, pos[q]) / zDenom);
      w *= expf(-fabsf(lp - lum709(iq)) / lDenom);
      s0 += w * iq.x;
      s1 += w * iq.y;
      s2 += w * iq.z;
      wsum += w;
    }
  }
  out[p] = wsum > 0.0f ? make_float4(s0 / wsum, s1 / wsum, s2 / wsum, 0.0f) : ip;
}

// denoise_atrous.cpp:199-208 write-back (valid pixels clamped at 0), into E
// (the first-bounce hit fraction in E.w stays)
__global__ void k_atrous_store(int n, const float4* __restrict__ nrm, const float4* __restrict__ irr,
                               float4* __restrict__ E) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float4 v = irr[p];
  if (atrousValid(nrm, p)) v = make_float4(fmaxf(0.0f, v.x), fmaxf(0.0f, v.y), fmaxf(0.0f, v.z), 0.0f);
  E[p] = make_float4(v.x, v.y, v.z, E[p].w);
}

// upsampleJointBilateral (pt1_upsample.hpp:35-176): per receiver of the
// render grid, the 2x2 reduced-grid neighbourhood weighted bilinear x
// max(0, N.Nh)^normalPow x exp(-|z - zh| / (depthScale z + 1e-6)); when all
// weights die, the nearest valid quad pixel, else (fallbackRadius > 0) the
// nearest valid pixel of the wider window; `patch` marks the pixels whose
// weight sum is below patchThresh or died (the rims re-gathered at full res).
__global__ void k_gi_upsample(FrameConst fc, FrameBufs b, GiField low, float normalPow, float depthScale,
                              int fallbackRadius, uint8_t* __restrict__ patch, float patchThresh) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= fc.W * fc.H) return;
  const int W = fc.W, H = fc.H, hw = low.w, hh = low.h;
  if (patch) patch[pix] = 0;
  const float4 nf = b.giNrm[pix];
  if (nf.w == 0.0f) {
    b.giE[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  const int px = pix % W, py = pix / W;
  const float sy = (static_cast<float>(py) + 0.5f) * static_cast<float>(hh) / static_cast<float>(H) - 0.5f;
  const int by = static_cast<int>(floorf(sy));
  const float fy = sy - static_cast<float>(by);
  const float sx = (static_cast<float>(px) + 0.5f) * static_cast<float>(hw) / static_cast<float>(W) - 0.5f;
  const int bx = static_cast<int>(floorf(sx));
  const float fx = sx - static_cast<float>(bx);
  const float zf = b.depth[pix];
  float wSum = 0.0f, e0 = 0.0f, e1 = 0.0f, e2 = 0.0f, oSum = 0.0f, bestBilin = -1.0f;
  int bestPix = -1;
  for (int dy = 0; dy < 2; ++dy)
    for (int dx = 0; dx < 2; ++dx) {
      int hx = bx + dx, hy = by + dy;
      if (hx < 0) hx = 0;
      if (hy < 0) hy = 0;
      if (hx >= hw) hx = hw - 1;
      if (hy >= hh) hy = hh - 1;
      const int hp = hy * hw + hx;
      const float4 nh = low.nrm[hp];
      if (nh.w == 0.0f) continue;
      const float wb = (dx ? fx : 1.0f - fx) * (dy ? fy : 1.0f - fy);
      if (wb > bestBilin) {
        bestBilin = wb;
        bestPix = hp;
      }
      const float wn = powf(fmaxf(0.0f, nf.x * nh.x + nf.y * nh.y + nf.z * nh.z), normalPow);
      const float zh = low.depth[hp];
      const float wz = expf(-fabsf(zf - zh) / (depthScale * zf + 1e-6f));
      const float w = wb * wn * wz;
      const float4 eh = low.E[hp];
      wSum += w;
      e0 += w * eh.x;
      e1 += w * eh.y;
      e2 += w * eh.z;
      oSum += w * eh.w;
    }
  float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (wSum >= 1e-6f) {
    if (patch && wSum < patchThresh) patch[pix] = 1;
    const float inv = 1.0f / wSum;
    out = make_float4(e0 * inv, e1 * inv, e2 * inv, oSum * inv);
  } else if (bestPix >= 0) {
    if (patch) patch[pix] = 1;
    out = low.E[bestPix];
  } else {
    if (patch) patch[pix] = 1;
    int found = -1;
    if (fallbackRadius > 0) {
      float bestD2 = 1e30f;
      for (int wy = by - fallbackRadius + 1; wy <= by + fallbackRadius; ++wy) {
        if (wy < 0 || wy >= hh) continue;
        for (int wx = bx - fallbackRadius + 1; wx <= bx + fallbackRadius; ++wx) {
          if (wx < 0 || wx >= hw) continue;
          const int hp = wy * hw + wx;
          if (low.nrm[hp].w == 0.0f) continue;
          const float ddx = static_cast<float>(wx) - sx, ddy = static_cast<float>(wy) - sy;
          const float d2 = ddx * ddx + ddy * ddy;
          if (d2 < bestD2) {
            bestD2 = d2;
            found = hp;
          }
        }
      }
    }
    if (found >= 0) out = low.E[found];
  }
  b.giE[pix] = out;
}

// smoothPatchedPixels (pt1_upsample.hpp:188-245), radius 2: a patched pixel's
// E (and hit fraction) becomes the guide-weighted mean of its receiver
// neighbourhood in the snapshot, itself at weight 1, taps below 1e-4 skipped.
__global__ void k_gi_smooth_patched(FrameConst fc, FrameBufs b, const uint8_t* __restrict__ patch, float normalPow,
                                    float depthScale, const float4* __restrict__ Esrc) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= fc.W * fc.H || !patch[pix]) return;
  const int W = fc.W, H = fc.H, px = pix % W, py = pix / W;
  const float4 nf = b.giNrm[pix];
  const float zf = b.depth[pix];
  const float4 es = Esrc[pix];
  float wSum = 1.0f, e0 = es.x, e1 = es.y, e2 = es.z, oSum = es.w;
  for (int dy = -2; dy <= 2; ++dy) {
    const int ny = py + dy;
    if (ny < 0 || ny >= H) continue;
    for (int dx = -2; dx <= 2; ++dx) {
      if (dx == 0 && dy == 0) continue;
      const int nx = px + dx;
      if (nx < 0 || nx >= W) continue;
      const int q = ny * W + nx;
      const float4 nn = b.giNrm[q];
      if (nn.w == 0.0f) continue;
      const float wn = powf(fmaxf(0.0f, nf.x * nn.x + nf.y * nn.y + nf.z * nn.z), normalPow);
      const float zn = b.depth[q];
      const float wz = expf(-fabsf(zf - zn) / (depthScale * zf + 1e-6f));
      const float w = wn * wz;
      if (w < 1e-4f) continue;
      const float4 eq = Esrc[q];
      wSum += w;
      e0 += w * eq.x;
      e1 += w * eq.y;
      e2 += w * eq.z;
      oSum += w * eq.w;
    }
  }
  const float inv = 1.0f / wSum;
  b.giE[pix] = make_float4(e0 * inv, e1 * inv, e2 * inv, oSum * inv);
}

// Closest hit for the lanes with `live`, called by every lane of the wave:
// the wave-vote BVH8 traversal, or the per-lane one on the BVH2 layout.
template <bool kCount, int kLayout>
__device__ __forceinline__ bool closestHitWave(const SceneView& s, bool live, V3 o, V3 d, float tn, float tf, Hit& h,
                                               uint32_t* nv, uint32_t* nt) {
  if (kLayout == kLayoutBvh8) return traceClosestWave8<kCount>(s, live, o, d, tn, tf, h, nv, nt);
  return live && trace<kLayout, false, kCount>(s, o, d, tn, tf, h, nv, nt);
}

// Any hit for the lanes with `live`, called by every lane of the wave.
template <bool kCount, int kLayout>
__device__ __forceinline__ bool anyHitWave(const SceneView& s, bool live, V3 o, V3 d, float tn, float tf, uint32_t* nv,
                                           uint32_t* nt) {
  if (kLayout == kLayoutBvh8) return traceAnyWave8<kCount>(s, live, o, d, tn, tf, nv, nt);
  Hit h;
  return live && trace<kLayout, true, kCount>(s, o, d, tn, tf, h, nv, nt);
}

// pt1GatherPoint (pt1_gather.hpp:228-485) for one listed pixel per lane: spp
// cosine paths from its first hit, each vertex adding its shadow-tested
// direct light (pt1EvalVertex, :117-212) times the path throughput, Russian
// roulette from the second vertex, the environment at an escape. A lane's
// samples run in order and are summed in order, so E is the oracle's to the
// last bit of the same float operations; E.w = the first-bounce hit
// fraction. The lanes advance as a state machine with one path segment per
// iteration, so every iteration's gather rays go through one wave-wide
// closest-hit traversal (node and primitive steps voted across the wave).
// Pixels are handed out as work: a lane that has finished its pixel goes
// idle, and once kGiRefill lanes of the wave are idle (or all of them), the
// wave takes that many consecutive list entries from a partition's cursor
// (one atomic), moving on to the next partition when one is used up; a wave
// leaves when every partition is used up and none of its lanes has work.
constexpr int kGiRefill = 2;

template <bool kCount, int kLayout>
__global__ __launch_bounds__(kTraceBlock) void k_gi_gather(FrameConst fc, SceneView s, FrameBufs b, GiField gf) {
  bool live = false;
  uint32_t nGi = 0, nNee = 0, nv = 0, nt = 0;
  // per-pixel constants
  uint32_t pix = 0, seed = 0, sBase = 0, sSeed = 0;
  V3 N = mk(0, 0, 1), Ng = N, O = mk(0, 0, 0);
  float eps = 0.0f, cpx = 0.0f, cpy = 0.0f;
  const int spp = gf.spp;
  // per-sample path state
  int si = 0, bo = 0, nOcc = 0;
  bool inSample = false;
  uint32_t s0 = 0, s1 = 0;
  V3 org = O, wi = mk(0, 0, 1), tp = mk(1, 1, 1), L = mk(0, 0, 0), Esum = mk(0, 0, 0);
  float tnear = eps;
  auto beginPixel = [&](uint32_t slot) {
    live = true;
    si = 0;
    nOcc = 0;
    inSample = false;
    Esum = mk(0.0f, 0.0f, 0.0f);
    pix = gf.list[slot];
    const V3 P = ld3(gf.pos[pix]);
    N = ld3(gf.nrm[pix]);
    // the render grid's G-buffer has no geometric normal: the horizon test
    // then uses the shading normal (gatherPt1Grid, pt1_gather.hpp:570-573)
    Ng = gf.geoN ? ld3(gf.geoN[pix]) : N;
    seed = hashU32(pix ^ fc.giSeedMix);
    eps = selfIntersectEps(P, N, gf.depth[pix]);
    O = P + N * eps;
    // pt2MakePixelSampler: blue noise = one global sequence, pixel sections
    // of nextPow2(spp) samples along the base-4 scrambled Morton curve;
    // sobol = an independent sequence per pixel
    sSeed = seed;
    if (fc.giPt2 && fc.giPattern == 1) {
      uint32_t pow2 = 1;
      while (pow2 < static_cast<uint32_t>(spp)) pow2 <<= 1;
      const uint32_t px = pix % static_cast<uint32_t>(gf.w), py = pix / static_cast<uint32_t>(gf.w);
      const uint32_t m = (interleaveBits16(px) << 1) | interleaveBits16(py);
      sBase = reverseBits32(owenHashRevBase4(reverseBits32(m), fc.giSeedMix)) * pow2;
      sSeed = fc.giSeedMix;
    }
    if (fc.giLd) {
      uint32_t c0 = seed, c1 = 0x9E3779B9u;
      tea2(c0, c1);
      cpx = u32ToUnorm(c0);
      cpy = u32ToUnorm(c1);
    }
  };
  const int curBase = gf.fillBase == kGiFill ? kGiCursor : kGiPatchCursor;
  uint32_t part = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) % kQueueParts;
  int tried = 0;
  bool drained = false;  // every partition is used up
  auto endSample = [&]() {
    if (fc.giClamp > 0.0f) {
      const float lum = 0.2126f * L.x + 0.7152f * L.y + 0.0722f * L.z;
      if (lum > fc.giClamp) {
        const float k = fc.giClamp / lum;
        L.x *= k;
        L.y *= k;
        L.z *= k;
      }
    }
    Esum.x += L.x;
    Esum.y += L.y;
    Esum.z += L.z;
    ++si;
    inSample = false;
  };
  for (;;) {
    const unsigned long long idle = __ballot(!live);
    const int nIdle = __popcll(idle);
    if (!drained && (nIdle >= kGiRefill || nIdle == 64)) {
      uint32_t base = 0, take = 0;
      while (take == 0u && !drained) {
        const uint32_t fill = queueFill(b.counters, gf.fillBase, static_cast<int>(part));
        unsigned long long* cur = &b.counters[curBase + kLineWords * part];
        uint32_t got = fill;
        if (__lane_id() == 0) {
          // a plain read first: a used-up partition costs no atomic
          const uint64_t seen = __hip_atomic_load(cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (seen < fill) got = static_cast<uint32_t>(atomicAdd(cur, static_cast<unsigned long long>(nIdle)));
        }
        got = __shfl(got, 0, 64);
        if (got < fill) {
          base = got;
          take = min(static_cast<uint32_t>(nIdle), fill - got);
        } else if (++tried == kQueueParts) {
          drained = true;
        } else {
          part = (part + 1u) % kQueueParts;
        }
      }
      const uint32_t rank = static_cast<uint32_t>(__popcll(idle & ((1ull << __lane_id()) - 1ull)));
      if (!live && rank < take) beginPixel(part * static_cast<uint32_t>(b.partCap) + base + rank);
    }
    // lanes between samples start the next one (a below-horizon draw is a
    // sample of L = 0 that traces nothing); a lane whose samples are done
    // stores E
    while (live && !inSample) {
      if (si == spp) {
        const float invN = 1.0f / static_cast<float>(spp);
        gf.E[pix] = make_float4(Esum.x * invN, Esum.y * invN, Esum.z * invN, static_cast<float>(nOcc) * invN);
        live = false;
        break;
      }
      s0 = seed;
      s1 = static_cast<uint32_t>(si);
      tea2(s0, s1);
      float u1, u2;
      if (fc.giPt2) {
        sobolBurley2D(sBase + static_cast<uint32_t>(si), sSeed, u1, u2);
      } else if (fc.giLd) {
        aoSample2d(true, 0u, si, spp, cpx, cpy, u1, u2);
      } else {
        u1 = u32ToUnorm(s0);
        u2 = u32ToUnorm(s1);
      }
      wi = cosineSample(u1, u2, frameFromNormal(N));
      L = mk(0.0f, 0.0f, 0.0f);
      if (!(dot(wi, Ng) > 0.0f)) {
        ++nOcc;
        endSample();
        continue;
      }
      org = O;
      tnear = eps;
      tp = mk(1.0f, 1.0f, 1.0f);
      bo = 1;
      inSample = true;
    }
    const bool ray = live && inSample;
    if (__ballot(ray) == 0ull) {  // no lane has work: refill, or leave
      if (drained) break;
      continue;
    }
    Hit h;
    const bool hit = closestHitWave<kCount, kLayout>(s, ray, org, wi, tnear, fc.giMaxDist, h, &nv, &nt);
    // the hit: an escape adds the environment, an outline decoration absorbs
    // the path, anything else is a path vertex (vert)
    bool vert = false;
    V3 Ny = mk(0.0f, 0.0f, 1.0f), Cy = mk(0.0f, 0.0f, 0.0f), NgS = Ny, Py = mk(0.0f, 0.0f, 0.0f);
    float kd = 0.0f, em = 0.0f, epsY = 0.0f;
    if (ray) {
      ++nGi;
      if (!hit) {
        // irradiance_cache.hpp:188-194
        const float w = 0.5f * (dot(wi, mk(fc.aoUpAxis[0], fc.aoUpAxis[1], fc.aoUpAxis[2])) + 1.0f);
        L.x += tp.x * (fc.giEnvScale[0] * (fc.giGround[0] + (fc.giSky[0] - fc.giGround[0]) * w));
        L.y += tp.y * (fc.giEnvScale[1] * (fc.giGround[1] + (fc.giSky[1] - fc.giGround[1]) * w));
        L.z += tp.z * (fc.giEnvScale[2] * (fc.giGround[2] + (fc.giSky[2] - fc.giGround[2]) * w));
        endSample();
      } else {
        if (bo == 1) ++nOcc;
        // pt1EvalVertex: the hit as a path vertex
        const uint32_t kind = h.bits >> kKindShift, id = h.bits & kIdMask;
        bool absorbed = false;
        if (kind == 0u) {
          const uint4 tri = reinterpret_cast<const uint4*>(s.triIdx)[id];
          const float4* vn = reinterpret_cast<const float4*>(s.vtxNrm);
          const float4* vc = reinterpret_cast<const float4*>(s.vtxCol);
          const float4 n0 = vn[tri.x], n1 = vn[tri.y], n2 = vn[tri.z];
          const float4 c0 = vc[tri.x], c1 = vc[tri.y], c2 = vc[tri.z];
          const float w0 = 1.0f - h.u - h.v;
          Ny = normalize(mk(w0 * n0.x + h.u * n1.x + h.v * n2.x, w0 * n0.y + h.u * n1.y + h.v * n2.y,
                            w0 * n0.z + h.u * n1.z + h.v * n2.z));
          Cy = mk(w0 * c0.x + h.u * c1.x + h.v * c2.x, w0 * c0.y + h.u * c1.y + h.v * c2.y,
                  w0 * c0.z + h.u * c1.z + h.v * c2.z);
          const DevMaterial& mm = s.meshMats[tri.w & kInfoMatMask];
          kd = diffuseWeight(mm);
          if (fc.giPt2 && fc.giEmissive) em = mm.emission;
          NgS = h.ng;
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
          absorbed = (info & kInfoFromEdge) != 0u;  // outline decoration absorbs the path
          Ny = safeNormalize(h.ng);
          Cy = mk(col.x, col.y, col.z);
          const DevMaterial& pm = s.primMats[info & kInfoMatMask];
          kd = diffuseWeight(pm);
          if (fc.giPt2 && fc.giEmissive) em = pm.emission;
          NgS = Ny;
        }
        if (absorbed) {
          endSample();
        } else {
          if (dot(Ny, wi) > 0.0f) Ny = neg(Ny);
          Py = mk(org.x + wi.x * h.t, org.y + wi.y * h.t, org.z + wi.z * h.t);
          epsY = selfIntersectEps(Py, wi, h.t);
          vert = true;
        }
      }
    }
    // next-event estimation at the vertices: every lane of the wave takes
    // part in each light's voted shadow traversal (lanes without a vertex or
    // facing away from the light as idle lanes)
    V3 E = mk(0.0f, 0.0f, 0.0f);
    for (int li = 0; li < fc.nLights; ++li) {
      const DevLight& l = fc.lights[li];
      const V3 Ld = mk(l.L[0], l.L[1], l.L[2]);
      const float ndl = vert ? dot(Ny, Ld) : 0.0f;
      const bool shadowRay = vert && ndl > 0.0f;
      V3 So = mk(0.0f, 0.0f, 0.0f), dir = Ld;
      if (shadowRay) {
        // computeShadow's single hard ray (secondary_rays.hpp:118-124), or
        // (pt2, an area light) one cone-jittered ray from the path's stream
        V3 ng = dot(NgS, Ny) < 0.0f ? neg(NgS) : NgS;
        ng = safeNormalize(ng, Ny);
        So = Py + ng * epsY;
        if (fc.giPt2 && l.radius > 0.0f) {
          tea2(s0, s1);
          dir = sampleLightDir(Ld, l.tanRadius, u32ToUnorm(s0), u32ToUnorm(s1));
        }
        ++nNee;
      }
      const bool occl = anyHitWave<kCount, kLayout>(s, shadowRay, So, dir, epsY, __builtin_inff(), &nv, &nt);
      if (shadowRay) {
        const float sh = occl ? 0.0f : 1.0f;
        E.x += ndl * l.color[0] * sh;
        E.y += ndl * l.color[1] * sh;
        E.z += ndl * l.color[2] * sh;
      }
    }
    if (!vert) continue;
    // radiance + emitted (pt1EvalVertex's split; emitted is 0 unless pt2
    // transports the vertex's emission)
    const V3 emd = em > 0.0f ? mk(em * Cy.x, em * Cy.y, em * Cy.z) : mk(0.0f, 0.0f, 0.0f);
    L.x += tp.x * (kd * Cy.x * E.x + emd.x);
    L.y += tp.y * (kd * Cy.y * E.y + emd.y);
    L.z += tp.z * (kd * Cy.z * E.z + emd.z);
    if (bo == fc.giBounces) {
      endSample();
      continue;
    }
    tp.x *= kd * Cy.x;
    tp.y *= kd * Cy.y;
    tp.z *= kd * Cy.z;
    if (bo >= 2) {
      tea2(s0, s1);
      float pc = fmaxf(tp.x, fmaxf(tp.y, tp.z));
      pc = fminf(0.95f, fmaxf(0.05f, pc));
      if (u32ToUnorm(s0) >= pc) {
        endSample();
        continue;
      }
      const float inv = 1.0f / pc;
      tp.x *= inv;
      tp.y *= inv;
      tp.z *= inv;
    }
    tea2(s0, s1);
    const Frame3 fv = frameFromNormal(Ny);
    wi = cosineSample(u32ToUnorm(s0), u32ToUnorm(s1), fv);
    const float epsV = selfIntersectEps(Py, wi, h.t);
    org = mk(Py.x + Ny.x * epsV, Py.y + Ny.y * epsV, Py.z + Ny.z * epsV);
    tnear = epsV;
    ++bo;
  }
  {
    const uint32_t g = waveSum(nGi), ne = waveSum(nNee);
    if (__lane_id() == 0) {
      if (g) atomicAdd(&statGroup(b.counters)[kCntGiRays], (unsigned long long)g);
      if (ne) atomicAdd(&statGroup(b.counters)[kCntGiNee], (unsigned long long)ne);
    }
  }
  if (kCount) addVisits(b.counters, 1, nv, nt);
}

// embree_renderer.cpp:1027-1049 (color += giIntensity * reflectance * E on
// receivers) then fog.cpp:17-36 over every hi-res pixel (the layer passes
// ran without the fused fog).
__global__ void k_gi_composite(FrameConst fc, FrameBufs b, float* __restrict__ indirect,
                               float* __restrict__ occlusion) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= fc.W * fc.H) return;
  float4 c = b.hiColor[i];
  V3 ind = mk(0.0f, 0.0f, 0.0f);
  float occ = 0.0f;
  if (b.giNrm[i].w != 0.0f) {
    const float4 E = b.giE[i];
    const float4 R = b.giRefl[i];
    ind = mk(fc.giIntensity * R.x * E.x, fc.giIntensity * R.y * E.y, fc.giIntensity * R.z * E.z);
    occ = E.w;
    c.x += ind.x;
    c.y += ind.y;
    c.z += ind.z;
  }
  const float vz = b.viewZ[i];
  if (fc.fogOn && vz > 0.0f) {
    const float f = fogFactor(fc, vz);
    if (fc.transparentBg) {
      c.w *= f;
    } else {
      c.x = fc.fogColor[0] * (1.0f - f) + c.x * f;
      c.y = fc.fogColor[1] * (1.0f - f) + c.y * f;
      c.z = fc.fogColor[2] * (1.0f - f) + c.z * f;
    }
  }
  b.hiColor[i] = c;
  if (indirect) {
    indirect[3 * static_cast<int64_t>(i) + 0] = ind.x;
    indirect[3 * static_cast<int64_t>(i) + 1] = ind.y;
    indirect[3 * static_cast<int64_t>(i) + 2] = ind.z;
  }
  if (occlusion) occlusion[i] = occ;
}

// ---- adaptive antialiasing (render/adaptive_aa.hpp, embree_renderer.cpp:1428-1674) ----

// One output pixel's centre result as the refinement predicates see it.
struct AaPx {
  float4 c;    // colour + alpha (before fog)
  float z;     // first-hit view z (0 on a miss)
  float4 n;    // first-hit shading normal xyz, (kind << 16 | group) bits; all ones on a miss
};

__device__ __forceinline__ AaPx aaLoad(const AaSet& cs, const float4* firstN, int i) {
  AaPx p;
  p.c = cs.color[i];
  p.z = cs.viewZ[i];
  p.n = firstN[i];
  return p;
}

// adaptive_aa.hpp:66-115 (aaPairNeedsRefine) with the default thresholds:
// normal delta 0.3, depth crack 3 x the smaller outer slope or 12 output-pixel
// world sizes (times the nearer |z| in perspective).
__device__ bool aaPair(const AaPx& a, const AaPx& b, float za0, float zb1, float pws, float thr, bool persp) {
  const uint32_t ia = __float_as_uint(a.n.w), ib = __float_as_uint(b.n.w);
  if (ia != ib) return true;  // first-hit geometry or section differs (background included)
  if (fabsf(a.c.x - b.c.x) > thr || fabsf(a.c.y - b.c.y) > thr || fabsf(a.c.z - b.c.z) > thr ||
      fabsf(a.c.w - b.c.w) > thr)
    return true;
  if (ia == 0xFFFFFFFFu) return false;  // both background
  const float la = a.n.x * a.n.x + a.n.y * a.n.y + a.n.z * a.n.z;
  const float lb = b.n.x * b.n.x + b.n.y * b.n.y + b.n.z * b.n.z;
  if (la > 0.0f && lb > 0.0f) {
    const float d = a.n.x * b.n.x + a.n.y * b.n.y + a.n.z * b.n.z;
    if (1.0f - d > 0.3f) return true;
  }
  const float step = fabsf(b.z - a.z);
  const float sa = fabsf(a.z - za0);
  const float sb = fabsf(zb1 - b.z);
  const float slope = sa < sb ? sa : sb;
  float absTerm = 12.0f * pws;
  if (persp) {
    const float zmin = fabsf(a.z) < fabsf(b.z) ? fabsf(a.z) : fabsf(b.z);
    absTerm *= zmin;
  }
  const float thresh = 3.0f * slope > absTerm ? 3.0f * slope : absTerm;
  return step > thresh;
}

// adaptive_aa.hpp:128-162: an output pixel is raw-flagged when any of the
// (up to four) 4-neighbour pairs it belongs to needs refinement; outer
// slopes come from the next pixel along the axis, or the pair's own linear
// extrapolation at the border.
__global__ void k_aa_raw(int Wf, int Hf, AaSet cs, const float4* __restrict__ firstN, float pws, float thr,
                         int persp, uint8_t* __restrict__ raw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Wf * Hf) return;
  const int x = i % Wf, y = i / Wf;
  bool f = false;
  // horizontal pairs (x0, x0 + 1) with x0 = x - 1 and x0 = x
  for (int x0 = x - 1; x0 <= x && !f; ++x0) {
    if (x0 < 0 || x0 + 1 >= Wf) continue;
    const AaPx a = aaLoad(cs, firstN, y * Wf + x0), b = aaLoad(cs, firstN, y * Wf + x0 + 1);
    const float za0 = x0 > 0 ? cs.viewZ[y * Wf + x0 - 1] : 2.0f * a.z - b.z;
    const float zb1 = x0 + 2 < Wf ? cs.viewZ[y * Wf + x0 + 2] : 2.0f * b.z - a.z;
    f = aaPair(a, b, za0, zb1, pws, thr, persp != 0);
  }
  for (int y0 = y - 1; y0 <= y && !f; ++y0) {
    if (y0 < 0 || y0 + 1 >= Hf) continue;
    const AaPx a = aaLoad(cs, firstN, y0 * Wf + x), b = aaLoad(cs, firstN, (y0 + 1) * Wf + x);
    const float za0 = y0 > 0 ? cs.viewZ[(y0 - 1) * Wf + x] : 2.0f * a.z - b.z;
    const float zb1 = y0 + 2 < Hf ? cs.viewZ[(y0 + 2) * Wf + x] : 2.0f * b.z - a.z;
    f = aaPair(a, b, za0, zb1, pws, thr, persp != 0);
  }
  raw[i] = f ? 1 : 0;
}

// adaptive_aa.hpp:36-57 (3x3 dilation, border-clamped) + the pixel lists:
// a flagged pixel gets the next slot of the flagged list, a replicated one
// (when its centre is re-evaluated) the next slot of the replicated list
// (one atomic per wave each).
__global__ void k_aa_dilate(int Wf, int Hf, const uint8_t* __restrict__ raw, uint8_t* __restrict__ mask,
                            uint32_t* __restrict__ slotOf, uint32_t* __restrict__ flagList,
                            uint32_t* __restrict__ repList, int wantRep, float* __restrict__ maskOut,
                            unsigned long long* cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < Wf * Hf;
  bool m = false;
  if (valid) {
    const int x = i % Wf, y = i / Wf;
    for (int dy = -1; dy <= 1 && !m; ++dy) {
      const int yy = y + dy;
      if (yy < 0 || yy >= Hf) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int xx = x + dx;
        if (xx < 0 || xx >= Wf) continue;
        if (raw[yy * Wf + xx]) {
          m = true;
          break;
        }
      }
    }
    mask[i] = m ? 1 : 0;
    if (maskOut) maskOut[i] = m ? 1.0f : 0.0f;
  }
  const unsigned long long lt = (1ull << __lane_id()) - 1ull;
  const unsigned long long mf = __ballot(valid && m);
  const unsigned long long mr = __ballot(valid && !m && wantRep);
  uint32_t bf = 0, br = 0;
  if (__lane_id() == 0) {
    if (mf) bf = static_cast<uint32_t>(atomicAdd(&cnt[kAaFlagFill], (unsigned long long)__popcll(mf)));
    if (mr) br = static_cast<uint32_t>(atomicAdd(&cnt[kAaRepFill], (unsigned long long)__popcll(mr)));
  }
  bf = __shfl(bf, 0, 64);
  br = __shfl(br, 0, 64);
  if (!valid) return;
  if (m) {
    const uint32_t slot = bf + static_cast<uint32_t>(__popcll(mf & lt));
    flagList[slot] = static_cast<uint32_t>(i);
    slotOf[i] = slot;
  } else if (wantRep) {
    const uint32_t slot = br + static_cast<uint32_t>(__popcll(mr & lt));
    repList[slot] = static_cast<uint32_t>(i);
    slotOf[i] = slot;
  }
}

// Phase 3 store (embree_renderer.cpp:1619-1673) + fog (fog.cpp:17-36, the
// float operations of finalizePixel): every hi-res pixel takes its sample
// set's result -- the replicated centre (re-evaluated or not), the centre
// itself, a refined subpixel, or the in-order box average of its k x k fine
// samples (colour, depth and view z averaged; id and AO of the first sample).
__global__ void k_aa_compose(FrameConst fc, AaSet cs, AaSet rs, AaSet fs, const uint8_t* __restrict__ mask,
                             const uint32_t* __restrict__ slotOf, int boosted, FrameBufs out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= fc.W * fc.H) return;
  const int ss = fc.smSs, k = fc.smK, cSub = fc.smCSub;
  const int px = i % fc.W, py = i / fc.W;
  const int X = px / ss, Y = py / ss, ci = px % ss, cj = py % ss;
  const int xy = Y * fc.smWf + X;
  AaSet src;
  int64_t si;
  int n = 1;
  if (!mask[xy]) {
    src = boosted ? rs : cs;
    si = boosted ? static_cast<int64_t>(slotOf[xy]) : xy;
  } else if (k == 1) {
    if (ci == cSub && cj == cSub) {
      src = cs;
      si = xy;
    } else {
      const int cell = cj * ss + ci, centre = cSub * ss + cSub;
      src = fs;
      si = static_cast<int64_t>(slotOf[xy]) * (ss * ss - 1) + (cell > centre ? cell - 1 : cell);
    }
  } else {
    src = fs;
    n = k * k;
    si = (static_cast<int64_t>(slotOf[xy]) * ss * ss + cj * ss + ci) * n;
  }
  float4 c = src.color[si];
  float depth = src.depth[si];
  float vz = src.viewZ[si];
  if (n > 1) {
    for (int f = 1; f < n; ++f) {
      const float4 v = src.color[si + f];
      c.x += v.x;
      c.y += v.y;
      c.z += v.z;
      c.w += v.w;
      depth += src.depth[si + f];
      vz += src.viewZ[si + f];
    }
    const float inv = 1.0f / static_cast<float>(n);
    c.x *= inv;
    c.y *= inv;
    c.z *= inv;
    c.w *= inv;
    depth *= inv;
    vz *= inv;
  }
  if (fc.fogOn && vz > 0.0f) {
    const float f = fogFactor(fc, vz);
    if (fc.transparentBg) {
      c.w *= f;
    } else {
      c.x = fc.fogColor[0] * (1.0f - f) + c.x * f;
      c.y = fc.fogColor[1] * (1.0f - f) + c.y * f;
      c.z = fc.fogColor[2] * (1.0f - f) + c.z * f;
    }
  }
  out.hiColor[i] = c;
  out.depth[i] = depth;
  out.viewZ[i] = vz;
  if (out.hitId) out.hitId[i] = src.hitId[si];
  if (out.aoOpen) out.aoOpen[i] = src.aoOpen[si];
}

// The learned denoiser's input features on the render grid, from the gather
// before any denoise: log(1 + E) rgb, the receiver's albedo (its diffuse
// reflectance), the normal AOV, and the first-hit distance over the frame's
// largest (k_gi_depth_max); with stride 12 (the network's input) two noise
// descriptors follow as zeros (the gather records no per-sample spread).
// Pixels of the Wp x Hp feature grid beyond W x H (the network's padding)
// repeat the nearest edge pixel.
__global__ void k_gi_depth_max(int64_t n, const float* __restrict__ depth, unsigned int* __restrict__ maxBits) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  unsigned int v = 0u;
  if (i < n) v = __float_as_uint(fmaxf(depth[i], 0.0f));  // non-negative floats order as their bits
  for (int o = 32; o > 0; o >>= 1) v = max(v, static_cast<unsigned int>(__shfl_xor(static_cast<int>(v), o, 64)));
  if (__lane_id() == 0 && v) atomicMax(maxBits, v);
}
__global__ void k_gi_features(int W, int H, int Wp, int Hp, const float4* __restrict__ E,
                              const float4* __restrict__ refl, const float4* __restrict__ nrm,
                              const float* __restrict__ depth, const unsigned int* __restrict__ maxBits,
                              int stride, float* __restrict__ feat) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<int64_t>(Wp) * Hp) return;
  const int x = min(static_cast<int>(i % Wp), W - 1), y = min(static_cast<int>(i / Wp), H - 1);
  const int64_t q = static_cast<int64_t>(y) * W + x;
  const float4 e = E[q], a = refl[q], nn = nrm[q];
  const bool recv = nn.w != 0.0f;
  const float dmax = __uint_as_float(*maxBits);
  float* f = feat + i * stride;
  f[0] = log1pf(fmaxf(e.x, 0.0f));
  f[1] = log1pf(fmaxf(e.y, 0.0f));
  f[2] = log1pf(fmaxf(e.z, 0.0f));
  f[3] = recv ? a.x : 0.0f;
  f[4] = recv ? a.y : 0.0f;
  f[5] = recv ? a.z : 0.0f;
  f[6] = nn.x;
  f[7] = nn.y;
  f[8] = nn.z;
  f[9] = dmax > 0.0f ? depth[q] / dmax : 0.0f;
  if (stride == 12) {
    f[10] = 0.0f;
    f[11] = 0.0f;
  }
}
// the network's log(1 + E) back onto the receivers' E (the gather's hit
// fraction in E.w is kept)
__global__ void k_gi_apply_net(int W, int H, int Wp, const float* __restrict__ pred, const float4* __restrict__ nrm,
                               float4* __restrict__ E) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<int64_t>(W) * H) return;
  if (nrm[i].w == 0.0f) return;
  const int x = static_cast<int>(i % W), y = static_cast<int>(i / W);
  const float* p = pred + (static_cast<int64_t>(y) * Wp + x) * 3;
  float4 e = E[i];
  e.x = fmaxf(expm1f(p[0]), 0.0f);
  e.y = fmaxf(expm1f(p[1]), 0.0f);
  e.z = fmaxf(expm1f(p[2]), 0.0f);
  E[i] = e;
}

// insideAnySolid (edges/object_space_edges.cpp:128-146) for the clip's sample
// points: whether the point lies strictly inside a sphere or a cylinder
// (flat caps) other than its own source, a thinner cylinder never clipping
// a sphere ring (ringRs > 0). Only the solids listed in the point's grid cell
// are tested; every solid whose box holds the point is listed there.
__global__ void k_solids_contain(int64_t n, const float4* __restrict__ pts, const int2* __restrict__ src,
                                 const float4* __restrict__ sph, const float4* __restrict__ cyl,
                                 const uint32_t* __restrict__ cellStart, const uint32_t* __restrict__ cellItems,
                                 float lox, float loy, float loz, float inv, int dx, int dy, int dz,
                                 uint8_t* __restrict__ inside) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 q = pts[i];
  const int2 sk = src[i];
  const V3 X = mk(q.x, q.y, q.z);
  const float ringRs = q.w;
  auto cellOf = [](float v, float lo, float iv, int dim) {
    const float f = floorf((v - lo) * iv);
    return min(max(static_cast<int>(f), 0), dim - 1);
  };
  const int cx = cellOf(X.x, lox, inv, dx), cy = cellOf(X.y, loy, inv, dy), cz = cellOf(X.z, loz, inv, dz);
  const uint32_t c = (static_cast<uint32_t>(cz) * static_cast<uint32_t>(dy) + static_cast<uint32_t>(cy)) *
                         static_cast<uint32_t>(dx) + static_cast<uint32_t>(cx);
  bool in = false;
  for (uint32_t k = cellStart[c], e = cellStart[c + 1]; k < e && !in; ++k) {
    const uint32_t item = cellItems[k];
    if (!(item & 0x80000000u)) {
      if (static_cast<int>(item) == sk.x) continue;
      const float4 sp = sph[item];
      in = len(X - mk(sp.x, sp.y, sp.z)) < sp.w;
    } else {
      const uint32_t j = item & 0x7FFFFFFFu;
      if (static_cast<int>(j) == sk.y) continue;
      const float4 a = cyl[2 * j], b4 = cyl[2 * j + 1];
      if (ringRs > 0.0f && a.w < ringRs - 1.0e-4f) continue;
      const V3 p0 = mk(a.x, a.y, a.z);
      const V3 ab = mk(b4.x, b4.y, b4.z) - p0;
      const float L = len(ab);
      if (L <= 0.0f) continue;
      const V3 u = ab * (1.0f / L);
      const float t = dot(X - p0, u);
      if (t < 0.0f || t > L) continue;
      const V3 perp = (X - p0) - u * t;
      in = len(perp) < a.w;
    }
  }
  inside[i] = in ? 1 : 0;
}
__global__ void k_zero_u32(unsigned int* c, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = 0u;
}
__global__ void k_zero(unsigned long long* c, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = 0ull;
}
// Zeroes the hit-queue fills, continuation-queue fills and AO cursors.
__global__ void k_reset_queues(unsigned long long* c) {
  if (threadIdx.x < 3 * kQueueParts) c[kHitFill + kLineWords * threadIdx.x] = 0ull;
}

static inline unsigned blocksFor(int64_t n, int bs) {
  return static_cast<unsigned>((n + bs - 1) / bs);
}
void launchGiFeatures(int W, int H, int Wp, int Hp, const FrameBufs& b, unsigned int* maxBits, int stride,
                      float* feat, hipStream_t st) {
  const int64_t n = static_cast<int64_t>(W) * H;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, st, maxBits, 1);
  hipLaunchKernelGGL(k_gi_depth_max, dim3(blocksFor(n, 256)), dim3(256), 0, st, n, b.depth, maxBits);
  const int64_t np = static_cast<int64_t>(Wp) * Hp;
  hipLaunchKernelGGL(k_gi_features, dim3(blocksFor(np, 256)), dim3(256), 0, st, W, H, Wp, Hp, b.giE, b.giRefl,
                     b.giNrm, b.depth, maxBits, stride, feat);
}
void launchGiApplyNet(int W, int H, int Wp, const float* pred, const FrameBufs& b, hipStream_t st) {
  const int64_t n = static_cast<int64_t>(W) * H;
  hipLaunchKernelGGL(k_gi_apply_net, dim3(blocksFor(n, 256)), dim3(256), 0, st, W, H, Wp, pred, b.giNrm, b.giE);
}
void launchSolidsContain(int64_t n, const float4* pts, const int2* src, const float4* sph, const float4* cyl,
                         const uint32_t* cellStart, const uint32_t* cellItems, const float lo[3], float inv,
                         const int dim[3], uint8_t* inside, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_solids_contain, dim3(static_cast<unsigned>(blocksFor(n, 256))), dim3(256), 0, st, n, pts,
                     src, sph, cyl, cellStart, cellItems, lo[0], lo[1], lo[2], inv, dim[0], dim[1], dim[2], inside);
}


void launchZeroCounters(unsigned long long* c, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_zero, dim3(blocksFor(n, 64)), dim3(64), 0, st, c, n);
}
void launchResetQueues(unsigned long long* c, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_queues, dim3(1), dim3(64), 0, st, c);
}
#define UBG_DISPATCH(KERNEL, GRID, COUNT, ...)                                                  \
  do {                                                                                          \
    const size_t lds = stackLdsBytes(s, kTraceBlock);                                           \
    if (s.layout == kLayoutBvh8) {                                                              \
      if (COUNT)                                                                                \
        hipLaunchKernelGGL((KERNEL<true, kLayoutBvh8>), dim3(GRID), dim3(kTraceBlock), lds, st, __VA_ARGS__); \
      else                                                                                      \
        hipLaunchKernelGGL((KERNEL<false, kLayoutBvh8>), dim3(GRID), dim3(kTraceBlock), lds, st, __VA_ARGS__); \
    } else {                                                                                    \
      if (COUNT)                                                                                \
        hipLaunchKernelGGL((KERNEL<true, kLayoutBvh2>), dim3(GRID), dim3(kTraceBlock), lds, st, __VA_ARGS__); \
      else                                                                                      \
        hipLaunchKernelGGL((KERNEL<false, kLayoutBvh2>), dim3(GRID), dim3(kTraceBlock), lds, st, __VA_ARGS__); \
    }                                                                                           \
  } while (0)

// Compute units of the current device (cached per device).
static int computeUnits() {
  static std::atomic<int> cu[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int n = cu[dev].load(std::memory_order_relaxed);
  if (n == 0) {
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cu[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

int queuePartCap(int W, int H) {
  const int64_t threads = static_cast<int64_t>((W + 7) / 8) * ((H + 7) / 8) * 64;
  const int64_t blocks = (threads + kTraceBlock - 1) / kTraceBlock;
  return static_cast<int>((blocks + kQueueParts - 1) / kQueueParts * kTraceBlock);
}

void launchPrimary(const FrameConst& fc, const SceneView& s, const FrameBufs& b, int iter,
                   const PartFill& fill, bool countVisits, hipStream_t st) {
  if (iter == 0) {
    int64_t n;
    if (fc.smMode == kSmGrid)  // one 64-lane wave per 8x8 tile of the hi-res grid
      n = static_cast<int64_t>((fc.W + 7) / 8) * ((fc.H + 7) / 8) * 64;
    else if (fc.smMode == kSmCenter && b.smList == nullptr)  // ... of the output grid
      n = static_cast<int64_t>((fc.smWf + 7) / 8) * ((fc.smHf + 7) / 8) * 64;
    else
      n = fc.nList;
    if (n <= 0) return;
    UBG_DISPATCH(k_primary, blocksFor(n, kTraceBlock), countVisits, fc, s, b, iter, fill, 1);
    return;
  }
  uint32_t mx = 0;
  for (int p = 0; p < kQueueParts; ++p) mx = std::max(mx, fill.n[p]);
  if (mx == 0) return;
  const int partBlocks = static_cast<int>(blocksFor(mx, kTraceBlock));
  UBG_DISPATCH(k_primary, partBlocks * kQueueParts, countVisits, fc, s, b, iter, fill, partBlocks);
}
void launchAo(const FrameConst& fc, const SceneView& s, const FrameBufs& b, bool countVisits, hipStream_t st) {
  if (fc.aoSamples <= 0) return;
  const int64_t slots = static_cast<int64_t>(kQueueParts) * b.partCap;
  if (fc.aoEnh) {
    UBG_DISPATCH(k_ao_quality, blocksFor(slots * fc.aoSamples, kTraceBlock), countVisits, fc, s, b);
    return;
  }
  if (s.layout == kLayoutBvh8) {
    // persistent grid: every wave slot of the chip, no more than the work
    const int64_t rays = slots * fc.aoSamples;
    const int64_t resident = (int64_t)computeUnits() * 4 * kAoWavesPerSimd * 64 / kTraceBlock;
    const int64_t blocks = std::min<int64_t>(blocksFor(rays, kTraceBlock), resident);
    const size_t lds = stackLdsBytes(s, kTraceBlock);
    if (countVisits)
      hipLaunchKernelGGL(k_ao_persistent<true>, dim3(blocks), dim3(kTraceBlock), lds, st, fc, s, b);
    else
      hipLaunchKernelGGL(k_ao_persistent<false>, dim3(blocks), dim3(kTraceBlock), lds, st, fc, s, b);
    return;
  }
  UBG_DISPATCH(k_ao, blocksFor(slots * fc.aoLanes, kTraceBlock), countVisits, fc, s, b, 0);
}

// ---- batched frames ----
size_t traceLdsBytes(const SceneView& s) { return stackLdsBytes(s, kTraceBlock); }
int64_t aoResidentBlocks(int wavesPerSimd) {
  return static_cast<int64_t>(computeUnits()) * 4 * wavesPerSimd * 64 / kTraceBlock;
}
void launchZeroBatch(const BatchItem* items, int n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_batch, dim3(8, static_cast<unsigned>(n)), dim3(256), 0, st, items, kCounterWords);
}
void launchResetBatch(const BatchItem* items, const BatchLaunch* L, int n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_reset_batch, dim3(static_cast<unsigned>(n)), dim3(64), 0, st, items, L);
}
void launchPrimaryBatch(const BatchItem* items, const BatchLaunch* L, int64_t blocks, int iter, size_t lds,
                        hipStream_t st) {
  if (blocks <= 0) return;
  hipLaunchKernelGGL(k_primary_batch, dim3(static_cast<unsigned>(blocks)), dim3(kTraceBlock), lds, st, items, L,
                     iter);
}
void launchAoBatch(const BatchItem* items, const BatchLaunch* L, int64_t blocks, size_t lds, hipStream_t st) {
  if (blocks <= 0) return;
  hipLaunchKernelGGL(k_ao_batch, dim3(static_cast<unsigned>(blocks)), dim3(kTraceBlock), lds, st, items, L);
}
void launchResolveBatch(const BatchItem* items, const BatchLaunch* L, int64_t blocks, int iter, size_t lds,
                        hipStream_t st) {
  if (blocks <= 0) return;
  hipLaunchKernelGGL(k_resolve_batch, dim3(static_cast<unsigned>(blocks)), dim3(kTraceBlock), lds, st, items, L,
                     iter);
}
void launchPostBatch(const BatchItem* items, const BatchLaunch* L, int64_t blocks, hipStream_t st) {
  if (blocks <= 0) return;
  hipLaunchKernelGGL(k_post_batch, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, items, L);
}
void launchFillsBatch(const BatchItem* items, int n, uint32_t* fills, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fills_batch, dim3(static_cast<unsigned>(n)), dim3(64), 0, st, items, fills);
}
void launchStatsBatch(const BatchItem* items, int n, unsigned long long* stats, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_stats_batch, dim3(static_cast<unsigned>(n)), dim3(64), 0, st, items, stats);
}
void launchResolve(const FrameConst& fc, const SceneView& s, const FrameBufs& b, int iter, bool countVisits,
                   hipStream_t st) {
  const int64_t slots = static_cast<int64_t>(kQueueParts) * b.partCap;
  // the principled evaluator is compiled into a kernel of its own, used only
  // for scenes that have a principled material
  if (b.principled)
    UBG_DISPATCH(k_resolve_pbr, blocksFor(slots, kTraceBlock), countVisits, fc, s, b, iter);
  else
    UBG_DISPATCH(k_resolve, blocksFor(slots, kTraceBlock), countVisits, fc, s, b, iter);
}
void launchGiList(const GiField& f, const uint8_t* mask, bool zeroOthers, const FrameBufs& b, hipStream_t st) {
  // (the same block -> partition map as a first primary launch over W x H:
  // every grid here has at most W x H pixels, so partitions never overflow)
  const int64_t n = static_cast<int64_t>(f.w) * f.h;
  hipLaunchKernelGGL(k_gi_list, dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), 0, st, f, mask,
                     zeroOthers ? 1 : 0, b);
}
void launchGiGbuffer(const FrameConst& fc, const SceneView& s, const GiField& f, float4* albedo,
                     unsigned long long* counters, hipStream_t st) {
  const int64_t n = static_cast<int64_t>(f.w) * f.h;
  const size_t lds = stackLdsBytes(s, kTraceBlock);
  if (s.layout == kLayoutBvh8)
    hipLaunchKernelGGL(k_gi_gbuf<kLayoutBvh8>, dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), lds, st, fc, s, f,
                       albedo, counters);
  else
    hipLaunchKernelGGL(k_gi_gbuf<kLayoutBvh2>, dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), lds, st, fc, s, f,
                       albedo, counters);
}
void launchGiGather(const FrameConst& fc, const SceneView& s, const FrameBufs& b, const GiField& f, bool countVisits,
                    hipStream_t st) {
  const int64_t slots = static_cast<int64_t>(kQueueParts) * b.partCap;
  UBG_DISPATCH(k_gi_gather, blocksFor(slots, kTraceBlock), countVisits, fc, s, b, f);
}
void launchGiDenoise(const GiField& f, int iters, float sigmaZ, float sigmaN, float sigmaL, float4* irrA,
                     float4* irrB, float* grad, float* var, hipStream_t st) {
  iters = iters < 0 ? 0 : iters;
  if (iters == 0) return;  // denoise_atrous.cpp:32-33
  const int n = f.w * f.h;
  const unsigned g = blocksFor(n, 256);
  sigmaZ = sigmaZ > 1.0e-4f ? sigmaZ : 1.0e-4f;
  sigmaL = sigmaL > 1.0e-4f ? sigmaL : 1.0e-4f;
  hipLaunchKernelGGL(k_atrous_scrub, dim3(g), dim3(256), 0, st, n, f.E, irrA);
  hipLaunchKernelGGL(k_atrous_stats, dim3(g), dim3(256), 0, st, f.w, f.h, f.nrm, f.pos, irrA, grad, var);
  float4* in = irrA;
  float4* out = irrB;
  for (int it = 0; it < iters; ++it) {
    hipLaunchKernelGGL(k_atrous_iter, dim3(g), dim3(256), 0, st, f.w, f.h, 1 << it, sigmaZ, sigmaN, sigmaL, f.nrm,
                       f.pos, in, grad, var, out);
    float4* t = in;
    in = out;
    out = t;
  }
  hipLaunchKernelGGL(k_atrous_store, dim3(g), dim3(256), 0, st, n, f.nrm, in, f.E);
}
void launchGiUpsample(const FrameConst& fc, const FrameBufs& b, const GiField& low, float normalPow,
                      float depthScale, int fallbackRadius, uint8_t* patch, float patchThresh, hipStream_t st) {
  const int64_t n = static_cast<int64_t>(fc.W) * fc.H;
  hipLaunchKernelGGL(k_gi_upsample, dim3(blocksFor(n, 256)), dim3(256), 0, st, fc, b, low, normalPow, depthScale,
                     fallbackRadius, patch, patchThresh);
}
void launchGiSmoothPatched(const FrameConst& fc, const FrameBufs& b, const uint8_t* patch, float normalPow,
                           float depthScale, const float4* Esrc, hipStream_t st) {
  const int64_t n = static_cast<int64_t>(fc.W) * fc.H;
  hipLaunchKernelGGL(k_gi_smooth_patched, dim3(blocksFor(n, 256)), dim3(256), 0, st, fc, b, patch, normalPow,
                     depthScale, Esrc);
}
void launchGiComposite(const FrameConst& fc, const FrameBufs& b, float* indirect, float* occlusion, hipStream_t st) {
  const int64_t n = static_cast<int64_t>(fc.W) * fc.H;
  hipLaunchKernelGGL(k_gi_composite, dim3(blocksFor(n, 256)), dim3(256), 0, st, fc, b, indirect, occlusion);
}
void launchAaMask(const FrameConst& fc, const AaSet& centers, const float4* firstN, float pixelWorldSize,
                  float threshold, uint8_t* raw, uint8_t* mask, uint32_t* slotOf, uint32_t* flagList,
                  uint32_t* repList, bool wantRep, float* maskOut, unsigned long long* counters, hipStream_t st) {
  const int n = fc.smWf * fc.smHf;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_aa_raw, dim3(blocksFor(n, 256)), dim3(256), 0, st, fc.smWf, fc.smHf, centers, firstN,
                     pixelWorldSize, threshold, fc.ortho ? 0 : 1, raw);
  hipLaunchKernelGGL(k_aa_dilate, dim3(blocksFor(n, 256)), dim3(256), 0, st, fc.smWf, fc.smHf, raw, mask, slotOf,
                     flagList, repList, wantRep ? 1 : 0, maskOut, counters);
}
void launchAaCompose(const FrameConst& fc, const AaSet& centers, const AaSet& rep, const AaSet& cells,
                     const uint8_t* mask, const uint32_t* slotOf, bool boosted, const FrameBufs& out,
                     hipStream_t st) {
  const int64_t n = static_cast<int64_t>(fc.W) * fc.H;
  hipLaunchKernelGGL(k_aa_compose, dim3(blocksFor(n, 256)), dim3(256), 0, st, fc, centers, rep, cells, mask, slotOf,
                     boosted ? 1 : 0, out);
}
void launchPost(const float4* hi, int W, int H, int ss, float gamma, float4* out, hipStream_t st) {
  const int n = (W / ss) * (H / ss);
  if (n <= 0) return;
  hipLaunchKernelGGL(k_post, dim3(blocksFor(n, 256)), dim3(256), 0, st, hi, W, H, ss, gamma, out);
}
void launchQuery(const SceneView& s, const void* rays, int64_t n, void* hits, uint8_t* occ,
                 bool any, hipStream_t st) {
  if (n <= 0) return;
  const size_t lds = stackLdsBytes(s, kTraceBlock);
  if (s.layout == kLayoutBvh8)
    hipLaunchKernelGGL(k_query<kLayoutBvh8>, dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), lds, st, s,
                       static_cast<const RayIn*>(rays), n, static_cast<HitOut*>(hits), occ, any ? 1 : 0);
  else
    hipLaunchKernelGGL(k_query<kLayoutBvh2>, dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), lds, st, s,
                       static_cast<const RayIn*>(rays), n, static_cast<HitOut*>(hits), occ, any ? 1 : 0);
}

void launchCoarseGbuf(const FrameConst& fc, const SceneView& s, const FrameBufs& b, bool countVisits,
                      hipStream_t st) {
  UBG_DISPATCH(k_coarse_gbuf, blocksFor(static_cast<int64_t>(fc.cW) * fc.cH, kTraceBlock), countVisits, fc, s, b);
}
void launchCoarseReduce(const FrameConst& fc, const FrameBufs& b, hipStream_t st) {
  const int64_t slots = static_cast<int64_t>(kQueueParts) * b.partCap;
  hipLaunchKernelGGL(k_coarse_reduce, dim3(blocksFor(slots, kTraceBlock)), dim3(kTraceBlock), 0, st, fc, b);
}

void launchSegments(const SceneView& s, const void* segs, int64_t n, const int32_t* excl, int64_t nExcl, float eps,
                    float grazeCosEps, float coplanarEps, uint8_t* occ, hipStream_t st) {
  if (n <= 0) return;
  const size_t lds = stackLdsBytes(s, kTraceBlock);
  if (s.layout == kLayoutBvh8)
    hipLaunchKernelGGL(k_segments<kLayoutBvh8>, dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), lds, st, s,
                       static_cast<const SegIn*>(segs), n, excl, nExcl, eps, grazeCosEps, coplanarEps, occ);
  else
    hipLaunchKernelGGL(k_segments<kLayoutBvh2>, dim3(blocksFor(n, kTraceBlock)), dim3(kTraceBlock), lds, st, s,
                       static_cast<const SegIn*>(segs), n, excl, nExcl, eps, grazeCosEps, coplanarEps, occ);
}

}  // namespace ubg

// ---------------------------------------------------------------------------
// Host test hook (ub_testing.h): the traversal above, run on the CPU.
#include <string>
#include <vector>

#include "ub_build.h"
#include "ub_testing.h"

extern "C" int ubgx_host_query(const ubg_scene_desc* desc, int32_t bvh_layout, int64_t n_rays, const ubg_ray* rays,
                               ubg_hit* hits, uint8_t* occ, uint64_t* stats) {
  using namespace ubg;
  if (!desc || (n_rays > 0 && !rays)) return UBG_ERR_INVALID_ARGUMENT;
  HostScene hs;
  std::string err;
  const int layout = bvh_layout == UBG_BVH2 ? kLayoutBvh2 : kLayoutBvh8;
  const int st = buildHostScene(desc, layout, hs, err);
  if (st != UBG_OK) return st;
  SceneView v{};
  v.nodes = hs.nodes.data();
  v.prims = hs.prims.data();
  v.nNodes = static_cast<int32_t>(hs.nNodes);
  v.nPrims = static_cast<int32_t>(hs.prims.size() / 3);
  v.layout = hs.layout;
  v.stackDepth = hs.stackDepth;
  v.rndPlanes = hs.rndPlanes.data();
  v.nodeKinds = hs.nodeKinds.empty() ? nullptr : hs.nodeKinds.data();
  std::vector<uint32_t> s0(static_cast<size_t>(hs.stackDepth) + 1), s1(static_cast<size_t>(hs.stackDepth) + 1);
  uint64_t visits = 0, tests = 0;
  for (int64_t i = 0; i < n_rays; ++i) {
    const ubg_ray& r = rays[i];
    const V3 o = mk(r.org[0], r.org[1], r.org[2]);
    const V3 d = mk(r.dir[0], r.dir[1], r.dir[2]);
    Hit h{};
    uint32_t nv = 0, nt = 0;
    if (hits) {
      const bool f = layout == kLayoutBvh8
                         ? traverse8<false, true>(v, o, d, r.tnear, r.tfar, h, s0.data(), s1.data(), 1, &nv, &nt)
                         : traverse2<false, true>(v, o, d, r.tnear, r.tfar, h, s0.data(), 1, &nv, &nt);
      visits += nv;
      tests += nt;
      if (layout == kLayoutBvh8 && v.nNodes > 0) {
        // the resumable closest-hit steps (as the kernels run them) must find
        // the same hit as the one-shot traversal
        Trav8 t;
        trav8Init(t, o, d, r.tnear, r.tfar);
        Hit h2{};
        bool f2 = false;
        uint32_t a = 0, b = 0;
        for (;;) {
          const int rs = t.tHits != 0u ? trav8PrimClosest<false>(v, t, h2, f2, &b)
                                       : trav8Node<false>(v, t, s0.data(), s1.data(), 1, &a);
          if (rs != 0) break;
        }
        if (f2 != f || (f && (f2u(h2.t) != f2u(h.t) || h2.bits != h.bits))) return UBG_ERR_INTERNAL;
      }
      ubg_hit& out = hits[i];
      if (f) {
        out.t = h.t;
        out.u = h.u;
        out.v = h.v;
        out.kind = h.bits >> kKindShift;
        out.prim_id = h.bits & kIdMask;
        out.ng[0] = h.ng.x;
        out.ng[1] = h.ng.y;
        out.ng[2] = h.ng.z;
      } else {
        out.t = __builtin_inff();
        out.u = out.v = 0.0f;
        out.kind = 0xFFFFFFFFu;
        out.prim_id = 0xFFFFFFFFu;
        out.ng[0] = out.ng[1] = out.ng[2] = 0.0f;
      }
    }
    if (occ) {
      uint32_t a = 0, b = 0;
      if (layout == kLayoutBvh8) {
        // the one-shot any-hit traversal and the resumable step machine must
        // agree; a disagreement is reported as 2
        const bool o1 = traverse8<true, false>(v, o, d, r.tnear, r.tfar, h, s0.data(), s1.data(), 1, &a, &b);
        bool o2 = false, o3 = false;
        if (v.nNodes > 0) {
          Trav8 t;
          trav8Init(t, o, d, r.tnear, r.tfar);
          int rs = 0;
          while ((rs = trav8StepAny<false>(v, t, s0.data(), s1.data(), 1, &a, &b)) == 0) {
          }
          o2 = rs > 0;
          // the AO kernel's order: primitives by kind (cones first here, to
          // differ from the node order), nodes with their kind planes
          trav8Init(t, o, d, r.tnear, r.tfar);
          for (;;) {
            if (t.tHits != 0u) {
              rs = trav8KindMask(t, 3u)   ? trav8PrimKind<3, false>(v, t, &b)
                   : trav8KindMask(t, 1u) ? trav8PrimKind<1, false>(v, t, &b)
                   : trav8KindMask(t, 2u) ? trav8PrimKind<2, false>(v, t, &b)
                                          : trav8PrimKind<0, false>(v, t, &b);
            } else {
              rs = trav8Node<false, true>(v, t, s0.data(), s1.data(), 1, &a);
            }
            if (rs != 0) break;
          }
          o3 = rs > 0;
        }
        occ[i] = (o1 == o2 && o1 == o3) ? (o1 ? 1 : 0) : 2;
      } else {
        occ[i] = traverse2<true, false>(v, o, d, r.tnear, r.tfar, h, s0.data(), 1, &a, &b) ? 1 : 0;
      }
    }
  }
  if (stats) {
    stats[0] = visits;
    stats[1] = tests;
    stats[2] = static_cast<uint64_t>(hs.depth);
    stats[3] = static_cast<uint64_t>(hs.nNodes);
    stats[4] = static_cast<uint64_t>(hs.stackDepth);
    stats[5] = static_cast<uint64_t>(hs.nodeBytes);
  }
  return UBG_OK;
}
