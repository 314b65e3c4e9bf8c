// ub_layout.h -- HBM layout of a built scene and the per-frame constants,
// shared by the host builder (ub_build.cpp), the C-ABI (ub_capi.cpp) and the
// gfx950 kernels (ub_kernels.hip).
//
// Scene layout (all arrays 16-byte aligned, structure-of-arrays by role):
//   nodes  : BVH2, 64 B per inner node = 4 x float4
//              [0] = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)
//              [1] = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//              [2] = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)
//              [3] = (child0, child1, -, -) as int; child >= 0 is an inner
//                    node, child < 0 is a leaf ~((first << 3) | (count - 1))
//   prims  : 48 B per primitive = 3 x float4, in leaf order, so a leaf is one
//            contiguous read. The last word holds (kind << 30) | primID.
//              sphere : (cx, cy, cz, r) | - | (-, -, -, bits)
//              cone   : (ax, ay, az, ra) | (bx, by, bz, rb) | (L, 1/L, (rb-ra)/L, bits)
//              round  : (ax, ay, az, ra) | (bx, by, bz, rb) | (-, -, -, bits)
//              tri    : (v0x, v0y, v0z, v1x) | (v1y, v1z, v2x, v2y) | (v2z, -, -, bits)
//   shading side tables, indexed by the per-kind primID (read once per hit,
//   never during traversal): color float4, info uint32 (material index in
//   bits 0..23, fromEdge in bit 31), opacity1 float; round curves add two
//   float4 joint planes (nL.xyz, hasL) and (nR.xyz, hasR); triangles carry
//   uint4 (i0, i1, i2, material) into per-vertex normal/color float4 arrays.
#pragma once

#include <stdint.h>

namespace ubg {

constexpr int kMaxDepth = 32;   // BVH2 builder guarantee; = traversal stack entries
constexpr int kMaxLeaf = 8;     // BVH2 primitives per leaf (3 count bits)
constexpr int kMaxLeaf8 = 3;    // BVH8 primitives per leaf slot (unary count in meta)
constexpr int kMaxLights = 16;  // lights carried in the kernel argument block

// Acceleration-structure layouts (ubg_scene_create_ex).
constexpr int kLayoutBvh2 = 0;  // full-precision BVH2, 64 B nodes
constexpr int kLayoutBvh8 = 1;  // compressed 8-wide BVH, 80 B nodes, 8-bit child boxes

constexpr uint32_t kKindShift = 30;
constexpr uint32_t kIdMask = (1u << 30) - 1u;
constexpr uint32_t kInfoFromEdge = 1u << 31;
constexpr uint32_t kInfoMatMask = (1u << 24) - 1u;

// Shading-ready material (precomputed Blinn exponent). 48 bytes.
struct DevMaterial {
  float ambient, diffuse, specular, brilliance;
  float phong, phongSize, reflection, emission;
  float blinnExp;
  uint32_t metallic;
  float roughness;
  uint32_t toon;  // Material::toonLike() (scene.hpp:190-196)
  // ShadingModel::Principled (scene.hpp:127-171): Material::Pbr, with the
  // anisotropy rotation's cosine and sine (angle = turns * 2 pi) taken on the
  // host once per material
  uint32_t principled;
  float pbrMetallic, pbrRoughness, pbrSpecular, pbrAnisotropy, pbrRotation;
  float rotCos, rotSin;
};

// Material::diffuseWeight (scene.hpp:171-174): the diffuse-lobe weight every
// diffuse consumer reads (GI reflectance, gather albedo)
__host__ __device__ inline float diffuseWeight(const DevMaterial& m) {
  return m.principled ? m.diffuse * (1.0f - m.pbrMetallic) : m.diffuse;
}

// One BVH8 node as the builders write it (80 bytes; the traversal reads it
// as 5 x uint4): box origin, per-axis scale exponents (scale = 2^(e - 127)),
// inner-child mask, first inner child, first leaf primitive, per-slot meta
// (0 empty; leaf: unary count << 5 | offset; inner: 1 << 5 | (24 + slot)),
// quantized child planes lo/hi per axis and slot.
struct Node8 {
  float p[3];
  uint8_t e[3];
  uint8_t imask;
  uint32_t childBase, primBase;
  uint8_t meta[8];
  uint8_t qlo[3][8];
  uint8_t qhi[3][8];
};
static_assert(sizeof(Node8) == 80, "BVH8 node is 80 bytes");

struct SceneView {
  const void* nodes;      // BVH2: float4 x 4 per inner node; BVH8: uint4 x 5 per node
  const void* nodeKinds;  // BVH8: uint2 per node, the kinds of its leaf primitives as two
                          // bit planes (bit i of .x / .y = bit 0 / 1 of the kind of primitive
                          // primBase + i); null for BVH2
  const void* prims;      // float4 x 3 per primitive, in the layout's leaf order
  int32_t nNodes;
  int32_t nPrims;
  int32_t layout;         // kLayoutBvh2 / kLayoutBvh8
  int32_t stackDepth;     // traversal stack entries the tree needs
  const void* sphColor;   // float4
  const uint32_t* sphInfo;
  const void* coneColor;  // float4
  const uint32_t* coneInfo;
  const float* coneOp1;
  const void* rndColor;   // float4
  const uint32_t* rndInfo;
  const float* rndOp1;
  const void* rndPlanes;  // float4 x 2
  const void* triIdx;     // uint4
  const void* vtxNrm;     // float4
  const void* vtxCol;     // float4
  const DevMaterial* primMats;
  const DevMaterial* meshMats;
  int64_t meshBaseTris;   // triangles per mesh instance copy (0 = no mesh)
  // section group per primitive (Sphere/Cylinder::group, Mesh::triGroupId by
  // base triangle), read only for the adaptive-AA first-hit record
  const uint16_t* sphGroup;
  const uint16_t* coneGroup;
  const uint16_t* rndGroup;
  const uint16_t* triGroup;
  // unit p0 -> p1 axis per cone / round segment (float4), the anisotropy
  // tangent of principled cylinders; null unless an anisotropic principled
  // cylinder exists
  const void* coneAxis;
  const void* rndAxis;
};

struct DevLight {
  float L[3];
  float color[3];
  float radius;    // radians
  float tanRadius;
  int32_t highlight;
  int32_t soft;    // shadowSamples > 1 && radius > 0
  int32_t pad[2];
};

// Per-frame constants, passed by value in the kernel argument block.
struct FrameConst {
  int32_t W, H;            // hi-res grid
  int32_t ortho;
  float pos[3], dir[3], right[3], up[3];
  float halfW, halfH, persHalfW, persHalfH;
  float clipNear, clipFar; // view-z planes; clipOn == 0 => [0, inf)
  int32_t clipOn;
  float bg[3];
  float amb[3];
  int32_t transparentBg;
  int32_t transparency;
  int32_t maxIters;
  int32_t aoSamples;
  float aoDist;
  float aoIntensity;
  int32_t aoLanes;         // lanes per hit in the AO kernel (power of two <= 64)
  int32_t shadows;
  int32_t shadowSamples;
  float specScale;
  int32_t fogOn;
  float fogColor[3];
  float fogStart, fogEnd;
  int32_t nLights;
  // enhanced AO (nearest-hit estimator); aoEnh == 0 = legacy binary AO
  int32_t aoEnh;
  float aoFalloff;
  int32_t aoMultiScale, aoBent, aoMultibounce, aoLowDisc;
  float aoDiffuse;
  float aoSky[3], aoGround[3], aoUpAxis[3];
  // coarse-grid AO: grid cW x cH (coarse != 0); seedW = width of the AO seed
  // lattice (base = px + py * seedW): the hi-res W, or cW for the grid gather
  int32_t coarse, cW, cH, seedW;
  // Ray lattice of the pass: primary rays go through the centres of a latW x
  // latH grid over the image plane (the hi-res W x H, or W*k x H*k for the
  // adaptive-AA fine samples), and (px, py) on it seed the AO / shadow
  // streams with seedW = latW. The sample map (smMode) turns a pass's output
  // index into its lattice position; see kSm* below.
  int32_t latW, latH;
  int32_t smMode;          // kSmGrid / kSmCenter / kSmCells / kSmFine
  int32_t smSs, smK, smCSub, smWf, smHf, smPer;
  uint32_t smBase;         // first sample of this chunk of the pass
  uint32_t nList;          // samples in this chunk (list modes)
  // pt1 GI (gi != 0): gather paths per pixel, path length, ray reach, seed
  // mix hashU32(pt1Seed), stratified first bounce, luminance clamp, and the
  // environment: radiance = envScale * (ground + (sky - ground) * w) with
  // w = 0.5 (dot(wi, aoUpAxis) + 1), envScale = giEnvIntensity * ambient
  int32_t gi, giSpp, giBounces, giLd;
  int32_t giPt2, giPattern, giEmissive;  // pt2: Sobol-Burley first bounce, area-light NEE, emission
  float giMaxDist, giIntensity, giClamp;
  uint32_t giSeedMix;
  float giEnvScale[3], giSky[3], giGround[3];
  DevLight lights[kMaxLights];
};

// Sample maps of a pass (output index s = smBase + local):
// * kSmGrid  : s = py * W + px on the hi-res lattice (the full grid);
// * kSmCenter: s -> output pixel XY (= s, or smList[s]); lattice position
//              (X * ss + cSub, Y * ss + cSub), cSub = ss / 2;
// * kSmCells : the ss^2 - 1 non-centre subpixels of flagged output pixel
//              smList[s / smPer] (smPer = ss^2 - 1), in row order;
// * kSmFine  : the k x k fine samples of every subpixel of flagged pixel
//              smList[s / smPer] (smPer = ss^2 k^2), subpixel-major, row order
//              inside; lattice W*k x H*k.
constexpr int kSmGrid = 0;
constexpr int kSmCenter = 1;
constexpr int kSmCells = 2;
constexpr int kSmFine = 3;

// Frame-time counters (u64), in one buffer of kCounterWords words:
// * statistics (ray and visit counts): kStatSpread groups of kNumCounters
//   slots, each group on its own 128-byte line; a wave adds into group
//   (block % kStatSpread) and the host sums the groups, so no single address
//   takes an atomic from every wave of a launch;
// * queue fills: the hit queue and the continuation queue are each split
//   into kQueueParts partitions (a producing block appends to partition
//   block % kQueueParts), each fill counter on its own line;
// * AO work cursors, one per hit-queue partition, each on its own line.
// Same-address device atomics serialize at the memory side (they come from
// all XCDs), so every counter that many waves touch is spread this way.
constexpr int kStatSpread = 64;
constexpr int kQueueParts = 16;
constexpr int kLineWords = 16;  // u64 words per 128-byte line

enum Counter {
  kCntPrimary = 0,
  kCntShaded = 1,
  kCntAoTraced = 2,
  kCntAoSampled = 3,
  kCntShadow = 4,
  kCntCapped = 5,
  kCntGiRays = 6,     // GI gather / continuation closest-hit queries
  kCntGiNee = 7,      // GI next-event shadow queries
  kCntNodes = 8,      // node visits per ray class: 8 primary, 9 AO, 10 shadow      // node visits per ray class: 8 primary, 9 AO, 10 shadow
  kCntPrimTests = 11, // primitive tests per ray class: 11, 12, 13
  kCntGiGbuf = 14,    // GI reduced-grid first-hit queries
  kNumCounters = 16
};

// word offsets; partition p's counter is at base + kLineWords * p
constexpr int kHitFill = kNumCounters * kStatSpread;
constexpr int kActiveFill = kHitFill + kLineWords * kQueueParts;
constexpr int kAoCursor = kActiveFill + kLineWords * kQueueParts;
// adaptive AA: fills of the flagged and the replicated output-pixel lists
constexpr int kAaFlagFill = kAoCursor + kLineWords * kQueueParts;
constexpr int kAaRepFill = kAaFlagFill + kLineWords;
// GI receiver list (hi-res pixels whose first hit gathers), partitioned like
// the hit queue
constexpr int kGiFill = kAaRepFill + kLineWords;
// GI rim-patch list (reduced gather grids), partitioned the same way
constexpr int kGiPatchFill = kGiFill + kLineWords * kQueueParts;
// gather work cursors, one per partition of the receiver / rim-patch list
constexpr int kGiCursor = kGiPatchFill + kLineWords * kQueueParts;
constexpr int kGiPatchCursor = kGiCursor + kLineWords * kQueueParts;
constexpr int kCounterWords = kGiPatchCursor + kLineWords * kQueueParts;

}  // namespace ubg
