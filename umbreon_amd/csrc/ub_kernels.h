// ub_kernels.h -- host-side launchers of the gfx950 wavefront kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ub_layout.h"

namespace ubg {

// Device buffers one frame's passes read and write (all owned by the context
// except the optional user outputs). The hit queue and the continuation
// queue are split into kQueueParts partitions of partCap slots each
// (partition p = slots [p * partCap, (p + 1) * partCap)); partCap is a
// multiple of the trace block size and kQueueParts * partCap covers every
// thread of the first primary launch.
struct FrameBufs {
  // hit queue (structure of arrays, filled by the primary pass)
  uint32_t* hitPix;   // seed base px + py * seedW of the hit's ray-lattice position
  uint32_t* hitOut;   // list passes: the hit's output index (null: = hitPix)
  uint32_t* hitPrim;  // the hit's primitive (kind << kKindShift | id); null unless an
                      // anisotropic principled primitive needs its tangent
  int32_t principled; // the scene has a principled material
  float* hitT;
  float4* hitO;       // offset origin xyz, eps
  float4* hitN;       // shading normal xyz, packed flags
  float4* hitC;       // pigment rgb, opacity
  uint32_t* hitAo;    // AO occluded count
  float* aoT;         // enhanced AO: nearest-hit distance per (slot, sample); -1 = below the horizon
  float* hitOpen;     // coarse AO: openness looked up for the hit (kFlagCoarse)
  float4* hitBent;    // coarse AO + enhanced estimator: looked-up bent normal
  // coarse AO grid (cW x cH): guide (shading normal, first-hit distance; 0 =
  // no AO cell), gathered openness, bent normal (enhanced estimator)
  float4* cGuide;
  float* cOpen;
  float4* cBent;
  // continuation (transparency) queue and per-pixel walk state
  uint32_t* activeIn;
  uint32_t* activeOut;
  float4* acc;        // Cf.rgb, Af
  float* tnear;
  // hi-res outputs
  float4* hiColor;
  float* depth;
  float* viewZ;
  uint32_t* hitId;    // may be null
  float* aoOpen;      // may be null
  // adaptive AA: first-hit shading normal and (kind << 16 | group), all ones
  // on a miss (may be null); the pass's output-pixel list (sample maps)
  float4* firstN;
  const uint32_t* smList;
  // pt1 GI (hi-res, first hit): position, shading normal + receiver flag (w),
  // diffuse reflectance; the gathered E + first-bounce hit fraction (w); the
  // receiver list (partitioned like the hit queue)
  float4* giPos;
  float4* giNrm;
  float4* giRefl;
  float4* giE;
  uint32_t* giList;
  unsigned long long* counters;
  int32_t partCap;
};

// Output buffers of one adaptive-AA sample set (any may be null but color).
struct AaSet {
  float4* color;      // linear RGBA before fog
  float* depth;
  float* viewZ;
  uint32_t* hitId;
  float* aoOpen;
};

// Adaptive-AA refinement (adaptive_aa.hpp): the raw pair predicates over the
// Wf x Hf centre results, then the 3x3 dilation, which also fills the
// flagged list (and, with `wantRep`, the replicated list) with one slot per
// output pixel (slotOf) and writes the float mask when `maskOut` is set.
void launchAaMask(const FrameConst& fc, const AaSet& centers, const float4* firstN, float pixelWorldSize,
                  float threshold, uint8_t* raw, uint8_t* mask, uint32_t* slotOf, uint32_t* flagList,
                  uint32_t* repList, bool wantRep, float* maskOut, unsigned long long* counters, hipStream_t st);
// Composes the W x H hi-res buffers (fog applied) from the centre, replicated
// and refined sample sets (embree_renderer.cpp:1619-1673 + fog.cpp:17-36).
void launchAaCompose(const FrameConst& fc, const AaSet& centers, const AaSet& rep, const AaSet& cells,
                     const uint8_t* mask, const uint32_t* slotOf, bool boosted, const FrameBufs& out,
                     hipStream_t st);

// One pt1 GI gather grid: the render grid (the layer pass's first-hit
// G-buffer) or a reduced grid with its own G-buffer (pt1_gbuffer.hpp).
struct GiField {
  int32_t w, h;
  const float4* pos;   // first-hit position xyz
  const float4* nrm;   // shading normal xyz (the denoise / upsample guide), w != 0: gathers
  const float4* geoN;  // geometric normal for the horizon test (null: the shading normal)
  const float* depth;  // primary-ray distance (the gather origin's epsilon scale)
  float4* E;           // gathered E rgb, first-bounce hit fraction in w
  uint32_t* list;      // gather list, partitioned like the hit queue
  int32_t fillBase;    // counter word of the list's partition fills
  int32_t spp;
};

// pt1 GI after the layer passes (embree_renderer.cpp:560-1049):
// * GiList: the pixels of `f` to gather (nrm.w != 0, or mask != 0) into the
//   list; with zeroOthers the rest get E = 0;
// * GiGbuffer: tracePt1GBuffer (pt1_gbuffer.hpp:86-200) over the reduced
//   grid `f` (albedo guide into `albedo`);
// * GiGather: pt1GatherPoint (pt1_gather.hpp:228-485) per listed pixel;
// * GiDenoise: denoisePt1E's a-trous filter on f.E (denoise_atrous.cpp;
//   irrA / irrB / grad / var: f.w x f.h scratch);
// * GiUpsample: upsampleJointBilateral (pt1_upsample.hpp:35-176) of the
//   reduced E to the render grid, the rim-patch mask into `patch`;
// * GiSmoothPatched: smoothPatchedPixels (pt1_upsample.hpp:188-245) over the
//   patched render-grid pixels, reading the snapshot `Esrc`;
// * GiComposite: colour += gi_intensity * reflectance * E, then the fog
//   (fog.cpp:17-36), and the optional AOVs.
void launchGiList(const GiField& f, const uint8_t* mask, bool zeroOthers, const FrameBufs& b, hipStream_t st);
void launchGiGbuffer(const FrameConst& fc, const SceneView& s, const GiField& f, float4* albedo,
                     unsigned long long* counters, hipStream_t st);
void launchGiGather(const FrameConst& fc, const SceneView& s, const FrameBufs& b, const GiField& f, bool countVisits,
                    hipStream_t st);
void launchGiDenoise(const GiField& f, int iters, float sigmaZ, float sigmaN, float sigmaL, float4* irrA,
                     float4* irrB, float* grad, float* var, hipStream_t st);
void launchGiUpsample(const FrameConst& fc, const FrameBufs& b, const GiField& low, float normalPow,
                      float depthScale, int fallbackRadius, uint8_t* patch, float patchThresh, hipStream_t st);
void launchGiSmoothPatched(const FrameConst& fc, const FrameBufs& b, const uint8_t* patch, float normalPow,
                           float depthScale, const float4* Esrc, hipStream_t st);
void launchGiComposite(const FrameConst& fc, const FrameBufs& b, float* indirect, float* occlusion, hipStream_t st);

// Queue slots per partition for a W x H hi-res grid.
int queuePartCap(int W, int H);


// Device address of this device's copy of the fixed frame scratch
// (UBG_NOALLOC_SCRATCH_BYTES, ub_static.hip); looked up once per device.
hipError_t staticScratchAddress(int device, void** out);

// Per-partition fills of the continuation queue a layer consumes.
struct PartFill {
  uint32_t n[kQueueParts];
};

// Batched frames (ubg_render_batch). One request of a batch: its frame
// constants, scene and buffers, the resolve parameters and output.
constexpr int kMaxBatch = 256;
// Waves per SIMD the persistent AO kernels are compiled for (their register
// budget), and so the most a batch's AO grid can keep resident.
constexpr int kAoWavesPerSimd = 5;
struct BatchItem {
  FrameConst fc;
  SceneView s;
  FrameBufs b;
  float4* out;     // final (downsampled, gamma) colour
  int32_t ss;
  float gamma;
};
// The requests one batched launch covers: batch item, inclusive prefix of
// the blocks, and for continuation layers the blocks per queue partition and
// the partition fills.
struct BatchLaunch {
  int32_t n;
  int32_t item[kMaxBatch];
  uint32_t blockEnd[kMaxBatch];
  int32_t partBlocks[kMaxBatch];
  PartFill fill[kMaxBatch];
};
size_t traceLdsBytes(const SceneView& s);
int64_t aoResidentBlocks(int wavesPerSimd);
void launchZeroBatch(const BatchItem* items, int n, hipStream_t st);
void launchResetBatch(const BatchItem* items, const BatchLaunch* L, int n, hipStream_t st);
void launchPrimaryBatch(const BatchItem* items, const BatchLaunch* L, int64_t blocks, int iter, size_t lds,
                        hipStream_t st);
void launchAoBatch(const BatchItem* items, const BatchLaunch* L, int64_t blocks, size_t lds, hipStream_t st);
void launchResolveBatch(const BatchItem* items, const BatchLaunch* L, int64_t blocks, int iter, size_t lds,
                        hipStream_t st);
void launchPostBatch(const BatchItem* items, const BatchLaunch* L, int64_t blocks, hipStream_t st);
void launchFillsBatch(const BatchItem* items, int n, uint32_t* fills, hipStream_t st);
void launchStatsBatch(const BatchItem* items, int n, unsigned long long* stats, hipStream_t st);

void launchZeroCounters(unsigned long long* c, int n, hipStream_t st);
void launchResetQueues(unsigned long long* c, hipStream_t st);
// iter == 0 traces the pass's samples (every hi-res pixel for kSmGrid, the
// output pixels for an unlisted kSmCenter pass, else fc.nList list samples);
// later layers trace the continuation queue whose partition fills are
// `fill` (read back by the host).
void launchPrimary(const FrameConst& fc, const SceneView& s, const FrameBufs& b, int iter,
                   const PartFill& fill, bool countVisits, hipStream_t st);
// AO and resolve read the hit-queue fills on the device (no host round trip).
void launchAo(const FrameConst& fc, const SceneView& s, const FrameBufs& b, bool countVisits, hipStream_t st);
// The learned denoiser's features of the render-grid gather (Wp x Hp x
// stride: 10 for the caller's copy, 12 = kUnetFeatures for the network with
// the two noise descriptors zero; edge-replicated beyond W x H; maxBits: one
// scratch word) and the network's log(1 + E) back onto the receivers' E.
void launchGiFeatures(int W, int H, int Wp, int Hp, const FrameBufs& b, unsigned int* maxBits, int stride,
                      float* feat, hipStream_t st);
void launchGiApplyNet(int W, int H, int Wp, const float* pred, const FrameBufs& b, hipStream_t st);
// Point-in-solid tests of the object-space edge clip over a uniform grid
// (spheres x, y, z, r; cylinders p0.xyz r, p1.xyz 0; points x, y, z, ring
// radius; sources = the skipped sphere / cylinder, -1 for none).
void launchSolidsContain(int64_t n, const float4* pts, const int2* src, const float4* sph, const float4* cyl,
                         const uint32_t* cellStart, const uint32_t* cellItems, const float lo[3], float inv,
                         const int dim[3], uint8_t* inside, hipStream_t st);
void launchResolve(const FrameConst& fc, const SceneView& s, const FrameBufs& b, int iter, bool countVisits,
                   hipStream_t st);
// Coarse-grid AO pre-pass: the first-hit G-buffer of the cW x cH grid (AO
// cells appended to the hit queue), then, after the AO kernels ran on that
// queue with seedW = cW, the per-cell reduction into cOpen / cBent.
void launchCoarseGbuf(const FrameConst& fc, const SceneView& s, const FrameBufs& b, bool countVisits,
                      hipStream_t st);
void launchCoarseReduce(const FrameConst& fc, const FrameBufs& b, hipStream_t st);
void launchPost(const float4* hi, int W, int H, int ss, float gamma, float4* out, hipStream_t st);
void launchQuery(const SceneView& s, const void* rays, int64_t n, void* hits, uint8_t* occ,
                 bool any, hipStream_t st);
// ubg_occluded_segments: `segs` = n ubg_segment records, `excl` = the flat
// exclusion list (nExcl entries; segment ranges are checked against it).
void launchSegments(const SceneView& s, const void* segs, int64_t n, const int32_t* excl, int64_t nExcl, float eps,
                    float grazeCosEps, float coplanarEps, uint8_t* occ, hipStream_t st);

}  // namespace ubg
