// ub_kernels.h -- host-side launchers of the gfx950 wavefront kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ub_layout.h"

namespace ubg {

// Device buffers one frame's passes read and write (all owned by the context
// except the optional user outputs). Capacity = hi-res pixel count.
struct FrameBufs {
  // hit queue (structure of arrays, filled by the primary pass)
  uint32_t* hitPix;
  float* hitT;
  float4* hitO;       // offset origin xyz, eps
  float4* hitN;       // shading normal xyz, packed flags
  float4* hitC;       // pigment rgb, opacity
  uint32_t* hitAo;    // AO occluded count
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
  unsigned long long* counters;
};

void launchZeroCounters(unsigned long long* c, int n, hipStream_t st);
void launchResetQueues(unsigned long long* c, hipStream_t st);
void launchPrimary(const FrameConst& fc, const SceneView& s, const FrameBufs& b, int iter,
                   int nActive, bool countVisits, hipStream_t st);
// nMax bounds the hit-queue fill; the kernels read the exact fill from
// counters[kCntHits] on the device, so no host round trip is needed.
void launchAo(const FrameConst& fc, const SceneView& s, const FrameBufs& b, int nMax,
              bool countVisits, hipStream_t st);
void launchResolve(const FrameConst& fc, const SceneView& s, const FrameBufs& b, int iter,
                   int nMax, bool countVisits, hipStream_t st);
void launchPost(const float4* hi, int W, int H, int ss, float gamma, float4* out, hipStream_t st);
void launchQuery(const SceneView& s, const void* rays, int64_t n, void* hits, uint8_t* occ,
                 bool any, hipStream_t st);

}  // namespace ubg
