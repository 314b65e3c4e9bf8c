// ub_devbuild.h -- BVH8 construction on the GPU (gfx950).
//
// Input: the primitive records (3 x float4 each, any order) and their
// conservative FP32 boxes, both already in device memory. Output: BVH8 nodes
// in the layout of ub_layout.h and the records reordered into leaf order, in
// freshly allocated device memory owned by the caller (hipFree).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace ubg {

struct DeviceBvh8 {
  void* nodes = nullptr;  // 80 B per node
  void* prims = nullptr;  // 48 B per primitive, leaf order
  void* kinds = nullptr;  // 8 B per node: leaf primitive kind bit planes
  int64_t nNodes = 0;
  int32_t depth = 0;      // levels of 8-wide nodes
  double sahCost = 0.0;   // SAH of the binary tree (traversal 1, primitive 1), as the host builder reports
  double seconds = 0.0;   // device build time (sort to last level), host wall clock
};

// Returns UBG_OK or an error status with `err` set. `recs` = 3 float4 per
// primitive, `boxes` = 2 float4 per primitive (lo.xyz, hi.xyz), n >= 1.
int deviceBuildBvh8(const float4* recs, const float4* boxes, int64_t n, hipStream_t st, DeviceBvh8& out,
                    std::string& err, int leafMax = 3);

}  // namespace ubg
