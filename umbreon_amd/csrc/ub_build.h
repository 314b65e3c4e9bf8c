// ub_build.h -- host-side scene flattening and BVH construction.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/umbreon_gpu.h"
#include "ub_layout.h"

namespace ubg {

struct F4 {
  float x, y, z, w;
};
struct U4 {
  uint32_t x, y, z, w;
};

// Everything that goes to HBM for one scene, still on the host.
struct HostScene {
  std::vector<F4> nodes;   // BVH2: 4 per inner node; BVH8: 5 per node
  std::vector<uint32_t> nodeKinds;  // BVH8: 2 per node (leaf primitive kind bit planes)
  std::vector<F4> prims;   // 3 per primitive, leaf order (record order when recordsOnly)
  std::vector<F4> boxes;   // recordsOnly: 2 per primitive (lo.xyz, hi.xyz), conservative FP32
  std::vector<F4> sphColor;
  std::vector<uint32_t> sphInfo;
  std::vector<F4> coneColor;
  std::vector<uint32_t> coneInfo;
  std::vector<float> coneOp1;
  std::vector<F4> rndColor;
  std::vector<uint32_t> rndInfo;
  std::vector<float> rndOp1;
  std::vector<F4> rndPlanes;  // 2 per round segment
  std::vector<int32_t> rndSource;  // source cylinder of each round segment
  std::vector<U4> triIdx;
  // section groups: per sphere, cone and round segment, and per base triangle
  std::vector<uint16_t> sphGroup, coneGroup, rndGroup, triGroup;
  // anisotropy tangent axes (curve_build.cpp:163-167, :243-245), built only
  // when an anisotropic principled cylinder exists
  std::vector<F4> coneAxis, rndAxis;
  std::vector<F4> vtxNrm;
  std::vector<F4> vtxCol;
  std::vector<DevMaterial> primMats;
  std::vector<DevMaterial> meshMats;
  int64_t nTris = 0, nSpheres = 0, nRound = 0, nCone = 0;
  int64_t meshBaseTris = 0;  // triangles of one instance copy
  bool aoReceivers = false;  // a mesh or a primitive that is not outline decoration
  bool pt2Emitters = false;  // an emissive-NEE source for the pt2 integrator
  bool pt2Reflective = false;  // a receiver that owns a traced reflection
  bool principled = false;     // a material with ShadingModel::Principled
  bool anisoPrims = false;     // an anisotropic principled sphere or cylinder
  int32_t layout = kLayoutBvh8;
  int32_t depth = 0;       // tree levels
  int32_t stackDepth = 0;  // traversal stack entries needed
  int64_t nNodes = 0;
  int32_t nodeBytes = 0;
  double sahCost = 0.0;
};

// Flatten `desc` (curve chaining, instance baking, side tables) and build the
// BVH in `layout` (kLayoutBvh2 / kLayoutBvh8). With recordsOnly the BVH is
// left to the caller: `prims` keeps the records in geometry order and
// `boxes` holds their bounds. leafMax8: primitives per BVH8 leaf slot
// (1..kMaxLeaf8). Returns UBG_OK or an error status with `err` set.
int buildHostScene(const ubg_scene_desc* desc, int layout, HostScene& out, std::string& err,
                   bool recordsOnly = false, int leafMax8 = kMaxLeaf8);

// shading.hpp:34-41
float blinnExponent(float roughness);

}  // namespace ubg
