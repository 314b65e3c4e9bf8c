// ub_pathtrace.h -- host-side interface of the physically based path tracer
// (ub_pathtrace.hip) behind ubg_pathtrace.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ub_layout.h"

namespace ubg {
namespace pt {

constexpr int kMatDiffuse = 0;
constexpr int kMatConductor = 1;
constexpr int kMatDielectric = 2;
constexpr int kMaxSpatial = 16;   // ReSTIR neighbours
constexpr int kMaxCandidates = 12;
constexpr int kPtFeatures = 12;  // the denoiser's input features per pixel
constexpr int kLightGroup = 8;  // lights per bounding sphere of the BSDF rays' light test
// Colour encoding of the denoiser's input and output: u = log(1 + 32 c) / 4.
// Its slope over c follows the sRGB transfer's within a factor of ~1.5 from
// c = 0.01 to 1 (the images are compared in sRGB), where log(1 + c) would
// weigh the dark colours 5-10 times less than the comparison does.
constexpr float kColourGain = 32.0f;
constexpr float kColourScale = 0.25f;

struct Light {      // a sphere of constant radiance
  float c[3];
  float r;
  float Le[3];
  float pad;
};
struct Mat {        // per material index of the scene
  int32_t type;
  float alpha;      // GGX alpha = roughness^2
  float ior;
  int32_t pad;
};
// One sample pass over the image.
struct Params {
  int32_t W, H;
  int32_t ortho;
  float pos[3], dir[3], right[3], up[3];
  float halfW, halfH, persHalfW, persHalfH;
  int32_t sample;        // index of this pass's sample
  int32_t bounces;       // shaded path vertices (BSDF scatterings)
  uint32_t seed;
  int32_t nLights;
  float env[3];
  int32_t restirM;       // light candidates per vertex
  int32_t restirK;       // spatial neighbours at the first vertex
  float restirRadius;    // pixels
  int32_t neeAll;        // 1: one cone sample of every light instead of RIS / ReSTIR
  const Light* lights;
  // bounding spheres of the lights in groups of kLightGroup (index order):
  // c[3], r; ceil(nLights / kLightGroup) of them
  const float4* lightGroups;
  const Mat* primMats;
  const Mat* meshMats;   // == primMats + nPrimMats (one table, read by offset)
  int32_t nPrimMats, nMeshMats;
  unsigned long long* rays;  // traced-ray counter
  unsigned int* depthMax;    // sample 0 with features: maximum first-hit distance (float bits)
};
// First-hit record: position + ray distance (w = 0: no surface), shading
// normal + material type, geometric normal + GGX alpha, base colour + IOR,
// direction to the eye.
struct GPix {
  float4 P, N, Ng, C, wo;
};
// The first vertex's light reservoir: selected light point + light index,
// weight sum, candidate count, unbiased contribution weight, target.
struct Reservoir {
  float4 Q;
  float wsum, M, W, target;
};

}  // namespace pt

void ptZero(float4* acc, int n, unsigned int* maxBits, unsigned long long* rays, hipStream_t st);
void ptSample(const pt::Params& p, const SceneView& s, pt::GPix* gb, pt::Reservoir* res, float4* acc, float4* first,
              hipStream_t st);
// colour = acc.xyz / spp (W x H; acc.w holds the sum of the samples' squared
// mean-of-rgb); with feat: the Wp x Hp edge-padded denoiser features as well
// (maxBits: the maximum first-hit distance k_pt_guides recorded in
// Params::depthMax).
void ptResolve(int W, int H, int Wp, int Hp, int spp, const float4* acc, const float4* first,
               const unsigned int* maxBits, float4* color, float* feat, hipStream_t st);
// The denoiser's self-ensemble: turn k (0..7; bit 2 transpose, then bit 0
// mirrored columns, bit 1 mirrored rows) of the Wp x Hp feature image with
// its normals, and the prediction on turn k turned back onto W x H (stride
// Wp) with weight wgt (k = 0 writes, the others add).
void ptFeatTurn(int k, int Wp, int Hp, const float* feat, float* turned, hipStream_t st);
void ptPredUnturn(int k, int W, int H, int Wp, int Hp, float wgt, const float* predK, float* sum, hipStream_t st);
void ptApplyNet(int W, int H, int Wp, const float* pred, float4* color, hipStream_t st);

}  // namespace ubg
