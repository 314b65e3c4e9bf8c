// ub_capi.cpp -- the extern "C" boundary of the gfx950 per-ray backend
// (declared in include/umbreon_gpu.h).
//
// Device context <-> rtcNewDevice/SharedDevice (embree_renderer.cpp:1267-1273,
// umbreon.cpp:21-39); scene <-> buildEmbreeScene (scene_build.cpp:133-170);
// ubg_render <-> the EmbreeRenderer::render grid path
// (embree_renderer.cpp:1260-1427) followed by renderFrame's fog, box
// downsample and assumed gamma (pipeline.cpp:113-129,182-254). Errors never
// abort: every failure returns a status and leaves a message for
// ubg_last_error(), which the C++ facade turns into std::runtime_error.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/umbreon_gpu.h"
#include "ub_build.h"
#include "ub_devbuild.h"
#include "ub_edges.h"
#include "ub_kernels.h"
#include "ub_pathtrace.h"
#include "ub_unet.h"
#include "ub_layout.h"
#include "ub_testing.h"

namespace {

thread_local std::string g_lastError;

int fail(int status, const std::string& msg) {
  g_lastError = msg;
  return status;
}

int hipFail(hipError_t e, const char* what) {
  const int st = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? UBG_ERR_OUT_OF_MEMORY
                                                                            : UBG_ERR_DEVICE;
  return fail(st, std::string(what) + ": " + hipGetErrorString(e));
}

#define UBG_HIP(call, what)                 \
  do {                                      \
    hipError_t e__ = (call);                \
    if (e__ != hipSuccess) return hipFail(e__, what); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct ubg_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<hipEvent_t> tev;  // per-kernel timing events (grow-only)
  std::mutex mu;
  // grow-only frame scratch
  DevBuf hitPix, hitT, hitO, hitN, hitC, hitAo, aoT, activeA, activeB, acc, tnear, hiColor, depth, viewZ,
      hitId, aoOpen, outColor, counters, query, queryOut, hitOpen, hitBent, cGuide, cOpen, cBent;
  // adaptive AA: hit output indices, the centre / replicated / refined sample
  // sets (32 B per sample), first-hit records, per-output-pixel mask state
  DevBuf hitOut, aaSetC, aaSetR, aaSetF, aaFirstN, aaPix, hitPrim;
  // pt1 GI: first-hit position / normal / reflectance, gathered E, receiver
  // list, staging of the indirect and occlusion AOVs
  DevBuf giPos, giNrm, giRefl, giE, giList, giInd, giOcc, dnA, dnB, dnGrad, dnVar, giLow, giPatch, giSnap;
  // object-space edge clip: points + sources, solids + grid, results
  DevBuf edgePts, edgeSolids, edgeOut;
  // learned denoiser staging of host inputs / outputs; the render's feature
  // AOV, the padded network input, its prediction and a max-depth word
  DevBuf unetIn, unetTgt, unetOut, giFeat, giFeatP, giPred, giMax;
  unsigned long long* hostCounters = nullptr;  // pinned
  // batched frames: the shared queue / frame-buffer pool, the request table,
  // launch tables and read-backs (device), their pinned host staging
  DevBuf batchPool, batchItems, batchLaunch, batchOut;
  // path tracer: accumulator, first-hit records, G-buffer, reservoirs,
  // lights + materials, padded features, prediction, colour staging, words
  DevBuf ptAcc, ptFirst, ptG, ptRes, ptTab, ptFeat, ptPred, ptColor, ptWords, ptFeatT, ptPredT;
  void* batchHost = nullptr;  // pinned: items + launch tables + read-backs
  size_t batchHostBytes = 0;
  // instantiated graphs of a batched layer (reset, primary, AO, resolve,
  // fill read-back), keyed by the launch shape; valid while the batch
  // buffers stay where they are
  std::map<std::vector<int64_t>, hipGraphExec_t> layerGraphs;
  void clearLayerGraphs() {
    for (auto& kv : layerGraphs) (void)hipGraphExecDestroy(kv.second);
    layerGraphs.clear();
  }
  void* staticScratch = nullptr;               // device address of the fixed scratch (ub_static.hip)
  hipError_t staticStatus = hipSuccess;
  std::vector<DevBuf*> all() {
    return {&hitPix, &hitT,   &hitO,    &hitN,  &hitC,  &hitAo,   &activeA,  &activeB,  &acc,     &tnear,
            &hiColor, &depth, &viewZ,   &hitId, &aoOpen, &outColor, &counters, &query, &queryOut, &aoT,
            &hitOpen, &hitBent, &cGuide, &cOpen, &cBent, &hitOut, &aaSetC, &aaSetR, &aaSetF, &aaFirstN, &hitPrim,
            &aaPix, &giPos, &giNrm, &giRefl, &giE, &giList, &giInd, &giOcc, &dnA, &dnB, &dnGrad,
            &dnVar, &giLow, &giPatch, &giSnap, &edgePts, &edgeSolids, &edgeOut, &unetIn,
            &unetTgt, &unetOut, &giFeat, &giFeatP, &giPred, &giMax, &batchPool, &batchItems, &batchLaunch,
            &batchOut, &ptAcc, &ptFirst, &ptG, &ptRes, &ptTab, &ptFeat, &ptPred, &ptColor, &ptWords, &ptFeatT, &ptPredT};
  }
};

struct ubg_unet {
  ubg_context* ctx = nullptr;
  ubg::Unet net;
};

struct ubg_scene {
  ubg_context* ctx = nullptr;
  ubg::SceneView view{};
  bool aoReceivers = false;
  bool pt2Emitters = false, pt2Reflective = false;
  bool principled = false, anisoPrims = false;
  std::vector<void*> allocs;
  ubg_scene_info info{};
};

namespace {

// Bump allocation of frame scratch from the fixed per-GPU array during a
// ubg_render_noalloc call (set on the calling thread for that call only).
struct StaticArena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
};
thread_local StaticArena* t_arena = nullptr;

int ensure(DevBuf& b, size_t bytes) {
  if (t_arena) {
    if (b.p && b.bytes >= bytes) return UBG_OK;
    const size_t want = (std::max<size_t>(bytes, 256) + 255) & ~size_t(255);
    if (t_arena->used + want > t_arena->cap)
      return fail(UBG_ERR_OUT_OF_MEMORY, "ubg_render_noalloc: the frame needs more scratch than the fixed " +
                                             std::to_string(t_arena->cap >> 20) + " MiB (" +
                                             std::to_string((t_arena->used + want) >> 20) + "+ MiB requested)");
    b.p = t_arena->base + t_arena->used;
    b.bytes = want;
    t_arena->used += want;
    return UBG_OK;
  }
  if (b.bytes >= bytes && b.p) return UBG_OK;
  if (b.p) {
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  const size_t want = std::max<size_t>(bytes, 256);
  UBG_HIP(hipMalloc(&b.p, want), "hipMalloc(frame scratch)");
  b.bytes = want;
  return UBG_OK;
}

template <class T>
int upload(ubg_scene* s, const std::vector<T>& v, const void** dst) {
  *dst = nullptr;
  if (v.empty()) return UBG_OK;
  void* p = nullptr;
  UBG_HIP(hipMalloc(&p, v.size() * sizeof(T)), "hipMalloc(scene)");
  s->allocs.push_back(p);
  UBG_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s->ctx->stream),
          "hipMemcpy(scene upload)");
  s->info.device_bytes += static_cast<int64_t>(v.size() * sizeof(T));
  *dst = p;
  return UBG_OK;
}

struct V3h {
  float x, y, z;
};
inline V3h scale(V3h a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(V3h a, V3h b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3h cross(V3h a, V3h b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3h normalize(V3h a) {
  const float l = std::sqrt(dot(a, a));
  return l > 0.0f ? scale(a, 1.0f / l) : a;
}

int copyOut(void* dst, const void* src, size_t bytes, int32_t memory, hipStream_t st) {
  if (!dst || bytes == 0) return UBG_OK;
  UBG_HIP(hipMemcpyAsync(dst, src, bytes, memory ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st),
          "hipMemcpy(frame output)");
  return UBG_OK;
}

}  // namespace

extern "C" {

int ubg_abi_version(void) { return UBG_ABI_VERSION; }

const char* ubg_status_string(int status) {
  switch (status) {
    case UBG_OK: return "ok";
    case UBG_ERR_INVALID_ARGUMENT: return "invalid argument";
    case UBG_ERR_DEVICE: return "device error";
    case UBG_ERR_OUT_OF_MEMORY: return "out of device memory";
    case UBG_ERR_CANCELLED: return "cancelled";
    case UBG_ERR_UNSUPPORTED: return "unsupported";
    case UBG_ERR_INTERNAL: return "internal error";
    default: return "unknown status";
  }
}

const char* ubg_last_error(void) { return g_lastError.c_str(); }

int ubg_context_create(int device_ordinal, ubg_context** out) {
  if (!out) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_context_create: out is null");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0)
    return fail(UBG_ERR_DEVICE, "ubg_context_create: no HIP device available");
  if (device_ordinal < 0 || device_ordinal >= n)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_context_create: device ordinal out of range");
  hipDeviceProp_t prop;
  UBG_HIP(hipGetDeviceProperties(&prop, device_ordinal), "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(UBG_ERR_DEVICE, std::string("ubg_context_create: device is ") + prop.gcnArchName +
                                    ", this library is built for gfx950 only");
  UBG_HIP(hipSetDevice(device_ordinal), "hipSetDevice");
  ubg_context* c = new ubg_context();
  c->device = device_ordinal;
  hipError_t se = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (se != hipSuccess) {
    delete c;
    return hipFail(se, "hipStreamCreate");
  }
  for (auto& ev : c->ev) {
    if (hipEventCreate(&ev) != hipSuccess) {
      ubg_context_destroy(c);
      return fail(UBG_ERR_DEVICE, "hipEventCreate failed");
    }
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&c->hostCounters), ubg::kCounterWords * sizeof(unsigned long long),
                    hipHostMallocDefault) != hipSuccess) {
    ubg_context_destroy(c);
    return fail(UBG_ERR_OUT_OF_MEMORY, "hipHostMalloc(counters) failed");
  }
  // this device's copy of the fixed scratch of ubg_render_noalloc (device
  // current); a failed lookup only disables that entry point
  c->staticStatus = ubg::staticScratchAddress(device_ordinal, &c->staticScratch);
  *out = c;
  return UBG_OK;
}

void ubg_context_destroy(ubg_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (DevBuf* b : ctx->all())
    if (b->p) (void)hipFree(b->p);
  for (auto& ev : ctx->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (hipEvent_t ev : ctx->tev) (void)hipEventDestroy(ev);
  if (ctx->hostCounters) (void)hipHostFree(ctx->hostCounters);
  ctx->clearLayerGraphs();
  if (ctx->batchHost) (void)hipHostFree(ctx->batchHost);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

static void freeSceneLocked(ubg_scene* s) {
  (void)hipStreamSynchronize(s->ctx->stream);
  for (void* p : s->allocs) (void)hipFree(p);
  delete s;
}

int ubg_scene_create_ex(ubg_context* ctx, const ubg_scene_desc* desc, int32_t bvh_layout, ubg_scene** out) {
  ubg_scene_options o{};
  o.bvh_layout = bvh_layout;
  return ubg_scene_create_opts(ctx, desc, &o, out);
}

int ubg_scene_create_opts(ubg_context* ctx, const ubg_scene_desc* desc, const ubg_scene_options* opts,
                          ubg_scene** out) {
  if (!ctx || !desc || !out || !opts) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_scene_create: null argument");
  const int32_t bvh_layout = opts->bvh_layout;
  if (opts->leaf_max < 0 || opts->leaf_max > ubg::kMaxLeaf8)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_scene_create: leaf_max must be 0 (default) or 1..3");
  const int leafMax = opts->leaf_max > 0 ? opts->leaf_max : ubg::kMaxLeaf8;
  *out = nullptr;
  if (bvh_layout != UBG_BVH2 && bvh_layout != UBG_BVH8 && bvh_layout != UBG_BVH8_GPU_BUILD)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_scene_create: unknown BVH layout");
  const bool gpuBuild = bvh_layout == UBG_BVH8_GPU_BUILD;
  std::lock_guard<std::mutex> lk(ctx->mu);
  UBG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  const auto t0 = std::chrono::steady_clock::now();
  ubg::HostScene hs;
  std::string err;
  const int st = ubg::buildHostScene(desc, bvh_layout == UBG_BVH2 ? ubg::kLayoutBvh2 : ubg::kLayoutBvh8, hs, err,
                                     gpuBuild, leafMax);
  if (st != UBG_OK) return fail(st, "scene build failed: " + err);
  ubg_scene* s = new ubg_scene();
  s->ctx = ctx;
  int r = UBG_OK;
  auto up = [&](auto& vec, const void** dst) {
    if (r == UBG_OK) r = upload(s, vec, dst);
  };
  const void* p = nullptr;
  if (gpuBuild && !hs.prims.empty()) {
    // records and boxes go up in geometry order; the GPU builds the BVH8 and
    // writes the records again in leaf order
    const void* recs = nullptr;
    const void* boxes = nullptr;
    up(hs.prims, &recs);
    up(hs.boxes, &boxes);
    ubg::DeviceBvh8 db;
    if (r == UBG_OK) {
      r = ubg::deviceBuildBvh8(static_cast<const float4*>(recs), static_cast<const float4*>(boxes),
                               static_cast<int64_t>(hs.prims.size() / 3), ctx->stream, db, err, leafMax);
      if (r != UBG_OK) (void)fail(r, "scene build failed: " + err);
    }
    if (r == UBG_OK) {
      s->allocs.push_back(db.nodes);
      s->allocs.push_back(db.prims);
      s->allocs.push_back(db.kinds);
      s->view.nodes = db.nodes;
      s->view.prims = db.prims;
      s->view.nodeKinds = db.kinds;
      hs.nNodes = db.nNodes;
      hs.depth = db.depth;
      hs.stackDepth = db.depth;
      hs.nodeBytes = 80;
      hs.sahCost = db.sahCost;
    }
    // the geometry-order copies are not needed any more
    (void)hipStreamSynchronize(ctx->stream);
    for (const void* tmp : {recs, boxes}) {
      auto it = std::find(s->allocs.begin(), s->allocs.end(), tmp);
      if (tmp && it != s->allocs.end()) {
        (void)hipFree(const_cast<void*>(tmp));
        s->allocs.erase(it);
      }
    }
    s->info.device_bytes -= static_cast<int64_t>((hs.prims.size() + hs.boxes.size()) * sizeof(ubg::F4));
    if (r == UBG_OK) s->info.device_bytes += db.nNodes * 88 + static_cast<int64_t>(hs.prims.size() / 3) * 48;
  } else {
    up(hs.nodes, &s->view.nodes);
    up(hs.prims, &s->view.prims);
    up(hs.nodeKinds, &s->view.nodeKinds);
  }
  up(hs.sphColor, &s->view.sphColor);
  up(hs.sphInfo, &p);
  s->view.sphInfo = static_cast<const uint32_t*>(p);
  up(hs.coneColor, &s->view.coneColor);
  up(hs.coneInfo, &p);
  s->view.coneInfo = static_cast<const uint32_t*>(p);
  up(hs.coneOp1, &p);
  s->view.coneOp1 = static_cast<const float*>(p);
  up(hs.rndColor, &s->view.rndColor);
  up(hs.rndInfo, &p);
  s->view.rndInfo = static_cast<const uint32_t*>(p);
  up(hs.rndOp1, &p);
  s->view.rndOp1 = static_cast<const float*>(p);
  up(hs.rndPlanes, &s->view.rndPlanes);
  up(hs.triIdx, &s->view.triIdx);
  up(hs.vtxNrm, &s->view.vtxNrm);
  up(hs.vtxCol, &s->view.vtxCol);
  up(hs.primMats, &p);
  s->view.primMats = static_cast<const ubg::DevMaterial*>(p);
  up(hs.meshMats, &p);
  s->view.meshMats = static_cast<const ubg::DevMaterial*>(p);
  up(hs.sphGroup, &p);
  s->view.sphGroup = static_cast<const uint16_t*>(p);
  up(hs.coneGroup, &p);
  s->view.coneGroup = static_cast<const uint16_t*>(p);
  up(hs.rndGroup, &p);
  s->view.rndGroup = static_cast<const uint16_t*>(p);
  up(hs.triGroup, &p);
  s->view.triGroup = static_cast<const uint16_t*>(p);
  up(hs.coneAxis, &s->view.coneAxis);
  up(hs.rndAxis, &s->view.rndAxis);
  if (r == UBG_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) r = hipFail(e, "scene upload");
  }
  if (r != UBG_OK) {
    freeSceneLocked(s);
    return r;
  }
  s->view.nNodes = static_cast<int32_t>(hs.nNodes);
  s->view.meshBaseTris = hs.meshBaseTris;
  s->aoReceivers = hs.aoReceivers;
  s->pt2Emitters = hs.pt2Emitters;
  s->pt2Reflective = hs.pt2Reflective;
  s->principled = hs.principled;
  s->anisoPrims = hs.anisoPrims;
  s->view.nPrims = static_cast<int32_t>(hs.prims.size() / 3);  // records in either order
  s->view.layout = hs.layout;
  s->view.stackDepth = std::max(1, hs.stackDepth);
  const auto t1 = std::chrono::steady_clock::now();
  s->info.n_triangles = hs.nTris;
  s->info.n_spheres = hs.nSpheres;
  s->info.n_round_segments = hs.nRound;
  s->info.n_cone_segments = hs.nCone;
  s->info.n_nodes = s->view.nNodes;
  s->info.n_leaf_prims = s->view.nPrims;
  s->info.bvh_depth = hs.depth;
  s->info.node_bytes = hs.nodeBytes;
  s->info.bvh_layout = bvh_layout;
  s->info.stack_depth = s->view.stackDepth;
  s->info.prim_bytes = 48;
  s->info.build_seconds = std::chrono::duration<double>(t1 - t0).count();
  s->info.sah_cost = hs.sahCost;
  *out = s;
  return UBG_OK;
}

int ubg_scene_create(ubg_context* ctx, const ubg_scene_desc* desc, ubg_scene** out) {
  return ubg_scene_create_ex(ctx, desc, UBG_BVH8, out);
}

void ubg_scene_destroy(ubg_scene* scene) {
  if (!scene) return;
  if (scene->ctx) {
    std::lock_guard<std::mutex> lk(scene->ctx->mu);
    (void)hipSetDevice(scene->ctx->device);
    freeSceneLocked(scene);
    return;
  }
  delete scene;
}

int ubg_scene_get_info(const ubg_scene* scene, ubg_scene_info* out) {
  if (!scene || !out) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_scene_get_info: null argument");
  *out = scene->info;
  return UBG_OK;
}

// Frame constants of one render (embree_renderer.cpp:375-386, 436-458): the
// camera basis, the per-ray options, the GI and coarse-AO plans and the
// lights; giDiv receives the GI gather-grid divisor. Fails on options the
// path does not implement.
static int frameConstants(ubg_context* ctx, const ubg_scene* scene, const ubg_render_params* p, const ubg_frame* fr,
                          int W, int H, int ss, ubg::FrameConst& fc, int& giDiv) {
  std::memset(&fc, 0, sizeof(fc));
  fc.W = W;
  fc.H = H;
  const ubg_camera& cam = p->camera;
  const float aspect = static_cast<float>(W) / static_cast<float>(H);
  const V3h dir = normalize(V3h{cam.direction[0], cam.direction[1], cam.direction[2]});
  const V3h right = normalize(cross(dir, V3h{cam.up[0], cam.up[1], cam.up[2]}));
  const V3h trueUp = normalize(cross(right, dir));
  fc.ortho = cam.orthographic ? 1 : 0;
  std::memcpy(fc.pos, cam.position, sizeof(fc.pos));
  fc.dir[0] = dir.x, fc.dir[1] = dir.y, fc.dir[2] = dir.z;
  fc.right[0] = right.x, fc.right[1] = right.y, fc.right[2] = right.z;
  fc.up[0] = trueUp.x, fc.up[1] = trueUp.y, fc.up[2] = trueUp.z;
  fc.halfH = cam.height * 0.5f;
  fc.halfW = fc.halfH * aspect;
  fc.persHalfH = std::tan(cam.fovy * 0.01745329252f * 0.5f);
  fc.persHalfW = fc.persHalfH * aspect;
  const float inf = std::numeric_limits<float>::infinity();
  fc.clipOn = (p->clip_near <= -inf && p->clip_far >= inf) ? 0 : 1;
  fc.clipNear = p->clip_near;
  fc.clipFar = p->clip_far;
  std::memcpy(fc.bg, p->background, sizeof(fc.bg));
  std::memcpy(fc.amb, p->ambient_color, sizeof(fc.amb));
  fc.transparentBg = p->transparent_background ? 1 : 0;
  fc.transparency = p->transparency ? 1 : 0;
  fc.maxIters = p->transparency ? (p->max_transparent_layers + 1) : 1;
  fc.aoSamples = p->ao_samples;
  fc.aoDist = p->ao_distance;
  fc.aoIntensity = p->ao_intensity;
  int lanes = 1;
  while (lanes < p->ao_samples && lanes < 64) lanes <<= 1;
  fc.aoLanes = lanes;
  fc.shadows = p->shadows ? 1 : 0;
  fc.shadowSamples = p->shadow_samples;
  fc.specScale = p->specular_scale;
  fc.fogOn = p->fog.enabled ? 1 : 0;
  std::memcpy(fc.fogColor, p->fog.color, sizeof(fc.fogColor));
  fc.fogStart = p->fog.start;
  fc.fogEnd = p->fog.end;
  // enhanced AO (render_options.hpp:79-82; embree_renderer.cpp:1320-1323 for the axis)
  fc.aoEnh = (p->ao_falloff_power > 0.0f || p->ao_multi_scale || p->ao_bent_normal || p->ao_multibounce ||
              p->ao_low_discrepancy || p->ao_diffuse_factor > 0.0f)
                 ? 1
                 : 0;
  fc.aoFalloff = p->ao_falloff_power;
  fc.aoMultiScale = p->ao_multi_scale ? 1 : 0;
  fc.aoBent = p->ao_bent_normal ? 1 : 0;
  fc.aoMultibounce = p->ao_multibounce ? 1 : 0;
  fc.aoLowDisc = p->ao_low_discrepancy ? 1 : 0;
  fc.aoDiffuse = p->ao_diffuse_factor;
  std::memcpy(fc.aoSky, p->ao_sky_color, sizeof(fc.aoSky));
  std::memcpy(fc.aoGround, p->ao_ground_color, sizeof(fc.aoGround));
  {
    V3h up = trueUp;
    if (!p->ao_use_camera_up) {
      const V3h a{p->ao_up[0], p->ao_up[1], p->ao_up[2]};
      const float d = dot(a, a);
      if (!(d < 1.0e-18f)) up = scale(a, 1.0f / std::sqrt(d));
    }
    fc.aoUpAxis[0] = up.x, fc.aoUpAxis[1] = up.y, fc.aoUpAxis[2] = up.z;
  }
  // pt1 GI (render_options.hpp:93-162): gather-grid divisor per
  // embree_renderer.cpp:716-718 (pipeline.cpp:55 resolves -1 to the
  // supersample factor)
  const bool gi = p->gi != 0;
  giDiv = 1;
  if (gi) {
    if (p->gi_integrator != 1 && p->gi_integrator != 2)
      return fail(UBG_ERR_UNSUPPORTED, "ubg_render: the irradiance-cache GI integrator (gi_integrator 0) is not implemented");
    if (p->gi_integrator == 2) {
      if (p->pt2_rounds > 0) return fail(UBG_ERR_UNSUPPORTED, "ubg_render: pt2 ReSTIR (pt2_rounds > 0) is not implemented");
      if (p->pt2_adaptive) return fail(UBG_ERR_UNSUPPORTED, "ubg_render: pt2 adaptive spp is not implemented");
      if (p->pt2_emissive && p->pt2_emissive_nee && scene->pt2Emitters)
        return fail(UBG_ERR_UNSUPPORTED, "ubg_render: pt2 emissive NEE (an emissive material with pt2_emissive_nee) "
                                         "is not implemented");
      if (p->pt2_reflect && scene->pt2Reflective)
        return fail(UBG_ERR_UNSUPPORTED, "ubg_render: pt2 traced reflection (a reflective material with "
                                         "pt2_reflect) is not implemented");
    }
    giDiv = p->pt1_gather_div < 0 ? ss : p->pt1_gather_div;
    if (giDiv == 0) giDiv = p->pt1_half_res ? 2 : 1;
    if (giDiv < 1) giDiv = 1;
    if ((p->pt1_denoise == 2 || fr->gi_features) && giDiv != 1)
      return fail(UBG_ERR_UNSUPPORTED, "ubg_render: the learned denoiser and its features need the full-resolution "
                                       "gather (pt1_gather_div resolving to 1)");
    if (p->pt1_denoise == 2 && (!p->pt1_denoise_net || p->pt1_denoise_net->ctx != ctx))
      return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render: pt1_denoise 2 needs pt1_denoise_net, a network on this context");
    if (p->pt1_spp > (1 << 20) || p->gi_bounces > (1 << 16) ||
        static_cast<int64_t>(std::max(1, p->pt1_spp)) * std::max(1, p->pt1_edge_patch_spp_mul) > (1 << 24))
      return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render: pt1_spp / gi_bounces / edge-patch samples out of range");
  }
  fc.gi = gi ? 1 : 0;
  fc.giPt2 = (gi && p->gi_integrator == 2) ? 1 : 0;
  fc.giPattern = p->pt2_pattern;
  fc.giEmissive = p->pt2_emissive ? 1 : 0;
  fc.giSpp = std::max(1, p->pt1_spp);
  fc.giBounces = std::max(1, p->gi_bounces);
  fc.giLd = p->pt1_ld ? 1 : 0;
  fc.giMaxDist = p->gi_max_distance > 0.0f ? p->gi_max_distance : std::numeric_limits<float>::infinity();
  fc.giIntensity = p->gi_intensity;
  fc.giClamp = p->pt1_clamp;
  {
    // irradiance_cache.hpp:176-184 of pt1Seed (gatherPt1Grid's seedMix)
    uint32_t x = p->pt1_seed;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    fc.giSeedMix = x;
  }
  for (int k = 0; k < 3; ++k) {
    // embree_renderer.cpp:589-600: sky = pt1SkyRadiance, ground = the sky or
    // (gradient mode) aoGroundColor; radiance scaled by giEnvIntensity * ambient
    fc.giEnvScale[k] = p->gi_env_intensity * p->ambient_color[k];
    fc.giSky[k] = p->pt1_sky_radiance[k];
    fc.giGround[k] = p->pt1_sky_mode == 1 ? p->ao_ground_color[k] : p->pt1_sky_radiance[k];
  }
  // coarse-grid AO (pipeline.cpp:68: -1 = the supersample factor;
  // embree_renderer.cpp:531-532: only with AO and something AO can darken)
  {
    const int div = p->ao_res_div < 0 ? ss : p->ao_res_div;
    // (pipeline.cpp:72-80: off under GI)
    fc.coarse = (div > 1 && p->ao_samples > 0 && scene->aoReceivers && !gi) ? 1 : 0;
    fc.cW = fc.coarse ? (W + div - 1) / div : 0;
    fc.cH = fc.coarse ? (H + div - 1) / div : 0;
    fc.seedW = W;
  }
  fc.latW = W;
  fc.latH = H;
  fc.smMode = ubg::kSmGrid;
  fc.nLights = p->n_lights;
  for (int i = 0; i < p->n_lights; ++i) {
    const ubg_light& dl = p->lights[i];
    ubg::DevLight& l = fc.lights[i];
    const V3h L = normalize(V3h{-dl.direction[0], -dl.direction[1], -dl.direction[2]});
    l.L[0] = L.x, l.L[1] = L.y, l.L[2] = L.z;
    l.color[0] = dl.color[0] * dl.intensity;
    l.color[1] = dl.color[1] * dl.intensity;
    l.color[2] = dl.color[2] * dl.intensity;
    l.highlight = dl.casts_highlight ? 1 : 0;
    l.radius = (p->gi_integrator == 2 && dl.angular_radius > 0.0f) ? dl.angular_radius
                                                                   : p->light_radius_deg * 0.01745329252f;
    l.tanRadius = std::tan(l.radius);
    l.soft = (p->shadow_samples > 1 && l.radius > 0.0f) ? 1 : 0;
  }

  return UBG_OK;
}

// One frame; haveLock: the caller already holds ctx->mu.
static int renderFrame(ubg_context* ctx, const ubg_scene* scene, const ubg_render_params* p, ubg_frame* fr,
                       bool haveLock) {
  if (!ctx || !scene || !p || !fr) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render: null argument");
  if (scene->ctx != ctx) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render: scene belongs to another context");
  if (p->width <= 0 || p->height <= 0) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render: width/height must be > 0");
  if (p->n_lights < 0 || (p->n_lights > 0 && !p->lights))
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render: bad light array");
  if (p->n_lights > ubg::kMaxLights)
    return fail(UBG_ERR_UNSUPPORTED, "ubg_render: more than 16 lights");
  if (p->ao_samples < 0 || p->shadow_samples < 0 || p->max_transparent_layers < 0)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render: negative sample or layer count");
  const int ss = std::max(1, p->supersample);
  const int64_t W64 = static_cast<int64_t>(p->width) * ss, H64 = static_cast<int64_t>(p->height) * ss;
  if (W64 * H64 >= (int64_t(1) << 31) || W64 * H64 * 4 >= (int64_t(1) << 40))
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render: frame too large");
  const int W = static_cast<int>(W64), H = static_cast<int>(H64);
  const int npx = W * H;
  const bool countVisits = (p->flags & UBG_FLAG_COUNT_VISITS) != 0u;

  std::unique_lock<std::mutex> lk(ctx->mu, std::defer_lock);
  if (!haveLock) lk.lock();
  UBG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t st = ctx->stream;

  // ---- frame constants (embree_renderer.cpp:375-386, 436-458) ----
  ubg::FrameConst fc;
  int giDiv = 1;
  {
    const int rs = frameConstants(ctx, scene, p, fr, W, H, ss, fc, giDiv);
    if (rs != UBG_OK) return rs;
  }
  const bool gi = p->gi != 0;
  const ubg_camera& cam = p->camera;
  const int giLw = (W + giDiv - 1) / giDiv, giLh = (H + giDiv - 1) / giDiv;
  const size_t giLn = static_cast<size_t>(giLw) * giLh;

  // ---- adaptive AA plan (embree_renderer.cpp:1379-1385, 1430-1434, 1549, 1597-1602) ----
  // (pipeline.cpp:60-65: adaptive AA falls back to the grid under GI)
  const bool adaptive = p->aa_mode == 1 && (ss > 1 || p->aa_depth > 1) && !gi;
  const int kFine = std::max(1, (p->aa_depth + ss - 1) / ss);
  const int64_t nOut = static_cast<int64_t>(p->width) * p->height;
  // stochastic effects a replicated centre is re-evaluated for, with ss^2
  // times their samples (AO not when the coarse grid supplies it)
  const bool softShadows = p->shadow_samples > 1 && p->light_radius_deg > 0.0f;
  const bool needBoost = adaptive && ss > 1 && ((p->ao_samples > 0 && !fc.coarse) || softShadows);
  const int aoMul = (needBoost && !fc.coarse) ? ss * ss : 1;
  const int shMul = (needBoost && softShadows) ? ss * ss : 1;
  int64_t maxRefined = 0;  // refined samples if every output pixel were flagged
  if (adaptive) {
    if (p->aa_mode != 1 || p->aa_depth < 0)
      return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render: bad adaptive AA settings");
    const int64_t per = kFine == 1 ? static_cast<int64_t>(ss) * ss - 1 : static_cast<int64_t>(ss) * ss * kFine * kFine;
    maxRefined = nOut * per;
    if (maxRefined >= (int64_t(1) << 31) || static_cast<int64_t>(W) * kFine >= (int64_t(1) << 31) ||
        static_cast<int64_t>(H) * kFine >= (int64_t(1) << 31))
      return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render: aa_depth too large for this frame");
    if (static_cast<int64_t>(p->ao_samples) * aoMul >= (int64_t(1) << 24) ||
        static_cast<int64_t>(p->shadow_samples) * shMul >= (int64_t(1) << 24))
      return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render: too many samples for adaptive AA");
  } else if (p->aa_mode != 0 && p->aa_mode != 1) {
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render: aa_mode must be 0 or 1");
  }

  // ---- frame scratch ----
  const size_t n = static_cast<size_t>(npx);
  int r = UBG_OK;
  auto need = [&](DevBuf& b, size_t bytes) {
    if (r == UBG_OK) r = ensure(b, bytes);
  };
  // queues: kQueueParts partitions of partCap slots
  const int partCap = ubg::queuePartCap(W, H);
  const size_t qn = static_cast<size_t>(ubg::kQueueParts) * static_cast<size_t>(partCap);
  need(ctx->hitPix, qn * 4);
  need(ctx->hitT, qn * 4);
  need(ctx->hitO, qn * 16);
  need(ctx->hitN, qn * 16);
  need(ctx->hitC, qn * 16);
  need(ctx->hitAo, qn * 4);
  if (fc.aoEnh && p->ao_samples > 0) need(ctx->aoT, qn * static_cast<size_t>(p->ao_samples) * aoMul * 4);
  if (fc.coarse) {
    const size_t cells = static_cast<size_t>(fc.cW) * fc.cH;
    need(ctx->hitOpen, qn * 4);
    need(ctx->cGuide, cells * 16);
    need(ctx->cOpen, cells * 4);
    if (fc.aoEnh) {
      need(ctx->hitBent, qn * 16);
      need(ctx->cBent, cells * 16);
    }
  }
  need(ctx->activeA, qn * 4);
  need(ctx->activeB, qn * 4);
  need(ctx->acc, n * 16);
  need(ctx->tnear, n * 4);
  need(ctx->hiColor, n * 16);
  need(ctx->depth, n * 4);
  need(ctx->viewZ, n * 4);
  if (fr->hit_id) need(ctx->hitId, n * 4);
  if (fr->ao_open) need(ctx->aoOpen, n * 4);
  const size_t outN = static_cast<size_t>(p->width) * p->height;
  need(ctx->outColor, outN * 16);
  need(ctx->counters, ubg::kCounterWords * sizeof(unsigned long long));
  if (scene->anisoPrims) need(ctx->hitPrim, qn * 4);
  if (adaptive) {
    need(ctx->hitOut, qn * 4);
    need(ctx->aaSetC, static_cast<size_t>(nOut) * 32);
    need(ctx->aaFirstN, static_cast<size_t>(nOut) * 16);
    need(ctx->aaPix, static_cast<size_t>(nOut) * 18);  // slot, flagged, replicated (u32), mask (f32), raw, mask (u8)
  }
  if (gi) {
    need(ctx->giPos, n * 16);
    need(ctx->giNrm, n * 16);
    need(ctx->giRefl, n * 16);
    need(ctx->giE, n * 16);
    need(ctx->giList, qn * 4);
    if (fr->gi_indirect) need(ctx->giInd, n * 12);
    if (fr->gi_features || p->pt1_denoise == 2) need(ctx->giMax, 256);
    if (fr->gi_features) need(ctx->giFeat, n * 40);
    if (p->pt1_denoise == 2) {
      const int m = 1 << (p->pt1_denoise_net->net.levels() - 1);
      const size_t np = static_cast<size_t>((W + m - 1) / m * m) * ((H + m - 1) / m * m);
      need(ctx->giFeatP, np * ubg::kUnetFeatures * 4);
      need(ctx->giPred, np * 12);
    }
    if (fr->gi_occlusion) need(ctx->giOcc, n * 4);
    if (p->pt1_denoise) {
      need(ctx->dnA, n * 16);
      need(ctx->dnB, n * 16);
      need(ctx->dnGrad, n * 4);
      need(ctx->dnVar, n * 4);
    }
    if (giDiv > 1) {
      // reduced grid: position, normal (+ hit flag), geometric normal, albedo,
      // depth, E; the rim-patch mask of the render grid and the E snapshot
      need(ctx->giLow, giLn * (5 * 16 + 4));
      need(ctx->giPatch, n);
      need(ctx->giSnap, n * 16);
    }
  }
  if (r != UBG_OK) return r;

  ubg::FrameBufs b;
  b.hitPix = static_cast<uint32_t*>(ctx->hitPix.p);
  b.hitT = static_cast<float*>(ctx->hitT.p);
  b.hitO = static_cast<float4*>(ctx->hitO.p);
  b.hitN = static_cast<float4*>(ctx->hitN.p);
  b.hitC = static_cast<float4*>(ctx->hitC.p);
  b.hitAo = static_cast<uint32_t*>(ctx->hitAo.p);
  b.aoT = (fc.aoEnh && p->ao_samples > 0) ? static_cast<float*>(ctx->aoT.p) : nullptr;
  b.hitOpen = fc.coarse ? static_cast<float*>(ctx->hitOpen.p) : nullptr;
  b.hitBent = (fc.coarse && fc.aoEnh) ? static_cast<float4*>(ctx->hitBent.p) : nullptr;
  b.cGuide = fc.coarse ? static_cast<float4*>(ctx->cGuide.p) : nullptr;
  b.cOpen = fc.coarse ? static_cast<float*>(ctx->cOpen.p) : nullptr;
  b.cBent = (fc.coarse && fc.aoEnh) ? static_cast<float4*>(ctx->cBent.p) : nullptr;
  b.activeIn = static_cast<uint32_t*>(ctx->activeA.p);
  b.activeOut = static_cast<uint32_t*>(ctx->activeB.p);
  b.acc = static_cast<float4*>(ctx->acc.p);
  b.tnear = static_cast<float*>(ctx->tnear.p);
  b.hiColor = static_cast<float4*>(ctx->hiColor.p);
  b.depth = static_cast<float*>(ctx->depth.p);
  b.viewZ = static_cast<float*>(ctx->viewZ.p);
  b.hitId = fr->hit_id ? static_cast<uint32_t*>(ctx->hitId.p) : nullptr;
  b.aoOpen = fr->ao_open ? static_cast<float*>(ctx->aoOpen.p) : nullptr;
  b.counters = static_cast<unsigned long long*>(ctx->counters.p);
  b.partCap = partCap;
  b.hitOut = nullptr;
  b.hitPrim = scene->anisoPrims ? static_cast<uint32_t*>(ctx->hitPrim.p) : nullptr;
  b.principled = scene->principled ? 1 : 0;
  b.firstN = nullptr;
  b.smList = nullptr;
  b.giPos = gi ? static_cast<float4*>(ctx->giPos.p) : nullptr;
  b.giNrm = gi ? static_cast<float4*>(ctx->giNrm.p) : nullptr;
  b.giRefl = gi ? static_cast<float4*>(ctx->giRefl.p) : nullptr;
  b.giE = gi ? static_cast<float4*>(ctx->giE.p) : nullptr;
  b.giList = gi ? static_cast<uint32_t*>(ctx->giList.p) : nullptr;

  fr->cancelled = 0;
  fr->layers = 0;
  for (int k = 0; k < 4; ++k) {
    fr->kernel_ms[k] = 0.0;
    fr->kernel_launches[k] = 0;
  }
  // Optional per-kernel stream events (UBG_FLAG_KERNEL_TIMES).
  const bool timing = (p->flags & UBG_FLAG_KERNEL_TIMES) != 0u;
  struct Mark {
    int kind;
    size_t e0, e1;
  };
  std::vector<Mark> marks;
  size_t nextEv = 0;
  auto evAt = [&](size_t i) -> hipEvent_t {
    while (ctx->tev.size() <= i) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      ctx->tev.push_back(e);
    }
    return ctx->tev[i];
  };
  auto timed = [&](int kind, auto&& launch) {
    if (!timing) {
      launch();
      return;
    }
    const size_t e0 = nextEv++, e1 = nextEv++;
    hipEvent_t a = evAt(e0), bb = evAt(e1);
    if (a && bb) (void)hipEventRecord(a, st);
    launch();
    if (a && bb) {
      (void)hipEventRecord(bb, st);
      marks.push_back({kind, e0, e1});
    }
  };
  ubg::launchZeroCounters(b.counters, ubg::kCounterWords, st);
  UBG_HIP(hipEventRecord(ctx->ev[0], st), "hipEventRecord");
  if (fc.coarse) {
    // coarse-grid AO pre-pass (embree_renderer.cpp:527-553): G-buffer cells
    // through the hit queue, the AO kernels on them with the coarse seed
    // lattice, then the per-cell reduction into the grid
    ubg::launchResetQueues(b.counters, st);
    timed(UBG_KERNEL_PRIMARY, [&] { ubg::launchCoarseGbuf(fc, scene->view, b, countVisits, st); });
    ubg::FrameConst fcCoarse = fc;
    fcCoarse.seedW = fc.cW;
    timed(UBG_KERNEL_AO, [&] { ubg::launchAo(fcCoarse, scene->view, b, countVisits, st); });
    timed(UBG_KERNEL_AO, [&] { ubg::launchCoarseReduce(fcCoarse, b, st); });
  }
  // One pass: the transparency layers of `nFirst` samples (the first layer
  // per the pass's sample map, then the continuation queue).
  bool cancelled = false;
  auto runLayers = [&](const ubg::FrameConst& f, ubg::FrameBufs bb, int64_t nFirst) -> int {
    int64_t nActive = nFirst;
    ubg::PartFill fill{};  // continuation-queue fills the next layer consumes
    for (int iter = 0; iter < f.maxIters && nActive > 0; ++iter) {
      if (p->cancel_flag && *p->cancel_flag) {
        cancelled = true;
        return UBG_OK;
      }
      ++fr->layers;
      ubg::launchResetQueues(bb.counters, st);
      timed(UBG_KERNEL_PRIMARY, [&] { ubg::launchPrimary(f, scene->view, bb, iter, fill, countVisits, st); });
      if (f.aoSamples > 0) timed(UBG_KERNEL_AO, [&] { ubg::launchAo(f, scene->view, bb, countVisits, st); });
      timed(UBG_KERNEL_RESOLVE, [&] { ubg::launchResolve(f, scene->view, bb, iter, countVisits, st); });
      UBG_HIP(hipGetLastError(), "kernel launch");
      if (iter + 1 >= f.maxIters) break;
      // One host round trip per transparency layer: the continuation fills.
      UBG_HIP(hipMemcpyAsync(ctx->hostCounters, bb.counters, ubg::kCounterWords * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, st),
              "hipMemcpy(counters)");
      UBG_HIP(hipStreamSynchronize(st), "layer sync");
      nActive = 0;
      for (int q = 0; q < ubg::kQueueParts; ++q) {
        fill.n[q] = static_cast<uint32_t>(ctx->hostCounters[ubg::kActiveFill + ubg::kLineWords * q]);
        nActive += fill.n[q];
      }
      std::swap(bb.activeIn, bb.activeOut);
    }
    return UBG_OK;
  };
  // Adaptive AA (embree_renderer.cpp:1428-1674): phase 1 shades the centre
  // subpixel of every output pixel, phase 2 builds the refinement mask from
  // those results, phase 3 shades the refined subpixels (or their k x k fine
  // samples) and re-evaluates the replicated centres with boosted sample
  // counts; a compose pass then writes the hi-res buffers and the fog. Every
  // pass is the per-layer wavefront pipeline over a sample map; passes larger
  // than the hi-res grid run in chunks of at most W x H samples.
  auto renderAdaptive = [&]() -> int {
    const int Wf = p->width, Hf = p->height;
    const int cSub = ss / 2;
    const bool wantHitId = fr->hit_id != nullptr, wantAo = fr->ao_open != nullptr;
    // a sample set of `cap` samples: colour 16 B, depth, view z, hit id, AO 4 B each
    auto setOf = [&](DevBuf& d, int64_t cap) {
      ubg::AaSet a;
      char* base = static_cast<char*>(d.p);
      a.color = reinterpret_cast<float4*>(base);
      a.depth = reinterpret_cast<float*>(base + 16 * cap);
      a.viewZ = reinterpret_cast<float*>(base + 20 * cap);
      a.hitId = wantHitId ? reinterpret_cast<uint32_t*>(base + 24 * cap) : nullptr;
      a.aoOpen = wantAo ? reinterpret_cast<float*>(base + 28 * cap) : nullptr;
      return a;
    };
    auto bufsFor = [&](const ubg::AaSet& a, int64_t off) {
      ubg::FrameBufs bb = b;
      bb.hiColor = a.color + off;
      bb.depth = a.depth + off;
      bb.viewZ = a.viewZ + off;
      bb.hitId = a.hitId ? a.hitId + off : nullptr;
      bb.aoOpen = a.aoOpen ? a.aoOpen + off : nullptr;
      bb.hitOut = static_cast<uint32_t*>(ctx->hitOut.p);
      bb.firstN = nullptr;
      return bb;
    };
    // a list pass over `total` samples, in chunks the queues can hold
    auto runList = [&](ubg::FrameConst f, const ubg::AaSet& a, const uint32_t* list, int64_t total) -> int {
      const int64_t chunk = npx;
      for (int64_t base = 0; base < total && !cancelled; base += chunk) {
        f.smBase = static_cast<uint32_t>(base);
        f.nList = static_cast<uint32_t>(std::min(chunk, total - base));
        ubg::FrameBufs bb = bufsFor(a, base);
        bb.smList = list;
        const int rr = runLayers(f, bb, f.nList);
        if (rr != UBG_OK) return rr;
      }
      return UBG_OK;
    };
    const ubg::AaSet cs = setOf(ctx->aaSetC, nOut);
    char* pixBase = static_cast<char*>(ctx->aaPix.p);
    uint32_t* slotOf = reinterpret_cast<uint32_t*>(pixBase);
    uint32_t* flagList = slotOf + nOut;
    uint32_t* repList = flagList + nOut;
    float* maskF = reinterpret_cast<float*>(repList + nOut);
    uint8_t* raw = reinterpret_cast<uint8_t*>(maskF + nOut);
    uint8_t* mask = raw + nOut;

    // phase 1: the centres, at their true hi-res seeds and plain sample counts
    ubg::FrameConst fa = fc;
    fa.fogOn = 0;  // fog is applied when the hi-res buffers are composed
    fa.smMode = ubg::kSmCenter;
    fa.smSs = ss;
    fa.smK = 1;
    fa.smCSub = cSub;
    fa.smWf = Wf;
    fa.smHf = Hf;
    fa.smPer = 1;
    fa.smBase = 0;
    fa.nList = static_cast<uint32_t>(nOut);
    {
      ubg::FrameBufs ba = bufsFor(cs, 0);
      ba.firstN = static_cast<float4*>(ctx->aaFirstN.p);
      const int rr = runLayers(fa, ba, nOut);
      if (rr != UBG_OK || cancelled) return rr;
    }
    // phase 2: the refinement mask and the pixel lists
    const float pws = cam.orthographic ? (cam.height / static_cast<float>(Hf))
                                       : (2.0f * fc.persHalfH / static_cast<float>(Hf));
    ubg::launchAaMask(fa, cs, static_cast<const float4*>(ctx->aaFirstN.p), pws, p->aa_threshold, raw, mask, slotOf,
                      flagList, repList, needBoost, fr->aa_mask ? maskF : nullptr, b.counters, st);
    UBG_HIP(hipGetLastError(), "kernel launch");
    if (fr->aa_mask) {
      const int rr = copyOut(fr->aa_mask, maskF, static_cast<size_t>(nOut) * 4, fr->memory, st);
      if (rr != UBG_OK) return rr;
    }
    UBG_HIP(hipMemcpyAsync(ctx->hostCounters, b.counters, ubg::kCounterWords * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, st),
            "hipMemcpy(counters)");
    UBG_HIP(hipStreamSynchronize(st), "mask sync");
    const int64_t nFlag = static_cast<int64_t>(ctx->hostCounters[ubg::kAaFlagFill]);
    const int64_t nRep = needBoost ? static_cast<int64_t>(ctx->hostCounters[ubg::kAaRepFill]) : 0;
    fr->aa_flagged = nFlag;
    // phase 3a: replicated centres re-evaluated with ss^2 x the stochastic samples
    ubg::AaSet rs{};
    if (nRep > 0) {
      int rr = ensure(ctx->aaSetR, static_cast<size_t>(nRep) * 32);
      if (rr != UBG_OK) return rr;
      rs = setOf(ctx->aaSetR, nRep);
      ubg::FrameConst fb = fa;
      fb.aoSamples = p->ao_samples * aoMul;
      int lanesB = 1;
      while (lanesB < fb.aoSamples && lanesB < 64) lanesB <<= 1;
      fb.aoLanes = lanesB;
      fb.shadowSamples = p->shadow_samples * shMul;
      rr = runList(fb, rs, repList, nRep);
      if (rr != UBG_OK || cancelled) return rr;
    }
    // phase 3b: every subpixel of the flagged pixels (k x k fine samples each when k > 1)
    ubg::AaSet fs{};
    if (nFlag > 0) {
      const int64_t per = kFine == 1 ? static_cast<int64_t>(ss) * ss - 1
                                     : static_cast<int64_t>(ss) * ss * kFine * kFine;
      const int64_t total = nFlag * per;
      int rr = ensure(ctx->aaSetF, static_cast<size_t>(total) * 32);
      if (rr != UBG_OK) return rr;
      fs = setOf(ctx->aaSetF, total);
      ubg::FrameConst ff = fa;
      ff.smMode = kFine == 1 ? ubg::kSmCells : ubg::kSmFine;
      ff.smK = kFine;
      ff.smPer = static_cast<int32_t>(per);
      ff.latW = W * kFine;
      ff.latH = H * kFine;
      ff.seedW = W * kFine;
      rr = runList(ff, fs, flagList, total);
      if (rr != UBG_OK || cancelled) return rr;
    }
    // compose the hi-res buffers (+ fog)
    ubg::FrameConst fcomp = fc;
    fcomp.smSs = ss;
    fcomp.smK = kFine;
    fcomp.smCSub = cSub;
    fcomp.smWf = Wf;
    fcomp.smHf = Hf;
    ubg::launchAaCompose(fcomp, cs, rs, fs, mask, slotOf, nRep > 0, b, st);
    UBG_HIP(hipGetLastError(), "kernel launch");
    return UBG_OK;
  };
  fr->aa_flagged = 0;
  if (adaptive) {
    r = renderAdaptive();
  } else if (gi) {
    // the layers without the fused fog, then the gather and the composite
    // (which applies the fog after adding the indirect, pipeline.cpp:126-129)
    ubg::FrameConst fl = fc;
    fl.fogOn = 0;
    r = runLayers(fl, b, npx);
    if (r == UBG_OK && !cancelled) {
      // the render grid as a gather field (no geometric normal)
      ubg::GiField full;
      full.w = W;
      full.h = H;
      full.pos = b.giPos;
      full.nrm = b.giNrm;
      full.geoN = nullptr;
      full.depth = b.depth;
      full.E = b.giE;
      full.list = b.giList;
      full.fillBase = ubg::kGiFill;
      full.spp = fc.giSpp;
      auto denoise = [&](const ubg::GiField& f) {
        if (p->pt1_denoise)
          ubg::launchGiDenoise(f, p->denoise_iters, p->denoise_sigma_z, p->denoise_sigma_n, p->denoise_sigma_l,
                               static_cast<float4*>(ctx->dnA.p), static_cast<float4*>(ctx->dnB.p),
                               static_cast<float*>(ctx->dnGrad.p), static_cast<float*>(ctx->dnVar.p), st);
      };
      if (giDiv <= 1) {
        // embree_renderer.cpp:959-1011: gather (and denoise) on the render grid
        ubg::launchGiList(full, nullptr, true, b, st);
        timed(UBG_KERNEL_AO, [&] { ubg::launchGiGather(fc, scene->view, b, full, countVisits, st); });
        if (fr->gi_features)
          ubg::launchGiFeatures(W, H, W, H, b, static_cast<unsigned int*>(ctx->giMax.p), 10,
                                static_cast<float*>(ctx->giFeat.p), st);
        if (p->pt1_denoise == 2) {
          // the learned denoiser in place of the a-trous filter: features on
          // the grid padded to the network's multiple, the prediction back
          // onto the receivers
          ubg::Unet& net = p->pt1_denoise_net->net;
          const int m = 1 << (net.levels() - 1);
          const int Wp = (W + m - 1) / m * m, Hp = (H + m - 1) / m * m;
          ubg::launchGiFeatures(W, H, Wp, Hp, b, static_cast<unsigned int*>(ctx->giMax.p), ubg::kUnetFeatures,
                                static_cast<float*>(ctx->giFeatP.p), st);
          UBG_HIP(net.forward(1, Hp, Wp, static_cast<const float*>(ctx->giFeatP.p), static_cast<float*>(ctx->giPred.p),
                              st),
                  "learned denoiser");
          ubg::launchGiApplyNet(W, H, Wp, static_cast<const float*>(ctx->giPred.p), b, st);
        } else {
          denoise(full);
        }
      } else {
        // embree_renderer.cpp:719-958: own G-buffer on the reduced grid, gather
        // and denoise there, joint-bilateral upsample, full-res rim patch
        char* lb = static_cast<char*>(ctx->giLow.p);
        ubg::GiField low;
        low.w = giLw;
        low.h = giLh;
        low.pos = reinterpret_cast<float4*>(lb);
        low.nrm = reinterpret_cast<float4*>(lb + 16 * giLn);
        low.geoN = reinterpret_cast<float4*>(lb + 32 * giLn);
        float4* lowAlb = reinterpret_cast<float4*>(lb + 48 * giLn);
        low.E = reinterpret_cast<float4*>(lb + 64 * giLn);
        low.depth = reinterpret_cast<float*>(lb + 80 * giLn);
        low.list = b.giList;
        low.fillBase = ubg::kGiFill;
        low.spp = fc.giSpp;
        (void)lowAlb;
        timed(UBG_KERNEL_PRIMARY, [&] { ubg::launchGiGbuffer(fc, scene->view, low, lowAlb, b.counters, st); });
        ubg::launchGiList(low, nullptr, true, b, st);
        timed(UBG_KERNEL_AO, [&] { ubg::launchGiGather(fc, scene->view, b, low, countVisits, st); });
        denoise(low);
        uint8_t* patch = p->pt1_edge_patch ? static_cast<uint8_t*>(ctx->giPatch.p) : nullptr;
        ubg::launchGiUpsample(fc, b, low, p->pt1_upsample_normal_pow, p->pt1_upsample_depth_scale,
                              giDiv <= 2 ? 0 : giDiv, patch, p->pt1_edge_patch_thresh, st);
        if (patch) {
          // the rims at full resolution with spp x pt1EdgePatchSppMul samples,
          // then the guided smoothing of exactly those pixels
          ubg::GiField pf = full;
          pf.fillBase = ubg::kGiPatchFill;
          pf.spp = fc.giSpp * std::max(1, p->pt1_edge_patch_spp_mul);
          ubg::launchGiList(pf, patch, false, b, st);
          timed(UBG_KERNEL_AO, [&] { ubg::launchGiGather(fc, scene->view, b, pf, countVisits, st); });
          UBG_HIP(hipMemcpyAsync(ctx->giSnap.p, b.giE, n * 16, hipMemcpyDeviceToDevice, st), "hipMemcpy(E snapshot)");
          ubg::launchGiSmoothPatched(fc, b, patch, p->pt1_upsample_normal_pow, p->pt1_upsample_depth_scale,
                                     static_cast<const float4*>(ctx->giSnap.p), st);
        }
      }
      ubg::launchGiComposite(fc, b, fr->gi_indirect ? static_cast<float*>(ctx->giInd.p) : nullptr,
                             fr->gi_occlusion ? static_cast<float*>(ctx->giOcc.p) : nullptr, st);
      UBG_HIP(hipGetLastError(), "kernel launch");
    }
  } else {
    r = runLayers(fc, b, npx);
  }
  if (r != UBG_OK) return r;
  UBG_HIP(hipEventRecord(ctx->ev[1], st), "hipEventRecord");
  if (cancelled) {
    UBG_HIP(hipStreamSynchronize(st), "cancel sync");
    fr->cancelled = 1;
    return fail(UBG_ERR_CANCELLED, "ubg_render: cancelled");
  }
  const bool hiresOnly = (p->flags & UBG_FLAG_HIRES_ONLY) != 0u;
  if (!hiresOnly)
    timed(UBG_KERNEL_POST, [&] {
      ubg::launchPost(b.hiColor, W, H, ss, p->assumed_gamma, static_cast<float4*>(ctx->outColor.p), st);
    });
  UBG_HIP(hipEventRecord(ctx->ev[2], st), "hipEventRecord");
  UBG_HIP(hipGetLastError(), "kernel launch");
  // ---- outputs ----
  fr->hi_w = W;
  fr->hi_h = H;
  fr->out_w = hiresOnly ? W : p->width;
  fr->out_h = hiresOnly ? H : p->height;
  r = copyOut(fr->color, hiresOnly ? ctx->hiColor.p : ctx->outColor.p,
              (hiresOnly ? n : outN) * 16, fr->memory, st);
  if (r == UBG_OK) r = copyOut(fr->depth, ctx->depth.p, n * 4, fr->memory, st);
  if (r == UBG_OK) r = copyOut(fr->view_z, ctx->viewZ.p, n * 4, fr->memory, st);
  if (r == UBG_OK && fr->hit_id) r = copyOut(fr->hit_id, ctx->hitId.p, n * 4, fr->memory, st);
  if (r == UBG_OK && fr->ao_open) r = copyOut(fr->ao_open, ctx->aoOpen.p, n * 4, fr->memory, st);
  if (r == UBG_OK && gi && fr->gi_indirect) r = copyOut(fr->gi_indirect, ctx->giInd.p, n * 12, fr->memory, st);
  if (r == UBG_OK && gi && fr->gi_occlusion) r = copyOut(fr->gi_occlusion, ctx->giOcc.p, n * 4, fr->memory, st);
  if (r == UBG_OK && gi && fr->gi_features) r = copyOut(fr->gi_features, ctx->giFeat.p, n * 40, fr->memory, st);
  if (r != UBG_OK) return r;
  UBG_HIP(hipMemcpyAsync(ctx->hostCounters, b.counters, ubg::kCounterWords * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, st),
          "hipMemcpy(counters)");
  UBG_HIP(hipStreamSynchronize(st), "frame sync");
  // statistics: sum of the per-block-group partial counts
  unsigned long long c[ubg::kNumCounters] = {0};
  for (int g = 0; g < ubg::kStatSpread; ++g)
    for (int k = 0; k < ubg::kNumCounters; ++k) c[k] += ctx->hostCounters[g * ubg::kNumCounters + k];
  float msTrace = 0.0f, msPost = 0.0f;
  UBG_HIP(hipEventElapsedTime(&msTrace, ctx->ev[0], ctx->ev[1]), "hipEventElapsedTime");
  UBG_HIP(hipEventElapsedTime(&msPost, ctx->ev[1], ctx->ev[2]), "hipEventElapsedTime");
  fr->render_seconds = msTrace * 1e-3;
  fr->post_seconds = msPost * 1e-3;
  fr->stats.primary_rays = c[ubg::kCntPrimary];
  fr->stats.shaded_hits = c[ubg::kCntShaded];
  fr->stats.ao_rays_traced = c[ubg::kCntAoTraced];
  fr->stats.ao_rays_sampled = c[ubg::kCntAoSampled];
  fr->stats.shadow_rays = c[ubg::kCntShadow];
  fr->stats.capped_rays = c[ubg::kCntCapped];
  fr->stats.gi_rays = c[ubg::kCntGiRays];
  fr->stats.gi_nee_rays = c[ubg::kCntGiNee];
  fr->stats.gi_gbuffer_rays = c[ubg::kCntGiGbuf];
  for (int k = 0; k < 3; ++k) {
    fr->stats.node_visits[k] = countVisits ? c[ubg::kCntNodes + k] : 0;
    fr->stats.prim_tests[k] = countVisits ? c[ubg::kCntPrimTests + k] : 0;
  }
  for (const Mark& m : marks) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ctx->tev[m.e0], ctx->tev[m.e1]) == hipSuccess) {
      fr->kernel_ms[m.kind] += ms;
      fr->kernel_launches[m.kind] += 1;
    }
  }
  return UBG_OK;
}

static int runQuery(ubg_context* ctx, const ubg_scene* scene, int64_t n, const ubg_ray* rays,
                    void* out, size_t outBytes, int32_t memory, bool any) {
  if (!ctx || !scene || (n > 0 && (!rays || !out)) || n < 0)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg query: bad argument");
  if (scene->ctx != ctx) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg query: scene belongs to another context");
  if (n == 0) return UBG_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  UBG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t st = ctx->stream;
  const void* dRays = rays;
  void* dOut = out;
  if (!memory) {
    int r = ensure(ctx->query, static_cast<size_t>(n) * sizeof(ubg_ray));
    if (r == UBG_OK) r = ensure(ctx->queryOut, static_cast<size_t>(n) * outBytes);
    if (r != UBG_OK) return r;
    UBG_HIP(hipMemcpyAsync(ctx->query.p, rays, static_cast<size_t>(n) * sizeof(ubg_ray),
                           hipMemcpyHostToDevice, st),
            "hipMemcpy(rays)");
    dRays = ctx->query.p;
    dOut = ctx->queryOut.p;
  }
  ubg::launchQuery(scene->view, dRays, n, any ? nullptr : dOut, any ? static_cast<uint8_t*>(dOut) : nullptr,
                   any, st);
  UBG_HIP(hipGetLastError(), "kernel launch");
  if (!memory)
    UBG_HIP(hipMemcpyAsync(out, dOut, static_cast<size_t>(n) * outBytes, hipMemcpyDeviceToHost, st),
            "hipMemcpy(query result)");
  UBG_HIP(hipStreamSynchronize(st), "query sync");
  return UBG_OK;
}

int ubg_render(ubg_context* ctx, const ubg_scene* scene, const ubg_render_params* p, ubg_frame* fr) {
  return renderFrame(ctx, scene, p, fr, false);
}

// ---- batched frames ---------------------------------------------------------
// n requests rendered by the same launches (ub_kernels.hip, k_*_batch): every
// request keeps its own scene, frame constants, queues and outputs, carved
// from one shared pool; per transparency layer one launch of each kernel
// covers every request still walking layers, and one read-back brings all
// their continuation fills. Each request's image and ray counts are those of
// ubg_render on it alone.
int ubg_render_batch(ubg_context* ctx, int32_t n, const ubg_scene* const* scenes, const ubg_render_params* params,
                     ubg_frame* frames, const ubg_batch_options* opts) {
  if (!ctx || n <= 0 || !scenes || !params || !frames)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render_batch: bad argument");
  if (n > ubg::kMaxBatch)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render_batch: more than " + std::to_string(ubg::kMaxBatch) +
                                              " requests in one batch");
  const int wavesPerSimd = opts && opts->ao_waves_per_simd > 0 ? std::min(opts->ao_waves_per_simd, ubg::kAoWavesPerSimd)
                                                                : ubg::kAoWavesPerSimd;
  const bool useGraph = !opts || opts->no_graph == 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  UBG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t st = ctx->stream;
  struct Req {
    int W, H, ss, partCap;
    size_t npx, qn, outN, off;
  };
  std::vector<Req> rq(static_cast<size_t>(n));
  std::vector<ubg::BatchItem> items(static_cast<size_t>(n));
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  size_t pool = 0;
  for (int i = 0; i < n; ++i) {
    const ubg_scene* sc = scenes[i];
    const ubg_render_params* p = &params[i];
    ubg_frame* fr = &frames[i];
    const std::string who = "ubg_render_batch: request " + std::to_string(i) + ": ";
    if (!sc || sc->ctx != ctx) return fail(UBG_ERR_INVALID_ARGUMENT, who + "scene missing or on another context");
    if (sc->view.layout != ubg::kLayoutBvh8 || sc->principled || sc->anisoPrims)
      return fail(UBG_ERR_UNSUPPORTED, who + "batches take BVH8 scenes with POV materials");
    if (p->width <= 0 || p->height <= 0 || p->n_lights < 0 || (p->n_lights > 0 && !p->lights) ||
        p->n_lights > ubg::kMaxLights || p->ao_samples < 0 || p->shadow_samples < 0 ||
        p->max_transparent_layers < 0)
      return fail(UBG_ERR_INVALID_ARGUMENT, who + "bad frame parameters");
    if (p->flags & (UBG_FLAG_COUNT_VISITS | UBG_FLAG_KERNEL_TIMES))
      return fail(UBG_ERR_UNSUPPORTED, who + "visit counting and kernel timing are single-frame options");
    Req& q = rq[i];
    q.ss = std::max(1, p->supersample);
    const int64_t W64 = static_cast<int64_t>(p->width) * q.ss, H64 = static_cast<int64_t>(p->height) * q.ss;
    if (W64 * H64 >= (int64_t(1) << 31)) return fail(UBG_ERR_INVALID_ARGUMENT, who + "frame too large");
    q.W = static_cast<int>(W64);
    q.H = static_cast<int>(H64);
    ubg::FrameConst& fc = items[i].fc;
    int giDiv = 1;
    const int rs = frameConstants(ctx, sc, p, fr, q.W, q.H, q.ss, fc, giDiv);
    if (rs != UBG_OK) return rs;
    const bool adaptive = p->aa_mode == 1 && (q.ss > 1 || p->aa_depth > 1);
    if (fc.gi || fc.coarse || fc.aoEnh || adaptive)
      return fail(UBG_ERR_UNSUPPORTED, who + "GI, coarse or enhanced AO and adaptive AA are single-frame options");
    q.partCap = ubg::queuePartCap(q.W, q.H);
    q.qn = static_cast<size_t>(ubg::kQueueParts) * q.partCap;
    q.npx = static_cast<size_t>(q.W) * q.H;
    q.outN = static_cast<size_t>(p->width) * p->height;
    q.off = pool;
    // queues (pix, t, O, N, C, AO count, two continuation lists), per-pixel
    // walk state and outputs, the final colour, the counters
    pool += al(q.qn * 4) * 4 + al(q.qn * 16) * 3 + al(q.qn * 4) + al(q.npx * 16) * 2 + al(q.npx * 4) * 5 +
            al(q.outN * 16) + al(ubg::kCounterWords * sizeof(unsigned long long));
  }
  const size_t launchBytes = al(sizeof(ubg::BatchLaunch));
  const size_t itemBytes = al(sizeof(ubg::BatchItem) * static_cast<size_t>(n));
  const size_t fillBytes = al(sizeof(uint32_t) * ubg::kQueueParts * n);
  const size_t statBytes = al(sizeof(unsigned long long) * ubg::kNumCounters * n);
  const void* before[4] = {ctx->batchPool.p, ctx->batchItems.p, ctx->batchLaunch.p, ctx->batchOut.p};
  int r = ensure(ctx->batchPool, pool);
  if (r == UBG_OK) r = ensure(ctx->batchItems, itemBytes);
  if (r == UBG_OK) r = ensure(ctx->batchLaunch, 4 * launchBytes);
  if (r == UBG_OK) r = ensure(ctx->batchOut, fillBytes + statBytes);
  if (r != UBG_OK) return r;
  if (before[0] != ctx->batchPool.p || before[1] != ctx->batchItems.p || before[2] != ctx->batchLaunch.p ||
      before[3] != ctx->batchOut.p)
    ctx->clearLayerGraphs();
  const size_t hostNeed = itemBytes + 4 * launchBytes + fillBytes + statBytes;
  if (ctx->batchHostBytes < hostNeed) {
    ctx->clearLayerGraphs();
    if (ctx->batchHost) (void)hipHostFree(ctx->batchHost);
    ctx->batchHost = nullptr;
    ctx->batchHostBytes = 0;
    UBG_HIP(hipHostMalloc(&ctx->batchHost, hostNeed, hipHostMallocDefault), "hipHostMalloc(batch staging)");
    ctx->batchHostBytes = hostNeed;
  }
  char* hb = static_cast<char*>(ctx->batchHost);
  auto* hItems = reinterpret_cast<ubg::BatchItem*>(hb);
  ubg::BatchLaunch* hL[4];
  for (int k = 0; k < 4; ++k) hL[k] = reinterpret_cast<ubg::BatchLaunch*>(hb + itemBytes + k * launchBytes);
  auto* hFills = reinterpret_cast<uint32_t*>(hb + itemBytes + 4 * launchBytes);
  auto* hStats = reinterpret_cast<unsigned long long*>(hb + itemBytes + 4 * launchBytes + fillBytes);
  const auto* dItems = static_cast<const ubg::BatchItem*>(ctx->batchItems.p);
  const ubg::BatchLaunch* dL[4];
  for (int k = 0; k < 4; ++k)
    dL[k] = reinterpret_cast<const ubg::BatchLaunch*>(static_cast<char*>(ctx->batchLaunch.p) + k * launchBytes);
  auto* dFills = static_cast<uint32_t*>(ctx->batchOut.p);
  auto* dStats = reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->batchOut.p) + fillBytes);

  for (int i = 0; i < n; ++i) {
    const Req& q = rq[i];
    char* base = static_cast<char*>(ctx->batchPool.p) + q.off;
    auto take = [&](size_t bytes) {
      char* at = base;
      base += al(bytes);
      return at;
    };
    ubg::BatchItem& it = items[i];
    it.s = scenes[i]->view;
    ubg::FrameBufs& b = it.b;
    std::memset(&b, 0, sizeof(b));
    b.hitPix = reinterpret_cast<uint32_t*>(take(q.qn * 4));
    b.hitT = reinterpret_cast<float*>(take(q.qn * 4));
    b.activeIn = reinterpret_cast<uint32_t*>(take(q.qn * 4));
    b.activeOut = reinterpret_cast<uint32_t*>(take(q.qn * 4));
    b.hitO = reinterpret_cast<float4*>(take(q.qn * 16));
    b.hitN = reinterpret_cast<float4*>(take(q.qn * 16));
    b.hitC = reinterpret_cast<float4*>(take(q.qn * 16));
    b.hitAo = reinterpret_cast<uint32_t*>(take(q.qn * 4));
    b.acc = reinterpret_cast<float4*>(take(q.npx * 16));
    b.hiColor = reinterpret_cast<float4*>(take(q.npx * 16));
    b.tnear = reinterpret_cast<float*>(take(q.npx * 4));
    b.depth = reinterpret_cast<float*>(take(q.npx * 4));
    b.viewZ = reinterpret_cast<float*>(take(q.npx * 4));
    uint32_t* hitId = reinterpret_cast<uint32_t*>(take(q.npx * 4));
    float* aoOpen = reinterpret_cast<float*>(take(q.npx * 4));
    b.hitId = frames[i].hit_id ? hitId : nullptr;
    b.aoOpen = frames[i].ao_open ? aoOpen : nullptr;
    it.out = reinterpret_cast<float4*>(take(q.outN * 16));
    b.counters = reinterpret_cast<unsigned long long*>(take(ubg::kCounterWords * sizeof(unsigned long long)));
    b.partCap = q.partCap;
    it.ss = q.ss;
    it.gamma = params[i].assumed_gamma;
    frames[i].cancelled = 0;
    frames[i].layers = 0;
  }
  auto pushItems = [&]() -> int {
    std::memcpy(hItems, items.data(), sizeof(ubg::BatchItem) * static_cast<size_t>(n));
    UBG_HIP(hipMemcpyAsync(ctx->batchItems.p, hItems, sizeof(ubg::BatchItem) * static_cast<size_t>(n),
                           hipMemcpyHostToDevice, st),
            "hipMemcpy(batch items)");
    return UBG_OK;
  };
  auto pushLaunch = [&](int k) -> int {
    UBG_HIP(hipMemcpyAsync(const_cast<ubg::BatchLaunch*>(dL[k]), hL[k], sizeof(ubg::BatchLaunch),
                           hipMemcpyHostToDevice, st),
            "hipMemcpy(batch launch)");
    return UBG_OK;
  };
  if ((r = pushItems()) != UBG_OK) return r;
  ubg::launchZeroBatch(dItems, n, st);
  UBG_HIP(hipEventRecord(ctx->ev[0], st), "hipEventRecord");
  // the requests still walking layers, and their continuation fills
  std::vector<int> live(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) live[i] = i;
  std::vector<ubg::PartFill> fill(static_cast<size_t>(n));
  const int64_t aoResident = ubg::aoResidentBlocks(wavesPerSimd);
  for (int iter = 0; !live.empty(); ++iter) {
    for (int i = 0; i < n; ++i)
      if (params[i].cancel_flag && *params[i].cancel_flag) frames[i].cancelled = 1;
    ubg::BatchLaunch& LP = *hL[0];
    ubg::BatchLaunch& LA = *hL[1];
    ubg::BatchLaunch& LR = *hL[2];
    LP.n = LA.n = LR.n = 0;
    uint32_t blocksP = 0, blocksR = 0;
    int64_t aoRays = 0;
    size_t lds = 0;
    for (int i : live) {
      const Req& q = rq[i];
      const ubg::FrameConst& fc = items[i].fc;
      uint32_t nb;
      int partBlocks = 1;
      if (iter == 0) {
        nb = static_cast<uint32_t>((static_cast<int64_t>((q.W + 7) / 8) * ((q.H + 7) / 8) * 64 + 127) / 128);
      } else {
        uint32_t mx = 0;
        for (int p = 0; p < ubg::kQueueParts; ++p) mx = std::max(mx, fill[i].n[p]);
        partBlocks = static_cast<int>((mx + 127) / 128);
        nb = static_cast<uint32_t>(partBlocks * ubg::kQueueParts);
      }
      const int e = LP.n++;
      LP.item[e] = i;
      LP.partBlocks[e] = partBlocks;
      LP.fill[e] = fill[i];
      blocksP += nb;
      LP.blockEnd[e] = blocksP;
      const int er = LR.n++;
      LR.item[er] = i;
      LR.partBlocks[er] = 1;
      blocksR += static_cast<uint32_t>(q.qn / 128);
      LR.blockEnd[er] = blocksR;
      if (fc.aoSamples > 0) {
        LA.item[LA.n] = i;
        LA.partBlocks[LA.n] = 1;
        LA.blockEnd[LA.n] = 0;
        ++LA.n;
        // rays this layer can spawn: every ray of the layer hits at most once
        const int64_t rays = iter == 0 ? static_cast<int64_t>(q.npx) : static_cast<int64_t>(nb) * 128;
        aoRays += rays * fc.aoSamples;
      }
      lds = std::max(lds, ubg::traceLdsBytes(items[i].s));
      ++frames[i].layers;
    }
    for (int k = 0; k < 3; ++k)
      if ((r = pushLaunch(k)) != UBG_OK) return r;
    const int64_t aoBlocks = LA.n > 0 ? std::min<int64_t>(aoResident, (aoRays + 127) / 128) : 0;
    // The layer: reset, primary, AO, resolve, then one read-back for the
    // whole batch (every continuation fill). Its launches depend only on the
    // shape below (the tables they read are uploaded before), so a layer of a
    // known shape replays an instantiated graph.
    auto layer = [&]() -> int {
      ubg::launchResetBatch(dItems, dL[0], LP.n, st);
      ubg::launchPrimaryBatch(dItems, dL[0], blocksP, iter, lds, st);
      if (aoBlocks > 0) ubg::launchAoBatch(dItems, dL[1], aoBlocks, lds, st);
      ubg::launchResolveBatch(dItems, dL[2], blocksR, iter, lds, st);
      ubg::launchFillsBatch(dItems, n, dFills, st);
      UBG_HIP(hipMemcpyAsync(hFills, dFills, sizeof(uint32_t) * ubg::kQueueParts * n, hipMemcpyDeviceToHost, st),
              "hipMemcpy(batch fills)");
      return UBG_OK;
    };
    if (useGraph) {
      // (the layer index is a kernel argument: the first layer and the last
      // permitted one take other branches)
      const std::vector<int64_t> key{iter, n, LP.n, blocksP, aoBlocks, blocksR, static_cast<int64_t>(lds)};
      auto it = ctx->layerGraphs.find(key);
      if (it == ctx->layerGraphs.end()) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        UBG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        const int lr = layer();
        const hipError_t ce = hipStreamEndCapture(st, &g);
        if (lr != UBG_OK) {
          if (g) (void)hipGraphDestroy(g);
          return lr;
        }
        UBG_HIP(ce, "hipStreamEndCapture");
        const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        UBG_HIP(ie, "hipGraphInstantiate");
        if (ctx->layerGraphs.size() >= 32) ctx->clearLayerGraphs();
        it = ctx->layerGraphs.emplace(key, ge).first;
      }
      UBG_HIP(hipGraphLaunch(it->second, st), "hipGraphLaunch");
    } else {
      if ((r = layer()) != UBG_OK) return r;
    }
    UBG_HIP(hipGetLastError(), "kernel launch");
    UBG_HIP(hipStreamSynchronize(st), "batch layer sync");
    std::vector<int> next;
    for (int i : live) {
      int64_t act = 0;
      for (int p = 0; p < ubg::kQueueParts; ++p) {
        fill[i].n[p] = hFills[i * ubg::kQueueParts + p];
        act += fill[i].n[p];
      }
      std::swap(items[i].b.activeIn, items[i].b.activeOut);
      if (act > 0 && iter + 1 < items[i].fc.maxIters && !frames[i].cancelled) next.push_back(i);
    }
    live.swap(next);
    if (!live.empty() && (r = pushItems()) != UBG_OK) return r;
  }
  UBG_HIP(hipEventRecord(ctx->ev[1], st), "hipEventRecord");
  // resolve every request's frame in one launch
  {
    ubg::BatchLaunch& LQ = *hL[3];
    LQ.n = 0;
    uint32_t blocks = 0;
    for (int i = 0; i < n; ++i) {
      if (params[i].flags & UBG_FLAG_HIRES_ONLY) continue;
      LQ.item[LQ.n] = i;
      LQ.partBlocks[LQ.n] = 1;
      blocks += static_cast<uint32_t>((rq[i].outN + 255) / 256);
      LQ.blockEnd[LQ.n] = blocks;
      ++LQ.n;
    }
    if (LQ.n > 0) {
      if ((r = pushLaunch(3)) != UBG_OK) return r;
      ubg::launchPostBatch(dItems, dL[3], blocks, st);
    }
  }
  UBG_HIP(hipEventRecord(ctx->ev[2], st), "hipEventRecord");
  ubg::launchStatsBatch(dItems, n, dStats, st);
  UBG_HIP(hipGetLastError(), "kernel launch");
  for (int i = 0; i < n && r == UBG_OK; ++i) {
    ubg_frame* fr = &frames[i];
    const Req& q = rq[i];
    const bool hiresOnly = (params[i].flags & UBG_FLAG_HIRES_ONLY) != 0u;
    const ubg::FrameBufs& b = items[i].b;
    fr->hi_w = q.W;
    fr->hi_h = q.H;
    fr->out_w = hiresOnly ? q.W : params[i].width;
    fr->out_h = hiresOnly ? q.H : params[i].height;
    r = copyOut(fr->color, hiresOnly ? static_cast<const void*>(b.hiColor) : static_cast<const void*>(items[i].out),
                (hiresOnly ? q.npx : q.outN) * 16, fr->memory, st);
    if (r == UBG_OK) r = copyOut(fr->depth, b.depth, q.npx * 4, fr->memory, st);
    if (r == UBG_OK) r = copyOut(fr->view_z, b.viewZ, q.npx * 4, fr->memory, st);
    if (r == UBG_OK && fr->hit_id) r = copyOut(fr->hit_id, b.hitId, q.npx * 4, fr->memory, st);
    if (r == UBG_OK && fr->ao_open) r = copyOut(fr->ao_open, b.aoOpen, q.npx * 4, fr->memory, st);
  }
  if (r != UBG_OK) return r;
  UBG_HIP(hipMemcpyAsync(hStats, dStats, sizeof(unsigned long long) * ubg::kNumCounters * n, hipMemcpyDeviceToHost,
                         st),
          "hipMemcpy(batch stats)");
  UBG_HIP(hipStreamSynchronize(st), "batch sync");
  float msTrace = 0.0f, msPost = 0.0f;
  UBG_HIP(hipEventElapsedTime(&msTrace, ctx->ev[0], ctx->ev[1]), "hipEventElapsedTime");
  UBG_HIP(hipEventElapsedTime(&msPost, ctx->ev[1], ctx->ev[2]), "hipEventElapsedTime");
  int status = UBG_OK;
  for (int i = 0; i < n; ++i) {
    ubg_frame* fr = &frames[i];
    const unsigned long long* c = hStats + static_cast<size_t>(i) * ubg::kNumCounters;
    fr->render_seconds = msTrace * 1e-3;  // the batch's trace window
    fr->post_seconds = msPost * 1e-3;
    std::memset(&fr->stats, 0, sizeof(fr->stats));
    fr->stats.primary_rays = c[ubg::kCntPrimary];
    fr->stats.shaded_hits = c[ubg::kCntShaded];
    fr->stats.ao_rays_traced = c[ubg::kCntAoTraced];
    fr->stats.ao_rays_sampled = c[ubg::kCntAoSampled];
    fr->stats.shadow_rays = c[ubg::kCntShadow];
    fr->stats.capped_rays = c[ubg::kCntCapped];
    for (int k = 0; k < 4; ++k) {
      fr->kernel_ms[k] = 0.0;
      fr->kernel_launches[k] = 0;
    }
    if (fr->cancelled) status = UBG_ERR_CANCELLED;
  }
  if (status != UBG_OK) return fail(status, "ubg_render_batch: cancelled");
  return UBG_OK;
}

// ---- physically based path tracing --------------------------------------------
int ubg_pathtrace(ubg_context* ctx, const ubg_scene* scene, const ubg_pt_params* p, ubg_pt_frame* fr) {
  if (!ctx || !scene || !p || !fr || !fr->color) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_pathtrace: null argument");
  if (scene->ctx != ctx) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_pathtrace: scene belongs to another context");
  if (scene->view.layout != ubg::kLayoutBvh8)
    return fail(UBG_ERR_UNSUPPORTED, "ubg_pathtrace: needs a BVH8 scene");
  if (p->width <= 0 || p->height <= 0 || static_cast<int64_t>(p->width) * p->height >= (int64_t(1) << 28))
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_pathtrace: bad image size");
  if (p->spp < 1 || p->spp > 65536 || p->bounces < 1 || p->bounces > 64)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_pathtrace: spp must be 1..65536 and bounces 1..64");
  if (p->n_lights < 0 || p->n_lights > 4096 || (p->n_lights > 0 && !p->lights))
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_pathtrace: bad light array (0..4096 lights)");
  const int M = p->restir_candidates > 0 ? p->restir_candidates : 8;
  if (M > ubg::pt::kMaxCandidates || p->restir_spatial < 0 || p->restir_spatial > ubg::pt::kMaxSpatial)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_pathtrace: restir_candidates 1..12, restir_spatial 0..16");
  if ((p->n_prim_materials > 0 && !p->prim_materials) || (p->n_mesh_materials > 0 && !p->mesh_materials) ||
      p->n_prim_materials < 0 || p->n_mesh_materials < 0)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_pathtrace: bad material arrays");
  for (int k = 0; k < p->n_prim_materials + p->n_mesh_materials; ++k) {
    const ubg_pt_material& m = k < p->n_prim_materials ? p->prim_materials[k] : p->mesh_materials[k - p->n_prim_materials];
    if (m.type < 0 || m.type > 2 || !(m.roughness >= 0.0f) || (m.type == UBG_PT_DIELECTRIC && !(m.ior > 0.0f)))
      return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_pathtrace: bad material");
  }
  ubg::Unet* net = nullptr;
  if (p->denoise_net) {
    if (p->denoise_net->ctx != ctx) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_pathtrace: network on another context");
    net = &p->denoise_net->net;
  }
  const int ens = p->denoise_ensemble;
  if (ens != 0 && ens != 1 && ens != 2 && ens != 4 && ens != 8)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_pathtrace: denoise_ensemble must be 0, 1, 2, 4 or 8");
  // the network's passes: 0 = 8 from 32 spp on, where the 7 extra passes cost
  // under 1 % of the sample passes, else 1
  const int passes = ens != 0 ? ens : (p->spp >= 32 ? 8 : 1);
  std::lock_guard<std::mutex> lk(ctx->mu);
  UBG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t st = ctx->stream;
  const int W = p->width, H = p->height;
  const size_t n = static_cast<size_t>(W) * H;
  int Wp = W, Hp = H;
  if (net) {
    const int m = 1 << (net->levels() - 1);
    Wp = (W + m - 1) / m * m;
    Hp = (H + m - 1) / m * m;
  }
  const size_t np = static_cast<size_t>(Wp) * Hp;
  static_assert(ubg::pt::kPtFeatures == ubg::kUnetFeatures, "the network reads the path tracer's features");
  const bool wantFeat = net || fr->features;
  const int nGroups = std::max(1, (p->n_lights + ubg::pt::kLightGroup - 1) / ubg::pt::kLightGroup);
  const size_t tabBytes = sizeof(ubg::pt::Light) * std::max(1, p->n_lights) +
                          sizeof(ubg::pt::Mat) * (std::max(1, p->n_prim_materials) + std::max(1, p->n_mesh_materials)) +
                          sizeof(float4) * nGroups;
  int r = ensure(ctx->ptAcc, n * 16);
  if (r == UBG_OK) r = ensure(ctx->ptFirst, n * 32);
  if (r == UBG_OK) r = ensure(ctx->ptG, n * sizeof(ubg::pt::GPix));
  if (r == UBG_OK) r = ensure(ctx->ptRes, n * sizeof(ubg::pt::Reservoir));
  if (r == UBG_OK) r = ensure(ctx->ptTab, tabBytes);
  if (r == UBG_OK) r = ensure(ctx->ptWords, 256);
  if (r == UBG_OK && wantFeat) r = ensure(ctx->ptFeat, np * ubg::pt::kPtFeatures * 4);
  if (r == UBG_OK && net) r = ensure(ctx->ptPred, np * 3 * 4);
  if (r == UBG_OK && net && passes > 1) r = ensure(ctx->ptFeatT, np * ubg::pt::kPtFeatures * 4);
  if (r == UBG_OK && net && passes > 1) r = ensure(ctx->ptPredT, np * 3 * 4);
  if (r == UBG_OK && fr->memory == 0) r = ensure(ctx->ptColor, n * 16);
  if (r != UBG_OK) return r;
  // lights and materials
  std::vector<ubg::pt::Light> lights(static_cast<size_t>(std::max(1, p->n_lights)));
  for (int i = 0; i < p->n_lights; ++i) {
    const ubg_sphere_light& l = p->lights[i];
    if (!(l.radius > 0.0f)) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_pathtrace: light radius must be > 0");
    lights[i] = ubg::pt::Light{{l.center[0], l.center[1], l.center[2]}, l.radius,
                               {l.radiance[0], l.radiance[1], l.radiance[2]}, 0.0f};
  }
  auto mats = [](const ubg_pt_material* src, int cnt) {
    std::vector<ubg::pt::Mat> out(static_cast<size_t>(std::max(1, cnt)), ubg::pt::Mat{ubg::pt::kMatDiffuse, 0.0f, 1.5f, 0});
    for (int i = 0; i < cnt; ++i) out[i] = ubg::pt::Mat{src[i].type, src[i].roughness * src[i].roughness, src[i].ior, 0};
    return out;
  };
  const std::vector<ubg::pt::Mat> pm = mats(p->prim_materials, p->n_prim_materials);
  const std::vector<ubg::pt::Mat> mm = mats(p->mesh_materials, p->n_mesh_materials);
  // bounding spheres of the lights in groups of kLightGroup (index order),
  // in double, widened by 1e-4 relative so that rounding in the device's
  // group test never culls a ray that the per-light test would accept
  std::vector<float4> groups(static_cast<size_t>(nGroups), make_float4(0.0f, 0.0f, 0.0f, 0.0f));
  for (int g = 0; g * ubg::pt::kLightGroup < p->n_lights; ++g) {
    const int l0 = g * ubg::pt::kLightGroup, l1 = std::min(p->n_lights, l0 + ubg::pt::kLightGroup);
    double c[3] = {0.0, 0.0, 0.0};
    for (int i = l0; i < l1; ++i)
      for (int k = 0; k < 3; ++k) c[k] += lights[i].c[k];
    for (int k = 0; k < 3; ++k) c[k] /= (l1 - l0);
    double R = 0.0;
    for (int i = l0; i < l1; ++i) {
      const double dx = lights[i].c[0] - c[0], dy = lights[i].c[1] - c[1], dz = lights[i].c[2] - c[2];
      R = std::max(R, std::sqrt(dx * dx + dy * dy + dz * dz) + lights[i].r);
    }
    groups[g] = make_float4(static_cast<float>(c[0]), static_cast<float>(c[1]), static_cast<float>(c[2]),
                            static_cast<float>(R * (1.0 + 1e-4) + 1e-6 * std::max({std::fabs(c[0]), std::fabs(c[1]),
                                                                                  std::fabs(c[2]), 1.0})));
  }
  char* tab = static_cast<char*>(ctx->ptTab.p);
  const size_t lb = sizeof(ubg::pt::Light) * lights.size(), pb = sizeof(ubg::pt::Mat) * pm.size();
  UBG_HIP(hipMemcpyAsync(tab, lights.data(), lb, hipMemcpyHostToDevice, st), "hipMemcpy(lights)");
  UBG_HIP(hipMemcpyAsync(tab + lb, pm.data(), pb, hipMemcpyHostToDevice, st), "hipMemcpy(materials)");
  UBG_HIP(hipMemcpyAsync(tab + lb + pb, mm.data(), sizeof(ubg::pt::Mat) * mm.size(), hipMemcpyHostToDevice, st),
          "hipMemcpy(materials)");
  const size_t gOff = lb + pb + sizeof(ubg::pt::Mat) * mm.size();  // 16-byte aligned (Light 32 B, Mat 16 B)
  UBG_HIP(hipMemcpyAsync(tab + gOff, groups.data(), sizeof(float4) * groups.size(), hipMemcpyHostToDevice, st),
          "hipMemcpy(light groups)");
  // camera basis (embree_renderer.cpp:375-386, on the image grid)
  ubg::pt::Params q;
  std::memset(&q, 0, sizeof(q));
  {
    const ubg_camera& cam = p->camera;
    const float aspect = static_cast<float>(W) / static_cast<float>(H);
    const V3h dir = normalize(V3h{cam.direction[0], cam.direction[1], cam.direction[2]});
    const V3h right = normalize(cross(dir, V3h{cam.up[0], cam.up[1], cam.up[2]}));
    const V3h up = normalize(cross(right, dir));
    q.W = W;
    q.H = H;
    q.ortho = cam.orthographic ? 1 : 0;
    std::memcpy(q.pos, cam.position, sizeof(q.pos));
    q.dir[0] = dir.x, q.dir[1] = dir.y, q.dir[2] = dir.z;
    q.right[0] = right.x, q.right[1] = right.y, q.right[2] = right.z;
    q.up[0] = up.x, q.up[1] = up.y, q.up[2] = up.z;
    q.halfH = cam.height * 0.5f;
    q.halfW = q.halfH * aspect;
    q.persHalfH = std::tan(cam.fovy * 0.01745329252f * 0.5f);
    q.persHalfW = q.persHalfH * aspect;
  }
  q.bounces = p->bounces;
  q.seed = p->seed;
  q.nLights = p->n_lights;
  std::memcpy(q.env, p->env_radiance, sizeof(q.env));
  q.restirM = M;
  q.restirK = p->restir_spatial;
  q.restirRadius = p->restir_radius > 0.0f ? p->restir_radius : 16.0f;
  q.neeAll = p->nee_all_lights ? 1 : 0;
  q.lights = reinterpret_cast<const ubg::pt::Light*>(tab);
  q.lightGroups = reinterpret_cast<const float4*>(tab + gOff);
  q.primMats = reinterpret_cast<const ubg::pt::Mat*>(tab + lb);
  q.meshMats = reinterpret_cast<const ubg::pt::Mat*>(tab + lb + pb);
  q.nPrimMats = static_cast<int32_t>(pm.size());
  q.nMeshMats = static_cast<int32_t>(mm.size());
  q.rays = reinterpret_cast<unsigned long long*>(ctx->ptWords.p) + 1;
  q.depthMax = static_cast<unsigned int*>(ctx->ptWords.p);
  auto* maxBits = static_cast<unsigned int*>(ctx->ptWords.p);
  auto* acc = static_cast<float4*>(ctx->ptAcc.p);
  auto* first = static_cast<float4*>(ctx->ptFirst.p);
  ubg::ptZero(acc, static_cast<int>(n), maxBits, q.rays, st);
  UBG_HIP(hipEventRecord(ctx->ev[0], st), "hipEventRecord");
  for (int smp = 0; smp < p->spp; ++smp) {
    q.sample = smp;
    ubg::ptSample(q, scene->view, static_cast<ubg::pt::GPix*>(ctx->ptG.p), static_cast<ubg::pt::Reservoir*>(ctx->ptRes.p),
                  acc, wantFeat ? first : nullptr, st);
  }
  UBG_HIP(hipGetLastError(), "kernel launch");
  UBG_HIP(hipEventRecord(ctx->ev[1], st), "hipEventRecord");
  float4* color = fr->memory ? reinterpret_cast<float4*>(fr->color) : static_cast<float4*>(ctx->ptColor.p);
  float* feat = wantFeat ? static_cast<float*>(ctx->ptFeat.p) : nullptr;
  ubg::ptResolve(W, H, Wp, Hp, p->spp, acc, first, maxBits, color, feat, st);
  if (net) {
    float* pred = static_cast<float*>(ctx->ptPred.p);
    if (passes == 1) {
      UBG_HIP(net->forward(1, Hp, Wp, feat, pred, st), "unet forward");
    } else {
      // self-ensemble: the network on each turn of the features, the
      // predictions turned back and averaged
      float* featT = static_cast<float*>(ctx->ptFeatT.p);
      float* predT = static_cast<float*>(ctx->ptPredT.p);
      for (int k = 0; k < passes; ++k) {
        const bool tr = (k & 4) != 0;
        if (k > 0) ubg::ptFeatTurn(k, Wp, Hp, feat, featT, st);
        UBG_HIP(net->forward(1, tr ? Wp : Hp, tr ? Hp : Wp, k > 0 ? featT : feat, predT, st), "unet forward");
        ubg::ptPredUnturn(k, W, H, Wp, Hp, 1.0f / static_cast<float>(passes), predT, pred, st);
      }
    }
    ubg::ptApplyNet(W, H, Wp, pred, color, st);
  }
  UBG_HIP(hipGetLastError(), "kernel launch");
  UBG_HIP(hipEventRecord(ctx->ev[2], st), "hipEventRecord");
  if (fr->memory == 0) r = copyOut(fr->color, color, n * 16, 0, st);
  if (r == UBG_OK && fr->features) {
    // the unpadded rows of the (possibly padded) feature image
    const size_t px = sizeof(float) * ubg::pt::kPtFeatures;
    UBG_HIP(hipMemcpy2DAsync(fr->features, static_cast<size_t>(W) * px, feat, static_cast<size_t>(Wp) * px,
                             static_cast<size_t>(W) * px, static_cast<size_t>(H),
                             fr->memory ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st),
            "hipMemcpy(features)");
  }
  if (r != UBG_OK) return r;
  unsigned long long rays = 0;
  UBG_HIP(hipMemcpyAsync(&rays, q.rays, sizeof(rays), hipMemcpyDeviceToHost, st), "hipMemcpy(rays)");
  UBG_HIP(hipStreamSynchronize(st), "pathtrace sync");
  float ms0 = 0.0f, ms1 = 0.0f;
  UBG_HIP(hipEventElapsedTime(&ms0, ctx->ev[0], ctx->ev[1]), "hipEventElapsedTime");
  UBG_HIP(hipEventElapsedTime(&ms1, ctx->ev[1], ctx->ev[2]), "hipEventElapsedTime");
  fr->trace_seconds = ms0 * 1e-3;
  fr->denoise_seconds = ms1 * 1e-3;
  fr->rays = rays;
  return UBG_OK;
}

int ubg_render_noalloc(ubg_context* ctx, const ubg_scene* scene, const ubg_render_params* p, ubg_frame* fr) {
  if (!ctx || !scene || !p || !fr) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render_noalloc: null argument");
  if (fr->memory != 1)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render_noalloc: the outputs must be device buffers (memory 1)");
  if (p->flags & UBG_FLAG_KERNEL_TIMES)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render_noalloc: UBG_FLAG_KERNEL_TIMES creates events; not allowed");
  if (p->gi && p->pt1_denoise == 2)
    return fail(UBG_ERR_UNSUPPORTED, "ubg_render_noalloc: the learned denoiser allocates its activations");
  if (ctx->staticStatus != hipSuccess || !ctx->staticScratch)
    return hipFail(ctx->staticStatus != hipSuccess ? ctx->staticStatus : hipErrorInvalidSymbol,
                   "ubg_render_noalloc: fixed scratch address");
  // The basic path's scratch (ray queues and per-pixel state) must fit before
  // anything is launched; a feature's extra arrays are checked as carved.
  const int ss = std::max(1, p->supersample);
  const int64_t W = static_cast<int64_t>(std::max(0, p->width)) * ss, H = static_cast<int64_t>(std::max(0, p->height)) * ss;
  if (W > 0 && H > 0 && W * H < (int64_t(1) << 31)) {
    const int64_t qn = static_cast<int64_t>(ubg::kQueueParts) * ubg::queuePartCap(static_cast<int>(W), static_cast<int>(H));
    const int64_t need = qn * 68 + W * H * 52 + static_cast<int64_t>(p->width) * p->height * 16;
    if (need > static_cast<int64_t>(UBG_NOALLOC_SCRATCH_BYTES))
      return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_render_noalloc: a " + std::to_string(W) + " x " + std::to_string(H) +
                                                " hi-res frame needs " + std::to_string(need >> 20) +
                                                " MiB of scratch, more than the fixed " +
                                                std::to_string(UBG_NOALLOC_SCRATCH_BYTES >> 20) + " MiB");
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  // The context's own (allocated) buffers step aside for the call and come
  // back afterwards: ensure() carves from the fixed array meanwhile.
  std::vector<DevBuf> saved;
  for (DevBuf* b : ctx->all()) {
    saved.push_back(*b);
    *b = DevBuf{};
  }
  StaticArena arena{static_cast<char*>(ctx->staticScratch), UBG_NOALLOC_SCRATCH_BYTES, 0};
  t_arena = &arena;
  const int st = renderFrame(ctx, scene, p, fr, true);
  t_arena = nullptr;
  const std::vector<DevBuf*> bufs = ctx->all();
  for (size_t i = 0; i < bufs.size(); ++i) *bufs[i] = saved[i];
  return st;
}

int ubg_intersect(ubg_context* ctx, const ubg_scene* scene, int64_t n_rays, const ubg_ray* rays,
                  ubg_hit* hits, int32_t memory) {
  return runQuery(ctx, scene, n_rays, rays, hits, sizeof(ubg_hit), memory, false);
}

int ubg_occluded(ubg_context* ctx, const ubg_scene* scene, int64_t n_rays, const ubg_ray* rays,
                 uint8_t* occluded, int32_t memory) {
  return runQuery(ctx, scene, n_rays, rays, occluded, 1, memory, true);
}

int ubg_occluded_segments(ubg_context* ctx, const ubg_scene* scene, int64_t n, const ubg_segment* segs,
                          const int32_t* excl, int64_t nExcl, float eps, float grazeCosEps, float coplanarEps,
                          uint8_t* occluded, int32_t memory) {
  if (!ctx || !scene || n < 0 || nExcl < 0 || (n > 0 && (!segs || !occluded)) || (nExcl > 0 && !excl))
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_occluded_segments: bad argument");
  if (scene->ctx != ctx)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_occluded_segments: scene belongs to another context");
  if (!memory)
    for (int64_t i = 0; i < n; ++i)
      if (segs[i].exclude_count < 0 || segs[i].exclude_first < 0 ||
          static_cast<int64_t>(segs[i].exclude_first) + segs[i].exclude_count > nExcl)
        return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_occluded_segments: exclusion range out of bounds");
  if (n == 0) return UBG_OK;
  std::lock_guard<std::mutex> lk(ctx->mu);
  UBG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t st = ctx->stream;
  const void* dSegs = segs;
  const void* dExcl = excl;
  void* dOut = occluded;
  if (!memory) {
    const size_t segBytes = static_cast<size_t>(n) * sizeof(ubg_segment);
    const size_t exclBytes = static_cast<size_t>(nExcl) * sizeof(int32_t);
    int r = ensure(ctx->query, segBytes + exclBytes);
    if (r == UBG_OK) r = ensure(ctx->queryOut, static_cast<size_t>(n));
    if (r != UBG_OK) return r;
    UBG_HIP(hipMemcpyAsync(ctx->query.p, segs, segBytes, hipMemcpyHostToDevice, st), "hipMemcpy(segments)");
    if (exclBytes)
      UBG_HIP(hipMemcpyAsync(static_cast<char*>(ctx->query.p) + segBytes, excl, exclBytes, hipMemcpyHostToDevice,
                             st),
              "hipMemcpy(exclusions)");
    dSegs = ctx->query.p;
    dExcl = exclBytes ? static_cast<const void*>(static_cast<char*>(ctx->query.p) + segBytes) : nullptr;
    dOut = ctx->queryOut.p;
  }
  ubg::launchSegments(scene->view, dSegs, n, static_cast<const int32_t*>(dExcl), nExcl, eps, grazeCosEps,
                      coplanarEps, static_cast<uint8_t*>(dOut), st);
  UBG_HIP(hipGetLastError(), "kernel launch");
  if (!memory)
    UBG_HIP(hipMemcpyAsync(occluded, dOut, static_cast<size_t>(n), hipMemcpyDeviceToHost, st),
            "hipMemcpy(segment result)");
  UBG_HIP(hipStreamSynchronize(st), "segment query sync");
  return UBG_OK;
}

int ubg_object_space_edges(ubg_context* ctx, const ubg_scene_desc* d, const ubg_camera* cam,
                           const ubg_edge_options* o, ubg_edges* out) {
  if (!d || !cam || !o || !out) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_object_space_edges: bad argument");
  *out = ubg_edges{};
  if ((d->n_spheres > 0 && !d->sphere_center_radius) ||
      (d->n_cylinders > 0 && (!d->cyl_p0 || !d->cyl_p1 || !d->cyl_radius)) ||
      (d->n_vertices > 0 && !d->positions) || (d->n_triangles > 0 && d->n_vertices < 3) || d->n_spheres < 0 ||
      d->n_cylinders < 0 || d->n_triangles < 0)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_object_space_edges: bad scene arrays");
  if (d->index)
    for (int64_t c = 0; c < 3 * d->n_triangles; ++c)
      if (d->index[c] >= static_cast<uint64_t>(d->n_vertices))
        return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_object_space_edges: mesh index out of range");
  if (!d->index && 3 * d->n_triangles > d->n_vertices)
    return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_object_space_edges: more corners than vertices");
  const bool meshEdges = d->n_vertices >= 3 && d->n_triangles > 0 &&
                         (o->mesh_silhouette || o->mesh_crease || o->mesh_border);
  if (o->visibility_clip && meshEdges)
    return fail(UBG_ERR_UNSUPPORTED, "object-space edges: the mesh hidden-line clip (visibilityClip) is not implemented");
  std::vector<ubg::EdgeChord> chords;
  ubg::analyticChords(d, *cam, o->segments, o->raise, chords);
  ubg::EdgeList edges;
  if (!o->clip) {
    for (const ubg::EdgeChord& c : chords) edges.push(c.a, c.b, *o, c.group);
  } else if (!chords.empty()) {
    // clipAndEmit (object_space_edges.cpp:151-182): the chords' sample points
    // go to the device for the point-in-solid tests, the runs outside every
    // other solid come back as edges
    std::vector<int64_t> first(chords.size() + 1, 0);
    std::vector<int> K(chords.size());
    for (size_t i = 0; i < chords.size(); ++i) {
      K[i] = ubg::chordSamples(chords[i], o->width);
      first[i + 1] = first[i] + (K[i] > 0 ? K[i] + 1 : 0);
    }
    const int64_t nPts = first.back();
    std::vector<float4> pts(static_cast<size_t>(nPts));
    std::vector<int2> src(static_cast<size_t>(nPts));
    for (size_t i = 0; i < chords.size(); ++i) {
      const ubg::EdgeChord& c = chords[i];
      const float ringRs = c.srcSphere >= 0 ? d->sphere_center_radius[4 * c.srcSphere + 3] : -1.0f;
      for (int k = 0; k <= K[i] && K[i] > 0; ++k) {
        float p[3];
        ubg::chordPoint(c, K[i], k, p);
        pts[static_cast<size_t>(first[i] + k)] = make_float4(p[0], p[1], p[2], ringRs);
        src[static_cast<size_t>(first[i] + k)] = make_int2(c.srcSphere, c.srcCyl);
      }
    }
    ubg::SolidGrid g;
    ubg::buildSolidGrid(d, g);
    std::vector<uint8_t> inside(static_cast<size_t>(nPts), 0);
    if (nPts > 0) {
      if (!ctx) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_object_space_edges: the clip needs a device context");
      std::lock_guard<std::mutex> lk(ctx->mu);
      UBG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
      hipStream_t st = ctx->stream;
      const size_t ptBytes = static_cast<size_t>(nPts) * 16, srcBytes = static_cast<size_t>(nPts) * 8;
      const size_t sphBytes = g.spheres.size() * 4, cylBytes = g.cyls.size() * 4;
      const size_t startBytes = g.cellStart.size() * 4, itemBytes = g.cellItems.size() * 4;
      int r = ensure(ctx->edgePts, ptBytes + srcBytes);
      if (r == UBG_OK) r = ensure(ctx->edgeSolids, sphBytes + cylBytes + startBytes + itemBytes + 64);
      if (r == UBG_OK) r = ensure(ctx->edgeOut, static_cast<size_t>(nPts));
      if (r != UBG_OK) return r;
      char* pb = static_cast<char*>(ctx->edgePts.p);
      char* sb = static_cast<char*>(ctx->edgeSolids.p);
      // 16-byte aligned sub-buffers: spheres, cylinders, cell starts, items
      const size_t oCyl = (sphBytes + 15) & ~size_t(15), oStart = oCyl + ((cylBytes + 15) & ~size_t(15));
      const size_t oItem = oStart + ((startBytes + 15) & ~size_t(15));
      UBG_HIP(hipMemcpyAsync(pb, pts.data(), ptBytes, hipMemcpyHostToDevice, st), "hipMemcpy(edge points)");
      UBG_HIP(hipMemcpyAsync(pb + ptBytes, src.data(), srcBytes, hipMemcpyHostToDevice, st), "hipMemcpy(edge sources)");
      if (sphBytes) UBG_HIP(hipMemcpyAsync(sb, g.spheres.data(), sphBytes, hipMemcpyHostToDevice, st), "hipMemcpy(spheres)");
      if (cylBytes)
        UBG_HIP(hipMemcpyAsync(sb + oCyl, g.cyls.data(), cylBytes, hipMemcpyHostToDevice, st), "hipMemcpy(cylinders)");
      UBG_HIP(hipMemcpyAsync(sb + oStart, g.cellStart.data(), startBytes, hipMemcpyHostToDevice, st),
              "hipMemcpy(grid)");
      if (itemBytes)
        UBG_HIP(hipMemcpyAsync(sb + oItem, g.cellItems.data(), itemBytes, hipMemcpyHostToDevice, st),
                "hipMemcpy(grid items)");
      ubg::launchSolidsContain(nPts, reinterpret_cast<const float4*>(pb), reinterpret_cast<const int2*>(pb + ptBytes),
                               reinterpret_cast<const float4*>(sb), reinterpret_cast<const float4*>(sb + oCyl),
                               reinterpret_cast<const uint32_t*>(sb + oStart),
                               reinterpret_cast<const uint32_t*>(sb + oItem), g.lo, g.inv, g.dim,
                               static_cast<uint8_t*>(ctx->edgeOut.p), st);
      UBG_HIP(hipGetLastError(), "kernel launch");
      UBG_HIP(hipMemcpyAsync(inside.data(), ctx->edgeOut.p, static_cast<size_t>(nPts), hipMemcpyDeviceToHost, st),
              "hipMemcpy(edge clip result)");
      UBG_HIP(hipStreamSynchronize(st), "edge clip sync");
    }
    for (size_t i = 0; i < chords.size(); ++i) {
      const int k1 = K[i];
      if (k1 <= 0) continue;
      const uint8_t* in = &inside[static_cast<size_t>(first[i])];
      auto emitRun = [&](int a, int b) {
        if (b <= a) return;
        float pa[3], pbp[3];
        ubg::chordPoint(chords[i], k1, a, pa);
        ubg::chordPoint(chords[i], k1, b, pbp);
        edges.push(pa, pbp, *o, chords[i].group);
      };
      int runStart = -1;
      for (int k = 0; k <= k1; ++k) {
        if (!in[k] && runStart < 0) runStart = k;
        if (in[k] && runStart >= 0) {
          emitRun(runStart, k - 1);
          runStart = -1;
        }
      }
      if (runStart >= 0) emitRun(runStart, k1);
    }
  }
  if (meshEdges) ubg::meshFeatureEdges(d, *cam, *o, edges);
  const int64_t n = edges.size();
  out->n = n;
  out->radius = o->width > 0.0f ? o->width : 0.0f;
  if (n > 0) {
    out->p0 = static_cast<float*>(std::malloc(static_cast<size_t>(n) * 12));
    out->p1 = static_cast<float*>(std::malloc(static_cast<size_t>(n) * 12));
    out->color = static_cast<float*>(std::malloc(static_cast<size_t>(n) * 16));
    out->opacity1 = static_cast<float*>(std::malloc(static_cast<size_t>(n) * 4));
    out->group = static_cast<uint16_t*>(std::malloc(static_cast<size_t>(n) * 2));
    if (!out->p0 || !out->p1 || !out->color || !out->opacity1 || !out->group) {
      ubg_edges_free(out);
      return fail(UBG_ERR_OUT_OF_MEMORY, "ubg_object_space_edges: host allocation");
    }
    std::memcpy(out->p0, edges.p0.data(), static_cast<size_t>(n) * 12);
    std::memcpy(out->p1, edges.p1.data(), static_cast<size_t>(n) * 12);
    std::memcpy(out->color, edges.color.data(), static_cast<size_t>(n) * 16);
    std::memcpy(out->opacity1, edges.opacity1.data(), static_cast<size_t>(n) * 4);
    std::memcpy(out->group, edges.group.data(), static_cast<size_t>(n) * 2);
  }
  return UBG_OK;
}

int ubg_unet_create(ubg_context* ctx, int32_t levels, const int32_t* channels, uint64_t seed, ubg_unet** out) {
  if (!ctx || !channels || !out) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_unet_create: bad argument");
  *out = nullptr;
  std::lock_guard<std::mutex> lk(ctx->mu);
  UBG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  auto* u = new ubg_unet;
  u->ctx = ctx;
  int ch[ubg::kUnetMaxLevels] = {0};
  for (int l = 0; l < levels && l < ubg::kUnetMaxLevels; ++l) ch[l] = channels[l];
  std::string err;
  if (u->net.init(levels, ch, seed, err) != 0) {
    delete u;
    return fail(UBG_ERR_INVALID_ARGUMENT, err);
  }
  *out = u;
  return UBG_OK;
}

void ubg_unet_destroy(ubg_unet* u) {
  if (!u) return;
  std::lock_guard<std::mutex> lk(u->ctx->mu);
  (void)hipSetDevice(u->ctx->device);
  delete u;
}

int64_t ubg_unet_param_count(const ubg_unet* u) { return u ? u->net.paramCount() : 0; }

int ubg_unet_set_loss(ubg_unet* u, int32_t mode) {
  if (!u || mode < 0 || mode > 1) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_unet_set_loss: mode must be 0 or 1");
  std::lock_guard<std::mutex> lk(u->ctx->mu);
  u->net.setLossL2(mode == 1);
  return UBG_OK;
}

int ubgx_unet_set_min_workgroups(ubg_unet* u, int64_t n) {
  if (!u || n < 0) return fail(UBG_ERR_INVALID_ARGUMENT, "ubgx_unet_set_min_workgroups: bad argument");
  std::lock_guard<std::mutex> lk(u->ctx->mu);
  u->net.setMinWorkgroups(n);
  return UBG_OK;
}

int ubg_unet_get_params(ubg_unet* u, float* params) {
  if (!u || !params) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_unet_get_params: bad argument");
  std::lock_guard<std::mutex> lk(u->ctx->mu);
  UBG_HIP(hipSetDevice(u->ctx->device), "hipSetDevice");
  UBG_HIP(u->net.getParams(params, u->ctx->stream), "unet parameters");
  return UBG_OK;
}

int ubg_unet_set_params(ubg_unet* u, const float* params) {
  if (!u || !params) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_unet_set_params: bad argument");
  std::lock_guard<std::mutex> lk(u->ctx->mu);
  UBG_HIP(hipSetDevice(u->ctx->device), "hipSetDevice");
  UBG_HIP(u->net.setParams(params, u->ctx->stream), "unet parameters");
  return UBG_OK;
}

namespace {
int unetShape(const ubg_unet* u, int32_t n, int32_t h, int32_t w) {
  const int m = 1 << (u->net.levels() - 1);
  if (n <= 0 || h <= 0 || w <= 0 || h % m != 0 || w % m != 0)
    return fail(UBG_ERR_INVALID_ARGUMENT, "unet: n, h, w must be positive and h, w multiples of 2^(levels - 1)");
  return UBG_OK;
}
}  // namespace

int ubg_unet_infer(ubg_unet* u, int32_t n, int32_t h, int32_t w, const float* feat, float* out, int32_t memory) {
  if (!u || !feat || !out) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_unet_infer: bad argument");
  int r = unetShape(u, n, h, w);
  if (r != UBG_OK) return r;
  ubg_context* ctx = u->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  UBG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t st = ctx->stream;
  const size_t P = static_cast<size_t>(n) * h * w;
  const float* dFeat = feat;
  float* dOut = out;
  if (!memory) {
    r = ensure(ctx->unetIn, P * ubg::kUnetFeatures * 4);
    if (r == UBG_OK) r = ensure(ctx->unetOut, P * 3 * 4);
    if (r != UBG_OK) return r;
    UBG_HIP(hipMemcpyAsync(ctx->unetIn.p, feat, P * ubg::kUnetFeatures * 4, hipMemcpyHostToDevice, st),
            "hipMemcpy(unet features)");
    dFeat = static_cast<const float*>(ctx->unetIn.p);
    dOut = static_cast<float*>(ctx->unetOut.p);
  }
  UBG_HIP(u->net.forward(n, h, w, dFeat, dOut, st), "unet forward");
  if (!memory) UBG_HIP(hipMemcpyAsync(out, dOut, P * 3 * 4, hipMemcpyDeviceToHost, st), "hipMemcpy(unet output)");
  UBG_HIP(hipStreamSynchronize(st), "unet sync");
  return UBG_OK;
}

int ubg_unet_train_step(ubg_unet* u, int32_t n, int32_t h, int32_t w, const float* feat, const float* target,
                        float lr, float* loss, float* grads, int32_t memory) {
  if (!u || !feat || !target) return fail(UBG_ERR_INVALID_ARGUMENT, "ubg_unet_train_step: bad argument");
  int r = unetShape(u, n, h, w);
  if (r != UBG_OK) return r;
  ubg_context* ctx = u->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  UBG_HIP(hipSetDevice(ctx->device), "hipSetDevice");
  hipStream_t st = ctx->stream;
  const size_t P = static_cast<size_t>(n) * h * w;
  const float* dFeat = feat;
  const float* dTgt = target;
  if (!memory) {
    r = ensure(ctx->unetIn, P * ubg::kUnetFeatures * 4);
    if (r == UBG_OK) r = ensure(ctx->unetTgt, P * 3 * 4);
    if (r != UBG_OK) return r;
    UBG_HIP(hipMemcpyAsync(ctx->unetIn.p, feat, P * ubg::kUnetFeatures * 4, hipMemcpyHostToDevice, st),
            "hipMemcpy(unet features)");
    UBG_HIP(hipMemcpyAsync(ctx->unetTgt.p, target, P * 3 * 4, hipMemcpyHostToDevice, st), "hipMemcpy(unet targets)");
    dFeat = static_cast<const float*>(ctx->unetIn.p);
    dTgt = static_cast<const float*>(ctx->unetTgt.p);
  }
  float l = 0.0f;
  UBG_HIP(u->net.gradients(n, h, w, dFeat, dTgt, &l, st), "unet gradients");
  if (loss) *loss = l;
  if (grads) UBG_HIP(u->net.copyGrads(grads, st), "unet gradient copy");
  if (lr > 0.0f) UBG_HIP(u->net.adam(lr, st), "unet adam");
  UBG_HIP(hipStreamSynchronize(st), "unet sync");
  return UBG_OK;
}

void ubg_edges_free(ubg_edges* e) {
  if (!e) return;
  std::free(e->p0);
  std::free(e->p1);
  std::free(e->color);
  std::free(e->opacity1);
  std::free(e->group);
  *e = ubg_edges{};
}

}  // extern "C"
