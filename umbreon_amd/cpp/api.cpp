// umbreon::render / renderAsync on the MI355X backend.
//
// Host orchestration of the reference's public flow (umbreon.cpp:141-341 and
// pipeline.cpp:20-256): option checks, the group-alpha multipass blend in the
// sRGB domain, and per pass the scene flattening into the C-ABI's
// structure-of-arrays description (umbreon_gpu.h), one ubg_scene_create_ex
// (upload + BVH) and one ubg_render (hi-res trace + shade + fog, downsample,
// gamma) on a process-wide device context. Failures surface as
// std::runtime_error with the C-ABI's message.

#include <algorithm>
#include <array>
#include <atomic>
#include <cstring>
#include <exception>
#include <future>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "postprocess/image_ops.hpp"
#include "umbreon.hpp"
#include "umbreon_gpu.h"

namespace umbreon {

namespace detail {
// Grants the backend the address of RenderProgress's cancel flag.
class ProgressAccess {
 public:
  static const volatile int32_t* cancelFlag(RenderProgress& p) {
    return reinterpret_cast<const volatile int32_t*>(&p.cancel_);
  }
};
}  // namespace detail

namespace {

static_assert(sizeof(std::atomic<int32_t>) == sizeof(int32_t), "lock-free int32 atomic expected");

[[noreturn]] void throwStatus(int st, const char* what) {
  throw std::runtime_error(std::string("umbreon (MI355X): ") + what + ": " + ubg_last_error() + " [" +
                           ubg_status_string(st) + "]");
}

// One gfx950 context for the process (the SharedDevice of the reference);
// the C-ABI serializes concurrent calls on it.
ubg_context* device() {
  static std::once_flag once;
  static ubg_context* ctx = nullptr;
  static int status = UBG_OK;
  static std::string msg;
  std::call_once(once, [] {
    status = ubg_context_create(0, &ctx);
    if (status != UBG_OK) msg = ubg_last_error();
  });
  if (status != UBG_OK)
    throw std::runtime_error("umbreon (MI355X): device init failed: " + msg + " [" + ubg_status_string(status) +
                             "]");
  return ctx;
}

ubg_material toC(const Material& m) {
  ubg_material c{};
  c.ambient = m.ambient;
  c.diffuse = m.diffuse;
  c.specular = m.specular;
  c.roughness = m.roughness;
  c.brilliance = m.brilliance;
  c.phong = m.phong;
  c.phong_size = m.phongSize;
  c.metallic = m.metallic ? 1 : 0;
  c.reflection = m.reflection;
  c.emission = m.emission;
  c.model = static_cast<int32_t>(m.model);
  c.pbr_metallic = m.pbr.metallic;
  c.pbr_roughness = m.pbr.roughness;
  c.pbr_specular = m.pbr.specular;
  c.pbr_anisotropy = m.pbr.anisotropy;
  c.pbr_anisotropy_rotation = m.pbr.anisotropyRotation;
  return c;
}

// The section groups one render pass draws: a group is drawn unless its
// entry in `hidden` is set. An empty filter draws everything.
struct GroupFilter {
  std::vector<uint8_t> hidden;
  bool draws(uint16_t g) const { return g >= hidden.size() || hidden[g] == 0; }
  bool drawsAll() const { return std::find(hidden.begin(), hidden.end(), 1) == hidden.end(); }
};

// Scene -> structure-of-arrays C description (owns the arrays), keeping only
// the primitives of the groups the filter draws. Triangles of hidden groups
// leave the index list (a filtered soup becomes an explicit index over the
// same vertices), so the vertex streams are shared by every pass.
struct FlatScene {
  std::vector<float> pos, nrm, col;
  std::vector<uint32_t> idx;
  std::vector<ubg_material> meshMats;
  std::vector<uint8_t> triMat;
  std::vector<uint16_t> triGroup;
  std::vector<float> sph, sphCol;
  std::vector<int32_t> sphMat;
  std::vector<uint16_t> sphGroup;
  std::vector<uint8_t> sphEdge;
  std::vector<float> c0, c1, crad, ccol, cop1;
  std::vector<int32_t> cmat;
  std::vector<uint16_t> cgroup;
  std::vector<uint8_t> copen, cedge;
  std::vector<ubg_material> primMats;
  std::vector<float> inst;
  std::map<std::array<uint32_t, 16>, int32_t> matIndex;
  int64_t nTris = 0;
  bool explicitIndex = false;
  ubg_scene_desc desc{};

  int32_t primMaterial(const Material& m) {
    const ubg_material c = toC(m);
    std::array<uint32_t, 16> key;
    static_assert(sizeof(ubg_material) == sizeof(key), "material key size");
    std::memcpy(key.data(), &c, sizeof(c));
    auto it = matIndex.find(key);
    if (it != matIndex.end()) return it->second;
    const int32_t i = static_cast<int32_t>(primMats.size());
    primMats.push_back(c);
    matIndex.emplace(key, i);
    return i;
  }

  FlatScene(const Scene& s, const GroupFilter& filter) {
    const Mesh& m = s.mesh;
    const std::size_t nv = m.positions.size();
    pos.reserve(3 * nv);
    for (std::size_t v = 0; v < nv; ++v) {
      const Vec3& p = m.positions[v];
      pos.insert(pos.end(), {p.x, p.y, p.z});
      const Vec3 n = v < m.normals.size() ? m.normals[v] : Vec3{};
      nrm.insert(nrm.end(), {n.x, n.y, n.z});
      const Vec4 c = v < m.colors.size() ? m.colors[v] : Vec4{};
      col.insert(col.end(), {c.x, c.y, c.z, c.w});
    }
    if (m.triMaterialId.empty()) {
      meshMats.push_back(toC(m.material));
    } else {
      for (const Material& mm : m.materials) meshMats.push_back(toC(mm));
    }
    const std::size_t nt = m.triangleCount();
    if (filter.drawsAll()) {
      idx = m.index;
      triMat = m.triMaterialId;
      triGroup = m.triGroupId;
      nTris = static_cast<int64_t>(nt);
      explicitIndex = !idx.empty();
    } else {
      for (std::size_t t = 0; t < nt; ++t) {
        if (!filter.draws(m.groupForTri(t))) continue;
        idx.insert(idx.end(), {m.cornerVertex(3 * t), m.cornerVertex(3 * t + 1), m.cornerVertex(3 * t + 2)});
        if (!m.triMaterialId.empty()) triMat.push_back(m.triMaterialId[t]);
        if (!m.triGroupId.empty()) triGroup.push_back(m.triGroupId[t]);
      }
      nTris = static_cast<int64_t>(idx.size() / 3);
      explicitIndex = true;
    }
    for (const Sphere& sp : s.spheres) {
      if (!filter.draws(sp.group)) continue;
      sph.insert(sph.end(), {sp.center.x, sp.center.y, sp.center.z, sp.radius});
      sphCol.insert(sphCol.end(), {sp.color.x, sp.color.y, sp.color.z, sp.color.w});
      sphMat.push_back(primMaterial(sp.material));
      sphGroup.push_back(sp.group);
      sphEdge.push_back(sp.fromEdgeMacro ? 1 : 0);
    }
    for (const Cylinder& cy : s.cylinders) {
      if (!filter.draws(cy.group)) continue;
      addCylinder(cy.p0, cy.p1, cy.radius, cy.color, cy.opacity1, primMaterial(cy.material), cy.group, cy.open,
                  cy.fromEdgeMacro);
    }
    for (const Vec3& o : s.instanceOffsets) inst.insert(inst.end(), {o.x, o.y, o.z});
    publish();
  }

  void addCylinder(const Vec3& p0, const Vec3& p1, float r, const Vec4& color, float op1, int32_t mat,
                   uint16_t group, bool open, bool fromEdge) {
    c0.insert(c0.end(), {p0.x, p0.y, p0.z});
    c1.insert(c1.end(), {p1.x, p1.y, p1.z});
    crad.push_back(r);
    ccol.insert(ccol.end(), {color.x, color.y, color.z, color.w});
    cop1.push_back(op1);
    cmat.push_back(mat);
    cgroup.push_back(group);
    copen.push_back(open ? 1 : 0);
    cedge.push_back(fromEdge ? 1 : 0);
  }

  // Generated object-space edges join the cylinders as open curves with the
  // flat outline material, not fromEdgeMacro (object_space_edges.cpp:240-262).
  void appendEdges(const ubg_edges& e) {
    const int32_t mat = primMaterial(Material::flatOutline());
    for (int64_t i = 0; i < e.n; ++i)
      addCylinder(Vec3{e.p0[3 * i], e.p0[3 * i + 1], e.p0[3 * i + 2]},
                  Vec3{e.p1[3 * i], e.p1[3 * i + 1], e.p1[3 * i + 2]}, e.radius,
                  Vec4{e.color[4 * i], e.color[4 * i + 1], e.color[4 * i + 2], e.color[4 * i + 3]}, e.opacity1[i],
                  mat, e.group[i], true, false);
    publish();
  }

  std::size_t effectiveTriangles() const {
    return static_cast<std::size_t>(nTris) * (inst.empty() ? 1 : inst.size() / 3);
  }

  // Point the C description at the arrays (again after they grew).
  void publish() {
    desc = ubg_scene_desc{};
    desc.n_vertices = static_cast<int64_t>(pos.size() / 3);
    desc.positions = pos.data();
    desc.normals = nrm.data();
    desc.colors = col.data();
    desc.n_triangles = nTris;
    desc.index = explicitIndex && !idx.empty() ? idx.data() : nullptr;
    desc.mesh_materials = meshMats.data();
    desc.n_mesh_materials = static_cast<int32_t>(meshMats.size());
    desc.tri_material = triMat.empty() ? nullptr : triMat.data();
    desc.tri_group = triGroup.empty() ? nullptr : triGroup.data();
    desc.n_spheres = static_cast<int64_t>(sphGroup.size());
    desc.sphere_center_radius = sph.data();
    desc.sphere_color = sphCol.data();
    desc.sphere_material = sphMat.data();
    desc.sphere_group = sphGroup.data();
    desc.sphere_from_edge = sphEdge.data();
    desc.n_cylinders = static_cast<int64_t>(crad.size());
    desc.cyl_p0 = c0.data();
    desc.cyl_p1 = c1.data();
    desc.cyl_radius = crad.data();
    desc.cyl_color = ccol.data();
    desc.cyl_opacity1 = cop1.data();
    desc.cyl_material = cmat.data();
    desc.cyl_group = cgroup.data();
    desc.cyl_open = copen.data();
    desc.cyl_from_edge = cedge.data();
    desc.prim_materials = primMats.data();
    desc.n_prim_materials = static_cast<int32_t>(primMats.size());
    desc.n_instances = static_cast<int64_t>(inst.size() / 3);
    desc.instance_offsets = inst.empty() ? nullptr : inst.data();
  }
};

// Options the per-ray GPU path does not implement: reject rather than return
// a different image.
void checkSupported(const Scene& s, const RenderOptions& o) {
  if (o.strokeEdges.enable && o.objectSpaceEdges.enable)
    throw std::runtime_error(
        "umbreon::render: strokeEdges and objectSpaceEdges are mutually exclusive; enable at most one");
  auto no = [](const char* what) {
    throw std::runtime_error(std::string("umbreon (MI355X): ") + what + " is not supported by the GPU backend");
  };
  if (o.gi && o.giIntegrator != 1 && o.giIntegrator != 2) no("the irradiance-cache GI integrator (giIntegrator 0)");
  if (o.aoWriteAov) no("AO AOVs (aoWriteAov)");
  if (o.denoiser != 0) no("the final-colour denoiser");
  if (o.envLights > 0) no("environment dome lights (envLights)");
  if (o.strokeEdges.enable) no("stroke edges");
  if (s.lights.size() > 16) no("more than 16 lights");
}

ubg_camera toCamera(const Camera& c) {
  ubg_camera k{};
  const float cp[3] = {c.position.x, c.position.y, c.position.z};
  const float cd[3] = {c.direction.x, c.direction.y, c.direction.z};
  const float cu[3] = {c.up.x, c.up.y, c.up.z};
  std::memcpy(k.position, cp, sizeof(cp));
  std::memcpy(k.direction, cd, sizeof(cd));
  std::memcpy(k.up, cu, sizeof(cu));
  k.fovy = c.fovy;
  k.orthographic = c.orthographic ? 1 : 0;
  k.height = c.height;
  return k;
}

// generateObjectSpaceEdges (object_space_edges.cpp:172-264) over the
// primitives a pass draws, appended to that pass's flat scene; the caller's
// Scene is never touched.
void addObjectSpaceEdges(FlatScene& flat, const Camera& camera, const ObjectSpaceEdgeOptions& o) {
  const ubg_camera cam = toCamera(camera);
  ubg_edge_options eo{};
  eo.width = o.width;
  eo.raise = o.raise;
  eo.segments = o.segments;
  for (int k = 0; k < 3; ++k) eo.color[k] = o.color[k];
  eo.clip = o.clip ? 1 : 0;
  eo.mesh_silhouette = o.meshSilhouette ? 1 : 0;
  eo.mesh_crease = o.meshCrease ? 1 : 0;
  eo.mesh_border = o.meshBorder ? 1 : 0;
  eo.visibility_clip = o.visibilityClip ? 1 : 0;
  eo.mesh_hard_edge_deg = o.meshHardEdgeDeg;
  eo.crease_angle_deg = o.creaseAngleDeg;
  eo.mesh_crease_smooth_veto_deg = o.meshCreaseSmoothVetoDeg;
  eo.mesh_crease_convex_only = o.meshCreaseConvexOnly ? 1 : 0;
  eo.mesh_border_coplanar_veto_deg = o.meshBorderCoplanarVetoDeg;
  eo.mesh_crease_max_degree = o.meshCreaseMaxDegree;
  ubg_edges e{};
  const int st = ubg_object_space_edges(device(), &flat.desc, &cam, &eo, &e);
  if (st != UBG_OK) throwStatus(st, "object-space edges");
  flat.appendEdges(e);
  ubg_edges_free(&e);
}

// renderFrame (pipeline.cpp:20-256) for one pass on the GPU.
FrameResult renderPass(const Scene& scene, const GroupFilter& groups, const RenderOptions& opt,
                       RenderProgress* progress) {
  ubg_context* ctx = device();
  const int ss = std::max(1, opt.supersample);
  if (progress) progress->beginPhase(RenderPhase::Setup);
  FlatScene flat(scene, groups);
  // pipeline.cpp:30-42: the object-space edges of what this pass draws
  if (opt.objectSpaceEdges.enable) addObjectSpaceEdges(flat, scene.camera, opt.objectSpaceEdges);
  logMessage(LogLevel::Info, "render %dx%d ss%d (grid %dx%d), %lld tris / %lld spheres / %lld cylinders%s", opt.width,
             opt.height, ss, opt.width * ss, opt.height * ss, static_cast<long long>(flat.desc.n_triangles),
             static_cast<long long>(flat.desc.n_spheres), static_cast<long long>(flat.desc.n_cylinders),
             opt.aoSamples > 0 ? ", AO" : "");
  FrameResult f;
  f.effectiveTriangles = flat.effectiveTriangles();
  if (progress && progress->cancelRequested()) {
    f.cancelled = true;
    return f;
  }
  ubg_scene* gs = nullptr;
  // the reference rebuilds its BVH on every render (scene_build.cpp:133-170);
  // here the BVH8 is built on the GPU (Morton order + PLOC + the 8-wide
  // collapse), the host only flattening the scene into primitive records
  int st = ubg_scene_create_ex(ctx, &flat.desc, UBG_BVH8_GPU_BUILD, &gs);
  if (st != UBG_OK) throwStatus(st, "scene build");
  ubg_scene_info info{};
  ubg_scene_get_info(gs, &info);

  ubg_render_params p{};
  p.width = opt.width;
  p.height = opt.height;
  p.supersample = ss;
  const Camera& c = scene.camera;
  const float cp[3] = {c.position.x, c.position.y, c.position.z};
  const float cd[3] = {c.direction.x, c.direction.y, c.direction.z};
  const float cu[3] = {c.up.x, c.up.y, c.up.z};
  std::memcpy(p.camera.position, cp, sizeof(cp));
  std::memcpy(p.camera.direction, cd, sizeof(cd));
  std::memcpy(p.camera.up, cu, sizeof(cu));
  p.camera.fovy = c.fovy;
  p.camera.orthographic = c.orthographic ? 1 : 0;
  p.camera.height = c.height;
  std::vector<ubg_light> lights;
  for (const DistantLight& l : scene.lights) {
    ubg_light L{};
    L.direction[0] = l.direction.x, L.direction[1] = l.direction.y, L.direction[2] = l.direction.z;
    L.color[0] = l.color.x, L.color[1] = l.color.y, L.color[2] = l.color.z;
    L.intensity = l.intensity;
    L.casts_highlight = l.castsHighlight ? 1 : 0;
    L.angular_radius = l.angularRadius;
    lights.push_back(L);
  }
  p.lights = lights.data();
  p.n_lights = static_cast<int32_t>(lights.size());
  p.ambient_color[0] = scene.ambientColor.x, p.ambient_color[1] = scene.ambientColor.y,
  p.ambient_color[2] = scene.ambientColor.z;
  p.background[0] = scene.background.x, p.background[1] = scene.background.y, p.background[2] = scene.background.z;
  p.fog.enabled = scene.fog.enabled ? 1 : 0;
  p.fog.color[0] = scene.fog.color.x, p.fog.color[1] = scene.fog.color.y, p.fog.color[2] = scene.fog.color.z;
  p.fog.start = scene.fog.start;
  p.fog.end = scene.fog.end;
  p.clip_near = scene.clipNear;
  p.clip_far = scene.clipFar;
  p.assumed_gamma = scene.assumedGamma;
  p.ao_samples = opt.aoSamples;
  p.ao_distance = opt.aoDistance;
  p.ao_intensity = opt.aoIntensity;
  p.shadows = opt.shadows ? 1 : 0;
  p.shadow_samples = opt.shadowSamples;
  p.light_radius_deg = opt.lightRadius;
  p.gi_integrator = opt.giIntegrator;
  p.specular_scale = opt.specularScale;
  p.transparency = opt.transparency ? 1 : 0;
  p.transparent_background = opt.transparentBackground ? 1 : 0;
  p.max_transparent_layers = opt.maxTransparentLayers;
  p.cancel_flag = progress ? detail::ProgressAccess::cancelFlag(*progress) : nullptr;
  p.ao_falloff_power = opt.aoFalloffPower;
  p.ao_multi_scale = opt.aoMultiScale ? 1 : 0;
  p.ao_bent_normal = opt.aoBentNormal ? 1 : 0;
  p.ao_use_camera_up = opt.aoUseCameraUp ? 1 : 0;
  for (int k = 0; k < 3; ++k) {
    p.ao_sky_color[k] = opt.aoSkyColor[k];
    p.ao_ground_color[k] = opt.aoGroundColor[k];
    p.ao_up[k] = opt.aoUp[k];
  }
  p.ao_multibounce = opt.aoMultibounce ? 1 : 0;
  p.ao_low_discrepancy = opt.aoLowDiscrepancy ? 1 : 0;
  p.ao_diffuse_factor = opt.aoDiffuseFactor;
  p.ao_res_div = opt.aoResDiv;
  p.aa_mode = opt.aaMode;
  p.aa_threshold = opt.aaThreshold;
  p.aa_depth = opt.aaDepth;
  p.gi = opt.gi ? 1 : 0;
  p.gi_bounces = opt.giBounces;
  p.gi_max_distance = opt.giMaxDistance;
  p.gi_intensity = opt.giIntensity;
  p.gi_env_intensity = opt.giEnvIntensity;
  p.pt1_spp = opt.pt1Spp;
  p.pt1_seed = opt.pt1Seed;
  p.pt1_sky_mode = opt.pt1SkyMode;
  for (int k = 0; k < 3; ++k) p.pt1_sky_radiance[k] = opt.pt1SkyRadiance[k];
  p.pt1_ld = opt.pt1Ld ? 1 : 0;
  p.pt1_clamp = opt.pt1Clamp;
  p.pt1_gather_div = opt.pt1GatherDiv;
  p.pt1_half_res = opt.pt1HalfRes ? 1 : 0;
  p.pt1_denoise = opt.pt1Denoise ? 1 : 0;
  p.denoise_iters = opt.denoiseIters;
  p.denoise_sigma_z = opt.denoiseSigmaZ;
  p.denoise_sigma_n = opt.denoiseSigmaN;
  p.denoise_sigma_l = opt.denoiseSigmaL;
  p.pt1_upsample_normal_pow = opt.pt1UpsampleNormalPow;
  p.pt1_upsample_depth_scale = opt.pt1UpsampleDepthScale;
  p.pt1_edge_patch = opt.pt1EdgePatch ? 1 : 0;
  p.pt1_edge_patch_spp_mul = opt.pt1EdgePatchSppMul;
  p.pt1_edge_patch_thresh = opt.pt1EdgePatchThresh;
  p.pt2_pattern = opt.pt2Pattern;
  p.pt2_emissive = opt.pt2Emissive ? 1 : 0;
  p.pt2_rounds = opt.pt2Rounds;
  p.pt2_adaptive = opt.pt2Adaptive ? 1 : 0;
  p.pt2_emissive_nee = opt.pt2EmissiveNee ? 1 : 0;
  p.pt2_reflect = opt.pt2Reflect ? 1 : 0;

  const std::size_t hiN = static_cast<std::size_t>(opt.width) * ss * opt.height * ss;
  f.width = opt.width;
  f.height = opt.height;
  f.color.assign(static_cast<std::size_t>(opt.width) * opt.height * 4, 0.0f);
  f.depth.assign(hiN, 0.0f);
  if (scene.fog.enabled) f.viewZ.assign(hiN, 0.0f);
  ubg_frame fr{};
  fr.color = f.color.data();
  fr.depth = f.depth.data();
  fr.view_z = f.viewZ.empty() ? nullptr : f.viewZ.data();
  // FrameResult::aaMask: Wf x Hf, only for an adaptive render with aaDebug
  // (embree_renderer.cpp:1379-1385, 1540)
  if (opt.aaMode == 1 && (ss > 1 || opt.aaDepth > 1) && opt.aaDebug)
    f.aaMask.assign(static_cast<std::size_t>(opt.width) * opt.height, 0.0f);
  fr.aa_mask = f.aaMask.empty() ? nullptr : f.aaMask.data();
  // GI AOVs (embree_renderer.cpp:463-519 sizes them on the render grid)
  if (opt.gi) {
    f.indirect.assign(hiN * 3, 0.0f);
    f.giOcclusion.assign(hiN, 0.0f);
  }
  fr.gi_indirect = f.indirect.empty() ? nullptr : f.indirect.data();
  fr.gi_occlusion = f.giOcclusion.empty() ? nullptr : f.giOcclusion.data();
  fr.memory = 0;
  if (progress) progress->beginPhase(RenderPhase::Primary);
  st = ubg_render(ctx, gs, &p, &fr);
  ubg_scene_destroy(gs);
  if (st == UBG_ERR_CANCELLED) {
    f.cancelled = true;
    return f;
  }
  if (st != UBG_OK) throwStatus(st, "render");
  if (fr.stats.capped_rays > 0)
    logMessage(LogLevel::Warning,
               "%llu ray(s) reached the transparent-layer cap (%d); transmission past that depth was dropped -- "
               "raise maxTransparentLayers if it is visible",
               static_cast<unsigned long long>(fr.stats.capped_rays), opt.maxTransparentLayers);
  if (progress) progress->beginPhase(RenderPhase::Postprocess);
  f.renderSeconds = fr.render_seconds;
  f.pt1Rays.gatherRays = fr.stats.gi_rays;
  f.pt1Rays.neeRays = fr.stats.gi_nee_rays;
  f.pt1Rays.gbufferRays = fr.stats.gi_gbuffer_rays;
  // pt1_denoise.cpp:61-66: without OIDN the E denoise is the a-trous filter
  // (DenoiserBackend::AtrousBilateral == 1)
  if (opt.gi && opt.pt1Denoise) f.pt1DenoiserUsed = 1;
  f.pt1Timing.bvhBuild = info.build_seconds;
  f.pt1Timing.primary = fr.render_seconds;
  return f;
}

// Group-alpha compositing (umbreon.cpp:141-254 defines the result): the
// frame is a weighted sum of passes taken on sRGB-encoded colour with linear
// alpha, then decoded. One pass draws the backdrop (every blended group
// hidden) with weight 1 - sum of the alphas; each blend entry adds a pass
// that also draws its group, with its alpha as weight.
struct BlendPass {
  GroupFilter groups;
  float weight = 0.0f;
};

std::vector<BlendPass> blendPlan(const std::vector<GroupBlend>& blend) {
  std::size_t span = 0;
  float backdrop = 1.0f;
  for (const GroupBlend& g : blend) {
    span = std::max<std::size_t>(span, static_cast<std::size_t>(g.group) + 1);
    backdrop -= g.alpha;
  }
  GroupFilter none;
  none.hidden.assign(span, 0);
  for (const GroupBlend& g : blend) none.hidden[g.group] = 1;
  std::vector<BlendPass> plan;
  plan.push_back({none, backdrop});
  for (const GroupBlend& g : blend) {
    BlendPass p{none, g.alpha};
    p.groups.hidden[g.group] = 0;
    plan.push_back(std::move(p));
  }
  return plan;
}

// Running weighted sum of RGBA frames: rgb in the sRGB-encoded domain,
// alpha linear.
class SrgbBlend {
 public:
  void add(const std::vector<float>& rgba, float w) {
    if (sum_.size() < rgba.size()) sum_.resize(rgba.size(), 0.0f);
    for (std::size_t i = 0; i < rgba.size(); ++i)
      sum_[i] += (i % 4 == 3) ? w * rgba[i] : w * srgbEncodeF(rgba[i]);
  }
  // Decoded frame: rgb = decode(max(0, sum)), alpha clamped to [0, 1].
  std::vector<float> take() {
    for (std::size_t i = 0; i < sum_.size(); ++i)
      sum_[i] = (i % 4 == 3) ? std::clamp(sum_[i], 0.0f, 1.0f) : srgbDecodeF(std::max(0.0f, sum_[i]));
    return std::move(sum_);
  }

 private:
  std::vector<float> sum_;
};

FrameResult renderImpl(const Scene& scene, const RenderOptions& opt, RenderProgress* progress) {
  checkSupported(scene, opt);
  if (scene.groupBlend.empty()) {
    FrameResult f = renderPass(scene, GroupFilter{}, opt, progress);
    if (progress && !f.cancelled) progress->markDone();
    return f;
  }
  const std::vector<BlendPass> plan = blendPlan(scene.groupBlend);
  SrgbBlend blend;
  FrameResult out;
  double traceSeconds = 0.0, buildSeconds = 0.0;
  for (std::size_t i = 0; i < plan.size() && !out.cancelled; ++i) {
    if (progress) progress->beginPass(i, plan.size());
    logMessage(LogLevel::Info, "group-alpha: pass %zu of %zu, weight %.3f", i + 1, plan.size(),
               static_cast<double>(plan[i].weight));
    FrameResult f = renderPass(scene, plan[i].groups, opt, progress);
    traceSeconds += f.renderSeconds;
    buildSeconds += f.pt1Timing.bvhBuild;
    blend.add(f.color, plan[i].weight);
    out = std::move(f);  // the last pass's buffers carry depth and the AOVs
  }
  out.color = blend.take();
  out.renderSeconds = traceSeconds;
  out.pt1Timing.bvhBuild = buildSeconds;
  out.pt1Timing.primary = traceSeconds;
  if (progress && !out.cancelled) progress->markDone();
  return out;
}

}  // namespace

FrameResult render(const Scene& scene, const RenderOptions& opt) { return renderImpl(scene, opt, nullptr); }

FrameResult render(const Scene& scene, const RenderOptions& opt, RenderProgress& progress) {
  return renderImpl(scene, opt, &progress);
}

// ---------------------------------------------------------------- async
// A background render is a packaged task on its own thread: the future
// carries the frame or the exception, the task's progress object is the one
// the render polls for cancellation.
struct RenderTask::Impl {
  Scene scene;
  RenderOptions opt;
  RenderProgress progress;
  std::future<FrameResult> frame;
  std::thread worker;
  bool collected = false;

  // finished: the frame was collected, or the future holds it
  bool ready() const {
    return collected || (frame.valid() && frame.wait_for(std::chrono::seconds(0)) == std::future_status::ready);
  }
  void stop() {
    progress.requestCancel();
    if (worker.joinable()) worker.join();
  }
};

RenderTask::RenderTask(std::unique_ptr<Impl> impl) noexcept : p_(std::move(impl)) {}
RenderTask::RenderTask(RenderTask&&) noexcept = default;
RenderTask& RenderTask::operator=(RenderTask&& o) noexcept {
  if (this != &o) {
    if (p_) p_->stop();
    p_ = std::move(o.p_);
  }
  return *this;
}
RenderTask::~RenderTask() {
  if (p_) p_->stop();
}

float RenderTask::progress() const noexcept { return p_ ? p_->progress.fraction() : 0.0f; }
RenderPhase RenderTask::phase() const noexcept { return p_ ? p_->progress.phase() : RenderPhase::Idle; }
bool RenderTask::done() const noexcept { return p_ && p_->ready(); }
void RenderTask::cancel() noexcept {
  if (p_) p_->progress.requestCancel();
}
bool RenderTask::wait_for(std::chrono::milliseconds timeout) const {
  return !p_ || !p_->frame.valid() || p_->frame.wait_for(timeout) == std::future_status::ready;
}
void RenderTask::wait() const {
  if (p_ && p_->frame.valid()) p_->frame.wait();
}
FrameResult RenderTask::get() {
  if (!p_) throw std::runtime_error("umbreon::RenderTask::get on an empty task");
  if (p_->collected) throw std::runtime_error("umbreon::RenderTask::get called twice");
  p_->collected = true;
  if (p_->worker.joinable()) p_->worker.join();
  return p_->frame.get();  // rethrows the render's exception
}

RenderTask renderAsync(Scene scene, RenderOptions opt) {
  auto impl = std::make_unique<RenderTask::Impl>();
  impl->scene = std::move(scene);
  impl->opt = std::move(opt);
  RenderTask::Impl* task = impl.get();
  std::packaged_task<FrameResult()> job([task] { return render(task->scene, task->opt, task->progress); });
  task->frame = job.get_future();
  task->worker = std::thread(std::move(job));
  return RenderTask(std::move(impl));
}

}  // namespace umbreon
