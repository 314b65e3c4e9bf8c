/*
 * umbreon_gpu.h -- C-ABI of the MI355X (gfx950) per-ray backend of libumbreon.
 *
 * This is the drop-in boundary. Every type below is plain C (no C++ or torch
 * types), every buffer is a plain pointer plus a count, and every call returns
 * an int status (UBG_OK == 0). The C++ facade (umbreon::render, include/umbreon/)
 * flattens umbreon::Scene into ubg_scene_desc and RenderOptions into
 * ubg_render_params and converts non-zero statuses into std::runtime_error.
 *
 * What each entry point replaces in the reference (CueMol/umbreon):
 *   ubg_context_create / _destroy  -> rtcNewDevice + SharedDevice reuse
 *                                     (src/umbreon/render/embree_renderer.cpp:1267-1273,
 *                                      src/umbreon/umbreon.cpp:21-39), releaseEmbree
 *                                     (embree_renderer.cpp:281-292)
 *   ubg_scene_create / _destroy    -> buildEmbreeScene (src/umbreon/render/scene_build.cpp:133-170)
 *                                     incl. buildTriangleMesh (:16-88), buildSpheres (:91-129),
 *                                     buildCylinderGeometry (src/umbreon/render/curve_build.cpp:41-267)
 *                                     and the BuiltScene side tables (scene_build.hpp:63-118)
 *   ubg_render                     -> EmbreeRenderer::render grid path
 *                                     (embree_renderer.cpp:1260-1427) and its adaptive-AA
 *                                     driver (:1428-1674, render/adaptive_aa.hpp) + applyFog
 *                                     (src/umbreon/postprocess/fog.cpp:17-36) + boxDownsample /
 *                                     applyAssumedGamma (src/umbreon/postprocess/image_ops.cpp:9-46),
 *                                     i.e. renderFrame's trace/fog/resolve steps
 *                                     (src/umbreon/render/pipeline.cpp:113-129,182-254)
 *   ubg_intersect / ubg_occluded   -> batched rtcIntersect1 / rtcOccluded1 (Embree 4 public API,
 *                                     call sites shading/transparency.hpp:382,
 *                                     shading/secondary_rays.hpp:95) and the edge-pass seam
 *                                     EmbreeRenderer::occluded (embree_renderer.hpp:68-70)
 *   ubg_last_error                 -> the reference's std::runtime_error messages
 *                                     (embree_renderer.cpp:1272, scene_build.cpp:164-168)
 */
#ifndef UMBREON_GPU_H
#define UMBREON_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UBG_ABI_VERSION 8

/* Status codes. */
enum {
  UBG_OK = 0,
  UBG_ERR_INVALID_ARGUMENT = 1,
  UBG_ERR_DEVICE = 2,          /* no usable gfx950 device / HIP runtime error */
  UBG_ERR_OUT_OF_MEMORY = 3,
  UBG_ERR_CANCELLED = 4,       /* cancel_flag was raised; frame is partial */
  UBG_ERR_UNSUPPORTED = 5,     /* option/material the GPU path does not implement */
  UBG_ERR_INTERNAL = 6
};

/* Geometry kinds. The numbering is the reference's GeomKind (scene_build.hpp),
 * which also forms the low two bits of the edge objectId. */
enum {
  UBG_KIND_MESH = 0,
  UBG_KIND_SPHERE = 1,
  UBG_KIND_ROUND_CURVE = 2,   /* `open` cylinders, chained (ROUND_LINEAR_CURVE) */
  UBG_KIND_CONE_CURVE = 3     /* capped cylinders (CONE_LINEAR_CURVE) */
};

/* Flattened umbreon::Material (scene.hpp:130-214). 64 bytes. */
typedef struct ubg_material {
  float ambient, diffuse, specular, roughness, brilliance, phong, phong_size;
  int32_t metallic;
  float reflection, emission;
  int32_t model; /* 0 = Pov, 1 = Principled, 2 = Toon */
  float pbr_metallic, pbr_roughness, pbr_specular, pbr_anisotropy,
      pbr_anisotropy_rotation;
} ubg_material;

/* Flattened umbreon::Scene geometry (scene.hpp:225-368). All arrays are
 * borrowed for the duration of ubg_scene_create only. Optional arrays may be
 * NULL (meaning: the reference default). */
typedef struct ubg_scene_desc {
  /* Mesh::positions/normals/colors (per vertex) and Mesh::index. */
  int64_t n_vertices;
  const float* positions;       /* 3 * n_vertices */
  const float* normals;         /* 3 * n_vertices */
  const float* colors;          /* 4 * n_vertices, rgb + opacity */
  int64_t n_triangles;
  const uint32_t* index;        /* 3 * n_triangles, NULL => corner c uses vertex c */
  /* Mesh materials: tri_material == NULL => every triangle uses
   * mesh_materials[0] (Mesh::material); else mesh_materials[tri_material[t]]
   * (Mesh::materials / Mesh::triMaterialId). */
  const ubg_material* mesh_materials;
  int32_t n_mesh_materials;
  const uint8_t* tri_material;  /* n_triangles or NULL */
  const uint16_t* tri_group;    /* n_triangles or NULL (group 0) */

  /* Scene::spheres. */
  int64_t n_spheres;
  const float* sphere_center_radius; /* 4 * n: x, y, z, radius */
  const float* sphere_color;         /* 4 * n: rgb + opacity */
  const int32_t* sphere_material;    /* n: index into prim_materials */
  const uint16_t* sphere_group;      /* n or NULL */
  const uint8_t* sphere_from_edge;   /* n or NULL (Sphere::fromEdgeMacro) */

  /* Scene::cylinders. cyl_open selects the curve kind exactly like
   * Cylinder::open: 1 => chained ROUND curve, 0 => capped CONE curve. */
  int64_t n_cylinders;
  const float* cyl_p0;               /* 3 * n */
  const float* cyl_p1;               /* 3 * n */
  const float* cyl_radius;           /* n */
  const float* cyl_color;            /* 4 * n: rgb + opacity at p0 */
  const float* cyl_opacity1;         /* n or NULL (=> -1, uniform) */
  const int32_t* cyl_material;       /* n: index into prim_materials */
  const uint16_t* cyl_group;         /* n or NULL */
  const uint8_t* cyl_open;           /* n or NULL (=> capped) */
  const uint8_t* cyl_from_edge;      /* n or NULL */

  const ubg_material* prim_materials; /* shared by spheres and cylinders */
  int32_t n_prim_materials;

  /* Scene::instanceOffsets, baked like the reference (scene_build.cpp:148-149). */
  int64_t n_instances;
  const float* instance_offsets;     /* 3 * n or NULL */
} ubg_scene_desc;

typedef struct ubg_camera {         /* umbreon::Camera (scene.hpp:380-388) */
  float position[3];
  float direction[3];
  float up[3];
  float fovy;                        /* degrees */
  int32_t orthographic;
  float height;                      /* ortho image-plane height */
} ubg_camera;

typedef struct ubg_light {          /* umbreon::DistantLight (scene.hpp:390-411) */
  float direction[3];                /* direction the light travels */
  float color[3];
  float intensity;
  int32_t casts_highlight;
  float angular_radius;              /* radians */
} ubg_light;

typedef struct ubg_fog {            /* umbreon::Fog (scene.hpp:413-428), linear GL fog */
  int32_t enabled;
  float color[3];
  float start;
  float end;
} ubg_fog;

/* ubg_render_params::flags */
#define UBG_FLAG_HIRES_ONLY 1u   /* stop after trace+fog: color is the hi-res
                                    (supersampled) linear buffer, no downsample
                                    and no assumed gamma */
#define UBG_FLAG_COUNT_VISITS 2u /* also count BVH node visits and primitive
                                    tests (instrumented kernels; slower) */
#define UBG_FLAG_KERNEL_TIMES 4u /* time each wavefront kernel with stream
                                    events (ubg_frame::kernel_ms) */

/* Everything one frame reads: scene-level constants (camera, lights, ambient,
 * background, fog, clip planes, assumed gamma) plus the per-ray subset of
 * RenderOptions (render/render_options.hpp:14-328). */
typedef struct ubg_render_params {
  int32_t width, height;             /* FINAL image size (RenderOptions::width/height) */
  int32_t supersample;               /* hi-res grid = width*ss x height*ss */
  ubg_camera camera;
  const ubg_light* lights;
  int32_t n_lights;
  float ambient_color[3];            /* Scene::ambientColor */
  float background[3];               /* Scene::background */
  ubg_fog fog;
  float clip_near, clip_far;         /* Scene::clipNear/clipFar (+-inf = off) */
  float assumed_gamma;               /* Scene::assumedGamma */
  int32_t ao_samples;
  float ao_distance;
  float ao_intensity;
  int32_t shadows;
  int32_t shadow_samples;
  float light_radius_deg;            /* RenderOptions::lightRadius */
  int32_t gi_integrator;             /* only selects the per-light radius rule */
  float specular_scale;
  int32_t transparency;
  int32_t transparent_background;
  int32_t max_transparent_layers;
  uint32_t flags;                    /* UBG_FLAG_* */
  /* Cooperative cancel (RenderProgress::requestCancel): polled between
   * wavefront passes; a non-zero value stops the frame (UBG_ERR_CANCELLED). */
  const volatile int32_t* cancel_flag;
  /* Enhanced AO (render_options.hpp:52-61, ambient_occlusion.hpp:98-267,
   * ao_shade.hpp:44-141). Any of falloff > 0, multi_scale, bent_normal,
   * multibounce, low_discrepancy or diffuse_factor > 0 switches AO to the
   * nearest-hit estimator; all zero keeps the legacy binary estimator. The sky
   * and ground colours and the gradient axis are read only with bent_normal;
   * ao_up is used when ao_use_camera_up == 0 (else the camera's true up). */
  float ao_falloff_power;
  int32_t ao_multi_scale;
  int32_t ao_bent_normal;
  float ao_sky_color[3];
  float ao_ground_color[3];
  int32_t ao_use_camera_up;
  float ao_up[3];
  int32_t ao_multibounce;
  int32_t ao_low_discrepancy;
  float ao_diffuse_factor;
  /* Coarse-grid AO (render_options.hpp:64-74, ao/ao_coarse.hpp): > 1 gathers
   * AO once per cell of the hi-res grid divided by this factor on a first-hit
   * G-buffer, and shading hits take the guided bilateral lookup of it, or the
   * exact inline gather where the guides reject; -1 = the supersample factor;
   * 0 or 1 = inline AO everywhere. */
  int32_t ao_res_div;
  /* Adaptive antialiasing (render_options.hpp:20-39, render/adaptive_aa.hpp,
   * embree_renderer.cpp:1428-1674). aa_mode 1 with supersample > 1 or
   * aa_depth > 1: shade the centre subpixel of every output pixel, flag the
   * output pixels whose 4-neighbour pairs differ (first-hit kind or group,
   * colour or alpha contrast > aa_threshold, shading-normal fold, depth
   * crack; then a 3x3 dilation), shade every subpixel of the flagged pixels
   * (aa_depth > supersample: ceil(aa_depth / ss)^2 box-averaged samples per
   * subpixel) and replicate the centre elsewhere, re-evaluated with ss^2
   * times the AO / soft-shadow samples when those are stochastic.
   * aa_mode 0 = the full supersample grid. */
  int32_t aa_mode;
  float aa_threshold;                /* reference default 0.1 */
  int32_t aa_depth;
  /* Path-traced diffuse GI, the pt1 integrator (render_options.hpp:93-162,
   * integrator/pt1/{pt1_gather,pt1_gbuffer,pt1_upsample}.hpp, pt1_denoise.cpp,
   * embree_renderer.cpp:560-1131). gi != 0 requires gi_integrator 1 or 2
   * (the irradiance cache, 0, is UBG_ERR_UNSUPPORTED). The gather grid is the render grid
   * divided by pt1_gather_div (-1 = the supersample factor, 0 = 2 with
   * pt1_half_res else 1); a divisor > 1 traces its own first-hit G-buffer,
   * gathers (and denoises) there, joint-bilateral upsamples E to the render
   * grid and, with pt1_edge_patch, re-gathers the rim pixels the upsample
   * cannot support at full resolution with pt1_edge_patch_spp_mul x the
   * samples. The E denoise is the a-trous filter (the reference's path when
   * OIDN is not built). Mesh hits and real (not outline) primitives become
   * receivers: their constant ambient is dropped (toon finishes keep unit
   * ambient) and pt1_spp cosine paths per hi-res pixel gather E (next-event
   * estimation at the distant lights, Russian roulette from the second
   * vertex, the sky/ground environment at escapes); the frame adds
   * gi_intensity * diffuse * pigment * E before fog. Under gi the reference
   * turns adaptive AA and coarse AO off, and so does this path. */
  int32_t gi;
  int32_t gi_bounces;
  float gi_max_distance;             /* <= 0: unbounded */
  float gi_intensity;
  float gi_env_intensity;
  int32_t pt1_spp;
  uint32_t pt1_seed;
  int32_t pt1_sky_mode;              /* 0 uniform sky; 1 gradient (ground = ao_ground_color) */
  float pt1_sky_radiance[3];
  int32_t pt1_ld;                    /* Hammersley first bounce with a per-pixel shift */
  float pt1_clamp;                   /* per-sample luminance clamp; 0 = off */
  int32_t pt1_gather_div;            /* -1 = supersample factor, 0 = pt1_half_res rule, k */
  int32_t pt1_half_res;
  int32_t pt1_denoise;
  int32_t denoise_iters;             /* a-trous iterations (reference default 5) */
  float denoise_sigma_z, denoise_sigma_n, denoise_sigma_l;  /* defaults 1, 128, 4 */
  float pt1_upsample_normal_pow;     /* default 32 */
  float pt1_upsample_depth_scale;    /* default 0.02 */
  int32_t pt1_edge_patch;            /* default 1 */
  int32_t pt1_edge_patch_spp_mul;    /* default 4 */
  float pt1_edge_patch_thresh;       /* default 0.3 */
  /* pt2 (gi_integrator 2, the reference default; render_options.hpp:184-253,
   * integrator/pt2/): the pt1 gather with the first bounce drawn from the
   * Owen-scrambled Sobol sequence (pt2_pattern 0 per pixel, 1 blue noise),
   * area lights sampled by one cone ray per path vertex, and (pt2_emissive)
   * emission transported at the vertices. ReSTIR (pt2_rounds > 0), adaptive
   * spp, emissive-triangle NEE (pt2_emissive_nee with an emissive material
   * in the scene) and traced reflections (pt2_reflect with a reflective
   * material) are not implemented: UBG_ERR_UNSUPPORTED. */
  int32_t pt2_pattern;               /* default 1 */
  int32_t pt2_emissive;              /* default 1 */
  int32_t pt2_rounds;                /* default 0 */
  int32_t pt2_adaptive;              /* default 0 */
  int32_t pt2_emissive_nee;          /* default 1 */
  int32_t pt2_reflect;               /* default 1 */
  /* pt1_denoise == 2: the learned denoiser (ubg_unet_*, a network on this
   * context) replaces the a-trous filter on the gathered irradiance; needs
   * the full-resolution gather (pt1_gather_div resolving to 1). */
  struct ubg_unet* pt1_denoise_net;
} ubg_render_params;

/* Ray accounting of one frame. A pure function of the scene and options:
 * identical for every launch configuration. */
typedef struct ubg_ray_stats {
  uint64_t primary_rays;      /* closest-hit queries, incl. transparency continuations */
  uint64_t shaded_hits;       /* hits shaded (one per layer walked) */
  uint64_t ao_rays_traced;    /* AO any-hit queries */
  uint64_t ao_rays_sampled;   /* below-horizon AO samples counted occluded, not traced */
  uint64_t shadow_rays;       /* shadow any-hit queries */
  uint64_t capped_rays;       /* primary rays that reached the transparent-layer cap */
  uint64_t gi_rays;           /* GI gather / continuation closest-hit queries */
  uint64_t gi_nee_rays;       /* GI next-event shadow queries */
  uint64_t gi_gbuffer_rays;   /* GI reduced-grid first-hit queries */
  /* UBG_FLAG_COUNT_VISITS only, per ray class [0] primary, [1] AO, [2] shadow: */
  uint64_t node_visits[3];    /* inner-node fetches */
  uint64_t prim_tests[3];     /* primitive tests */
} ubg_ray_stats;

/* ubg_frame::kernel_ms slots (UBG_FLAG_KERNEL_TIMES) */
enum { UBG_KERNEL_PRIMARY = 0, UBG_KERNEL_AO = 1, UBG_KERNEL_RESOLVE = 2, UBG_KERNEL_POST = 3 };

/* Output buffers, allocated by the caller. memory == 0: host pointers (the
 * library stages through its own device buffers); memory == 1: device
 * pointers on the context's GPU (no host copy). Optional outputs may be NULL. */
typedef struct ubg_frame {
  float* color;        /* out_w * out_h * 4 linear HDR RGBA (top-left origin) */
  float* depth;        /* hi_w * hi_h first-hit ray distance (0 on miss) */
  float* view_z;       /* hi_w * hi_h first-hit linear view-z (0 on miss) */
  uint32_t* hit_id;    /* hi_w * hi_h first-hit (kind << 30) | primID, 0xFFFFFFFF on miss */
  float* ao_open;      /* hi_w * hi_h first-hit AO openness (1 when no AO) */
  int32_t memory;
  /* Filled by ubg_render: */
  int32_t out_w, out_h, hi_w, hi_h;
  ubg_ray_stats stats;
  double render_seconds;   /* trace + shade + fog window (device events) */
  double post_seconds;     /* downsample + gamma */
  int32_t cancelled;
  int32_t layers;          /* transparency layers walked (wavefront passes) */
  double kernel_ms[4];     /* UBG_FLAG_KERNEL_TIMES: summed time per kernel kind */
  int32_t kernel_launches[4];
  /* Adaptive AA refinement mask (FrameResult::aaMask): out_w * out_h floats,
   * 1 = refined, 0 = replicated centre; may be NULL; left untouched when the
   * frame is not adaptive. Same memory space as the other outputs. */
  float* aa_mask;
  int64_t aa_flagged;      /* adaptive AA: output pixels refined (0 otherwise) */
  /* GI AOVs (FrameResult::indirect / giOcclusion), hi_w * hi_h, may be NULL:
   * the composited indirect rgb (3 floats per pixel) and the first-bounce hit
   * fraction of the gather; zero on pixels that are not receivers. */
  float* gi_indirect;
  float* gi_occlusion;
  /* The learned denoiser's input features of the gathered (not yet
   * denoised) irradiance, hi_w * hi_h * 10: log(1 + E) rgb, albedo rgb,
   * normal xyz, first-hit distance / its frame maximum; may be NULL; needs
   * the full-resolution gather. The network (ubg_unet_*) reads 12 features
   * per pixel: these 10, then two zeros for the noise descriptors. */
  float* gi_features;
} ubg_frame;

/* Acceleration-structure layouts (ubg_scene_create_ex). */
enum {
  UBG_BVH2 = 0,          /* binary BVH, full fp32 child boxes, 64 B nodes */
  UBG_BVH8 = 1,          /* 8-wide BVH, 8-bit quantized child boxes, 80 B nodes (default) */
  UBG_BVH8_GPU_BUILD = 2 /* the UBG_BVH8 layout, built on the GPU (Morton order + PLOC
                            clustering + the same SAH-optimal 8-wide collapse) */
};

typedef struct ubg_scene_info {
  int64_t n_triangles, n_spheres, n_round_segments, n_cone_segments;
  int64_t n_nodes;          /* BVH nodes */
  int64_t n_leaf_prims;     /* primitive records in leaf order */
  int32_t bvh_depth;        /* tree levels */
  int32_t node_bytes;       /* bytes per BVH node */
  int32_t bvh_layout;       /* UBG_BVH2 / UBG_BVH8 / UBG_BVH8_GPU_BUILD */
  int32_t stack_depth;      /* traversal stack entries per lane */
  int32_t prim_bytes;       /* bytes per leaf primitive record */
  double build_seconds;     /* upload + BVH build wall time */
  double sah_cost;
  int64_t device_bytes;     /* device memory held by the scene */
} ubg_scene_info;

/* One ray of the batched query API: origin, direction, [tnear, tfar]. */
typedef struct ubg_ray {
  float org[3];
  float tnear;
  float dir[3];
  float tfar;
} ubg_ray;

/* Closest hit of one ray (RTCHit analogue). */
typedef struct ubg_hit {
  float t;             /* +inf on miss */
  float u, v;          /* triangle barycentrics / curve axial fraction in u */
  uint32_t kind;       /* UBG_KIND_*, 0xFFFFFFFF on miss */
  uint32_t prim_id;    /* per-kind primitive id, 0xFFFFFFFF on miss */
  float ng[3];         /* unnormalized geometric normal */
} ubg_hit;

typedef struct ubg_context ubg_context;
typedef struct ubg_scene ubg_scene;

int ubg_abi_version(void);
const char* ubg_status_string(int status);

/* Device context: one gfx950 device, one HIP stream, reusable scratch.
 * Calls on one context are serialized by an internal mutex; use one context
 * per host thread for concurrent renders. */
int ubg_context_create(int device_ordinal, ubg_context** out);
void ubg_context_destroy(ubg_context* ctx);
/* Last error message of the calling thread (never NULL). */
const char* ubg_last_error(void);

/* Upload geometry and build the BVH (UBG_BVH8). */
int ubg_scene_create(ubg_context* ctx, const ubg_scene_desc* desc, ubg_scene** out);
/* Same with an explicit acceleration-structure layout (UBG_BVH2 / UBG_BVH8 /
 * UBG_BVH8_GPU_BUILD); all give the same images (up to which of two
 * primitives hit at exactly the same distance wins), the layout and the
 * builder only change build and traversal speed. */
int ubg_scene_create_ex(ubg_context* ctx, const ubg_scene_desc* desc, int32_t bvh_layout, ubg_scene** out);
/* Scene build options. leaf_max: primitives per BVH8 leaf slot, 1..3
 * (0 = 3); a tuning knob -- images do not depend on it (primitives hit at
 * the same distance resolve by (kind, primID), whatever the tree). */
typedef struct ubg_scene_options {
  int32_t bvh_layout;   /* UBG_BVH2 / UBG_BVH8 / UBG_BVH8_GPU_BUILD */
  int32_t leaf_max;
  int32_t reserved[6];
} ubg_scene_options;
int ubg_scene_create_opts(ubg_context* ctx, const ubg_scene_desc* desc, const ubg_scene_options* opts,
                          ubg_scene** out);
void ubg_scene_destroy(ubg_scene* scene);
int ubg_scene_get_info(const ubg_scene* scene, ubg_scene_info* out);

/* Render one frame (hi-res trace + shade + fog, then resolve). */
int ubg_render(ubg_context* ctx, const ubg_scene* scene,
               const ubg_render_params* params, ubg_frame* frame);

/* Tuning of a batched render (may be NULL: the defaults). Any choice gives
 * the same images; it only changes speed. */
typedef struct ubg_batch_options {
  int32_t ao_waves_per_simd;   /* persistent AO grid: waves per SIMD, 1..5 (0 = 5; larger
                                  values are clamped to 5, the kernel's register budget) */
  int32_t no_graph;            /* 1: launch each layer's kernels directly instead of
                                  replaying its instantiated HIP graph */
  int32_t reserved[6];
} ubg_batch_options;

/* Batched render (a server's path; the reference serves one request per
 * call, umbreon.cpp:322-341 renderAsync): n <= 256 requests, each with its
 * own scene, frame parameters and outputs (frames[i].memory per request),
 * traced by the same wavefront launches -- one launch of each kernel per
 * transparency layer covers every request still walking layers, the AO pass
 * is one persistent grid that drains the requests' hit queues in turn, and
 * the queues and frame buffers are carved from one shared pool. A layer's
 * launch sequence (queue reset, primary, AO, resolve, fill read-back) is
 * captured once per launch shape into a HIP graph kept by the context and
 * replayed by later layers and batches of that shape. Every
 * request's image, AOVs and ray counts equal those of ubg_render on it alone
 * (the determinism contract of docs/api/libumbreon.md:586-587 extended to
 * batch composition). Supported: BVH8 scenes with POV materials, legacy AO,
 * hard / soft shadows, transparency, fog, supersampling; GI, coarse or
 * enhanced AO, adaptive AA, visit counting and kernel timing return
 * UBG_ERR_UNSUPPORTED. render_seconds is the batch's trace window. */
int ubg_render_batch(ubg_context* ctx, int32_t n, const ubg_scene* const* scenes,
                     const ubg_render_params* params, ubg_frame* frames, const ubg_batch_options* opts);

/* ---- Physically based path tracing (ubg_pathtrace) ----
 * No reference counterpart: the reference's GI integrators (pt1 / pt2) are
 * diffuse-only and refraction is outside its scope (docs/api/libumbreon.md:33).
 * GGX conductors (Schlick Fresnel, F0 = the primitive's colour) and GGX
 * dielectrics (exact Fresnel, reflection + refraction, Walter et al. 2007),
 * Lambertian surfaces, spherical area lights reached by next-event
 * estimation only (resampled importance sampling over M candidates at every
 * vertex, ReSTIR spatial reuse over K neighbours at the first), a constant
 * environment for rays that leave the scene. */
typedef struct ubg_sphere_light {
  float center[3];
  float radius;
  float radiance[3];
  float reserved;
} ubg_sphere_light;

enum { UBG_PT_DIFFUSE = 0, UBG_PT_CONDUCTOR = 1, UBG_PT_DIELECTRIC = 2 };

typedef struct ubg_pt_material {    /* indexed like the scene's material indices */
  int32_t type;                      /* UBG_PT_* */
  float roughness;                   /* GGX alpha = roughness^2; < 0.02: smooth */
  float ior;                         /* dielectric: inside / outside */
  int32_t reserved;
} ubg_pt_material;

typedef struct ubg_pt_params {
  int32_t width, height;             /* image (1 camera ray per pixel per sample) */
  ubg_camera camera;
  int32_t spp;                       /* samples per pixel, 1..65536 */
  int32_t bounces;                   /* shaded path vertices, 1..64 */
  uint32_t seed;
  const ubg_sphere_light* lights;    /* host array */
  int32_t n_lights;                  /* 0..4096 */
  float env_radiance[3];
  const ubg_pt_material* prim_materials;  /* host arrays; NULL / 0: diffuse */
  int32_t n_prim_materials;
  const ubg_pt_material* mesh_materials;
  int32_t n_mesh_materials;
  int32_t restir_candidates;         /* M, 1..12 (0 = 8) */
  int32_t restir_spatial;            /* K, 0..16 */
  float restir_radius;               /* pixels (0 = 16) */
  int32_t nee_all_lights;            /* 1: one cone sample of every light at every vertex */
  struct ubg_unet* denoise_net;      /* optional: the colour is the network's denoising, its
                                        prediction u decoded as c = expm1(4 u) / 32 */
  int32_t denoise_ensemble;          /* 1, 2, 4 or 8: u is the mean of the network's
                                        predictions on that many flips / quarter turns of the
                                        features (identity, mirrored columns, mirrored rows,
                                        both, then the same after a transpose), each turned
                                        back; 0 = 8 from 32 spp on (the extra network passes
                                        then cost < 1 % of the frame), else 1 */
  int32_t reserved[7];
} ubg_pt_params;

typedef struct ubg_pt_frame {
  float* color;           /* width * height * 4 (linear HDR, alpha 1) */
  float* features;        /* width * height * 12 or NULL: log(1 + 32 c) / 4 of the noisy
                             colour c (rgb),
                             albedo guide rgb, normal guide xyz (camera space: right,
                             up, towards the eye), first-hit distance / frame max,
                             1 / sqrt(spp), and the standard error of the pixel's mean
                             (r + g + b) / 3 over its spp samples (unbiased variance,
                             0 at 1 spp) times the encoding's slope at that mean,
                             8 / (1 + 32 m), at most 4.
                             The guides are the first hit's, or for a dielectric
                             first hit F * white + (1 - F) * the tinted colour of the
                             first other surface along the ideal refraction (up to 4
                             interfaces) and that surface's normal */
  int32_t memory;         /* 0 host buffers, 1 device */
  double trace_seconds;   /* the spp sample passes (device events) */
  double denoise_seconds; /* features + network + output */
  uint64_t rays;          /* camera, continuation and shadow rays traced */
  int32_t reserved[4];
} ubg_pt_frame;

int ubg_pathtrace(ubg_context* ctx, const ubg_scene* scene, const ubg_pt_params* params, ubg_pt_frame* frame);

/* Size of the fixed frame scratch of ubg_render_noalloc: one device array
 * of this many bytes per GPU the library runs on, held until the process
 * ends. */
#define UBG_NOALLOC_SCRATCH_BYTES (320u << 20)

/* ubg_render without device allocation. It has no counterpart in the
 * reference; it is for callers that cannot allocate device memory during a
 * frame (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same kernels and same images as ubg_render, but every
 * frame-scratch buffer (ray queues, walk state, hi-res colour, depth and view
 * z, counters, adaptive-AA and GI work arrays) is carved from the library's
 * fixed per-GPU scratch array of UBG_NOALLOC_SCRATCH_BYTES (320 MiB: a
 * 1920x1080 frame at 1 spp with AO, shadows and transparency fits; about
 * 150 bytes per hi-res pixel are needed), so it calls no hipMalloc / hipFree.
 * The outputs must be device buffers (frame->memory == 1); a frame that needs
 * more scratch, UBG_FLAG_KERNEL_TIMES (its events are created on demand) and
 * the learned denoiser (its activations are its own allocations) are
 * rejected. Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU (from two contexts, or concurrently with another process's
 * context on the same device) race; calls on one context are serialized. */
int ubg_render_noalloc(ubg_context* ctx, const ubg_scene* scene,
                       const ubg_render_params* params, ubg_frame* frame);

/* Batched closest-hit and any-hit queries (host or device buffers, per
 * `memory` as in ubg_frame). occluded[i] = 1 when any primitive lies in
 * [tnear, tfar] along ray i. */
int ubg_intersect(ubg_context* ctx, const ubg_scene* scene, int64_t n_rays,
                  const ubg_ray* rays, ubg_hit* hits, int32_t memory);
int ubg_occluded(ubg_context* ctx, const ubg_scene* scene, int64_t n_rays,
                 const ubg_ray* rays, uint8_t* occluded, int32_t memory);

/* One segment of the edge pass's visibility query, from p to q. Its
 * exclusion list (the edge's own incident faces, as base mesh triangle ids)
 * is exclude_faces[exclude_first .. exclude_first + exclude_count). */
typedef struct ubg_segment {
  float p[3];
  float q[3];
  int32_t exclude_first;
  int32_t exclude_count;
} ubg_segment;

/* Batched EmbreeRenderer::occluded(p, q, excludeFaces, nExclude, eps,
 * grazeCosEps, coplanarEps) (render/embree_renderer.cpp:294-357, with the
 * argument filter of :48-131): occluded[i] = 1 when a primitive hit lies
 * inside the trimmed segment (max(eps*len, eps), len*(1-eps)) and the filter
 * does not reject it. The filter rejects hits on the segment's excluded mesh
 * faces (instance copies map to their base triangle), hits on faces grazed
 * nearly edge-on (|cos(dir, Ng)| <= graze_cos_eps, when > 0) and hits whose
 * face plane passes within coplanar_eps of p (when > 0); after a rejected
 * hit the primitive's next root along the segment is tried. exclude_faces
 * may be NULL when every exclude_count is 0; all buffers per `memory`. */
int ubg_occluded_segments(ubg_context* ctx, const ubg_scene* scene, int64_t n_segments,
                          const ubg_segment* segments, const int32_t* exclude_faces, int64_t n_exclude_faces,
                          float eps, float graze_cos_eps, float coplanar_eps, uint8_t* occluded,
                          int32_t memory);

/* ObjectSpaceEdgeOptions (render/edge_types.hpp:311-409) without `enable`
 * (calling the entry point is the enable). */
typedef struct ubg_edge_options {
  float width;                        /* edge cylinder radius (default 0.03) */
  float raise;                        /* outward contour offset (0) */
  int32_t segments;                   /* ring tessellation (48, at least 3) */
  float color[3];                     /* linear rgb (0, 0, 0) */
  int32_t clip;                       /* union-boundary clip of the analytic silhouettes (1) */
  int32_t mesh_silhouette;            /* smooth n.v == 0 contour + hard-edge straddle (1) */
  int32_t mesh_crease;                /* dihedral creases (1) */
  int32_t mesh_border;                /* open boundary edges (1) */
  int32_t visibility_clip;            /* mesh hidden-line clip: must be 0 (UBG_ERR_UNSUPPORTED) */
  float mesh_hard_edge_deg;           /* 40 */
  float crease_angle_deg;             /* 30 */
  float mesh_crease_smooth_veto_deg;  /* 0 = off */
  int32_t mesh_crease_convex_only;    /* 0 */
  float mesh_border_coplanar_veto_deg;/* 0 = off */
  int32_t mesh_crease_max_degree;     /* 0 = off */
} ubg_edge_options;

/* The generated edges: n open cylinders of radius max(width, 0) with the
 * flat outline material, in emission order (analytic sphere rings, cylinder
 * side lines / cap rims, then the mesh silhouette, crease and border edges).
 * Arrays are owned by the library until ubg_edges_free. */
typedef struct ubg_edges {
  int64_t n;
  float* p0;          /* 3 * n */
  float* p1;          /* 3 * n */
  float* color;       /* 4 * n: rgb + opacity at p0 */
  float* opacity1;    /* n: opacity at p1, -1 = uniform */
  uint16_t* group;    /* n: the source primitive's / triangle's section */
  float radius;
} ubg_edges;

/* generateObjectSpaceEdges (edges/object_space_edges.cpp:172-264) for the
 * scene's spheres, cylinders and mesh seen by `camera`: the n.v == 0
 * silhouettes of the analytic primitives (fromEdgeMacro ones skipped), each
 * chord split into its runs outside every other solid when `clip` is set
 * (the point-in-solid tests run on the device), then the mesh's feature
 * edges. The caller appends them to Scene::cylinders (open, flat outline,
 * not fromEdgeMacro) before building the scene. The instance offsets are
 * ignored, as in the reference. ctx may be NULL when no clip runs (clip = 0,
 * or no sphere or cylinder silhouettes): the rest is host work. */
int ubg_object_space_edges(ubg_context* ctx, const ubg_scene_desc* desc, const ubg_camera* camera,
                           const ubg_edge_options* options, ubg_edges* out);
void ubg_edges_free(ubg_edges* edges);

/* The learned denoiser: a U-Net over the GI irradiance field and its guides
 * (it replaces denoisePt1E / OIDN, pt1_denoise.cpp:20-81), with bf16 MFMA
 * convolutions and fp32 accumulation, trained on the device. levels 2..6,
 * channels[l] in {32, 64, 96, 128}. Features per pixel (NHWC fp32, 12):
 * log(1 + E) rgb, albedo rgb, normal xyz, depth / max depth, and two noise
 * descriptors (the path tracer's 1 / sqrt(spp) and colour standard error,
 * ubg_pt_frame::features; zero for the GI gather); outputs and targets (3):
 * log(1 + E). h and w must be multiples of 2^(levels - 1).
 * Parameters are fp32, per convolution W[9][cout][cin] then b[cout]. */
typedef struct ubg_unet ubg_unet;
int ubg_unet_create(ubg_context* ctx, int32_t levels, const int32_t* channels, uint64_t seed, ubg_unet** out);
void ubg_unet_destroy(ubg_unet* net);
int64_t ubg_unet_param_count(const ubg_unet* net);
int ubg_unet_get_params(ubg_unet* net, float* params);       /* host buffer */
int ubg_unet_set_params(ubg_unet* net, const float* params); /* host buffer */
/* Denoised log(1 + E) of n images (features and out per `memory`). */
int ubg_unet_infer(ubg_unet* net, int32_t n, int32_t h, int32_t w, const float* features, float* out,
                   int32_t memory);
/* One training step on a batch: the mean loss on log(1 + E) into *loss (L1
 * unless ubg_unet_set_loss chose the squared error), the gradients into
 * `grads` (host, may be NULL), then an Adam step (beta 0.9 / 0.999, eps 1e-8)
 * when lr > 0. Features and targets per `memory`. */
int ubg_unet_train_step(ubg_unet* net, int32_t n, int32_t h, int32_t w, const float* features,
                        const float* targets, float lr, float* loss, float* grads, int32_t memory);
/* Training loss of the network: 0 = mean absolute error (the default), 1 =
 * mean squared error (the PSNR's measure). No counterpart in the reference
 * (OIDN is pretrained there). */
int ubg_unet_set_loss(ubg_unet* net, int32_t mode);

#ifdef __cplusplus
}
#endif

#endif /* UMBREON_GPU_H */
