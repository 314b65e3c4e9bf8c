/*
 * oracle.h -- CPU restatement of libumbreon's per-ray path (TEST INFRASTRUCTURE).
 *
 * This library is the correctness oracle and the CPU baseline of the MI355X
 * backend. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product (umbreon_amd) never links or calls it.
 *
 * It consumes the same flattened scene / parameter structs as the GPU C-ABI
 * (include/umbreon_gpu.h) so both sides see byte-identical inputs, and it
 * implements everything itself: its own curve chaining, its own binned-SAH
 * BVH (or brute force), its own intersectors and the restated shading chain.
 * precision 32 follows the reference's float-op order; precision 64 is the
 * FP64 oracle of the same integrator.
 */
#ifndef UMBREON_ORACLE_H
#define UMBREON_ORACLE_H

#include "../include/umbreon_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_options {
  int32_t precision;    /* 32 or 64 */
  int32_t threads;      /* 0 = hardware concurrency */
  int32_t row_begin;    /* hi-res rows [row_begin, row_end) with step row_step; */
  int32_t row_end;      /*   row_end <= 0 means all rows */
  int32_t row_step;     /*   <= 0 means 1 */
  int32_t brute_force;  /* 1 = test every primitive (no BVH) */
} orc_options;

/* Keep the last built scene (geometry + BVH) for later orc_render calls on
 * an unchanged description (keyed by a hash of its arrays); 0 turns the
 * reuse off and frees it. Off by default. */
void orc_set_scene_cache(int32_t enable);

/* Full frame (renderFrame trace/fog/resolve). Outputs are host pointers. */
int orc_render(const ubg_scene_desc* desc, const ubg_render_params* params,
               ubg_frame* frame, const orc_options* opt);

/* Batched rtcIntersect1 / rtcOccluded1 analogues. */
int orc_intersect(const ubg_scene_desc* desc, int64_t n, const ubg_ray* rays,
                  ubg_hit* hits, const orc_options* opt);
int orc_occluded(const ubg_scene_desc* desc, int64_t n, const ubg_ray* rays,
                 uint8_t* occluded, const orc_options* opt);

/* Batched EmbreeRenderer::occluded (embree_renderer.cpp:294-357) with its
 * face-exclusion / grazing / coincident-plane filter, FP32, brute force. */
int orc_occluded_segments(const ubg_scene_desc* desc, int64_t n, const ubg_segment* segs,
                          const int32_t* exclude_faces, int64_t n_exclude_faces, float eps,
                          float graze_cos_eps, float coplanar_eps, uint8_t* occluded);

/* Curve chaining result (curve_build.cpp:88-199): per round segment, the
 * source cylinder index and the neighbor flags (bit0 LEFT, bit1 RIGHT), in
 * primID order. Returns the segment count, or -1 if `cap` is too small. */
int64_t orc_chain_curves(const ubg_scene_desc* desc, int64_t cap,
                         int32_t* src_cylinder, uint8_t* flags);

/* Adaptive-AA refinement mask (render/adaptive_aa.hpp:66-194, edges mode
 * excluded) over hand-built centre results of a wf x hf output grid: rgba
 * (4 per pixel), view z, shading normal (3 per pixel), first-hit geometry id
 * (0xFFFFFFFF = background) and group (-1 = background). mask = 0/1 per pixel. */
void orc_aa_mask(int32_t wf, int32_t hf, const float* rgba, const float* view_z, const float* normal,
                 const uint32_t* geom, const int32_t* group, float pixel_world_size, float threshold,
                 int32_t perspective, uint8_t* mask);

/* pt1GatherPoint (integrator/pt1/pt1_gather.hpp:228-485) at one shading
 * point, FP32, brute force over the scene: the lights, ambient, environment
 * (pt1_sky_*, gi_env_intensity, ao up axis), gi_bounces, gi_max_distance,
 * pt1_ld and pt1_clamp come from `params`. E = E_stored (3 floats), occlusion
 * = first-bounce hit fraction, gi_rays = closest-hit queries (may be NULL). */
int orc_pt1_gather_point(const ubg_scene_desc* desc, const ubg_render_params* params, const float* P,
                         const float* N, const float* Ng, int32_t spp, uint32_t seed, float eps_t, float* E,
                         float* occlusion, uint64_t* gi_rays);

/* denoisePt1E without OIDN (pt1_denoise.cpp:20-81): NaN/Inf scrub, then the
 * a-trous filter (denoise_atrous.cpp) on E (w * h * 3, in place) with the
 * normal and position guides (w * h * 3 each). */
void orc_pt1_denoise(int32_t w, int32_t h, float* E, const float* normal, const float* position, int32_t iters,
                     float sigma_z, float sigma_n, float sigma_l);

/* pt2 first-bounce sampler (integrator/pt2/pt2_sampler.hpp): the 2D
 * shuffled Owen-scrambled Sobol point `index` under `seed`, and the per-pixel
 * arrangement (pattern 0 per-pixel sequence, 1 blue noise; spp rounded up to
 * a power of two for the section length). */
void orc_pt2_sobol_2d(uint32_t index, uint32_t seed, float* u1, float* u2);
void orc_pt2_pixel_sampler(int32_t pattern, uint32_t px, uint32_t py, uint32_t frame_seed, uint32_t pixel_hash,
                           uint32_t spp, uint32_t* index_base, uint32_t* seed);

/* Reference TEA-8 (secondary_rays.hpp:51-58), exposed for known-answer tests. */
void orc_tea2(uint32_t* v0, uint32_t* v1);

/* Object-space silhouette edges of the spheres and cylinders
 * (generateObjectSpaceEdges without the mesh feature edges), the union clip
 * tested against every solid. Free with orc_edges_free. */
int orc_object_space_edges(const ubg_scene_desc* desc, const ubg_camera* cam, const ubg_edge_options* opt,
                           ubg_edges* out);
void orc_edges_free(ubg_edges* e);

/* The physically based path tracer of ubg_pathtrace (no reference
 * counterpart), restated: color = width * height * 4 (the mean over spp),
 * first = width * height * 8 or NULL (first hit of sample 0: albedo guide rgb,
 * t, camera-space normal guide xyz, 1; zeros where the camera ray misses; a
 * dielectric first hit's guides follow the ideal refraction to the first
 * other surface), rays = rays traced. The denoise_net field is ignored. */
int orc_pathtrace(const ubg_scene_desc* desc, const ubg_pt_params* params, float* color, float* first,
                  uint64_t* rays, const orc_options* opt);
/* Its BSDF in the local frame (n = +z), FP64, over n inputs: f(wo, wi)
 * |cos wi| of a material (type, roughness, ior, base colour C); and one
 * sample of wi for each wo from three uniforms (ok = 0 where the path
 * ends), weight = f |cos| / pdf. Vectors are 3 doubles each. */
void orc_pt_bsdf_eval(int32_t type, float roughness, float ior, const float* C, int64_t n, const double* wo,
                      const double* wi, double* out);
void orc_pt_bsdf_sample(int32_t type, float roughness, float ior, const float* C, int64_t n, const double* wo,
                        const double* u, double* wi, double* weight, uint8_t* ok);
/* Its counter-based uniform for (seed, pixel, sample, dimension). */
/* Its BSDF sampling density (solid angle) of wi for wo, FP64, over n pairs
 * (0 where bsdf_sample never draws wi; smooth lobes are not evaluated). */
void orc_pt_bsdf_pdf(int32_t type, float roughness, float ior, const float* C, int64_t n, const double* wo,
                     const double* wi, double* out);
float orc_pt_rnd(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim);

/* Post-process helpers (fog.cpp:7-36, image_ops.cpp:9-79). */
float orc_fog_factor(const ubg_fog* fog, float z);
void orc_box_downsample(const float* src, int w, int h, int channels, int ss,
                        float* dst);
float orc_srgb_encode(float v);
float orc_srgb_decode(float s);

#ifdef __cplusplus
}
#endif

#endif
