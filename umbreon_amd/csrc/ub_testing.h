// ub_testing.h -- host-side test hook of libumbreon_gpu.so (not part of the
// public C-ABI). Builds a scene exactly as ubg_scene_create does and runs the
// library's own traversal code (the __host__ __device__ functions of
// ub_device.h) on the CPU, so the BVH layouts and the traversal can be checked
// against the CPU oracle on a machine without a GPU.
#pragma once

#include <stdint.h>

#include "../../include/umbreon_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* hits and/or occ may be NULL. For BVH8, occ[i] is 2 when the one-shot and
 * the resumable (per-step) any-hit traversals disagree; the call returns
 * UBG_ERR_INTERNAL when the resumable closest-hit steps find another hit than
 * the one-shot closest-hit traversal. stats (may be NULL) receives
 * [0] node visits, [1] primitive tests (closest-hit queries),
 * [2] tree levels, [3] node count, [4] stack entries, [5] node bytes. */
int ubgx_host_query(const ubg_scene_desc* desc, int32_t bvh_layout, int64_t n_rays, const ubg_ray* rays,
                    ubg_hit* hits, uint8_t* occ, uint64_t* stats);

/* The U-Net convolutions' launch-shape threshold (ubg::Unet::setMinWorkgroups):
 * 0 makes every convolution take the widest pixel tiles its rows allow, so
 * small test images run the kernel instantiations that large frames use. */
typedef struct ubg_unet ubg_unet;
int ubgx_unet_set_min_workgroups(ubg_unet* net, int64_t n);

#ifdef __cplusplus
}
#endif
