// Internal (C++) entry points shared between the kernel files and the host
// classes: the C-ABI functions with optional timing events around the
// dominant kernel of each stage.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "libfusion_amd.h"

// lf_update_weighted / lf_update with events recorded immediately before and
// after the integration kernel (either may be null).  With samples (the
// weighted update of a DeviceImage frame) the integration kernel reads the
// frame's packed per-pixel samples instead of nmap and rgb; depth is still
// the raw depth of block allocation.
int lf_update_ev(const lf_map_params *p, const lf_map_storage *s, const float *depth, const float *nmap,
                 const unsigned char *rgb, int cols, int rows, const float *pose24, const lf_intrinsics *K,
                 hipStream_t st, bool weighted, hipEvent_t e0, hipEvent_t e1, const float *samples = nullptr);
// lf_deintegrate_weighted, optionally from a frame's packed samples.
int lf_deintegrate_ev(const lf_map_params *p, const lf_map_storage *s, const float *depth, const float *nmap,
                      const unsigned char *rgb, int cols, int rows, const float *pose24, const lf_intrinsics *K,
                      hipStream_t st, const float *samples);
// Keyframe re-fusion of n keyframes (raw depth and packed samples of one
// size, float pose24 of each keyframe's old and new pose, device pointers).
int lf_refuse_ev(const lf_map_params *p, const lf_map_storage *s, const float *const *depth,
                 const float *const *samples, const float *const *old_pose24, const float *const *new_pose24, int n,
                 int cols, int rows, const lf_intrinsics *K, int colour, hipStream_t st);
// lf_select_visible without the count read-back.
int lf_select_visible_ev(const lf_map_params *p, const lf_map_storage *s, const float *pose24, const lf_intrinsics *K,
                         int cols, int rows, hipStream_t st);
// lf_raycast with events around the ray-marching kernel.
int lf_raycast_ev(const lf_map_params *p, const lf_map_storage *s, float *vmap, unsigned char *rgb_out, int cols,
                  int rows, const float *pose24, const lf_intrinsics *K, int check_colour, hipStream_t st,
                  hipEvent_t e0, hipEvent_t e1);

// Device buffers of a DeviceImage: per-level pointers (up to 8 levels).
struct LfImageLevels {
    float *depth_raw;                    // level 0, metres
    unsigned char *rgb;                  // level 0 colour copy
    float *depth[8], *intensity[8], *dx[8], *dy[8];
    float *vmap[8], *nmap[8];            // float4 per pixel
    float *samples;                      // level 0 integration samples, float4 per pixel (may be null)
    int half_maps;                       // intensity / dx / dy are FP16, nmap 4 x FP16 (vmap stays float4)
};

// DeviceImage::upload (device_image.cpp:146-181) in levels + 1 launches.  Every
// level must be exactly half of the one above (K[l] of each level).
int lf_pyramid_upload(const uint16_t *depth_u16, const unsigned char *rgb, float depth_scale, int levels,
                    const lf_intrinsics *K, const LfImageLevels &m, hipStream_t st);
// DeviceImage::resize_device_map (device_image.cpp:188-199): vertex pyramid of
// vmap[0] and the normals of every level, in (levels + 1) / 2 launches.
int lf_pyramid_refresh(int levels, const lf_intrinsics *K, const LfImageLevels &m, hipStream_t st);

// The same passes planned instead of launched, for a batch of images: each
// pass is an lf_pass_bytes() device descriptor appended to out, and pass p of
// n images runs as one launch (lf_pass_batch over n consecutive descriptors in
// device memory; max_wg = the largest lf_pass_workgroups of them).  An image
// with fewer passes fills its place with lf_pass_skip.
size_t lf_pass_bytes(void);
int lf_pyramid_upload_plan(const uint16_t *depth_u16, const unsigned char *rgb, float depth_scale, int levels,
                           const lf_intrinsics *K, const LfImageLevels &m, std::vector<unsigned char> &out);
int lf_pyramid_refresh_plan(int levels, const lf_intrinsics *K, const LfImageLevels &m,
                            std::vector<unsigned char> &out);
int lf_pass_workgroups(const void *pass);
void lf_pass_skip(void *pass);
int lf_pass_batch(const void *passes, int n, int max_wg, hipStream_t st);

// lf_render_scene over float4 maps, or a 4 x FP16 normal map (half_maps).
int lf_render_scene_maps(const void *vmap, const void *nmap, int half_maps, const unsigned char *rgb, int cols,
                         int rows, unsigned char *rgba_out, hipStream_t st);

// One map's weighted update in a batch (the System frame's form: raw depth for
// block allocation, packed samples for integration, colour or null); a map
// with enabled = 0 is left alone.  lf_update_batch_make writes n descriptors
// of lf_fuse_job_bytes() each to jobs_host with the launch dimensions;
// lf_update_batch_launch runs block allocation, the visible lists and the
// integration of every map as three launches (y index = map) from a device
// copy of the descriptors.  Each map's volume equals lf_update_ev's.
struct LfFuseItem {
    const lf_map_params *p;
    const lf_map_storage *s;
    const float *depth;
    const float *samples;
    const unsigned char *rgb;
    const float *pose24;
    const lf_intrinsics *K;
    int cols, rows, enabled;
};
struct LfFuseBatchDims {
    int n, any, mixed, tiles_max, vis_grid_max, integrate_grid;
};
size_t lf_fuse_job_bytes(void);
int lf_update_batch_make(const LfFuseItem *items, int n, void *jobs_host, LfFuseBatchDims *dims);
// e0 / e1 (either may be null) are recorded immediately around the integration launch.
int lf_update_batch_launch(const void *jobs_dev, const LfFuseBatchDims &d, hipStream_t st, hipEvent_t e0 = nullptr,
                           hipEvent_t e1 = nullptr);

// One raycast of a batch: a map, its target vertex map (and colour, or null),
// the float T_wc / T_cw pose and the image size.
struct LfRayItem {
    const lf_map_params *p;
    const lf_map_storage *s;
    float *vmap;
    unsigned char *rgb_out;
    const float *pose24;
    const lf_intrinsics *K;
    int cols, rows;
};
// Bytes of one item's device descriptor; jobs_host (pinned) and jobs_dev
// (device) hold n of them.  The host copy is read by an asynchronous copy on
// st: it must not be rewritten before that copy has run.
size_t lf_ray_job_bytes(void);
// z-range and raycast of n maps in one launch per stage on st; each item's
// result equals lf_raycast_ev's.
int lf_raycast_batch_ev(const LfRayItem *items, int n, void *jobs_host, void *jobs_dev, int check_colour,
                        hipStream_t st);
// The same in two parts: the n descriptors written to jobs_host with the
// launch dimensions, then (once a copy of them is in device memory) the launches.
struct LfRayBatchDims {
    int n, mixed, tiles_max, zgrid_max, grid_max;
    size_t zrange_lds;
};
int lf_raycast_batch_make(const LfRayItem *items, int n, void *jobs_host, LfRayBatchDims *dims);
// e0 / e1 (either may be null) are recorded immediately around the ray-marching launch.
int lf_raycast_batch_launch(const void *jobs_dev, const LfRayBatchDims &d, int check_colour, hipStream_t st,
                            hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
