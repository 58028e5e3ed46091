/*
 * libfusion_amd — C-ABI of the MI355X (gfx950) dense-reconstruction hot path.
 *
 * Every entry point takes plain device pointers, sizes, POD parameter structs
 * and an explicit hipStream_t (passed as void* so this header needs no HIP
 * include), and returns an int status (LF_OK = 0).  Nothing here allocates
 * per call except the functions whose name ends in _create.
 *
 * The cut line is the set of free functions libfusion's host classes call
 * (reference: include/voxel_hashing/map_proc.h:14-102,
 * include/tracking/pose_estimator.h:13-39, include/tracking/m_estimator.h:11-24,
 * include/tracking/cuda_imgproc.h:11-39).  Each declaration below names the
 * reference function it replaces.
 *
 * Conventions shared by all entry points
 *  - Images are dense row-major (pitch == width * element size).
 *  - float4 maps (vmap / nmap) are 4 consecutive floats per pixel.
 *  - RGB images are 3 consecutive uint8 per pixel, channel 0 = R.
 *  - A device pose block ("pose24") is 24 floats: T_world_camera as a row-major
 *    3x4 matrix followed by T_camera_world (its inverse) as a row-major 3x4.
 *    The reference passes Sophus::SE3d by value and casts inside
 *    (depth_fusion.cu:462,471); here the caller supplies both float matrices
 *    on the device so a frame step never needs a host round trip.
 */
#ifndef LIBFUSION_AMD_H
#define LIBFUSION_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LF_OK 0
#define LF_ERR_ARG 1         /* bad argument (null pointer, bad size) */
#define LF_ERR_HIP 2         /* a HIP runtime call failed */
#define LF_ERR_CAPACITY 3    /* an explicit capacity was exceeded */
#define LF_ERR_NOT_READY 4   /* state missing (e.g. raycast before update) */

/* Voxel block edge (voxels) — reference BLOCK_SIZE (map_struct.h:9). */
#define LF_BLOCK 8
#define LF_BLOCK3 512

/* Overflow bits reported in counters[LF_CNT_OVERFLOW]. */
#define LF_OVF_TABLE 1u      /* no free hash slot on a probe path */
#define LF_OVF_POOL 2u       /* voxel-block pool exhausted */
#define LF_OVF_TRIANGLES 4u  /* mesh triangle capacity exceeded */

/* Indices into the device counter array of a map. */
#define LF_CNT_ALLOCATED 0   /* blocks taken from the pool (may exceed pool on overflow) */
#define LF_CNT_VISIBLE 1     /* visible blocks after the last update */
#define LF_CNT_OVERFLOW 2    /* LF_OVF_* bits, sticky until reset */
#define LF_CNT_TRIANGLES 3   /* triangles requested by the last mesh extraction */
#define LF_CNT_VIS_SUM_LO 4  /* running sum of visible counts (stats), low word */
#define LF_CNT_VIS_SUM_HI 5  /* high word */
#define LF_CNT_UPDATES 6     /* number of updates since reset */
#define LF_CNT_RENDER_BLOCKS 7 /* visible blocks with a non-empty tile rectangle, last raycast */
#define LF_CNT_FUSE_PASSES 8 /* block-selection passes (updates and de-integrations) since reset:
                                the per-pass seed of a mixed map's stochastic weight rounding */
#define LF_NUM_COUNTERS 16

/* Camera intrinsics — mirrors fusion::IntrinsicMatrix (intrinsic_matrix.h:12-37).
 * invfx/invfy must equal 1.0f/fx and 1.0f/fy (lf_intrinsics_make fills them). */
typedef struct lf_intrinsics {
    int width, height;
    float fx, fy, cx, cy, invfx, invfy;
} lf_intrinsics;

/* Map configuration — mirrors fusion::MapState (map_struct.h:17-51) with the
 * truncation distance made a parameter (the reference hard-codes 5 voxels,
 * map_struct.cu:120-123). */
typedef struct lf_map_params {
    int num_buckets;          /* home buckets of computeHash (map_struct.cu:143-148) */
    int num_entries;          /* total hash slots (buckets + excess), >= num_buckets */
    int num_blocks;           /* voxel-block pool capacity */
    int max_triangles;        /* mesh capacity, reference num_max_mesh_triangles_ */
    float zmin_raycast, zmax_raycast;
    float zmin_update, zmax_update;
    float voxel_size;         /* metres */
    float truncation;         /* metres */
    int precision;            /* LF_PRECISION_FP32 or LF_PRECISION_MIXED */
} lf_map_params;

/* Storage precision of a map and of the images and tracking of a System.
 *   FP32  the reference's method: int16 TSDF, float32 weights, float32 vertex /
 *         normal / intensity / gradient maps.
 *   MIXED FP16 TSDF with 8-bit weight codes (lf_weight_code, stored with
 *         stochastic rounding: lf_weight_code_dither) packed with the colour
 *         into one word per voxel (6 bytes a voxel); FP16 normal,
 *         intensity and gradient maps (vertex maps stay float4: FP16's 2 mm
 *         step at 2-4 m biased the ICP fit); the same tracking arithmetic as
 *         FP32 (FP32 rows, their normal equations accumulated in FP64; FP64
 *         solve and pose updates).  Results are stated against the FP64
 *         method, not bit for bit. */
#define LF_PRECISION_FP32 0
#define LF_PRECISION_MIXED 1

/* Device storage of one voxel-hash map — the counterpart of fusion::MapStorage
 * (map_struct.h:69-77).  Layout (all arrays in HBM):
 *   keys[num_entries]        uint64 packed block coordinate, ~0 = empty slot
 *   slot_block[num_entries]  pool block index of the slot, -1 = none
 *   slot_flag[num_entries]   1 if the slot's block was touched by the current update
 *   block_pos[4*num_blocks]  {x, y, z, slot} of every pool block in allocation order
 *   sdf[512*num_blocks]      FP32 precision: int16 TSDF, value = sdf/32767
 *                            truncation units; -32768 marks a voxel whose weight is 0.
 *                            Mixed: FP16 TSDF in truncation units; 0xFFFF (a NaN)
 *                            marks a voxel whose weight is 0
 *   weight[512*num_blocks]   FP32 precision: float32 fusion weight, clamped to 255.
 *                            Mixed: one 32-bit word per voxel, the 8-bit weight
 *                            code (lf_weight_value; the weight clamped to 255
 *                            first) in bits 0-7 and the colour R, G, B in bits
 *                            8-15, 16-23, 24-31, so weight and colour are one access
 *   rgb[3*512*num_blocks]    FP32 precision: uint8 colour, planar per block:
 *                            R[512] G[512] B[512].  Mixed: unused (may be null)
 *   visible[num_blocks]      pool block indices visible in the last update
 *   counters[LF_NUM_COUNTERS]
 *   occupancy[LF_OCC_WORDS]  occupancy bits of 4x4x4-block regions: bit
 *                            lf_occupancy_bit(bx, by, bz) is set when a block of
 *                            the region is inserted (a clear bit proves every
 *                            block of the region absent; the raycast skips the
 *                            hash probe there)
 *   visible_pos[4*num_blocks] {x, y, z, pool block} of visible[i]: the list's
 *                            block coordinates beside their indices, so a
 *                            kernel walking the list needs one load per block
 *   block_dir[LF_DIR_CELLS]  pool block of every block coordinate with each axis in
 *                            [-LF_DIR_DIM/2, LF_DIR_DIM/2), -1 = none: a direct
 *                            directory beside the hash table (4x4x4-block bricks of
 *                            64 cells, cell lf_dir_cell(bx, by, bz)); the raycast
 *                            resolves blocks with one load, blocks outside it
 *                            through the hash table
 * Voxel i of block b lives at b*512 + z*64 + y*8 + x (map_struct.cu:657-660). */
#define LF_OCC_WORDS 4096      /* 2^17 occupancy bits */
#define LF_DIR_DIM 256         /* block directory: 256^3 cells (64 MB), 10.24 m per axis at 5 mm voxels */
#define LF_DIR_CELLS (LF_DIR_DIM * LF_DIR_DIM * LF_DIR_DIM)

typedef struct lf_map_storage {
    unsigned long long *keys;
    int *slot_block;
    unsigned char *slot_flag;
    int *block_pos;
    short *sdf;
    float *weight;
    unsigned char *rgb;
    int *visible;
    unsigned int *counters;
    float *zrange;            /* raycast z-range tiles, 2 * tiles_x * tiles_y floats */
    unsigned int *occupancy;  /* LF_OCC_WORDS words */
    int *visible_pos;         /* 4 * num_blocks ints */
    int *block_dir;           /* LF_DIR_CELLS ints */
    int tiles_x, tiles_y;
} lf_map_storage;

/* ---------------------------------------------------------------- general */

/* Occupancy bit of block (bx, by, bz): its 4x4x4-block region (floor
 * division) (rx, ry, rz) wrapped onto a 64 x 64 x 32 torus,
 * (rz mod 32) * 4096 + (ry mod 64) * 64 + (rx mod 64). */
int lf_occupancy_bit(int bx, int by, int bz);
/* Directory cell of block (bx, by, bz), or -1 outside the directory. */
int lf_dir_cell(int bx, int by, int bz);
/* 8-bit weight code of a mixed-precision map: an unsigned float with 4
 * exponent bits (bias 7) and 4 mantissa bits; code 0 is weight 0, codes
 * 1..15 are m * 2^-10, code e<<4|m (e >= 1) is (1 + m/16) * 2^(e-7), up to 496.
 * lf_weight_code rounds a weight >= 0 to the nearest code (ties to the even
 * code); weights above 496 give 0xFF. */
int lf_weight_code(float w);
float lf_weight_value(int code);
/* The code of w >= 0 rounded stochastically with the random bits rnd: the
 * next code up with probability equal to w's fractional position between its
 * two neighbouring codes (rnd's low 19 bits dither weights >= 2^-6, its top 24
 * bits smaller ones), so the code's value equals w in expectation.  This is
 * how a mixed map stores fused and de-integrated weights, with rnd a hash of
 * the voxel and the map's fusion pass (LF_CNT_FUSE_PASSES): round-to-nearest
 * would stall a running sum of small increments at a fraction of its value. */
int lf_weight_code_dither(float w, unsigned int rnd);

/* Library version string and the gfx target the kernels were built for. */
const char *lf_version(void);
/* Fill an lf_intrinsics the way IntrinsicMatrix's constructor does
 * (intrinsic_matrix.h:27-31). */
void lf_intrinsics_make(int width, int height, float fx, float fy, float cx, float cy,
                        lf_intrinsics *out);
/* IntrinsicMatrix::pyr_down (intrinsic_matrix.h:33-37). */
void lf_intrinsics_pyr_down(const lf_intrinsics *in, lf_intrinsics *out);
/* computeHash (map_struct.cu:143-148): ((x*73856093)^(y*19349669)^(z*83492791)) mod n,
 * int32 wrap-around, result in [0, n). */
int lf_compute_hash(int x, int y, int z, int num_buckets);
/* Copy the marching-cubes tables compiled into the library
 * (edge[256], tri[256*16] int32) — reference device_tables.h:9-280. */
void lf_mc_tables(int *edge_out, int *tri_out);
/* Number of HIP devices visible (0 on a machine without a GPU). */
int lf_device_count(void);
/* Last error message of this thread ("" if none). */
const char *lf_last_error(void);

/* -------------------------------------------------------------- voxel map */

/* MapStruct<true>::create (map_struct.cu:391-414): allocate device storage for
 * params, z-range tiles for a (width x height) raycast. */
int lf_map_create(const lf_map_params *p, int width, int height, lf_map_storage *out);
/* MapStruct<true>::release (map_struct.cu:424-449). */
int lf_map_release(lf_map_storage *s);
/* MapStruct<true>::reset (map_struct.cu:74-93): empty table, empty pool,
 * zero voxels and counters. */
int lf_map_reset(const lf_map_params *p, const lf_map_storage *s, void *stream);
/* Copy the device counters to host memory (synchronises the stream). */
int lf_map_read_counters(const lf_map_storage *s, unsigned int *out16, void *stream);

/* fusion::cuda::update_weighted (depth_fusion.cu:502-575): block allocation along
 * the truncation band of every valid pixel (create_blocks_kernel, :121-228),
 * visible-block selection (check_visibility_flag_kernel, :49-83, plus the
 * compaction :85-93,549-557) and TSDF + colour fusion with weight
 * |sin(n_z)|/d^2 (update_map_weighted_kernel, :356-428).
 *   depth  float metres [rows*cols], nmap float4 [rows*cols], rgb uint8x3.
 *   pose24 device pose block of the frame.
 * The visible count stays on the device (counters[LF_CNT_VISIBLE]). */
int lf_update_weighted(const lf_map_params *p, const lf_map_storage *s,
                       const float *depth, const float *nmap, const unsigned char *rgb,
                       int cols, int rows, const float *pose24, const lf_intrinsics *K,
                       void *stream);
/* fusion::cuda::update (depth_fusion.cu:430-500): same allocation and visibility,
 * TSDF + colour fusion with weight 1/d^2 and no normal (update_map_with_colour_kernel,
 * :288-354). */
int lf_update(const lf_map_params *p, const lf_map_storage *s,
              const float *depth, const unsigned char *rgb,
              int cols, int rows, const float *pose24, const lf_intrinsics *K, void *stream);

/* fusion::cuda::create_rendering_blocks + raycast_with_colour (ray_tracing.cu:223-277,
 * 478-564,611-645) when check_colour != 0, or + raycast (ray_tracing.cu:346-425,
 * 577-609) when check_colour == 0.  Casts rays of a (cols x rows) camera at pose24
 * against the visible blocks of the last update, writing vmap float4 (camera
 * frame, NaN where nothing is hit) and, if rgb_out is non-null, the nearest
 * voxel colour (255 where nothing is hit). */
int lf_raycast(const lf_map_params *p, const lf_map_storage *s,
               float *vmap, unsigned char *rgb_out, int cols, int rows,
               const float *pose24, const lf_intrinsics *K, int check_colour, void *stream);

/* count_visible_entry (ray_tracing.cu:647-807): the map's visible list becomes
 * every allocated block with a corner inside a cols x rows image at the pose
 * (T_world_camera, then its inverse) with z in [zmin_update, zmax_update]; the
 * count is returned when count_out is non-null (synchronising).  The
 * reference's functor never initialises its entry count (ray_tracing.cu:733);
 * this is the intended semantic over allocated blocks. */
int lf_select_visible(const lf_map_params *p, const lf_map_storage *s, const float *pose24,
                      const lf_intrinsics *K, int cols, int rows, unsigned int *count_out, void *stream);
/* raycast_check_visibility (voxel_hashing.cpp:121-133): lf_select_visible at the
 * pose, then lf_raycast from it — renders the map from any pose, not only the
 * last update's (relocalisation). */
int lf_raycast_check_visibility(const lf_map_params *p, const lf_map_storage *s, float *vmap,
                                unsigned char *rgb_out, int cols, int rows, const float *pose24,
                                const lf_intrinsics *K, int check_colour, void *stream);

/* fusion::cuda::create_mesh_vertex_only / _with_normal / _with_colour
 * (scene_mesh.cu:260-301,308-351,600-643): marching cubes over every allocated
 * block.  vertex_out: 9 floats per triangle (metres); normal_out: 9 floats per
 * triangle (flat normal repeated), may be null; colour_out: 9 uint8 per triangle,
 * may be null.  Triangle order is unspecified (the reference appends with an
 * atomic counter).  *triangles_out receives min(count, max_triangles); the
 * return is LF_ERR_CAPACITY if triangles were dropped. */
int lf_create_mesh(const lf_map_params *p, const lf_map_storage *s,
                   float *vertex_out, float *normal_out, unsigned char *colour_out,
                   int max_triangles, unsigned int *triangles_out, void *stream);

/* ------------------------------------------------- checkpoint / transfer */

/* MapStruct::writeToDisk via DenseMapping::writeMapToDisk (map_struct.cu:529-566;
 * voxel_hashing.cpp:189-196): the reference's binary layout (12-byte voxels of
 * the whole pool, 20-byte chained hash entries, heap, bucket mutexes, heap and
 * excess counters) sized by the map's capacities, plus "<path>.txt" holding
 * {num_blocks, num_hash_entries, num_buckets}.  LF_ERR_CAPACITY when the
 * reference table's excess area cannot hold the map's bucket collisions. */
int lf_map_write(const lf_map_params *p, const lf_map_storage *s, const char *path, void *stream);
/* MapStruct::readFromDisk via DenseMapping::readMapFromDisk (map_struct.cu:568-613;
 * voxel_hashing.cpp:198-205): resets the map and loads every live entry of a
 * reference-layout file (sizes from "<path>.txt", else the map's capacities),
 * compacted in pool order.  LF_ERR_CAPACITY when the pool or table is too small. */
int lf_map_read(const lf_map_params *p, const lf_map_storage *s, const char *path, void *stream);
/* MapStruct::copyTo (map_struct.cu:451-469): src into dst (same parameters). */
int lf_map_copy(const lf_map_params *p, const lf_map_storage *src, const lf_map_storage *dst, void *stream);

/* --------------------------------------------------------- image pipeline */

/* Depth conversion of System::process_images (system.cpp:43) with the scale as a
 * parameter: out = in * scale. */
int lf_depth_to_float(const uint16_t *in, float *out, int cols, int rows, float scale,
                      void *stream);
/* filterDepthBilateral (cuda_imgproc.cu:197-200): cv::cuda::bilateralFilter with
 * ksize 5, sigma_colour 1, sigma_space 1 on metres. */
int lf_bilateral_filter(const float *in, float *out, int cols, int rows, void *stream);
/* pyrDownDepth (cuda_imgproc.cu:202-205): resize x0.5, linear, 2x2 mean. */
int lf_pyr_down_depth(const float *in, float *out, int cols, int rows, void *stream);
/* pyrDownVMap (cuda_imgproc.cu:212-215): same on float4 maps. */
int lf_pyr_down_vmap(const float *in, float *out, int cols, int rows, void *stream);
/* pyrDownImage (cuda_imgproc.cu:207-210): 5x5 Gaussian, keep even pixels,
 * output ((cols+1)/2, (rows+1)/2). */
int lf_pyr_down_image(const float *in, float *out, int cols, int rows, void *stream);
/* cvtColor(RGB2GRAY) on the float image (device_image.cpp:159-162). */
int lf_rgb_to_intensity(const unsigned char *rgb, float *out, int cols, int rows,
                        void *stream);
/* computeDerivative (cuda_imgproc.cu:217-247); last row/column written as 0. */
int lf_compute_derivative(const float *image, float *dx, float *dy, int cols, int rows,
                          void *stream);
/* backProjectDepth (cuda_imgproc.cu:249-274). */
int lf_back_project_depth(const float *depth, float *vmap, int cols, int rows,
                          const lf_intrinsics *K, void *stream);
/* computeNMap (cuda_imgproc.cu:276-305); last row/column written as NaN. */
int lf_compute_nmap(const float *vmap, float *nmap, int cols, int rows, void *stream);

/* renderScene / renderSceneTextured (cuda_imgproc.cu:17-108): RGBA8 Phong
 * shading of a camera-frame vertex / normal map (float4 each) with the light
 * at (5, 5, 5); rgb (uint8 x3) textures it, null renders grey.  NaN vertices
 * get (4, 2, 2, 255). */
int lf_render_scene(const float *vmap, const float *nmap, const unsigned char *rgb, int cols, int rows,
                    unsigned char *rgba_out, void *stream);
/* ToSemiDenseImageK (cuda_imgproc.cu:110-128): image where dx > th_dx or
 * dy > th_dy, else 255. */
int lf_semi_dense(const float *image, const float *dx, const float *dy, int cols, int rows, float th_dx,
                  float th_dy, float *semi_out, void *stream);
/* warp_image (cuda_imgproc.cu:155-195): dst(x) = src bilinearly sampled where
 * vmap_dst(x) lands under pose12 (float 3x4, host) and K, inside a one-pixel
 * margin; 0 elsewhere.  src / dst uint8 x3. */
int lf_warp_image(const unsigned char *src, const float *vmap_dst, int cols, int rows, const float *pose12,
                  const lf_intrinsics *K, unsigned char *dst, void *stream);

/* ------------------------------------------- frame pyramid (DeviceImage) */

/* DeviceImage (device_image.h:18-88): one RGB-D frame's device pyramid.  Every
 * level is exactly half of the one above (even sizes down to the last level).
 * The images of a mixed-precision System store INTENSITY, DX and DY as FP16
 * and NMAP as 4 x FP16; every other plane is as listed. */
typedef struct lf_image lf_image;

enum {
    LF_PLANE_DEPTH = 0,      /* float, metres, bilateral-filtered (get_depth)      */
    LF_PLANE_INTENSITY = 1,  /* float grey (get_intensity)                         */
    LF_PLANE_DX = 2,         /* float (get_intensity_dx)                           */
    LF_PLANE_DY = 3,         /* float (get_intensity_dy)                           */
    LF_PLANE_VMAP = 4,       /* float4 vertices (get_vmap)                         */
    LF_PLANE_NMAP = 5,       /* float4 normals (get_nmap)                          */
    LF_PLANE_RAW_DEPTH = 6,  /* float, metres, level 0 only (get_raw_depth)        */
    LF_PLANE_RGB = 7,        /* uint8 RGB copy of the input, level 0 (get_image)   */
    LF_PLANE_SAMPLES = 8     /* float4 integration samples, level 0: {raw depth, or
                                NaN where the normal is invalid; |sin(n_z)|/d^2;
                                colour r | g<<8 | b<<16 as bits; 0}              */
};

/* DeviceImage(K, NUM_PYRS) (device_image.cpp:8-21); work is queued on stream. */
int lf_image_create(const lf_intrinsics *K, int num_pyr, void *stream, lf_image **out);
int lf_image_destroy(lf_image *img);
/* upload (device_image.cpp:146-181) from device-resident depth (uint16) and
 * RGB (uint8 x3): depth to metres, bilateral filter, intensity, the depth and
 * intensity pyramids, and every level's derivatives, vertex and normal maps. */
int lf_image_upload(lf_image *img, const uint16_t *depth, const unsigned char *rgb, float depth_scale);
/* resize_device_map (device_image.cpp:188-199): vertex pyramid of the level-0
 * vertex map (e.g. written by lf_raycast) and the normals of every level. */
int lf_image_resize_device_map(lf_image *img);
/* Device pointer of a plane (LF_PLANE_*) of a level; rows are packed. */
int lf_image_plane(lf_image *img, int level, int plane, void **out);
/* get_rendered_image / get_rendered_scene_textured (device_image.cpp:119-129):
 * render the level-0 vertex / normal maps (textured with the frame's colour
 * when textured != 0) into the image's RGBA8 plane; *rgba_out receives it. */
int lf_image_render(lf_image *img, int textured, void **rgba_out);
/* De-integration (not in the reference; deleteBlock, map_struct.cu:327-363, has
 * no callers): the inverse of lf_update_weighted for the same frame and pose,
 * over the same voxels (no block is allocated).  weight' = weight - w_frame,
 * sdf and colour un-averaged; a weight that reaches 0 returns the voxel to
 * unseen.  Exact up to int16 sdf / uint8 colour rounding except where the
 * weight saturated at 255. */
int lf_deintegrate_weighted(const lf_map_params *p, const lf_map_storage *s, const float *depth,
                            const float *nmap, const unsigned char *rgb, int cols, int rows,
                            const float *pose24, const lf_intrinsics *K, void *stream);
/* De-integration of a DeviceImage frame (the inverse of lf_update_image). */
int lf_deintegrate_image(const lf_map_params *p, const lf_map_storage *s, lf_image *img, const float *pose24,
                         int colour, void *stream);
/* Keyframe re-fusion: n keyframes (DeviceImage frames of one size and
 * intrinsics) leave the map at their old poses and are fused at their new
 * (corrected) poses, in one batched pass of three launches.  No single
 * reference counterpart: the reference re-tracks keyframes against the model
 * (keyframe_graph.cpp:162-229) and has a deleteBlock with no callers
 * (map_struct.cu:327-363); this is lf_deintegrate_image of keyframes 0..n-1 at
 * old_pose24[k] followed by lf_update_image of keyframes 0..n-1 at
 * new_pose24[k], replayed per voxel in registers with one store (no
 * re-quantisation between the steps).  old_pose24[k] / new_pose24[k]: device
 * pointers to float T_wc, T_cw (24 floats).  1 <= n <= LF_REFUSE_MAX_KEYFRAMES. */
#define LF_REFUSE_MAX_KEYFRAMES 31
int lf_refuse_keyframes(const lf_map_params *p, const lf_map_storage *s, lf_image *const *frames,
                        const float *const *old_pose24, const float *const *new_pose24, int n, int colour,
                        void *stream);
/* update(RgbdImagePtr frame) (voxel_hashing.cpp:44-65) into a caller's map:
 * lf_update_weighted on the frame's raw depth, level-0 normals and (colour != 0)
 * colour, read through the frame's packed integration samples.  Same result
 * as lf_update_weighted on those planes. */
int lf_update_image(const lf_map_params *p, const lf_map_storage *s, lf_image *img, const float *pose24,
                    int colour, void *stream);

/* ------------------------------------------------------------- tracking */

/* DenseTracking(K, NUM_PYR) (icp_tracker.h:28-88): a tracker that owns source
 * (current) and reference pyramids, for re-tracking keyframes against
 * raycast or stored maps (keyframe_graph.cpp:162-229). */
typedef struct lf_tracker lf_tracker;

/* TrackingResult (icp_tracker.h:13-18): update = estimate^-1, 3x4 row-major. */
typedef struct lf_tracking_result {
    int success;
    float icp_error;
    double update[12];
} lf_tracking_result;

int lf_tracker_create(const lf_intrinsics *K, int num_pyr, void *stream, lf_tracker **out);
int lf_tracker_destroy(lf_tracker *t);
/* set_source_* / set_reference_* (icp_tracker.cpp:164-208): level-0 device
 * planes; vmap float4 (pyrDownVMap pyramid and normals, build_pyramid.cpp:55-76),
 * image RGB u8 (grey, pyrDown pyramid, derivatives), depth float (bilateral
 * and depth pyramid), intensity float. */
int lf_tracker_set_source_vmap(lf_tracker *t, const float *vmap);
int lf_tracker_set_source_image(lf_tracker *t, const unsigned char *rgb);
int lf_tracker_set_source_depth(lf_tracker *t, const float *depth);
int lf_tracker_set_source_intensity(lf_tracker *t, const float *intensity);
int lf_tracker_set_reference_image(lf_tracker *t, const unsigned char *rgb);
int lf_tracker_set_reference_intensity(lf_tracker *t, const float *intensity);
int lf_tracker_set_reference_vmap(lf_tracker *t, const float *vmap);
/* swap_intensity_pyr (icp_tracker.cpp:158-162). */
int lf_tracker_swap_intensity(lf_tracker *t);
/* compute_transform(ctx) (icp_tracker.cpp:210-314) on the set_* pyramids:
 * iterations[level] Gauss-Newton steps per level (num_levels <= NUM_PYR),
 * initial estimate 3x4 (null: identity).  Synchronising.  As in the
 * reference, success is always 1 and icp_error is last_icp_error. */
int lf_tracker_compute_transform(lf_tracker *t, const double *initial_estimate12, const int *iterations,
                                 int num_levels, lf_tracking_result *out);
/* compute_transform(reference, current, ctx) (icp_tracker.cpp:23-156) on two
 * frames: success requires finite errors and |log(T^-1 T0)| <= 0.1.
 * Synchronising. */
int lf_tracker_compute_transform_images(lf_tracker *t, lf_image *reference, lf_image *current,
                                        const double *initial_estimate12, const int *iterations, int num_levels,
                                        lf_tracking_result *out);
/* A plane (LF_PLANE_*) of the tracker's source (source != 0) or reference pyramid. */
int lf_tracker_plane(lf_tracker *t, int source, int level, int plane, void **out);
/* The normal-equation sums of the tracker's last Gauss-Newton step.  No
 * reference counterpart: the reference sums JtJ / Jtr per term in FP32 and
 * combines them on the host (icp_tracker.cpp:50-96); here they are combined
 * per pixel and accumulated in FP64 on the device, and this read exposes them
 * for checking.  sums31[0..26]: A = w H_icp + H_rgb and b in create_jtjjtr<6,7>
 * order (reduce_sum.h:48-61), w = 1e6 (icp_tracker.cpp:89-90); [27] / [28]:
 * ICP sum r^2 / count; [29] / [30]: RGB sum r^2 / count.  Synchronising. */
int lf_tracker_get_sums(lf_tracker *t, double *sums31);

/* Scratch for the reductions: lf_reduce_scratch_floats() floats of device memory. */
size_t lf_reduce_scratch_floats(void);
/* icp_reduce (pose_estimator.cu:267-307): point-to-plane ICP normal equations of
 * the reference frame (last_*) against the current frame (curr_*) under the
 * float 3x4 pose (12 floats, device) mapping reference-camera points into the
 * current camera.  Writes 29 floats to sum29 (device): the 28 upper-triangle
 * products of [J|r] followed by the correspondence count. */
int lf_icp_reduce(const float *curr_vmap, const float *curr_nmap,
                  const float *last_vmap, const float *last_nmap,
                  int cols, int rows, const float *pose12, const lf_intrinsics *K,
                  float *scratch, float *sum29, void *stream);
/* lf_icp_reduce without the scratch argument, allocating nothing: no
 * counterpart in the reference; for callers that cannot allocate (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Its
 * partial rows go to a fixed device array of 512 x 58 floats (116 KiB) that
 * the library holds on each GPU it runs on until the process ends; any image
 * size fits (the reduction's grid is capped at 512 workgroups).  All calls on
 * one GPU share that array: calls that overlap on one GPU (other streams or
 * threads) race, so serialise them.  Same result as lf_icp_reduce. */
int lf_icp_reduce_noalloc(const float *curr_vmap, const float *curr_nmap,
                          const float *last_vmap, const float *last_nmap,
                          int cols, int rows, const float *pose12, const lf_intrinsics *K,
                          float *sum29, void *stream);
/* rgb_step (m_estimator.cu:437-482): Huber-weighted photometric normal equations
 * with stddev_estimate; same output layout as lf_icp_reduce. */
int lf_rgb_step(const float *curr_intensity, const float *last_intensity,
                const float *last_vmap, const float *intensity_dx, const float *intensity_dy,
                int cols, int rows, float stddev_estimate, const float *pose12,
                const lf_intrinsics *K, float *scratch, float *sum29, void *stream);
/* create_jtjjtr<6,7> (reduce_sum.h:48-61) plus residual[2] = {sum r^2, count}:
 * unpack 29 sums into a symmetric 6x6 (row-major == column-major) and a 6-vector. */
void lf_unpack_jtjjtr(const float *sum29, float *jtj36, float *jtr6, float *residual2);

/* ------------------------------------------------- pipeline (System facade) */

typedef struct lf_system lf_system;

typedef struct lf_system_config {
    lf_intrinsics K;          /* level-0 intrinsics */
    lf_map_params map;
    int num_pyr;              /* pyramid levels (reference System(K, NUM_PYR)) */
    int iterations[8];        /* Gauss-Newton iterations per level, [0] = finest */
    float depth_scale;        /* metres per raw depth unit */
    int integrate_colour;     /* fuse RGB into the volume */
} lf_system_config;

/* System::System (system.cpp:12-29). */
int lf_system_create(const lf_system_config *cfg, lf_system **out);
int lf_system_destroy(lf_system *sys);
/* System::process_images (system.cpp:40-114) on device-resident input:
 * depth uint16 [rows*cols] and rgb uint8x3.  Enqueues the whole frame on the
 * system's stream without host synchronisation. */
int lf_system_process_images(lf_system *sys, const uint16_t *depth, const unsigned char *rgb);
/* System::process_images(const cv::Mat depth, const cv::Mat image)
 * (system.cpp:40-114) on host-resident input, as the reference takes it:
 * depth uint16 and rgb uint8x3 in host memory are copied to device staging on
 * the system's stream, then the frame runs as lf_system_process_images.  With
 * pinned host buffers (hipHostMalloc) the copies are asynchronous and the host
 * buffers must stay unchanged until the frame has been synchronised. */
int lf_system_process_images_host(lf_system *sys, const uint16_t *depth, const unsigned char *rgb);
/* Wait for all enqueued frames. */
int lf_system_synchronize(lf_system *sys);
/* Camera pose of the last processed frame, row-major 4x4 doubles (synchronises). */
int lf_system_get_pose(lf_system *sys, double *out16);
/* Tracking result flags of the last frame: success, icp correspondences. */
int lf_system_get_tracking(lf_system *sys, int *success, float *icp_error);
/* System::restart (system.cpp:148-157). */
int lf_system_restart(lf_system *sys);
/* System::fetch_mesh_* (system.cpp:168-181); normal_out / colour_out may be null. */
int lf_system_fetch_mesh(lf_system *sys, float *vertex_out, float *normal_out,
                         unsigned char *colour_out, int max_triangles,
                         unsigned int *triangles_out);
/* Access to the system's map (for inspection and checkpointing). */
int lf_system_map(lf_system *sys, lf_map_params *p, lf_map_storage *s);
/* Device pointers of the reference frame's level-l vmap / nmap after the frame. */
int lf_system_maps(lf_system *sys, int level, const float **vmap, const float **nmap);
/* get_shaded_depth (which 0: the odometry's current frame), get_rendered_scene
 * (1: the reference frame, i.e. the last raycast) and get_rendered_scene_textured
 * (2) of System (system.cpp:132-146): *rgba_out receives the RGBA8 image
 * (device); synchronising. */
int lf_system_render(lf_system *sys, int which, void **rgba_out);
/* System::writeMapToDisk / readMapFromDisk (system.cpp:202-210). */
int lf_system_write_map(lf_system *sys, const char *path);
int lf_system_read_map(lf_system *sys, const char *path);
/* The system's HIP stream (as void*). */
void *lf_system_stream(lf_system *sys);
/* Event timing of the pipeline stages: stages is a bit mask, bit s brackets
 * stage s with HIP events on the system's stream (0 integrate kernel,
 * 1 raycast kernel, 2 tracking, 4 whole frame); 0 turns timing off.  Each
 * event record costs a few microseconds of stream time, so time only the
 * stages you read.  read_timing returns {launches, total_ms} of a stage. */
#define LF_TIMING_INTEGRATE 0x1
#define LF_TIMING_RAYCAST 0x2
#define LF_TIMING_TRACKING 0x4
#define LF_TIMING_FRAME 0x10
int lf_system_set_timing(lf_system *sys, int stages);
/* Replay steady-state frames from HIP graphs (enable != 0): the first tracked
 * frame of each of the two frame-buffer parities is captured (its whole device
 * work; the inputs are first copied into fixed buffers) and later frames
 * replay it.  Results are those of the eager path; frames with timing on, the
 * first frame and frames with mapping off run eagerly. */
int lf_system_set_graphs(lf_system *sys, int enable);
int lf_system_read_timing(lf_system *sys, int stage, int *launches, double *total_ms);

/* ---------------------------------------------------------------- batch */

/* Several independent systems stepped together (no counterpart in the
 * reference, which runs one System per process; for callers that serve many
 * RGB-D streams on one GPU).  System i is made from cfgs[i] (sizes and voxel
 * sizes may differ).  lf_batch_process takes one frame per system (device
 * pointers, uint16 depth and uint8 RGB of system i's size): pass p of every
 * system's image pyramid runs in one launch, Gauss-Newton step s of every
 * system in one launch pair, each map is fused on its own stream, then one
 * batched z-range + raycast launch per stage covers every map, then the
 * re-pyramids.  All systems share one precision (LF_ERR_ARG otherwise).  The
 * tracking reductions of n systems take 1/n of a lone system's grid each, so
 * each system's poses, tracking results and volume equal those of the same
 * system run alone up to FP64 summation order in the tracking sums (exactly
 * when n is 1).
 * Asynchronous; lf_batch_synchronize waits for all. */
typedef struct lf_batch lf_batch;
int lf_batch_create(const lf_system_config *cfgs, int n, lf_batch **out);
int lf_batch_destroy(lf_batch *b);
int lf_batch_process(lf_batch *b, const uint16_t *const *depth, const unsigned char *const *rgb);
/* lf_batch_process from host memory (System::process_images, system.cpp:40-114,
 * takes host images): each system's frame is copied into the batch's device
 * staging on the upload stream, which overlaps the previous step's fusion and
 * raycast.  With pinned host buffers (hipHostMalloc) the copies are
 * asynchronous; the host buffers must stay unchanged until the step has been
 * synchronised. */
int lf_batch_process_host(lf_batch *b, const uint16_t *const *depth, const unsigned char *const *rgb);
int lf_batch_synchronize(lf_batch *b);
/* Event timing of a batch's steps on its batch stream: stages is a bit mask,
 * bit s brackets stage s (0 the batched integration kernel, 1 the batched
 * ray-marching kernel, 2 the Gauss-Newton launches of a step, 3 block
 * allocation + visible lists + integration, 4 the batch stream's whole step;
 * the uploads overlap on another stream and are not in stage 4); 0 turns
 * timing off.  read_timing returns {launches, total_ms} of a stage
 * (synchronising). */
#define LF_BATCH_TIMING_INTEGRATE 0x1
#define LF_BATCH_TIMING_RAYCAST 0x2
#define LF_BATCH_TIMING_TRACKING 0x4
#define LF_BATCH_TIMING_FUSION 0x8
#define LF_BATCH_TIMING_STEP 0x10
int lf_batch_set_timing(lf_batch *b, int stages);
int lf_batch_read_timing(lf_batch *b, int stage, int *launches, double *total_ms);
/* Launch options of a batch (no reference counterpart: they shape this
 * build's batched launches).  Each system's poses, volume and maps stay those
 * of the default shapes up to FP64 summation order in the tracking sums.
 *   LF_BATCH_OPT_UPLOAD_OVERLAP: 1 (default) a step's uploads run on the
 *     upload stream beside the previous step's fusion and raycast; 0 on the
 *     batch stream ahead of the step's tracking.
 *   LF_BATCH_OPT_TRACK_WORKGROUPS: workgroups per system of each Gauss-Newton
 *     step, 1..512 (0, the default: 512 / n, at most the level's pixels / 512).
 *   LF_BATCH_OPT_INTEGRATE_WORKGROUPS: workgroups per map of the batched
 *     integration, 1..65535 (0, the default: max(2048, 8192 / n)).
 *   LF_BATCH_OPT_ZRANGE_WORKGROUPS: workgroups per map of the batched z-range
 *     fill, 1..65535 (0, the default: 1/n of a lone map's, at least 64). */
#define LF_BATCH_OPT_UPLOAD_OVERLAP 0
#define LF_BATCH_OPT_TRACK_WORKGROUPS 1
#define LF_BATCH_OPT_INTEGRATE_WORKGROUPS 2
#define LF_BATCH_OPT_ZRANGE_WORKGROUPS 3
int lf_batch_set_option(lf_batch *b, int option, int value);
int lf_batch_get_option(lf_batch *b, int option, int *value);
/* Replay the batch's steady-state steps from HIP graphs (enable != 0; no
 * reference counterpart): a step's table copy and uploads, its Gauss-Newton
 * launches and its fusion / raycast / re-pyramid launches are captured as
 * three graphs per table buffer and launch shape on first use and replayed
 * after; the events joining the upload and batch streams stay outside them.
 * Results are those of the eager steps; steps with timing on, or in which a
 * system starts a frame, run eagerly. */
int lf_batch_set_graphs(lf_batch *b, int enable);
/* Online autotuning of the launch options above (no reference counterpart).
 * The batch's next steps try each option's candidate values in turn, one
 * option at a time (integration workgroups {default, 512, 1024, 4096},
 * tracking workgroups {default, 16, 48, 64}, z-range workgroups {default,
 * 32, 128, 256}, upload overlap {1, 0}); each candidate runs one settling step
 * and steps_per_trial steps timed by events on the batch stream, in `rounds`
 * interleaved rounds (odd rounds in reverse order, so a steady drift of the
 * step time cancels), and the fastest median is kept before the next option
 * when it beats the default's by more than 1% (one synchronisation per
 * option).  The steps are the caller's own frames:
 * only launch shapes change (results up to FP64 summation order in the
 * tracking sums).  The choice is cached per batch shape (every system's image
 * size, levels, voxel size and precision) for the process, and appended to
 * the file set by lf_batch_tune_file; a batch of a cached shape starts with it.
 * steps_per_trial 1..64 (0 stops), rounds 1..16.  lf_batch_tuning: 1 while
 * tuning runs. */
int lf_batch_autotune(lf_batch *b, int steps_per_trial, int rounds);
int lf_batch_tuning(lf_batch *b);
/* Read tuned choices from path (lines "<shape> <values>", later lines win) and
 * append new ones to it; null or "" stops using a file. */
int lf_batch_tune_file(const char *path);
/* System i's camera pose (4x4 row-major doubles; synchronising), last
 * tracking result and map counters. */
int lf_batch_get_pose(lf_batch *b, int i, double *out16);
int lf_batch_get_tracking(lf_batch *b, int i, int *success, float *icp_error);
int lf_batch_read_counters(lf_batch *b, int i, unsigned int *out16);
/* System i's map (the arrays of its volume, as lf_system_map). */
int lf_batch_map(lf_batch *b, int i, lf_map_params *p, lf_map_storage *s);

#ifdef __cplusplus
}
#endif

#endif /* LIBFUSION_AMD_H */
