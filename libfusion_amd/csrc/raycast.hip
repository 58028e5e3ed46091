// Raycasting of the hashed TSDF volume.
//
// Reference: src/voxel_hashing/ray_tracing.cu — create_rendering_blocks
// (RenderingBlockDelegate :18-211, host :223-277), MapRenderingDelegate
// (:279-564: read_sdf :291-305, read_sdf_interped :307-339, operator() :346-425,
// raycast_with_colour :478-564), launchers :577-645.
//
// MI355X design:
//  * z-range tiles (one per 8x8 pixels) take the min / max depth of every
//    visible block's projected rectangle.  All depths are positive, so IEEE
//    ordering equals signed-integer ordering and plain 32-bit atomicMin/Max
//    replace the reference's CAS loops; the rectangle is filled directly (the
//    reference's 16x16 sub-rectangle list covers exactly the same tiles and only
//    bounds the work at 100k entries).
//  * One wavefront marches the 64 rays of one 8x8 tile: those rays share the
//    tile's z-range and neighbouring directions, so their loop trip counts and
//    the voxel blocks they touch are nearly the same (little divergence, good
//    L2 locality).
//  * Each lane caches the last block it resolved: consecutive samples step at
//    most a few voxels, so most samples and most trilinear taps skip the
//    lookup entirely.  A new block is tested against the occupancy bits in LDS,
//    then looked up in the dense block directory (one 4-byte load; the cells
//    of a 4x4x4-block brick share two cache lines, so a wavefront's lookups
//    mostly hit L2, where hash-table slots of neighbouring blocks are
//    scattered over 15 MB).
#include "lf_common.h"
#include "lf_internal.h"

using namespace lf;

namespace {

struct RayDev {
    const unsigned long long *keys;
    const int *slot_block;
    const short *sdf;
    const float *weight;
    const unsigned char *rgb;
    const int *visible;
    const int *visible_pos;  // {x, y, z, pool block} of visible[i]
    const int *block_dir;    // LF_DIR_CELLS pool blocks, -1 = none
    const int *block_pos;
    unsigned int *counters;
    const unsigned int *occupancy;  // LF_OCC_WORDS words
    float *zx, *zy;
    int tiles_x, tiles_y;
    int num_buckets, num_entries;
    float voxel_size, truncation, zmin_raycast, zmax_raycast;
    // 1 / voxel_size and raycast_step_scale() / 2 = truncation / voxel_size / 2,
    // by the same float operations on the host
    float inv_voxel_size, half_step_scale;
    int pool_fits_lds;  // pool indices fit the 24-bit field of the LDS cache
};

// A map read from a descriptor table, its pointers asserted global.
LF_DEV RayDev global_ray(const RayDev &d) {
    RayDev r = d;
    r.keys = global_ptr(d.keys); r.slot_block = global_ptr(d.slot_block); r.sdf = global_ptr(d.sdf);
    r.weight = global_ptr(d.weight); r.rgb = global_ptr(d.rgb); r.visible = global_ptr(d.visible);
    r.visible_pos = global_ptr(d.visible_pos); r.block_dir = global_ptr(d.block_dir);
    r.block_pos = global_ptr(d.block_pos); r.counters = global_ptr(d.counters); r.occupancy = global_ptr(d.occupancy);
    r.zx = global_ptr(d.zx); r.zy = global_ptr(d.zy);
    return r;
}

// One raycast's inputs and outputs: the map, the target image and the pose
// (single launches take it by value, batched launches read job blockIdx.y).
struct RayJob {
    RayDev r;
    float *vmap;
    unsigned char *rgb_out;
    const float *pose24;
    lf_intrinsics K;
    int cols, rows, wg_x, wg_y;
};

// Workgroup wg of nwg: the z-range tiles back to k_zrange_clear's values.
LF_DEV void zrange_clear_wg(const RayDev &r, int wg, int nwg) {
    const int n = r.tiles_x * r.tiles_y;
    if (wg == 0 && threadIdx.x == 0) r.counters[LF_CNT_RENDER_BLOCKS] = 0u;
    for (int i = wg * blockDim.x + threadIdx.x; i < n; i += nwg * blockDim.x) {
        r.zx[i] = 100.0f;
        r.zy[i] = 0.0f;
    }
}

__global__ void k_zrange_clear(RayDev r) { zrange_clear_wg(r, blockIdx.x, gridDim.x); }

__global__ void k_zrange_clear_batch(const RayJob *__restrict__ jobs) {
    zrange_clear_wg(global_ray(jobs[blockIdx.y].r), blockIdx.x, gridDim.x);
}

LF_DEV int sat16(float v) {
    if (!(v == v)) return 0;
    if (v > 32767.0f) return 32767;
    if (v < -32768.0f) return -32768;
    return (int)v;
}

constexpr int kZThreads = 512;  // lanes of a k_zrange_fill workgroup: 8 per visible block

// RenderingBlockDelegate::create_rendering_block (ray_tracing.cu:63-126) and
// the tile fill (:187-210).  Eight lanes take one visible block, one corner
// each, and combine their projections with three butterfly steps: the
// reference's compare-and-set loop over the corners, starting from the image
// bounds, is min / max with NaN ignored followed by the monotone 16-bit
// saturation, which is what fminf / fmaxf and sat16 compute.  The 8 lanes then
// share the rectangle's tiles (lane c takes columns c, c + 8, ... of each
// row).  kLds: the min / max go to an LDS copy of the tile grid first (a
// workgroup's blocks are consecutive visible-list entries, so their
// rectangles overlap heavily), then the tiles it changed, inside the bounding
// box of its rectangles, are flushed to the global grid with one atomic each.
// Min and max do not depend on order, so the tiles are the same as with
// per-block global atomics.
// Workgroup wg of nwg.  s_tile: kLds: [0, T) near z bits, [T, 2T) far z bits;
// s_box: ulx, uly, lrx, lry of the workgroup's rectangles; s_rend: its
// rendering blocks.  kFast (mixed precision): the projection with the
// hardware reciprocal.
template <bool kLds, bool kFast = false>
LF_DEV void zrange_fill_wg(const RayDev &r, const float *__restrict__ pose24, const lf_intrinsics &K, int wg, int nwg,
                           int *s_tile, int *s_box, unsigned int &s_rend) {
    constexpr int kPer = kZThreads / 8;  // blocks per workgroup pass
    const int tid = threadIdx.x, c = tid & 7;
    const int nvis = (int)r.counters[LF_CNT_VISIBLE];
    if (wg * kPer >= nvis) return;  // whole workgroup
    const int T = r.tiles_x * r.tiles_y;
    const int near0 = __float_as_int(100.0f), far0 = __float_as_int(0.0f);  // k_zrange_clear's values
    if (tid == 0) s_rend = 0u;
    if (kLds) {
        for (int i = tid; i < T; i += blockDim.x) {
            s_tile[i] = near0;
            s_tile[T + i] = far0;
        }
        if (tid == 0) {
            s_box[0] = s_box[1] = INT_MAX;
            s_box[2] = s_box[3] = -1;
        }
    }
    __syncthreads();
    float Tcw[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) Tcw[i] = pose24[12 + i];
    const float scale = 8.0f * r.voxel_size, qnan = __int_as_float(0x7fc00000);
    // the trip count is uniform over the workgroup, so every lane reaches the shuffles
    for (int i0 = wg * kPer; i0 < nvis; i0 += nwg * kPer) {
        const int i = i0 + tid / 8;
        const bool have = i < nvis;
        float fu = qnan, cu = qnan, fv = qnan, cv = qnan, zlo = qnan, zhi = qnan;  // NaN: ignored
        if (have) {
            const int4 bp = reinterpret_cast<const int4 *>(r.visible_pos)[i];
            const float wx = (float)(bp.x + (c & 1)) * scale, wy = (float)(bp.y + ((c >> 1) & 1)) * scale,
                        wz = (float)(bp.z + ((c >> 2) & 1)) * scale;
            const f3 pt = xform(Tcw, f3{wx, wy, wz});
            float u, v;
            if (kFast) {
                const float iz = __builtin_amdgcn_rcpf(pt.z);
                u = (K.fx * pt.x * iz + K.cx) * 0.125f;
                v = (K.fy * pt.y * iz + K.cy) * 0.125f;
            } else {
                u = (K.fx * pt.x / pt.z + K.cx) / 8.0f;
                v = (K.fy * pt.y / pt.z + K.cy) / 8.0f;
            }
            fu = floorf(u); cu = ceilf(u); fv = floorf(v); cv = ceilf(v);
            zlo = zhi = pt.z;
        }
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
            fu = fminf(fu, __shfl_xor(fu, o));
            cu = fmaxf(cu, __shfl_xor(cu, o));
            fv = fminf(fv, __shfl_xor(fv, o));
            cv = fmaxf(cv, __shfl_xor(cv, o));
            zlo = fminf(zlo, __shfl_xor(zlo, o));
            zhi = fmaxf(zhi, __shfl_xor(zhi, o));
        }
        int ulx = sat16(fminf((float)r.tiles_x, fu)), lrx = sat16(fmaxf(-1.0f, cu));
        int uly = sat16(fminf((float)r.tiles_y, fv)), lry = sat16(fmaxf(-1.0f, cv));
        float zr0 = fminf(r.zmax_raycast, zlo);
        const float zr1 = fmaxf(r.zmin_raycast, zhi);
        if (ulx < 0) ulx = 0;
        if (uly < 0) uly = 0;
        if (lrx >= r.tiles_x) lrx = r.tiles_x - 1;
        if (lry >= r.tiles_y) lry = r.tiles_y - 1;
        if (zr0 < r.zmin_raycast) zr0 = r.zmin_raycast;
        const bool renders = have && !(ulx > lrx || uly > lry) && !(zr1 < r.zmin_raycast);
        // counted per wavefront in LDS (one lane per block); one global atomic per workgroup below
        const unsigned long long rend = __ballot(renders && c == 0);
        if (lane_id() == 0 && rend) atomicAdd(&s_rend, (unsigned int)__popcll(rend));
        if (!renders) continue;
        if (kLds && c == 0) {
            atomicMin(&s_box[0], ulx);
            atomicMin(&s_box[1], uly);
            atomicMax(&s_box[2], lrx);
            atomicMax(&s_box[3], lry);
        }
        const int zi0 = __float_as_int(zr0), zi1 = __float_as_int(zr1);
        for (int y = uly; y <= lry; ++y)
            for (int x = ulx + c; x <= lrx; x += 8) {
                const int t = y * r.tiles_x + x;
                if (kLds) {
                    atomicMin(&s_tile[t], zi0);
                    atomicMax(&s_tile[T + t], zi1);
                } else {
                    atomicMin(reinterpret_cast<int *>(r.zx) + t, zi0);
                    atomicMax(reinterpret_cast<int *>(r.zy) + t, zi1);
                }
            }
    }
    __syncthreads();
    if (tid == 0 && s_rend) atomicAdd(&r.counters[LF_CNT_RENDER_BLOCKS], s_rend);
    if (kLds) {
        const int bx0 = s_box[0], by0 = s_box[1], bw = s_box[2] - bx0 + 1, bh = s_box[3] - by0 + 1;
        if (bw <= 0 || bh <= 0) return;
        for (int j = tid; j < bw * bh; j += blockDim.x) {
            const int t = (by0 + j / bw) * r.tiles_x + bx0 + j % bw;
            if (s_tile[t] != near0) atomicMin(reinterpret_cast<int *>(r.zx) + t, s_tile[t]);
            if (s_tile[T + t] != far0) atomicMax(reinterpret_cast<int *>(r.zy) + t, s_tile[T + t]);
        }
    }
}

template <bool kLds, bool kFast>
__global__ __launch_bounds__(kZThreads) void k_zrange_fill(RayDev r, const float *__restrict__ pose24,
                                                           lf_intrinsics K) {
    extern __shared__ int s_tile[];
    __shared__ int s_box[4];
    __shared__ unsigned int s_rend;
    zrange_fill_wg<kLds, kFast>(r, pose24, K, blockIdx.x, gridDim.x, s_tile, s_box, s_rend);
}

template <bool kLds, bool kFast>
__global__ __launch_bounds__(kZThreads) void k_zrange_fill_batch(const RayJob *__restrict__ jobs) {
    extern __shared__ int s_tile[];
    __shared__ int s_box[4];
    __shared__ unsigned int s_rend;
    const RayJob &j = jobs[blockIdx.y];
    zrange_fill_wg<kLds, kFast>(global_ray(j.r), global_ptr(j.pose24), j.K, blockIdx.x, gridDim.x, s_tile, s_box, s_rend);
}

// Workgroup-shared, direct-mapped cache of block lookups in LDS, negative
// results included.  An entry is one 64-bit word: valid bit, three 13-bit
// block coordinates and a 24-bit pool index (all ones = not allocated), so a
// single LDS store / load keeps it consistent without atomics.  The 4 tiles of
// a workgroup march through nearly the same blocks, and most samples of a ray
// lie in unallocated blocks, so most global hash probes disappear.
constexpr int kLdsCache = 512;
constexpr unsigned long long kNoBlock = 0xFFFFFFull;

struct BlockCache {
    int bx, by, bz, b;
    unsigned long long *lds;  // null: no shared cache (coordinates / pool out of range)
    const unsigned int *occ;  // the map's occupancy bits, copied to LDS
    int box;                  // b < 0: 2 when the block's whole 4x4x4-block region is unallocated, else 0
};

LF_DEV bool fits13(int v) { return v >= -4096 && v < 4096; }

LF_DEV unsigned long long lds_tag(int bx, int by, int bz) {
    return (1ull << 63) | ((unsigned long long)(bx & 0x1FFF) << 50) | ((unsigned long long)(by & 0x1FFF) << 37) |
           ((unsigned long long)(bz & 0x1FFF) << 24);
}

// Every sample of a ray lies within a few voxels of the segment between its
// start and end points, so one check per ray (all coordinates finite and
// within kRayRange voxels) lets the per-sample code use a plain floor and
// shift: no NaN or overflow guard per sample (to_i3's), no kBad block.  A ray
// outside the range has no sample in any representable block (block keys hold
// 21 bits per axis) and finds nothing.
constexpr float kRayRange = 1.0e9f;
// floor to int in one instruction (the floor-rounding conversion), equal to
// (int)floorf(v) for every |v| < 2^31
LF_DEV int ray_floor(float v) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}
// 1 when v's sign bit is clear, else 0 (for a v neither NaN nor zero: v > 0)
LF_DEV unsigned int positive(float v) { return ~__float_as_uint(v) >> 31; }

// The pool block of a block outside the block directory: the workgroup's LDS
// cache of lookups, then the hash table.  (Inside the directory its load is
// faster than the LDS cache round trip plus the loads of its misses: 250 ->
// 234 us at C1.)
LF_DEV int resolve_far(const RayDev &r, unsigned long long *lds, int bx, int by, int bz) {
    if (lds && fits13(bx) && fits13(by) && fits13(bz)) {
        const unsigned long long tag = lds_tag(bx, by, bz);
        const unsigned int h = ((unsigned int)bx * 73856093u ^ (unsigned int)by * 19349669u ^
                                (unsigned int)bz * 83492791u) & (kLdsCache - 1);
        const unsigned long long e = lds[h];
        if ((e & ~kNoBlock) == tag) {
            const unsigned long long idx = e & kNoBlock;
            return idx == kNoBlock ? -1 : (int)idx;
        }
        const int b = find_block(r.keys, r.slot_block, bx, by, bz, r.num_buckets, r.num_entries);
        lds[h] = tag | (b < 0 ? kNoBlock : (unsigned long long)b);
        return b;
    }
    return find_block(r.keys, r.slot_block, bx, by, bz, r.num_buckets, r.num_entries);
}

LF_DEV bool region_occupied(const BlockCache &bc, int bx, int by, int bz) {
    const unsigned int ob = occupancy_bit(bx, by, bz);
    return (bc.occ[ob >> 5] >> (ob & 31)) & 1u;
}

// Resolve the pool block of block coordinate (bx, by, bz): the lane's last
// block, then the occupancy bits (a clear bit: the whole 4x4x4-block region
// is unallocated, which is most of the free space a ray crosses), then one
// load from the block directory, outside it resolve_far.
LF_DEV int resolve(const RayDev &r, BlockCache &bc, int bx, int by, int bz) {
    if (bx == bc.bx && by == bc.by && bz == bc.bz) return bc.b;
    bc.bx = bx; bc.by = by; bc.bz = bz;
    const bool empty_region = !region_occupied(bc, bx, by, bz);
    bc.box = empty_region ? 2 : 0;
    if (empty_region) {
        bc.b = -1;
        return -1;
    }
    const int cell = dir_cell(bx, by, bz);
    bc.b = cell >= 0 ? r.block_dir[cell] : resolve_far(r, bc.lds, bx, by, bz);
    return bc.b;
}

// The block of a trilinear corner across a block face from the sample's own
// block, without the lane's cache (which keeps the sample's block): inside
// the directory its one load (no occupancy test: a block next to a sampled
// voxel's block seldom lies in an empty region), outside it the occupancy
// bits and resolve_far.
LF_DEV int corner_block(const RayDev &r, const BlockCache &bc, int bx, int by, int bz) {
    const int cell = dir_cell(bx, by, bz);
    if (cell >= 0) return r.block_dir[cell];
    return region_occupied(bc, bx, by, bz) ? resolve_far(r, bc.lds, bx, by, bz) : -1;
}

// The 2-byte sdf of voxel vi, addressed by a 32-bit byte offset (a pool of
// at most 2^31 / 512 blocks keeps 2 * vi below 2^32).
LF_DEV unsigned short sdf_bits(const RayDev &r, int vi) { return at_byte(reinterpret_cast<const unsigned short *>(r.sdf), 2u * (unsigned int)vi); }

// findVoxel + read_sdf (ray_tracing.cu:291-305) at a sample of an in-range
// ray.  A voxel is valid when its weight is non-zero, which the sdf plane
// encodes (kSdfUnseen; kHalfUnseen, a NaN, in mixed precision), so one 2-byte
// load per sample decides both validity and value.
// floor_out (or null) receives floor(p).
template <bool kMixed>
LF_DEV float read_sdf(const RayDev &r, BlockCache &bc, f3 p, bool &valid, int *voxel_out = nullptr,
                      i3 *floor_out = nullptr) {
    const int vx = ray_floor(p.x), vy = ray_floor(p.y), vz = ray_floor(p.z);
    if (floor_out) *floor_out = i3{vx, vy, vz};
    const int b = resolve(r, bc, vx >> 3, vy >> 3, vz >> 3);
    valid = false;
    if (b < 0) return __int_as_float(0x7fc00000);
    const int vi = b * LF_BLOCK3 + lidx(vx, vy, vz);
    if (voxel_out) *voxel_out = vi;
    if constexpr (kMixed) {
        const float sv = half_bits_to_float(sdf_bits(r, vi));
        valid = !isnan(sv);
        return sv;
    }
    const short sd = (short)sdf_bits(r, vi);
    if (sd == kSdfUnseen) return __int_as_float(0x7fc00000);
    valid = true;
    return unpack_sdf(sd);
}

// read_sdf_interped (ray_tracing.cu:307-339).  Corner k is the voxel of
// p + offset_k with offsets 0 or 1 per axis, so each axis has only two
// coordinates, floor(p) and floor(p + 1), computed once, and the corners lie
// in at most two blocks per axis.  Corner k's block is B[k & m], where m has
// a bit for each axis whose two coordinates lie in different blocks and B[e]
// is the block one step across the faces of e: B[0] is p's own block (the
// lane's cache), and the others are looked up in one loop over the nonzero
// subsets of m (one code path whatever faces a lane's corners straddle; none
// in the common case).  Then all 8 sdf loads are issued in one round trip
// (a corner in an unallocated block reads voxel 0 of pool block 0, and NaN
// replaces corner 0's value).  An invalid corner is NaN and the corner values are
// finite otherwise, so the result is valid exactly when it is not NaN.  The
// arithmetic is the reference's order.
// f0: floor(p) when the caller has it from the sample read at p, whose
// block the lane's cache then holds; else computed and resolved here.
template <bool kMixed>
LF_DEV float read_interp(const RayDev &r, BlockCache &bc, f3 p, bool &valid, const i3 *f0 = nullptr) {
    // floor(p) as int and as float (exact: an int from a float converts back
    // to that float)
    const i3 fl = f0 ? *f0 : i3{ray_floor(p.x), ray_floor(p.y), ray_floor(p.z)};
    const int cx[2] = {fl.x, ray_floor(p.x + 1.0f)}, cy[2] = {fl.y, ray_floor(p.y + 1.0f)},
              cz[2] = {fl.z, ray_floor(p.z + 1.0f)};
    const float fx = (float)cx[0], fy = (float)cy[0], fz = (float)cz[0];
    const f3 xyz = {p.x - fx, p.y - fy, p.z - fz};
    // corner order of the reference: (0,0,0) (1,0,0) (0,1,0) (1,1,0) (0,0,1) (1,0,1) (0,1,1) (1,1,1)
    const int bx0 = cx[0] >> 3, bx1 = cx[1] >> 3, by0 = cy[0] >> 3, by1 = cy[1] >> 3, bz0 = cz[0] >> 3,
              bz1 = cz[1] >> 3;
    const int m = (bx1 != bx0 ? 1 : 0) | (by1 != by0 ? 2 : 0) | (bz1 != bz0 ? 4 : 0);
    int bb[8];
    bb[0] = f0 ? bc.b : resolve(r, bc, bx0, by0, bz0);
#pragma unroll
    for (int k = 1; k < 8; ++k) bb[k] = bb[0];
    for (int e = m & -m; e != 0; e = (e - m) & m) {  // the nonzero subsets of m, increasing
        const int b = corner_block(r, bc, (e & 1) ? bx1 : bx0, (e & 2) ? by1 : by0, (e & 4) ? bz1 : bz0);
#pragma unroll
        for (int k = 1; k < 8; ++k)
            if ((k & m) == e) bb[k] = b;
    }
    // byte offsets of the corners' sdf: block << 10 | 2 * lidx
    const unsigned int ox[2] = {(unsigned int)(cx[0] & 7) << 1, (unsigned int)(cx[1] & 7) << 1},
                       oy[2] = {(unsigned int)(cy[0] & 7) << 4, (unsigned int)(cy[1] & 7) << 4},
                       oz[2] = {(unsigned int)(cz[0] & 7) << 7, (unsigned int)(cz[1] & 7) << 7};
    const unsigned short *sdf16 = reinterpret_cast<const unsigned short *>(r.sdf);
    unsigned short c[8];
    int bmin = bb[0];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        c[k] = at_byte(sdf16, ((unsigned int)max(bb[k], 0) << 10) | ox[k & 1] | oy[(k >> 1) & 1] | oz[k >> 2]);
        bmin = min(bmin, bb[k]);
    }
    const float qnan = __int_as_float(0x7fc00000);
    float s[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if constexpr (kMixed) s[k] = half_bits_to_float(c[k]);  // kHalfUnseen is a NaN already
        else s[k] = (short)c[k] == kSdfUnseen ? qnan : unpack_sdf((short)c[k]);
    }
    if (bmin < 0) s[0] = qnan;  // a corner in an unallocated block
    float r0 = (1.0f - xyz.x) * s[0] + xyz.x * s[1];
    float r1 = (1.0f - xyz.x) * s[2] + xyz.x * s[3];
    const float r2 = (1.0f - xyz.y) * r0 + xyz.y * r1;
    r0 = (1.0f - xyz.x) * s[4] + xyz.x * s[5];
    r1 = (1.0f - xyz.x) * s[6] + xyz.x * s[7];
    const float r3 = (1.0f - xyz.y) * r0 + xyz.y * r1;
    const float v = (1.0f - xyz.z) * r2 + xyz.z * r3;
    valid = !isnan(v);
    return v;
}

// MapRenderingDelegate::operator() (check_colour == false) and
// ::raycast_with_colour (check_colour == true): ray_tracing.cu:346-425,478-564.
// Workgroup wg (its index within this raycast's grid; the grid's x extent is
// a multiple of 8, so wg % 8 is the XCD the workgroup runs on).  kMixed: FP16
// voxels in (the vertices out are float4 in both precisions).
template <bool kCheckColour, bool kMixed>
LF_DEV void raycast_wg(const RayDev &r, float *__restrict__ vmap, unsigned char *__restrict__ rgb_out, int cols,
                       int rows, const float *__restrict__ pose24, const lf_intrinsics &K, int wg_x, int wg_y, int wg,
                       unsigned long long *s_cache, uint4 *s_occ) {
    // DenseMapping::raycast leaves the map untouched when nothing is visible
    // (voxel_hashing.cpp:92-93) or no rendering block was made (:105).
    if (r.counters[LF_CNT_VISIBLE] == 0u || r.counters[LF_CNT_RENDER_BLOCKS] == 0u) return;
    for (int i = threadIdx.x; i < kLdsCache; i += blockDim.x) s_cache[i] = 0ull;
    const uint4 *occ4 = reinterpret_cast<const uint4 *>(r.occupancy);
    for (int i = threadIdx.x; i < LF_OCC_WORDS / 4; i += blockDim.x) s_occ[i] = occ4[i];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // XCD-aware placement: workgroups are dealt to the 8 XCDs in turn, so
    // workgroup g (XCD g % 8) takes the (g / 8)-th workgroup tile of band g % 8:
    // each XCD marches a contiguous band of the image, and the blocks its rays
    // read stay in that XCD's L2.
    const int nwg = wg_x * wg_y, per_xcd = (nwg + kXcds - 1) / kXcds;
    const int L = (wg % kXcds) * per_xcd + wg / kXcds;
    if (L >= nwg) return;
    const int tile_x = (L % wg_x) * 2 + (wave & 1), tile_y = (L / wg_x) * 2 + (wave >> 1);
    const int x = tile_x * 8 + (lane & 7), y = tile_y * 8 + (lane >> 3);
    if (x >= cols || y >= rows) return;
    const size_t px = (size_t)y * cols + x;
    const float qnan = __int_as_float(0x7fc00000);
    float4 res4 = make_float4(qnan, qnan, qnan, qnan);
    unsigned char col[3] = {255, 255, 255};
    do {
        if (tile_x >= r.tiles_x || tile_y >= r.tiles_y) break;
        const int t = tile_y * r.tiles_x + tile_x;
        const float z0 = r.zx[t], z1 = r.zy[t];
        if (z1 < 1e-3f || z0 < 1e-3f || isnan(z0) || isnan(z1)) break;
        const float *Twc = pose24;
        const float inv_vs = r.inv_voxel_size;  // the map's (scalar registers)
        f3 pt = {((float)x - K.cx) * K.invfx * z0, ((float)y - K.cy) * K.invfy * z0, z0};
        float dist_s = norm(pt) * inv_vs;
        const f3 bs = mul(xform(Twc, pt), inv_vs);
        pt = f3{((float)x - K.cx) * K.invfx * z1, ((float)y - K.cy) * K.invfy * z1, z1};
        const float dist_e = norm(pt) * inv_vs;
        const f3 be = mul(xform(Twc, pt), inv_vs);
        const f3 dir = normalised(sub(be, bs));
        // the per-ray range check of the sample code (kRayRange); NaN fails it
        const float reach = fmaxf(fmaxf(fmaxf(fabsf(bs.x), fabsf(bs.y)), fmaxf(fabsf(bs.z), fabsf(be.x))),
                                  fmaxf(fabsf(be.y), fabsf(be.z)));
        if (!(reach < kRayRange) || isnan(bs.x + bs.y + bs.z + be.x + be.y + be.z) || isnan(dir.x) || isnan(dir.y) ||
            isnan(dir.z))
            break;
        f3 res = bs;
        BlockCache bc = {kBad, kBad, kBad, -1, r.pool_fits_lds ? s_cache : nullptr,
                         reinterpret_cast<const unsigned int *>(s_occ), 0};
        // A sample in an unallocated block (or empty 4x4x4-block region) is
        // followed by the reference's 2-voxel steps for as long as they stay
        // inside that box, where every read finds nothing: those steps are
        // taken without the reads, by the same float operations.  The count
        // stops 0.05 voxels short of where the ray leaves the box, so that no
        // sample the box does not surely hold is passed over.
        const f3 inv_dir = {__builtin_amdgcn_rcpf(dir.x), __builtin_amdgcn_rcpf(dir.y), __builtin_amdgcn_rcpf(dir.z)};
        // bit a set: the ray runs towards + along axis a (inv_dir's sign bit
        // clear, +inf included), so a box's exit face on that axis is its far side
        const unsigned int fwd = positive(inv_dir.x) | positive(inv_dir.y) << 1 | positive(inv_dir.z) << 2;
        // The march keeps 2 dir (exact), which a quiet step adds as it is; a
        // step of s voxels adds 2 dir * (s / 2), the same float as dir * s
        // (both factors are exact power-of-two scalings of the reference's).
        const f3 dir2 = mul(dir, 2.0f);
        const float half_scale = r.half_step_scale;
        float sdf = 1.0f, last_sdf, step;
        bool valid = false, back_face = false;
        while (dist_s < dist_e) {
            last_sdf = sdf;
            i3 fl;
            sdf = read_sdf<kMixed>(r, bc, res, valid, nullptr, &fl);
            int quiet = 0;  // further samples surely inside the same unallocated box
            if (bc.b < 0) {
                const int k = bc.box, sh = k + 3;  // region: 32 voxels a side; block: 8
                const float tx = ((float)(((bc.bx >> k) + (int)(fwd & 1u)) << sh) - res.x) * inv_dir.x;
                const float ty = ((float)(((bc.by >> k) + (int)((fwd >> 1) & 1u)) << sh) - res.y) * inv_dir.y;
                const float tz = ((float)(((bc.bz >> k) + (int)(fwd >> 2)) << sh) - res.z) * inv_dir.z;
                const float texit = fminf(tx, fminf(ty, tz));  // NaN / inf terms drop out
                // samples at 2k (k >= 1) with 2k < texit - 0.05
                if (texit < 1e4f) quiet = max(0, (int)ceilf(0.5f * (texit - 0.05f)) - 1);
            }
            if (sdf <= 0.5f && sdf >= -0.5f) sdf = read_interp<kMixed>(r, bc, res, valid, &fl);
            if (sdf <= 0.0f) break;
            if (sdf >= 0.0f && last_sdf < 0.0f) { back_face = true; break; }
            // half the reference's step max(sdf * step_scale, 1) (or 2), exactly
            const float half_step = valid ? fmaxf(sdf * half_scale, 0.5f) : 1.0f;
            res = add(res, mul(dir2, half_step));
            dist_s = __builtin_fmaf(2.0f, half_step, dist_s);  // dist_s + step, one rounding
            for (int k = 0; k < quiet && dist_s < dist_e; ++k) {  // each: a read that finds nothing, a 2-voxel step
                res = add(res, dir2);
                dist_s += 2.0f;
            }
        }
        if (back_face || !(sdf <= 0.0f)) break;
        const f3 dir1 = mul(dir2, 0.5f);  // dir again, exactly
        const float step_scale = 2.0f * half_scale;
        step = sdf * step_scale;
        res = add(res, mul(dir1, step));
        sdf = read_interp<kMixed>(r, bc, res, valid);
        step = sdf * step_scale;
        res = add(res, mul(dir1, step));
        if (!valid) break;
        if (kCheckColour) {  // read_colour (ray_tracing.cu:427-441)
            bool cv;
            int vi = -1;
            read_sdf<kMixed>(r, bc, res, cv, &vi);
            if (!cv) break;
            if constexpr (kMixed) {  // colour packed beside the weight code
                const unsigned int word = reinterpret_cast<const unsigned int *>(r.weight)[vi];
                col[0] = (unsigned char)word_channel(word, 0);
                col[1] = (unsigned char)word_channel(word, 1);
                col[2] = (unsigned char)word_channel(word, 2);
            } else {
                const int b = vi / LF_BLOCK3, l = vi % LF_BLOCK3;
                const unsigned char *c = r.rgb + (size_t)b * (3 * LF_BLOCK3) + l;
                col[0] = c[0]; col[1] = c[LF_BLOCK3]; col[2] = c[2 * LF_BLOCK3];
            }
        }
        const float *Tcw = pose24 + 12;
        const f3 q = xform(Tcw, mul(res, r.voxel_size));
        res4 = make_float4(q.x, q.y, q.z, 1.0f);
    } while (false);
    reinterpret_cast<float4 *>(vmap)[px] = res4;
    if (rgb_out) {
        unsigned char *o = rgb_out + 3 * ((size_t)y * cols + x);
        o[0] = col[0]; o[1] = col[1]; o[2] = col[2];
    }
}

template <bool kCheckColour, bool kMixed>
__global__ __launch_bounds__(256) void k_raycast(RayDev r, float *__restrict__ vmap, unsigned char *__restrict__ rgb_out,
                                                 int cols, int rows, const float *__restrict__ pose24,
                                                 lf_intrinsics K, int wg_x, int wg_y) {
    __shared__ unsigned long long s_cache[kLdsCache];
    __shared__ uint4 s_occ[LF_OCC_WORDS / 4];
    raycast_wg<kCheckColour, kMixed>(r, vmap, rgb_out, cols, rows, pose24, K, wg_x, wg_y, blockIdx.x, s_cache, s_occ);
}

// Many maps' rays fill the chip many times over, so this kernel is held to 64
// VGPRs for 8 waves per SIMD; the single-map kernel (about 5 waves per SIMD
// in all) keeps the compiler's own allocation, whose sample chain is shorter.
template <bool kCheckColour, bool kMixed>
__global__ __launch_bounds__(256, 8) void k_raycast_batch(const RayJob *__restrict__ jobs) {
    __shared__ unsigned long long s_cache[kLdsCache];
    __shared__ uint4 s_occ[LF_OCC_WORDS / 4];
    const RayJob &j = jobs[blockIdx.y];
    const int g = kXcds * ((j.wg_x * j.wg_y + kXcds - 1) / kXcds);
    if ((int)blockIdx.x >= g) return;  // this raycast's grid is smaller than the batch's
    raycast_wg<kCheckColour, kMixed>(global_ray(j.r), global_ptr(j.vmap), global_ptr(j.rgb_out), j.cols, j.rows,
                                     global_ptr(j.pose24), j.K, j.wg_x, j.wg_y, blockIdx.x, s_cache, s_occ);
}

// Dynamic LDS available to one workgroup on this device (less the static use).
size_t zrange_lds_limit() {
    static size_t limit = 0;
    if (!limit) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            v = 32768;
        limit = v > 1024 ? (size_t)v - 1024 : 0;
    }
    return limit;
}

}  // namespace

static RayDev make_raydev(const lf_map_params *p, const lf_map_storage *s) {
    RayDev r;
    r.keys = s->keys; r.slot_block = s->slot_block; r.sdf = s->sdf; r.weight = s->weight; r.rgb = s->rgb;
    r.visible = s->visible; r.visible_pos = s->visible_pos; r.block_dir = s->block_dir; r.block_pos = s->block_pos;
    r.counters = s->counters;
    r.occupancy = s->occupancy;
    r.zx = s->zrange; r.zy = s->zrange + (size_t)s->tiles_x * s->tiles_y;
    r.tiles_x = s->tiles_x; r.tiles_y = s->tiles_y;
    r.num_buckets = p->num_buckets; r.num_entries = p->num_entries;
    r.voxel_size = p->voxel_size; r.truncation = p->truncation;
    r.inv_voxel_size = 1.0f / p->voxel_size;
    r.half_step_scale = 0.5f * (p->truncation * r.inv_voxel_size);  // exact: a normal float (checked with the params)
    r.zmin_raycast = p->zmin_raycast; r.zmax_raycast = p->zmax_raycast;
    r.pool_fits_lds = p->num_blocks < (int)kNoBlock ? 1 : 0;
    return r;
}

static int zrange_grid(const lf_map_params *p) {
    return std::max(1, std::min(1024, (p->num_blocks + kZThreads / 8 - 1) / (kZThreads / 8)));
}

int lf_raycast_ev(const lf_map_params *p, const lf_map_storage *s, float *vmap, unsigned char *rgb_out, int cols,
                  int rows, const float *pose24, const lf_intrinsics *K, int check_colour, hipStream_t st,
                  hipEvent_t e0, hipEvent_t e1) {
    if (!p || !s || !vmap || !pose24 || !K) return lf_set_error(LF_ERR_ARG, "null argument to raycast");
    if (int rc = lf_check_storage(s, p)) return rc;
    if (cols / 8 != s->tiles_x || rows / 8 != s->tiles_y)
        return lf_set_error(LF_ERR_ARG, "raycast size %dx%d does not match the map's %dx%d tiles", cols, rows,
                            s->tiles_x, s->tiles_y);
    const RayDev r = make_raydev(p, s);
    const int T = s->tiles_x * s->tiles_y;
    k_zrange_clear<<<std::max(1, std::min(1024, (T + 255) / 256)), 256, 0, st>>>(r);
    LF_LAUNCH_CHECK();
    // the per-workgroup tile grid in LDS when it fits (C1: 4800 tiles, 38 KB)
    const size_t lds = 2 * (size_t)T * sizeof(int);
    const int zgrid = zrange_grid(p);
    const bool fast = p->precision == LF_PRECISION_MIXED;
    if (lds <= zrange_lds_limit()) {
        if (fast) k_zrange_fill<true, true><<<zgrid, kZThreads, lds, st>>>(r, pose24, *K);
        else k_zrange_fill<true, false><<<zgrid, kZThreads, lds, st>>>(r, pose24, *K);
    } else {
        if (fast) k_zrange_fill<false, true><<<zgrid, kZThreads, 0, st>>>(r, pose24, *K);
        else k_zrange_fill<false, false><<<zgrid, kZThreads, 0, st>>>(r, pose24, *K);
    }
    LF_LAUNCH_CHECK();
    const int tx = (cols + 7) / 8, ty = (rows + 7) / 8;
    const int wg_x = (tx + 1) / 2, wg_y = (ty + 1) / 2;  // 2x2 tiles per workgroup
    const int g = kXcds * ((wg_x * wg_y + kXcds - 1) / kXcds);
    if (e0) LF_HIP_CHECK(hipEventRecord(e0, st));
    if (p->precision == LF_PRECISION_MIXED) {  // FP16 voxels
        if (check_colour) k_raycast<true, true><<<g, 256, 0, st>>>(r, vmap, rgb_out, cols, rows, pose24, *K, wg_x, wg_y);
        else k_raycast<false, true><<<g, 256, 0, st>>>(r, vmap, rgb_out, cols, rows, pose24, *K, wg_x, wg_y);
    } else {
        if (check_colour) k_raycast<true, false><<<g, 256, 0, st>>>(r, vmap, rgb_out, cols, rows, pose24, *K, wg_x, wg_y);
        else k_raycast<false, false><<<g, 256, 0, st>>>(r, vmap, rgb_out, cols, rows, pose24, *K, wg_x, wg_y);
    }
    LF_LAUNCH_CHECK();
    if (e1) LF_HIP_CHECK(hipEventRecord(e1, st));
    return LF_OK;
}

size_t lf_ray_job_bytes(void) { return sizeof(RayJob); }

// Several raycasts (their own maps, images and poses) in one launch per stage:
// a 2-D grid whose y index is the raycast, x the workgroup within it.  Each
// raycast's workgroups run exactly the single-launch code, so its tiles and
// vertex map equal those of lf_raycast_ev; the batch packs the raycasts'
// long-ray tails onto one grid instead of leaving each tail alone on the chip.
int lf_raycast_batch_make(const LfRayItem *items, int n, void *jobs_host, LfRayBatchDims *dims) {
    if (!items || n <= 0 || !jobs_host || !dims) return lf_set_error(LF_ERR_ARG, "bad raycast batch arguments");
    if (n > 65535) return lf_set_error(LF_ERR_ARG, "raycast batch of %d exceeds the grid's y extent", n);
    RayJob *J = static_cast<RayJob *>(jobs_host);
    LfRayBatchDims d = {};
    d.n = n;
    d.mixed = items[0].p && items[0].p->precision == LF_PRECISION_MIXED;
    size_t lds_max = 0;
    for (int i = 0; i < n; ++i) {
        const LfRayItem &it = items[i];
        if (!it.p || !it.s || !it.vmap || !it.pose24 || !it.K)
            return lf_set_error(LF_ERR_ARG, "null argument in raycast batch item %d", i);
        if (it.p->precision != items[0].p->precision)
            return lf_set_error(LF_ERR_ARG, "raycast batch item %d: maps of one batch share one precision", i);
        if (int rc = lf_check_storage(it.s, it.p)) return rc;
        if (it.cols / 8 != it.s->tiles_x || it.rows / 8 != it.s->tiles_y)
            return lf_set_error(LF_ERR_ARG, "batch item %d: raycast size %dx%d does not match the map's %dx%d tiles",
                                i, it.cols, it.rows, it.s->tiles_x, it.s->tiles_y);
        J[i].r = make_raydev(it.p, it.s);
        J[i].vmap = it.vmap;
        J[i].rgb_out = it.rgb_out;
        J[i].pose24 = it.pose24;
        J[i].K = *it.K;
        J[i].cols = it.cols;
        J[i].rows = it.rows;
        const int tx = (it.cols + 7) / 8, ty = (it.rows + 7) / 8;
        J[i].wg_x = (tx + 1) / 2;
        J[i].wg_y = (ty + 1) / 2;
        const int T = it.s->tiles_x * it.s->tiles_y;
        d.tiles_max = std::max(d.tiles_max, T);
        d.zgrid_max = std::max(d.zgrid_max, zrange_grid(it.p));
        d.grid_max = std::max(d.grid_max, kXcds * ((J[i].wg_x * J[i].wg_y + kXcds - 1) / kXcds));
        lds_max = std::max(lds_max, 2 * (size_t)T * sizeof(int));
    }
    d.zrange_lds = lds_max;
    // n maps share the z-range launch: each takes 1/n of a lone map's grid
    // (at least 64 workgroups), so that a workgroup's LDS tile grid, cleared
    // and flushed once, collects several passes of blocks
    d.zgrid_max = std::max(std::min(d.zgrid_max, 64), d.zgrid_max / n);
    *dims = d;
    return LF_OK;
}

int lf_raycast_batch_launch(const void *jobs_dev, const LfRayBatchDims &d, int check_colour, hipStream_t st,
                            hipEvent_t e0, hipEvent_t e1) {
    if (!jobs_dev || d.n <= 0) return lf_set_error(LF_ERR_ARG, "bad raycast batch launch");
    const RayJob *jobs = static_cast<const RayJob *>(jobs_dev);
    const int n = d.n;
    k_zrange_clear_batch<<<dim3(std::max(1, std::min(1024, (d.tiles_max + 255) / 256)), n), 256, 0, st>>>(jobs);
    LF_LAUNCH_CHECK();
    const dim3 zg(d.zgrid_max, n);
    if (d.zrange_lds <= zrange_lds_limit()) {
        if (d.mixed) k_zrange_fill_batch<true, true><<<zg, kZThreads, d.zrange_lds, st>>>(jobs);
        else k_zrange_fill_batch<true, false><<<zg, kZThreads, d.zrange_lds, st>>>(jobs);
    } else {
        if (d.mixed) k_zrange_fill_batch<false, true><<<zg, kZThreads, 0, st>>>(jobs);
        else k_zrange_fill_batch<false, false><<<zg, kZThreads, 0, st>>>(jobs);
    }
    LF_LAUNCH_CHECK();
    const dim3 grid(d.grid_max, n);
    if (e0) LF_HIP_CHECK(hipEventRecord(e0, st));
    if (d.mixed) {
        if (check_colour) k_raycast_batch<true, true><<<grid, 256, 0, st>>>(jobs);
        else k_raycast_batch<false, true><<<grid, 256, 0, st>>>(jobs);
    } else {
        if (check_colour) k_raycast_batch<true, false><<<grid, 256, 0, st>>>(jobs);
        else k_raycast_batch<false, false><<<grid, 256, 0, st>>>(jobs);
    }
    LF_LAUNCH_CHECK();
    if (e1) LF_HIP_CHECK(hipEventRecord(e1, st));
    return LF_OK;
}

int lf_raycast_batch_ev(const LfRayItem *items, int n, void *jobs_host, void *jobs_dev, int check_colour,
                        hipStream_t st) {
    LfRayBatchDims d;
    if (int rc = lf_raycast_batch_make(items, n, jobs_host, &d)) return rc;
    if (!jobs_dev) return lf_set_error(LF_ERR_ARG, "bad raycast batch arguments");
    LF_HIP_CHECK(hipMemcpyAsync(jobs_dev, jobs_host, (size_t)n * sizeof(RayJob), hipMemcpyHostToDevice, st));
    return lf_raycast_batch_launch(jobs_dev, d, check_colour, st);
}

// raycast_check_visibility (voxel_hashing.cpp:121-133): the blocks visible from
// the pose (count_visible_entry, ray_tracing.cu:647-807), then the raycast.
extern "C" int lf_raycast_check_visibility(const lf_map_params *p, const lf_map_storage *s, float *vmap,
                                           unsigned char *rgb_out, int cols, int rows, const float *pose24,
                                           const lf_intrinsics *K, int check_colour, void *stream) {
    if (!p || !s || !vmap || !pose24 || !K) return lf_set_error(LF_ERR_ARG, "null argument to raycast_check_visibility");
    hipStream_t st = lf_stream(stream);
    if (int rc = lf_select_visible_ev(p, s, pose24, K, cols, rows, st)) return rc;
    return lf_raycast_ev(p, s, vmap, rgb_out, cols, rows, pose24, K, check_colour, st, nullptr, nullptr);
}

extern "C" int lf_raycast(const lf_map_params *p, const lf_map_storage *s, float *vmap, unsigned char *rgb_out,
                          int cols, int rows, const float *pose24, const lf_intrinsics *K, int check_colour,
                          void *stream) {
    return lf_raycast_ev(p, s, vmap, rgb_out, cols, rows, pose24, K, check_colour, lf_stream(stream), nullptr,
                         nullptr);
}
