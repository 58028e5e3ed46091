// Voxel-hash block allocation, visible-block selection and TSDF integration.
//
// Reference: src/voxel_hashing/depth_fusion.cu (create_blocks_kernel :121-228,
// check_visibility_flag_kernel :49-83, copy_visible_block_kernel :85-93,
// update_map_kernel / update_map_with_colour_kernel / update_map_weighted_kernel
// :230-428, host update / update_weighted :430-575) and the hash-map primitives
// of src/data_struct/map_struct.cu (computeHash :143-148, createBlock :280-325,
// findEntry :365-381).
//
// MI355X design (not a translation of the reference's scheme):
//  * The table is open addressing with linear probing from computeHash's home
//    bucket over num_entries slots.  A slot holds a 63-bit packed block
//    coordinate claimed with one 64-bit CAS — no bucket locks, so no request is
//    ever dropped (the reference drops an allocation when a bucket lock is
//    contended, map_struct.cu:303-324).
//  * Pool blocks are handed out by an atomic counter and their coordinates are
//    recorded in allocation order (block_pos), so visibility runs over the
//    allocated blocks, O(allocated), instead of scanning every table entry.
//  * A 16x16-pixel workgroup first deduplicates its requests in an LDS hash set
//    (ds_cmpst_b64): neighbouring pixels request the same few blocks, so only
//    unique keys reach HBM.
//  * The visible list is appended with one atomic per wavefront (ballot + popc).
//  * Integration walks the visible list with a fixed grid (grid-stride), one
//    256-thread workgroup per 8^3 block: each lane owns two voxels of the
//    block's contiguous 512-voxel run, so every voxel plane is read and written
//    with fully coalesced 128-byte wavefront accesses.
#include "lf_common.h"
#include "lf_internal.h"

using namespace lf;

namespace {

struct MapDev {
    unsigned long long *keys;
    int *slot_block;
    unsigned char *slot_flag;
    int *block_pos;
    short *sdf;
    float *weight;
    unsigned char *rgb;
    int *visible;
    int *visible_pos;  // {x, y, z, pool block} of visible[i]
    int *block_dir;    // LF_DIR_CELLS pool blocks, -1 = none
    unsigned int *counters;
    unsigned int *occupancy;
    int num_buckets, num_entries, num_blocks;
    float voxel_size, truncation, zmin_update, zmax_update;
    float inv_truncation;  // 1 / truncation, computed on the host (the same correctly rounded division)
};

// A map read from a descriptor table, its pointers asserted global.
LF_DEV MapDev global_dev(const MapDev &d) {
    MapDev m = d;
    m.keys = global_ptr(d.keys); m.slot_block = global_ptr(d.slot_block); m.slot_flag = global_ptr(d.slot_flag);
    m.block_pos = global_ptr(d.block_pos); m.sdf = global_ptr(d.sdf); m.weight = global_ptr(d.weight);
    m.rgb = global_ptr(d.rgb); m.visible = global_ptr(d.visible); m.visible_pos = global_ptr(d.visible_pos);
    m.block_dir = global_ptr(d.block_dir); m.counters = global_ptr(d.counters); m.occupancy = global_ptr(d.occupancy);
    return m;
}

MapDev make_dev(const lf_map_params *p, const lf_map_storage *s) {
    MapDev d;
    d.keys = s->keys; d.slot_block = s->slot_block; d.slot_flag = s->slot_flag; d.block_pos = s->block_pos;
    d.sdf = s->sdf; d.weight = s->weight; d.rgb = s->rgb; d.visible = s->visible; d.counters = s->counters;
    d.occupancy = s->occupancy;
    d.visible_pos = s->visible_pos;
    d.block_dir = s->block_dir;
    d.num_buckets = p->num_buckets; d.num_entries = p->num_entries; d.num_blocks = p->num_blocks;
    d.voxel_size = p->voxel_size; d.truncation = p->truncation;
    d.inv_truncation = 1.0f / p->truncation;
    d.zmin_update = p->zmin_update; d.zmax_update = p->zmax_update;
    return d;
}

// createBlock (map_struct.cu:280-325) as a lock-free insert.  Marks the slot as
// touched by this update (the reference's flag = 2, depth_fusion.cu:157).
// Returns the block's hash slot (-1: the table is full); *claimed: this call
// created the entry.
LF_DEV int insert_block(const MapDev &m, int bx, int by, int bz, bool *claimed_out = nullptr) {
    const unsigned long long key = pack_key(bx, by, bz);
    int slot = compute_hash(bx, by, bz, m.num_buckets);
    for (int i = 0; i < m.num_entries; ++i) {
        // A plain load may see a stale empty slot (another XCD's L2); the CAS is
        // authoritative.  A non-empty key never changes until the next reset.
        unsigned long long k = m.keys[slot];
        bool claimed = false;
        if (k == kEmptyKey) {
            k = atomicCAS(&m.keys[slot], kEmptyKey, key);
            claimed = (k == kEmptyKey);
        }
        if (claimed) {
            const unsigned int ob = occupancy_bit(bx, by, bz);
            atomicOr(&m.occupancy[ob >> 5], 1u << (ob & 31));
            unsigned int b = atomicAdd(&m.counters[LF_CNT_ALLOCATED], 1u);
            if (b < (unsigned int)m.num_blocks) {
                m.slot_block[slot] = (int)b;
                reinterpret_cast<int4 *>(m.block_pos)[b] = make_int4(bx, by, bz, slot);
                const int cell = dir_cell(bx, by, bz);
                if (cell >= 0) m.block_dir[cell] = (int)b;
            } else {
                atomicOr(&m.counters[LF_CNT_OVERFLOW], LF_OVF_POOL);
            }
            m.slot_flag[slot] = 1;
            if (claimed_out) *claimed_out = true;
            return slot;
        }
        if (k == key) {
            m.slot_flag[slot] = 1;
            return slot;
        }
        if (++slot == m.num_entries) slot = 0;
    }
    atomicOr(&m.counters[LF_CNT_OVERFLOW], LF_OVF_TABLE);
    return -1;
}

// One pixel of create_blocks_kernel (depth_fusion.cu:121-228): the 3-D
// Bresenham walk over the blocks between the near and far ends of the
// truncation band around the measured depth.  kFast (mixed precision): the
// voxel coordinates by a product with the reciprocal voxel size.
template <bool kFast, class F>
LF_DEV void band_blocks(int x, int y, float z, const float *Twc, const lf_intrinsics &K, const MapDev &m, F emit) {
    const float zmin = m.zmin_update, zmax = m.zmax_update;
    if (isnan(z) || z < zmin || z > zmax) return;
    const float half = m.truncation * 0.5f;
    const float zn = fmaxf(zmin, z - half);
    const float zf = fminf(zmax, z + half);
    if (zn >= zf) return;
    const float vs = m.voxel_size;
    const float ux = K.invfx * ((float)x - K.cx), uy = K.invfy * ((float)y - K.cy);
    f3 pn = xform(Twc, f3{ux * zn, uy * zn, zn});
    f3 pf = xform(Twc, f3{ux * zf, uy * zf, zf});
    i3 vn, vf;
    if (kFast) {
        const float iv = 1.0f / vs;
        vn = to_i3(mul(pn, iv));
        vf = to_i3(mul(pf, iv));
    } else {
        vn = to_i3(dvd(pn, vs));
        vf = to_i3(dvd(pf, vs));
    }
    if (vn.x == kBad || vn.y == kBad || vn.z == kBad || vf.x == kBad || vf.y == kBad || vf.z == kBad) return;
    i3 a = {blk(vn.x), blk(vn.y), blk(vn.z)}, b = {blk(vf.x), blk(vf.y), blk(vf.z)};
    const int dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
    const int sx = dx < 0 ? -1 : 1, sy = dy < 0 ? -1 : 1, sz = dz < 0 ? -1 : 1;
    const int ax = abs(dx), ay = abs(dy), az = abs(dz);
    const int ex = ax << 1, ey = ay << 1, ez = az << 1;
    emit(a.x, a.y, a.z);
    if (ax >= ay && ax >= az) {
        int e1 = ey - 1, e2 = ez - 1;
        for (int i = 0; i < ax; ++i) {
            if (e1 > 0) { a.y += sy; e1 -= ex; }
            if (e2 > 0) { a.z += sz; e2 -= ex; }
            e1 += ey; e2 += ez; a.x += sx;
            emit(a.x, a.y, a.z);
        }
    } else if (ay >= ax && ay >= az) {
        int e1 = ex - 1, e2 = ez - 1;
        for (int i = 0; i < ay; ++i) {
            if (e1 > 0) { a.x += sx; e1 -= ey; }
            if (e2 > 0) { a.z += sz; e2 -= ey; }
            e1 += ex; e2 += ez; a.y += sy;
            emit(a.x, a.y, a.z);
        }
    } else {
        int e1 = ey - 1, e2 = ex - 1;
        for (int i = 0; i < az; ++i) {
            if (e1 > 0) { a.y += sy; e1 -= ez; }
            if (e2 > 0) { a.x += sx; e2 -= ez; }
            e1 += ey; e2 += ex; a.z += sz;
            emit(a.x, a.y, a.z);
        }
    }
}

constexpr int kDedup = 1024;  // LDS request set per 16x16 tile (8 KiB)

LF_DEV unsigned int key_mix(unsigned long long k) {
    k ^= k >> 29;
    k *= 0xBF58476D1CE4E5B9ull;
    k ^= k >> 32;
    return (unsigned int)k;
}

// Mark an existing block as touched by the current frame (no insertion).
// The mask bit of a block that keyframe re-fusion's band walks created.
constexpr unsigned long long kRefuseFresh = 1ull << 63;

LF_DEV int touch_block(const MapDev &m, int bx, int by, int bz) {
    const int b = find_block(m.keys, m.slot_block, bx, by, bz, m.num_buckets, m.num_entries);
    if (b < 0) return -1;
    const int slot = reinterpret_cast<const int4 *>(m.block_pos)[b].w;
    m.slot_flag[slot] = 1;
    return slot;
}

// kInsert: create_blocks_kernel.  !kInsert (de-integration): the same band
// walk only marks the blocks that exist, so the visible list of a frame can be
// rebuilt without allocating.
// 16x16 pixel tile (tx, ty) of the frame.  slot_mask (keyframe re-fusion):
// each block the walk marks also gets `bit` in its hash slot's mask, and
// kRefuseFresh when the walk created it.
template <bool kInsert, bool kFast = false>
LF_DEV void alloc_wg(const MapDev &m, const float *__restrict__ depth, int cols, int rows,
                     const float *__restrict__ pose24, const lf_intrinsics &K, int tx, int ty,
                     unsigned long long *s_set, unsigned long long *slot_mask = nullptr,
                     unsigned long long bit = 0ull) {
    const int tid = threadIdx.x;
    const int x = tx * 16 + (tid & 15), y = ty * 16 + (tid >> 4);
    const bool inside = x < cols && y < rows;
    // the pixel's depth is in flight while the request set is cleared
    const float z = inside ? depth[(size_t)y * cols + x] : 0.0f;
    for (int i = tid; i < kDedup; i += 256) s_set[i] = kEmptyKey;
    if (tx == 0 && ty == 0 && tid == 0) {
        m.counters[LF_CNT_VISIBLE] = 0u;
        // one more fusion pass: read by the integration / de-integration that
        // follows this kernel (the seed of the mixed weights' dither)
        m.counters[LF_CNT_FUSE_PASSES] = m.counters[LF_CNT_FUSE_PASSES] + 1u;
    }
    float Twc[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) Twc[i] = pose24[i];
    __syncthreads();
    if (!inside) return;
    band_blocks<kFast>(x, y, z, Twc, K, m, [&](int bx, int by, int bz) {
        const unsigned long long key = pack_key(bx, by, bz);
        unsigned int h = key_mix(key) & (kDedup - 1);
        for (int probe = 0; probe < kDedup; ++probe) {
            unsigned long long old = atomicCAS(&s_set[h], kEmptyKey, key);
            if (old == kEmptyKey) {
                bool fresh = false;
                const int slot = kInsert ? insert_block(m, bx, by, bz, &fresh) : touch_block(m, bx, by, bz);
                if (slot_mask && slot >= 0) atomicOr(&slot_mask[slot], bit | (fresh ? kRefuseFresh : 0ull));
                return;
            }
            if (old == key) return;  // another lane of this tile owns the request
            h = (h + 1) & (kDedup - 1);
        }
        // LDS set full: go straight to the table
        bool fresh = false;
        const int slot = kInsert ? insert_block(m, bx, by, bz, &fresh) : touch_block(m, bx, by, bz);
        if (slot_mask && slot >= 0) atomicOr(&slot_mask[slot], bit | (fresh ? kRefuseFresh : 0ull));
    });
}

template <bool kInsert, bool kFast = false>
__global__ __launch_bounds__(256) void k_alloc(MapDev m, const float *__restrict__ depth, int cols, int rows,
                                               const float *__restrict__ pose24, lf_intrinsics K) {
    __shared__ unsigned long long s_set[kDedup];
    alloc_wg<kInsert, kFast>(m, depth, cols, rows, pose24, K, blockIdx.x, blockIdx.y, s_set);
}

// is_block_visible (depth_fusion.cu:27-47): any of the 8 block corners inside
// the image with z in [zmin, zmax] of the update.  kFast (mixed precision):
// the projection with the hardware reciprocal.
template <bool kFast = false>
LF_DEV bool block_visible(int bx, int by, int bz, const float *Tcw, const lf_intrinsics &K, int cols, int rows,
                          float zmin, float zmax, float scale) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float wx = (float)(bx + (c & 1)) * scale, wy = (float)(by + ((c >> 1) & 1)) * scale,
                    wz = (float)(bz + ((c >> 2) & 1)) * scale;
        f3 pt = xform(Tcw, f3{wx, wy, wz});
        float u, v;
        if (kFast) {
            const float iz = __builtin_amdgcn_rcpf(pt.z);
            u = K.fx * pt.x * iz + K.cx;
            v = K.fy * pt.y * iz + K.cy;
        } else {
            u = K.fx * pt.x / pt.z + K.cx;
            v = K.fy * pt.y / pt.z + K.cy;
        }
        if (!(u < 0.0f || v < 0.0f || u > (float)(cols - 1) || v > (float)(rows - 1) || pt.z < zmin || pt.z > zmax))
            return true;
    }
    return false;
}

constexpr int kVisThreads = 1024;  // lanes of a k_visible workgroup
constexpr int kVisGrid = 512;      // at most this many workgroups (grid-stride beyond)

// check_visibility_flag_kernel + exclusive scan + copy (depth_fusion.cu:49-93,
// 549-557) over the allocated blocks: touched-this-update, or frustum-visible.
// kGeometric: count_visible_entry (ray_tracing.cu:647-807), frustum-visible
// only, touched flags left alone.  The list is appended with one global atomic
// per workgroup (a wavefront ballot, then the wavefront counts prefixed in
// LDS): same-address atomics serialise, so one per wavefront cost most of the
// kernel.  The list order is unspecified, as with the per-wavefront append.
// Workgroup wg of nwg.
template <bool kGeometric, bool kFast = false>
LF_DEV void visible_wg(const MapDev &m, const float *__restrict__ pose24, const lf_intrinsics &K, int cols, int rows,
                       int wg, int nwg, unsigned int *s_cnt, unsigned int &s_base) {
    float Tcw[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) Tcw[i] = pose24[12 + i];
    const unsigned int n = min(m.counters[LF_CNT_ALLOCATED], (unsigned int)m.num_blocks);
    const float scale = 8.0f * m.voxel_size;
    const unsigned int stride = nwg * blockDim.x;
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    // the trip count is uniform over the workgroup (its barriers need every lane)
    for (unsigned int b0 = wg * blockDim.x; b0 < n; b0 += stride) {
        const unsigned int b = b0 + threadIdx.x;
        bool vis = false;
        int4 p = make_int4(0, 0, 0, 0);
        if (b < n) {
            p = reinterpret_cast<const int4 *>(m.block_pos)[b];
            if (!kGeometric && m.slot_flag[p.w]) {
                vis = true;
                m.slot_flag[p.w] = 0;
            } else {
                vis = block_visible<kFast>(p.x, p.y, p.z, Tcw, K, cols, rows, m.zmin_update, m.zmax_update, scale);
            }
        }
        const unsigned long long ballot = __ballot(vis);
        if (lane == 0) s_cnt[wave] = (unsigned int)__popcll(ballot);
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned int total = 0;
            for (int w = 0; w < kVisThreads / kWave; ++w) {
                const unsigned int c = s_cnt[w];
                s_cnt[w] = total;
                total += c;
            }
            s_base = total ? atomicAdd(&m.counters[LF_CNT_VISIBLE], total) : 0u;
        }
        __syncthreads();
        if (vis) {
            const unsigned int slot = s_base + s_cnt[wave] + __popcll(ballot & ((1ull << lane) - 1ull));
            m.visible[slot] = (int)b;
            reinterpret_cast<int4 *>(m.visible_pos)[slot] = make_int4(p.x, p.y, p.z, (int)b);
        }
        __syncthreads();  // s_cnt and s_base are rewritten by the next iteration
    }
}

template <bool kGeometric = false, bool kFast = false>
__global__ __launch_bounds__(kVisThreads) void k_visible(MapDev m, const float *__restrict__ pose24, lf_intrinsics K,
                                                         int cols, int rows) {
    __shared__ unsigned int s_cnt[kVisThreads / kWave];
    __shared__ unsigned int s_base;
    visible_wg<kGeometric, kFast>(m, pose24, K, cols, rows, blockIdx.x, gridDim.x, s_cnt, s_base);
}

int visible_grid(const lf_map_params *p) {
    return std::min(kVisGrid, std::max(1, (p->num_blocks + kVisThreads - 1) / kVisThreads));
}

// update_map_weighted_kernel (kWeighted, depth_fusion.cu:356-428) and
// update_map_with_colour_kernel (!kWeighted, :288-354).  kPacked (weighted
// only): the pixel's depth, weight and colour come from one 16-byte
// integration sample (DeviceImage frames), whose weight is the same
// expression evaluated when the sample was made; otherwise depth, normal and
// colour are gathered separately.  kMixed: FP16 sdf and 8-bit weight codes,
// and the projection and the weighted means use one reciprocal each instead
// of correctly rounded divisions.
// Workgroup wg of nwg (grid-stride over the visible list).
// The 256 weight-code values (weight_value) in LDS: a mixed voxel's weight
// is one LDS read instead of the decode's arithmetic.  Every thread of the
// workgroup calls it.
LF_DEV void fill_weight_table(float *s) {
    for (int i = threadIdx.x; i < 256; i += blockDim.x) s[i] = weight_value((unsigned int)i);
    __syncthreads();
}

// kTable: the weight codes' values from wtab, an LDS array of 256 floats that
// this function fills (every thread of the workgroup calls it).
template <bool kWeighted, bool kColour, bool kPacked, bool kMixed, bool kTable = false>
LF_DEV void integrate_wg(const MapDev &m, const float *__restrict__ depth, const float *__restrict__ nmap,
                         const unsigned char *__restrict__ rgb, int cols, int rows, const float *__restrict__ pose24,
                         const lf_intrinsics &K, const float4 *__restrict__ samples, int wg, int nwg,
                         float *wtab = nullptr) {
    static_assert(!kPacked || kWeighted, "packed samples carry the weighted update's weight");
    float Tcw[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) Tcw[i] = pose24[12 + i];
    const int nvis = (int)m.counters[LF_CNT_VISIBLE];
    const unsigned int pass = kMixed ? m.counters[LF_CNT_FUSE_PASSES] : 0u;
    // after the loads above: they stay scalar loads ahead of the barrier
    if constexpr (kTable) fill_weight_table(wtab);
    if (wg == 0 && threadIdx.x == 0) {
        atomicAdd(&m.counters[LF_CNT_UPDATES], 1u);
        atomicAdd(reinterpret_cast<unsigned long long *>(&m.counters[LF_CNT_VIS_SUM_LO]), (unsigned long long)nvis);
    }
    const float vs = m.voxel_size, trunc = m.truncation, inv_trunc = m.inv_truncation;
    const float zmin = m.zmin_update, zmax = m.zmax_update;
    const int t = threadIdx.x;
    const int lx = t & 7, ly = (t >> 3) & 7, lz0 = t >> 6;
    const float fcols = (float)cols, frows = (float)rows;
    // A lane owns voxels t and t + 256 of the block.  Three phases, each
    // issuing both voxels' loads before anything waits on them: the
    // projections and input gathers, then (inside the truncation band) the
    // voxel's weight, sdf and colour together, then the fusion and stores.
    // The block's list entry is the same for every lane (scalar registers,
    // loaded one iteration ahead); voxel and pixel addresses are 32-bit byte
    // offsets from a scalar base.
    // the voxel terms of the stochastic rounding's hash, once per lane
    const unsigned int rand_local[2] = {(unsigned int)t * kRandLocalMul, (unsigned int)(t + 256) * kRandLocalMul};
    const int4 *vpos = reinterpret_cast<const int4 *>(m.visible_pos);  // x, y, z, pool block
    int4 p_next = wg < nvis ? vpos[wg] : make_int4(0, 0, 0, 0);
    for (int i = wg; i < nvis; i += nwg) {
        const int4 p = make_int4(__builtin_amdgcn_readfirstlane(p_next.x), __builtin_amdgcn_readfirstlane(p_next.y),
                                 __builtin_amdgcn_readfirstlane(p_next.z), __builtin_amdgcn_readfirstlane(p_next.w));
        p_next = vpos[min(i + nwg, nvis - 1)];  // unconditional: no wait at a join
        const size_t b = (size_t)p.w;
        short *sdf_b = m.sdf + b * LF_BLOCK3;
        float *w_b = m.weight + b * LF_BLOCK3;
        unsigned int *wd_b = reinterpret_cast<unsigned int *>(m.weight) + b * LF_BLOCK3;  // mixed: voxel words
        unsigned char *c_b = m.rgb + b * (3 * LF_BLOCK3);
        bool ok[2];
        unsigned int pix[2];
        float ptz[2], dist[2], nz[2], wpk[2];
        unsigned int cpk[2];
        f3 pt0 = {0.f, 0.f, 0.f};
        if constexpr (kMixed) {
            const float wx = (float)(p.x * 8 + lx) * vs, wy = (float)(p.y * 8 + ly) * vs, wz = (float)(p.z * 8 + lz0) * vs;
            pt0 = xform(Tcw, f3{wx, wy, wz});
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int lz = lz0 + 4 * h;
            f3 pt;
            if constexpr (kMixed) {
                // mixed: the second voxel, 4 voxels along z, is one step of the transform's z column
                pt = h == 0 ? pt0
                            : f3{pt0.x + Tcw[2] * (4.0f * vs), pt0.y + Tcw[6] * (4.0f * vs), pt0.z + Tcw[10] * (4.0f * vs)};
            } else {
                const float wx = (float)(p.x * 8 + lx) * vs, wy = (float)(p.y * 8 + ly) * vs,
                            wz = (float)(p.z * 8 + lz) * vs;
                pt = xform(Tcw, f3{wx, wy, wz});
            }
            // The pixel is (floor(tu), floor(tv)) (__float2int_rd of the
            // rounded projection); it lies in the image exactly when
            // 0 <= tu < cols and 0 <= tv < rows (NaN fails), where the floor
            // is the truncating conversion.
            float tu, tv;
            if constexpr (kMixed) {
                const float iz = __builtin_amdgcn_rcpf(pt.z);
                tu = K.fx * pt.x * iz + K.cx + 0.5f;
                tv = K.fy * pt.y * iz + K.cy + 0.5f;
            } else {
                tu = K.fx * pt.x / pt.z + K.cx + 0.5f;
                tv = K.fy * pt.y / pt.z + K.cy + 0.5f;
            }
            ok[h] = tu >= 0.0f && tu < fcols && tv >= 0.0f && tv < frows;
            pix[h] = ok[h] ? __umul24((unsigned int)tv, (unsigned int)cols) + (unsigned int)tu : 0u;
            ptz[h] = pt.z;
            dist[h] = 0.0f; nz[h] = 0.0f; wpk[h] = 0.0f; cpk[h] = 0u;
        }
        // The gathers and voxel loads below are unconditional, so that both
        // voxels' loads issue before the first wait: a lane with nothing to
        // load reads pixel 0, or voxel 0 of its block, and its values are
        // never used.  Mixed precision reads every voxel's sdf and word in
        // the same round as the gathers (one memory round trip before the
        // fusion instead of two, for the bytes of the voxels the band leaves
        // out).
        unsigned int word[2], vo[2];
        short sp[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            vo[h] = vgpr_offset((unsigned int)(t + 256 * h));
            if constexpr (kMixed) {
                word[h] = at_byte(wd_b, 4u * vo[h]);  // weight code and colour together
                sp[h] = at_byte(sdf_b, 2u * vo[h]);
            }
        }
        float nx[2] = {0.0f, 0.0f};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (kPacked) {
                const float4 smp = at_byte(samples, vgpr_offset(16u * pix[h]));
                dist[h] = smp.x;  // NaN also when the normal is invalid
                wpk[h] = smp.y;
                cpk[h] = __float_as_uint(smp.z);
            } else {
                dist[h] = at_byte(depth, 4u * pix[h]);
                if (kWeighted) {
                    nx[h] = at_byte(nmap, 16u * pix[h]);
                    nz[h] = at_byte(nmap, 16u * pix[h] + 8u);
                }
            }
        }
        float sdf[2], wraw[2];
        unsigned char cold[2][3];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float d = dist[h];
            bool in = ok[h];
            if (kPacked) in = in && !(isnan(d) || d > zmax || d < zmin);
            else if (kWeighted) in = in && !(isnan(d) || isnan(nx[h]) || d > zmax || d < zmin);
            else in = in && !(isnan(d) || (double)d < 1e-2 || d > zmax || d < zmin);
            sdf[h] = d - ptz[h];
            in = in && !(sdf[h] < -trunc);
            ok[h] = in;
            if constexpr (!kMixed) {
                const unsigned int o = in ? vo[h] : 0u;
                wraw[h] = at_byte(w_b, 4u * o);
                if (kColour) {
                    cold[h][0] = at_byte(c_b, o);
                    cold[h][1] = at_byte(c_b, o + LF_BLOCK3);
                    cold[h][2] = at_byte(c_b, o + 2 * LF_BLOCK3);
                }
                sp[h] = at_byte(sdf_b, 2u * o);
            }
        }
        float wp[2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
            wp[h] = kMixed ? (kTable ? wtab[word_code(word[h])] : weight_value(word_code(word[h]))) : wraw[h];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (!ok[h]) continue;
            const float d = dist[h];
            const float sdf_h = fminf(1.0f, sdf[h] * inv_trunc);
            const unsigned int o = vo[h];
            const float wt = kPacked ? wpk[h] : kWeighted ? fabsf(sinf(nz[h])) / (d * d) : 1.0f / (d * d);
            unsigned char cin[3];
            if (kColour) {
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    cin[k] = kPacked ? (unsigned char)(cpk[h] >> (8 * k)) : rgb[3 * pix[h] + k];
            }
            if constexpr (kMixed) {
                unsigned short *sdf16 = reinterpret_cast<unsigned short *>(sdf_b);
                // the stored weight is rounded stochastically (unbiased running sums)
                const unsigned int rnd = voxel_rand_mixed(p.x, p.y, p.z, rand_local[h], pass);
                // without colour fusion the word's colour bits are kept
                unsigned int colour = word[h] & ~0xFFu;
                if (wp[h] == 0.0f) {
                    // a first observation with a positive weight is seen, as in
                    // the reference (the smallest code stands for weights below
                    // it); a zero weight leaves the voxel unseen
                    const unsigned int code = wt > 0.0f ? max(1u, weight_code_sr(fminf(wt, 255.0f), rnd)) : 0u;
                    at_byte(sdf16, vgpr_offset(2u * o)) = code == 0u ? kHalfUnseen : float_to_half_bits(sdf_h);
                    if (kColour) colour = word_colour_bits(cin[0], cin[1], cin[2]);
                    at_byte(wd_b, vgpr_offset(4u * o)) = code | colour;
                    continue;
                }
                const float inv = __builtin_amdgcn_rcpf(wp[h] + wt);
                const float spf = half_bits_to_float((unsigned short)sp[h]);
                at_byte(sdf16, vgpr_offset(2u * o)) = float_to_half_bits((spf * wp[h] + sdf_h * wt) * inv);
                const unsigned int code = weight_code_sr(fminf(wp[h] + wt, 255.0f), rnd);
                if (kColour) {
                    unsigned int ch[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float mc = ((float)word_channel(word[h], k) * wp[h] + (float)cin[k] * wt) * inv;
                        ch[k] = (unsigned int)(int)fminf(mc, 255.0f);
                    }
                    colour = word_colour_bits(ch[0], ch[1], ch[2]);
                }
                at_byte(wd_b, vgpr_offset(4u * o)) = code | colour;
                continue;
            }
            if (wp[h] == 0.0f) {
                // weight stays 0 only when wt is 0: keep the voxel's sdf code unseen
                at_byte(sdf_b, 2u * o) = wt == 0.0f ? kSdfUnseen : pack_sdf(sdf_h);
                at_byte(w_b, 4u * o) = wt > 255.0f ? 255.0f : wt;
                if (kColour) {
                    at_byte(c_b, o) = cin[0]; at_byte(c_b, o + LF_BLOCK3) = cin[1];
                    at_byte(c_b, o + 2 * LF_BLOCK3) = cin[2];
                }
                continue;
            }
            const float spf = unpack_sdf(sp[h]);
            const float fused = (spf * wp[h] + sdf_h * wt) / (wp[h] + wt);
            at_byte(sdf_b, 2u * o) = pack_sdf(fused);
            const float wn = wp[h] + wt;
            at_byte(w_b, 4u * o) = wn > 255.0f ? 255.0f : wn;
            if (kColour) {
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const float mixed = ((float)cold[h][k] * wp[h] + (float)cin[k] * wt) / (wp[h] + wt);
                    at_byte(c_b, o + k * LF_BLOCK3) = (unsigned char)(int)mixed;
                }
            }
        }
    }
}

template <bool kWeighted, bool kColour, bool kPacked = false, bool kMixed = false>
__global__ __launch_bounds__(256) void k_integrate(MapDev m, const float *__restrict__ depth,
                                                   const float *__restrict__ nmap,
                                                   const unsigned char *__restrict__ rgb, int cols, int rows,
                                                   const float *__restrict__ pose24, lf_intrinsics K,
                                                   const float4 *__restrict__ samples = nullptr) {
    __shared__ float s_wv[kMixed ? 256 : 1];
    integrate_wg<kWeighted, kColour, kPacked, kMixed, kMixed>(m, depth, nmap, rgb, cols, rows, pose24, K, samples,
                                                              blockIdx.x, gridDim.x, s_wv);
}

// One map's weighted update from a frame's packed samples within a batch
// (each a 2-D launch whose y index is the map): the arguments of the three
// kernels, and each map's grids.
struct FuseJob {
    MapDev m;
    const float *depth;
    const float4 *samples;
    const unsigned char *rgb;  // null: no colour
    const float *pose24;
    lf_intrinsics K;
    int cols, rows, tiles_x, tiles_y, vis_grid;
};

template <bool kFast>
__global__ __launch_bounds__(256) void k_alloc_batch(const FuseJob *__restrict__ jobs) {
    __shared__ unsigned long long s_set[kDedup];
    const FuseJob &j = jobs[blockIdx.y];
    const int t = blockIdx.x;
    if (t >= j.tiles_x * j.tiles_y) return;
    alloc_wg<true, kFast>(global_dev(j.m), global_ptr(j.depth), j.cols, j.rows, global_ptr(j.pose24), j.K, t % j.tiles_x,
                   t / j.tiles_x, s_set);
}

template <bool kFast>
__global__ __launch_bounds__(kVisThreads) void k_visible_batch(const FuseJob *__restrict__ jobs) {
    __shared__ unsigned int s_cnt[kVisThreads / kWave];
    __shared__ unsigned int s_base;
    const FuseJob &j = jobs[blockIdx.y];
    if ((int)blockIdx.x >= j.vis_grid) return;
    visible_wg<false, kFast>(global_dev(j.m), global_ptr(j.pose24), j.K, j.cols, j.rows, blockIdx.x, j.vis_grid, s_cnt,
                      s_base);
}

template <bool kMixed>
__global__ __launch_bounds__(256) void k_integrate_batch(const FuseJob *__restrict__ jobs) {
    __shared__ float s_wv[256];
    const FuseJob &j = jobs[blockIdx.y];
    if (j.tiles_x == 0) return;  // a map without an update this step (the whole workgroup)
    const MapDev m = global_dev(j.m);
    const float *depth = global_ptr(j.depth), *pose24 = global_ptr(j.pose24);
    const float4 *samples = global_ptr(j.samples);
    const unsigned char *rgb = global_ptr(j.rgb);
    if (rgb)
        integrate_wg<true, true, true, kMixed, kMixed>(m, depth, nullptr, rgb, j.cols, j.rows, pose24, j.K, samples,
                                                       blockIdx.x, gridDim.x, s_wv);
    else
        integrate_wg<true, false, true, kMixed, kMixed>(m, depth, nullptr, rgb, j.cols, j.rows, pose24, j.K, samples,
                                                        blockIdx.x, gridDim.x, s_wv);
}

// De-integration: the inverse of the weighted update for one frame, over the
// same voxels (the selection depends only on depth and pose).  With w the
// voxel's weight and w_f the frame's: w' = w - w_f, sdf' = (sdf w - sdf_f w_f) / w',
// colour likewise; w' <= 0 returns the voxel to unseen.  Exact up to the int16
// sdf and uint8 colour rounding, except where the weight saturated at 255.
// Not in the reference (deleteBlock, map_struct.cu:327-363, has no callers).
// kMixed: FP16 sdf and 8-bit weight codes (exact up to their rounding).
template <bool kColour, bool kPacked, bool kMixed = false>
__global__ __launch_bounds__(256) void k_deintegrate(MapDev m, const float *__restrict__ depth,
                                                     const float *__restrict__ nmap,
                                                     const unsigned char *__restrict__ rgb, int cols, int rows,
                                                     const float *__restrict__ pose24, lf_intrinsics K,
                                                     const float4 *__restrict__ samples) {
    float Tcw[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) Tcw[i] = pose24[12 + i];
    const int nvis = (int)m.counters[LF_CNT_VISIBLE];
    const unsigned int pass = kMixed ? m.counters[LF_CNT_FUSE_PASSES] : 0u;
    const float vs = m.voxel_size, trunc = m.truncation, inv_trunc = m.inv_truncation;
    const float zmin = m.zmin_update, zmax = m.zmax_update;
    const int t = threadIdx.x;
    const int lx = t & 7, ly = (t >> 3) & 7, lz0 = t >> 6;
    for (int i = blockIdx.x; i < nvis; i += gridDim.x) {
        const int4 p = reinterpret_cast<const int4 *>(m.visible_pos)[i];  // x, y, z, pool block
        const int b = p.w;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int lz = lz0 + 4 * h;
            const float wx = (float)(p.x * 8 + lx) * vs, wy = (float)(p.y * 8 + ly) * vs, wz = (float)(p.z * 8 + lz) * vs;
            const f3 pt = xform(Tcw, f3{wx, wy, wz});
            const int u = proj_floor(K.fx * pt.x / pt.z + K.cx + 0.5f);
            const int v = proj_floor(K.fy * pt.y / pt.z + K.cy + 0.5f);
            if (u < 0 || v < 0 || u > cols - 1 || v > rows - 1) continue;
            const size_t pix = (size_t)v * cols + u;
            float dist, wt;
            unsigned char cin[3] = {0, 0, 0};
            if (kPacked) {
                const float4 smp = samples[pix];
                dist = smp.x;
                wt = smp.y;
                const unsigned int c = __float_as_uint(smp.z);
                cin[0] = (unsigned char)c; cin[1] = (unsigned char)(c >> 8); cin[2] = (unsigned char)(c >> 16);
                if (isnan(dist) || dist > zmax || dist < zmin) continue;
            } else {
                dist = depth[pix];
                const float nx = nmap[4 * pix], nz = nmap[4 * pix + 2];
                if (isnan(dist) || isnan(nx) || dist > zmax || dist < zmin) continue;
                wt = fabsf(sinf(nz)) / (dist * dist);
                if (kColour) { cin[0] = rgb[3 * pix]; cin[1] = rgb[3 * pix + 1]; cin[2] = rgb[3 * pix + 2]; }
            }
            float sdf = dist - pt.z;
            if (sdf < -trunc) continue;
            sdf = fminf(1.0f, sdf * inv_trunc);
            const size_t vi = (size_t)b * LF_BLOCK3 + (size_t)(t + 256 * h);
            unsigned short *sdf16 = reinterpret_cast<unsigned short *>(m.sdf);
            unsigned int *wd = reinterpret_cast<unsigned int *>(m.weight);  // mixed: voxel words
            const unsigned int word = kMixed ? wd[vi] : 0u;
            const float w0 = kMixed ? weight_value(word_code(word)) : m.weight[vi];
            if (w0 == 0.0f) continue;  // nothing fused here
            unsigned char *c = kMixed ? nullptr : m.rgb + (size_t)b * (3 * LF_BLOCK3) + (size_t)(t + 256 * h);
            const float wr = w0 - wt;
            // mixed: the stored w0 is within one code step (at most 1/16 of
            // it) of the exact sum, so a remainder below that is the rounding
            // of the fused weights and the voxel returns to unseen; a larger
            // one is rounded stochastically, as fusion stores it
            const bool keep = kMixed ? wr > w0 * (1.0f / 16.0f) : wr > 0.0f;
            const unsigned int wr_code =
                kMixed && keep ? max(1u, weight_code_sr(wr, voxel_rand(p.x, p.y, p.z, (unsigned int)(t + 256 * h), pass)))
                               : 0u;
            if (!keep) {
                if (kMixed) {
                    sdf16[vi] = kHalfUnseen;
                    wd[vi] = kColour ? 0u : word & ~0xFFu;  // weight 0; colour cleared with colour fusion
                } else {
                    m.sdf[vi] = kSdfUnseen;
                    m.weight[vi] = 0.0f;
                    if (kColour) { c[0] = 0; c[LF_BLOCK3] = 0; c[2 * LF_BLOCK3] = 0; }
                }
                continue;
            }
            if (kMixed) {
                const float sr = (half_bits_to_float(sdf16[vi]) * w0 - sdf * wt) / wr;
                sdf16[vi] = float_to_half_bits(fmaxf(-1.0f, fminf(1.0f, sr)));
                unsigned int colour = word & ~0xFFu;
                if (kColour) {
                    unsigned int ch[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float cr = ((float)word_channel(word, k) * w0 - (float)cin[k] * wt) / wr;
                        ch[k] = (unsigned int)(int)fmaxf(0.0f, fminf(255.0f, cr));
                    }
                    colour = word_colour_bits(ch[0], ch[1], ch[2]);
                }
                wd[vi] = wr_code | colour;
            } else {
                const float sr = (unpack_sdf(m.sdf[vi]) * w0 - sdf * wt) / wr;
                m.sdf[vi] = pack_sdf(fmaxf(-1.0f, fminf(1.0f, sr)));
                m.weight[vi] = wr;
                if (kColour) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        const float cr = ((float)c[k * LF_BLOCK3] * w0 - (float)cin[k] * wt) / wr;
                        c[k * LF_BLOCK3] = (unsigned char)(int)fmaxf(0.0f, fminf(255.0f, cr));
                    }
                }
            }
        }
    }
}

// ------------------------------------------------------------ keyframe re-fusion
//
// n keyframes (each a frame's raw depth and packed integration samples, all of
// one size and intrinsics) leave the map at their old poses and are fused at
// their new poses in one pass: block allocation for every keyframe at its new
// pose (one launch), the union of the blocks any old or new pose sees (one
// launch), then one lane per voxel replays, in registers, the de-integration
// of keyframes 0..n-1 at their old poses followed by their weighted update at
// the new poses, in that order, and stores once.  Each step follows
// k_deintegrate / the weighted update (depth_fusion.cu:356-428) for that
// voxel, so the result is that of the 2n launches in sequence except that the
// voxel is not re-quantised between the steps.  A keyframe's step touches the
// blocks its own launches would have listed: those its pose sees
// (is_block_visible, depth_fusion.cu:27-47) and those the band walk of its
// depth at that pose marks (create_blocks_kernel, depth_fusion.cu:121-228;
// the old poses' walks only mark blocks that exist), recorded per hash slot
// as one bit per keyframe and pose.  A block this call creates did not exist
// for the de-integrations, nor for the updates before that of the first
// keyframe whose walk reaches it (in sequence it would be allocated there).
constexpr int kMaxRefuse = LF_REFUSE_MAX_KEYFRAMES;  // 2 poses each: one wavefront's lanes
static_assert(2 * kMaxRefuse < 64, "a wavefront tests every pose of a block; bit 63 marks new blocks");

struct RefuseArgs {
    MapDev m;
    unsigned long long *slot_mask;  // per hash slot: bit k old pose k, bit n + k new pose k
    const float *depth[kMaxRefuse];
    const float4 *samples[kMaxRefuse];
    const float *old24[kMaxRefuse];
    const float *new24[kMaxRefuse];
    int n, cols, rows;
    lf_intrinsics K;
};

__global__ void k_fill_u64(unsigned long long *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Row y < n of the grid: keyframe y's band walk at its old pose (existing
// blocks only); row n + k: keyframe k's at its new pose, allocating.
template <bool kFast>
__global__ __launch_bounds__(256) void k_refuse_alloc(RefuseArgs a) {
    __shared__ unsigned long long s_set[kDedup];
    const int tiles_x = (a.cols + 15) / 16;
    const int y = blockIdx.y, t = blockIdx.x;
    if (y < a.n)
        alloc_wg<false, kFast>(a.m, a.depth[y], a.cols, a.rows, a.old24[y], a.K, t % tiles_x, t / tiles_x, s_set,
                               a.slot_mask, 1ull << y);
    else
        alloc_wg<true, kFast>(a.m, a.depth[y - a.n], a.cols, a.rows, a.new24[y - a.n], a.K, t % tiles_x,
                              t / tiles_x, s_set, a.slot_mask, 1ull << y);
}

// Pose j of the 2n (j < n: keyframe j's old pose, else keyframe j - n's new
// pose) as a float T_cw.
LF_DEV const float *refuse_pose(const RefuseArgs &a, int j) {
    return (j < a.n ? a.old24[j] : a.new24[j - a.n]) + 12;
}

// The visible list becomes every allocated block that a keyframe's old or new
// pose sees or whose slot the band walks marked (the slot flags, left set
// for k_refuse to clear).
template <bool kFast>
__global__ __launch_bounds__(kVisThreads) void k_refuse_visible(RefuseArgs a) {
    __shared__ unsigned int s_cnt[kVisThreads / kWave];
    __shared__ unsigned int s_base;
    const MapDev &m = a.m;
    const unsigned int nb = min(m.counters[LF_CNT_ALLOCATED], (unsigned int)m.num_blocks);
    const float scale = 8.0f * m.voxel_size;
    const unsigned int stride = gridDim.x * blockDim.x;
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    for (unsigned int b0 = blockIdx.x * blockDim.x; b0 < nb; b0 += stride) {
        const unsigned int b = b0 + threadIdx.x;
        bool vis = false;
        int4 p = make_int4(0, 0, 0, 0);
        if (b < nb) {
            p = reinterpret_cast<const int4 *>(m.block_pos)[b];
            vis = m.slot_flag[p.w] != 0 || a.slot_mask[p.w] != 0ull;
            for (int j = 0; j < 2 * a.n && !vis; ++j) {
                float Tcw[12];
#pragma unroll
                for (int i = 0; i < 12; ++i) Tcw[i] = refuse_pose(a, j)[i];
                vis = block_visible<kFast>(p.x, p.y, p.z, Tcw, a.K, a.cols, a.rows, m.zmin_update, m.zmax_update,
                                           scale);
            }
        }
        const unsigned long long ballot = __ballot(vis);
        if (lane == 0) s_cnt[wave] = (unsigned int)__popcll(ballot);
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned int total = 0;
            for (int w = 0; w < kVisThreads / kWave; ++w) {
                const unsigned int c = s_cnt[w];
                s_cnt[w] = total;
                total += c;
            }
            s_base = total ? atomicAdd(&m.counters[LF_CNT_VISIBLE], total) : 0u;
        }
        __syncthreads();
        if (vis) {
            const unsigned int slot = s_base + s_cnt[wave] + __popcll(ballot & ((1ull << lane) - 1ull));
            m.visible[slot] = (int)b;
            reinterpret_cast<int4 *>(m.visible_pos)[slot] = make_int4(p.x, p.y, p.z, (int)b);
        }
        __syncthreads();
    }
}

// Keyframe k's observation of a voxel at world point (wx, wy, wz) under the
// float 3x4 T_cw: the weighted update's selection (pixel in the image, depth
// within the update range, sdf >= -truncation), its clamped sdf in truncation
// units, its weight and colour.  kFast: the hardware reciprocal projection of
// the mixed-precision update.
template <bool kFast>
LF_DEV bool refuse_observe(const RefuseArgs &a, int k, const float *Tcw, float wx, float wy, float wz, float &sdf,
                           float &wt, unsigned int &rgb) {
    const MapDev &m = a.m;
    const f3 pt = xform(Tcw, f3{wx, wy, wz});
    int u, v;
    if (kFast) {
        const float iz = __builtin_amdgcn_rcpf(pt.z);
        u = proj_floor(a.K.fx * pt.x * iz + a.K.cx + 0.5f);
        v = proj_floor(a.K.fy * pt.y * iz + a.K.cy + 0.5f);
    } else {
        u = proj_floor(a.K.fx * pt.x / pt.z + a.K.cx + 0.5f);
        v = proj_floor(a.K.fy * pt.y / pt.z + a.K.cy + 0.5f);
    }
    if (u < 0 || v < 0 || u > a.cols - 1 || v > a.rows - 1) return false;
    const float4 smp = a.samples[k][(size_t)v * a.cols + u];
    const float d = smp.x;
    if (isnan(d) || d > m.zmax_update || d < m.zmin_update) return false;
    const float s = d - pt.z;
    if (s < -m.truncation) return false;
    sdf = fminf(1.0f, s * m.inv_truncation);
    wt = smp.y;
    rgb = __float_as_uint(smp.z);
    return true;
}

// One lane per voxel (two voxels per lane, 256 lanes per block) of the visible
// list.  kMixed: FP16 sdf and the packed weight-code / colour word, the
// weights rounded stochastically once at the end (as the update stores them).
template <bool kColour, bool kMixed>
__global__ __launch_bounds__(256) void k_refuse(RefuseArgs a) {
    const MapDev &m = a.m;
    const int nvis = (int)m.counters[LF_CNT_VISIBLE];
    const unsigned int pass = kMixed ? m.counters[LF_CNT_FUSE_PASSES] : 0u;
    const float vs = m.voxel_size, scale = 8.0f * vs;
    const int t = threadIdx.x, lane = lane_id();
    const int lx = t & 7, ly = (t >> 3) & 7, lz0 = t >> 6;
    for (int i = blockIdx.x; i < nvis; i += gridDim.x) {
        const int4 p = reinterpret_cast<const int4 *>(m.visible_pos)[i];
        const int b = p.w;
        // which poses see the block: lane j < 2n of each wavefront tests pose j
        bool sees = false;
        if (lane < 2 * a.n) {
            float Tcw[12];
#pragma unroll
            for (int c = 0; c < 12; ++c) Tcw[c] = refuse_pose(a, lane)[c];
            sees = block_visible<kMixed>(p.x, p.y, p.z, Tcw, a.K, a.cols, a.rows, m.zmin_update, m.zmax_update,
                                         scale);
        }
        const int slot = reinterpret_cast<const int4 *>(m.block_pos)[b].w;
        const bool flagged = m.slot_flag[slot] != 0;
        const unsigned long long marks = a.slot_mask[slot], all = (1ull << a.n) - 1ull;
        const unsigned long long seen = __ballot(sees) | (marks & ~kRefuseFresh);
        unsigned long long old_mask = seen & all, new_mask = (seen >> a.n) & all;
        if (marks & kRefuseFresh) {
            // created here: by the first keyframe whose new-pose walk reached it
            const unsigned long long walks = (marks >> a.n) & all;
            old_mask = 0ull;
            new_mask &= ~((walks & (0ull - walks)) - 1ull);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int lz = lz0 + 4 * h;
            const float wx = (float)(p.x * 8 + lx) * vs, wy = (float)(p.y * 8 + ly) * vs, wz = (float)(p.z * 8 + lz) * vs;
            const unsigned int o = (unsigned int)(t + 256 * h);
            const size_t vi = (size_t)b * LF_BLOCK3 + o;
            unsigned short *sdf16 = reinterpret_cast<unsigned short *>(m.sdf);
            unsigned int *wd = reinterpret_cast<unsigned int *>(m.weight);
            unsigned char *c = m.rgb + (size_t)b * (3 * LF_BLOCK3) + o;
            // the voxel's running mean (sdf, colour) and weight
            float w, sdf, col[3] = {0.0f, 0.0f, 0.0f};
            unsigned int word = 0u;
            if constexpr (kMixed) {
                word = wd[vi];
                w = weight_value(word_code(word));
                sdf = w > 0.0f ? half_bits_to_float(sdf16[vi]) : 0.0f;
#pragma unroll
                for (int k = 0; k < 3; ++k) col[k] = (float)word_channel(word, k);
            } else {
                w = m.weight[vi];
                sdf = w > 0.0f ? unpack_sdf(m.sdf[vi]) : 0.0f;
                if (kColour) {
                    col[0] = c[0]; col[1] = c[LF_BLOCK3]; col[2] = c[2 * LF_BLOCK3];
                }
            }
            bool changed = false;
            // de-integration at the old poses (k_deintegrate's steps)
            for (unsigned long long mk = old_mask; mk; mk &= mk - 1ull) {
                const int k = __builtin_ctzll(mk);
                float s, wt;
                unsigned int rgb;
                if (w == 0.0f || !refuse_observe<kMixed>(a, k, a.old24[k] + 12, wx, wy, wz, s, wt, rgb)) continue;
                changed = true;
                const float wr = w - wt;
                const bool keep = kMixed ? wr > w * (1.0f / 16.0f) : wr > 0.0f;
                if (!keep) {
                    w = 0.0f;
                    sdf = 0.0f;
                    if (kColour) col[0] = col[1] = col[2] = 0.0f;
                    continue;
                }
                sdf = fmaxf(-1.0f, fminf(1.0f, (sdf * w - s * wt) / wr));
                if (kColour) {
#pragma unroll
                    for (int q = 0; q < 3; ++q)
                        col[q] = fmaxf(0.0f, fminf(255.0f, (col[q] * w - (float)((rgb >> (8 * q)) & 0xFFu) * wt) / wr));
                }
                w = wr;
            }
            // the weighted update at the new poses
            for (unsigned long long mk = new_mask; mk; mk &= mk - 1ull) {
                const int k = __builtin_ctzll(mk);
                float s, wt;
                unsigned int rgb;
                if (!refuse_observe<kMixed>(a, k, a.new24[k] + 12, wx, wy, wz, s, wt, rgb)) continue;
                changed = true;
                if (w == 0.0f) {
                    sdf = s;
                    w = fminf(wt, 255.0f);
                    if (kColour) {
#pragma unroll
                        for (int q = 0; q < 3; ++q) col[q] = (float)((rgb >> (8 * q)) & 0xFFu);
                    }
                    continue;
                }
                const float wn = w + wt;
                sdf = (sdf * w + s * wt) / wn;
                if (kColour) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) col[q] = (col[q] * w + (float)((rgb >> (8 * q)) & 0xFFu) * wt) / wn;
                }
                w = fminf(wn, 255.0f);
            }
            if (!changed) continue;
            if constexpr (kMixed) {
                unsigned int colour = word & ~0xFFu;
                if (kColour)
                    colour = word_colour_bits((unsigned int)(int)col[0], (unsigned int)(int)col[1],
                                              (unsigned int)(int)col[2]);
                // a voxel left with a positive weight stays seen (the smallest
                // code stands for weights below it); none left: unseen
                const unsigned int code =
                    w > 0.0f ? max(1u, weight_code_sr(fminf(w, 255.0f), voxel_rand(p.x, p.y, p.z, o, pass))) : 0u;
                sdf16[vi] = code ? float_to_half_bits(sdf) : kHalfUnseen;
                wd[vi] = code | (code ? colour : (kColour ? 0u : colour));
            } else {
                m.sdf[vi] = w > 0.0f ? pack_sdf(sdf) : kSdfUnseen;
                m.weight[vi] = w;
                if (kColour) {
                    c[0] = (unsigned char)(int)col[0];
                    c[LF_BLOCK3] = (unsigned char)(int)col[1];
                    c[2 * LF_BLOCK3] = (unsigned char)(int)col[2];
                }
            }
        }
        __syncthreads();  // every lane has read the slot flag
        if (t == 0 && flagged) m.slot_flag[slot] = 0;
    }
}

// one map at 640x480 / 5 mm (~40k visible blocks): 4096 workgroups measured
// 70 us, 8192 and 16384 both 66 us
int integrate_grid() { return 8192; }

int check_update_args(const lf_map_params *p, const lf_map_storage *s, const float *depth, int cols, int rows,
                      const float *pose24, const lf_intrinsics *K) {
    if (!p || !s || !depth || !pose24 || !K) return lf_set_error(LF_ERR_ARG, "null argument to update");
    if (cols <= 0 || rows <= 0) return lf_set_error(LF_ERR_ARG, "bad image size %dx%d", cols, rows);
    return LF_OK;
}

}  // namespace

int lf_update_ev(const lf_map_params *p, const lf_map_storage *s, const float *depth, const float *nmap,
                 const unsigned char *rgb, int cols, int rows, const float *pose24, const lf_intrinsics *K,
                 hipStream_t st, bool weighted, hipEvent_t e0, hipEvent_t e1, const float *samples) {
    if (int rc = lf_check_storage(s, p)) return rc;
    const MapDev m = make_dev(p, s);
    dim3 g((cols + 15) / 16, (rows + 15) / 16);
    const bool mixed = p->precision == LF_PRECISION_MIXED;
    if (mixed) k_alloc<true, true><<<g, 256, 0, st>>>(m, depth, cols, rows, pose24, *K);
    else k_alloc<true, false><<<g, 256, 0, st>>>(m, depth, cols, rows, pose24, *K);
    LF_LAUNCH_CHECK();
    if (mixed) k_visible<false, true><<<visible_grid(p), kVisThreads, 0, st>>>(m, pose24, *K, cols, rows);
    else k_visible<false, false><<<visible_grid(p), kVisThreads, 0, st>>>(m, pose24, *K, cols, rows);
    LF_LAUNCH_CHECK();
    const int G = integrate_grid();
    if (e0) LF_HIP_CHECK(hipEventRecord(e0, st));
    const float4 *smp = reinterpret_cast<const float4 *>(samples);
    if (p->precision == LF_PRECISION_MIXED) {
        // the frame's packed samples (DeviceImage) or separate depth / normal / colour
        if (weighted && smp) {
            if (rgb) k_integrate<true, true, true, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K, smp);
            else k_integrate<true, false, true, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K, smp);
        } else if (weighted) {
            if (rgb) k_integrate<true, true, false, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K);
            else k_integrate<true, false, false, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K);
        } else {
            if (rgb) k_integrate<false, true, false, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K);
            else k_integrate<false, false, false, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K);
        }
    } else if (weighted && smp) {
        if (rgb) k_integrate<true, true, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K, smp);
        else k_integrate<true, false, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K, smp);
    } else if (weighted) {
        if (rgb) k_integrate<true, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K);
        else k_integrate<true, false><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K);
    } else {
        if (rgb) k_integrate<false, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K);
        else k_integrate<false, false><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K);
    }
    LF_LAUNCH_CHECK();
    if (e1) LF_HIP_CHECK(hipEventRecord(e1, st));
    return LF_OK;
}

int lf_deintegrate_ev(const lf_map_params *p, const lf_map_storage *s, const float *depth, const float *nmap,
                      const unsigned char *rgb, int cols, int rows, const float *pose24, const lf_intrinsics *K,
                      hipStream_t st, const float *samples) {
    if (int rc = lf_check_storage(s, p)) return rc;
    const MapDev m = make_dev(p, s);
    dim3 g((cols + 15) / 16, (rows + 15) / 16);
    const bool mixed = p->precision == LF_PRECISION_MIXED;
    if (mixed) k_alloc<false, true><<<g, 256, 0, st>>>(m, depth, cols, rows, pose24, *K);
    else k_alloc<false, false><<<g, 256, 0, st>>>(m, depth, cols, rows, pose24, *K);
    LF_LAUNCH_CHECK();
    if (mixed) k_visible<false, true><<<visible_grid(p), kVisThreads, 0, st>>>(m, pose24, *K, cols, rows);
    else k_visible<false, false><<<visible_grid(p), kVisThreads, 0, st>>>(m, pose24, *K, cols, rows);
    LF_LAUNCH_CHECK();
    const int G = integrate_grid();
    const float4 *smp = reinterpret_cast<const float4 *>(samples);
    if (p->precision == LF_PRECISION_MIXED) {
        if (smp) {
            if (rgb) k_deintegrate<true, true, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K, smp);
            else k_deintegrate<false, true, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K, smp);
        } else {
            if (rgb) k_deintegrate<true, false, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K, smp);
            else k_deintegrate<false, false, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K, smp);
        }
    } else if (smp) {
        if (rgb) k_deintegrate<true, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K, smp);
        else k_deintegrate<false, true><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K, smp);
    } else {
        if (rgb) k_deintegrate<true, false><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K, smp);
        else k_deintegrate<false, false><<<G, 256, 0, st>>>(m, depth, nmap, rgb, cols, rows, pose24, *K, smp);
    }
    LF_LAUNCH_CHECK();
    return LF_OK;
}

size_t lf_fuse_job_bytes(void) { return sizeof(FuseJob); }

int lf_update_batch_make(const LfFuseItem *items, int n, void *jobs_host, LfFuseBatchDims *dims) {
    if (!items || n <= 0 || n > 65535 || !jobs_host || !dims) return lf_set_error(LF_ERR_ARG, "bad update batch");
    FuseJob *J = static_cast<FuseJob *>(jobs_host);
    LfFuseBatchDims d = {};
    d.n = n;
    for (int i = 0; i < n; ++i) {
        const LfFuseItem &it = items[i];
        FuseJob &j = J[i];
        j = FuseJob{};
        if (!it.enabled) continue;  // tiles 0: every kernel skips this map
        if (!it.p || !it.s || !it.depth || !it.samples || !it.pose24 || !it.K || it.cols <= 0 || it.rows <= 0)
            return lf_set_error(LF_ERR_ARG, "bad update batch item %d", i);
        if (int rc = lf_check_storage(it.s, it.p)) return rc;
        const int mixed = it.p->precision == LF_PRECISION_MIXED;
        if (d.any && mixed != d.mixed)
            return lf_set_error(LF_ERR_ARG, "update batch item %d: maps of one batch share one precision", i);
        d.any = 1;
        d.mixed = mixed;
        j.m = make_dev(it.p, it.s);
        j.depth = it.depth;
        j.samples = reinterpret_cast<const float4 *>(it.samples);
        j.rgb = it.rgb;
        j.pose24 = it.pose24;
        j.K = *it.K;
        j.cols = it.cols;
        j.rows = it.rows;
        j.tiles_x = (it.cols + 15) / 16;
        j.tiles_y = (it.rows + 15) / 16;
        j.vis_grid = visible_grid(it.p);
        d.tiles_max = std::max(d.tiles_max, j.tiles_x * j.tiles_y);
        d.vis_grid_max = std::max(d.vis_grid_max, j.vis_grid);
    }
    // the single launch's 8192 workgroups shared out, at least 2048 per map
    // (16 mixed maps, frames 10-49, 5 alternating runs: 1024 / 2048 / 4096
    // per map gave 7376 / 7433 / 7429 frames/s aggregate)
    d.integrate_grid = std::max(2048, integrate_grid() / n);
    *dims = d;
    return LF_OK;
}

int lf_update_batch_launch(const void *jobs_dev, const LfFuseBatchDims &d, hipStream_t st, hipEvent_t e0,
                           hipEvent_t e1) {
    if (!jobs_dev || d.n <= 0) return lf_set_error(LF_ERR_ARG, "bad update batch launch");
    if (!d.any) return LF_OK;
    const FuseJob *J = static_cast<const FuseJob *>(jobs_dev);
    if (d.mixed) {
        k_alloc_batch<true><<<dim3(d.tiles_max, d.n), 256, 0, st>>>(J);
        LF_LAUNCH_CHECK();
        k_visible_batch<true><<<dim3(d.vis_grid_max, d.n), kVisThreads, 0, st>>>(J);
    } else {
        k_alloc_batch<false><<<dim3(d.tiles_max, d.n), 256, 0, st>>>(J);
        LF_LAUNCH_CHECK();
        k_visible_batch<false><<<dim3(d.vis_grid_max, d.n), kVisThreads, 0, st>>>(J);
    }
    LF_LAUNCH_CHECK();
    if (e0) LF_HIP_CHECK(hipEventRecord(e0, st));
    if (d.mixed) k_integrate_batch<true><<<dim3(d.integrate_grid, d.n), 256, 0, st>>>(J);
    else k_integrate_batch<false><<<dim3(d.integrate_grid, d.n), 256, 0, st>>>(J);
    LF_LAUNCH_CHECK();
    if (e1) LF_HIP_CHECK(hipEventRecord(e1, st));
    return LF_OK;
}

// The three launches of a re-fusion (the mask array cleared before them).
static int refuse_launch(const lf_map_params *p, const RefuseArgs &a, int colour, hipStream_t st) {
    LF_LAUNCH_CHECK();  // the mask clear
    const bool mixed = p->precision == LF_PRECISION_MIXED;
    const dim3 g(((a.cols + 15) / 16) * ((a.rows + 15) / 16), 2 * a.n);
    if (mixed) k_refuse_alloc<true><<<g, 256, 0, st>>>(a);
    else k_refuse_alloc<false><<<g, 256, 0, st>>>(a);
    LF_LAUNCH_CHECK();
    if (mixed) k_refuse_visible<true><<<visible_grid(p), kVisThreads, 0, st>>>(a);
    else k_refuse_visible<false><<<visible_grid(p), kVisThreads, 0, st>>>(a);
    LF_LAUNCH_CHECK();
    const int G = integrate_grid();
    if (mixed) {
        if (colour) k_refuse<true, true><<<G, 256, 0, st>>>(a);
        else k_refuse<false, true><<<G, 256, 0, st>>>(a);
    } else {
        if (colour) k_refuse<true, false><<<G, 256, 0, st>>>(a);
        else k_refuse<false, false><<<G, 256, 0, st>>>(a);
    }
    LF_LAUNCH_CHECK();
    return LF_OK;
}

int lf_refuse_ev(const lf_map_params *p, const lf_map_storage *s, const float *const *depth,
                 const float *const *samples, const float *const *old_pose24, const float *const *new_pose24, int n,
                 int cols, int rows, const lf_intrinsics *K, int colour, hipStream_t st) {
    if (!p || !s || !depth || !samples || !old_pose24 || !new_pose24 || !K)
        return lf_set_error(LF_ERR_ARG, "null argument to keyframe re-fusion");
    if (n < 1 || n > kMaxRefuse)
        return lf_set_error(LF_ERR_ARG, "keyframe re-fusion takes 1..%d keyframes, not %d", kMaxRefuse, n);
    if (cols <= 0 || rows <= 0) return lf_set_error(LF_ERR_ARG, "bad image size %dx%d", cols, rows);
    if (int rc = lf_check_storage(s, p)) return rc;
    if (colour && p->precision != LF_PRECISION_MIXED && !s->rgb)
        return lf_set_error(LF_ERR_ARG, "colour re-fusion needs the map's colour planes");
    RefuseArgs a = {};
    a.m = make_dev(p, s);
    for (int k = 0; k < n; ++k) {
        if (!depth[k] || !samples[k] || !old_pose24[k] || !new_pose24[k])
            return lf_set_error(LF_ERR_ARG, "null input of keyframe %d", k);
        a.depth[k] = depth[k];
        a.samples[k] = reinterpret_cast<const float4 *>(samples[k]);
        a.old24[k] = old_pose24[k];
        a.new24[k] = new_pose24[k];
    }
    a.n = n;
    a.cols = cols;
    a.rows = rows;
    a.K = *K;
    // the per-slot masks: stream-ordered device memory for the call's length
    const size_t E = (size_t)p->num_entries;
    void *mask = nullptr;
    LF_HIP_CHECK(hipMallocAsync(&mask, E * sizeof(unsigned long long), st));
    a.slot_mask = static_cast<unsigned long long *>(mask);
    k_fill_u64<<<(int)std::min<size_t>(4096, (E + 255) / 256), 256, 0, st>>>(a.slot_mask, E);
    const int rc = refuse_launch(p, a, colour, st);
    const hipError_t ef = hipFreeAsync(mask, st);
    if (rc) return rc;
    if (ef != hipSuccess) return lf_set_error(LF_ERR_HIP, "hipFreeAsync failed: %s", hipGetErrorString(ef));
    return LF_OK;
}

extern "C" int lf_deintegrate_weighted(const lf_map_params *p, const lf_map_storage *s, const float *depth,
                                       const float *nmap, const unsigned char *rgb, int cols, int rows,
                                       const float *pose24, const lf_intrinsics *K, void *stream) {
    int rc = check_update_args(p, s, depth, cols, rows, pose24, K);
    if (rc) return rc;
    if (!nmap) return lf_set_error(LF_ERR_ARG, "deintegrate_weighted needs a normal map");
    return lf_deintegrate_ev(p, s, depth, nmap, rgb, cols, rows, pose24, K, lf_stream(stream), nullptr);
}

__global__ void k_reset_visible(MapDev m) {
    if (threadIdx.x == 0 && blockIdx.x == 0) m.counters[LF_CNT_VISIBLE] = 0u;
}

int lf_select_visible_ev(const lf_map_params *p, const lf_map_storage *s, const float *pose24, const lf_intrinsics *K,
                         int cols, int rows, hipStream_t st) {
    if (int rc = lf_check_storage(s, p)) return rc;
    const MapDev m = make_dev(p, s);
    k_reset_visible<<<1, 64, 0, st>>>(m);
    LF_LAUNCH_CHECK();
    if (p->precision == LF_PRECISION_MIXED)
        k_visible<true, true><<<visible_grid(p), kVisThreads, 0, st>>>(m, pose24, *K, cols, rows);
    else
        k_visible<true, false><<<visible_grid(p), kVisThreads, 0, st>>>(m, pose24, *K, cols, rows);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" int lf_select_visible(const lf_map_params *p, const lf_map_storage *s, const float *pose24,
                                 const lf_intrinsics *K, int cols, int rows, unsigned int *count_out, void *stream) {
    if (!p || !s || !pose24 || !K || cols <= 0 || rows <= 0)
        return lf_set_error(LF_ERR_ARG, "bad arguments to lf_select_visible");
    hipStream_t st = lf_stream(stream);
    if (int rc = lf_select_visible_ev(p, s, pose24, K, cols, rows, st)) return rc;
    if (count_out) {
        LF_HIP_CHECK(hipMemcpyAsync(count_out, s->counters + LF_CNT_VISIBLE, sizeof(unsigned int),
                                    hipMemcpyDeviceToHost, st));
        LF_HIP_CHECK(hipStreamSynchronize(st));
    }
    return LF_OK;
}

extern "C" int lf_update_weighted(const lf_map_params *p, const lf_map_storage *s, const float *depth,
                                  const float *nmap, const unsigned char *rgb, int cols, int rows,
                                  const float *pose24, const lf_intrinsics *K, void *stream) {
    int rc = check_update_args(p, s, depth, cols, rows, pose24, K);
    if (rc) return rc;
    if (!nmap) return lf_set_error(LF_ERR_ARG, "update_weighted needs a normal map");
    return lf_update_ev(p, s, depth, nmap, rgb, cols, rows, pose24, K, lf_stream(stream), true, nullptr, nullptr);
}

extern "C" int lf_update(const lf_map_params *p, const lf_map_storage *s, const float *depth,
                         const unsigned char *rgb, int cols, int rows, const float *pose24, const lf_intrinsics *K,
                         void *stream) {
    int rc = check_update_args(p, s, depth, cols, rows, pose24, K);
    if (rc) return rc;
    return lf_update_ev(p, s, depth, nullptr, rgb, cols, rows, pose24, K, lf_stream(stream), false, nullptr,
                        nullptr);
}
