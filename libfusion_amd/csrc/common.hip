// Library-wide entry points: errors, intrinsics, hash, tables, map storage.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "lf_common.h"
#include "mc_tables.h"

static thread_local char g_err[512] = "";

int lf_set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

extern "C" const char *lf_last_error(void) { return g_err; }

extern "C" const char *lf_version(void) { return "libfusion_amd 0.1 (gfx950, wave64)"; }

extern "C" void lf_intrinsics_make(int width, int height, float fx, float fy, float cx, float cy,
                                   lf_intrinsics *out) {
    out->width = width;
    out->height = height;
    out->fx = fx;
    out->fy = fy;
    out->cx = cx;
    out->cy = cy;
    out->invfx = 1.0f / fx;
    out->invfy = 1.0f / fy;
}

extern "C" void lf_intrinsics_pyr_down(const lf_intrinsics *in, lf_intrinsics *out) {
    const float s = 0.5f;
    lf_intrinsics_make((int)(s * (float)in->width), (int)(s * (float)in->height), s * in->fx, s * in->fy,
                       s * in->cx, s * in->cy, out);
}

extern "C" int lf_compute_hash(int x, int y, int z, int num_buckets) {
    return lf::compute_hash(x, y, z, num_buckets);
}

extern "C" int lf_occupancy_bit(int bx, int by, int bz) { return (int)lf::occupancy_bit(bx, by, bz); }

extern "C" int lf_dir_cell(int bx, int by, int bz) { return lf::dir_cell(bx, by, bz); }

extern "C" void lf_mc_tables(int *edge_out, int *tri_out) {
    for (int i = 0; i < 256; ++i) {
        edge_out[i] = lf_mc_edge_table[i];
        for (int j = 0; j < 16; ++j) tri_out[i * 16 + j] = lf_mc_tri_table[i][j];
    }
}

extern "C" int lf_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// ------------------------------------------------------------ map storage

static int check_params(const lf_map_params *p) {
    if (!p) return lf_set_error(LF_ERR_ARG, "null params");
    if (p->num_buckets <= 0 || p->num_entries < p->num_buckets || p->num_blocks <= 0)
        return lf_set_error(LF_ERR_ARG, "bad map capacities (buckets %d, entries %d, blocks %d)", p->num_buckets,
                            p->num_entries, p->num_blocks);
    if ((long long)p->num_blocks * LF_BLOCK3 > (long long)INT_MAX)
        return lf_set_error(LF_ERR_ARG, "pool of %d blocks exceeds 32-bit voxel indexing", p->num_blocks);
    if (!(p->voxel_size > 0) || !(p->truncation > 0) || !(p->zmin_raycast > 0) || !(p->zmin_update >= 0))
        return lf_set_error(LF_ERR_ARG, "bad voxel size / truncation / z ranges");
    // the raycast's step scale (truncation in voxels) and its half are normal floats
    const float step_scale = p->truncation * (1.0f / p->voxel_size);
    if (!(step_scale >= 0x1p-125f && step_scale <= 0x1p100f))
        return lf_set_error(LF_ERR_ARG, "truncation / voxel size %g out of range", (double)step_scale);
    if (p->precision != LF_PRECISION_FP32 && p->precision != LF_PRECISION_MIXED)
        return lf_set_error(LF_ERR_ARG, "unknown map precision %d", p->precision);
    return LF_OK;
}

// Bytes of one voxel's weight word: float32, or in mixed precision the 8-bit
// weight code with the voxel's colour packed beside it (32 bits).
size_t lf_weight_bytes(const lf_map_params *p) { return sizeof(float); }

// Every array of a map storage present and the tile grid non-empty; entry
// points call this before any launch, so storage filled in by hand with an
// array missing is an error rather than a kernel writing through null.  The
// colour planes (rgb) are required only of an FP32 map (p given): a mixed
// map packs its colour into the weight words.
int lf_check_storage(const lf_map_storage *s, const lf_map_params *p) {
    if (!s) return lf_set_error(LF_ERR_ARG, "null map storage");
    if (p)
        if (int rc = check_params(p)) return rc;
    const bool need_rgb = p && p->precision != LF_PRECISION_MIXED;
    const struct { const void *ptr; const char *name; } arrays[] = {
        {s->keys, "keys"}, {s->slot_block, "slot_block"}, {s->slot_flag, "slot_flag"},
        {s->block_pos, "block_pos"}, {s->sdf, "sdf"}, {s->weight, "weight"}, {need_rgb ? static_cast<const void *>(s->rgb) : static_cast<const void *>(s), "rgb"},
        {s->visible, "visible"}, {s->counters, "counters"}, {s->zrange, "zrange"},
        {s->occupancy, "occupancy"}, {s->visible_pos, "visible_pos"}, {s->block_dir, "block_dir"}};
    for (const auto &a : arrays)
        if (!a.ptr) return lf_set_error(LF_ERR_ARG, "map storage array %s is null", a.name);
    if (s->tiles_x <= 0 || s->tiles_y <= 0)
        return lf_set_error(LF_ERR_ARG, "map storage tile grid %dx%d", s->tiles_x, s->tiles_y);
    return LF_OK;
}

extern "C" int lf_map_create(const lf_map_params *p, int width, int height, lf_map_storage *s) {
    int rc = check_params(p);
    if (rc) return rc;
    if (!s || width <= 0 || height <= 0) return lf_set_error(LF_ERR_ARG, "bad map storage / size");
    memset(s, 0, sizeof(*s));
    const size_t E = (size_t)p->num_entries, B = (size_t)p->num_blocks;
    s->tiles_x = width / 8;
    s->tiles_y = height / 8;
    if (s->tiles_x <= 0 || s->tiles_y <= 0) return lf_set_error(LF_ERR_ARG, "raycast size below 8x8");
    LF_HIP_CHECK(hipMalloc(&s->keys, E * sizeof(unsigned long long)));
    LF_HIP_CHECK(hipMalloc(&s->slot_block, E * sizeof(int)));
    LF_HIP_CHECK(hipMalloc(&s->slot_flag, E));
    LF_HIP_CHECK(hipMalloc(&s->block_pos, B * 4 * sizeof(int)));
    LF_HIP_CHECK(hipMalloc(&s->sdf, B * LF_BLOCK3 * sizeof(short)));
    LF_HIP_CHECK(hipMalloc(&s->weight, B * LF_BLOCK3 * lf_weight_bytes(p)));
    if (p->precision != LF_PRECISION_MIXED) LF_HIP_CHECK(hipMalloc(&s->rgb, B * LF_BLOCK3 * 3));
    LF_HIP_CHECK(hipMalloc(&s->visible, B * sizeof(int)));
    LF_HIP_CHECK(hipMalloc(&s->counters, LF_NUM_COUNTERS * sizeof(unsigned int)));
    LF_HIP_CHECK(hipMalloc(&s->zrange, 2 * (size_t)s->tiles_x * s->tiles_y * sizeof(float)));
    LF_HIP_CHECK(hipMalloc(&s->occupancy, LF_OCC_WORDS * sizeof(unsigned int)));
    LF_HIP_CHECK(hipMalloc(&s->visible_pos, B * 4 * sizeof(int)));
    LF_HIP_CHECK(hipMalloc(&s->block_dir, (size_t)LF_DIR_CELLS * sizeof(int)));
    return LF_OK;
}

extern "C" int lf_map_release(lf_map_storage *s) {
    if (!s) return LF_OK;
    void *ptrs[] = {s->keys, s->slot_block, s->slot_flag, s->block_pos, s->sdf, s->weight,
                    s->rgb, s->visible, s->counters, s->zrange, s->occupancy, s->visible_pos, s->block_dir};
    for (void *q : ptrs)
        if (q) LF_HIP_CHECK(hipFree(q));
    memset(s, 0, sizeof(*s));
    return LF_OK;
}

__global__ void k_reset_table(unsigned long long *keys, int *slot_block, unsigned char *slot_flag, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        keys[i] = lf::kEmptyKey;
        slot_block[i] = -1;
        slot_flag[i] = 0;
    }
}

__global__ void k_zero_u32(unsigned int *p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

__global__ void k_fill_u128(uint4 *p, size_t n, unsigned int v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(v, v, v, v);
}

// MapStruct::reset (map_struct.cu:74-93).  Zero-fill kernels rather than
// memset nodes keep the reset capturable in a HIP graph.
extern "C" int lf_map_reset(const lf_map_params *p, const lf_map_storage *s, void *stream) {
    int rc = check_params(p);
    if (rc) return rc;
    if ((rc = lf_check_storage(s, p))) return rc;
    hipStream_t st = lf_stream(stream);
    const int E = p->num_entries;
    k_reset_table<<<std::min(4096, (E + 255) / 256), 256, 0, st>>>(s->keys, s->slot_block, s->slot_flag, E);
    LF_LAUNCH_CHECK();
    k_zero_u32<<<1, 64, 0, st>>>(s->counters, LF_NUM_COUNTERS);
    LF_LAUNCH_CHECK();
    k_zero_u32<<<LF_OCC_WORDS / 256, 256, 0, st>>>(s->occupancy, LF_OCC_WORDS);
    LF_LAUNCH_CHECK();
    const size_t B = (size_t)p->num_blocks;
    // sdf (2 B, every voxel starts unseen), weight word (4 B, 0: weight 0 and,
    // in a mixed map, colour 0), rgb (3 B, 0; FP32 maps) per voxel, and the
    // block directory (every cell -1); every array is a multiple of 512 bytes.
    const bool mixed = p->precision == LF_PRECISION_MIXED;
    const unsigned int unseen2 = mixed ? 0xFFFFFFFFu   // two FP16 kHalfUnseen codes
                                       : 0x80008000u;  // two int16 kSdfUnseen codes
    struct { void *ptr; size_t bytes; unsigned int v; } arrs[] = {
        {s->sdf, B * 1024, unseen2}, {s->weight, B * LF_BLOCK3 * lf_weight_bytes(p), 0u},
        {mixed ? nullptr : s->rgb, B * 1536, 0u}, {s->block_dir, (size_t)LF_DIR_CELLS * sizeof(int), 0xFFFFFFFFu}};
    for (auto &a : arrs) {
        if (!a.ptr) continue;
        size_t n16 = a.bytes / 16;
        int grid = (int)std::min<size_t>(8192, (n16 + 255) / 256);
        k_fill_u128<<<grid, 256, 0, st>>>(reinterpret_cast<uint4 *>(a.ptr), n16, a.v);
        LF_LAUNCH_CHECK();
    }
    return LF_OK;
}

__global__ void k_dir_fill(const int4 *block_pos, const unsigned int *counters, int num_blocks, int *dir) {
    const unsigned int n = min(counters[LF_CNT_ALLOCATED], (unsigned int)num_blocks);
    for (unsigned int b = blockIdx.x * blockDim.x + threadIdx.x; b < n; b += gridDim.x * blockDim.x) {
        const int4 p = block_pos[b];
        const int c = lf::dir_cell(p.x, p.y, p.z);
        if (c >= 0) dir[c] = (int)b;
    }
}

int lf_rebuild_dir(const lf_map_params *p, const lf_map_storage *s, hipStream_t st) {
    if (int rc = lf_check_storage(s, p)) return rc;
    const size_t n16 = (size_t)LF_DIR_CELLS * sizeof(int) / 16;
    k_fill_u128<<<(int)std::min<size_t>(8192, (n16 + 255) / 256), 256, 0, st>>>(
        reinterpret_cast<uint4 *>(s->block_dir), n16, 0xFFFFFFFFu);
    LF_LAUNCH_CHECK();
    k_dir_fill<<<std::max(1, std::min(2048, (p->num_blocks + 255) / 256)), 256, 0, st>>>(
        reinterpret_cast<const int4 *>(s->block_pos), s->counters, p->num_blocks, s->block_dir);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" int lf_weight_code(float w) { return (int)lf::weight_code(w); }

extern "C" int lf_weight_code_dither(float w, unsigned int rnd) { return (int)lf::weight_code_sr(w, rnd); }

extern "C" float lf_weight_value(int code) { return lf::weight_value((unsigned int)code & 0xFFu); }

extern "C" int lf_map_read_counters(const lf_map_storage *s, unsigned int *out16, void *stream) {
    if (int rc = lf_check_storage(s)) return rc;
    if (!out16) return lf_set_error(LF_ERR_ARG, "null counter output");
    hipStream_t st = lf_stream(stream);
    LF_HIP_CHECK(hipMemcpyAsync(out16, s->counters, LF_NUM_COUNTERS * sizeof(unsigned int),
                                hipMemcpyDeviceToHost, st));
    LF_HIP_CHECK(hipStreamSynchronize(st));
    return LF_OK;
}
