// Map checkpoint / resume in the reference's on-disk layout, and map copies.
//
// Reference: MapStruct::writeToDisk / readFromDisk (src/data_struct/map_struct.cu:529-613),
// DenseMapping::writeMapToDisk / readMapFromDisk (src/voxel_hashing/voxel_hashing.cpp:189-205),
// MapStruct::copyTo (map_struct.cu:451-469).
//
// The reference file is its host arrays written back to back:
//   voxels     num_blocks * 512 records of 12 bytes {int16 sdf, 2 pad, float32 weight, uint8 rgb[3], 1 pad}
//   hash table num_hash_entries records of 20 bytes {int32 ptr (voxel offset = block * 512, -1 = free),
//              int32 offset (1-based link into the excess area, <= 0 = end of chain), int32 pos[3]}
//   heap       num_blocks int32 (free-block stack), then num_buckets int32 bucket mutexes,
//              then int32 heap counter and int32 excess counter
// and "<file>.txt" holds the raw MapSize {int32 num_blocks, num_hash_entries, num_buckets}.
//
// This library's map is an open-addressing table with a dense block pool and
// planar voxel arrays, so writing converts: blocks keep their pool index (the
// reference's heap hands out blocks in allocation order too: heap[i] = B-1-i,
// taken from the top), and the chained table is rebuilt by inserting the
// blocks in pool order the way createBlock (map_struct.cu:280-325) does on an
// uncontended table.  Reading accepts any file whose live entries (ptr != -1)
// name distinct blocks: it compacts them in pool order and rebuilds this
// library's table.  The sdf of a weight-0 voxel is not observable in the
// reference; it is written as 0 and read as the unseen code.  A mixed-
// precision map converts on the way: FP16 sdf to the reference's int16
// (packFloat's truncation) and 8-bit weight codes to float32, and back to the
// nearest FP16 and nearest code on reading.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

#include "lf_common.h"

namespace {

struct RefVoxel {
    int16_t sdf;
    int16_t pad0;
    float weight;
    uint8_t rgb[3];
    uint8_t pad1;
};
static_assert(sizeof(RefVoxel) == 12, "reference Voxel record");

struct RefEntry {
    int32_t ptr, offset, x, y, z;
};
static_assert(sizeof(RefEntry) == 20, "reference HashEntry record");

struct RefSize {
    int32_t num_blocks, num_hash_entries, num_buckets;
};

struct HostMap {
    int n = 0;
    std::vector<int> pos;  // 4 per block
    std::vector<short> sdf;
    std::vector<float> weight;
    std::vector<unsigned char> rgb;
};

int download(const lf_map_params *p, const lf_map_storage *s, hipStream_t st, HostMap &h) {
    unsigned int c[LF_NUM_COUNTERS];
    LF_HIP_CHECK(hipMemcpyAsync(c, s->counters, sizeof(c), hipMemcpyDeviceToHost, st));
    LF_HIP_CHECK(hipStreamSynchronize(st));
    h.n = (int)std::min<unsigned int>(c[LF_CNT_ALLOCATED], (unsigned int)p->num_blocks);
    const size_t n = (size_t)h.n;
    h.pos.resize(4 * n);
    h.sdf.resize(n * LF_BLOCK3);
    h.weight.resize(n * LF_BLOCK3);
    h.rgb.resize(3 * n * LF_BLOCK3);
    const bool mixed = p->precision == LF_PRECISION_MIXED;
    std::vector<unsigned int> words(mixed ? n * LF_BLOCK3 : 0);
    if (n) {
        LF_HIP_CHECK(hipMemcpyAsync(h.pos.data(), s->block_pos, 4 * n * sizeof(int), hipMemcpyDeviceToHost, st));
        LF_HIP_CHECK(hipMemcpyAsync(h.sdf.data(), s->sdf, n * LF_BLOCK3 * sizeof(short), hipMemcpyDeviceToHost, st));
        if (mixed) {
            LF_HIP_CHECK(hipMemcpyAsync(words.data(), s->weight, n * LF_BLOCK3 * sizeof(unsigned int),
                                        hipMemcpyDeviceToHost, st));
        } else {
            LF_HIP_CHECK(hipMemcpyAsync(h.weight.data(), s->weight, n * LF_BLOCK3 * sizeof(float), hipMemcpyDeviceToHost, st));
            LF_HIP_CHECK(hipMemcpyAsync(h.rgb.data(), s->rgb, 3 * n * LF_BLOCK3, hipMemcpyDeviceToHost, st));
        }
    }
    LF_HIP_CHECK(hipStreamSynchronize(st));
    if (mixed)  // to the FP32 layout: int16 sdf (unseen kept), float weights, colour planes
        for (size_t v = 0; v < n * LF_BLOCK3; ++v) {
            h.weight[v] = lf::weight_value(lf::word_code(words[v]));
            const unsigned short hb = (unsigned short)h.sdf[v];
            h.sdf[v] = hb == lf::kHalfUnseen ? lf::kSdfUnseen : lf::pack_sdf(lf::half_bits_to_float(hb));
            const size_t b = v / LF_BLOCK3, i = v % LF_BLOCK3;
            for (int k = 0; k < 3; ++k)
                h.rgb[b * 3 * LF_BLOCK3 + (size_t)k * LF_BLOCK3 + i] = (unsigned char)lf::word_channel(words[v], k);
        }
    return LF_OK;
}

struct File {
    FILE *f = nullptr;
    explicit File(const std::string &path, const char *mode) { f = std::fopen(path.c_str(), mode); }
    ~File() { if (f) std::fclose(f); }
};

}  // namespace

// DenseMapping::writeMapToDisk: the reference layout, sized by the map's
// capacities (num_blocks, num_entries, num_buckets), plus "<path>.txt".
extern "C" int lf_map_write(const lf_map_params *p, const lf_map_storage *s, const char *path, void *stream) {
    if (!p || !s || !path) return lf_set_error(LF_ERR_ARG, "null argument to lf_map_write");
    if (int rc = lf_check_storage(s, p)) return rc;
    hipStream_t st = lf_stream(stream);
    HostMap h;
    if (int rc = download(p, s, st, h)) return rc;
    const int B = p->num_blocks, E = p->num_entries, NB = p->num_buckets, X = E - NB;
    // chained table: heads at computeHash, collisions appended to the excess area
    std::vector<RefEntry> table((size_t)E, RefEntry{-1, -1, 0, 0, 0});
    int excess_used = 0;
    for (int b = 0; b < h.n; ++b) {
        const int x = h.pos[4 * b], y = h.pos[4 * b + 1], z = h.pos[4 * b + 2];
        const int bucket = lf::compute_hash(x, y, z, NB);
        RefEntry *cur = &table[bucket];
        if (cur->ptr == -1) {
            *cur = RefEntry{b * LF_BLOCK3, cur->offset, x, y, z};
            continue;
        }
        while (cur->offset > 0) cur = &table[(size_t)NB + cur->offset - 1];
        if (excess_used >= X)
            return lf_set_error(LF_ERR_CAPACITY, "the reference table's %d excess entries cannot hold the map's "
                                "bucket collisions", X);
        table[(size_t)NB + excess_used] = RefEntry{b * LF_BLOCK3, 0, x, y, z};
        cur->offset = ++excess_used;
    }
    File f(path, "wb");
    if (!f.f) return lf_set_error(LF_ERR_ARG, "cannot open %s for writing", path);
    // voxels, streamed a block at a time (unused pool blocks are zero records)
    std::vector<RefVoxel> blk(LF_BLOCK3);
    for (int b = 0; b < B; ++b) {
        std::memset(blk.data(), 0, blk.size() * sizeof(RefVoxel));
        if (b < h.n) {
            for (int i = 0; i < LF_BLOCK3; ++i) {
                const size_t v = (size_t)b * LF_BLOCK3 + i;
                RefVoxel &r = blk[i];
                r.weight = h.weight[v];
                r.sdf = (r.weight == 0.0f || h.sdf[v] == lf::kSdfUnseen) ? (int16_t)0 : h.sdf[v];
                for (int k = 0; k < 3; ++k) r.rgb[k] = h.rgb[(size_t)b * 3 * LF_BLOCK3 + (size_t)k * LF_BLOCK3 + i];
            }
        }
        if (std::fwrite(blk.data(), sizeof(RefVoxel), LF_BLOCK3, f.f) != (size_t)LF_BLOCK3)
            return lf_set_error(LF_ERR_ARG, "write failed on %s", path);
    }
    if (std::fwrite(table.data(), sizeof(RefEntry), table.size(), f.f) != table.size())
        return lf_set_error(LF_ERR_ARG, "write failed on %s", path);
    // heap as the reference leaves it after h.n allocations from a reset map
    std::vector<int32_t> ints((size_t)B);
    for (int i = 0; i < B; ++i) ints[i] = B - i - 1;
    if (std::fwrite(ints.data(), sizeof(int32_t), ints.size(), f.f) != ints.size())
        return lf_set_error(LF_ERR_ARG, "write failed on %s", path);
    ints.assign((size_t)NB, 0);
    if (std::fwrite(ints.data(), sizeof(int32_t), ints.size(), f.f) != ints.size())
        return lf_set_error(LF_ERR_ARG, "write failed on %s", path);
    // the next excess entry createBlock takes is NB + counter - 1
    const int32_t counters[2] = {B - 1 - h.n, excess_used ? excess_used + 1 : 0};
    if (std::fwrite(counters, sizeof(int32_t), 2, f.f) != 2) return lf_set_error(LF_ERR_ARG, "write failed on %s", path);
    File t(std::string(path) + ".txt", "wb");
    if (!t.f) return lf_set_error(LF_ERR_ARG, "cannot open %s.txt for writing", path);
    const RefSize sz{B, E, NB};
    if (std::fwrite(&sz, sizeof(sz), 1, t.f) != 1) return lf_set_error(LF_ERR_ARG, "write failed on %s.txt", path);
    return LF_OK;
}

// DenseMapping::readMapFromDisk: the file's sizes come from "<path>.txt" (or
// the map's capacities when it is absent, as the reference reads with its own
// state); the map is reset and refilled.
extern "C" int lf_map_read(const lf_map_params *p, const lf_map_storage *s, const char *path, void *stream) {
    if (!p || !s || !path) return lf_set_error(LF_ERR_ARG, "null argument to lf_map_read");
    if (int rc = lf_check_storage(s, p)) return rc;
    hipStream_t st = lf_stream(stream);
    RefSize sz{p->num_blocks, p->num_entries, p->num_buckets};
    {
        File t(std::string(path) + ".txt", "rb");
        if (t.f && std::fread(&sz, sizeof(sz), 1, t.f) != 1)
            return lf_set_error(LF_ERR_ARG, "short size record in %s.txt", path);
    }
    if (sz.num_blocks <= 0 || sz.num_hash_entries <= 0 || sz.num_buckets <= 0 || sz.num_buckets > sz.num_hash_entries)
        return lf_set_error(LF_ERR_ARG, "bad sizes in %s.txt", path);
    File f(path, "rb");
    if (!f.f) return lf_set_error(LF_ERR_ARG, "cannot open %s", path);
    const long long expect = (long long)sz.num_blocks * LF_BLOCK3 * 12 + (long long)sz.num_hash_entries * 20 +
                             4LL * (sz.num_blocks + sz.num_buckets + 2);
    std::fseek(f.f, 0, SEEK_END);
    const long long len = std::ftell(f.f);
    if (len != expect)
        return lf_set_error(LF_ERR_ARG, "%s holds %lld bytes; its sizes (%d blocks, %d entries, %d buckets) need %lld",
                            path, len, sz.num_blocks, sz.num_hash_entries, sz.num_buckets, expect);
    // live entries, compacted in pool (ptr) order
    std::vector<RefEntry> table((size_t)sz.num_hash_entries);
    std::fseek(f.f, (long)((long long)sz.num_blocks * LF_BLOCK3 * 12), SEEK_SET);
    if (std::fread(table.data(), sizeof(RefEntry), table.size(), f.f) != table.size())
        return lf_set_error(LF_ERR_ARG, "short hash table in %s", path);
    std::vector<RefEntry> live;
    for (const RefEntry &e : table) {
        if (e.ptr == -1) continue;
        if (e.ptr < 0 || e.ptr % LF_BLOCK3 || e.ptr / LF_BLOCK3 >= sz.num_blocks)
            return lf_set_error(LF_ERR_ARG, "entry with voxel offset %d outside the %d-block pool", e.ptr, sz.num_blocks);
        live.push_back(e);
    }
    std::sort(live.begin(), live.end(), [](const RefEntry &a, const RefEntry &b) { return a.ptr < b.ptr; });
    for (size_t i = 1; i < live.size(); ++i)
        if (live[i].ptr == live[i - 1].ptr)
            return lf_set_error(LF_ERR_ARG, "two entries share pool block %d", live[i].ptr / LF_BLOCK3);
    const int n = (int)live.size();
    if (n > p->num_blocks) return lf_set_error(LF_ERR_CAPACITY, "%d blocks exceed the pool of %d", n, p->num_blocks);
    // this library's table, filled in pool order
    const int E = p->num_entries;
    std::vector<unsigned long long> keys((size_t)E, lf::kEmptyKey);
    std::vector<int> slot_block((size_t)E, -1);
    std::vector<int> pos(4 * (size_t)n);
    std::vector<unsigned int> occ(LF_OCC_WORDS, 0u);
    for (int b = 0; b < n; ++b) {
        const RefEntry &e = live[b];
        const unsigned long long key = lf::pack_key(e.x, e.y, e.z);
        int slot = lf::compute_hash(e.x, e.y, e.z, p->num_buckets), probes = 0;
        while (keys[slot] != lf::kEmptyKey && keys[slot] != key && probes < E) {
            if (++slot == E) slot = 0;
            ++probes;
        }
        if (probes == E) return lf_set_error(LF_ERR_CAPACITY, "hash table of %d entries is full", E);
        if (keys[slot] == key) return lf_set_error(LF_ERR_ARG, "block (%d, %d, %d) appears twice", e.x, e.y, e.z);
        keys[slot] = key;
        slot_block[slot] = b;
        pos[4 * b] = e.x; pos[4 * b + 1] = e.y; pos[4 * b + 2] = e.z; pos[4 * b + 3] = slot;
        const unsigned int ob = lf::occupancy_bit(e.x, e.y, e.z);
        occ[ob >> 5] |= 1u << (ob & 31);
    }
    // voxels of the live blocks
    std::vector<short> sdf((size_t)n * LF_BLOCK3);
    std::vector<float> weight((size_t)n * LF_BLOCK3);
    std::vector<unsigned char> rgb((size_t)n * 3 * LF_BLOCK3);
    std::vector<RefVoxel> blk(LF_BLOCK3);
    for (int b = 0; b < n; ++b) {
        std::fseek(f.f, (long)((long long)live[b].ptr * 12), SEEK_SET);
        if (std::fread(blk.data(), sizeof(RefVoxel), LF_BLOCK3, f.f) != (size_t)LF_BLOCK3)
            return lf_set_error(LF_ERR_ARG, "short voxel block in %s", path);
        for (int i = 0; i < LF_BLOCK3; ++i) {
            const size_t v = (size_t)b * LF_BLOCK3 + i;
            const RefVoxel &r = blk[i];
            weight[v] = r.weight;
            // weight 0 is the unseen code; a stored -32768 (outside packFloat's range) is clamped
            sdf[v] = r.weight == 0.0f ? lf::kSdfUnseen : std::max<short>(r.sdf, (short)-32767);
            for (int k = 0; k < 3; ++k) rgb[(size_t)b * 3 * LF_BLOCK3 + (size_t)k * LF_BLOCK3 + i] = r.rgb[k];
        }
    }
    if (int rc = lf_map_reset(p, s, stream)) return rc;
    LF_HIP_CHECK(hipMemcpyAsync(s->keys, keys.data(), keys.size() * sizeof(unsigned long long), hipMemcpyHostToDevice, st));
    LF_HIP_CHECK(hipMemcpyAsync(s->slot_block, slot_block.data(), slot_block.size() * sizeof(int), hipMemcpyHostToDevice, st));
    LF_HIP_CHECK(hipMemcpyAsync(s->occupancy, occ.data(), occ.size() * sizeof(unsigned int), hipMemcpyHostToDevice, st));
    std::vector<unsigned int> words;
    if (p->precision == LF_PRECISION_MIXED) {  // nearest FP16 sdf; weight code and colour in one word
        words.resize(weight.size());
        for (size_t v = 0; v < weight.size(); ++v) {
            const unsigned int code = lf::weight_code(std::min(weight[v], 255.0f));
            sdf[v] = (short)(code == 0 ? lf::kHalfUnseen : lf::float_to_half_bits(lf::unpack_sdf(sdf[v])));
            const size_t b = v / LF_BLOCK3, i = v % LF_BLOCK3;
            const unsigned char *c = &rgb[b * 3 * LF_BLOCK3 + i];
            words[v] = code | lf::word_colour_bits(c[0], c[LF_BLOCK3], c[2 * LF_BLOCK3]);
        }
    }
    if (n) {
        LF_HIP_CHECK(hipMemcpyAsync(s->block_pos, pos.data(), pos.size() * sizeof(int), hipMemcpyHostToDevice, st));
        LF_HIP_CHECK(hipMemcpyAsync(s->sdf, sdf.data(), sdf.size() * sizeof(short), hipMemcpyHostToDevice, st));
        if (words.empty()) {
            LF_HIP_CHECK(hipMemcpyAsync(s->weight, weight.data(), weight.size() * sizeof(float), hipMemcpyHostToDevice, st));
            LF_HIP_CHECK(hipMemcpyAsync(s->rgb, rgb.data(), rgb.size(), hipMemcpyHostToDevice, st));
        } else {
            LF_HIP_CHECK(hipMemcpyAsync(s->weight, words.data(), words.size() * sizeof(unsigned int),
                                        hipMemcpyHostToDevice, st));
        }
    }
    const unsigned int allocated = (unsigned int)n;
    LF_HIP_CHECK(hipMemcpyAsync(s->counters + LF_CNT_ALLOCATED, &allocated, sizeof(allocated), hipMemcpyHostToDevice, st));
    if (int rc = lf_rebuild_dir(p, s, st)) return rc;
    LF_HIP_CHECK(hipStreamSynchronize(st));
    return LF_OK;
}

// MapStruct::copyTo: every array of src into dst (same parameters).
extern "C" int lf_map_copy(const lf_map_params *p, const lf_map_storage *src, const lf_map_storage *dst,
                           void *stream) {
    if (!p || !src || !dst) return lf_set_error(LF_ERR_ARG, "null argument to lf_map_copy");
    if (int rc = lf_check_storage(src, p)) return rc;
    if (int rc = lf_check_storage(dst, p)) return rc;
    if (src->tiles_x != dst->tiles_x || src->tiles_y != dst->tiles_y)
        return lf_set_error(LF_ERR_ARG, "maps made for different raycast sizes");
    hipStream_t st = lf_stream(stream);
    const size_t E = (size_t)p->num_entries, B = (size_t)p->num_blocks;
    struct { void *d; const void *s; size_t n; } parts[] = {
        {dst->keys, src->keys, E * sizeof(unsigned long long)},
        {dst->slot_block, src->slot_block, E * sizeof(int)},
        {dst->slot_flag, src->slot_flag, E},
        {dst->block_pos, src->block_pos, 4 * B * sizeof(int)},
        {dst->sdf, src->sdf, B * LF_BLOCK3 * sizeof(short)},
        {dst->weight, src->weight, B * LF_BLOCK3 * lf_weight_bytes(p)},
        {dst->rgb, src->rgb, p->precision == LF_PRECISION_MIXED ? 0 : 3 * B * LF_BLOCK3},  // mixed: in the words
        {dst->visible, src->visible, B * sizeof(int)},
        {dst->visible_pos, src->visible_pos, B * 4 * sizeof(int)},
        {dst->block_dir, src->block_dir, (size_t)LF_DIR_CELLS * sizeof(int)},
        {dst->counters, src->counters, LF_NUM_COUNTERS * sizeof(unsigned int)},
        {dst->occupancy, src->occupancy, LF_OCC_WORDS * sizeof(unsigned int)},
    };
    for (auto &c : parts)
        if (c.n) LF_HIP_CHECK(hipMemcpyAsync(c.d, c.s, c.n, hipMemcpyDeviceToDevice, st));
    return LF_OK;
}
