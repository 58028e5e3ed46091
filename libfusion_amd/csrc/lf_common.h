// Shared device helpers of the gfx950 kernels.
//
// Float arithmetic is written in the operation order of libfusion's POD math
// (include/math/vector_type.h, include/math/matrix_type.h) and every HIP file
// is compiled with -ffp-contract=off, so each product and sum is rounded on its
// own exactly as on the CPU: block coordinates, projections and visibility
// tests agree bit for bit with a CPU restatement on the same inputs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <climits>
#include <type_traits>

#include "libfusion_amd.h"

#define LF_DEV __device__ __forceinline__
#define LF_HD __host__ __device__ __forceinline__

namespace lf {

constexpr unsigned long long kEmptyKey = ~0ull;
constexpr int kBad = INT_MIN;
constexpr int kWave = 64;
constexpr int kXcds = 8;  // MI355X: 8 XCDs, each with its own L2; workgroups are dealt to them in turn

struct f3 { float x, y, z; };
LF_HD f3 add(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
LF_HD f3 sub(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
LF_HD f3 mul(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
LF_HD f3 dvd(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
LF_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
LF_HD f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
LF_HD float norm(f3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
LF_HD f3 normalised(f3 a) { return dvd(a, norm(a)); }

// Matrix3x4 applied to a point / rotated (matrix_type.h:86-98).
LF_HD f3 xform(const float *m, f3 v) {
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3], m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7],
            m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11]};
}
LF_HD f3 rotate(const float *m, f3 v) {
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
            m[8] * v.x + m[9] * v.y + m[10] * v.z};
}

struct i3 { int x, y, z; };

// ToVector3i (vector_type.h:218-225) with NaN / out-of-range sent to kBad:
// floor, one range test on |floor(v)| (false for NaN), one conversion.  Equal
// to the truncate-and-adjust form (b = (int)v; b - ((float)b > v)) for every
// float inside the range, checked over all 2^32 bit patterns.
LF_HD int floor_i(float v) {
    const float f = floorf(v);
    return fabsf(f) < 2147483520.0f ? (int)f : kBad;
}
LF_HD i3 to_i3(f3 v) { return {floor_i(v.x), floor_i(v.y), floor_i(v.z)}; }
// __float2int_rd of a projected coordinate; non-finite / huge values leave every image.
LF_HD int proj_floor(float t) {
    if (!(t > -1e9f && t < 1e9f)) return kBad;
    return (int)floorf(t);
}
// voxelPosToBlockPos (map_struct.cu:626-636): floor division by 8.
LF_HD int blk(int v) { return v == kBad ? kBad : (v >> 3); }
LF_HD int lidx(int x, int y, int z) { return ((z & 7) << 6) | ((y & 7) << 3) | (x & 7); }

// computeHash (map_struct.cu:143-148).
LF_HD int compute_hash(int x, int y, int z, int n) {
    uint32_t a = (uint32_t)x * 73856093u, b = (uint32_t)y * 19349669u, c = (uint32_t)z * 83492791u;
    int res = (int)(a ^ b ^ c);
    res %= n;
    return res < 0 ? res + n : res;
}

// Block coordinate packed into 63 bits (21 bits per axis, two's complement).
LF_HD unsigned long long pack_key(int x, int y, int z) {
    return ((unsigned long long)(x & 0x1FFFFF) << 42) | ((unsigned long long)(y & 0x1FFFFF) << 21) |
           (unsigned long long)(z & 0x1FFFFF);
}

// sdf code of a voxel whose weight is 0 (never observed, or only with weight 0).
// packFloat never produces it (|v| <= 1 maps into [-32767, 32767]), so the
// raycast decides validity from the 2-byte sdf alone, without the weight.
constexpr short kSdfUnseen = -32768;

// v / 32767 correctly rounded, as a product with the rounded reciprocal and one
// fused correction from the exact residual: bit-identical to the IEEE
// division for all 65536 inputs (checked exhaustively), in 3 instructions
// instead of the division's 11.
LF_HD float unpack_sdf(short v) {
    const float x = (float)v, r = 1.0f / 32767.0f;
    const float q = x * r;
    return fmaf(fmaf(-q, 32767.0f, x), r, q);
}
LF_HD short pack_sdf(float v) { return (short)(v * 32767.0f); }

// Mixed-precision voxels: FP16 sdf in truncation units, kHalfUnseen (a NaN)
// where the weight is 0, so one 2-byte load decides validity and value.
constexpr unsigned short kHalfUnseen = 0xFFFF;
LF_HD float half_bits_to_float(unsigned short h) { return (float)__builtin_bit_cast(_Float16, h); }
LF_HD unsigned short float_to_half_bits(float v) { return __builtin_bit_cast(unsigned short, (_Float16)v); }

// 8-bit weight code: an unsigned float with 4 exponent bits (bias 7) and 4
// mantissa bits.  Code 0 is 0; codes 1..15 are m * 2^-10; code e<<4 | m with
// e >= 1 is (1 + m/16) * 2^(e-7), up to 496.
// For c < 256, ((e + 120) << 23) | (m << 19) is (c << 19) + (120 << 23),
// and below 16 the code is m itself.
LF_HD float weight_value(unsigned int c) {
    const float sub = (float)c * (1.0f / 1024.0f);
    const float nrm = __builtin_bit_cast(float, (c << 19) + (120u << 23));
    return c < 16u ? sub : nrm;  // both computed: a select, not a branch
}
// Nearest code of w >= 0, ties to the even code; above 496, 0xFF.  Below 2^-6
// the code is round(w * 1024) (exact scaling; 16 rounds up into code 0x10 =
// 2^-6).  Above, the float's exponent is re-biased (127 -> 7) and its 23
// mantissa bits rounded to 4 by adding just under half a unit plus the kept
// part's last bit; a carry moves into the exponent as it should.
LF_HD unsigned int weight_code(float w) {
    // every case computed and selected (no branch); r wraps for w < 2^-6,
    // whose code is the scaled one
    const unsigned int small = (unsigned int)__builtin_rintf(fminf(fmaxf(w, 0.0f), 0.015625f) * 1024.0f);
    const unsigned int r = __builtin_bit_cast(unsigned int, w) - (120u << 23);
    const unsigned int c = (r + 0x3FFFFu + ((r >> 19) & 1u)) >> 19;
    const unsigned int big = c > 0xFFu ? 0xFFu : c;
    return !(w > 0.0f) ? 0u : w < 0.015625f ? small : big;
}

// A mixed map's voxel word (its weight array): the 8-bit weight code in bits
// 0-7 and the colour R, G, B in bits 8-15, 16-23 and 24-31, so a voxel's
// weight and colour are one 4-byte access (6 bytes per voxel with the FP16 sdf).
LF_HD unsigned int word_code(unsigned int word) { return word & 0xFFu; }
LF_HD unsigned int word_channel(unsigned int word, int k) { return (word >> (8 + 8 * k)) & 0xFFu; }
LF_HD unsigned int word_colour_bits(unsigned int r, unsigned int g, unsigned int b) {
    return (r << 8) | (g << 16) | (b << 24);
}

// The same code rounded stochastically: up to the next code with probability
// equal to w's fractional position between its two neighbouring codes, so
// the stored weight equals w in expectation.  rnd supplies the random bits
// (its low 19 bits dither the normal range, its top 24 the range below 2^-6).
// A running sum of increments far below half a code step would stall at a
// fraction of its true value under round-to-nearest; this keeps it unbiased.
// Callers clamp w to 255 first, so the result stays below 0xF1.
LF_HD unsigned int weight_code_sr(float w, unsigned int rnd) {
    const float d = (float)(rnd >> 8) * (1.0f / 16777216.0f);  // [0, 1)
    const unsigned int small = (unsigned int)floorf(fminf(fmaxf(w, 0.0f), 0.015625f) * 1024.0f + d);
    const unsigned int r = __builtin_bit_cast(unsigned int, w) - (120u << 23);
    const unsigned int c = (r + (rnd & 0x7FFFFu)) >> 19;
    const unsigned int big = c > 0xFFu ? 0xFFu : c;
    return !(w > 0.0f) ? 0u : w < 0.015625f ? small : big;
}

// 32 random bits of (voxel, fusion pass): an integer hash (xor-shift /
// multiply finaliser) of the voxel's block coordinate, its index in the block
// and the map's pass counter, so every voxel and every pass dithers
// independently.  The voxel is named by where it is, not by its pool slot
// (which the allocation race decides), so a map's values do not depend on
// the order its blocks were allocated in.
constexpr unsigned int kRandLocalMul = 0x27D4EB2Fu;
// voxel_rand with its voxel term local * kRandLocalMul given (mod 2^32,
// (local + (pass << 9)) * k = local * k + (pass << 9) * k, so a caller that
// visits many blocks computes each lane's product once).
LF_HD unsigned int voxel_rand_mixed(int bx, int by, int bz, unsigned int local_mul, unsigned int pass) {
    unsigned int x = (unsigned int)bx * 0x9E3779B1u ^ (unsigned int)by * 0x85EBCA77u ^
                     (unsigned int)bz * 0xC2B2AE3Du ^ (local_mul + (pass << 9) * kRandLocalMul);
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}
LF_HD unsigned int voxel_rand(int bx, int by, int bz, unsigned int local, unsigned int pass) {
    return voxel_rand_mixed(bx, by, bz, local * kRandLocalMul, pass);
}

// *(T *)((char *)base + off): a 32-bit byte offset from a (scalar) base, so
// the access takes the offset in one VGPR instead of 64-bit address math.
template <class T> LF_DEV T &at_byte(T *base, unsigned int off) {
    using C = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return *reinterpret_cast<T *>(reinterpret_cast<C *>(base) + off);
}

// A 32-bit byte offset held in a vector register at this point of the code
// (an empty asm statement, no instruction): the compiler can neither hoist it
// out of a loop nor widen it to 64 bits early, so a load or store at_byte(p,
// off) with a uniform p keeps the scalar-base + 32-bit-offset address form
// instead of a 64-bit vector add per access.
LF_DEV unsigned int vgpr_offset(unsigned int v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Occupancy bit of a block: its 4x4x4-block region (rx, ry, rz) = (bx, by, bz)
// >> 2 wrapped onto a 64 x 64 x 32 torus, bit (rz mod 32) * 4096 + (ry mod
// 64) * 64 + (rx mod 64), with world z (the vertical) on the short axis.
// Regions less than 64 regions apart in x and y and 32 in z never share a bit
// (at 5 mm voxels a 10.24 x 10.24 x 5.12 m box), and the bit costs three
// bit-field extracts and two shift-ors (a multiplicative hash: three
// quarter-rate 32-bit multiplies).
LF_HD unsigned int occupancy_bit(int bx, int by, int bz) {
    static_assert(LF_OCC_WORDS * 32 == (1 << 17), "the torus has 2^17 regions");
    const uint32_t rx = (uint32_t)(bx >> 2) & 63u, ry = (uint32_t)(by >> 2) & 63u, rz = (uint32_t)(bz >> 2) & 31u;
    return (rz << 12) | (ry << 6) | rx;
}

// Block directory cell (4x4x4-block bricks of 64 consecutive cells, bricks in
// x-fastest order), or -1 for a block outside the directory.
LF_HD int dir_cell(int bx, int by, int bz) {
    constexpr int h = LF_DIR_DIM / 2, nb = LF_DIR_DIM / 4;
    const unsigned int x = (unsigned int)bx + h, y = (unsigned int)by + h, z = (unsigned int)bz + h;
    if (x >= (unsigned int)LF_DIR_DIM || y >= (unsigned int)LF_DIR_DIM || z >= (unsigned int)LF_DIR_DIM) return -1;
    const unsigned int brick = ((z >> 2) * nb + (y >> 2)) * nb + (x >> 2);
    return (int)((brick << 6) | ((z & 3u) << 4) | ((y & 3u) << 2) | (x & 3u));
}

// Linear-probing lookup from the computeHash home bucket; returns the pool
// block of (bx, by, bz) or -1.  The probe is bounded by the table size.  A
// slot's key and block index are loaded together (one memory round trip per
// probed slot; most lookups end at the home slot).
LF_DEV int find_block(const unsigned long long *__restrict__ keys, const int *__restrict__ slot_block,
                      int bx, int by, int bz, int num_buckets, int num_entries) {
    if (bx == kBad || by == kBad || bz == kBad) return -1;
    const unsigned long long key = pack_key(bx, by, bz);
    int slot = compute_hash(bx, by, bz, num_buckets);
    for (int i = 0; i < num_entries; ++i) {
        const unsigned long long k = keys[slot];
        const int b = slot_block[slot];
        if (k == key) return b;
        if (k == kEmptyKey) return -1;
        if (++slot == num_entries) slot = 0;
    }
    return -1;
}

LF_DEV int lane_id() { return __lane_id(); }

// A device pointer asserted to address global memory.  Pointers read from a
// descriptor table in memory are generic, and every access through them
// compiles to a flat_ instruction (which also waits on the LDS / scalar
// counter); through the result they compile to global_ instructions, as
// kernel arguments do.
template <class T>
LF_DEV T *global_ptr(T *p) {
    return (T *)((__attribute__((address_space(1))) T *)reinterpret_cast<uintptr_t>(p));
}

}  // namespace lf

// Error reporting (common.hip).
int lf_set_error(int code, const char *fmt, ...);
// LF_OK when every array of the storage is present (common.hip); the colour
// planes are required only when p is given and the map is FP32.
int lf_check_storage(const lf_map_storage *s, const lf_map_params *p = nullptr);
// Bytes of one voxel's weight word (common.hip): float32, or a mixed map's
// 8-bit weight code with its colour packed beside it.
size_t lf_weight_bytes(const lf_map_params *p);
// Block directory rebuilt from the pool (clear, then every allocated block's
// cell), for storage whose pool was written directly (common.hip).
int lf_rebuild_dir(const lf_map_params *p, const lf_map_storage *s, hipStream_t st);
#define LF_HIP_CHECK(call)                                                               \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return lf_set_error(LF_ERR_HIP, "%s failed: %s (%s:%d)", #call,              \
                                hipGetErrorString(e_), __FILE__, __LINE__);              \
    } while (0)
#define LF_LAUNCH_CHECK() LF_HIP_CHECK(hipGetLastError())

static inline hipStream_t lf_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }
