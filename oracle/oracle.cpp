// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of libfusion's dense-reconstruction hot path, templated on
// the arithmetic type (float = the method as the GPU runs it, double = the FP64
// accuracy oracle).  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker or the
// reported CPU baseline — never as part of the product path.
//
// Every function names the reference file:line it restates (paths relative to
// the libfusion source tree).  Deliberate deviations from the reference are
// marked "DEVIATION" and listed in the design notes.
//
// Parity pins (where each function is checked independently of the GPU):
//   computeHash            known answers (tests/test_oracle.py)
//   block allocation       brute-force dense ray sampling superset + exact Bresenham
//   integrate              fronto-parallel plane closed form
//   raycast                analytic plane / sphere hit within a voxel fraction
//   marching cubes         table digests, single-cube cases, sphere watertightness
//   image ops              numpy/scipy restatements and closed forms on ramps/planes
//   reductions             FP64 numpy J^T J on synthetic correspondences
//   SE(3) / LDLT           exp/log invariants, numpy.linalg.solve
// Third-party numerics (OpenCV bilateral/resize/pyrDown, Thrust scan, Sophus
// float casts, Eigen LDLT pivoting) are "parity unpinned": the oracle follows
// the documented choices written next to each function.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../libfusion_amd/csrc/mc_tables.h"

namespace orc {

struct Params {  // layout of lf_map_params
    int num_buckets, num_entries, num_blocks, max_triangles;
    float zmin_raycast, zmax_raycast, zmin_update, zmax_update, voxel_size, truncation;
};
struct Intr {  // layout of lf_intrinsics
    int width, height;
    float fx, fy, cx, cy, invfx, invfy;
};

// Intrinsics in the arithmetic type R.  For float these are the stored floats;
// for double invfx is recomputed as 1/fx in double.
template <class R> struct K_ {
    int w, h;
    R fx, fy, cx, cy, invfx, invfy;
    explicit K_(const Intr &k) : w(k.width), h(k.height), fx(k.fx), fy(k.fy), cx(k.cx), cy(k.cy),
        invfx(sizeof(R) == 4 ? (R)k.invfx : (R)1 / (R)k.fx), invfy(sizeof(R) == 4 ? (R)k.invfy : (R)1 / (R)k.fy) {}
};

// ------------------------------------------------------------------ math
// Operation order follows include/math/vector_type.h and matrix_type.h exactly.
template <class R> struct V3 { R x, y, z; };
template <class R> inline V3<R> add(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> sub(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> mul(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }   // vector_type.h:169-172,181-184
template <class R> inline V3<R> dvd(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }   // vector_type.h:87-90
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {                                    // vector_type.h:107-110
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> inline R norm(V3<R> a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }  // :97-100
template <class R> inline V3<R> normalised(V3<R> a) { return dvd(a, norm(a)); }                      // :198-201

// Matrix3x4::operator() and rotate (matrix_type.h:86-98): row dot (x, y, z, 1).
template <class R> struct M34 {
    R m[12];
    V3<R> operator()(V3<R> v) const {
        return {m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3] * (R)1,
                m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7] * (R)1,
                m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11] * (R)1};
    }
    V3<R> rotate(V3<R> v) const {
        return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[4] * v.x + m[5] * v.y + m[6] * v.z,
                m[8] * v.x + m[9] * v.y + m[10] * v.z};
    }
};

struct I3 { int x, y, z; };
const int kBad = std::numeric_limits<int>::min();

// ToVector3i per component (vector_type.h:218-225): truncate, then step down
// when the truncation rounded up.  DEVIATION: NaN and |v| >= 2^31 (undefined
// behaviour in C++) map to kBad, a coordinate no block ever has.
template <class R> inline int floor_i(R v) {
    if (!(v > (R)-2147483520.0 && v < (R)2147483520.0)) return kBad;
    int b = (int)v;
    return (R)b > v ? b - 1 : b;
}
template <class R> inline I3 to_i3(V3<R> v) { return {floor_i(v.x), floor_i(v.y), floor_i(v.z)}; }

// __float2int_rd of a projected coordinate (depth_fusion.cu:385-386): floor,
// with non-finite / huge values sent out of any image.
template <class R> inline int proj_floor(R t) {
    if (!(t > (R)-1e9 && t < (R)1e9)) return kBad;
    return (int)std::floor(t);
}

// voxelPosToBlockPos (map_struct.cu:626-636): floor division by 8.
inline int blk(int v) { return v == kBad ? kBad : (v < 0 ? (v - 7) / 8 : v / 8); }
inline I3 voxel_to_block(I3 v) { return {blk(v.x), blk(v.y), blk(v.z)}; }
// voxelPosToLocalPos + localPosToLocalIdx (map_struct.cu:643-673).
inline int local_idx(I3 v) {
    int x = v.x % 8, y = v.y % 8, z = v.z % 8;
    if (x < 0) x += 8;
    if (y < 0) y += 8;
    if (z < 0) z += 8;
    return z * 64 + y * 8 + x;
}

// computeHash (map_struct.cu:143-148), int32 wrap-around made explicit.
inline int compute_hash(int x, int y, int z, int n) {
    uint32_t a = (uint32_t)x * 73856093u, b = (uint32_t)y * 19349669u, c = (uint32_t)z * 83492791u;
    int res = (int)(a ^ b ^ c);
    res %= n;
    return res < 0 ? res + n : res;
}

// ------------------------------------------------------------------ map
// The reference stores blocks in a bucket + excess-list table behind per-bucket
// spin locks (map_struct.cu:280-325); a lock failure drops the allocation.
// The oracle restates the *result* of createBlock — an idempotent insert of the
// block coordinate — with a plain open-addressing table, so the allocated set
// is exactly the requested set (DEVIATION from the racy drop, by design).
template <class R> struct Map {
    Params p;
    std::vector<uint64_t> tk;
    std::vector<int> tv;
    size_t mask = 0;
    std::vector<I3> pos;
    std::vector<int16_t> sdf;
    std::vector<R> weight;
    std::vector<uint8_t> rgb;
    std::vector<uint8_t> touched;
    std::vector<int> visible;
    unsigned overflow = 0;
    // instrumentation for the bytes model
    std::vector<uint8_t> ray_touched;
    // raycast instrumentation: samples, samples in unallocated blocks, trilinear
    // reads, block changes along rays, rays, hits
    long long ray_stats[6] = {0, 0, 0, 0, 0, 0};
    // per pixel: samples, samples in allocated blocks, trilinear reads
    std::vector<int> ray_px;

    explicit Map(const Params &pp) : p(pp) { reset(); }
    static uint64_t key(I3 b) {
        return ((uint64_t)(b.x & 0x1FFFFF) << 42) | ((uint64_t)(b.y & 0x1FFFFF) << 21) | (uint64_t)(b.z & 0x1FFFFF);
    }
    static size_t mix(uint64_t k) {
        k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull; k ^= k >> 33;
        return (size_t)k;
    }
    void reset() {  // MapStruct::reset (map_struct.cu:74-93)
        size_t cap = 1024;
        while (cap < (size_t)p.num_blocks * 2 + 16) cap <<= 1;
        tk.assign(cap, ~0ull);
        tv.assign(cap, -1);
        mask = cap - 1;
        pos.clear(); sdf.clear(); weight.clear(); rgb.clear(); touched.clear(); visible.clear();
        overflow = 0;
    }
    int find(I3 b) const {  // findEntry (map_struct.cu:365-381)
        if (b.x == kBad || b.y == kBad || b.z == kBad) return -1;
        uint64_t k = key(b);
        for (size_t i = mix(k) & mask;; i = (i + 1) & mask) {
            if (tk[i] == k) return tv[i];
            if (tk[i] == ~0ull) return -1;
        }
    }
    int insert(I3 b) {  // createBlock (map_struct.cu:280-325) without lock failure
        uint64_t k = key(b);
        size_t i = mix(k) & mask;
        for (;; i = (i + 1) & mask) {
            if (tk[i] == k) return tv[i];
            if (tk[i] == ~0ull) break;
        }
        if ((int)pos.size() >= p.num_blocks) { overflow |= 2u; return -1; }
        int id = (int)pos.size();
        tk[i] = k; tv[i] = id;
        pos.push_back(b);
        sdf.resize(sdf.size() + 512, 0);
        weight.resize(weight.size() + 512, (R)0);
        rgb.resize(rgb.size() + 1536, 0);
        touched.push_back(0);
        return id;
    }
};

// unpackFloat / packFloat (voxel.h:36-44).
template <class R> inline R unpack(int16_t v) { return (R)v / (R)32767; }
template <class R> inline int16_t pack(R v) { return (int16_t)(v * (R)32767); }

template <class R> inline bool isnan_(R v) { return !(v == v); }

// ------------------------------------------------ allocation (create_blocks)
// unproject (depth_fusion.cu:101-105): (invfx*(x-cx))*z.
template <class R> inline V3<R> unproject_df(int x, int y, R z, const K_<R> &K) {
    return {K.invfx * ((R)x - K.cx) * z, K.invfy * ((R)y - K.cy) * z, z};
}

// One pixel of create_blocks_kernel (depth_fusion.cu:121-228): the blocks a
// 3-D Bresenham walk visits between the near and far ends of the truncation
// band.  Returns the visited blocks through `emit`.
template <class R, class F> void band_blocks(int x, int y, R z, const M34<R> &pose, const K_<R> &K,
                                             const Params &p, F emit) {
    const R zmin = (R)p.zmin_update, zmax = (R)p.zmax_update;
    if (isnan_(z) || z < zmin || z > zmax) return;
    const R half = (R)p.truncation * (R)0.5;
    const R z_near = std::max(zmin, z - half);
    const R z_far = std::min(zmax, z + half);
    if (z_near >= z_far) return;
    const R vs = (R)p.voxel_size;
    I3 a = voxel_to_block(to_i3(dvd(pose(unproject_df(x, y, z_near, K)), vs)));
    I3 b = voxel_to_block(to_i3(dvd(pose(unproject_df(x, y, z_far, K)), vs)));
    if (a.x == kBad || a.y == kBad || a.z == kBad || b.x == kBad || b.y == kBad || b.z == kBad) return;
    I3 d = {b.x - a.x, b.y - a.y, b.z - a.z};
    I3 inc = {d.x < 0 ? -1 : 1, d.y < 0 ? -1 : 1, d.z < 0 ? -1 : 1};
    I3 ab = {std::abs(d.x), std::abs(d.y), std::abs(d.z)};
    I3 e2 = {ab.x << 1, ab.y << 1, ab.z << 1};
    int err1, err2;
    emit(a);
    if (ab.x >= ab.y && ab.x >= ab.z) {
        err1 = e2.y - 1; err2 = e2.z - 1;
        for (int i = 0; i < ab.x; ++i) {
            if (err1 > 0) { a.y += inc.y; err1 -= e2.x; }
            if (err2 > 0) { a.z += inc.z; err2 -= e2.x; }
            err1 += e2.y; err2 += e2.z; a.x += inc.x;
            emit(a);
        }
    } else if (ab.y >= ab.x && ab.y >= ab.z) {
        err1 = e2.x - 1; err2 = e2.z - 1;
        for (int i = 0; i < ab.y; ++i) {
            if (err1 > 0) { a.x += inc.x; err1 -= e2.y; }
            if (err2 > 0) { a.z += inc.z; err2 -= e2.y; }
            err1 += e2.x; err2 += e2.z; a.y += inc.y;
            emit(a);
        }
    } else {
        err1 = e2.y - 1; err2 = e2.x - 1;
        for (int i = 0; i < ab.z; ++i) {
            if (err1 > 0) { a.y += inc.y; err1 -= e2.z; }
            if (err2 > 0) { a.x += inc.x; err2 -= e2.z; }
            err1 += e2.y; err2 += e2.x; a.z += inc.z;
            emit(a);
        }
    }
}

template <class R> void allocate(Map<R> &m, const R *depth, int cols, int rows, const M34<R> &pose, const K_<R> &K) {
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x)
            band_blocks<R>(x, y, depth[(size_t)y * cols + x], pose, K, m.p, [&](I3 b) {
                int id = m.insert(b);
                if (id >= 0) m.touched[id] = 1;
            });
}

// ----------------------------------------------------------- visibility
// is_vertex_visible / is_block_visible (depth_fusion.cu:15-47).
template <class R> bool block_visible(I3 bp, const M34<R> &inv, const K_<R> &K, int cols, int rows, const Params &p) {
    const R scale = (R)8 * (R)p.voxel_size;
    for (int c = 0; c < 8; ++c) {
        I3 t = {bp.x + (c & 1), bp.y + ((c >> 1) & 1), bp.z + ((c >> 2) & 1)};
        V3<R> w = {(R)t.x * scale, (R)t.y * scale, (R)t.z * scale};
        V3<R> pt = inv(w);
        R u = K.fx * pt.x / pt.z + K.cx, v = K.fy * pt.y / pt.z + K.cy;
        if (!(u < 0 || v < 0 || u > (R)(cols - 1) || v > (R)(rows - 1) || pt.z < (R)p.zmin_update ||
              pt.z > (R)p.zmax_update))
            return true;
    }
    return false;
}

// check_visibility_flag_kernel + scan + copy (depth_fusion.cu:49-93,549-557).
// DEVIATION: the count includes the last entry (the reference reads the
// exclusive scan's last element) and the flag buffer starts at zero.
// geometric_only: count_visible_entry (ray_tracing.cu:647-807) over the
// allocated blocks, touched flags left alone.
template <class R> void select_visible(Map<R> &m, const M34<R> &inv, const K_<R> &K, int cols, int rows,
                                       bool geometric_only = false) {
    const int n = (int)m.pos.size();
    std::vector<uint8_t> vis(n, 0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        vis[i] = (!geometric_only && m.touched[i]) ? 1 : (block_visible(m.pos[i], inv, K, cols, rows, m.p) ? 1 : 0);
    m.visible.clear();
    for (int i = 0; i < n; ++i) {
        if (vis[i]) m.visible.push_back(i);
        if (!geometric_only) m.touched[i] = 0;
    }
}

// ------------------------------------------------------------ integrate
// update_map_weighted_kernel (depth_fusion.cu:356-428) when nmap != null,
// update_map_with_colour_kernel (:288-354) when nmap == null.
template <class R> void integrate(Map<R> &m, const R *depth, const R *nmap, const uint8_t *rgb, int cols,
                                  int rows, const M34<R> &inv, const K_<R> &K) {
    const Params &p = m.p;
    const R vs = (R)p.voxel_size, trunc = (R)p.truncation, inv_trunc = (R)1 / trunc;
    const R zmin = (R)p.zmin_update, zmax = (R)p.zmax_update;
    const int nv = (int)m.visible.size();
#pragma omp parallel for schedule(dynamic, 16)
    for (int vi = 0; vi < nv; ++vi) {
        const int b = m.visible[vi];
        const I3 bp = m.pos[b];
        for (int lz = 0; lz < 8; ++lz)
            for (int ly = 0; ly < 8; ++ly)
                for (int lx = 0; lx < 8; ++lx) {
                    I3 vp = {bp.x * 8 + lx, bp.y * 8 + ly, bp.z * 8 + lz};
                    V3<R> w = {(R)vp.x * vs, (R)vp.y * vs, (R)vp.z * vs};
                    V3<R> pt = inv(w);
                    int u = proj_floor(K.fx * pt.x / pt.z + K.cx + (R)0.5);
                    int v = proj_floor(K.fy * pt.y / pt.z + K.cy + (R)0.5);
                    if (u < 0 || v < 0 || u > cols - 1 || v > rows - 1) continue;
                    const size_t pix = (size_t)v * cols + u;
                    R dist = depth[pix];
                    R nz = 0;
                    if (nmap) {
                        R nx = nmap[4 * pix];
                        nz = nmap[4 * pix + 2];
                        if (isnan_(dist) || isnan_(nx) || dist > zmax || dist < zmin) continue;
                    } else {
                        if (isnan_(dist) || (double)dist < 1e-2 || dist > zmax || dist < zmin) continue;
                    }
                    R sdf = dist - pt.z;
                    if (sdf < -trunc) continue;
                    sdf = std::min((R)1, sdf * inv_trunc);
                    const size_t vx = (size_t)b * 512 + (size_t)(lz * 64 + ly * 8 + lx);
                    R sdf_p = unpack<R>(m.sdf[vx]);
                    R w_p = m.weight[vx];
                    R wt = nmap ? std::fabs(std::sin(nz)) / (dist * dist) : (R)1 / (dist * dist);
                    uint8_t *c = &m.rgb[(size_t)b * 1536 + (size_t)(lz * 64 + ly * 8 + lx)];
                    const uint8_t cn[3] = {rgb[3 * pix], rgb[3 * pix + 1], rgb[3 * pix + 2]};
                    if (w_p == 0) {
                        m.sdf[vx] = pack(sdf);
                        m.weight[vx] = wt > (R)255 ? (R)255 : wt;
                        c[0] = cn[0]; c[512] = cn[1]; c[1024] = cn[2];
                        continue;
                    }
                    sdf_p = (sdf_p * w_p + sdf * wt) / (w_p + wt);
                    m.sdf[vx] = pack(sdf_p);
                    R wn = w_p + wt;
                    m.weight[vx] = wn > (R)255 ? (R)255 : wn;
                    for (int k = 0; k < 3; ++k) {
                        R mixed = ((R)c[512 * k] * w_p + (R)cn[k] * wt) / (w_p + wt);
                        c[512 * k] = (uint8_t)(int)mixed;  // ToVector3c (vector_type.h:213-216)
                    }
                }
    }
}

// -------------------------------------------------------------- raycast
template <class R> struct Ray {
    const Map<R> &m;
    bool instrument;
    // read_sdf (ray_tracing.cu:291-305)
    R read_sdf(V3<R> pt, bool &valid) const {
        I3 v = to_i3(pt);
        int b = m.find(voxel_to_block(v));
        if (b >= 0) {
            size_t i = (size_t)b * 512 + local_idx(v);
            if (instrument) const_cast<Map<R> &>(m).ray_touched[b] = 1;
            if (m.weight[i] != 0) { valid = true; return unpack<R>(m.sdf[i]); }
        }
        valid = false;
        return std::numeric_limits<R>::quiet_NaN();
    }
    // read_sdf_interped (ray_tracing.cu:307-339)
    R read_interp(V3<R> pt, bool &valid) const {
        V3<R> f = {std::floor(pt.x), std::floor(pt.y), std::floor(pt.z)};
        V3<R> xyz = sub(pt, f);
        R s0, s1, r0, r1, r2, r3;
        bool vp, v;
        s0 = read_sdf(pt, vp);
        s1 = read_sdf(add(pt, V3<R>{1, 0, 0}), v); vp &= v;
        r0 = ((R)1 - xyz.x) * s0 + xyz.x * s1;
        s0 = read_sdf(add(pt, V3<R>{0, 1, 0}), v); vp &= v;
        s1 = read_sdf(add(pt, V3<R>{1, 1, 0}), v); vp &= v;
        r1 = ((R)1 - xyz.x) * s0 + xyz.x * s1;
        r2 = ((R)1 - xyz.y) * r0 + xyz.y * r1;
        s0 = read_sdf(add(pt, V3<R>{0, 0, 1}), v); vp &= v;
        s1 = read_sdf(add(pt, V3<R>{1, 0, 1}), v); vp &= v;
        r0 = ((R)1 - xyz.x) * s0 + xyz.x * s1;
        s0 = read_sdf(add(pt, V3<R>{0, 1, 1}), v); vp &= v;
        s1 = read_sdf(add(pt, V3<R>{1, 1, 1}), v); vp &= v;
        r1 = ((R)1 - xyz.x) * s0 + xyz.x * s1;
        r3 = ((R)1 - xyz.y) * r0 + xyz.y * r1;
        valid = vp;
        return ((R)1 - xyz.z) * r2 + xyz.z * r3;
    }
};

// RenderingBlockDelegate::create_rendering_block (ray_tracing.cu:63-126) +
// fill_rendering_blocks (:187-210): min/max z of every visible block over its
// projected tile rectangle.  The 16x16 sub-rectangle split (:128-152) covers
// the rectangle exactly and only matters for its 100k capacity, which is not
// modelled (DEVIATION: unbounded).  Tile coordinates saturate to int16 like the
// reference's Vector2s, without its wrap-around on huge projections (DEVIATION).
template <class R> int zrange_tiles(const Map<R> &m, const M34<R> &inv, const K_<R> &K, int tx, int ty,
                                    std::vector<R> &zx, std::vector<R> &zy) {
    int made = 0;
    const Params &p = m.p;
    zx.assign((size_t)tx * ty, (R)100);
    zy.assign((size_t)tx * ty, (R)0);
    const R scale = (R)8 * (R)p.voxel_size;
    auto sat = [](R v) -> int {
        if (!(v == v)) return 0;
        if (v > (R)32767) return 32767;
        if (v < (R)-32768) return -32768;
        return (int)v;
    };
    for (int b : m.visible) {
        I3 bp = m.pos[b];
        int ulx = tx, uly = ty, lrx = -1, lry = -1;
        R zr0 = (R)p.zmax_raycast, zr1 = (R)p.zmin_raycast;
        for (int c = 0; c < 8; ++c) {
            I3 t = {bp.x + (c & 1), bp.y + ((c >> 1) & 1), bp.z + ((c >> 2) & 1)};
            V3<R> w = {(R)t.x * scale, (R)t.y * scale, (R)t.z * scale};
            V3<R> pt = inv(w);
            R u = (K.fx * pt.x / pt.z + K.cx) / (R)8, v = (K.fy * pt.y / pt.z + K.cy) / (R)8;
            R fu = std::floor(u), cu = std::ceil(u), fv = std::floor(v), cv = std::ceil(v);
            if ((R)ulx > fu) ulx = sat(fu);
            if ((R)lrx < cu) lrx = sat(cu);
            if ((R)uly > fv) uly = sat(fv);
            if ((R)lry < cv) lry = sat(cv);
            if (zr0 > pt.z) zr0 = pt.z;
            if (zr1 < pt.z) zr1 = pt.z;
        }
        if (ulx < 0) ulx = 0;
        if (uly < 0) uly = 0;
        if (lrx >= tx) lrx = tx - 1;
        if (lry >= ty) lry = ty - 1;
        if (ulx > lrx || uly > lry) continue;
        if (zr0 < (R)p.zmin_raycast) zr0 = (R)p.zmin_raycast;
        if (zr1 < (R)p.zmin_raycast) continue;
        ++made;
        for (int y = uly; y <= lry; ++y)
            for (int x = ulx; x <= lrx; ++x) {
                size_t i = (size_t)y * tx + x;
                zx[i] = std::min(zx[i], zr0);
                zy[i] = std::max(zy[i], zr1);
            }
    }
    return made;
}

// MapRenderingDelegate::operator() / raycast_with_colour (ray_tracing.cu:346-425,478-564).
template <class R> void raycast(Map<R> &m, R *vmap, uint8_t *rgb_out, int cols, int rows, const M34<R> &pose,
                                const M34<R> &inv, const K_<R> &K, int check_colour, bool instrument) {
    const Params &p = m.p;
    const int tx = cols / 8, ty = rows / 8;
    std::vector<R> zx, zy;
    // DenseMapping::raycast leaves the vertex map untouched when no block is
    // visible or no rendering block was made (voxel_hashing.cpp:92-93,105).
    if (m.visible.empty()) return;
    if (zrange_tiles(m, inv, K, tx, ty, zx, zy) == 0) return;
    if (instrument) {
        m.ray_touched.assign(m.pos.size(), 0);
        for (auto &v : m.ray_stats) v = 0;
        m.ray_px.assign((size_t)3 * cols * rows, 0);
    }
    Ray<R> ray{m, instrument};
    const R inv_vs = (R)1 / (R)p.voxel_size;
    const R step_scale = (R)p.truncation * inv_vs;
    const R vs = (R)p.voxel_size;
    const R qnan = std::numeric_limits<R>::quiet_NaN();
#pragma omp parallel for schedule(dynamic, 2) if (!instrument)
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x) {
            R *out = vmap + 4 * ((size_t)y * cols + x);
            out[0] = out[1] = out[2] = out[3] = qnan;
            if (rgb_out) rgb_out[3 * ((size_t)y * cols + x)] = rgb_out[3 * ((size_t)y * cols + x) + 1] =
                             rgb_out[3 * ((size_t)y * cols + x) + 2] = 255;
            int lx = x / 8, ly = y / 8;
            if (lx >= tx || ly >= ty) continue;  // DEVIATION: reference reads past the tile grid
            R z0 = zx[(size_t)ly * tx + lx], z1 = zy[(size_t)ly * tx + lx];
            if (z1 < (R)1e-3 || z0 < (R)1e-3 || isnan_(z0) || isnan_(z1)) continue;
            R sdf = 1, last_sdf;
            V3<R> pt = {((R)x - K.cx) * K.invfx * z0, ((R)y - K.cy) * K.invfy * z0, z0};
            R dist_s = norm(pt) * inv_vs;
            V3<R> bs = mul(pose(pt), inv_vs);
            pt = {((R)x - K.cx) * K.invfx * z1, ((R)y - K.cy) * K.invfy * z1, z1};
            R dist_e = norm(pt) * inv_vs;
            V3<R> be = mul(pose(pt), inv_vs);
            V3<R> dir = normalised(sub(be, bs));
            // DEVIATION: a ray with a coordinate beyond 1e9 voxels (or NaN) finds
            // nothing (no representable block lies there); the reference's
            // ToVector3i is undefined for such values
            const R reach = std::max({std::fabs(bs.x), std::fabs(bs.y), std::fabs(bs.z), std::fabs(be.x),
                                      std::fabs(be.y), std::fabs(be.z)});
            if (!(reach < (R)1e9) || isnan_(bs.x + bs.y + bs.z + be.x + be.y + be.z) || isnan_(dir.x) ||
                isnan_(dir.y) || isnan_(dir.z))
                continue;
            V3<R> res = bs;
            bool valid = false, found = false;
            R step;
            bool back_face = false;
            I3 lastb = {kBad, kBad, kBad};
            if (instrument) m.ray_stats[4]++;
            while (dist_s < dist_e) {
                last_sdf = sdf;
                sdf = ray.read_sdf(res, valid);
                if (instrument) {
                    I3 bb = voxel_to_block(to_i3(res));
                    m.ray_stats[0]++;
                    int *px = &m.ray_px[3 * ((size_t)y * cols + x)];
                    px[0]++;
                    if (m.find(bb) < 0) m.ray_stats[1]++;
                    else px[1]++;
                    if (bb.x != lastb.x || bb.y != lastb.y || bb.z != lastb.z) m.ray_stats[3]++;
                    lastb = bb;
                }
                if (sdf <= (R)0.5 && sdf >= (R)-0.5) {
                    sdf = ray.read_interp(res, valid);
                    if (instrument) {
                        m.ray_stats[2]++;
                        m.ray_px[3 * ((size_t)y * cols + x) + 2]++;
                    }
                }
                if (sdf <= (R)0) break;
                if (sdf >= (R)0 && last_sdf < (R)0) { back_face = true; break; }
                if (valid) step = std::max(sdf * step_scale, (R)1);
                else step = 2;
                res = add(res, mul(dir, step));
                dist_s += step;
            }
            if (back_face) continue;
            if (sdf <= (R)0) {
                step = sdf * step_scale;
                res = add(res, mul(dir, step));
                sdf = ray.read_interp(res, valid);
                step = sdf * step_scale;
                res = add(res, mul(dir, step));
                if (valid) found = true;
            }
            if (!found) continue;
            uint8_t col[3] = {255, 255, 255};
            if (check_colour) {  // read_colour (ray_tracing.cu:427-441)
                I3 v = to_i3(res);
                int b = m.find(voxel_to_block(v));
                if (b < 0) continue;
                size_t i = (size_t)b * 512 + local_idx(v);
                if (m.weight[i] == 0) continue;
                const uint8_t *c = &m.rgb[(size_t)b * 1536 + (size_t)local_idx(v)];
                col[0] = c[0]; col[1] = c[512]; col[2] = c[1024];
            }
            if (instrument) m.ray_stats[5]++;
            V3<R> q = inv(mul(res, vs));
            out[0] = q.x; out[1] = q.y; out[2] = q.z; out[3] = 1;
            if (rgb_out) {
                uint8_t *o = rgb_out + 3 * ((size_t)y * cols + x);
                o[0] = col[0]; o[1] = col[1]; o[2] = col[2];
            }
        }
}

// ------------------------------------------------------- marching cubes
// BuildVertexArray / BuildVertexAndColourArray (scene_mesh.cu:14-248,353-588).
// mode 0: vertices only, 1: + flat normals, 2: + corner colours.
template <class R> long mesh(const Map<R> &m, int mode, float *tri, float *nrm, uint8_t *col, long max_tri) {
    static const int co[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    long count = 0;
    const R factor = (R)m.p.voxel_size;
    auto interp = [](R v1, R v2) -> R {  // interpolate_sdf (scene_mesh.cu:108-117)
        if (std::fabs((double)((R)0 - v1)) < 1e-6) return 0;
        if (std::fabs((double)((R)0 - v2)) < 1e-6) return 1;
        if (std::fabs((double)(v1 - v2)) < 1e-6) return 0;
        return ((R)0 - v1) / (v2 - v1);
    };
    for (size_t b = 0; b < m.pos.size(); ++b) {
        const I3 bp = m.pos[b];
        for (int lz = 0; lz < 8; ++lz)
            for (int ly = 0; ly < 8; ++ly)
                for (int lx = 0; lx < 8; ++lx) {
                    I3 p0 = {bp.x * 8 + lx, bp.y * 8 + ly, bp.z * 8 + lz};
                    R s[8];
                    uint8_t c8[8][3];
                    bool ok = true;
                    for (int k = 0; k < 8 && ok; ++k) {
                        I3 v = {p0.x + co[k][0], p0.y + co[k][1], p0.z + co[k][2]};
                        int blk_id = m.find(voxel_to_block(v));
                        if (blk_id < 0) { ok = false; break; }
                        size_t i = (size_t)blk_id * 512 + local_idx(v);
                        R w = m.weight[i];
                        bool valid = mode == 2 ? ((double)w > 1e-3) : (w != 0);
                        if (!valid) { ok = false; break; }
                        s[k] = unpack<R>(m.sdf[i]);
                        const uint8_t *c = &m.rgb[(size_t)blk_id * 1536 + (size_t)local_idx(v)];
                        c8[k][0] = c[0]; c8[k][1] = c[512]; c8[k][2] = c[1024];
                    }
                    if (!ok) continue;
                    int ci = 0;
                    for (int k = 0; k < 8; ++k)
                        if (s[k] < 0) ci |= 1 << k;
                    const int et = lf_mc_edge_table[ci];
                    if (et == 0 || ci <= 0) continue;
                    V3<R> va[12];
                    const R px = (R)p0.x, py = (R)p0.y, pz = (R)p0.z;
                    static const int ea[12] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 1, 2, 3};
                    static const int eb[12] = {1, 2, 3, 0, 5, 6, 7, 4, 4, 5, 6, 7};
                    for (int e = 0; e < 12; ++e) {
                        if (!(et & (1 << e))) continue;
                        R t = interp(s[ea[e]], s[eb[e]]);
                        V3<R> o;
                        switch (e) {
                            case 0: o = {t, 0, 0}; break;
                            case 1: o = {1, t, 0}; break;
                            case 2: o = {(R)1 - t, 1, 0}; break;
                            case 3: o = {0, (R)1 - t, 0}; break;
                            case 4: o = {t, 0, 1}; break;
                            case 5: o = {1, t, 1}; break;
                            case 6: o = {(R)1 - t, 1, 1}; break;
                            case 7: o = {0, (R)1 - t, 1}; break;
                            case 8: o = {0, 0, t}; break;
                            case 9: o = {1, 0, t}; break;
                            case 10: o = {1, 1, t}; break;
                            default: o = {0, 1, t}; break;
                        }
                        va[e] = {px + o.x, py + o.y, pz + o.z};
                    }
                    for (int i = 0; lf_mc_tri_table[ci][i] != -1; i += 3) {
                        long id = count++;
                        if (id >= max_tri) continue;
                        V3<R> t3[3];
                        for (int k = 0; k < 3; ++k) {
                            int e = lf_mc_tri_table[ci][i + k];
                            t3[k] = mul(va[e], factor);
                            if (tri) { tri[9 * id + 3 * k] = (float)t3[k].x; tri[9 * id + 3 * k + 1] = (float)t3[k].y; tri[9 * id + 3 * k + 2] = (float)t3[k].z; }
                            if (col && mode == 2) {
                                int cidx = e < 8 ? e : e - 8;
                                for (int q = 0; q < 3; ++q) col[9 * id + 3 * k + q] = c8[cidx][q];
                            }
                        }
                        if (nrm && mode == 1) {
                            V3<R> n = normalised(cross(sub(t3[1], t3[0]), sub(t3[2], t3[0])));
                            for (int k = 0; k < 3; ++k) { nrm[9 * id + 3 * k] = (float)n.x; nrm[9 * id + 3 * k + 1] = (float)n.y; nrm[9 * id + 3 * k + 2] = (float)n.z; }
                        }
                    }
                }
    }
    return count;
}

// ------------------------------------------------------------- image ops
inline int reflect101(int i, int n) {
    if (n == 1) return 0;
    while (i < 0 || i >= n) i = i < 0 ? -i : 2 * n - 2 - i;
    return i;
}

// cv::cuda::bilateralFilter(src, dst, 5, 1, 1) (cuda_imgproc.cu:199) — parity
// unpinned (OpenCV not available).  Documented choice: 5x5 square window,
// reflect-101 border, w = exp(-(dx^2+dy^2)/2 - (v-c)^2/2), zeros included.
template <class R> void bilateral(const R *in, R *out, int w, int h) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const R c = in[(size_t)y * w + x];
            R s1 = 0, s2 = 0;
            for (int dy = -2; dy <= 2; ++dy)
                for (int dx = -2; dx <= 2; ++dx) {
                    R v = in[(size_t)reflect101(y + dy, h) * w + reflect101(x + dx, w)];
                    R d = v - c;
                    R wt = std::exp((R)(dx * dx + dy * dy) * (R)-0.5 + d * d * (R)-0.5);
                    s1 += v * wt;
                    s2 += wt;
                }
            out[(size_t)y * w + x] = s1 / s2;
        }
}

// cv::cuda::resize(x0.5, INTER_LINEAR) (cuda_imgproc.cu:204,214) — parity
// unpinned.  Documented choice: pixel-centre convention, i.e. the mean of the
// 2x2 source block, ((a + b) + (c + d)) * 0.25, NaN propagating; even sizes.
template <class R> void resize_half(const R *in, R *out, int w, int h, int ch) {
    const int ow = w / 2, oh = h / 2;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < oh; ++y)
        for (int x = 0; x < ow; ++x)
            for (int k = 0; k < ch; ++k) {
                R a = in[((size_t)(2 * y) * w + 2 * x) * ch + k], b = in[((size_t)(2 * y) * w + 2 * x + 1) * ch + k];
                R c = in[((size_t)(2 * y + 1) * w + 2 * x) * ch + k], d = in[((size_t)(2 * y + 1) * w + 2 * x + 1) * ch + k];
                out[((size_t)y * ow + x) * ch + k] = ((a + b) + (c + d)) * (R)0.25;
            }
}

// cv::cuda::pyrDown (cuda_imgproc.cu:209): 5x5 [1 4 6 4 1]^2 / 256 at even
// pixels, reflect-101 border (OpenCV's documented kernel and default border).
// Summation order: each row (t[-2]+t[2]) + 4(t[-1]+t[1]) + 6 t[0], then the
// same over the five row sums, times 1/256.
template <class R> void pyr_down(const R *in, R *out, int w, int h) {
    const int ow = (w + 1) / 2, oh = (h + 1) / 2;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < oh; ++y)
        for (int x = 0; x < ow; ++x) {
            R rs[5];
            for (int j = 0; j < 5; ++j) {
                const R *row = in + (size_t)reflect101(2 * y + j - 2, h) * w;
                R t0 = row[reflect101(2 * x - 2, w)], t1 = row[reflect101(2 * x - 1, w)], t2 = row[2 * x];
                R t3 = row[reflect101(2 * x + 1, w)], t4 = row[reflect101(2 * x + 2, w)];
                rs[j] = (t0 + t4) + (R)4 * (t1 + t3) + (R)6 * t2;
            }
            out[(size_t)y * ow + x] = ((rs[0] + rs[4]) + (R)4 * (rs[1] + rs[3]) + (R)6 * rs[2]) * ((R)1 / (R)256);
        }
}

// cvtColor(RGB2GRAY) on float (device_image.cpp:161-162): 0.299 R + 0.587 G + 0.114 B.
template <class R> void grey(const uint8_t *rgb, R *out, int w, int h) {
    for (size_t i = 0; i < (size_t)w * h; ++i)
        out[i] = (R)rgb[3 * i] * (R)0.299f + (R)rgb[3 * i + 1] * (R)0.587f + (R)rgb[3 * i + 2] * (R)0.114f;
}

// computeDerivativeK (cuda_imgproc.cu:217-234).  DEVIATION: the last row and
// column (unwritten in the reference) are 0.
template <class R> void derivative(const R *I, R *dx, R *dy, int w, int h) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            if (x >= w - 1 || y >= h - 1) { dx[i] = 0; dy[i] = 0; continue; }
            int x10 = std::max(x - 1, 0), x01 = std::min(x + 1, w), y10 = std::max(y - 1, 0), y01 = std::min(y + 1, h);
            dx[i] = (I[(size_t)y * w + x01] - I[(size_t)y * w + x10]) * (R)0.5;
            dy[i] = (I[(size_t)y01 * w + x] - I[(size_t)y10 * w + x]) * (R)0.5;
        }
}

// backProjectDepthK (cuda_imgproc.cu:249-263): 0.3 < z < 5.
template <class R> void back_project(const R *d, R *vm, int w, int h, const K_<R> &K) {
    const R qnan = std::numeric_limits<R>::quiet_NaN();
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            R z = d[i];
            if (z > (R)0.3f && z < (R)5.0f) {
                vm[4 * i] = z * ((R)x - K.cx) * K.invfx;
                vm[4 * i + 1] = z * ((R)y - K.cy) * K.invfy;
                vm[4 * i + 2] = z;
                vm[4 * i + 3] = 1;
            } else {
                vm[4 * i] = vm[4 * i + 1] = vm[4 * i + 2] = qnan;
                vm[4 * i + 3] = -1;
            }
        }
}

// computeNMapK (cuda_imgproc.cu:276-294).  DEVIATION: last row and column NaN.
template <class R> void compute_nmap(const R *vm, R *nm, int w, int h) {
    const R qnan = std::numeric_limits<R>::quiet_NaN();
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            if (x >= w - 1 || y >= h - 1) { nm[4 * i] = nm[4 * i + 1] = nm[4 * i + 2] = nm[4 * i + 3] = qnan; continue; }
            int x10 = std::max(x - 1, 0), x01 = std::min(x + 1, w), y10 = std::max(y - 1, 0), y01 = std::min(y + 1, h);
            auto at = [&](int xx, int yy) { const R *q = vm + 4 * ((size_t)yy * w + xx); return V3<R>{q[0], q[1], q[2]}; };
            V3<R> n = normalised(cross(sub(at(x01, y), at(x10, y)), sub(at(x, y01), at(x, y10))));
            nm[4 * i] = n.x; nm[4 * i + 1] = n.y; nm[4 * i + 2] = n.z; nm[4 * i + 3] = 1;
        }
}

// ------------------------------------------------------------ render / warp
// renderPoint (cuda_imgproc.cu:17-48): Phong shading, light at light_pos in the
// camera frame; pixel = (ambient, diffuse, specular) colour.  Float arithmetic
// in the reference's order; saturate-then-truncate to uint8.
inline void render_point(const float *p, const float *nrm, const float *pixel, uint8_t *out) {
    float c[3] = {4.f / 255.f, 2.f / 255.f, 2.f / 255.f};
    if (!(p[0] != p[0])) {
        const float Ka = 0.3f, Kd = 0.5f, Ks = 0.2f, n = 20.f, Lx = 1.f;
        const float Ax = pixel[0], Dx = pixel[1], Sx = pixel[2];
        V3<float> P{p[0], p[1], p[2]}, N{nrm[0], nrm[1], nrm[2]}, light{5.f, 5.f, 5.f}, zero{0.f, 0.f, 0.f};
        V3<float> L = normalised(sub(light, P));
        V3<float> V = normalised(sub(zero, P));
        V3<float> Rv = normalised(sub(mul(mul(N, 2.0f), dot(N, L)), L));
        const float Ix = Ax * Ka * Dx + Lx * Kd * Dx * std::max(0.f, dot(N, L)) +
                         Lx * Ks * Sx * std::pow(std::max(0.f, dot(Rv, V)), n);
        c[0] = c[1] = c[2] = Ix;
    }
    for (int k = 0; k < 3; ++k) {
        float v = c[k];
        v = v != v ? 0.f : std::min(1.f, std::max(0.f, v));  // __saturatef
        out[k] = (uint8_t)(v * 255.f);
    }
    out[3] = 255;
}

// renderSceneK / renderSceneTexturedK (cuda_imgproc.cu:50-108); rgb null: plain.
inline void render_scene(const float *vm, const float *nm, const uint8_t *rgb, int w, int h, uint8_t *out) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < w * h; ++i) {
        float pixel[3] = {1.f, 1.f, 1.f};
        if (rgb)
            for (int k = 0; k < 3; ++k) pixel[k] = (float)rgb[3 * i + k] / 255.f;
        render_point(vm + 4 * (size_t)i, nm + 4 * (size_t)i, pixel, out + 4 * (size_t)i);
    }
}

// interpolate_bilinear + warp_image_kernel (cuda_imgproc.cu:155-185).
inline void warp_image(const uint8_t *src, const float *vm, int w, int h, const float *P, const K_<float> &K,
                       uint8_t *dst) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            uint8_t o[3] = {0, 0, 0};
            const float *v = vm + 4 * i;
            const float qx = P[0] * v[0] + P[1] * v[1] + P[2] * v[2] + P[3];
            const float qy = P[4] * v[0] + P[5] * v[1] + P[6] * v[2] + P[7];
            const float qz = P[8] * v[0] + P[9] * v[1] + P[10] * v[2] + P[11];
            const float u = K.fx * qx / qz + K.cx, vv = K.fy * qy / qz + K.cy;
            if (u >= 1 && vv >= 1 && u < (float)(w - 1) && vv < (float)(h - 1)) {
                const int u0 = (int)std::floor(u), v0 = (int)std::floor(vv);
                const float cx = u - (float)u0, cy = vv - (float)v0;
                const uint8_t *a = src + 3 * ((size_t)v0 * w + u0), *b = a + 3 * (size_t)w;
                for (int k = 0; k < 3; ++k) {
                    const float r = ((float)a[k] * (1 - cx) + (float)a[3 + k] * cx) * (1 - cy) +
                                    ((float)b[k] * (1 - cx) + (float)b[3 + k] * cx) * cy;
                    o[k] = (uint8_t)(int)r;
                }
            }
            dst[3 * i] = o[0]; dst[3 * i + 1] = o[1]; dst[3 * i + 2] = o[2];
        }
}

// ------------------------------------------------------------ reductions
// Rows are accumulated in 64 fixed chunks (each summed in pixel order), and the
// chunks in order, so the result does not depend on the thread count.
template <class R, class RowFn> void reduce29(int n, RowFn rowfn, R *out) {
    const int chunks = 64;
    std::vector<R> part((size_t)chunks * 29, (R)0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < chunks; ++c) {
        int lo = (int)((long)n * c / chunks), hi = (int)((long)n * (c + 1) / chunks);
        R *s = &part[(size_t)c * 29];
        R row[7];
        for (int i = lo; i < hi; ++i) {
            bool found = rowfn(i, row);
            if (!found) continue;  // rows of rejected pixels are all zero
            int k = 0;
            for (int a = 0; a < 7; ++a)
                for (int b = a; b < 7; ++b) s[k++] += row[a] * row[b];
            s[28] += (R)1;
        }
    }
    for (int k = 0; k < 29; ++k) out[k] = 0;
    for (int c = 0; c < chunks; ++c)
        for (int k = 0; k < 29; ++k) out[k] += part[(size_t)c * 29 + k];
}

// Every pixel's row instead of their sums (test infrastructure: the exact sum
// of the method's own rows measures a summation's error): 8 values per pixel,
// the 7 row values (0 when rejected) and 1 if the pixel was accepted.
template <class R, class RowFn> void rows8(int n, RowFn rowfn, R *out) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        R row[7] = {0, 0, 0, 0, 0, 0, 0};
        const bool found = rowfn(i, row);
        for (int k = 0; k < 7; ++k) out[8 * (size_t)i + k] = found ? row[k] : (R)0;
        out[8 * (size_t)i + 7] = found ? (R)1 : (R)0;
    }
}

// The 29 sums of rowfn's rows (out) or every pixel's row (rows8, when rows8_out
// is given).
template <class R, class RowFn> void sums_or_rows(int n, RowFn rowfn, R *out, R *rows8_out) {
    if (rows8_out) rows8<R>(n, rowfn, rows8_out);
    else reduce29<R>(n, rowfn, out);
}

// ICPReduction (pose_estimator.cu:166-222): point-to-plane rows.
template <class R> void icp_reduce(const R *curr_v, const R *curr_n, const R *last_v, const R *last_n, int cols,
                                   int rows, const M34<R> &pose, const K_<R> &K, R *out, R *rows8_out = nullptr) {
    const R angle = sizeof(R) == 4 ? (R)(float)std::cos(30 * 3.14 / 180) : (R)std::cos(30 * 3.14 / 180);
    const R dthr = sizeof(R) == 4 ? (R)0.01f : (R)0.01;
    sums_or_rows<R>(cols * rows, [&](int i, R *row) -> bool {
        const R *lv = last_v + 4 * (size_t)i;
        V3<R> vl = {lv[0], lv[1], lv[2]};
        if (isnan_(vl.x)) return false;
        V3<R> vg = pose(vl);
        R invz = (R)1 / vg.z;
        int u = proj_floor(vg.x * invz * K.fx + K.cx + (R)0.5);
        int v = proj_floor(vg.y * invz * K.fy + K.cy + (R)0.5);
        if (u < 0 || v < 0 || u >= cols || v >= rows) return false;
        const size_t j = (size_t)v * cols + u;
        V3<R> vc = {curr_v[4 * j], curr_v[4 * j + 1], curr_v[4 * j + 2]};
        const R *ln = last_n + 4 * (size_t)i;
        V3<R> ng = pose.rotate(V3<R>{ln[0], ln[1], ln[2]});
        V3<R> nc = {curr_n[4 * j], curr_n[4 * j + 1], curr_n[4 * j + 2]};
        R dist = norm(sub(vg, vc));
        R sine = norm(cross(nc, ng));
        if (!(sine < angle && dist <= dthr && !isnan_(nc.x) && !isnan_(ng.x))) return false;
        V3<R> c = cross(vg, ng);
        row[0] = ng.x; row[1] = ng.y; row[2] = ng.z;
        row[3] = c.x; row[4] = c.y; row[5] = c.z;
        row[6] = dot(ng, sub(vc, vg));
        return true;
    }, out, rows8_out);
}

template <class R> inline R interp2(const R *img, int w, R x, R y) {  // m_estimator.cu:340-346
    int u = (int)std::floor(x), v = (int)std::floor(y);
    R cx = x - (R)u, cy = y - (R)v;
    return (img[(size_t)v * w + u] * ((R)1 - cx) + img[(size_t)v * w + u + 1] * cx) * ((R)1 - cy) +
           (img[(size_t)(v + 1) * w + u] * ((R)1 - cx) + img[(size_t)(v + 1) * w + u + 1] * cx) * cy;
}

// RgbReduction2 (m_estimator.cu:313-391): Huber-weighted photometric rows.
template <class R> void rgb_step(const R *curr_I, const R *last_I, const R *last_v, const R *dIdx, const R *dIdy,
                                 int cols, int rows, R stddev, const M34<R> &pose, const K_<R> &K, R *out,
                                 R *rows8_out = nullptr) {
    const double sd = (double)stddev;
    const R th = (R)(1.345 * sd);
    sums_or_rows<R>(cols * rows, [&](int i, R *row) -> bool {
        const R *pt = last_v + 4 * (size_t)i;
        if (pt[3] < 0 || isnan_(pt[0])) return false;
        R il = last_I[i];
        if (!std::isfinite(il)) return false;
        V3<R> p = pose(V3<R>{pt[0], pt[1], pt[2]});
        R u0 = p.x / p.z * K.fx + K.cx, v0 = p.y / p.z * K.fy + K.cy;
        if (!(u0 >= 2 && u0 < (R)(cols - 2) && v0 >= 2 && v0 < (R)(rows - 2))) return false;
        R ic = interp2(curr_I, cols, u0, v0), dx = interp2(dIdx, cols, u0, v0), dy = interp2(dIdy, cols, u0, v0);
        if (!((dx > 2 || dy > 2) && std::isfinite(ic) && std::isfinite(dx) && std::isfinite(dy))) return false;
        R zi = (R)1 / p.z;
        V3<R> l;
        l.x = dx * K.fx * zi;
        l.y = dy * K.fy * zi;
        l.z = -(l.x * p.x + l.y * p.y) * zi;
        R r = ic - il;
        if (sd > 10e-5) r /= stddev;
        R w = 1;
        if (std::fabs(r) > th && sd > 10e-6) w = std::sqrt(th / std::fabs(r));
        V3<R> wl = {l.x * w, l.y * w, l.z * w};
        V3<R> c = cross(p, l);
        row[6] = w * (-r);
        row[0] = wl.x; row[1] = wl.y; row[2] = wl.z;
        row[3] = c.x * w; row[4] = c.y * w; row[5] = c.z * w;
        return true;
    }, out, rows8_out);
}

// create_jtjjtr<6,7> (reduce_sum.h:48-61).
template <class R> void unpack29(const R *s, double *A, double *b) {
    int k = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j) {
            double v = (double)s[k++];
            if (j == 6) b[i] = v;
            else A[j * 6 + i] = A[i * 6 + j] = v;
        }
}

// --------------------------------------------------------------- SE(3)
// Closed-form SE(3) exponential with tangent [upsilon; omega] (Sophus
// ordering), Barfoot, State Estimation for Robotics, sec. 7.1.
struct SE3 {
    double R[9], t[3];
    SE3() { for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0) ? 1 : 0; t[0] = t[1] = t[2] = 0; }
};
inline SE3 se3_mul(const SE3 &a, const SE3 &b) {
    SE3 c;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) c.R[i * 3 + j] = a.R[i * 3] * b.R[j] + a.R[i * 3 + 1] * b.R[3 + j] + a.R[i * 3 + 2] * b.R[6 + j];
        c.t[i] = a.R[i * 3] * b.t[0] + a.R[i * 3 + 1] * b.t[1] + a.R[i * 3 + 2] * b.t[2] + a.t[i];
    }
    return c;
}
inline SE3 se3_inv(const SE3 &a) {
    SE3 c;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) c.R[i * 3 + j] = a.R[j * 3 + i];
    for (int i = 0; i < 3; ++i) c.t[i] = -(c.R[i * 3] * a.t[0] + c.R[i * 3 + 1] * a.t[1] + c.R[i * 3 + 2] * a.t[2]);
    return c;
}
inline SE3 se3_exp(const double *xi) {
    const double *u = xi, *w = xi + 3;
    double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2], th = std::sqrt(th2);
    double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0}, W2[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) W2[i * 3 + j] = W[i * 3] * W[j] + W[i * 3 + 1] * W[3 + j] + W[i * 3 + 2] * W[6 + j];
    double A, B, C;
    if (th < 1e-8) { A = 1 - th2 / 6; B = 0.5 - th2 / 24; C = 1.0 / 6 - th2 / 120; }
    else { A = std::sin(th) / th; B = (1 - std::cos(th)) / th2; C = (th - std::sin(th)) / (th2 * th); }
    SE3 T;
    double V[9];
    for (int i = 0; i < 9; ++i) {
        double I = (i % 4 == 0) ? 1 : 0;
        T.R[i] = I + A * W[i] + B * W2[i];
        V[i] = I + B * W[i] + C * W2[i];
    }
    for (int i = 0; i < 3; ++i) T.t[i] = V[i * 3] * u[0] + V[i * 3 + 1] * u[1] + V[i * 3 + 2] * u[2];
    return T;
}
// SE(3) logarithm (Barfoot sec. 7.1.3); returns [upsilon; omega].
inline void se3_log(const SE3 &T, double *xi) {
    double c = (T.R[0] + T.R[4] + T.R[8] - 1) * 0.5;
    c = std::max(-1.0, std::min(1.0, c));
    double th = std::acos(c);
    double w[3];
    double v[3] = {T.R[7] - T.R[5], T.R[2] - T.R[6], T.R[3] - T.R[1]};
    if (th < 1e-8) { for (int i = 0; i < 3; ++i) w[i] = 0.5 * v[i]; }
    else if (M_PI - th < 1e-6) {
        // near pi: axis from the diagonal
        int k = 0;
        for (int i = 1; i < 3; ++i) if (T.R[i * 4] > T.R[k * 4]) k = i;
        double ax[3];
        ax[k] = std::sqrt(std::max(0.0, (T.R[k * 4] + 1) * 0.5));
        for (int i = 0; i < 3; ++i) if (i != k) ax[i] = T.R[k * 3 + i] / (2 * ax[k]);
        for (int i = 0; i < 3; ++i) w[i] = ax[i] * th;
    } else { double s = th / (2 * std::sin(th)); for (int i = 0; i < 3; ++i) w[i] = s * v[i]; }
    double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    th = std::sqrt(th2);
    double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0}, W2[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) W2[i * 3 + j] = W[i * 3] * W[j] + W[i * 3 + 1] * W[3 + j] + W[i * 3 + 2] * W[6 + j];
    double D = th < 1e-8 ? 1.0 / 12 : (1 - th * std::sin(th) / (2 * (1 - std::cos(th)))) / th2;
    double Vi[9];
    for (int i = 0; i < 9; ++i) Vi[i] = ((i % 4 == 0) ? 1 : 0) - 0.5 * W[i] + D * W2[i];
    for (int i = 0; i < 3; ++i) xi[i] = Vi[i * 3] * T.t[0] + Vi[i * 3 + 1] * T.t[1] + Vi[i * 3 + 2] * T.t[2];
    for (int i = 0; i < 3; ++i) xi[3 + i] = w[i];
}

// Solve A x = b for a symmetric positive definite 6x6 by LDL^T without pivoting
// (the reference uses Eigen's pivoted LDLT, icp_tracker.cpp:92 — same solution
// for SPD A up to rounding; parity unpinned at the last bits).
inline bool ldlt6(const double *A, const double *b, double *x) {
    double L[36] = {0}, D[6];
    for (int j = 0; j < 6; ++j) {
        double d = A[j * 6 + j];
        for (int k = 0; k < j; ++k) d -= L[j * 6 + k] * L[j * 6 + k] * D[k];
        D[j] = d;
        L[j * 6 + j] = 1;
        for (int i = j + 1; i < 6; ++i) {
            double s = A[i * 6 + j];
            for (int k = 0; k < j; ++k) s -= L[i * 6 + k] * L[j * 6 + k] * D[k];
            L[i * 6 + j] = d != 0 ? s / d : 0;
        }
    }
    double y[6];
    for (int i = 0; i < 6; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= L[i * 6 + k] * y[k];
        y[i] = s;
    }
    for (int i = 0; i < 6; ++i) y[i] = D[i] != 0 ? y[i] / D[i] : 0;
    for (int i = 5; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < 6; ++k) s -= L[k * 6 + i] * x[k];
        x[i] = s;
    }
    return true;
}

// Float 3x4 of an SE3 (the reference's pose.cast<float>().matrix3x4()).
template <class R> M34<R> to_m34(const SE3 &T) {
    M34<R> m;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) m.m[i * 4 + j] = (R)T.R[i * 3 + j];
        m.m[i * 4 + 3] = (R)T.t[i];
    }
    return m;
}

// --------------------------------------------------------------- frames
// DeviceImage (device_image.cpp:65-199) on the CPU.
template <class R> struct Frame {
    int levels;
    std::vector<Intr> Ks;
    std::vector<std::vector<R>> depth, inten, dx, dy, vmap, nmap;
    std::vector<R> depth_raw;
    std::vector<uint8_t> rgb;
    Frame(const Intr &K0, int L) : levels(L) {
        Ks.push_back(K0);
        for (int l = 1; l < L; ++l) {  // IntrinsicMatrix::pyr_down (intrinsic_matrix.h:33-37)
            const Intr &k = Ks.back();
            Intr n;
            float s = 0.5f;
            n.width = (int)(s * (float)k.width); n.height = (int)(s * (float)k.height);
            n.fx = s * k.fx; n.fy = s * k.fy; n.cx = s * k.cx; n.cy = s * k.cy;
            n.invfx = 1.0f / n.fx; n.invfy = 1.0f / n.fy;
            Ks.push_back(n);
        }
        depth.resize(L); inten.resize(L); dx.resize(L); dy.resize(L); vmap.resize(L); nmap.resize(L);
        for (int l = 0; l < L; ++l) {
            size_t n = (size_t)Ks[l].width * Ks[l].height;
            depth[l].resize(n); inten[l].resize(n); dx[l].resize(n); dy[l].resize(n);
            vmap[l].resize(4 * n); nmap[l].resize(4 * n);
        }
    }
    // DeviceImage::upload (device_image.cpp:146-181)
    void upload(const uint16_t *d16, const uint8_t *rgb8, double scale) {
        const int W = Ks[0].width, H = Ks[0].height;
        const size_t n = (size_t)W * H;
        depth_raw.resize(n);
        const R sc = (R)scale;
        for (size_t i = 0; i < n; ++i) depth_raw[i] = (R)d16[i] * sc;
        rgb.assign(rgb8, rgb8 + 3 * n);
        grey(rgb8, inten[0].data(), W, H);
        bilateral(depth_raw.data(), depth[0].data(), W, H);
        for (int l = 1; l < levels; ++l) {
            resize_half(depth[l - 1].data(), depth[l].data(), Ks[l - 1].width, Ks[l - 1].height, 1);
            pyr_down(inten[l - 1].data(), inten[l].data(), Ks[l - 1].width, Ks[l - 1].height);
        }
        for (int l = 0; l < levels; ++l) {
            K_<R> K(Ks[l]);
            derivative(inten[l].data(), dx[l].data(), dy[l].data(), Ks[l].width, Ks[l].height);
            back_project(depth[l].data(), vmap[l].data(), Ks[l].width, Ks[l].height, K);
            compute_nmap(vmap[l].data(), nmap[l].data(), Ks[l].width, Ks[l].height);
        }
    }
    // DeviceImage::resize_device_map (device_image.cpp:188-199)
    void resize_device_map() {
        for (int l = 1; l < levels; ++l) resize_half(vmap[l - 1].data(), vmap[l].data(), Ks[l - 1].width, Ks[l - 1].height, 4);
        for (int l = 0; l < levels; ++l) compute_nmap(vmap[l].data(), nmap[l].data(), Ks[l].width, Ks[l].height);
    }
};

inline SE3 se3_from12(const double *m) {
    SE3 T;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) T.R[i * 3 + j] = m[i * 4 + j];
        T.t[i] = m[i * 4 + 3];
    }
    return T;
}
inline void se3_to12(const SE3 &T, double *m) {
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) m[i * 4 + j] = T.R[i * 3 + j];
        m[i * 4 + 3] = T.t[i];
    }
}

// set_source_vmap / set_source_image (icp_tracker.cpp:164-176; build_pyramid.cpp:
// 17-76): a frame's pyramids from a level-0 vertex map and an RGB image.
template <class R> void frame_from_maps(Frame<R> &f, const R *vmap0, const uint8_t *rgb) {
    const int W = f.Ks[0].width, H = f.Ks[0].height;
    std::copy(vmap0, vmap0 + 4 * (size_t)W * H, f.vmap[0].begin());
    f.resize_device_map();
    grey(rgb, f.inten[0].data(), W, H);
    for (int l = 1; l < f.levels; ++l) pyr_down(f.inten[l - 1].data(), f.inten[l].data(), f.Ks[l - 1].width, f.Ks[l - 1].height);
    for (int l = 0; l < f.levels; ++l) derivative(f.inten[l].data(), f.dx[l].data(), f.dy[l].data(), f.Ks[l].width, f.Ks[l].height);
}

// DenseTracking::compute_transform(ref, cur, ctx) (icp_tracker.cpp:23-156);
// with context_mode, compute_transform(ctx) (icp_tracker.cpp:210-314): NaN
// errors are not flagged and success is unconditional.  Starts at init.
template <class R> struct Tracker {
    int levels;
    std::vector<int> iters;
    bool success = true;
    float last_icp_error = 0;
    SE3 compute(const Frame<R> &ref, const Frame<R> &cur, const SE3 &init = SE3(), bool context_mode = false) {
        SE3 est = init, prev = init;  // Revertable<SE3d> (revertable.h:7-45)
        bool invalid = false;
        for (int level = levels - 1; level >= 0; --level) {
            const Intr &Ki = ref.Ks[level];
            K_<R> K(Ki);
            const int W = Ki.width, H = Ki.height;
            R icp_error = std::numeric_limits<R>::max(), rgb_error = std::numeric_limits<R>::max();
            int icp_count = 0, rgb_count = 0;
            R stddev = 0;
            for (int it = 0; it < iters[level]; ++it) {
                SE3 last = est;
                R last_icp = icp_error, last_rgb = rgb_error;
                M34<R> P = to_m34<R>(last);
                R si[29], sr[29];
                icp_reduce(cur.vmap[level].data(), cur.nmap[level].data(), ref.vmap[level].data(),
                           ref.nmap[level].data(), W, H, P, K, si);
                rgb_step(cur.inten[level].data(), ref.inten[level].data(), ref.vmap[level].data(),
                         cur.dx[level].data(), cur.dy[level].data(), W, H, stddev, P, K, sr);
                stddev = std::sqrt(sr[27] / (sr[28] - (R)6));
                double Ai[36], bi[6], Ar[36], br[6], A[36], b[6], x[6];
                unpack29(si, Ai, bi);
                unpack29(sr, Ar, br);
                // A = 1e6 * H_icp + H_rgb in the float type (icp_tracker.cpp:89-90)
                for (int k = 0; k < 36; ++k) A[k] = (double)((R)1e6 * (R)Ai[k] + (R)Ar[k]);
                for (int k = 0; k < 6; ++k) b[k] = (double)((R)1e6 * (R)bi[k] + (R)br[k]);
                ldlt6(A, b, x);
                prev = est;
                est = se3_mul(se3_exp(x), last);
                icp_error = std::sqrt(si[27]) / si[28];
                last_icp_error = (float)last_icp;
                if (isnan_(icp_error) && !context_mode) { invalid = true; break; }
                if (icp_error > last_icp) {
                    if (icp_count >= 2) { est = prev; break; }
                    icp_count++;
                    icp_error = last_icp;
                } else icp_count = 0;
                rgb_error = std::sqrt(sr[27]) / sr[28];
                if (isnan_(rgb_error) && !context_mode) { invalid = true; break; }
                if (rgb_error > last_rgb) {
                    if (rgb_count >= 2) { est = prev; break; }
                    rgb_count++;
                    rgb_error = last_rgb;
                } else rgb_count = 0;
            }
        }
        if (context_mode) { success = true; return se3_inv(est); }
        double xi[6];
        se3_log(se3_mul(se3_inv(est), init), xi);
        double nrm = 0;
        for (int k = 0; k < 6; ++k) nrm += xi[k] * xi[k];
        nrm = std::sqrt(nrm);
        if (invalid || nrm > 0.1) { success = false; return SE3(); }
        success = true;
        return se3_inv(est);
    }
};

// System::process_images + DenseOdometry::trackFrame + DenseMapping (system.cpp:40-114,
// rgbd_odometry.cpp:18-48, voxel_hashing.cpp:44-119).
template <class R> struct System {
    Map<R> map;
    Frame<R> a, b;
    Frame<R> *cur, *ref;
    Tracker<R> tracker;
    SE3 pose, last_pose;
    bool initialized = false;
    double depth_scale;
    int frame_id = 0;
    System(const Params &p, const Intr &K, int L, const int *iters, double scale)
        : map(p), a(K, L), b(K, L), cur(&a), ref(&b), depth_scale(scale) {
        tracker.levels = L;
        tracker.iters.assign(iters, iters + L);
    }
    void process(const uint16_t *d16, const uint8_t *rgb8) {
        cur->upload(d16, rgb8, depth_scale);
        if (!initialized) {
            pose = SE3();
            initialized = true;
        } else {
            SE3 upd = tracker.compute(*ref, *cur);
            pose = se3_mul(last_pose, upd);
        }
        last_pose = pose;
        std::swap(cur, ref);
        // DenseMapping::update(frame): raw depth, level-0 nmap, rgb, pose
        const Intr &K0 = ref->Ks[0];
        K_<R> K(K0);
        M34<R> Tw = to_m34<R>(pose), Tc = to_m34<R>(se3_inv(pose));
        allocate(map, ref->depth_raw.data(), K0.width, K0.height, Tw, K);
        select_visible(map, Tc, K, K0.width, K0.height);
        integrate(map, ref->depth_raw.data(), ref->nmap[0].data(), ref->rgb.data(), K0.width, K0.height, Tc, K);
        // DenseMapping::raycast returns early when nothing is visible (voxel_hashing.cpp:92-93)
        if (!map.visible.empty())
            raycast(map, ref->vmap[0].data(), (uint8_t *)nullptr, K0.width, K0.height, Tw, Tc, K, 1, false);
        ref->resize_device_map();
        ++frame_id;
    }
};

inline void se3_to16(const SE3 &P, double *T16) {
    for (int i = 0; i < 3; ++i) { for (int j = 0; j < 3; ++j) T16[i * 4 + j] = P.R[i * 3 + j]; T16[i * 4 + 3] = P.t[i]; }
    T16[12] = T16[13] = T16[14] = 0; T16[15] = 1;
}

template <class R> M34<R> m34_from(const R *p) { M34<R> m; for (int i = 0; i < 12; ++i) m.m[i] = p[i]; return m; }

}  // namespace orc

using namespace orc;

extern "C" void orc_render_scene(const float *vm, const float *nm, const uint8_t *rgb, int w, int h, uint8_t *out) {
    render_scene(vm, nm, rgb, w, h, out);
}
extern "C" void orc_warp_image(const uint8_t *src, const float *vm, int w, int h, const float *pose12, const Intr *K,
                               uint8_t *dst) {
    warp_image(src, vm, w, h, pose12, K_<float>(*K), dst);
}

#define ORC_INSTANTIATE(R, S)                                                                                     \
    extern "C" void *orc_map_create_##S(const Params *p) { return new Map<R>(*p); }                            \
    extern "C" void orc_map_free_##S(void *m) { delete static_cast<Map<R> *>(m); }                             \
    extern "C" void orc_map_reset_##S(void *m) { static_cast<Map<R> *>(m)->reset(); }                          \
    extern "C" int orc_map_num_blocks_##S(void *m) { return (int)static_cast<Map<R> *>(m)->pos.size(); }       \
    extern "C" int orc_map_num_visible_##S(void *m) { return (int)static_cast<Map<R> *>(m)->visible.size(); }  \
    extern "C" unsigned orc_map_overflow_##S(void *m) { return static_cast<Map<R> *>(m)->overflow; }           \
    extern "C" void orc_map_blocks_##S(void *m, int *xyz) {                                                    \
        auto *M = static_cast<Map<R> *>(m);                                                                    \
        for (size_t i = 0; i < M->pos.size(); ++i) { xyz[3 * i] = M->pos[i].x; xyz[3 * i + 1] = M->pos[i].y; xyz[3 * i + 2] = M->pos[i].z; } \
    }                                                                                                          \
    extern "C" void orc_map_visible_##S(void *m, int *xyz) {                                                   \
        auto *M = static_cast<Map<R> *>(m);                                                                    \
        for (size_t i = 0; i < M->visible.size(); ++i) { I3 p = M->pos[M->visible[i]]; xyz[3 * i] = p.x; xyz[3 * i + 1] = p.y; xyz[3 * i + 2] = p.z; } \
    }                                                                                                          \
    extern "C" void orc_map_voxels_##S(void *m, int16_t *sdf, R *w, uint8_t *rgb) {                            \
        auto *M = static_cast<Map<R> *>(m);                                                                    \
        std::memcpy(sdf, M->sdf.data(), M->sdf.size() * sizeof(int16_t));                                     \
        std::memcpy(w, M->weight.data(), M->weight.size() * sizeof(R));                                        \
        if (rgb) std::memcpy(rgb, M->rgb.data(), M->rgb.size());                                              \
    }                                                                                                          \
    extern "C" int orc_map_set_blocks_##S(void *m, int n, const int *xyz, const int16_t *sdf, const R *w,      \
                                          const uint8_t *rgb) {                                                \
        auto *M = static_cast<Map<R> *>(m);                                                                    \
        for (int i = 0; i < n; ++i) {                                                                          \
            int id = M->insert(I3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]});                                \
            if (id < 0) return -1;                                                                             \
            std::memcpy(&M->sdf[(size_t)id * 512], sdf + (size_t)i * 512, 512 * sizeof(int16_t));             \
            std::memcpy(&M->weight[(size_t)id * 512], w + (size_t)i * 512, 512 * sizeof(R));                   \
            if (rgb) std::memcpy(&M->rgb[(size_t)id * 1536], rgb + (size_t)i * 1536, 1536);                   \
        }                                                                                                      \
        return 0;                                                                                              \
    }                                                                                                          \
    extern "C" void orc_allocate_##S(void *m, const R *depth, int cols, int rows, const R *pose24,             \
                                     const Intr *K) {                                                          \
        allocate(*static_cast<Map<R> *>(m), depth, cols, rows, m34_from(pose24), K_<R>(*K));                   \
    }                                                                                                          \
    extern "C" void orc_update_weighted_##S(void *m, const R *depth, const R *nmap, const uint8_t *rgb,        \
                                            int cols, int rows, const R *pose24, const Intr *K) {              \
        auto &M = *static_cast<Map<R> *>(m);                                                                   \
        K_<R> k(*K);                                                                                           \
        allocate(M, depth, cols, rows, m34_from(pose24), k);                                                   \
        select_visible(M, m34_from(pose24 + 12), k, cols, rows);                                                           \
        integrate(M, depth, nmap, rgb, cols, rows, m34_from(pose24 + 12), k);                                  \
    }                                                                                                          \
    extern "C" int orc_select_visible_##S(void *m, const R *pose24, const Intr *K, int cols, int rows) {         \
        auto &M = *static_cast<Map<R> *>(m);                                                                   \
        select_visible(M, m34_from(pose24 + 12), K_<R>(*K), cols, rows, true);                                 \
        return (int)M.visible.size();                                                                          \
    }                                                                                                          \
    extern "C" void orc_raycast_##S(void *m, R *vmap, uint8_t *rgb, int cols, int rows, const R *pose24,       \
                                    const Intr *K, int check_colour, int instrument) {                        \
        raycast(*static_cast<Map<R> *>(m), vmap, rgb, cols, rows, m34_from(pose24), m34_from(pose24 + 12),     \
                K_<R>(*K), check_colour, instrument != 0);                                                     \
    }                                                                                                          \
    extern "C" void orc_ray_stats_##S(void *m, long long *out6) {                                              \
        for (int k = 0; k < 6; ++k) out6[k] = static_cast<Map<R> *>(m)->ray_stats[k];                          \
    }                                                                                                          \
    extern "C" void orc_ray_px_##S(void *m, int *out) {                                                        \
        auto &M = *static_cast<Map<R> *>(m);                                                                   \
        std::memcpy(out, M.ray_px.data(), M.ray_px.size() * sizeof(int));                                      \
    }                                                                                                          \
    extern "C" int orc_ray_blocks_touched_##S(void *m) {                                                       \
        auto &M = *static_cast<Map<R> *>(m);                                                                   \
        int n = 0;                                                                                             \
        for (uint8_t t : M.ray_touched) n += t;                                                                \
        return n;                                                                                              \
    }                                                                                                          \
    extern "C" void orc_zrange_##S(void *m, R *zx, R *zy, int cols, int rows, const R *pose24, const Intr *K) { \
        std::vector<R> a, b;                                                                                   \
        zrange_tiles(*static_cast<Map<R> *>(m), m34_from(pose24 + 12), K_<R>(*K), cols / 8, rows / 8, a, b);   \
        std::memcpy(zx, a.data(), a.size() * sizeof(R));                                                      \
        std::memcpy(zy, b.data(), b.size() * sizeof(R));                                                      \
    }                                                                                                          \
    extern "C" long orc_mesh_##S(void *m, int mode, float *tri, float *nrm, uint8_t *col, long max_tri) {      \
        return mesh(*static_cast<Map<R> *>(m), mode, tri, nrm, col, max_tri);                                  \
    }                                                                                                          \
    extern "C" void orc_bilateral_##S(const R *in, R *out, int w, int h) { bilateral(in, out, w, h); }         \
    extern "C" void orc_resize_half_##S(const R *in, R *out, int w, int h, int ch) { resize_half(in, out, w, h, ch); } \
    extern "C" void orc_pyr_down_##S(const R *in, R *out, int w, int h) { pyr_down(in, out, w, h); }           \
    extern "C" void orc_grey_##S(const uint8_t *rgb, R *out, int w, int h) { grey(rgb, out, w, h); }           \
    extern "C" void orc_derivative_##S(const R *I, R *dx, R *dy, int w, int h) { derivative(I, dx, dy, w, h); } \
    extern "C" void orc_back_project_##S(const R *d, R *vm, int w, int h, const Intr *K) {                     \
        back_project(d, vm, w, h, K_<R>(*K));                                                                  \
    }                                                                                                          \
    extern "C" void orc_nmap_##S(const R *vm, R *nm, int w, int h) { compute_nmap(vm, nm, w, h); }             \
    extern "C" void orc_icp_reduce_##S(const R *cv, const R *cn, const R *lv, const R *ln, int w, int h,       \
                                       const R *pose12, const Intr *K, R *out) {                               \
        icp_reduce(cv, cn, lv, ln, w, h, m34_from(pose12), K_<R>(*K), out);                                    \
    }                                                                                                          \
    extern "C" void orc_rgb_step_##S(const R *ci, const R *li, const R *lv, const R *dx, const R *dy, int w,   \
                                     int h, R stddev, const R *pose12, const Intr *K, R *out) {                \
        rgb_step(ci, li, lv, dx, dy, w, h, stddev, m34_from(pose12), K_<R>(*K), out);                          \
    }                                                                                                          \
    extern "C" void orc_icp_rows_##S(const R *cv, const R *cn, const R *lv, const R *ln, int w, int h,         \
                                     const R *pose12, const Intr *K, R *rows8) {                               \
        icp_reduce(cv, cn, lv, ln, w, h, m34_from(pose12), K_<R>(*K), (R *)nullptr, rows8);                    \
    }                                                                                                          \
    extern "C" void orc_rgb_rows_##S(const R *ci, const R *li, const R *lv, const R *dx, const R *dy, int w,   \
                                     int h, R stddev, const R *pose12, const Intr *K, R *rows8) {              \
        rgb_step(ci, li, lv, dx, dy, w, h, stddev, m34_from(pose12), K_<R>(*K), (R *)nullptr, rows8);          \
    }                                                                                                          \
    extern "C" int orc_track_maps_##S(const R *ref_vmap, const uint8_t *ref_rgb, const R *cur_vmap,            \
                                      const uint8_t *cur_rgb, const Intr *K, int L, const int *iters,          \
                                      const double *init12, int context_mode, double *update12) {              \
        Frame<R> ref(*K, L), cur(*K, L);                                                                       \
        frame_from_maps(ref, ref_vmap, ref_rgb);                                                               \
        frame_from_maps(cur, cur_vmap, cur_rgb);                                                               \
        Tracker<R> t;                                                                                          \
        t.levels = L;                                                                                          \
        t.iters.assign(iters, iters + L);                                                                      \
        se3_to12(t.compute(ref, cur, se3_from12(init12), context_mode != 0), update12);                        \
        return t.success ? 1 : 0;                                                                              \
    }                                                                                                          \
    extern "C" void *orc_system_create_##S(const Params *p, const Intr *K, int L, const int *iters,            \
                                           double scale) {                                                     \
        return new System<R>(*p, *K, L, iters, scale);                                                         \
    }                                                                                                          \
    extern "C" void orc_system_free_##S(void *s) { delete static_cast<System<R> *>(s); }                       \
    extern "C" void orc_system_process_##S(void *s, const uint16_t *d, const uint8_t *rgb) {                   \
        static_cast<System<R> *>(s)->process(d, rgb);                                                          \
    }                                                                                                          \
    extern "C" void orc_system_pose_##S(void *s, double *T16) {                                                \
        se3_to16(static_cast<System<R> *>(s)->pose, T16);                                                      \
    }                                                                                                          \
    extern "C" int orc_system_success_##S(void *s) { return static_cast<System<R> *>(s)->tracker.success ? 1 : 0; } \
    extern "C" void *orc_system_map_##S(void *s) { return &static_cast<System<R> *>(s)->map; }                 \
    extern "C" void orc_system_ref_maps_##S(void *s, int level, R *vmap, R *nmap) {                            \
        auto *S_ = static_cast<System<R> *>(s);                                                                \
        std::memcpy(vmap, S_->ref->vmap[level].data(), S_->ref->vmap[level].size() * sizeof(R));               \
        std::memcpy(nmap, S_->ref->nmap[level].data(), S_->ref->nmap[level].size() * sizeof(R));               \
    }

ORC_INSTANTIATE(float, f)
ORC_INSTANTIATE(double, d)

extern "C" int orc_compute_hash(int x, int y, int z, int n) { return compute_hash(x, y, z, n); }
extern "C" void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
extern "C" int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
extern "C" void orc_se3_exp(const double *xi, double *T16) {
    SE3 T = se3_exp(xi);
    for (int i = 0; i < 3; ++i) { for (int j = 0; j < 3; ++j) T16[i * 4 + j] = T.R[i * 3 + j]; T16[i * 4 + 3] = T.t[i]; }
    T16[12] = T16[13] = T16[14] = 0; T16[15] = 1;
}
extern "C" void orc_se3_log(const double *T16, double *xi) {
    SE3 T;
    for (int i = 0; i < 3; ++i) { for (int j = 0; j < 3; ++j) T.R[i * 3 + j] = T16[i * 4 + j]; T.t[i] = T16[i * 4 + 3]; }
    se3_log(T, xi);
}
extern "C" void orc_ldlt6(const double *A, const double *b, double *x) { ldlt6(A, b, x); }
