// Depth / intensity pre-processing and the vertex / normal pyramids.
//
// Reference: src/tracking/cuda_imgproc.cu (filterDepthBilateral :197-200,
// pyrDownDepth :202-205, pyrDownImage :207-210, pyrDownVMap :212-215,
// computeDerivativeK :217-247, backProjectDepthK :249-274, computeNMapK
// :276-305) and src/tracking/device_image.cpp (upload :146-181, grey
// conversion :159-162).  The OpenCV-CUDA calls of the reference become
// hand-written kernels with the documented OpenCV semantics noted per kernel.
//
// All kernels are one lane per output pixel on 16x16 workgroups (4 wavefronts of
// 16x4 pixels): every row segment a wavefront reads or writes is contiguous.
#include <vector>

#include "lf_common.h"
#include "lf_internal.h"

using namespace lf;

namespace {

LF_DEV int reflect101(int i, int n) {
    if (n == 1) return 0;
    while (i < 0 || i >= n) i = i < 0 ? -i : 2 * n - 2 - i;
    return i;
}

LF_DEV float depth_value(uint16_t d, float scale) { return (float)d * scale; }

// cv::cuda::bilateralFilter(src, dst, 5, 1, 1): 5x5 window, reflect-101 border,
// w = exp(-(dx^2+dy^2)/(2*1^2) - (v-c)^2/(2*1^2)); zero depths are not masked.
// at(x, y) returns the source value of an in-range pixel.
// kFast (mixed precision): the hardware exponential (v_exp_f32 on a base-2
// argument) instead of the correctly rounded expf.
template <class At, bool kFast = false>
LF_DEV float bilateral_at(const At &at, int x, int y, int w, int h) {
    const float c = at(x, y);
    float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
        const int yy = reflect101(y + dy, h);
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const float v = at(reflect101(x + dx, w), yy);
            const float d = v - c;
            const float arg = (float)(dx * dx + dy * dy) * -0.5f + d * d * -0.5f;
            const float wt = kFast ? __expf(arg) : expf(arg);
            s1 += v * wt;
            s2 += wt;
        }
    }
    return s1 / s2;
}

// cv::cuda::resize(x0.5, INTER_LINEAR) with the pixel-centre convention: the
// mean of the 2x2 source block, NaN propagating.
LF_DEV float mean4(float a, float b, float c, float d) { return ((a + b) + (c + d)) * 0.25f; }
LF_DEV float resize_at(const float *in, int w, int x, int y) {
    const float *r0 = in + (size_t)(2 * y) * w + 2 * x, *r1 = r0 + w;
    return mean4(r0[0], r0[1], r1[0], r1[1]);
}
// V(x, y) returns the vertex of an in-range pixel of the full-size map.
template <class V>
LF_DEV float4 resize4_at(const V &in, int x, int y) {
    const float4 a = in(2 * x, 2 * y), b = in(2 * x + 1, 2 * y), c = in(2 * x, 2 * y + 1), d = in(2 * x + 1, 2 * y + 1);
    return make_float4(mean4(a.x, b.x, c.x, d.x), mean4(a.y, b.y, c.y, d.y), mean4(a.z, b.z, c.z, d.z),
                       mean4(a.w, b.w, c.w, d.w));
}

// cv::cuda::pyrDown: [1 4 6 4 1]^2/256 Gaussian at even pixels, reflect-101.
// I(x, y) returns the source value of an in-range pixel.
template <class I>
LF_DEV float pyr_down_at(const I &in, int w, int h, int x, int y) {
    const int x0 = reflect101(2 * x - 2, w), x1 = reflect101(2 * x - 1, w), x2 = 2 * x, x3 = reflect101(2 * x + 1, w),
              x4 = reflect101(2 * x + 2, w);
    float rs[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        const int r = reflect101(2 * y + j - 2, h);
        rs[j] = (in(x0, r) + in(x4, r)) + 4.0f * (in(x1, r) + in(x3, r)) + 6.0f * in(x2, r);
    }
    return ((rs[0] + rs[4]) + 4.0f * (rs[1] + rs[3]) + 6.0f * rs[2]) * (1.0f / 256.0f);
}

// cvtColor(RGB2GRAY) on float: 0.299 R + 0.587 G + 0.114 B.
LF_DEV float grey_of(unsigned char r, unsigned char g, unsigned char b) {
    return (float)r * 0.299f + (float)g * 0.587f + (float)b * 0.114f;
}

// computeDerivativeK (cuda_imgproc.cu:217-234); last row / column 0.
template <class I>
LF_DEV void derivative_at(const I &in, int w, int h, int x, int y, float &dx, float &dy) {
    if (x >= w - 1 || y >= h - 1) {
        dx = 0.0f;
        dy = 0.0f;
        return;
    }
    const int x10 = max(x - 1, 0), x01 = min(x + 1, w), y10 = max(y - 1, 0), y01 = min(y + 1, h);
    dx = (in(x01, y) - in(x10, y)) * 0.5f;
    dy = (in(x, y01) - in(x, y10)) * 0.5f;
}

// backProjectDepthK (cuda_imgproc.cu:249-263).
LF_DEV float4 back_project_at(float z, int x, int y, const lf_intrinsics &K) {
    const float qnan = __int_as_float(0x7fc00000);
    if (z > 0.3f && z < 5.0f) return make_float4(z * ((float)x - K.cx) * K.invfx, z * ((float)y - K.cy) * K.invfy, z, 1.0f);
    return make_float4(qnan, qnan, qnan, -1.0f);
}

// computeNMapK (cuda_imgproc.cu:276-294); last row / column NaN.  V(x, y)
// returns the vertex of an in-range pixel.
// kFast (mixed precision): normalised with the hardware reciprocal square root.
template <class V, bool kFast = false>
LF_DEV float4 nmap_at(const V &vert, int w, int h, int x, int y) {
    const float qnan = __int_as_float(0x7fc00000);
    if (x >= w - 1 || y >= h - 1) return make_float4(qnan, qnan, qnan, qnan);
    const int x10 = max(x - 1, 0), x01 = min(x + 1, w), y10 = max(y - 1, 0), y01 = min(y + 1, h);
    const float4 a = vert(x01, y), b = vert(x10, y), c = vert(x, y01), e = vert(x, y10);
    const f3 cr = cross(f3{a.x - b.x, a.y - b.y, a.z - b.z}, f3{c.x - e.x, c.y - e.y, c.z - e.z});
    const f3 n = kFast ? mul(cr, __builtin_amdgcn_rsqf(dot(cr, cr))) : normalised(cr);
    return make_float4(n.x, n.y, n.z, 1.0f);
}

// The integration sample of a pixel, read by the weighted TSDF update as one
// 16-byte gather: x = raw depth, or NaN when the normal is invalid (both skip
// the pixel); y = the weight |sin(n_z)| / d^2 of update_map_weighted_kernel
// (depth_fusion.cu:356-428) in its operation order; z = colour r | g << 8 |
// b << 16 as bits.
template <bool kFast = false>
LF_DEV float4 integration_sample(float dist, float4 n, const unsigned char *c) {
    const float qnan = __int_as_float(0x7fc00000);
    // kFast (mixed precision): hardware sine and reciprocal
    const float wt = kFast ? fabsf(__sinf(n.z)) * __builtin_amdgcn_rcpf(dist * dist) : fabsf(sinf(n.z)) / (dist * dist);
    const unsigned int rgba = (unsigned int)c[0] | ((unsigned int)c[1] << 8) | ((unsigned int)c[2] << 16);
    return make_float4(isnan(n.x) ? qnan : dist, wt, __uint_as_float(rgba), 0.0f);
}

struct FloatAt {
    const float *p; int w;
    LF_DEV float operator()(int x, int y) const { return p[(size_t)y * w + x]; }
};
struct VertexAt {
    const float4 *p; int w;
    LF_DEV float4 operator()(int x, int y) const { return p[(size_t)y * w + x]; }
};
// a plane of FP32 values, or of FP16 values (kHalf)
template <bool kHalf>
struct PlaneAt {
    const void *p; int w;
    LF_DEV float operator()(int x, int y) const {
        const size_t i = (size_t)y * w + x;
        if constexpr (kHalf) return half_bits_to_float(static_cast<const unsigned short *>(p)[i]);
        else return static_cast<const float *>(p)[i];
    }
};
template <bool kHalf>
LF_DEV void store1(void *p, size_t i, float v) {
    if constexpr (kHalf) static_cast<unsigned short *>(p)[i] = float_to_half_bits(v);
    else static_cast<float *>(p)[i] = v;
}
template <bool kHalf>
LF_DEV void store4(void *p, size_t i, float4 v) {
    if constexpr (kHalf)
        static_cast<uint2 *>(p)[i] = make_uint2(float_to_half_bits(v.x) | ((unsigned int)float_to_half_bits(v.y) << 16),
                                                float_to_half_bits(v.z) | ((unsigned int)float_to_half_bits(v.w) << 16));
    else
        static_cast<float4 *>(p)[i] = v;
}
// vertex (x, y) of the quarter-size map, through the four half-size vertices
// resized on the fly from the full-size map (same arithmetic as two stored
// resizes)
template <class V>
LF_DEV float4 resize4_quarter_at(const V &full, int x, int y) {
    const float4 a = resize4_at(full, 2 * x, 2 * y), b = resize4_at(full, 2 * x + 1, 2 * y),
                 c = resize4_at(full, 2 * x, 2 * y + 1), d = resize4_at(full, 2 * x + 1, 2 * y + 1);
    return make_float4(mean4(a.x, b.x, c.x, d.x), mean4(a.y, b.y, c.y, d.y), mean4(a.z, b.z, c.z, d.z),
                       mean4(a.w, b.w, c.w, d.w));
}

// ------------------------------------------------------------ stand-alone kernels

__global__ void k_depth_to_float(const uint16_t *__restrict__ in, float *__restrict__ out, int n, float scale) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = depth_value(in[i], scale);
}

__global__ void k_bilateral(const float *__restrict__ in, float *__restrict__ out, int w, int h) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= w || y >= h) return;
    out[(size_t)y * w + x] = bilateral_at(FloatAt{in, w}, x, y, w, h);
}

// ch = 1 (depth) or 4 (vmap).
template <int kCh>
__global__ void k_resize_half(const float *__restrict__ in, float *__restrict__ out, int w, int h) {
    const int ow = w / 2, oh = h / 2;
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= ow || y >= oh) return;
    if (kCh == 1) out[(size_t)y * ow + x] = resize_at(in, w, x, y);
    else reinterpret_cast<float4 *>(out)[(size_t)y * ow + x] = resize4_at(VertexAt{reinterpret_cast<const float4 *>(in), w}, x, y);
}

__global__ void k_pyr_down(const float *__restrict__ in, float *__restrict__ out, int w, int h) {
    const int ow = (w + 1) / 2, oh = (h + 1) / 2;
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= ow || y >= oh) return;
    out[(size_t)y * ow + x] = pyr_down_at(FloatAt{in, w}, w, h, x, y);
}

__global__ void k_grey(const unsigned char *__restrict__ rgb, float *__restrict__ out, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = grey_of(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
}

__global__ void k_derivative(const float *__restrict__ I, float *__restrict__ dx, float *__restrict__ dy, int w, int h) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * w + x;
    derivative_at(FloatAt{I, w}, w, h, x, y, dx[i], dy[i]);
}

__global__ void k_back_project(const float *__restrict__ d, float *__restrict__ vm, int w, int h, lf_intrinsics K) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * w + x;
    reinterpret_cast<float4 *>(vm)[i] = back_project_at(d[i], x, y, K);
}

__global__ void k_nmap(const float *__restrict__ vm, float *__restrict__ nm, int w, int h) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= w || y >= h) return;
    reinterpret_cast<float4 *>(nm)[(size_t)y * w + x] = nmap_at(VertexAt{reinterpret_cast<const float4 *>(vm), w}, w, h, x, y);
}

// ------------------------------------------------------------ fused frame passes
//
// A pass is up to kMaxParts independent parts in one launch; each part tiles
// its own image with 16x16 workgroups, one lane per output pixel, and the
// workgroups of the parts follow each other in the 1-D grid.

constexpr int kMaxParts = 4;
enum PartKind { kFrameIn, kPyramid, kLevelMaps, kNmap };

struct Part {
    int kind, w, h, wg_begin;  // w, h: the part's output size
    int half;                  // FP16 intensity / gradient / vertex / normal planes
    const void *in0, *in1, *in2, *in3;
    void *out0, *out1, *out2, *out3, *out4;
    lf_intrinsics K;
    float scale;
};
struct Pass {
    Part part[kMaxParts];
    int nparts;
    int total_wg;  // workgroups of all parts
};

// The windowed parts (frame, pyramid, level maps, normal maps) of a 16x16
// output tile with their windows staged in LDS: each source pixel of the tile
// and its apron is read (and converted or back-projected) once instead of by
// every lane whose window holds it, and the window taps need no border
// arithmetic.  The values and the operation order are those of the per-pixel
// functions (bilateral_at, pyr_down_at, back_project_at, nmap_at), so the
// outputs are the same.
constexpr int kTile = 16;
constexpr int kBilApron = 2, kBilSide = kTile + 2 * kBilApron;  // 5x5 bilateral window
constexpr int kNrmApron = 1, kNrmSide = kTile + 2 * kNrmApron;  // 3x3 (cross) normal window
constexpr int kPyrSide = 2 * kTile + 4;  // the 2x source rows / columns of a pyrDown tile, 5-tap apron
union TileLds {
    float depth[kBilSide * kBilSide];
    float4 vert[kNrmSide * kNrmSide];
    float src[kPyrSide * kPyrSide];
};

// bilateral_at over the staged depths: the window of tile pixel (cx, cy) in
// the apron's coordinates, reflect-101 already applied when staging.
template <bool kFast>
LF_DEV float bilateral_tile(const float *s, int cx, int cy) {
    const float c = s[cy * kBilSide + cx];
    float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const float v = s[(cy + dy) * kBilSide + cx + dx];
            const float d = v - c;
            const float arg = (float)(dx * dx + dy * dy) * -0.5f + d * d * -0.5f;
            const float wt = kFast ? __expf(arg) : expf(arg);
            s1 += v * wt;
            s2 += wt;
        }
    }
    return s1 / s2;
}

// kFrameIn for the tile at (x0, y0).
template <bool kHalf>
LF_DEV void frame_in_tile(const Part &p, int x0, int y0, TileLds &lds) {
    const int w = p.w, h = p.h;
    const uint16_t *d16 = static_cast<const uint16_t *>(p.in0);
    for (int k = threadIdx.x; k < kBilSide * kBilSide; k += blockDim.x) {
        const int sx = reflect101(x0 - kBilApron + k % kBilSide, w), sy = reflect101(y0 - kBilApron + k / kBilSide, h);
        lds.depth[k] = depth_value(d16[(size_t)sy * w + sx], p.scale);
    }
    __syncthreads();
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4, x = x0 + lx, y = y0 + ly;
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * w + x;
    const unsigned char *rgb = static_cast<const unsigned char *>(p.in1);
    static_cast<float *>(p.out0)[i] = lds.depth[(ly + kBilApron) * kBilSide + lx + kBilApron];
    static_cast<float *>(p.out1)[i] = bilateral_tile<kHalf>(lds.depth, lx + kBilApron, ly + kBilApron);
    const unsigned char r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];
    store1<kHalf>(p.out2, i, grey_of(r, g, b));
    unsigned char *c = static_cast<unsigned char *>(p.out3) + 3 * i;
    c[0] = r; c[1] = g; c[2] = b;
}

// kPyramid for the tile at (x0, y0): the depth resize (resize_at), the
// intensity pyrDown from the tile's staged source window (reflect-101 applied
// when staging), in pyr_down_at's operation order.
template <bool kHalf>
LF_DEV void pyramid_tile(const Part &p, int x0, int y0, TileLds &lds) {
    const int w = p.w, h = p.h, sw = 2 * w, sh = 2 * h;
    const PlaneAt<kHalf> in{p.in1, sw};
    for (int k = threadIdx.x; k < kPyrSide * kPyrSide; k += blockDim.x) {
        const int sx = reflect101(2 * x0 - 2 + k % kPyrSide, sw), sy = reflect101(2 * y0 - 2 + k / kPyrSide, sh);
        lds.src[k] = in(sx, sy);
    }
    __syncthreads();
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4, x = x0 + lx, y = y0 + ly;
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * w + x;
    static_cast<float *>(p.out0)[i] = resize_at(static_cast<const float *>(p.in0), sw, x, y);
    const float *s = lds.src + (2 * ly) * kPyrSide + 2 * lx;  // tap (j, k): s[j * kPyrSide + k]
    float rs[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        const float *r = s + j * kPyrSide;
        rs[j] = (r[0] + r[4]) + 4.0f * (r[1] + r[3]) + 6.0f * r[2];
    }
    store1<kHalf>(p.out1, i, ((rs[0] + rs[4]) + 4.0f * (rs[1] + rs[3]) + 6.0f * rs[2]) * (1.0f / 256.0f));
}

// nmap_at of pixel (x, y), tile pixel (lx, ly), from the staged vertices:
// its neighbours (min(x+1, w), y), (max(x-1, 0), y), (x, min(y+1, h)),
// (x, max(y-1, 0)) are the tile's, staged with their coordinates clamped
// into the image.
template <bool kFast>
LF_DEV float4 normal_tile(const float4 *v, int lx, int ly, int x, int y, int w, int h) {
    const float qnan = __int_as_float(0x7fc00000);
    if (x >= w - 1 || y >= h - 1) return make_float4(qnan, qnan, qnan, qnan);
    const int cx = lx + kNrmApron, cy = ly + kNrmApron;
    const float4 a = v[cy * kNrmSide + cx + 1], b = v[cy * kNrmSide + cx - 1], c = v[(cy + 1) * kNrmSide + cx],
                 e = v[(cy - 1) * kNrmSide + cx];
    const f3 cr = cross(f3{a.x - b.x, a.y - b.y, a.z - b.z}, f3{c.x - e.x, c.y - e.y, c.z - e.z});
    const f3 n = kFast ? mul(cr, __builtin_amdgcn_rsqf(dot(cr, cr))) : normalised(cr);
    return make_float4(n.x, n.y, n.z, 1.0f);
}

// kNmap for the tile at (x0, y0): the stored vertices of the tile and a
// one-pixel apron staged once.  out1 / out2 (when set): the tile's 8x8
// half-size and 4x4 quarter-size vertices (kVmapDown's and kVmapDown2's
// values) from the same staged vertices, so the vertex map is read once.
template <bool kHalf>
LF_DEV void nmap_tile(const Part &p, int x0, int y0, TileLds &lds) {
    const int w = p.w, h = p.h;
    const float4 *vm = static_cast<const float4 *>(p.in0);
    for (int k = threadIdx.x; k < kNrmSide * kNrmSide; k += blockDim.x) {
        const int sx = min(max(x0 - kNrmApron + k % kNrmSide, 0), w - 1),
                  sy = min(max(y0 - kNrmApron + k / kNrmSide, 0), h - 1);
        lds.vert[k] = vm[(size_t)sy * w + sx];
    }
    __syncthreads();
    // tile pixel (tx, ty) of the staged vertices (in-range pixels only)
    const auto at = [&](int tx, int ty) { return lds.vert[(ty + kNrmApron) * kNrmSide + tx + kNrmApron]; };
    const int t = threadIdx.x;
    if (p.out1 && t < 64) {  // half-size pixel (x0 / 2 + i, y0 / 2 + j)
        const int i = t & 7, j = t >> 3, hx = x0 / 2 + i, hy = y0 / 2 + j, hw = w / 2, hh = h / 2;
        if (hx < hw && hy < hh)
            store4<false>(p.out1, (size_t)hy * hw + hx, resize4_at(at, i, j));
    } else if (p.out2 && t >= 64 && t < 80) {  // quarter-size pixel (x0 / 4 + i, y0 / 4 + j)
        const int i = (t - 64) & 3, j = (t - 64) >> 2, qx = x0 / 4 + i, qy = y0 / 4 + j, qw = w / 4, qh = h / 4;
        if (qx < qw && qy < qh)
            store4<false>(p.out2, (size_t)qy * qw + qx, resize4_quarter_at(at, i, j));
    }
    const int lx = t & 15, ly = t >> 4, x = x0 + lx, y = y0 + ly;
    if (x >= w || y >= h) return;
    store4<kHalf>(p.out0, (size_t)y * w + x, normal_tile<kHalf>(lds.vert, lx, ly, x, y, w, h));
}

// kLevelMaps for the tile at (x0, y0): the vertices of the tile and a
// one-pixel apron (coordinates clamped into the image, as nmap_at's
// neighbours are) back-projected once into LDS.
template <bool kHalf>
LF_DEV void level_maps_tile(const Part &p, int x0, int y0, TileLds &lds) {
    const int w = p.w, h = p.h;
    const float *d = static_cast<const float *>(p.in0);
    for (int k = threadIdx.x; k < kNrmSide * kNrmSide; k += blockDim.x) {
        const int sx = min(max(x0 - kNrmApron + k % kNrmSide, 0), w - 1),
                  sy = min(max(y0 - kNrmApron + k / kNrmSide, 0), h - 1);
        lds.vert[k] = back_project_at(d[(size_t)sy * w + sx], sx, sy, p.K);
    }
    __syncthreads();
    const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4, x = x0 + lx, y = y0 + ly;
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * w + x;
    float gx, gy;
    derivative_at(PlaneAt<kHalf>{p.in1, w}, w, h, x, y, gx, gy);
    store1<kHalf>(p.out0, i, gx);
    store1<kHalf>(p.out1, i, gy);
    store4<false>(p.out2, i, lds.vert[(ly + kNrmApron) * kNrmSide + lx + kNrmApron]);  // vertices stay FP32
    const float4 n = normal_tile<kHalf>(lds.vert, lx, ly, x, y, w, h);
    store4<kHalf>(p.out3, i, n);
    if (p.out4) {  // level 0: the pixel's integration sample
        const float dist = static_cast<const float *>(p.in2)[i];
        const unsigned char *col = static_cast<const unsigned char *>(p.in3) + 3 * i;
        static_cast<float4 *>(p.out4)[i] = integration_sample<kHalf>(dist, n, col);
    }
}

// Part p's workgroup t (the tile it covers); every lane reaches the part's
// barrier.  p.half (mixed precision): the intensity, gradient and normal
// planes are FP16 (read and written; depth, vertices, colour and samples are
// not), and the bilateral weights, normals and sample weights use the
// hardware exponential, reciprocal square root, sine and reciprocal.
LF_DEV void part_wg(const Part &p, int t, TileLds &lds) {
    const int tiles_x = (p.w + 15) >> 4;
    const int x0 = (t % tiles_x) * kTile, y0 = (t / tiles_x) * kTile;
    if (p.kind == kFrameIn) {
        if (p.half) frame_in_tile<true>(p, x0, y0, lds);
        else frame_in_tile<false>(p, x0, y0, lds);
        return;
    }
    if (p.kind == kLevelMaps) {
        if (p.half) level_maps_tile<true>(p, x0, y0, lds);
        else level_maps_tile<false>(p, x0, y0, lds);
        return;
    }
    if (p.kind == kNmap) {
        if (p.half) nmap_tile<true>(p, x0, y0, lds);
        else nmap_tile<false>(p, x0, y0, lds);
        return;
    }
    if (p.kind == kPyramid) {
        if (p.half) pyramid_tile<true>(p, x0, y0, lds);
        else pyramid_tile<false>(p, x0, y0, lds);
    }
}

// Workgroup wg of a pass.
LF_DEV void frame_pass_wg(const Pass &ps, int wg, TileLds &lds) {
    int pi = 0;
#pragma unroll
    for (int k = 1; k < kMaxParts; ++k)
        if (k < ps.nparts && wg >= ps.part[k].wg_begin) pi = k;
    const Part &p = ps.part[pi];
    part_wg(p, wg - p.wg_begin, lds);
}

__global__ __launch_bounds__(256) void k_frame_pass(Pass ps) {
    __shared__ TileLds lds;
    frame_pass_wg(ps, blockIdx.x, lds);
}

// One pass of each of gridDim.y images: row y of the grid runs table[y] (its
// part's pointers asserted global: they are read from memory).
__global__ __launch_bounds__(256) void k_frame_pass_batch(const Pass *__restrict__ table) {
    __shared__ TileLds lds;
    const Pass &ps = table[blockIdx.y];
    const int wg = blockIdx.x;
    if (wg >= ps.total_wg) return;
    int pi = 0;
#pragma unroll
    for (int k = 1; k < kMaxParts; ++k)
        if (k < ps.nparts && wg >= ps.part[k].wg_begin) pi = k;
    Part p = ps.part[pi];
    p.in0 = global_ptr(p.in0); p.in1 = global_ptr(p.in1); p.in2 = global_ptr(p.in2); p.in3 = global_ptr(p.in3);
    p.out0 = global_ptr(p.out0); p.out1 = global_ptr(p.out1); p.out2 = global_ptr(p.out2);
    p.out3 = global_ptr(p.out3); p.out4 = global_ptr(p.out4);
    part_wg(p, wg - p.wg_begin, lds);
}

dim3 grid16(int w, int h) { return dim3((w + 15) / 16, (h + 15) / 16); }
int grid1d(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>(4096, (n + 255) / 256)); }

int check_img(const void *a, const void *b, int w, int h) {
    if (!a || !b || w <= 0 || h <= 0) return lf_set_error(LF_ERR_ARG, "bad image arguments (%dx%d)", w, h);
    return LF_OK;
}

Part make_part(int kind, int w, int h) {
    Part p = {};
    p.kind = kind;
    p.w = w;
    p.h = h;
    return p;
}

// Workgroup offsets of the parts and the pass's total.
void seal_pass(Pass &ps, int half) {
    int wg = 0;
    for (int k = 0; k < ps.nparts; ++k) {
        ps.part[k].half = half;
        ps.part[k].wg_begin = wg;
        wg += ((ps.part[k].w + 15) / 16) * ((ps.part[k].h + 15) / 16);
    }
    ps.total_wg = wg;
}

int launch_pass(const Pass &ps, hipStream_t st) {
    if (ps.total_wg == 0) return LF_OK;
    k_frame_pass<<<ps.total_wg, 256, 0, st>>>(ps);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

int check_pyramid(int levels, const lf_intrinsics *K) {
    if (levels < 1 || levels > 8 || !K) return lf_set_error(LF_ERR_ARG, "bad pyramid (%d levels)", levels);
    for (int l = 0; l < levels; ++l) {
        if (K[l].width <= 0 || K[l].height <= 0) return lf_set_error(LF_ERR_ARG, "bad level %d size", l);
        if (l > 0 && (K[l].width * 2 != K[l - 1].width || K[l].height * 2 != K[l - 1].height))
            return lf_set_error(LF_ERR_ARG, "level %d is not half of level %d (%dx%d, %dx%d)", l, l - 1, K[l].width,
                                K[l].height, K[l - 1].width, K[l - 1].height);
    }
    return LF_OK;
}

// DeviceImage::upload as levels + 1 passes: the frame pass (raw depth,
// bilateral depth, intensity, colour copy), then per level its maps
// (derivatives, vertices, normals) together with the next level's pyramids.
int plan_upload(const uint16_t *depth_u16, const unsigned char *rgb, float depth_scale, int levels,
                const lf_intrinsics *K, const LfImageLevels &m, std::vector<Pass> &out) {
    if (int rc = check_pyramid(levels, K)) return rc;
    if (!depth_u16 || !rgb) return lf_set_error(LF_ERR_ARG, "null frame input");
    Pass ps = {};
    Part f = make_part(kFrameIn, K[0].width, K[0].height);
    f.in0 = depth_u16; f.in1 = rgb; f.scale = depth_scale;
    f.out0 = m.depth_raw; f.out1 = m.depth[0]; f.out2 = m.intensity[0]; f.out3 = m.rgb;
    ps.part[ps.nparts++] = f;
    seal_pass(ps, m.half_maps);
    out.push_back(ps);
    for (int l = 0; l < levels; ++l) {
        ps = Pass{};
        Part p = make_part(kLevelMaps, K[l].width, K[l].height);
        p.in0 = m.depth[l]; p.in1 = m.intensity[l]; p.K = K[l];
        p.out0 = m.dx[l]; p.out1 = m.dy[l]; p.out2 = m.vmap[l]; p.out3 = m.nmap[l];
        if (l == 0 && m.samples) {
            p.in2 = m.depth_raw; p.in3 = m.rgb; p.out4 = m.samples;
        }
        ps.part[ps.nparts++] = p;
        if (l + 1 < levels) {
            Part q = make_part(kPyramid, K[l + 1].width, K[l + 1].height);
            q.in0 = m.depth[l]; q.in1 = m.intensity[l]; q.out0 = m.depth[l + 1]; q.out1 = m.intensity[l + 1];
            ps.part[ps.nparts++] = q;
        }
        seal_pass(ps, m.half_maps);
        out.push_back(ps);
    }
    return LF_OK;
}

// DeviceImage::resize_device_map: the vertex pyramid of a raycast level-0 map
// and every level's normals, two levels per pass (a pass with base level b
// writes the normals still missing below b + 1 and, from the base level's
// staged tiles, the vertices of levels b + 1 and b + 2, the latter resized
// twice).
int plan_refresh(int levels, const lf_intrinsics *K, const LfImageLevels &m, std::vector<Pass> &out) {
    if (int rc = check_pyramid(levels, K)) return rc;
    int nmap_done = 0;
    for (int b = 0; nmap_done < levels; b += 2) {
        Pass ps = {};
        for (; nmap_done <= b && nmap_done < levels; ++nmap_done) {
            Part p = make_part(kNmap, K[nmap_done].width, K[nmap_done].height);
            p.in0 = m.vmap[nmap_done]; p.out0 = m.nmap[nmap_done];
            if (nmap_done == b) {  // the base level's tiles also write levels b + 1 and b + 2
                if (b + 1 < levels) p.out1 = m.vmap[b + 1];
                if (b + 2 < levels) p.out2 = m.vmap[b + 2];
            }
            ps.part[ps.nparts++] = p;
        }
        seal_pass(ps, m.half_maps);
        out.push_back(ps);
    }
    return LF_OK;
}

int launch_all(const std::vector<Pass> &passes, hipStream_t st) {
    for (const Pass &ps : passes)
        if (int rc = launch_pass(ps, st)) return rc;
    return LF_OK;
}

void append_bytes(const std::vector<Pass> &passes, std::vector<unsigned char> &out) {
    const unsigned char *b = reinterpret_cast<const unsigned char *>(passes.data());
    out.insert(out.end(), b, b + passes.size() * sizeof(Pass));
}

}  // namespace

int lf_pyramid_upload(const uint16_t *depth_u16, const unsigned char *rgb, float depth_scale, int levels,
                      const lf_intrinsics *K, const LfImageLevels &m, hipStream_t st) {
    std::vector<Pass> passes;
    if (int rc = plan_upload(depth_u16, rgb, depth_scale, levels, K, m, passes)) return rc;
    return launch_all(passes, st);
}

int lf_pyramid_refresh(int levels, const lf_intrinsics *K, const LfImageLevels &m, hipStream_t st) {
    std::vector<Pass> passes;
    if (int rc = plan_refresh(levels, K, m, passes)) return rc;
    return launch_all(passes, st);
}

size_t lf_pass_bytes(void) { return sizeof(Pass); }

int lf_pyramid_upload_plan(const uint16_t *depth_u16, const unsigned char *rgb, float depth_scale, int levels,
                           const lf_intrinsics *K, const LfImageLevels &m, std::vector<unsigned char> &out) {
    std::vector<Pass> passes;
    if (int rc = plan_upload(depth_u16, rgb, depth_scale, levels, K, m, passes)) return rc;
    append_bytes(passes, out);
    return LF_OK;
}

int lf_pyramid_refresh_plan(int levels, const lf_intrinsics *K, const LfImageLevels &m, std::vector<unsigned char> &out) {
    std::vector<Pass> passes;
    if (int rc = plan_refresh(levels, K, m, passes)) return rc;
    append_bytes(passes, out);
    return LF_OK;
}

int lf_pass_workgroups(const void *pass) { return static_cast<const Pass *>(pass)->total_wg; }

void lf_pass_skip(void *pass) { *static_cast<Pass *>(pass) = Pass{}; }

int lf_pass_batch(const void *passes, int n, int max_wg, hipStream_t st) {
    if (!passes || n <= 0 || n > 65535 || max_wg < 0) return lf_set_error(LF_ERR_ARG, "bad pass batch of %d", n);
    if (max_wg == 0) return LF_OK;
    k_frame_pass_batch<<<dim3(max_wg, n), 256, 0, st>>>(static_cast<const Pass *>(passes));
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" int lf_depth_to_float(const uint16_t *in, float *out, int cols, int rows, float scale, void *stream) {
    if (int rc = check_img(in, out, cols, rows)) return rc;
    const size_t n = (size_t)cols * rows;
    k_depth_to_float<<<grid1d(n), 256, 0, lf_stream(stream)>>>(in, out, (int)n, scale);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" int lf_bilateral_filter(const float *in, float *out, int cols, int rows, void *stream) {
    if (int rc = check_img(in, out, cols, rows)) return rc;
    k_bilateral<<<grid16(cols, rows), 256, 0, lf_stream(stream)>>>(in, out, cols, rows);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" int lf_pyr_down_depth(const float *in, float *out, int cols, int rows, void *stream) {
    if (int rc = check_img(in, out, cols, rows)) return rc;
    if ((cols & 1) || (rows & 1)) return lf_set_error(LF_ERR_ARG, "pyr_down_depth needs even sizes (%dx%d)", cols, rows);
    k_resize_half<1><<<grid16(cols / 2, rows / 2), 256, 0, lf_stream(stream)>>>(in, out, cols, rows);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" int lf_pyr_down_vmap(const float *in, float *out, int cols, int rows, void *stream) {
    if (int rc = check_img(in, out, cols, rows)) return rc;
    if ((cols & 1) || (rows & 1)) return lf_set_error(LF_ERR_ARG, "pyr_down_vmap needs even sizes (%dx%d)", cols, rows);
    k_resize_half<4><<<grid16(cols / 2, rows / 2), 256, 0, lf_stream(stream)>>>(in, out, cols, rows);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" int lf_pyr_down_image(const float *in, float *out, int cols, int rows, void *stream) {
    if (int rc = check_img(in, out, cols, rows)) return rc;
    k_pyr_down<<<grid16((cols + 1) / 2, (rows + 1) / 2), 256, 0, lf_stream(stream)>>>(in, out, cols, rows);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" int lf_rgb_to_intensity(const unsigned char *rgb, float *out, int cols, int rows, void *stream) {
    if (int rc = check_img(rgb, out, cols, rows)) return rc;
    const size_t n = (size_t)cols * rows;
    k_grey<<<grid1d(n), 256, 0, lf_stream(stream)>>>(rgb, out, (int)n);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" int lf_compute_derivative(const float *image, float *dx, float *dy, int cols, int rows, void *stream) {
    if (int rc = check_img(image, dx, cols, rows)) return rc;
    if (!dy) return lf_set_error(LF_ERR_ARG, "null dy");
    k_derivative<<<grid16(cols, rows), 256, 0, lf_stream(stream)>>>(image, dx, dy, cols, rows);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" int lf_back_project_depth(const float *depth, float *vmap, int cols, int rows, const lf_intrinsics *K,
                                     void *stream) {
    if (int rc = check_img(depth, vmap, cols, rows)) return rc;
    if (!K) return lf_set_error(LF_ERR_ARG, "null intrinsics");
    k_back_project<<<grid16(cols, rows), 256, 0, lf_stream(stream)>>>(depth, vmap, cols, rows, *K);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" int lf_compute_nmap(const float *vmap, float *nmap, int cols, int rows, void *stream) {
    if (int rc = check_img(vmap, nmap, cols, rows)) return rc;
    k_nmap<<<grid16(cols, rows), 256, 0, lf_stream(stream)>>>(vmap, nmap, cols, rows);
    LF_LAUNCH_CHECK();
    return LF_OK;
}
