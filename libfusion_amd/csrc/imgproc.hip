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
#include "lf_common.h"

using namespace lf;

namespace {

LF_DEV int reflect101(int i, int n) {
    if (n == 1) return 0;
    while (i < 0 || i >= n) i = i < 0 ? -i : 2 * n - 2 - i;
    return i;
}

__global__ void k_depth_to_float(const uint16_t *__restrict__ in, float *__restrict__ out, int n, float scale) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = (float)in[i] * scale;
}

// cv::cuda::bilateralFilter(src, dst, 5, 1, 1): 5x5 window, reflect-101 border,
// w = exp(-(dx^2+dy^2)/(2*1^2) - (v-c)^2/(2*1^2)); zero depths are not masked.
__global__ void k_bilateral(const float *__restrict__ in, float *__restrict__ out, int w, int h) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= w || y >= h) return;
    const float c = in[(size_t)y * w + x];
    float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
        const float *row = in + (size_t)reflect101(y + dy, h) * w;
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            const float v = row[reflect101(x + dx, w)];
            const float d = v - c;
            const float wt = expf((float)(dx * dx + dy * dy) * -0.5f + d * d * -0.5f);
            s1 += v * wt;
            s2 += wt;
        }
    }
    out[(size_t)y * w + x] = s1 / s2;
}

// cv::cuda::resize(x0.5, INTER_LINEAR) with the pixel-centre convention: the
// mean of the 2x2 source block, NaN propagating.  ch = 1 (depth) or 4 (vmap).
template <int kCh>
__global__ void k_resize_half(const float *__restrict__ in, float *__restrict__ out, int w, int h) {
    const int ow = w / 2, oh = h / 2;
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= ow || y >= oh) return;
#pragma unroll
    for (int k = 0; k < kCh; ++k) {
        const float a = in[((size_t)(2 * y) * w + 2 * x) * kCh + k], b = in[((size_t)(2 * y) * w + 2 * x + 1) * kCh + k];
        const float c = in[((size_t)(2 * y + 1) * w + 2 * x) * kCh + k],
                    d = in[((size_t)(2 * y + 1) * w + 2 * x + 1) * kCh + k];
        out[((size_t)y * ow + x) * kCh + k] = ((a + b) + (c + d)) * 0.25f;
    }
}

// cv::cuda::pyrDown: [1 4 6 4 1]^2/256 Gaussian at even pixels, reflect-101.
__global__ void k_pyr_down(const float *__restrict__ in, float *__restrict__ out, int w, int h) {
    const int ow = (w + 1) / 2, oh = (h + 1) / 2;
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= ow || y >= oh) return;
    const int x0 = reflect101(2 * x - 2, w), x1 = reflect101(2 * x - 1, w), x2 = 2 * x, x3 = reflect101(2 * x + 1, w),
              x4 = reflect101(2 * x + 2, w);
    float rs[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
        const float *row = in + (size_t)reflect101(2 * y + j - 2, h) * w;
        rs[j] = (row[x0] + row[x4]) + 4.0f * (row[x1] + row[x3]) + 6.0f * row[x2];
    }
    out[(size_t)y * ow + x] = ((rs[0] + rs[4]) + 4.0f * (rs[1] + rs[3]) + 6.0f * rs[2]) * (1.0f / 256.0f);
}

// cvtColor(RGB2GRAY) on float: 0.299 R + 0.587 G + 0.114 B.
__global__ void k_grey(const unsigned char *__restrict__ rgb, float *__restrict__ out, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        out[i] = (float)rgb[3 * i] * 0.299f + (float)rgb[3 * i + 1] * 0.587f + (float)rgb[3 * i + 2] * 0.114f;
}

// computeDerivativeK (cuda_imgproc.cu:217-234); last row / column 0.
__global__ void k_derivative(const float *__restrict__ I, float *__restrict__ dx, float *__restrict__ dy, int w, int h) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * w + x;
    if (x >= w - 1 || y >= h - 1) {
        dx[i] = 0.0f;
        dy[i] = 0.0f;
        return;
    }
    const int x10 = max(x - 1, 0), x01 = min(x + 1, w), y10 = max(y - 1, 0), y01 = min(y + 1, h);
    dx[i] = (I[(size_t)y * w + x01] - I[(size_t)y * w + x10]) * 0.5f;
    dy[i] = (I[(size_t)y01 * w + x] - I[(size_t)y10 * w + x]) * 0.5f;
}

// backProjectDepthK (cuda_imgproc.cu:249-263).
__global__ void k_back_project(const float *__restrict__ d, float *__restrict__ vm, int w, int h, lf_intrinsics K) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * w + x;
    const float z = d[i];
    const float qnan = __int_as_float(0x7fc00000);
    float4 o = make_float4(qnan, qnan, qnan, -1.0f);
    if (z > 0.3f && z < 5.0f) o = make_float4(z * ((float)x - K.cx) * K.invfx, z * ((float)y - K.cy) * K.invfy, z, 1.0f);
    reinterpret_cast<float4 *>(vm)[i] = o;
}

// computeNMapK (cuda_imgproc.cu:276-294); last row / column NaN.
__global__ void k_nmap(const float *__restrict__ vm, float *__restrict__ nm, int w, int h) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= w || y >= h) return;
    const size_t i = (size_t)y * w + x;
    const float qnan = __int_as_float(0x7fc00000);
    if (x >= w - 1 || y >= h - 1) {
        reinterpret_cast<float4 *>(nm)[i] = make_float4(qnan, qnan, qnan, qnan);
        return;
    }
    const float4 *V = reinterpret_cast<const float4 *>(vm);
    const int x10 = max(x - 1, 0), x01 = min(x + 1, w), y10 = max(y - 1, 0), y01 = min(y + 1, h);
    const float4 a = V[(size_t)y * w + x01], b = V[(size_t)y * w + x10], c = V[(size_t)y01 * w + x],
                 e = V[(size_t)y10 * w + x];
    const f3 n = normalised(cross(f3{a.x - b.x, a.y - b.y, a.z - b.z}, f3{c.x - e.x, c.y - e.y, c.z - e.z}));
    reinterpret_cast<float4 *>(nm)[i] = make_float4(n.x, n.y, n.z, 1.0f);
}

dim3 grid16(int w, int h) { return dim3((w + 15) / 16, (h + 15) / 16); }
int grid1d(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>(4096, (n + 255) / 256)); }

int check_img(const void *a, const void *b, int w, int h) {
    if (!a || !b || w <= 0 || h <= 0) return lf_set_error(LF_ERR_ARG, "bad image arguments (%dx%d)", w, h);
    return LF_OK;
}

}  // namespace

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
