// Shaded / textured scene render, image warp and semi-dense images.
//
// Reference: src/tracking/cuda_imgproc.cu (renderPoint / renderSceneK /
// renderSceneTexturedK :17-108, ToSemiDenseImageK / build_semi_dense_pyramid
// :110-153, interpolate_bilinear / warp_image_kernel :155-195), consumed by
// DeviceImage::get_rendered_image / get_rendered_scene_textured
// (device_image.cpp:119-129) and System::get_shaded_depth / get_rendered_scene /
// get_rendered_scene_textured (system.cpp:132-146).  Arithmetic in the
// reference's operation order (the HIP files are built with -ffp-contract=off).
#include "lf_common.h"
#include "lf_internal.h"

using namespace lf;

namespace {

// renderPoint (cuda_imgproc.cu:17-48): Phong shading with a light at
// light_pos in the camera frame; pixel = (ambient, diffuse, specular) colour.
LF_DEV uchar4 render_point(f3 point, f3 normal, f3 pixel, f3 light_pos) {
    f3 colour = {4.f / 255.f, 2.f / 255.f, 2.f / 255.f};
    if (!isnan(point.x)) {
        const float Ka = 0.3f, Kd = 0.5f, Ks = 0.2f, n = 20.f;
        const float Ax = pixel.x, Dx = pixel.y, Sx = pixel.z, Lx = 1.f;
        const f3 L = normalised(sub(light_pos, point));
        const f3 V = normalised(sub(f3{0.f, 0.f, 0.f}, point));
        const f3 R = normalised(sub(mul(mul(normal, 2.0f), dot(normal, L)), L));
        const float Ix = Ax * Ka * Dx + Lx * Kd * Dx * fmaxf(0.f, dot(normal, L)) +
                         Lx * Ks * Sx * powf(fmaxf(0.f, dot(R, V)), n);
        colour = f3{Ix, Ix, Ix};
    }
    return make_uchar4((unsigned char)(__saturatef(colour.x) * 255.f), (unsigned char)(__saturatef(colour.y) * 255.f),
                       (unsigned char)(__saturatef(colour.z) * 255.f), 255);
}

LF_DEV float4 half4_at(const void *p, size_t i) {
    const uint2 u = static_cast<const uint2 *>(p)[i];
    return make_float4(half_bits_to_float((unsigned short)u.x), half_bits_to_float((unsigned short)(u.x >> 16)),
                       half_bits_to_float((unsigned short)u.y), half_bits_to_float((unsigned short)(u.y >> 16)));
}

// renderSceneK / renderSceneTexturedK (cuda_imgproc.cu:50-108); rgb null: plain.
// kHalf: the normal map is 4 x FP16.
template <bool kHalf>
__global__ void k_render(const void *__restrict__ vmap, const void *__restrict__ nmap,
                         const unsigned char *__restrict__ rgb, int cols, int rows, uchar4 *__restrict__ out) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= cols || y >= rows) return;
    const size_t i = (size_t)y * cols + x;
    const float4 v = static_cast<const float4 *>(vmap)[i];  // vertices are FP32 in both precisions
    const float4 n = kHalf ? half4_at(nmap, i) : static_cast<const float4 *>(nmap)[i];
    f3 pixel = {1.f, 1.f, 1.f};
    if (rgb) pixel = f3{(float)rgb[3 * i] / 255.f, (float)rgb[3 * i + 1] / 255.f, (float)rgb[3 * i + 2] / 255.f};
    out[i] = render_point(f3{v.x, v.y, v.z}, f3{n.x, n.y, n.z}, pixel, f3{5.f, 5.f, 5.f});
}

// ToSemiDenseImageK (cuda_imgproc.cu:110-128).
__global__ void k_semi_dense(const float *__restrict__ image, const float *__restrict__ dx,
                             const float *__restrict__ dy, int cols, int rows, float th_dx, float th_dy,
                             float *__restrict__ semi) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= cols || y >= rows) return;
    const size_t i = (size_t)y * cols + x;
    semi[i] = (dx[i] > th_dx || dy[i] > th_dy) ? image[i] : 255.0f;
}

// interpolate_bilinear + warp_image_kernel (cuda_imgproc.cu:155-185): the
// source image sampled where the destination's vertices project under pose.
__global__ void k_warp(const unsigned char *__restrict__ src, const float4 *__restrict__ vmap_dst, int cols, int rows,
                       lf_intrinsics K, float p0, float p1, float p2, float p3, float p4, float p5, float p6,
                       float p7, float p8, float p9, float p10, float p11, unsigned char *__restrict__ dst) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= cols || y >= rows) return;
    const float P[12] = {p0, p1, p2, p3, p4, p5, p6, p7, p8, p9, p10, p11};
    const size_t i = (size_t)y * cols + x;
    const float4 v = vmap_dst[i];
    const f3 q = xform(P, f3{v.x, v.y, v.z});
    const float u = K.fx * q.x / q.z + K.cx, w = K.fy * q.y / q.z + K.cy;
    unsigned char out[3] = {0, 0, 0};
    if (u >= 1 && w >= 1 && u < (float)(cols - 1) && w < (float)(rows - 1)) {
        const int u0 = (int)floorf(u), v0 = (int)floorf(w);
        const float cx = u - (float)u0, cy = w - (float)v0;
        const unsigned char *a = src + 3 * ((size_t)v0 * cols + u0), *b = a + 3 * (size_t)cols;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float r = ((float)a[k] * (1 - cx) + (float)a[3 + k] * cx) * (1 - cy) +
                            ((float)b[k] * (1 - cx) + (float)b[3 + k] * cx) * cy;
            out[k] = (unsigned char)(int)r;
        }
    }
    dst[3 * i] = out[0]; dst[3 * i + 1] = out[1]; dst[3 * i + 2] = out[2];
}

dim3 grid16(int w, int h) { return dim3((w + 15) / 16, (h + 15) / 16); }

}  // namespace

int lf_render_scene_maps(const void *vmap, const void *nmap, int half_maps, const unsigned char *rgb, int cols,
                         int rows, unsigned char *rgba_out, hipStream_t st) {
    if (!vmap || !nmap || !rgba_out || cols <= 0 || rows <= 0) return lf_set_error(LF_ERR_ARG, "bad render arguments");
    uchar4 *out = reinterpret_cast<uchar4 *>(rgba_out);
    if (half_maps) k_render<true><<<grid16(cols, rows), 256, 0, st>>>(vmap, nmap, rgb, cols, rows, out);
    else k_render<false><<<grid16(cols, rows), 256, 0, st>>>(vmap, nmap, rgb, cols, rows, out);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" int lf_render_scene(const float *vmap, const float *nmap, const unsigned char *rgb, int cols, int rows,
                               unsigned char *rgba_out, void *stream) {
    return lf_render_scene_maps(vmap, nmap, 0, rgb, cols, rows, rgba_out, lf_stream(stream));
}

extern "C" int lf_semi_dense(const float *image, const float *dx, const float *dy, int cols, int rows, float th_dx,
                             float th_dy, float *semi_out, void *stream) {
    if (!image || !dx || !dy || !semi_out || cols <= 0 || rows <= 0)
        return lf_set_error(LF_ERR_ARG, "bad semi-dense arguments");
    k_semi_dense<<<grid16(cols, rows), 256, 0, lf_stream(stream)>>>(image, dx, dy, cols, rows, th_dx, th_dy, semi_out);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" int lf_warp_image(const unsigned char *src, const float *vmap_dst, int cols, int rows, const float *pose12,
                             const lf_intrinsics *K, unsigned char *dst, void *stream) {
    if (!src || !vmap_dst || !pose12 || !K || !dst || cols <= 0 || rows <= 0)
        return lf_set_error(LF_ERR_ARG, "bad warp arguments");
    const float *p = pose12;
    k_warp<<<grid16(cols, rows), 256, 0, lf_stream(stream)>>>(src, reinterpret_cast<const float4 *>(vmap_dst), cols,
                                                             rows, *K, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7],
                                                             p[8], p[9], p[10], p[11], dst);
    LF_LAUNCH_CHECK();
    return LF_OK;
}
