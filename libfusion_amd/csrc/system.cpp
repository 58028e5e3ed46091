// Host classes of the drop-in C++ layer and the lf_system_* C-ABI.
//
// Reference: src/tracking/device_image.cpp, src/voxel_hashing/voxel_hashing.cpp,
// src/tracking/icp_tracker.cpp, src/tracking/rgbd_odometry.cpp, src/system.cpp.
// Each method below enqueues kernels on its object's stream; nothing waits for
// the GPU inside a frame (the reference synchronises twice per Gauss-Newton
// iteration and once per integrate / raycast).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <new>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fusion_amd.hpp"
#include "lf_common.h"
#include "lf_internal.h"
#include "lf_track.h"

#define LF_RC(call)                 \
    do {                            \
        int rc_ = (call);           \
        if (rc_ != LF_OK) return rc_; \
    } while (0)

namespace fusion {

// ------------------------------------------------------------ small types

IntrinsicMatrix::IntrinsicMatrix() { lf_intrinsics_make(0, 0, 1, 1, 0, 0, this); invfx = invfy = 0; fx = fy = 0; }
IntrinsicMatrix::IntrinsicMatrix(int cols, int rows, float fx_, float fy_, float cx_, float cy_) {
    lf_intrinsics_make(cols, rows, fx_, fy_, cx_, cy_, this);
}
IntrinsicMatrix IntrinsicMatrix::pyr_down() const {
    IntrinsicMatrix k;
    lf_intrinsics_pyr_down(this, &k);
    return k;
}
void BuildIntrinsicPyramid(IntrinsicMatrix base, std::vector<IntrinsicMatrix> &pyr, int n) {
    pyr.clear();
    pyr.push_back(base);
    for (int i = 0; i < n - 1; ++i) pyr.push_back(pyr[i].pyr_down());
}

SE3d SE3d::exp(const double *xi) {
    const double *u = xi, *w = xi + 3;
    const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2], th = std::sqrt(th2);
    const double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    double W2[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) W2[i * 3 + j] = W[i * 3] * W[j] + W[i * 3 + 1] * W[3 + j] + W[i * 3 + 2] * W[6 + j];
    double A, B, C;
    if (th < 1e-8) { A = 1 - th2 / 6; B = 0.5 - th2 / 24; C = 1.0 / 6 - th2 / 120; }
    else { A = std::sin(th) / th; B = (1 - std::cos(th)) / th2; C = (th - std::sin(th)) / (th2 * th); }
    SE3d T;
    double V[9];
    for (int i = 0; i < 9; ++i) {
        const double I = (i % 4 == 0) ? 1 : 0;
        T.R[i] = I + A * W[i] + B * W2[i];
        V[i] = I + B * W[i] + C * W2[i];
    }
    for (int i = 0; i < 3; ++i) T.t[i] = V[i * 3] * u[0] + V[i * 3 + 1] * u[1] + V[i * 3 + 2] * u[2];
    return T;
}

void SE3d::log(double *xi) const {
    double c = std::max(-1.0, std::min(1.0, (R[0] + R[4] + R[8] - 1) * 0.5));
    double th = std::acos(c);
    const double v[3] = {R[7] - R[5], R[2] - R[6], R[3] - R[1]};
    double w[3];
    if (th < 1e-8) {
        for (int i = 0; i < 3; ++i) w[i] = 0.5 * v[i];
    } else if (M_PI - th < 1e-6) {
        int k = 0;
        for (int i = 1; i < 3; ++i)
            if (R[i * 4] > R[k * 4]) k = i;
        double ax[3];
        ax[k] = std::sqrt(std::max(0.0, (R[k * 4] + 1) * 0.5));
        for (int i = 0; i < 3; ++i)
            if (i != k) ax[i] = R[k * 3 + i] / (2 * ax[k]);
        for (int i = 0; i < 3; ++i) w[i] = ax[i] * th;
    } else {
        const double s = th / (2 * std::sin(th));
        for (int i = 0; i < 3; ++i) w[i] = s * v[i];
    }
    const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    th = std::sqrt(th2);
    const double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    double W2[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) W2[i * 3 + j] = W[i * 3] * W[j] + W[i * 3 + 1] * W[3 + j] + W[i * 3 + 2] * W[6 + j];
    const double D = th < 1e-8 ? 1.0 / 12 : (1 - th * std::sin(th) / (2 * (1 - std::cos(th)))) / th2;
    for (int i = 0; i < 3; ++i) {
        double s = 0;
        for (int j = 0; j < 3; ++j) s += ((i == j ? 1.0 : 0.0) - 0.5 * W[i * 3 + j] + D * W2[i * 3 + j]) * t[j];
        xi[i] = s;
        xi[3 + i] = w[i];
    }
}

SE3d SE3d::inverse() const {
    SE3d o;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) o.R[i * 3 + j] = R[j * 3 + i];
    for (int i = 0; i < 3; ++i) o.t[i] = -(o.R[i * 3] * t[0] + o.R[i * 3 + 1] * t[1] + o.R[i * 3 + 2] * t[2]);
    return o;
}

SE3d SE3d::operator*(const SE3d &b) const {
    SE3d c;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) c.R[i * 3 + j] = R[i * 3] * b.R[j] + R[i * 3 + 1] * b.R[3 + j] + R[i * 3 + 2] * b.R[6 + j];
        c.t[i] = R[i * 3] * b.t[0] + R[i * 3 + 1] * b.t[1] + R[i * 3 + 2] * b.t[2] + t[i];
    }
    return c;
}

void SE3d::matrix3x4(double *o) const {
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) o[i * 4 + j] = R[i * 3 + j];
        o[i * 4 + 3] = t[i];
    }
}

SE3d SE3d::from3x4(const double *m) {
    SE3d T;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) T.R[i * 3 + j] = m[i * 4 + j];
        T.t[i] = m[i * 4 + 3];
    }
    return T;
}

GpuImage::~GpuImage() {
    if (data_) (void)hipFree(data_);
}
GpuImage::GpuImage(GpuImage &&o) noexcept : cols(o.cols), rows(o.rows), elem(o.elem), data_(o.data_) { o.data_ = nullptr; }
GpuImage &GpuImage::operator=(GpuImage &&o) noexcept {
    if (this != &o) {
        if (data_) (void)hipFree(data_);
        data_ = o.data_; cols = o.cols; rows = o.rows; elem = o.elem;
        o.data_ = nullptr;
    }
    return *this;
}
void GpuImage::create(int c, int r, int e) {
    if (data_ && c == cols && r == rows && e == elem) return;
    if (data_) (void)hipFree(data_);
    data_ = nullptr;
    cols = c; rows = r; elem = e;
    if (hipMalloc(&data_, (size_t)c * r * e) != hipSuccess) {
        data_ = nullptr;
        throw std::bad_alloc();
    }
}

// ------------------------------------------------------------ DeviceImage

DeviceImage::DeviceImage(const IntrinsicMatrix &K, int num_pyr, hipStream_t stream, bool half_maps)
    : stream_(stream), half_maps_(half_maps) {
    if (num_pyr < 1 || num_pyr > 8) throw std::invalid_argument("DeviceImage: 1 to 8 pyramid levels");
    BuildIntrinsicPyramid(K, K_, num_pyr);
    depth_raw_.create(K.width, K.height, 4);
    image_.create(K.width, K.height, 3);
    samples_.create(K.width, K.height, 16);
    rendered_.create(K.width, K.height, 4);
    depth_.resize(num_pyr); intensity_.resize(num_pyr); dx_.resize(num_pyr); dy_.resize(num_pyr);
    vmap_.resize(num_pyr); nmap_.resize(num_pyr);
    for (int l = 0; l < num_pyr; ++l) {
        const int w = K_[l].width, h = K_[l].height;
        const int e1 = half_maps ? 2 : 4, e4 = 4 * e1;  // FP16 or FP32 components; vertices always FP32
        depth_[l].create(w, h, 4); intensity_[l].create(w, h, e1); dx_[l].create(w, h, e1); dy_[l].create(w, h, e1);
        vmap_[l].create(w, h, 16); nmap_[l].create(w, h, e4);
    }
}

LfImageLevels DeviceImage::levels_view() {
    LfImageLevels m = {};
    m.depth_raw = depth_raw_.ptr<float>();
    m.rgb = image_.ptr<unsigned char>();
    m.samples = samples_.ptr<float>();
    m.half_maps = half_maps_ ? 1 : 0;
    for (int l = 0; l < levels(); ++l) {
        m.depth[l] = depth_[l].ptr<float>();
        m.intensity[l] = intensity_[l].ptr<float>();
        m.dx[l] = dx_[l].ptr<float>();
        m.dy[l] = dy_[l].ptr<float>();
        m.vmap[l] = vmap_[l].ptr<float>();
        m.nmap[l] = nmap_[l].ptr<float>();
    }
    return m;
}

// upload (device_image.cpp:146-181): depth to metres, colour copy, intensity,
// bilateral filter, depth / intensity pyramids, then per level derivatives,
// vertices and normals — fused into levels + 1 launches.
int DeviceImage::upload(const uint16_t *depth_u16, const unsigned char *rgb, float depth_scale) {
    std::vector<lf_intrinsics> K(K_.begin(), K_.end());
    return lf_pyramid_upload(depth_u16, rgb, depth_scale, levels(), K.data(), levels_view(), stream_);
}

// resize_device_map (device_image.cpp:188-199): vertex pyramid of the level-0
// map and the normals of every level.
int DeviceImage::resize_device_map() {
    std::vector<lf_intrinsics> K(K_.begin(), K_.end());
    return lf_pyramid_refresh(levels(), K.data(), levels_view(), stream_);
}

int DeviceImage::plan_upload(const uint16_t *depth_u16, const unsigned char *rgb, float depth_scale,
                             std::vector<unsigned char> &out) {
    std::vector<lf_intrinsics> K(K_.begin(), K_.end());
    return lf_pyramid_upload_plan(depth_u16, rgb, depth_scale, levels(), K.data(), levels_view(), out);
}

int DeviceImage::plan_refresh(std::vector<unsigned char> &out) {
    std::vector<lf_intrinsics> K(K_.begin(), K_.end());
    return lf_pyramid_refresh_plan(levels(), K.data(), levels_view(), out);
}

const GpuImage &DeviceImage::get_rendered_image() {
    lf_render_scene_maps(vmap_[0].ptr<void>(), nmap_[0].ptr<void>(), half_maps_, nullptr, K_[0].width, K_[0].height,
                         rendered_.ptr<unsigned char>(), stream_);
    return rendered_;
}

const GpuImage &DeviceImage::get_rendered_scene_textured() {
    lf_render_scene_maps(vmap_[0].ptr<void>(), nmap_[0].ptr<void>(), half_maps_, image_.ptr<unsigned char>(),
                         K_[0].width, K_[0].height, rendered_.ptr<unsigned char>(), stream_);
    return rendered_;
}

int DeviceImage::set_vmap(const float *vmap) {
    if (!vmap) return lf_set_error(LF_ERR_ARG, "null vertex map");
    if (half_maps_) return lf_set_error(LF_ERR_ARG, "set_vmap takes FP32 images only");
    const size_t n = (size_t)K_[0].width * K_[0].height;
    LF_HIP_CHECK(hipMemcpyAsync(vmap_[0].ptr<float>(), vmap, n * 16, hipMemcpyDeviceToDevice, stream_));
    return resize_device_map();
}

int DeviceImage::intensity_pyramid() {
    void *st = stream_;
    for (int l = 1; l < levels(); ++l)
        LF_RC(lf_pyr_down_image(intensity_[l - 1].ptr<float>(), intensity_[l].ptr<float>(), K_[l - 1].width,
                                K_[l - 1].height, st));
    for (int l = 0; l < levels(); ++l)
        LF_RC(lf_compute_derivative(intensity_[l].ptr<float>(), dx_[l].ptr<float>(), dy_[l].ptr<float>(),
                                    K_[l].width, K_[l].height, st));
    return LF_OK;
}

int DeviceImage::set_intensity(const float *intensity) {
    if (!intensity) return lf_set_error(LF_ERR_ARG, "null intensity");
    if (half_maps_) return lf_set_error(LF_ERR_ARG, "set_intensity takes FP32 images only");
    const size_t n = (size_t)K_[0].width * K_[0].height;
    LF_HIP_CHECK(hipMemcpyAsync(intensity_[0].ptr<float>(), intensity, n * 4, hipMemcpyDeviceToDevice, stream_));
    return intensity_pyramid();
}

int DeviceImage::set_image(const unsigned char *rgb) {
    if (!rgb) return lf_set_error(LF_ERR_ARG, "null image");
    if (half_maps_) return lf_set_error(LF_ERR_ARG, "set_image takes FP32 images only");
    LF_RC(lf_rgb_to_intensity(rgb, intensity_[0].ptr<float>(), K_[0].width, K_[0].height, stream_));
    return intensity_pyramid();
}

int DeviceImage::set_depth(const float *depth) {
    if (!depth) return lf_set_error(LF_ERR_ARG, "null depth");
    if (half_maps_) return lf_set_error(LF_ERR_ARG, "set_depth takes FP32 images only");
    void *st = stream_;
    LF_RC(lf_bilateral_filter(depth, depth_[0].ptr<float>(), K_[0].width, K_[0].height, st));
    for (int l = 1; l < levels(); ++l)
        LF_RC(lf_pyr_down_depth(depth_[l - 1].ptr<float>(), depth_[l].ptr<float>(), K_[l - 1].width,
                                K_[l - 1].height, st));
    return LF_OK;
}

// ------------------------------------------------------------ DenseMapping

DenseMapping::DenseMapping(const IntrinsicMatrix &K, const lf_map_params &params, hipStream_t stream)
    : K_(K), params_(params), stream_(stream) {
    memset(&storage_, 0, sizeof(storage_));
    if (lf_map_create(&params_, K.width, K.height, &storage_) != LF_OK) throw std::bad_alloc();
    reset_mapping();
}

DenseMapping::~DenseMapping() { lf_map_release(&storage_); }

int DenseMapping::update(DeviceImage &frame, const float *pose24, bool colour, hipEvent_t e0, hipEvent_t e1) {
    return lf_update_ev(&params_, &storage_, frame.get_raw_depth().ptr<float>(), frame.get_nmap(0).ptr<float>(),
                        colour ? frame.get_image().ptr<unsigned char>() : nullptr, K_.width, K_.height, pose24, &K_,
                        stream_, true, e0, e1, frame.get_integration_samples().ptr<float>());
}

int DenseMapping::update(const float *depth, const unsigned char *rgb, const float *pose24) {
    return lf_update_ev(&params_, &storage_, depth, nullptr, rgb, K_.width, K_.height, pose24, &K_, stream_, false,
                        nullptr, nullptr);
}

int DenseMapping::raycast(float *vmap, const float *pose24, hipEvent_t e0, hipEvent_t e1) {
    return lf_raycast_ev(&params_, &storage_, vmap, nullptr, K_.width, K_.height, pose24, &K_, 1, stream_, e0, e1);
}

int DenseMapping::reset_mapping() { return lf_map_reset(&params_, &storage_, stream_); }

int DenseMapping::raycast_check_visibility(float *vmap, const float *pose24) {
    return lf_raycast_check_visibility(&params_, &storage_, vmap, nullptr, K_.width, K_.height, pose24, &K_, 1,
                                       stream_);
}

int DenseMapping::deintegrate(DeviceImage &frame, const float *pose24, bool colour) {
    return lf_deintegrate_ev(&params_, &storage_, frame.get_raw_depth().ptr<float>(), frame.get_nmap(0).ptr<float>(),
                             colour ? frame.get_image().ptr<unsigned char>() : nullptr, K_.width, K_.height, pose24,
                             &K_, stream_, frame.get_integration_samples().ptr<float>());
}

int DenseMapping::writeMapToDisk(const std::string &file_name) {
    return lf_map_write(&params_, &storage_, file_name.c_str(), stream_);
}

int DenseMapping::readMapFromDisk(const std::string &file_name) {
    return lf_map_read(&params_, &storage_, file_name.c_str(), stream_);
}

size_t DenseMapping::fetch_mesh_vertex_only(float *vertex, int max_tri) {
    unsigned int n = 0;
    lf_create_mesh(&params_, &storage_, vertex, nullptr, nullptr, max_tri, &n, stream_);
    return n;
}
size_t DenseMapping::fetch_mesh_with_normal(float *vertex, float *normal, int max_tri) {
    unsigned int n = 0;
    lf_create_mesh(&params_, &storage_, vertex, normal, nullptr, max_tri, &n, stream_);
    return n;
}
size_t DenseMapping::fetch_mesh_with_colour(float *vertex, unsigned char *colour, int max_tri) {
    unsigned int n = 0;
    lf_create_mesh(&params_, &storage_, vertex, nullptr, colour, max_tri, &n, stream_);
    return n;
}

// ------------------------------------------------------------ DenseTracking

DenseTracking::DenseTracking(hipStream_t stream) : stream_(stream) {
    if (hipMalloc(&state_, sizeof(LfTrackState)) != hipSuccess) throw std::bad_alloc();
    if (hipMalloc(&scratch_, lf_reduce_partials() * sizeof(float)) != hipSuccess) throw std::bad_alloc();
    // every field defined before the first frame (errors 0, success 0, counts 0)
    if (hipMemsetAsync(state_, 0, sizeof(LfTrackState), stream_) != hipSuccess)
        throw std::runtime_error("hipMemsetAsync failed");
    const double I[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    lf_track_set_pose(state_, I, stream_);
}

DenseTracking::DenseTracking(const IntrinsicMatrix &K, int num_pyr, hipStream_t stream) : DenseTracking(stream) {
    src_.reset(new DeviceImage(K, num_pyr, stream));
    ref_.reset(new DeviceImage(K, num_pyr, stream));
}

void DenseTracking::swap_intensity_pyr() {
    if (src_ && ref_) src_->swap_intensity(*ref_);
}

int DenseTracking::compute_transform(const TrackingContext &c) {
    if (!src_ || !ref_) return lf_set_error(LF_ERR_ARG, "tracker made without pyramids (use DenseTracking(K, NUM_PYR))");
    TrackingContext k = c;
    if (k.intrinsics_pyr_.empty()) k.intrinsics_pyr_ = ref_->intrinsics();
    return enqueue(*ref_, *src_, k, true);
}

DenseTracking::~DenseTracking() {
    (void)hipFree(state_);
    (void)hipFree(scratch_);
}

// compute_transform(ref, cur, ctx) (icp_tracker.cpp:23-156): levels coarse to
// fine, max_iterations_[level] Gauss-Newton steps each, all on the device.  The
// first step starts the estimate and the step that ends the frame also runs
// the finish (success test and odometry chain), so a frame is two launches
// per step.
int DenseTracking::compute_transform(DeviceImage &ref, DeviceImage &cur, const TrackingContext &c) {
    return enqueue(ref, cur, c, false);
}

int DenseTracking::plan(DeviceImage &ref, DeviceImage &cur, const TrackingContext &c,
                        std::vector<unsigned char> &jobs, std::vector<int> &grids) {
    return enqueue(ref, cur, c, false, &jobs, &grids);
}

int DenseTracking::enqueue(DeviceImage &ref, DeviceImage &cur, const TrackingContext &c, bool context_mode,
                           std::vector<unsigned char> *jobs, std::vector<int> *grids) {
    context_mode_ = context_mode;
    LfStepFlags f = {};
    f.context_mode = context_mode ? 1 : 0;
    (c.use_initial_guess_ ? c.initial_estimate_ : SE3d()).matrix3x4(f.init);
    struct Step { int level, it; };
    std::vector<Step> plan;
    for (int level = (int)c.max_iterations_.size() - 1; level >= 0; --level) {
        if (level >= ref.levels()) continue;
        for (int it = 0; it < c.max_iterations_[level]; ++it) plan.push_back({level, it});
    }
    if (plan.empty()) {
        LF_RC(lf_track_begin(state_, f.init, stream_));
        return lf_track_finish(state_, stream_);
    }
    const int finest = plan.back().level;
    for (size_t s = 0; s < plan.size(); ++s) {
        const int level = plan[s].level;
        LfTrackLevel lv;
        lv.curr_vmap = cur.get_vmap(level).ptr<float>();
        lv.curr_nmap = cur.get_nmap(level).ptr<float>();
        lv.last_vmap = ref.get_vmap(level).ptr<float>();
        lv.last_nmap = ref.get_nmap(level).ptr<float>();
        lv.curr_I = cur.get_intensity(level).ptr<float>();
        lv.last_I = ref.get_intensity(level).ptr<float>();
        lv.dIdx = cur.get_intensity_dx(level).ptr<float>();
        lv.dIdy = cur.get_intensity_dy(level).ptr<float>();
        const IntrinsicMatrix &K = c.intrinsics_pyr_.empty() ? ref.intrinsics()[level] : c.intrinsics_pyr_[level];
        lv.cols = K.width;
        lv.rows = K.height;
        lv.K = K;
        lv.half_maps = cur.half_maps() ? 1 : 0;
        f.first_of_level = plan[s].it == 0;
        f.first_of_frame = s == 0;
        f.last_level = level == finest;
        f.last_of_frame = s + 1 == plan.size();
        if (jobs) {
            const size_t J = lf_track_job_bytes();
            jobs->resize(jobs->size() + J);
            int G = 0;
            LF_RC(lf_track_job_make(lv, scratch_, state_, f, icp_weight, batch_share_, batch_wgs_, jobs->data() + jobs->size() - J,
                                    &G));
            grids->push_back(G);
        } else {
            LF_RC(lf_track_iteration(lv, scratch_, state_, f, icp_weight, stream_));
        }
    }
    return LF_OK;
}

TrackingResult DenseTracking::result() {
    LfTrackState h;
    TrackingResult r;
    if (hipMemcpyAsync(&h, state_, sizeof(h), hipMemcpyDeviceToHost, stream_) != hipSuccess) return r;
    if (hipStreamSynchronize(stream_) != hipSuccess) return r;
    r.sucess = h.success != 0;
    // compute_transform(ctx) reports last_icp_error (icp_tracker.cpp:310)
    r.icp_error = context_mode_ ? h.last_icp_error : h.icp_error;
    r.update = r.sucess ? SE3d::from3x4(h.est).inverse() : SE3d();
    return r;
}

int DenseTracking::read_sums(double *sums31) {
    double h[32];
    LF_HIP_CHECK(hipMemcpyAsync(h, state_->sums, sizeof(h), hipMemcpyDeviceToHost, stream_));
    LF_HIP_CHECK(hipStreamSynchronize(stream_));
    for (int k = 0; k < 31; ++k) sums31[k] = h[k];
    return LF_OK;
}

// ------------------------------------------------------------ DenseOdometry

DenseOdometry::DenseOdometry(const IntrinsicMatrix &base, int num_pyr, std::vector<int> iterations,
                             hipStream_t stream, bool half_maps)
    : iterations_(std::move(iterations)), stream_(stream), tracker_(stream) {
    BuildIntrinsicPyramid(base, K_, num_pyr);
    curr_.reset(new DeviceImage(base, num_pyr, stream, half_maps));
    ref_.reset(new DeviceImage(base, num_pyr, stream, half_maps));
    spare_.reset(new DeviceImage(base, num_pyr, stream, half_maps));
}

const float *DenseOdometry::pose24() const { return tracker_.state()->pose24; }

int DenseOdometry::trackFrame(const uint16_t *depth, const unsigned char *rgb, float depth_scale,
                              const SE3d &initial_pose) {
    LF_RC(curr_->upload(depth, rgb, depth_scale));
    if (!initialized_) {
        double T[12];
        initial_pose.matrix3x4(T);
        LF_RC(lf_track_set_pose(tracker_.state(), T, stream_));
        advance();
        initialized_ = true;
        return LF_OK;
    }
    TrackingContext c;
    c.use_initial_guess_ = true;
    c.initial_estimate_ = SE3d();
    c.intrinsics_pyr_ = K_;
    c.max_iterations_ = iterations_;
    // compute_transform ends with the finish, which also chains the odometry pose
    LF_RC(tracker_.compute_transform(*ref_, *curr_, c));
    advance();
    trackingLost = false;
    return LF_OK;
}

int DenseOdometry::planFrame(const uint16_t *depth, const unsigned char *rgb, float depth_scale,
                             const SE3d &initial_pose, std::vector<unsigned char> &jobs, std::vector<int> &grids,
                             std::vector<unsigned char> *upload) {
    if (upload)
        LF_RC(curr_->plan_upload(depth, rgb, depth_scale, *upload));
    else
        LF_RC(curr_->upload(depth, rgb, depth_scale));
    if (!initialized_) {
        double T[12];
        initial_pose.matrix3x4(T);
        LF_RC(lf_track_set_pose(tracker_.state(), T, stream_));
        advance();
        initialized_ = true;
        return LF_OK;
    }
    TrackingContext c;
    c.use_initial_guess_ = true;
    c.initial_estimate_ = SE3d();
    c.intrinsics_pyr_ = K_;
    c.max_iterations_ = iterations_;
    LF_RC(tracker_.plan(*ref_, *curr_, c, jobs, grids));
    advance();
    trackingLost = false;
    return LF_OK;
}

void DenseOdometry::reset() {
    initialized_ = false;
    trackingLost = false;
}

// ------------------------------------------------------------ System

System::System(const lf_system_config &cfg) : cfg_(cfg) {
    if (hipStreamCreate(&stream_) != hipSuccess) throw std::bad_alloc();
    IntrinsicMatrix K(cfg.K.width, cfg.K.height, cfg.K.fx, cfg.K.fy, cfg.K.cx, cfg.K.cy);
    mapping_.reset(new DenseMapping(K, cfg.map, stream_));
    std::vector<int> its(cfg.iterations, cfg.iterations + std::max(1, std::min(8, cfg.num_pyr)));
    odometry_.reset(new DenseOdometry(K, cfg.num_pyr, its, stream_, cfg.map.precision == LF_PRECISION_MIXED));
}

System::~System() {
    drop_graphs();
    mapping_.reset();
    odometry_.reset();
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
    for (auto &e : events_) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    if (stream_) (void)hipStreamDestroy(stream_);
}

hipEvent_t System::take_event() {
    hipEvent_t e = nullptr;
    if (!pool_.empty()) {
        e = pool_.back();
        pool_.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        e = nullptr;
    }
    return e;
}

void System::set_timing(int stages) {
    timing_ = stages;
    for (auto &e : events_) { pool_.push_back(e.a); pool_.push_back(e.b); }
    events_.clear();
}

int System::read_timing(int stage, int *launches, double *total_ms) {
    LF_HIP_CHECK(hipStreamSynchronize(stream_));
    int n = 0;
    double t = 0;
    for (auto &e : events_) {
        if (e.stage != stage) continue;
        float ms = 0;
        LF_HIP_CHECK(hipEventElapsedTime(&ms, e.a, e.b));
        t += ms;
        ++n;
    }
    *launches = n;
    *total_ms = t;
    return LF_OK;
}

// process_images (system.cpp:40-114): track, fuse, raycast, re-pyramid.
void System::drop_graphs() {
    for (hipGraphExec_t &e : exec_) {
        if (e) (void)hipGraphExecDestroy(e);
        e = nullptr;
    }
}

void System::set_graphs(bool on) {
    graphs_ = on;
    if (!on) drop_graphs();
}

int System::enqueue_before_raycast(const uint16_t *depth, const unsigned char *rgb) {
    DenseOdometry &od = *odometry_;
    if (!od.trackingLost) LF_RC(od.trackFrame(depth, rgb, cfg_.depth_scale, initial_pose_));
    DeviceImage &ref = od.get_reference_image();
    if (mapping_enabled) LF_RC(mapping_->update(ref, od.pose24(), cfg_.integrate_colour != 0));
    return LF_OK;
}

int System::plan_tracking(const uint16_t *depth, const unsigned char *rgb, std::vector<unsigned char> &jobs,
                          std::vector<int> &grids, std::vector<unsigned char> *upload) {
    DenseOdometry &od = *odometry_;
    if (!od.trackingLost) LF_RC(od.planFrame(depth, rgb, cfg_.depth_scale, initial_pose_, jobs, grids, upload));
    return LF_OK;
}

int System::enqueue_fusion() {
    DenseOdometry &od = *odometry_;
    if (mapping_enabled) LF_RC(mapping_->update(od.get_reference_image(), od.pose24(), cfg_.integrate_colour != 0));
    return LF_OK;
}

float *System::raycast_vmap() { return odometry_->get_reference_image().get_vmap(0).ptr<float>(); }

const float *System::raycast_pose24() { return odometry_->pose24(); }

int System::enqueue_after_raycast() { return odometry_->get_reference_image().resize_device_map(); }

int System::plan_fusion(LfFuseItem &it) {
    DeviceImage &ref = odometry_->get_reference_image();
    DenseMapping &m = *mapping_;
    it = LfFuseItem{&m.params(), &m.storage(), ref.get_raw_depth().ptr<float>(),
                    ref.get_integration_samples().ptr<float>(),
                    cfg_.integrate_colour ? ref.get_image().ptr<unsigned char>() : nullptr, odometry_->pose24(),
                    &m.intrinsics(), m.intrinsics().width, m.intrinsics().height, mapping_enabled ? 1 : 0};
    return LF_OK;
}

int System::plan_after_raycast(std::vector<unsigned char> &passes) {
    return odometry_->get_reference_image().plan_refresh(passes);
}

// The frame's device work (trackFrame, update, raycast, re-pyramid) without events.
int System::enqueue_frame(const uint16_t *depth, const unsigned char *rgb) {
    LF_RC(enqueue_before_raycast(depth, rgb));
    LF_RC(mapping_->raycast(raycast_vmap(), raycast_pose24()));
    return enqueue_after_raycast();
}

// ------------------------------------------------------------------- Batch

namespace {
struct TunedOption {
    int option;
    std::vector<int> values;  // candidates; the first is the default
};
// The options tuned, in order, with their candidate values (0: the built-in
// rule, which the sweeps behind it chose for 16 maps of 640x480).
const std::vector<TunedOption> &tuned_options() {
    static const std::vector<TunedOption> o = {
        {LF_BATCH_OPT_INTEGRATE_WORKGROUPS, {0, 512, 1024, 4096}},
        {LF_BATCH_OPT_TRACK_WORKGROUPS, {0, 16, 48, 64}},
        {LF_BATCH_OPT_ZRANGE_WORKGROUPS, {0, 32, 128, 256}},
        {LF_BATCH_OPT_UPLOAD_OVERLAP, {1, 0}},
    };
    return o;
}
std::mutex g_tune_mutex;
std::map<std::string, std::vector<int>> g_tune_cache;  // batch shape -> tuned option values
std::string g_tune_file;
}  // namespace

Batch::Batch(const lf_system_config *cfgs, int n) {
    auto fail = [&]() {
        release();
        throw std::bad_alloc();
    };
    for (int i = 0; i < n; ++i) {
        // batched launches run one map format
        if (cfgs[i].map.precision != cfgs[0].map.precision) throw std::invalid_argument("Batch: mixed precisions");
        sys_.emplace_back(new System(cfgs[i]));
        sys_.back()->odometry().tracker().set_batch_share(n);
    }
    half_maps_ = n > 0 && cfgs[0].map.precision == LF_PRECISION_MIXED;
    if (hipStreamCreate(&stream_) != hipSuccess || hipStreamCreate(&aux_) != hipSuccess) fail();
    ready_.assign(n, nullptr);
    for (hipEvent_t &e : ready_)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) fail();
    if (hipEventCreateWithFlags(&uploaded_, hipEventDisableTiming) != hipSuccess) fail();
    for (int k = 0; k < 2; ++k)
        if (hipEventCreateWithFlags(&table_read_[k], hipEventDisableTiming) != hipSuccess) fail();
    // a shape tuned before in this process (or read from the tuning file) starts with its choice
    std::vector<int> cached;
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tune_cache.find(tune_key());
        if (it != g_tune_cache.end()) cached = it->second;
    }
    for (size_t t = 0; t < cached.size(); ++t) (void)set_option(tuned_options()[t].option, cached[t]);
}

Batch::~Batch() { release(); }

// ------------------------------------------------------------ autotuning


int Batch::set_tune_file(const std::string &path) {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tune_file = path;
    if (path.empty()) return LF_OK;
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {  // "<shape key> v0 v1 ..." per line; later lines win
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key)) continue;
        std::vector<int> v;
        int x;
        while (ls >> x) v.push_back(x);
        if (v.size() == tuned_options().size()) g_tune_cache[key] = v;
    }
    return LF_OK;
}

// The batch's shape: every system's image size, pyramid levels, voxel size
// and precision, in order.
std::string Batch::tune_key() const {
    std::ostringstream k;
    k << "n" << sys_.size();
    for (const auto &s : sys_) {
        const lf_intrinsics &K = s->mapping().intrinsics();
        const lf_map_params &p = s->mapping().params();
        k << "_" << K.width << "x" << K.height << "l" << s->odometry().get_reference_image().levels() << "v"
          << (int)std::lround(p.voxel_size * 1e5f) << "p" << p.precision;
    }
    return k.str();
}

int Batch::autotune(int steps_per_trial, int rounds) {
    for (auto &e : tune_.pending) {
        pool_.push_back(e.second.first);
        pool_.push_back(e.second.second);
    }
    if (tune_.a) pool_.push_back(tune_.a);
    tune_ = Tuner{};
    if (steps_per_trial <= 0) return LF_OK;
    if (steps_per_trial > 64 || rounds < 1 || rounds > 16)
        return lf_set_error(LF_ERR_ARG, "autotune: 1..64 steps per trial and 1..16 rounds");
    tune_.per_trial = steps_per_trial;
    tune_.rounds = rounds;
    tune_.ms.assign(tuned_options()[0].values.size(), {});
    return LF_OK;
}

// The candidate of the current trial: odd rounds visit the candidates in
// reverse order, so a steady drift of the step time (maps grow) cancels in
// the medians instead of favouring the candidates tried first.
static int trial_candidate(int round, int cand, int count) { return round % 2 ? count - 1 - cand : cand; }

// Before a step: the current trial's candidate value (the first candidate of
// each option is its built-in default).
void Batch::tune_begin_step() {
    if (!tuning()) return;
    const TunedOption &o = tuned_options()[tune_.opt];
    (void)set_option(o.option, o.values[trial_candidate(tune_.round, tune_.cand, (int)o.values.size())]);
}

// After a step: mark the trial's timing events, advance, and when an option's
// trials are done, read them, keep the fastest candidate and go on.
int Batch::tune_end_step() {
    if (!tuning()) return LF_OK;
    if (tune_.step == 0) {  // the settling step ends: the timed steps start here
        tune_.a = take_event();
        if (!tune_.a) return lf_set_error(LF_ERR_HIP, "hipEventCreate failed");
        LF_HIP_CHECK(hipEventRecord(tune_.a, stream_));
    }
    if (++tune_.step <= tune_.per_trial) return LF_OK;
    hipEvent_t b = take_event();
    if (!b) return lf_set_error(LF_ERR_HIP, "hipEventCreate failed");
    LF_HIP_CHECK(hipEventRecord(b, stream_));
    const TunedOption &o = tuned_options()[tune_.opt];
    tune_.pending.push_back({trial_candidate(tune_.round, tune_.cand, (int)o.values.size()), {tune_.a, b}});
    tune_.a = nullptr;
    tune_.step = 0;
    if (++tune_.cand < (int)o.values.size()) return LF_OK;
    tune_.cand = 0;
    if (++tune_.round < tune_.rounds) return LF_OK;
    // every trial of this option is enqueued: read them
    LF_HIP_CHECK(hipEventSynchronize(tune_.pending.back().second.second));
    for (auto &e : tune_.pending) {
        float ms = 0.0f;
        LF_HIP_CHECK(hipEventElapsedTime(&ms, e.second.first, e.second.second));
        tune_.ms[e.first].push_back(ms / tune_.per_trial);
        pool_.push_back(e.second.first);
        pool_.push_back(e.second.second);
    }
    tune_.pending.clear();
    // the fastest median, kept only when it beats the built-in default (the
    // first candidate) by more than the trials' noise (1%)
    std::vector<double> med(tune_.ms.size());
    for (size_t c = 0; c < tune_.ms.size(); ++c) {
        std::vector<double> v = tune_.ms[c];
        std::sort(v.begin(), v.end());
        med[c] = v[v.size() / 2];
    }
    int best = 0;
    for (size_t c = 1; c < med.size(); ++c)
        if (med[c] < med[best]) best = (int)c;
    if (!(med[best] < 0.99 * med[0])) best = 0;
    LF_RC(set_option(o.option, o.values[best]));
    tune_.round = 0;
    if (++tune_.opt < (int)tuned_options().size()) {
        tune_.ms.assign(tuned_options()[tune_.opt].values.size(), {});
        return LF_OK;
    }
    // done: cache the choice for this shape
    std::vector<int> chosen;
    for (const TunedOption &t : tuned_options()) {
        int v = 0;
        LF_RC(get_option(t.option, &v));
        chosen.push_back(v);
    }
    tune_ = Tuner{};
    const std::string key = tune_key();
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tune_cache[key] = chosen;
    if (!g_tune_file.empty()) {
        std::ofstream out(g_tune_file, std::ios::app);
        out << key;
        for (int v : chosen) out << " " << v;
        out << "\n";
    }
    return LF_OK;
}

void Batch::StepGraphs::drop() {
    for (hipGraphExec_t *e : {&up, &track, &fuse}) {
        if (*e) (void)hipGraphExecDestroy(*e);
        *e = nullptr;
    }
    sig.clear();
}

void Batch::set_graphs(bool on) {
    graphs_ = on;
    if (!on) {
        if (stream_) (void)hipStreamSynchronize(stream_);
        if (aux_) (void)hipStreamSynchronize(aux_);
        for (StepGraphs &g : graphs_k_) g.drop();
    }
}

void Batch::release() {
    if (aux_) (void)hipStreamSynchronize(aux_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (StepGraphs &g : graphs_k_) g.drop();
    for (auto &s : sys_)
        if (s) (void)hipStreamSynchronize(s->stream());
    for (hipEvent_t e : ready_)
        if (e) (void)hipEventDestroy(e);
    ready_.clear();
    if (uploaded_) (void)hipEventDestroy(uploaded_);
    uploaded_ = nullptr;
    for (int k = 0; k < 2; ++k) {
        if (table_host_[k]) (void)hipHostFree(table_host_[k]);
        if (table_dev_[k]) (void)hipFree(table_dev_[k]);
        if (table_read_[k]) (void)hipEventDestroy(table_read_[k]);
        table_host_[k] = table_dev_[k] = nullptr;
        table_read_[k] = nullptr;
    }
    if (stream_) (void)hipStreamDestroy(stream_);
    stream_ = nullptr;
    if (aux_) (void)hipStreamDestroy(aux_);
    aux_ = nullptr;
    for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
    pool_.clear();
    for (auto &e : events_) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    events_.clear();
    depth_in_.clear();
    rgb_in_.clear();
    sys_.clear();
}

// Room for `bytes` of step descriptors per buffer.  The tables are read on
// the upload stream and on the batch stream: growing waits for both.
int Batch::reserve_table(size_t bytes) {
    if (bytes <= table_cap_) return LF_OK;
    LF_HIP_CHECK(hipStreamSynchronize(aux_));
    LF_HIP_CHECK(hipStreamSynchronize(stream_));
    const size_t cap = std::max(bytes, 2 * table_cap_);
    for (int k = 0; k < 2; ++k) {
        if (table_host_[k]) LF_HIP_CHECK(hipHostFree(table_host_[k]));
        if (table_dev_[k]) LF_HIP_CHECK(hipFree(table_dev_[k]));
        table_host_[k] = table_dev_[k] = nullptr;
        LF_HIP_CHECK(hipHostMalloc(&table_host_[k], cap, hipHostMallocDefault));
        LF_HIP_CHECK(hipMalloc(&table_dev_[k], cap));
    }
    table_cap_ = cap;
    return LF_OK;
}

hipEvent_t Batch::take_event() {
    hipEvent_t e = nullptr;
    if (!pool_.empty()) {
        e = pool_.back();
        pool_.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        e = nullptr;
    }
    return e;
}

int Batch::set_option(int option, int value) {
    switch (option) {
    case LF_BATCH_OPT_UPLOAD_OVERLAP:
        if (value != 0 && value != 1) return lf_set_error(LF_ERR_ARG, "upload overlap is 0 or 1, not %d", value);
        upload_overlap_ = value;
        return LF_OK;
    case LF_BATCH_OPT_TRACK_WORKGROUPS:
        if (value < 0 || value > 512) return lf_set_error(LF_ERR_ARG, "tracking workgroups %d not in 0..512", value);
        track_wgs_ = value;
        for (auto &s : sys_) s->odometry().tracker().set_batch_workgroups(value);
        return LF_OK;
    case LF_BATCH_OPT_INTEGRATE_WORKGROUPS:
        if (value < 0 || value > 65535)
            return lf_set_error(LF_ERR_ARG, "integration workgroups %d not in 0..65535", value);
        integrate_wgs_ = value;
        return LF_OK;
    case LF_BATCH_OPT_ZRANGE_WORKGROUPS:
        if (value < 0 || value > 65535)
            return lf_set_error(LF_ERR_ARG, "z-range workgroups %d not in 0..65535", value);
        zrange_wgs_ = value;
        return LF_OK;
    default:
        return lf_set_error(LF_ERR_ARG, "unknown batch option %d", option);
    }
}

int Batch::get_option(int option, int *value) const {
    switch (option) {
    case LF_BATCH_OPT_UPLOAD_OVERLAP: *value = upload_overlap_; return LF_OK;
    case LF_BATCH_OPT_TRACK_WORKGROUPS: *value = track_wgs_; return LF_OK;
    case LF_BATCH_OPT_INTEGRATE_WORKGROUPS: *value = integrate_wgs_; return LF_OK;
    case LF_BATCH_OPT_ZRANGE_WORKGROUPS: *value = zrange_wgs_; return LF_OK;
    default: return lf_set_error(LF_ERR_ARG, "unknown batch option %d", option);
    }
}

void Batch::set_timing(int stages) {
    timing_ = stages;
    for (auto &e : events_) { pool_.push_back(e.a); pool_.push_back(e.b); }
    events_.clear();
}

int Batch::read_timing(int stage, int *launches, double *total_ms) {
    LF_HIP_CHECK(hipStreamSynchronize(stream_));
    int n = 0;
    double t = 0;
    for (auto &e : events_) {
        if (e.stage != stage) continue;
        float ms = 0;
        LF_HIP_CHECK(hipEventElapsedTime(&ms, e.a, e.b));
        t += ms;
        ++n;
    }
    *launches = n;
    *total_ms = t;
    return LF_OK;
}

int Batch::process(const uint16_t *const *depth, const unsigned char *const *rgb) {
    tune_begin_step();
    LF_RC(step(depth, rgb, false));
    return tune_end_step();
}

int Batch::process_host(const uint16_t *const *depth, const unsigned char *const *rgb) {
    tune_begin_step();
    LF_RC(step(depth, rgb, true));
    return tune_end_step();
}

int Batch::step(const uint16_t *const *depth_in, const unsigned char *const *rgb_in, bool from_host) {
    const int n = size();
    const size_t PB = lf_pass_bytes(), J = lf_track_job_bytes(), RJ = lf_ray_job_bytes(), FJ = lf_fuse_job_bytes();
    // host frames are read by the upload passes from the batch's staging
    // buffers, written by copies on the upload stream ahead of those passes
    std::vector<const uint16_t *> dstage(n);
    std::vector<const unsigned char *> cstage(n);
    if (from_host) {
        if (depth_in_.size() != (size_t)n) {
            depth_in_.resize(n);
            rgb_in_.resize(n);
        }
        for (int i = 0; i < n; ++i) {
            const lf_intrinsics &K = sys_[i]->mapping().intrinsics();
            try {
                depth_in_[i].create(K.width, K.height, 2);
                rgb_in_[i].create(K.width, K.height, 3);
            } catch (const std::bad_alloc &) {
                return lf_set_error(LF_ERR_HIP, "device allocation of the batch's input staging failed");
            }
            dstage[i] = depth_in_[i].ptr<uint16_t>();
            cstage[i] = rgb_in_[i].ptr<unsigned char>();
        }
    }
    const uint16_t *const *depth = from_host ? dstage.data() : depth_in;
    const unsigned char *const *rgb = from_host ? cstage.data() : rgb_in;
    // every system's upload and re-pyramid planned as frame passes, its
    // Gauss-Newton steps as jobs
    std::vector<std::vector<unsigned char>> jobs(n), up(n), after(n);
    std::vector<std::vector<int>> grids(n);
    size_t steps = 0, n_up = 0, n_after = 0;
    bool any_start = false;  // a system starts its frame on its own stream (no steps)
    for (int i = 0; i < n; ++i) {
        DenseOdometry &od = sys_[i]->odometry();
        const bool tracks = od.initialized() && !od.trackingLost;
        LF_RC(sys_[i]->plan_tracking(depth[i], rgb[i], jobs[i], grids[i], &up[i]));
        LF_RC(sys_[i]->plan_after_raycast(after[i]));
        steps = std::max(steps, grids[i].size());
        n_up = std::max(n_up, up[i].size() / PB);
        n_after = std::max(n_after, after[i].size() / PB);
        // planning launches on the system's stream only to start a frame
        // without steps (first frame: the initial pose); the batch's tracking
        // waits for that, and needs no event otherwise
        if (!tracks || grids[i].empty()) {
            any_start = true;
            LF_HIP_CHECK(hipEventRecord(ready_[i], sys_[i]->stream()));
            LF_HIP_CHECK(hipStreamWaitEvent(stream_, ready_[i], 0));
        }
    }
    // One table per step, copied in one transfer: [pass p][system] (uploads,
    // then re-pyramids), [step s][system] tracking jobs, [system] updates,
    // [system] raycasts.
    const size_t n_pass = n_up + n_after;
    const size_t off_track = n_pass * n * PB, off_fuse = off_track + steps * n * J, off_ray = off_fuse + n * FJ,
                 bytes = off_ray + n * RJ;
    // the buffers of two steps ago must have been copied before they are rewritten
    const int k = parity_;
    if (table_used_[k]) LF_HIP_CHECK(hipEventSynchronize(table_read_[k]));
    LF_RC(reserve_table(bytes));
    unsigned char *host = static_cast<unsigned char *>(table_host_[k]);
    unsigned char *dev = static_cast<unsigned char *>(table_dev_[k]);
    std::vector<int> wmax(n_pass, 0), gmax(steps, 0);
    for (size_t p = 0; p < n_pass; ++p)
        for (int i = 0; i < n; ++i) {
            const std::vector<unsigned char> &src = p < n_up ? up[i] : after[i];
            const size_t q = p < n_up ? p : p - n_up;
            unsigned char *dst = host + (p * n + i) * PB;
            if (q < src.size() / PB) {
                std::memcpy(dst, src.data() + q * PB, PB);
                wmax[p] = std::max(wmax[p], lf_pass_workgroups(dst));
            } else {
                lf_pass_skip(dst);
            }
        }
    for (size_t s = 0; s < steps; ++s)
        for (int i = 0; i < n; ++i) {
            unsigned char *dst = host + off_track + (s * n + i) * J;
            if (s < grids[i].size()) {
                std::memcpy(dst, jobs[i].data() + s * J, J);
                gmax[s] = std::max(gmax[s], grids[i][s]);
            } else {
                lf_track_job_skip(dst);
            }
        }
    std::vector<LfFuseItem> fuse(n);
    std::vector<LfRayItem> items(n);
    for (int i = 0; i < n; ++i) {
        LF_RC(sys_[i]->plan_fusion(fuse[i]));
        DenseMapping &m = sys_[i]->mapping();
        items[i] = LfRayItem{&m.params(), &m.storage(), sys_[i]->raycast_vmap(), nullptr,
                             sys_[i]->raycast_pose24(), &m.intrinsics(), m.intrinsics().width,
                             m.intrinsics().height};
    }
    LfFuseBatchDims fd;
    LF_RC(lf_update_batch_make(fuse.data(), n, host + off_fuse, &fd));
    if (integrate_wgs_ > 0) fd.integrate_grid = integrate_wgs_;
    LfRayBatchDims ray;
    LF_RC(lf_raycast_batch_make(items.data(), n, host + off_ray, &ray));
    if (zrange_wgs_ > 0) ray.zgrid_max = zrange_wgs_;
    // Upload stream: the table, then the uploads.  The uploads rewrite each
    // system's frame buffer that the step before the previous one tracked
    // against (three buffers per system), and the table buffer was last read
    // by that same step, whose end the host waited for above: nothing on the
    // device is left to wait for, so the uploads run as soon as they are
    // enqueued, beside the previous step's tracking (latency-bound launches)
    // when the host keeps ahead.  Without the overlap (upload_overlap_ 0) they
    // run on the batch stream ahead of the step's tracking.
    hipStream_t up_st = upload_overlap_ ? aux_ : stream_;
    // The step's device work in three parts: the table copy and uploads (on
    // up_st), the Gauss-Newton launches, then fusion, raycast and re-pyramids
    // (on the batch stream).  Run eagerly, or captured once per table buffer
    // and shape as three graphs and replayed (set_graphs); the cross-stream
    // events stay outside the graphs, so the overlap is the same.
    auto enqueue_up = [&](hipStream_t st) -> int {
        LF_HIP_CHECK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, st));
        for (size_t p = 0; p < n_up; ++p) LF_RC(lf_pass_batch(dev + p * n * PB, n, wmax[p], st));
        return LF_OK;
    };
    auto enqueue_track = [&]() -> int {
        for (size_t s = 0; s < steps; ++s)
            LF_RC(lf_track_batch_step(dev + off_track + s * n * J, n, gmax[s], half_maps_, stream_));
        return LF_OK;
    };
    auto enqueue_fuse = [&](hipEvent_t i0, hipEvent_t i1, hipEvent_t f0, hipEvent_t f1, hipEvent_t r0,
                            hipEvent_t r1) -> int {
        if (f0) LF_HIP_CHECK(hipEventRecord(f0, stream_));
        LF_RC(lf_update_batch_launch(dev + off_fuse, fd, stream_, i0, i1));
        if (f1) LF_HIP_CHECK(hipEventRecord(f1, stream_));
        LF_RC(lf_raycast_batch_launch(dev + off_ray, ray, 1, stream_, r0, r1));
        for (size_t p = n_up; p < n_pass; ++p) LF_RC(lf_pass_batch(dev + p * n * PB, n, wmax[p], stream_));
        return LF_OK;
    };
    if (graphs_ && timing_ == 0 && !any_start) {
        // the launch shapes and the table's address decide the graphs
        std::vector<long long> sig = {n, (long long)n_up, (long long)n_after, (long long)steps, (long long)bytes,
                                      (long long)(uintptr_t)dev, half_maps_ ? 1 : 0, upload_overlap_,
                                      fd.n, fd.any, fd.mixed, fd.tiles_max, fd.vis_grid_max, fd.integrate_grid,
                                      ray.n, ray.mixed, ray.tiles_max, ray.zgrid_max, ray.grid_max,
                                      (long long)ray.zrange_lds};
        sig.insert(sig.end(), gmax.begin(), gmax.end());
        sig.insert(sig.end(), wmax.begin(), wmax.end());
        StepGraphs &g = graphs_k_[k];
        if (g.sig != sig) {
            g.drop();
            auto capture = [&](hipStream_t st, hipGraphExec_t &exec, auto &&enqueue) -> int {
                hipGraph_t graph = nullptr;
                LF_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                const int rc = enqueue();
                const hipError_t ec = hipStreamEndCapture(st, &graph);
                if (rc != LF_OK) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return rc;
                }
                if (ec != hipSuccess)
                    return lf_set_error(LF_ERR_HIP, "batch step capture failed: %s", hipGetErrorString(ec));
                const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (ei != hipSuccess) {
                    exec = nullptr;
                    return lf_set_error(LF_ERR_HIP, "batch step graph instantiation failed: %s",
                                        hipGetErrorString(ei));
                }
                return LF_OK;
            };
            LF_RC(capture(up_st, g.up, [&]() { return enqueue_up(up_st); }));
            LF_RC(capture(stream_, g.track, [&]() { return enqueue_track(); }));
            LF_RC(capture(stream_, g.fuse,
                          [&]() { return enqueue_fuse(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr); }));
            g.sig = sig;
        }
        if (from_host)
            for (int i = 0; i < n; ++i) {
                LF_HIP_CHECK(hipMemcpyAsync(depth_in_[i].ptr<void>(), depth_in[i], depth_in_[i].bytes(),
                                            hipMemcpyHostToDevice, up_st));
                LF_HIP_CHECK(hipMemcpyAsync(rgb_in_[i].ptr<void>(), rgb_in[i], rgb_in_[i].bytes(),
                                            hipMemcpyHostToDevice, up_st));
            }
        LF_HIP_CHECK(hipGraphLaunch(g.up, up_st));
        if (upload_overlap_) {
            LF_HIP_CHECK(hipEventRecord(uploaded_, aux_));
            LF_HIP_CHECK(hipStreamWaitEvent(stream_, uploaded_, 0));
        }
        LF_HIP_CHECK(hipGraphLaunch(g.track, stream_));
        LF_HIP_CHECK(hipGraphLaunch(g.fuse, stream_));
        LF_HIP_CHECK(hipEventRecord(table_read_[k], stream_));
        table_used_[k] = true;
        parity_ ^= 1;
        return LF_OK;
    }
    LF_HIP_CHECK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, up_st));
    if (from_host) {
        // the staging buffers were last read by the previous step's upload
        // passes, earlier on this same stream
        for (int i = 0; i < n; ++i) {
            LF_HIP_CHECK(hipMemcpyAsync(depth_in_[i].ptr<void>(), depth_in[i], depth_in_[i].bytes(),
                                        hipMemcpyHostToDevice, up_st));
            LF_HIP_CHECK(hipMemcpyAsync(rgb_in_[i].ptr<void>(), rgb_in[i], rgb_in_[i].bytes(), hipMemcpyHostToDevice,
                                        up_st));
        }
    }
    for (size_t p = 0; p < n_up; ++p) LF_RC(lf_pass_batch(dev + p * n * PB, n, wmax[p], up_st));
    if (upload_overlap_) {
        LF_HIP_CHECK(hipEventRecord(uploaded_, aux_));
        LF_HIP_CHECK(hipStreamWaitEvent(stream_, uploaded_, 0));
    }
    // stage events (set_timing), recorded on the batch stream
    Ev ev[5];
    for (int sid = 0; sid < 5; ++sid) {
        ev[sid] = Ev{sid, nullptr, nullptr};
        if (timing_ & (1 << sid)) {
            ev[sid].a = take_event();
            ev[sid].b = take_event();
            if (!ev[sid].a || !ev[sid].b) return lf_set_error(LF_ERR_HIP, "hipEventCreate failed");
        }
    }
    if (ev[4].a) LF_HIP_CHECK(hipEventRecord(ev[4].a, stream_));
    if (ev[2].a) LF_HIP_CHECK(hipEventRecord(ev[2].a, stream_));
    LF_RC(enqueue_track());
    if (ev[2].b) LF_HIP_CHECK(hipEventRecord(ev[2].b, stream_));
    // block allocation, visible lists and integration of every map, the
    // z-range and raycast, the re-pyramids of the raycast maps
    LF_RC(enqueue_fuse(ev[0].a, ev[0].b, ev[3].a, ev[3].b, ev[1].a, ev[1].b));
    if (ev[4].b) LF_HIP_CHECK(hipEventRecord(ev[4].b, stream_));
    for (const Ev &e : ev) {
        // an event pair whose launch did not happen (nothing to fuse) is not a timing
        if (!e.a) continue;
        if (e.stage == 0 && !fd.any) { pool_.push_back(e.a); pool_.push_back(e.b); continue; }
        events_.push_back(e);
    }
    LF_HIP_CHECK(hipEventRecord(table_read_[k], stream_));
    table_used_[k] = true;
    parity_ ^= 1;
    return LF_OK;
}

int Batch::synchronize() {
    for (auto &s : sys_) LF_HIP_CHECK(hipStreamSynchronize(s->stream()));
    LF_HIP_CHECK(hipStreamSynchronize(aux_));
    LF_HIP_CHECK(hipStreamSynchronize(stream_));
    return LF_OK;
}

int System::process_images(const uint16_t *depth, const unsigned char *rgb) {
    DenseOdometry &od = *odometry_;
    if (graphs_ && timing_ == 0 && mapping_enabled && od.initialized() && !od.trackingLost) {
        // steady state: the inputs go to fixed buffers, the frame is a graph of
        // this buffer parity (captured on its first use, replayed after)
        const int W = cfg_.K.width, H = cfg_.K.height;
        if (depth_in_.empty()) {
            depth_in_.create(W, H, 2);
            rgb_in_.create(W, H, 3);
        }
        LF_HIP_CHECK(hipMemcpyAsync(depth_in_.ptr<void>(), depth, (size_t)W * H * 2, hipMemcpyDeviceToDevice, stream_));
        LF_HIP_CHECK(hipMemcpyAsync(rgb_in_.ptr<void>(), rgb, (size_t)W * H * 3, hipMemcpyDeviceToDevice, stream_));
        const int k = od.buffer_parity();
        if (exec_[k]) {
            LF_HIP_CHECK(hipGraphLaunch(exec_[k], stream_));
            od.replayed_frame();
            return LF_OK;
        }
        hipGraph_t g = nullptr;
        LF_HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_frame(depth_in_.ptr<uint16_t>(), rgb_in_.ptr<unsigned char>());
        const hipError_t ec = hipStreamEndCapture(stream_, &g);
        if (rc != LF_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return lf_set_error(LF_ERR_HIP, "frame capture failed: %s", hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&exec_[k], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            exec_[k] = nullptr;
            return lf_set_error(LF_ERR_HIP, "frame graph instantiation failed: %s", hipGetErrorString(ei));
        }
        // the capture swapped the frame buffers on the host; the graph of the
        // parity the frame started in runs now
        LF_HIP_CHECK(hipGraphLaunch(exec_[k], stream_));
        return LF_OK;
    }
    // events of the stages selected by set_timing (stage s: bit s)
    Ev fr{4, nullptr, nullptr}, ig{0, nullptr, nullptr}, rc{1, nullptr, nullptr}, tr{2, nullptr, nullptr};
    const bool t_fr = timing_ & LF_TIMING_FRAME, t_ig = (timing_ & LF_TIMING_INTEGRATE) && mapping_enabled,
               t_rc = timing_ & LF_TIMING_RAYCAST, t_tr = (timing_ & LF_TIMING_TRACKING) && !odometry_->trackingLost;
    for (Ev *e : {&fr, &ig, &rc, &tr}) {
        const bool on = e == &fr ? t_fr : e == &ig ? t_ig : e == &rc ? t_rc : t_tr;
        if (on) { e->a = take_event(); e->b = take_event(); }
    }
    if (t_fr) LF_HIP_CHECK(hipEventRecord(fr.a, stream_));
    if (!od.trackingLost) {
        if (t_tr) LF_HIP_CHECK(hipEventRecord(tr.a, stream_));
        LF_RC(od.trackFrame(depth, rgb, cfg_.depth_scale, initial_pose_));
        if (t_tr) LF_HIP_CHECK(hipEventRecord(tr.b, stream_));
    }
    DeviceImage &ref = od.get_reference_image();
    if (mapping_enabled) LF_RC(mapping_->update(ref, od.pose24(), cfg_.integrate_colour != 0, ig.a, ig.b));
    LF_RC(mapping_->raycast(ref.get_vmap(0).ptr<float>(), od.pose24(), rc.a, rc.b));
    LF_RC(ref.resize_device_map());
    if (t_fr) LF_HIP_CHECK(hipEventRecord(fr.b, stream_));
    for (Ev *e : {&fr, &tr, &ig, &rc})
        if (e->a) events_.push_back(*e);
    return LF_OK;
}

int System::process_images_host(const uint16_t *depth, const unsigned char *rgb) {
    const int W = cfg_.K.width, H = cfg_.K.height;
    try {
        host_depth_.create(W, H, 2);
        host_rgb_.create(W, H, 3);
    } catch (const std::bad_alloc &) {
        return lf_set_error(LF_ERR_HIP, "device allocation of the input staging failed");
    }
    // the staging was last read by the previous frame's upload, earlier on this stream
    LF_HIP_CHECK(hipMemcpyAsync(host_depth_.ptr<void>(), depth, host_depth_.bytes(), hipMemcpyHostToDevice, stream_));
    LF_HIP_CHECK(hipMemcpyAsync(host_rgb_.ptr<void>(), rgb, host_rgb_.bytes(), hipMemcpyHostToDevice, stream_));
    return process_images(host_depth_.ptr<uint16_t>(), host_rgb_.ptr<unsigned char>());
}

int System::writeMapToDisk(const std::string &file_name) { return mapping_->writeMapToDisk(file_name); }

int System::readMapFromDisk(const std::string &file_name) { return mapping_->readMapFromDisk(file_name); }

int System::restart() {  // system.cpp:148-157
    double T[16];
    LF_RC(get_camera_pose(T));
    double T12[12];
    for (int i = 0; i < 12; ++i) T12[i] = T[i];
    initial_pose_ = SE3d::from3x4(T12);
    LF_RC(mapping_->reset_mapping());
    odometry_->reset();
    return LF_OK;
}

size_t System::fetch_mesh_vertex_only(float *v, int max_tri) { return mapping_->fetch_mesh_vertex_only(v, max_tri); }
size_t System::fetch_mesh_with_normal(float *v, float *n, int max_tri) {
    return mapping_->fetch_mesh_with_normal(v, n, max_tri);
}
size_t System::fetch_mesh_with_colour(float *v, unsigned char *c, int max_tri) {
    return mapping_->fetch_mesh_with_colour(v, c, max_tri);
}

int System::get_camera_pose(double *out16) {
    LfTrackState h;
    LF_HIP_CHECK(hipMemcpyAsync(&h, odometry_->tracker().state(), sizeof(h), hipMemcpyDeviceToHost, stream_));
    LF_HIP_CHECK(hipStreamSynchronize(stream_));
    for (int i = 0; i < 12; ++i) out16[i] = h.pose[i];
    out16[12] = out16[13] = out16[14] = 0;
    out16[15] = 1;
    return LF_OK;
}

}  // namespace fusion

// ------------------------------------------------------------ C-ABI

// ------------------------------------------------------------ lf_image_*

struct lf_image {
    fusion::DeviceImage *img;
};

static int check_halving(const lf_intrinsics *K, int num_pyr) {
    if (num_pyr < 1 || num_pyr > 8) return lf_set_error(LF_ERR_ARG, "num_pyr must be 1..8");
    if (K->width <= 0 || K->height <= 0) return lf_set_error(LF_ERR_ARG, "bad image size %dx%d", K->width, K->height);
    for (int l = 0; l + 1 < num_pyr; ++l) {
        const int sh = 2 << l;
        if (K->width % sh != 0 || K->height % sh != 0)
            return lf_set_error(LF_ERR_ARG, "%dx%d does not halve evenly over %d levels", K->width, K->height, num_pyr);
    }
    return LF_OK;
}

extern "C" int lf_image_create(const lf_intrinsics *K, int num_pyr, void *stream, lf_image **out) {
    if (!K || !out) return lf_set_error(LF_ERR_ARG, "null argument to lf_image_create");
    if (int rc = check_halving(K, num_pyr)) return rc;
    fusion::IntrinsicMatrix k(K->width, K->height, K->fx, K->fy, K->cx, K->cy);
    try {
        *out = new lf_image{new fusion::DeviceImage(k, num_pyr, lf_stream(stream))};
    } catch (const std::bad_alloc &) {
        return lf_set_error(LF_ERR_HIP, "device allocation failed in lf_image_create");
    }
    return LF_OK;
}

extern "C" int lf_image_destroy(lf_image *im) {
    if (!im) return LF_OK;
    delete im->img;
    delete im;
    return LF_OK;
}

extern "C" int lf_image_upload(lf_image *im, const uint16_t *depth, const unsigned char *rgb, float depth_scale) {
    if (!im || !depth || !rgb) return lf_set_error(LF_ERR_ARG, "null argument to lf_image_upload");
    return im->img->upload(depth, rgb, depth_scale);
}

extern "C" int lf_image_resize_device_map(lf_image *im) {
    if (!im) return lf_set_error(LF_ERR_ARG, "null image");
    return im->img->resize_device_map();
}

static int image_plane(fusion::DeviceImage &d, int level, int plane, void **out) {
    if (level < 0 || level >= d.levels()) return lf_set_error(LF_ERR_ARG, "bad level %d", level);
    if ((plane == LF_PLANE_RAW_DEPTH || plane == LF_PLANE_RGB || plane == LF_PLANE_SAMPLES) && level != 0)
        return lf_set_error(LF_ERR_ARG, "plane %d exists at level 0 only", plane);
    switch (plane) {
    case LF_PLANE_DEPTH: *out = d.get_depth(level).ptr<void>(); break;
    case LF_PLANE_INTENSITY: *out = d.get_intensity(level).ptr<void>(); break;
    case LF_PLANE_DX: *out = d.get_intensity_dx(level).ptr<void>(); break;
    case LF_PLANE_DY: *out = d.get_intensity_dy(level).ptr<void>(); break;
    case LF_PLANE_VMAP: *out = d.get_vmap(level).ptr<void>(); break;
    case LF_PLANE_NMAP: *out = d.get_nmap(level).ptr<void>(); break;
    case LF_PLANE_RAW_DEPTH: *out = d.get_raw_depth().ptr<void>(); break;
    case LF_PLANE_RGB: *out = d.get_image().ptr<void>(); break;
    case LF_PLANE_SAMPLES: *out = d.get_integration_samples().ptr<void>(); break;
    default: return lf_set_error(LF_ERR_ARG, "bad plane %d", plane);
    }
    return LF_OK;
}

extern "C" int lf_image_render(lf_image *im, int textured, void **rgba_out) {
    if (!im || !rgba_out) return lf_set_error(LF_ERR_ARG, "null argument to lf_image_render");
    const fusion::GpuImage &g = textured ? im->img->get_rendered_scene_textured() : im->img->get_rendered_image();
    *rgba_out = g.ptr<void>();
    return LF_OK;
}

extern "C" int lf_image_plane(lf_image *im, int level, int plane, void **out) {
    if (!im || !out) return lf_set_error(LF_ERR_ARG, "null argument to lf_image_plane");
    return image_plane(*im->img, level, plane, out);
}

// ------------------------------------------------------------ lf_tracker_*

struct lf_tracker {
    fusion::DenseTracking *t;
};

extern "C" int lf_tracker_create(const lf_intrinsics *K, int num_pyr, void *stream, lf_tracker **out) {
    if (!K || !out) return lf_set_error(LF_ERR_ARG, "null argument to lf_tracker_create");
    if (int rc = check_halving(K, num_pyr)) return rc;
    fusion::IntrinsicMatrix k(K->width, K->height, K->fx, K->fy, K->cx, K->cy);
    try {
        *out = new lf_tracker{new fusion::DenseTracking(k, num_pyr, lf_stream(stream))};
    } catch (const std::bad_alloc &) {
        return lf_set_error(LF_ERR_HIP, "device allocation failed in lf_tracker_create");
    }
    return LF_OK;
}

extern "C" int lf_tracker_destroy(lf_tracker *t) {
    if (!t) return LF_OK;
    delete t->t;
    delete t;
    return LF_OK;
}

#define LF_TRACKER_SET(name, type)                                                        \
    extern "C" int lf_tracker_##name(lf_tracker *t, const type *plane) {                 \
        if (!t || !plane) return lf_set_error(LF_ERR_ARG, "null argument to lf_tracker_" #name); \
        return t->t->name(plane);                                                         \
    }
LF_TRACKER_SET(set_source_vmap, float)
LF_TRACKER_SET(set_source_image, unsigned char)
LF_TRACKER_SET(set_source_depth, float)
LF_TRACKER_SET(set_source_intensity, float)
LF_TRACKER_SET(set_reference_image, unsigned char)
LF_TRACKER_SET(set_reference_intensity, float)
LF_TRACKER_SET(set_reference_vmap, float)
#undef LF_TRACKER_SET

extern "C" int lf_tracker_swap_intensity(lf_tracker *t) {
    if (!t) return lf_set_error(LF_ERR_ARG, "null tracker");
    t->t->swap_intensity_pyr();
    return LF_OK;
}

static int tracking_context(const double *init12, const int *iterations, int num_levels, int max_levels,
                            fusion::TrackingContext &c) {
    if (!iterations || num_levels < 1 || num_levels > max_levels)
        return lf_set_error(LF_ERR_ARG, "need 1..%d levels of iterations", max_levels);
    c.use_initial_guess_ = init12 != nullptr;
    if (init12) c.initial_estimate_ = fusion::SE3d::from3x4(init12);
    c.max_iterations_.assign(iterations, iterations + num_levels);
    return LF_OK;
}

extern "C" int lf_tracker_get_sums(lf_tracker *t, double *sums31) {
    if (!t || !sums31) return lf_set_error(LF_ERR_ARG, "null argument to lf_tracker_get_sums");
    return t->t->read_sums(sums31);
}

static int read_result(fusion::DenseTracking &t, lf_tracking_result *out) {
    const fusion::TrackingResult r = t.result();
    out->success = r.sucess ? 1 : 0;
    out->icp_error = r.icp_error;
    r.update.matrix3x4(out->update);
    return LF_OK;
}

extern "C" int lf_tracker_compute_transform(lf_tracker *t, const double *init12, const int *iterations,
                                            int num_levels, lf_tracking_result *out) {
    if (!t || !out) return lf_set_error(LF_ERR_ARG, "null argument to lf_tracker_compute_transform");
    fusion::TrackingContext c;
    if (int rc = tracking_context(init12, iterations, num_levels, t->t->source()->levels(), c)) return rc;
    LF_RC(t->t->compute_transform(c));
    return read_result(*t->t, out);
}

extern "C" int lf_tracker_compute_transform_images(lf_tracker *t, lf_image *reference, lf_image *current,
                                                   const double *init12, const int *iterations, int num_levels,
                                                   lf_tracking_result *out) {
    if (!t || !reference || !current || !out)
        return lf_set_error(LF_ERR_ARG, "null argument to lf_tracker_compute_transform_images");
    const int levels = std::min(reference->img->levels(), current->img->levels());
    fusion::TrackingContext c;
    if (int rc = tracking_context(init12, iterations, num_levels, levels, c)) return rc;
    c.intrinsics_pyr_ = reference->img->intrinsics();
    LF_RC(t->t->compute_transform(*reference->img, *current->img, c));
    return read_result(*t->t, out);
}

extern "C" int lf_tracker_plane(lf_tracker *t, int source, int level, int plane, void **out) {
    if (!t || !out) return lf_set_error(LF_ERR_ARG, "null argument to lf_tracker_plane");
    return image_plane(source ? *t->t->source() : *t->t->reference(), level, plane, out);
}

// update(RgbdImagePtr frame) (voxel_hashing.cpp:44-65) on a caller's map: the
// frame's raw depth for allocation, its packed integration samples (depth,
// weight from the level-0 normals, colour) for the weighted update.
extern "C" int lf_update_image(const lf_map_params *p, const lf_map_storage *s, lf_image *im, const float *pose24,
                               int colour, void *stream) {
    if (!p || !s || !im || !pose24) return lf_set_error(LF_ERR_ARG, "null argument to lf_update_image");
    fusion::DeviceImage &d = *im->img;
    const lf_intrinsics &K = d.intrinsics()[0];
    return lf_update_ev(p, s, d.get_raw_depth().ptr<float>(), d.get_nmap(0).ptr<float>(),
                        colour ? d.get_image().ptr<unsigned char>() : nullptr, K.width, K.height, pose24, &K,
                        lf_stream(stream), true, nullptr, nullptr, d.get_integration_samples().ptr<float>());
}

extern "C" int lf_deintegrate_image(const lf_map_params *p, const lf_map_storage *s, lf_image *im,
                                    const float *pose24, int colour, void *stream) {
    if (!p || !s || !im || !pose24) return lf_set_error(LF_ERR_ARG, "null argument to lf_deintegrate_image");
    fusion::DeviceImage &d = *im->img;
    const lf_intrinsics &K = d.intrinsics()[0];
    return lf_deintegrate_ev(p, s, d.get_raw_depth().ptr<float>(), d.get_nmap(0).ptr<float>(),
                             colour ? d.get_image().ptr<unsigned char>() : nullptr, K.width, K.height, pose24, &K,
                             lf_stream(stream), d.get_integration_samples().ptr<float>());
}

extern "C" int lf_refuse_keyframes(const lf_map_params *p, const lf_map_storage *s, lf_image *const *frames,
                                   const float *const *old_pose24, const float *const *new_pose24, int n, int colour,
                                   void *stream) {
    if (!p || !s || !frames || !old_pose24 || !new_pose24)
        return lf_set_error(LF_ERR_ARG, "null argument to lf_refuse_keyframes");
    if (n < 1 || n > LF_REFUSE_MAX_KEYFRAMES)
        return lf_set_error(LF_ERR_ARG, "lf_refuse_keyframes takes 1..%d keyframes, not %d", LF_REFUSE_MAX_KEYFRAMES, n);
    const float *depth[LF_REFUSE_MAX_KEYFRAMES], *samples[LF_REFUSE_MAX_KEYFRAMES];
    const lf_intrinsics *K0 = nullptr;
    for (int k = 0; k < n; ++k) {
        if (!frames[k]) return lf_set_error(LF_ERR_ARG, "null keyframe %d", k);
        fusion::DeviceImage &d = *frames[k]->img;
        const lf_intrinsics &K = d.intrinsics()[0];
        if (!K0) K0 = &d.intrinsics()[0];
        else if (K.width != K0->width || K.height != K0->height || K.fx != K0->fx || K.fy != K0->fy ||
                 K.cx != K0->cx || K.cy != K0->cy)
            return lf_set_error(LF_ERR_ARG, "keyframe %d's size or intrinsics differ from keyframe 0's", k);
        depth[k] = d.get_raw_depth().ptr<float>();
        samples[k] = d.get_integration_samples().ptr<float>();
    }
    return lf_refuse_ev(p, s, depth, samples, old_pose24, new_pose24, n, K0->width, K0->height, K0, colour,
                        lf_stream(stream));
}

// ------------------------------------------------------------ lf_system_*

struct lf_system {
    fusion::System *sys;
};

extern "C" int lf_system_create(const lf_system_config *cfg, lf_system **out) {
    if (!cfg || !out) return lf_set_error(LF_ERR_ARG, "null argument to lf_system_create");
    if (int rc = check_halving(&cfg->K, cfg->num_pyr)) return rc;
    try {
        *out = new lf_system{new fusion::System(*cfg)};
    } catch (const std::bad_alloc &) {
        return lf_set_error(LF_ERR_HIP, "device allocation failed in lf_system_create: %s", lf_last_error());
    }
    return LF_OK;
}

extern "C" int lf_system_destroy(lf_system *s) {
    if (!s) return LF_OK;
    delete s->sys;
    delete s;
    return LF_OK;
}

extern "C" int lf_system_process_images(lf_system *s, const uint16_t *depth, const unsigned char *rgb) {
    if (!s || !depth || !rgb) return lf_set_error(LF_ERR_ARG, "null argument to process_images");
    return s->sys->process_images(depth, rgb);
}

extern "C" int lf_system_process_images_host(lf_system *s, const uint16_t *depth, const unsigned char *rgb) {
    if (!s || !depth || !rgb) return lf_set_error(LF_ERR_ARG, "null argument to process_images_host");
    return s->sys->process_images_host(depth, rgb);
}

extern "C" int lf_system_synchronize(lf_system *s) {
    LF_HIP_CHECK(hipStreamSynchronize(s->sys->stream()));
    return LF_OK;
}

extern "C" int lf_system_get_pose(lf_system *s, double *out16) { return s->sys->get_camera_pose(out16); }

extern "C" int lf_system_get_tracking(lf_system *s, int *success, float *icp_error) {
    fusion::TrackingResult r = s->sys->odometry().tracker().result();
    if (success) *success = r.sucess ? 1 : 0;
    if (icp_error) *icp_error = r.icp_error;
    return LF_OK;
}

extern "C" int lf_system_restart(lf_system *s) { return s->sys->restart(); }

extern "C" int lf_system_fetch_mesh(lf_system *s, float *vertex_out, float *normal_out, unsigned char *colour_out,
                                    int max_triangles, unsigned int *triangles_out) {
    const fusion::DenseMapping &m = s->sys->mapping();
    return lf_create_mesh(&m.params(), &m.storage(), vertex_out, normal_out, colour_out, max_triangles,
                          triangles_out, s->sys->stream());
}

extern "C" int lf_system_map(lf_system *s, lf_map_params *p, lf_map_storage *st) {
    if (p) *p = s->sys->mapping().params();
    if (st) *st = s->sys->mapping().storage();
    return LF_OK;
}

extern "C" int lf_system_maps(lf_system *s, int level, const float **vmap, const float **nmap) {
    fusion::DeviceImage &ref = s->sys->odometry().get_reference_image();
    if (level < 0 || level >= ref.levels()) return lf_set_error(LF_ERR_ARG, "bad level %d", level);
    if (vmap) *vmap = ref.get_vmap(level).ptr<float>();
    if (nmap) *nmap = ref.get_nmap(level).ptr<float>();
    return LF_OK;
}

extern "C" void *lf_system_stream(lf_system *s) { return s->sys->stream(); }

extern "C" int lf_system_render(lf_system *s, int which, void **rgba_out) {
    if (!s || !rgba_out || which < 0 || which > 2) return lf_set_error(LF_ERR_ARG, "bad arguments to lf_system_render");
    fusion::DenseOdometry &od = s->sys->odometry();
    fusion::DeviceImage &img = which == 0 ? od.get_current_image() : od.get_reference_image();
    const fusion::GpuImage &g = which == 2 ? img.get_rendered_scene_textured() : img.get_rendered_image();
    LF_HIP_CHECK(hipStreamSynchronize(s->sys->stream()));
    *rgba_out = g.ptr<void>();
    return LF_OK;
}

extern "C" int lf_system_write_map(lf_system *s, const char *path) {
    if (!s || !path) return lf_set_error(LF_ERR_ARG, "null argument to lf_system_write_map");
    return s->sys->writeMapToDisk(path);
}

extern "C" int lf_system_read_map(lf_system *s, const char *path) {
    if (!s || !path) return lf_set_error(LF_ERR_ARG, "null argument to lf_system_read_map");
    return s->sys->readMapFromDisk(path);
}

extern "C" int lf_system_set_graphs(lf_system *s, int enable) {
    if (!s) return lf_set_error(LF_ERR_ARG, "null system");
    s->sys->set_graphs(enable != 0);
    return LF_OK;
}

extern "C" int lf_system_set_timing(lf_system *s, int stages) {
    s->sys->set_timing(stages);
    return LF_OK;
}

extern "C" int lf_system_read_timing(lf_system *s, int stage, int *launches, double *total_ms) {
    return s->sys->read_timing(stage, launches, total_ms);
}

// ------------------------------------------------------------ batch C-ABI

struct lf_batch {
    fusion::Batch *b;
};

extern "C" int lf_batch_create(const lf_system_config *cfgs, int n, lf_batch **out) {
    if (!cfgs || !out || n <= 0) return lf_set_error(LF_ERR_ARG, "bad arguments to lf_batch_create");
    for (int i = 0; i < n; ++i) {
        if (int rc = check_halving(&cfgs[i].K, cfgs[i].num_pyr)) return rc;
        if (cfgs[i].map.precision != cfgs[0].map.precision)
            return lf_set_error(LF_ERR_ARG, "the systems of a batch share one precision (system %d differs)", i);
    }
    try {
        *out = new lf_batch{new fusion::Batch(cfgs, n)};
    } catch (const std::exception &e) {
        return lf_set_error(LF_ERR_HIP, "batch creation failed: %s", e.what());
    }
    return LF_OK;
}

extern "C" int lf_batch_destroy(lf_batch *b) {
    if (!b) return LF_OK;
    delete b->b;
    delete b;
    return LF_OK;
}

extern "C" int lf_batch_process(lf_batch *b, const uint16_t *const *depth, const unsigned char *const *rgb) {
    if (!b || !depth || !rgb) return lf_set_error(LF_ERR_ARG, "null argument to lf_batch_process");
    for (int i = 0; i < b->b->size(); ++i)
        if (!depth[i] || !rgb[i]) return lf_set_error(LF_ERR_ARG, "null frame for batch system %d", i);
    return b->b->process(depth, rgb);
}

extern "C" int lf_batch_process_host(lf_batch *b, const uint16_t *const *depth, const unsigned char *const *rgb) {
    if (!b || !depth || !rgb) return lf_set_error(LF_ERR_ARG, "null argument to lf_batch_process_host");
    for (int i = 0; i < b->b->size(); ++i)
        if (!depth[i] || !rgb[i]) return lf_set_error(LF_ERR_ARG, "null frame for batch system %d", i);
    return b->b->process_host(depth, rgb);
}

extern "C" int lf_batch_autotune(lf_batch *b, int steps_per_trial, int rounds) {
    if (!b) return lf_set_error(LF_ERR_ARG, "null batch");
    return b->b->autotune(steps_per_trial, rounds);
}

extern "C" int lf_batch_tuning(lf_batch *b) { return b && b->b->tuning() ? 1 : 0; }

extern "C" int lf_batch_tune_file(const char *path) { return fusion::Batch::set_tune_file(path ? path : ""); }

extern "C" int lf_batch_set_graphs(lf_batch *b, int enable) {
    if (!b) return lf_set_error(LF_ERR_ARG, "null batch");
    b->b->set_graphs(enable != 0);
    return LF_OK;
}

extern "C" int lf_batch_set_option(lf_batch *b, int option, int value) {
    if (!b) return lf_set_error(LF_ERR_ARG, "null batch");
    return b->b->set_option(option, value);
}

extern "C" int lf_batch_get_option(lf_batch *b, int option, int *value) {
    if (!b || !value) return lf_set_error(LF_ERR_ARG, "null argument to lf_batch_get_option");
    return b->b->get_option(option, value);
}

extern "C" int lf_batch_set_timing(lf_batch *b, int stages) {
    if (!b) return lf_set_error(LF_ERR_ARG, "null batch");
    b->b->set_timing(stages);
    return LF_OK;
}

extern "C" int lf_batch_read_timing(lf_batch *b, int stage, int *launches, double *total_ms) {
    if (!b || !launches || !total_ms) return lf_set_error(LF_ERR_ARG, "null argument to lf_batch_read_timing");
    return b->b->read_timing(stage, launches, total_ms);
}

extern "C" int lf_batch_synchronize(lf_batch *b) {
    if (!b) return lf_set_error(LF_ERR_ARG, "null batch");
    return b->b->synchronize();
}

static int batch_index(lf_batch *b, int i) {
    if (!b) return lf_set_error(LF_ERR_ARG, "null batch");
    if (i < 0 || i >= b->b->size()) return lf_set_error(LF_ERR_ARG, "batch system %d of %d", i, b->b->size());
    return LF_OK;
}

extern "C" int lf_batch_get_pose(lf_batch *b, int i, double *out16) {
    if (int rc = batch_index(b, i)) return rc;
    if (!out16) return lf_set_error(LF_ERR_ARG, "null pose output");
    return b->b->system(i).get_camera_pose(out16);
}

extern "C" int lf_batch_get_tracking(lf_batch *b, int i, int *success, float *icp_error) {
    if (int rc = batch_index(b, i)) return rc;
    fusion::TrackingResult r = b->b->system(i).odometry().tracker().result();
    if (success) *success = r.sucess ? 1 : 0;
    if (icp_error) *icp_error = r.icp_error;
    return LF_OK;
}

extern "C" int lf_batch_map(lf_batch *b, int i, lf_map_params *p, lf_map_storage *st) {
    if (int rc = batch_index(b, i)) return rc;
    fusion::System &s = b->b->system(i);
    if (p) *p = s.mapping().params();
    if (st) *st = s.mapping().storage();
    return LF_OK;
}

extern "C" int lf_batch_read_counters(lf_batch *b, int i, unsigned int *out16) {
    if (int rc = batch_index(b, i)) return rc;
    fusion::System &s = b->b->system(i);
    return lf_map_read_counters(&s.mapping().storage(), out16, s.stream());
}
