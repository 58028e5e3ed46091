// C++ drop-in layer: libfusion's host classes (DenseMapping, DeviceImage,
// DenseTracking, DenseOdometry, System) re-implemented over the C-ABI of
// libfusion_amd.h.  Method names and meanings follow the reference headers
// (include/voxel_hashing/voxel_hashing.h:12-51, include/tracking/device_image.h:18-80,
// include/tracking/icp_tracker.h:13-88, include/tracking/rgbd_odometry.h:13-41,
// include/system.h:19-92).  Types the reference takes from third-party
// libraries change because those libraries are not part of this build:
//   cv::cuda::GpuMat -> fusion::GpuImage  (owning dense device buffer)
//   Sophus::SE3d     -> fusion::SE3d      (FP64 rigid transform)
//   cv::Mat input    -> device pointers to uint16 depth / uint8x3 RGB
// Every method enqueues work on the owning object's HIP stream and returns
// without waiting, except the ones documented as synchronising.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "libfusion_amd.h"

struct LfTrackState;
struct LfImageLevels;
struct LfFuseItem;

namespace fusion {

// IntrinsicMatrix (intrinsic_matrix.h:12-47).
struct IntrinsicMatrix : lf_intrinsics {
    IntrinsicMatrix();
    IntrinsicMatrix(int cols, int rows, float fx, float fy, float cx, float cy);
    IntrinsicMatrix pyr_down() const;
};
void BuildIntrinsicPyramid(IntrinsicMatrix base, std::vector<IntrinsicMatrix> &pyr, int num_pyr);

// FP64 rigid transform in place of Sophus::SE3d.
struct SE3d {
    double R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double t[3] = {0, 0, 0};
    static SE3d exp(const double *xi);  // tangent [upsilon; omega]
    void log(double *xi) const;
    SE3d inverse() const;
    SE3d operator*(const SE3d &o) const;
    void matrix3x4(double *out12) const;
    static SE3d from3x4(const double *in12);
};

// Owning dense device buffer in place of cv::cuda::GpuMat.
class GpuImage {
public:
    GpuImage() = default;
    GpuImage(int cols, int rows, int elem_bytes) { create(cols, rows, elem_bytes); }
    ~GpuImage();
    GpuImage(const GpuImage &) = delete;
    GpuImage &operator=(const GpuImage &) = delete;
    GpuImage(GpuImage &&o) noexcept;
    GpuImage &operator=(GpuImage &&o) noexcept;
    void create(int cols, int rows, int elem_bytes);
    bool empty() const { return data_ == nullptr; }
    size_t bytes() const { return (size_t)cols * rows * elem; }
    template <class T> T *ptr() const { return static_cast<T *>(data_); }
    int cols = 0, rows = 0, elem = 0;

private:
    void *data_ = nullptr;
};

// DeviceImage (device_image.h:18-80): one RGB-D frame and its pyramids.
class DeviceImage {
public:
    // half_maps: intensity, gradients, vertices and normals stored as FP16
    // (mixed precision); the set_* builders below take FP32 images only.
    DeviceImage(const IntrinsicMatrix &K, int num_pyr, hipStream_t stream, bool half_maps = false);
    bool half_maps() const { return half_maps_; }
    // DeviceImage::upload (device_image.cpp:146-181) from device-resident input.
    int upload(const uint16_t *depth_u16, const unsigned char *rgb, float depth_scale);
    // DeviceImage::resize_device_map (device_image.cpp:188-199).
    int resize_device_map();
    // upload / resize_device_map planned as frame-pass descriptors appended
    // to out (launched by a Batch together with other images' passes).
    int plan_upload(const uint16_t *depth_u16, const unsigned char *rgb, float depth_scale,
                    std::vector<unsigned char> &out);
    int plan_refresh(std::vector<unsigned char> &out);
    // Pyramids from given level-0 device planes (build_pyramid.cpp:7-76), as
    // DenseTracking::set_source_* / set_reference_* build them:
    int set_vmap(const float *vmap);                 // vertices, then every level's vertices and normals
    int set_intensity(const float *intensity);       // intensity pyramid and derivatives
    int set_image(const unsigned char *rgb);         // grey conversion, then set_intensity
    int set_depth(const float *depth);               // bilateral filter and depth pyramid
    // get_rendered_image / get_rendered_scene_textured (device_image.cpp:119-129):
    // RGBA8 shading of the level-0 vertex / normal maps (enqueued)
    const GpuImage &get_rendered_image();
    const GpuImage &get_rendered_scene_textured();
    // exchange the intensity pyramids of two frames (gradients stay)
    void swap_intensity(DeviceImage &o) { std::swap(intensity_, o.intensity_); }
    const GpuImage &get_raw_depth() const { return depth_raw_; }
    const GpuImage &get_image() const { return image_; }
    // level-0 integration samples: float4 {depth or NaN, weight, colour bits, 0} per pixel
    const GpuImage &get_integration_samples() const { return samples_; }
    GpuImage &get_depth(int l = 0) { return depth_[l]; }
    GpuImage &get_vmap(int l = 0) { return vmap_[l]; }
    GpuImage &get_nmap(int l = 0) { return nmap_[l]; }
    GpuImage &get_intensity(int l = 0) { return intensity_[l]; }
    GpuImage &get_intensity_dx(int l = 0) { return dx_[l]; }
    GpuImage &get_intensity_dy(int l = 0) { return dy_[l]; }
    const std::vector<IntrinsicMatrix> &intrinsics() const { return K_; }
    int levels() const { return (int)K_.size(); }

private:
    LfImageLevels levels_view();
    std::vector<IntrinsicMatrix> K_;
    hipStream_t stream_;
    int intensity_pyramid();
    GpuImage depth_raw_, image_, samples_, rendered_;
    std::vector<GpuImage> depth_, intensity_, dx_, dy_, vmap_, nmap_;
    bool half_maps_ = false;
};

// DenseMapping (voxel_hashing.h:12-51).
class DenseMapping {
public:
    DenseMapping(const IntrinsicMatrix &K, const lf_map_params &params, hipStream_t stream);
    ~DenseMapping();
    DenseMapping(const DenseMapping &) = delete;
    DenseMapping &operator=(const DenseMapping &) = delete;
    // update(RgbdImagePtr) (voxel_hashing.cpp:44-65): raw depth, level-0 normals, colour.
    int update(DeviceImage &frame, const float *pose24_dev, bool colour, hipEvent_t e0 = nullptr,
               hipEvent_t e1 = nullptr);
    // update(depth, image, pose) (voxel_hashing.cpp:67-85): weight 1/d^2, no normals.
    int update(const float *depth, const unsigned char *rgb, const float *pose24_dev);
    // raycast (voxel_hashing.cpp:87-119) into a float4 vertex map.
    int raycast(float *vmap, const float *pose24_dev, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
    // raycast_check_visibility (voxel_hashing.cpp:121-133): visible blocks from the pose, then raycast.
    int raycast_check_visibility(float *vmap, const float *pose24_dev);
    int reset_mapping();
    // Remove a frame fused with update(frame, pose, colour) (inverse weighted update).
    int deintegrate(DeviceImage &frame, const float *pose24_dev, bool colour);
    // writeMapToDisk / readMapFromDisk (voxel_hashing.cpp:189-205): the reference's file layout.
    int writeMapToDisk(const std::string &file_name);
    int readMapFromDisk(const std::string &file_name);
    // fetch_mesh_* (voxel_hashing.cpp:142-187); synchronising; return triangle counts.
    size_t fetch_mesh_vertex_only(float *vertex, int max_triangles);
    size_t fetch_mesh_with_normal(float *vertex, float *normal, int max_triangles);
    size_t fetch_mesh_with_colour(float *vertex, unsigned char *colour, int max_triangles);
    const lf_map_params &params() const { return params_; }
    const IntrinsicMatrix &intrinsics() const { return K_; }
    const lf_map_storage &storage() const { return storage_; }

private:
    IntrinsicMatrix K_;
    lf_map_params params_;
    lf_map_storage storage_;
    hipStream_t stream_;
};

struct TrackingResult {  // icp_tracker.h:13-18 (read back with DenseTracking::result)
    bool sucess = false;
    float icp_error = 0;
    SE3d update;
};

struct TrackingContext {  // icp_tracker.h:20-26
    bool use_initial_guess_ = true;
    std::vector<IntrinsicMatrix> intrinsics_pyr_;
    std::vector<int> max_iterations_;
    SE3d initial_estimate_;
};

// DenseTracking (icp_tracker.h:28-88).  The Gauss-Newton loop runs on the
// device; compute_transform only enqueues it.
class DenseTracking {
public:
    explicit DenseTracking(hipStream_t stream);
    // DenseTracking(K, NUM_PYR) (icp_tracker.cpp:8-21): owns source and reference
    // pyramids for the set_* / compute_transform(ctx) interface.
    DenseTracking(const IntrinsicMatrix &K, int num_pyr, hipStream_t stream);
    ~DenseTracking();
    DenseTracking(const DenseTracking &) = delete;
    DenseTracking &operator=(const DenseTracking &) = delete;
    // compute_transform(reference, current, ctx) (icp_tracker.cpp:23-156)
    int compute_transform(DeviceImage &reference, DeviceImage &current, const TrackingContext &c);
    // The same frame as launch arguments instead of launches (for a Batch):
    // one job of lf_track_job_bytes() bytes per Gauss-Newton step appended to
    // jobs, its grid to grids.  A plan without steps is launched at once.
    int plan(DeviceImage &reference, DeviceImage &current, const TrackingContext &c,
             std::vector<unsigned char> &jobs, std::vector<int> &grids);
    // Planned steps of a tracker that shares its launches with share - 1
    // others take 1/share of the reduction grid (a Batch sets it).
    void set_batch_share(int share) { batch_share_ = share < 1 ? 1 : share; }
    // wgs > 0: planned steps take exactly that many workgroups (at most a
    // lone system's grid) instead of 1/share of it.
    void set_batch_workgroups(int wgs) { batch_wgs_ = wgs < 0 ? 0 : wgs; }
    // compute_transform(ctx) (icp_tracker.cpp:210-314) on the set_* pyramids
    // (source = current frame, reference = last); success is unconditional.
    int compute_transform(const TrackingContext &c);
    // set_source_* / set_reference_* (icp_tracker.cpp:164-208) from device planes
    // at level-0 size: vmap float4, image RGB u8, depth and intensity float.
    int set_source_vmap(const float *vmap) { return src_->set_vmap(vmap); }
    int set_source_image(const unsigned char *rgb) { return src_->set_image(rgb); }
    int set_source_depth(const float *depth) { return src_->set_depth(depth); }
    int set_source_intensity(const float *intensity) { return src_->set_intensity(intensity); }
    int set_reference_image(const unsigned char *rgb) { return ref_->set_image(rgb); }
    int set_reference_intensity(const float *intensity) { return ref_->set_intensity(intensity); }
    int set_reference_vmap(const float *vmap) { return ref_->set_vmap(vmap); }
    // swap_intensity_pyr (icp_tracker.cpp:158-162)
    void swap_intensity_pyr();
    DeviceImage *source() { return src_.get(); }
    DeviceImage *reference() { return ref_.get(); }
    LfTrackState *state() const { return state_; }
    // Synchronising read of the last result (update = estimate^-1).
    TrackingResult result();
    // Synchronising read of the last Gauss-Newton step's FP64 sums (31 values).
    int read_sums(double *sums31);
    float icp_weight = 1e6f;  // icp_tracker.cpp:89-90

private:
    int enqueue(DeviceImage &reference, DeviceImage &current, const TrackingContext &c, bool context_mode,
                std::vector<unsigned char> *jobs = nullptr, std::vector<int> *grids = nullptr);
    hipStream_t stream_;
    LfTrackState *state_ = nullptr;
    float *scratch_ = nullptr;
    bool context_mode_ = false;
    int batch_share_ = 1;
    int batch_wgs_ = 0;
    std::unique_ptr<DeviceImage> src_, ref_;
};

// DenseOdometry (rgbd_odometry.h:13-41).
class DenseOdometry {
public:
    DenseOdometry(const IntrinsicMatrix &base, int num_pyr, std::vector<int> iterations, hipStream_t stream,
                  bool half_maps = false);
    // trackFrame (rgbd_odometry.cpp:18-48) on device-resident input.
    int trackFrame(const uint16_t *depth, const unsigned char *rgb, float depth_scale, const SE3d &initial_pose);
    // trackFrame with the Gauss-Newton steps planned into jobs / grids instead
    // of launched (the host bookkeeping happens as in trackFrame); the upload
    // is launched, or planned into *upload when that is given.
    int planFrame(const uint16_t *depth, const unsigned char *rgb, float depth_scale, const SE3d &initial_pose,
                  std::vector<unsigned char> &jobs, std::vector<int> &grids,
                  std::vector<unsigned char> *upload = nullptr);
    void reset();
    // the frame before the reference: the buffer the reference's two-buffer
    // swap leaves as current after a frame (rgbd_odometry.cpp:41,50)
    DeviceImage &get_current_image() { return *spare_; }
    DeviceImage &get_reference_image() { return *ref_; }
    DenseTracking &tracker() { return tracker_; }
    const float *pose24() const;
    bool initialized() const { return initialized_; }
    // which of the three frame buffers is current (0, 1, 2 in turn, one step
    // per tracked frame)
    int buffer_parity() const { return parity_; }
    // the host side of a tracked frame whose device work was replayed from a graph
    void replayed_frame() { advance(); trackingLost = false; }
    bool trackingLost = false;

private:
    // After a frame: it becomes the reference, and the next frame is uploaded
    // into the buffer of the frame before the reference, which the tracking
    // of the reference read last.  With three buffers the next upload does
    // not overwrite anything the current frame's tracking reads, so it can
    // run beside that tracking.
    void advance() {
        std::swap(ref_, spare_);
        std::swap(curr_, ref_);
        parity_ = (parity_ + 1) % 3;
    }
    int parity_ = 0;
    std::vector<IntrinsicMatrix> K_;
    std::vector<int> iterations_;
    hipStream_t stream_;
    std::unique_ptr<DeviceImage> curr_, ref_, spare_;
    DenseTracking tracker_;
    bool initialized_ = false;
};

// System (system.h:19-92).
class System {
public:
    System(const lf_system_config &cfg);
    ~System();
    int process_images(const uint16_t *depth, const unsigned char *rgb);
    // process_images (system.cpp:40-114) from host memory, as the reference
    // takes its cv::Mat frames: the frame is copied into device staging on the
    // system's stream first (asynchronous when the host buffers are pinned).
    int process_images_host(const uint16_t *depth, const unsigned char *rgb);
    int restart();
    // writeMapToDisk / readMapFromDisk (system.cpp:202-210)
    int writeMapToDisk(const std::string &file_name);
    int readMapFromDisk(const std::string &file_name);
    size_t fetch_mesh_vertex_only(float *vertex, int max_triangles);
    size_t fetch_mesh_with_normal(float *vertex, float *normal, int max_triangles);
    size_t fetch_mesh_with_colour(float *vertex, unsigned char *colour, int max_triangles);
    int get_camera_pose(double *out16);  // synchronising
    hipStream_t stream() const { return stream_; }
    DenseMapping &mapping() { return *mapping_; }
    DenseOdometry &odometry() { return *odometry_; }
    // Event timing of stages: 0 integrate kernel, 1 raycast kernel, 2 tracking,
    // 3 image pipeline, 4 whole frame.
    // stage mask of LF_TIMING_* bits (0: off)
    void set_timing(int stages);
    // Replay steady-state frames from HIP graphs: each of the two frame-buffer
    // parities is captured once (the frame's whole device work, inputs copied
    // into fixed buffers first) and replayed afterwards.  Frames with timing,
    // before initialisation or with mapping off run eagerly.
    void set_graphs(bool on);
    int read_timing(int stage, int *launches, double *total_ms);
    bool mapping_enabled = true;
    // The frame of process_images in three parts, for a Batch: tracking and
    // fusion, then (by the caller) the raycast into raycast_vmap() at
    // raycast_pose24(), then the re-pyramid.  No timing, no graphs.
    int enqueue_before_raycast(const uint16_t *depth, const unsigned char *rgb);
    // enqueue_before_raycast in two parts for a Batch: the upload with the
    // tracking steps planned (launched by the batch; the upload too when
    // upload is given), then the fusion.
    int plan_tracking(const uint16_t *depth, const unsigned char *rgb, std::vector<unsigned char> &jobs,
                      std::vector<int> &grids, std::vector<unsigned char> *upload = nullptr);
    int enqueue_fusion();
    float *raycast_vmap();
    const float *raycast_pose24();
    int enqueue_after_raycast();
    // enqueue_after_raycast planned as frame passes, and enqueue_fusion as one
    // map's item of a batched update (both after plan_tracking).
    int plan_after_raycast(std::vector<unsigned char> &passes);
    int plan_fusion(LfFuseItem &item);

private:
    lf_system_config cfg_;
    hipStream_t stream_ = nullptr;
    std::unique_ptr<DenseMapping> mapping_;
    std::unique_ptr<DenseOdometry> odometry_;
    SE3d initial_pose_;
    int timing_ = 0;
    struct Ev { int stage; hipEvent_t a, b; };
    std::vector<Ev> events_;
    std::vector<hipEvent_t> pool_;
    int enqueue_frame(const uint16_t *depth, const unsigned char *rgb);
    void drop_graphs();
    bool graphs_ = false;
    hipGraphExec_t exec_[3] = {nullptr, nullptr, nullptr};  // per frame buffer
    GpuImage depth_in_, rgb_in_;
    GpuImage host_depth_, host_rgb_;  // device staging of process_images_host
    hipEvent_t take_event();
};

// Several independent systems stepped together (no counterpart in the
// reference, which runs one System per process).  A step uploads every
// system's frame, pass p of all systems in one launch, on an upload stream
// that overlaps the previous step's fusion and raycast; runs the Gauss-Newton
// steps of all systems on the batch's stream, step s of every system in one
// launch pair; fuses each map on its own stream; runs one batched z-range +
// raycast launch per stage for all maps; then the re-pyramids, again one
// launch per pass for all systems.  The batched launches replace 2 x 18 tracking
// launches and 6 pyramid launches per system and pack the raycasts' long-ray
// tails onto one grid.
class Batch {
public:
    Batch(const lf_system_config *cfgs, int n);
    ~Batch();
    int process(const uint16_t *const *depth, const unsigned char *const *rgb);
    // process() from host memory (the reference's process_images takes host
    // images): each frame is copied into the batch's device staging buffers on
    // the upload stream, which overlaps the previous step's fusion and
    // raycast; pinned host buffers make the copies asynchronous.
    int process_host(const uint16_t *const *depth, const unsigned char *const *rgb);
    int synchronize();
    int size() const { return (int)sys_.size(); }
    System &system(int i) { return *sys_[i]; }
    hipStream_t stream() const { return stream_; }
    // Event timing on the batch stream (LF_BATCH_TIMING_* bits; 0: off) and
    // {launches, total ms} of a stage (synchronising).
    void set_timing(int stages);
    int read_timing(int stage, int *launches, double *total_ms);
    // Launch options (LF_BATCH_OPT_*): shapes of the batched launches, which
    // leave every system's results unchanged up to summation order.
    int set_option(int option, int value);
    int get_option(int option, int *value) const;
    // Replay steady-state steps from HIP graphs: the step's table copy and
    // uploads, its Gauss-Newton launches, and its fusion / raycast /
    // re-pyramid launches, each captured once per table buffer and launch
    // shape.  Steps with timing on, or where a system starts a frame, run
    // eagerly.  Results are those of the eager steps.
    void set_graphs(bool on);
    // Online autotuning of the launch options: the next steps try each
    // option's candidate values in turn (one option at a time, the best kept
    // before the next), each candidate for one settling step then
    // steps_per_trial timed steps, in `rounds` interleaved rounds, timed by
    // events on the batch stream; the medians decide.  The choice is cached
    // per batch shape (process-wide, and in the file of set_tune_file), and a
    // batch of a cached shape starts with it.  0 stops a running tuning.
    int autotune(int steps_per_trial, int rounds = 2);
    bool tuning() const { return tune_.per_trial > 0; }
    // The file the tuned choices are read from and appended to (empty: none).
    static int set_tune_file(const std::string &path);

private:
    struct Tuner {
        int per_trial = 0, rounds = 2;
        int opt = 0;         // index into the tuned options
        int round = 0, cand = 0, step = 0;  // position in the current option's trials
        std::vector<std::vector<double>> ms;  // per candidate: ms per step of each round
        hipEvent_t a = nullptr;  // end of the trial's settling step
        std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;  // (candidate, trial events)
    };
    Tuner tune_;
    std::string tune_key() const;
    void tune_begin_step();
    int tune_end_step();
    struct StepGraphs {
        hipGraphExec_t up = nullptr, track = nullptr, fuse = nullptr;
        std::vector<long long> sig;  // the launch shapes they were captured with
        void drop();
    };
    bool graphs_ = false;
    StepGraphs graphs_k_[2];  // per table buffer
    int step(const uint16_t *const *depth, const unsigned char *const *rgb, bool from_host);
    int upload_overlap_ = 1;  // uploads on aux_ beside the previous step (1) or on stream_ (0)
    int track_wgs_ = 0;       // tracking workgroups per system (0: the share rule)
    int integrate_wgs_ = 0;   // integration workgroups per map (0: the map-count rule)
    int zrange_wgs_ = 0;      // z-range workgroups per map (0: the map-count rule)
    void release();
    std::vector<std::unique_ptr<System>> sys_;
    hipStream_t stream_ = nullptr;
    std::vector<hipEvent_t> ready_;  // per system: its tracking and fusion are enqueued
    hipStream_t aux_ = nullptr;      // the step's descriptor table and uploads
    hipEvent_t uploaded_ = nullptr;  // the step's uploads
    bool half_maps_ = false;  // the systems are mixed precision
    int parity_ = 0;
    // the step's descriptors (frame passes, tracking jobs, raycast jobs),
    // double-buffered in pinned host and device memory
    void *table_host_[2] = {nullptr, nullptr}, *table_dev_[2] = {nullptr, nullptr};
    hipEvent_t table_read_[2] = {nullptr, nullptr};
    bool table_used_[2] = {false, false};
    size_t table_cap_ = 0;  // bytes per buffer
    int reserve_table(size_t bytes);
    // device staging of host frames (process_host), per system
    std::vector<GpuImage> depth_in_, rgb_in_;
    // stage timing: an event pair per timed stage and step
    int timing_ = 0;
    struct Ev { int stage; hipEvent_t a, b; };
    std::vector<Ev> events_;
    std::vector<hipEvent_t> pool_;
    hipEvent_t take_event();
};

}  // namespace fusion
