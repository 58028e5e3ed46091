// Device-resident Gauss-Newton tracking state shared by reduce.hip (kernels)
// and system.cpp (the host classes that sequence them).
#pragma once

#include <hip/hip_runtime.h>

#include "libfusion_amd.h"

// One tracker's state in device memory.  Poses are 3x4 row-major doubles.
struct LfTrackState {
    double est[12];      // current estimate T (reference camera -> current camera)
    double prev[12];     // Revertable<SE3d>::prev (revertable.h:7-45)
    double init[12];     // initial estimate (identity in DenseOdometry)
    float pose_f[12];    // float 3x4 of est, read by the next reduction
    float icp_error, rgb_error, stddev;
    float last_icp_error;  // icp error before the last step (last_icp_error of compute_transform(ctx))
    int icp_count, rgb_count;
    int active;          // 0 once the current level broke out of its loop
    int invalid;         // an error became NaN this frame
    int success;         // result of the last finished frame
    int iterations;      // Gauss-Newton steps taken this frame
    // odometry chain (DenseOdometry::trackFrame, rgbd_odometry.cpp:18-48)
    double pose[12];     // T_world_camera of the current frame
    float pose24[24];    // float T_world_camera and T_camera_world for the mapping kernels
    int pad[3];
    // the last Gauss-Newton step's FP64 sums: A = w H_icp + H_rgb and b in
    // create_jtjjtr<6,7> order (0-26), ICP r^2 / count (27 / 28), RGB r^2 /
    // count (29 / 30)
    double sums[32];
};

// Per-level inputs of one tracking iteration.
struct LfTrackLevel {
    const float *curr_vmap, *curr_nmap, *last_vmap, *last_nmap;
    const float *curr_I, *last_I, *dIdx, *dIdy;
    int cols, rows;
    lf_intrinsics K;
    int half_maps;  // the maps hold FP16 (nmap 4 x FP16, intensity and gradients FP16; vmaps stay float4)
};

// Scratch floats needed by the reductions (partial rows of the workgroups).
size_t lf_reduce_partials(void);

// Schedule position of one Gauss-Newton step within a tracked frame.
struct LfStepFlags {
    int first_of_level;  // first iteration of a pyramid level (resets the error history)
    int first_of_frame;  // the estimate starts at init (replaces lf_track_begin)
    int last_level;      // a step of the finest level: breaking out of its loop ends the frame
    int last_of_frame;   // the frame's last scheduled step: finish after it
    int context_mode;    // compute_transform(ctx) (icp_tracker.cpp:210-314): NaN errors are not
                         // flagged and the finish always succeeds with update = estimate^-1
    double init[12];     // initial estimate, 3x4 row-major (first_of_frame)
};

// One Gauss-Newton iteration of DenseTracking::compute_transform
// (icp_tracker.cpp:50-140) in two launches: fused ICP + RGB reduction, FP64
// solve, SE3 update, divergence checks and, when the frame ends at this step,
// the finish of lf_track_finish — entirely on the device.
int lf_track_iteration(const LfTrackLevel &lv, float *scratch, LfTrackState *st, const LfStepFlags &f,
                       float icp_weight, hipStream_t stream);
// Start a frame's estimate at T (3x4 doubles, host): est = prev = init = T.
// Only needed when a frame schedules no step.
int lf_track_begin(LfTrackState *st, const double *T12, hipStream_t stream);
// End of a tracked frame: success test, update = est^-1 (or identity), chain the
// odometry pose and refresh pose24.  Only needed when a frame schedules no step.
int lf_track_finish(LfTrackState *st, hipStream_t stream);
// Set the odometry pose directly (first frame / restart): T given as 3x4 doubles (host).
int lf_track_set_pose(LfTrackState *st, const double *T12, hipStream_t stream);

// Batched tracking: one step of several systems in one launch pair.  A job is
// the launch arguments of one system's step (lf_track_job_bytes each);
// lf_track_job_make fills one and reports its grid, lf_track_job_skip marks a
// system without a step at this index.  lf_track_batch_step runs step jobs[0..n)
// (device memory) with a grid of gmax workgroups per system (the largest job
// grid).  A job made with share n takes 1/n of lf_track_iteration's
// workgroups (at least one), so n systems together fill the resident grid
// once; each lane then sums several pixels, and the system's sums equal
// lf_track_iteration's up to FP32 summation order (exactly, for share 1).
// The jobs of one batch share one map format (half_maps as in their LfTrackLevel).
size_t lf_track_job_bytes(void);
// share: systems sharing the launch (the job's grid is 1/share of a lone
// system's); wgs > 0 sets the grid instead.
int lf_track_job_make(const LfTrackLevel &lv, float *scratch, LfTrackState *st, const LfStepFlags &f,
                      float icp_weight, int share, int wgs, void *job, int *grid);
void lf_track_job_skip(void *job);
int lf_track_batch_step(const void *jobs, int n, int gmax, int half_maps, hipStream_t stream);
