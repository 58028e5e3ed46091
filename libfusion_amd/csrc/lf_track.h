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
    int icp_count, rgb_count;
    int active;          // 0 once the current level broke out of its loop
    int invalid;         // an error became NaN this frame
    int success;         // result of the last finished frame
    int iterations;      // Gauss-Newton steps taken this frame
    float last_icp_sums[29], last_rgb_sums[29];
    // odometry chain (DenseOdometry::trackFrame, rgbd_odometry.cpp:18-48)
    double pose[12];     // T_world_camera of the current frame
    float pose24[24];    // float T_world_camera and T_camera_world for the mapping kernels
    int pad[4];
};

// Per-level inputs of one tracking iteration.
struct LfTrackLevel {
    const float *curr_vmap, *curr_nmap, *last_vmap, *last_nmap;
    const float *curr_I, *last_I, *dIdx, *dIdy;
    int cols, rows;
    lf_intrinsics K;
};

// Scratch floats needed by the reductions.
size_t lf_reduce_partials(void);

// One Gauss-Newton iteration of DenseTracking::compute_transform
// (icp_tracker.cpp:50-140): fused ICP + RGB reduction, FP64 solve, SE3 update,
// divergence checks — entirely on the device.
int lf_track_iteration(const LfTrackLevel &lv, float *scratch, LfTrackState *st, int first_iter_of_level,
                       float icp_weight, hipStream_t stream);
// Start a frame's estimate at T (3x4 doubles, host): est = prev = init = T.
int lf_track_begin(LfTrackState *st, const double *T12, hipStream_t stream);
// End of a tracked frame: success test, update = est^-1 (or identity), chain the
// odometry pose and refresh pose24.
int lf_track_finish(LfTrackState *st, hipStream_t stream);
// Set the odometry pose directly (first frame / restart): T given as 3x4 doubles (host).
int lf_track_set_pose(LfTrackState *st, const double *T12, hipStream_t stream);
