// Internal (C++) entry points shared between the kernel files and the host
// classes: the C-ABI functions with optional timing events around the
// dominant kernel of each stage.
#pragma once

#include <hip/hip_runtime.h>

#include "libfusion_amd.h"

// lf_update_weighted / lf_update with events recorded immediately before and
// after the integration kernel (either may be null).
int lf_update_ev(const lf_map_params *p, const lf_map_storage *s, const float *depth, const float *nmap,
                 const unsigned char *rgb, int cols, int rows, const float *pose24, const lf_intrinsics *K,
                 hipStream_t st, bool weighted, hipEvent_t e0, hipEvent_t e1);
// lf_raycast with events around the ray-marching kernel.
int lf_raycast_ev(const lf_map_params *p, const lf_map_storage *s, float *vmap, unsigned char *rgb_out, int cols,
                  int rows, const float *pose24, const lf_intrinsics *K, int check_colour, hipStream_t st,
                  hipEvent_t e0, hipEvent_t e1);
