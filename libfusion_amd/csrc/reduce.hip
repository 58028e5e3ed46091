// ICP and photometric normal-equation reductions and the on-device
// Gauss-Newton step.
//
// Reference: src/tracking/pose_estimator.cu (ICPReduction :166-260,
// icp_reduce :267-307), src/tracking/m_estimator.cu (RgbReduction2 :313-430,
// rgb_step :437-482), include/tracking/reduce_sum.h (WarpReduce / BlockReduce
// :11-46, create_jtjjtr :48-61), src/tracking/icp_tracker.cpp
// (compute_transform :23-156).
//
// MI355X design:
//  * One kernel evaluates both the ICP row and the RGB row of every reference
//    pixel (they share the projected vertex), accumulating 2 x 29 sums per lane.
//  * Wavefront reduction uses 64-lane shuffles, then 4 wavefront partials meet
//    in LDS; each workgroup writes one partial row.
//  * One lane per reference pixel, so a whole pyramid level is in flight at
//    once (the pass is bound by memory latency, not bandwidth); each pixel's
//    loads are issued in two rounds.
//  * A tracking iteration is two launches: the partial rows, then one
//    workgroup that sums them in a fixed order (deterministic), after which
//    one lane builds A = 1e6 H_icp + H_rgb, solves it with an FP64 LDL^T,
//    applies T <- exp(dx) T in FP64 and runs the reference's divergence /
//    revert logic; the first step of a frame starts the estimate and the step
//    that ends the frame runs the finish.  The next iteration reads the
//    updated pose from device memory: a whole frame of tracking is a fixed
//    sequence of launches with no host round trip (the reference downloads
//    the sums and solves on the host twice per iteration).
//  * The stand-alone reductions of the C-ABI (lf_icp_reduce, lf_rgb_step) are
//    a partial-sum launch and a one-workgroup final-sum launch.
#include <cfloat>
#include <climits>
#include <mutex>

#include "lf_common.h"
#include "lf_track.h"

using namespace lf;

namespace {

// Maximum workgroups of a reduction: the workgroups resident at once on
// MI355X (256 CUs x 2 workgroups of 512 lanes at the step kernel's register
// count).  A larger level loops over pixels with the next pixel's loads in
// flight, rather than waiting for a second round of workgroups.
constexpr int kGrid = 512;
constexpr int kThreads = 512;  // lanes per workgroup (8 wavefronts)
constexpr int kWaves = kThreads / 64;
constexpr int kVals = 58;      // 29 ICP + 29 RGB sums

struct ReduceArgs {
    LfTrackLevel lv;
    const float *pose12;   // float 3x4 (stand-alone reductions)
    float stddev;          // stand-alone RGB reduction
    float angle, dthr;
    float *partials;       // kGrid * kVals floats
    float *out29;          // modes 1/2
    LfTrackState *st;      // tracking step
    int first_iter;        // first iteration of a level
    float icp_weight;
    int first_of_frame;    // tracking step: the estimate starts at init
    int last_level;        // tracking step: the finest level
    int last_of_frame;     // tracking step: the frame's last scheduled step
    int context_mode;      // tracking step: compute_transform(ctx) semantics
    double init[12];       // tracking step: initial estimate (first_of_frame)
    float init_f[12];      // the same in FP32 (the first step's pose, uniform across the grid)
};

// create_jtjjtr<6,7> plus {r^2, count} of one row, summed in FP32: the
// reference's sums (the stand-alone reductions lf_icp_reduce / lf_rgb_step).
LF_DEV void accumulate(float *s, const float *row) {
    int k = 0;
#pragma unroll
    for (int i = 0; i < 7; ++i)
#pragma unroll
        for (int j = i; j < 7; ++j) s[k++] += row[i] * row[j];
    s[28] += 1.0f;
}

struct FloatSums {  // the 29 ICP and 29 RGB sums of the stand-alone reductions
    float si[29], sr[29];
    LF_DEV void clear() {
#pragma unroll
        for (int k = 0; k < 29; ++k) { si[k] = 0.0f; sr[k] = 0.0f; }
    }
    LF_DEV void icp(const float *row) { accumulate(si, row); }
    LF_DEV void rgb(const float *row) { accumulate(sr, row); }
};

// The tracking step's sums: the normal equations A = w H_icp + H_rgb,
// b = w g_icp + g_rgb of compute_transform (icp_tracker.cpp:89-96) combined
// per pixel and accumulated in FP64, plus each term's {r^2, count}.  A
// product of two FP32 row values is exact in FP64, the ICP weight costs one
// rounding at 2^-53 and every fused multiply-add one more, so a lane's serial
// sum and the tree above it stay within about 1e-13 of the exact sum of the
// FP32 rows (relative to the sum of magnitudes), where FP32 sums (the
// reference's, summed per term and combined in FP32) carry about 1e-6.
// Layout (kTrackVals): 0-26 the create_jtjjtr<6,7> order of A and b, 27 / 28
// ICP r^2 / count, 29 / 30 RGB r^2 / count.
constexpr int kTrackVals = 31;
struct CombinedSums {
    double s[32];  // kTrackVals, padded to a power of two for the wavefront transpose
    double w;      // the ICP weight
    unsigned n_icp, n_rgb;  // the counts, moved into s[28] / s[30] for the reduction
    LF_DEV void clear() {
#pragma unroll
        for (int k = 0; k < 32; ++k) s[k] = 0.0;
        n_icp = n_rgb = 0u;
    }
    LF_DEV void icp(const float *row) {
        double d[7], wd[6];
#pragma unroll
        for (int i = 0; i < 7; ++i) d[i] = (double)row[i];
#pragma unroll
        for (int i = 0; i < 6; ++i) wd[i] = w * d[i];
        int k = 0;
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = i; j < 7; ++j) s[k] = __builtin_fma(wd[i], d[j], s[k]), ++k;
        s[27] = __builtin_fma(d[6], d[6], s[27]);
        ++n_icp;
    }
    LF_DEV void rgb(const float *row) {
        double d[7];
#pragma unroll
        for (int i = 0; i < 7; ++i) d[i] = (double)row[i];
        int k = 0;
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = i; j < 7; ++j) s[k] = __builtin_fma(d[i], d[j], s[k]), ++k;
        s[29] = __builtin_fma(d[6], d[6], s[29]);
        ++n_rgb;
    }
};

// Both rows of one reference pixel — ICPReduction::searchPoint + getRow
// (pose_estimator.cu:168-222) and RgbReduction2::find_corresp +
// compute_jacobian with interp2 (m_estimator.cu:315-391) — with all loads in
// two rounds: the pixel's own vertex / normal / intensity first, then every
// gather (the ICP correspondence and the 12 bilinear taps of the photometric
// term) together.  The arithmetic follows the reference's operation order.
// A reference pixel's own inputs (its vertex, normal and intensity): they do
// not depend on the pose, so they are loaded before the pose is known.
struct Own {
    float4 lv, ln;
    float il;
};

// Element i of a map of float4 or of 4 x FP16 (kHalf), and of a plane of
// float or FP16: the mixed-precision maps widen to FP32 as they are loaded.
template <bool kHalf>
LF_DEV float4 map4(const float *p, size_t i) {
    if constexpr (kHalf) {
        const uint2 u = reinterpret_cast<const uint2 *>(p)[i];
        return make_float4(half_bits_to_float((unsigned short)u.x), half_bits_to_float((unsigned short)(u.x >> 16)),
                           half_bits_to_float((unsigned short)u.y), half_bits_to_float((unsigned short)(u.y >> 16)));
    } else {
        return reinterpret_cast<const float4 *>(p)[i];
    }
}
template <bool kHalf>
LF_DEV float plane1(const float *p, size_t i) {
    if constexpr (kHalf) return half_bits_to_float(reinterpret_cast<const unsigned short *>(p)[i]);
    else return p[i];
}

template <bool kIcp, bool kRgb, bool kHalf = false>
LF_DEV Own load_own(const LfTrackLevel &L, int i) {
    Own o;
    o.lv = map4<false>(L.last_vmap, i);  // vertices are FP32 in both precisions
    o.ln = make_float4(0.f, 0.f, 0.f, 0.f);
    o.il = 0.0f;
    if (kIcp) o.ln = map4<kHalf>(L.last_nmap, i);
    if (kRgb) o.il = plane1<kHalf>(L.last_I, i);
    return o;
}

// kHalf (mixed precision) also takes one hardware reciprocal of the depth for
// every projection and the hardware square root, where the FP32 path keeps
// the reference's correctly rounded divisions and roots.
LF_DEV float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

template <bool kIcp, bool kRgb, bool kHalf, class Sums>
LF_DEV void pixel_rows(const ReduceArgs &a, const Own &o, const float *P, float stddev, float th, Sums &sums) {
    const LfTrackLevel &L = a.lv;
    const float4 lv = o.lv, ln = o.ln;
    const float il = o.il;
    const f3 vg = xform(P, f3{lv.x, lv.y, lv.z});
    const float iz = kHalf ? __builtin_amdgcn_rcpf(vg.z) : 0.0f;
    // ICP correspondence (pose_estimator.cu:168-193)
    bool icp_ok = false;
    size_t j = 0;
    if (kIcp) {
        const float invz = kHalf ? iz : 1.0f / vg.z;
        const int u = proj_floor(vg.x * invz * L.K.fx + L.K.cx + 0.5f);
        const int v = proj_floor(vg.y * invz * L.K.fy + L.K.cy + 0.5f);
        icp_ok = !isnan(lv.x) && u >= 0 && v >= 0 && u < L.cols && v < L.rows;
        if (icp_ok) j = (size_t)v * L.cols + u;
    }
    // photometric correspondence (m_estimator.cu:315-338)
    bool rgb_ok = false;
    float u0 = 0.0f, v0 = 0.0f;
    size_t t0 = 0;
    if (kRgb) {
        u0 = (kHalf ? vg.x * iz : vg.x / vg.z) * L.K.fx + L.K.cx;
        v0 = (kHalf ? vg.y * iz : vg.y / vg.z) * L.K.fy + L.K.cy;
        rgb_ok = !(lv.w < 0.0f || isnan(lv.x)) && isfinite(il) && u0 >= 2.0f && u0 < (float)(L.cols - 2) &&
                 v0 >= 2.0f && v0 < (float)(L.rows - 2);
        if (rgb_ok) t0 = (size_t)(int)floorf(v0) * L.cols + (int)floorf(u0);
    }
    // gathers
    float4 cv = make_float4(0.f, 0.f, 0.f, 0.f), cn = cv;
    if (kIcp) {
        cv = map4<false>(L.curr_vmap, j);
        cn = map4<kHalf>(L.curr_nmap, j);
    }
    float tI[4], tx[4], ty[4];
    if (kRgb) {
        const size_t o[4] = {t0, t0 + 1, t0 + L.cols, t0 + L.cols + 1};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            tI[k] = plane1<kHalf>(L.curr_I, o[k]);
            tx[k] = plane1<kHalf>(L.dIdx, o[k]);
            ty[k] = plane1<kHalf>(L.dIdy, o[k]);
        }
    }
    float row[7];
    if (kIcp && icp_ok) {
        const f3 vc = {cv.x, cv.y, cv.z};
        const f3 ng = rotate(P, f3{ln.x, ln.y, ln.z});
        const f3 nc = {cn.x, cn.y, cn.z};
        const f3 dv = sub(vg, vc), cn3 = cross(nc, ng);
        const float dist = kHalf ? fast_sqrt(dot(dv, dv)) : norm(dv);
        const float sine = kHalf ? fast_sqrt(dot(cn3, cn3)) : norm(cn3);
        if (sine < a.angle && dist <= a.dthr && !isnan(nc.x) && !isnan(ng.x)) {
            const f3 c = cross(vg, ng);
            row[0] = ng.x; row[1] = ng.y; row[2] = ng.z;
            row[3] = c.x; row[4] = c.y; row[5] = c.z;
            row[6] = dot(ng, sub(vc, vg));
            sums.icp(row);
        }
    }
    if (kRgb && rgb_ok) {
        const float cx = u0 - floorf(u0), cy = v0 - floorf(v0);
        const float ic = (tI[0] * (1.0f - cx) + tI[1] * cx) * (1.0f - cy) + (tI[2] * (1.0f - cx) + tI[3] * cx) * cy;
        const float dx = (tx[0] * (1.0f - cx) + tx[1] * cx) * (1.0f - cy) + (tx[2] * (1.0f - cx) + tx[3] * cx) * cy;
        const float dy = (ty[0] * (1.0f - cx) + ty[1] * cx) * (1.0f - cy) + (ty[2] * (1.0f - cx) + ty[3] * cx) * cy;
        if ((dx > 2.0f || dy > 2.0f) && isfinite(ic) && isfinite(dx) && isfinite(dy)) {
            const float zi = kHalf ? iz : 1.0f / vg.z;
            f3 l;
            l.x = dx * L.K.fx * zi;
            l.y = dy * L.K.fy * zi;
            l.z = -(l.x * vg.x + l.y * vg.y) * zi;
            float r = ic - il;
            const double sd = (double)stddev;
            if (sd > 10e-5) r = kHalf ? r * __builtin_amdgcn_rcpf(stddev) : r / stddev;
            float w = 1.0f;
            if (fabsf(r) > th && sd > 10e-6)
                w = kHalf ? fast_sqrt(th * __builtin_amdgcn_rcpf(fabsf(r))) : sqrtf(th / fabsf(r));
            const f3 c = cross(vg, l);
            row[6] = w * (-r);
            row[0] = l.x * w; row[1] = l.y * w; row[2] = l.z * w;
            row[3] = c.x * w; row[4] = c.y * w; row[5] = c.z * w;
            sums.rgb(row);
        }
    }
}

// ---------------------------------------------------------- FP64 SE(3) helpers
// The FP64 helpers let the compiler fuse multiply-adds (one rounding instead
// of two, half the instructions of the serial step); the FP32 row arithmetic
// above keeps the reference's separate roundings.
LF_DEV void m34_identity(double *T) {
    for (int i = 0; i < 12; ++i) T[i] = (i == 0 || i == 5 || i == 10) ? 1.0 : 0.0;
}
LF_DEV void m34_mul(const double *A, const double *B, double *C) {
#pragma clang fp contract(fast)
    double t[12];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) t[i * 4 + j] = A[i * 4] * B[j] + A[i * 4 + 1] * B[4 + j] + A[i * 4 + 2] * B[8 + j];
        t[i * 4 + 3] = A[i * 4] * B[3] + A[i * 4 + 1] * B[7] + A[i * 4 + 2] * B[11] + A[i * 4 + 3];
    }
    for (int i = 0; i < 12; ++i) C[i] = t[i];
}
LF_DEV void m34_inv(const double *A, double *C) {
#pragma clang fp contract(fast)
    double t[12];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) t[i * 4 + j] = A[j * 4 + i];
    for (int i = 0; i < 3; ++i) t[i * 4 + 3] = -(t[i * 4] * A[3] + t[i * 4 + 1] * A[7] + t[i * 4 + 2] * A[11]);
    for (int i = 0; i < 12; ++i) C[i] = t[i];
}
// SE(3) exponential, tangent [upsilon; omega] (Sophus order); Barfoot sec. 7.1.
// The coefficients A = sin t / t, B = (1 - cos t) / t^2, C = (t - sin t) / t^3:
// below t = 0.02 (every step of a converging track) their Maclaurin series to
// the t^8 term, whose remainder is below 1e-20 relative there and which avoids
// the cancellation of 1 - cos t; above it one sincos and one division.
LF_DEV void se3_exp(const double *xi, double *T) {
#pragma clang fp contract(fast)
    const double *u = xi, *w = xi + 3;
    const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    double W2[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) W2[i * 3 + j] = W[i * 3] * W[j] + W[i * 3 + 1] * W[3 + j] + W[i * 3 + 2] * W[6 + j];
    double A, B, C;
    if (th2 < 4e-4) {
        const double t = th2;
        A = 1.0 - t * (1.0 / 6) * (1.0 - t * (1.0 / 20) * (1.0 - t * (1.0 / 42) * (1.0 - t * (1.0 / 72))));
        B = 0.5 * (1.0 - t * (1.0 / 12) * (1.0 - t * (1.0 / 30) * (1.0 - t * (1.0 / 56) * (1.0 - t * (1.0 / 90)))));
        C = (1.0 / 6) * (1.0 - t * (1.0 / 20) * (1.0 - t * (1.0 / 42) * (1.0 - t * (1.0 / 72) * (1.0 - t * (1.0 / 110)))));
    } else {
        const double th = sqrt(th2), r = 1.0 / th;
        double s, c;
        sincos(th, &s, &c);
        A = s * r;
        B = (1.0 - c) * r * r;
        C = (th - s) * r * r * r;
    }
    double V[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double I = i == j ? 1.0 : 0.0;
            T[i * 4 + j] = I + A * W[i * 3 + j] + B * W2[i * 3 + j];
            V[i * 3 + j] = I + B * W[i * 3 + j] + C * W2[i * 3 + j];
        }
    for (int i = 0; i < 3; ++i) T[i * 4 + 3] = V[i * 3] * u[0] + V[i * 3 + 1] * u[1] + V[i * 3 + 2] * u[2];
}
// Norm of the SE(3) logarithm (Barfoot sec. 7.1.3).
LF_DEV double se3_log_norm(const double *T) {
#pragma clang fp contract(fast)
    double c = (T[0] + T[5] + T[10] - 1) * 0.5;
    c = fmax(-1.0, fmin(1.0, c));
    double th = acos(c);
    const double v[3] = {T[9] - T[6], T[2] - T[8], T[4] - T[1]};
    double w[3];
    if (th < 1e-8) {
        for (int i = 0; i < 3; ++i) w[i] = 0.5 * v[i];
    } else if (M_PI - th < 1e-6) {
        // k = index of the largest diagonal entry (first one on ties); the
        // row of R is selected without a runtime array index
        const int k1 = T[5] > T[0] ? 1 : 0;
        const int k = T[10] > (k1 ? T[5] : T[0]) ? 2 : k1;
        const double r0 = k == 0 ? T[0] : k == 1 ? T[4] : T[8];
        const double r1 = k == 0 ? T[1] : k == 1 ? T[5] : T[9];
        const double r2 = k == 0 ? T[2] : k == 1 ? T[6] : T[10];
        const double rkk = k == 0 ? r0 : k == 1 ? r1 : r2;
        const double axk = sqrt(fmax(0.0, (rkk + 1) * 0.5));
        const double ax0 = k == 0 ? axk : r0 / (2 * axk);
        const double ax1 = k == 1 ? axk : r1 / (2 * axk);
        const double ax2 = k == 2 ? axk : r2 / (2 * axk);
        w[0] = ax0 * th;
        w[1] = ax1 * th;
        w[2] = ax2 * th;
    } else {
        const double s = th / (2 * sin(th));
        for (int i = 0; i < 3; ++i) w[i] = s * v[i];
    }
    const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    th = sqrt(th2);
    const double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    double W2[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) W2[i * 3 + j] = W[i * 3] * W[j] + W[i * 3 + 1] * W[3 + j] + W[i * 3 + 2] * W[6 + j];
    const double D = th < 1e-8 ? 1.0 / 12 : (1 - th * sin(th) / (2 * (1 - cos(th)))) / th2;
    double n2 = th2;
    for (int i = 0; i < 3; ++i) {
        double s = 0;
        for (int j = 0; j < 3; ++j) {
            const double Vi = (i == j ? 1.0 : 0.0) - 0.5 * W[i * 3 + j] + D * W2[i * 3 + j];
            s += Vi * T[j * 4 + 3];
        }
        n2 += s * s;
    }
    return sqrt(n2);
}
// LDL^T solve of a symmetric positive definite 6x6 (no pivoting); one
// division per pivot (its reciprocal scales the column and the solution).
LF_DEV void ldlt6(const double *A, const double *b, double *x) {
#pragma clang fp contract(fast)
    double L[36], D[6], Dinv[6];
#pragma unroll
    for (int i = 0; i < 36; ++i) L[i] = 0.0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        // row j of L D, shared by the updates of the rows below it
        double LD[6];
        double d = A[j * 6 + j];
#pragma unroll
        for (int k = 0; k < j; ++k) {
            LD[k] = L[j * 6 + k] * D[k];
            d -= L[j * 6 + k] * LD[k];
        }
        D[j] = d;
        Dinv[j] = d != 0.0 ? 1.0 / d : 0.0;
        L[j * 6 + j] = 1.0;
#pragma unroll
        for (int i = j + 1; i < 6; ++i) {
            double s = A[i * 6 + j];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= L[i * 6 + k] * LD[k];
            L[i * 6 + j] = s * Dinv[j];
        }
    }
    double y[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double s = b[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s -= L[i * 6 + k] * y[k];
        y[i] = s;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) y[i] *= Dinv[i];
#pragma unroll
    for (int i = 5; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int k = i + 1; k < 6; ++k) s -= L[k * 6 + i] * x[k];
        x[i] = s;
    }
}
// create_jtjjtr<6,7> (reduce_sum.h:48-61) of the combined sums: A and b.
LF_DEV void unpack27(const double *s, double *A, double *b) {
    int k = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j) {
            const double v = s[k++];
            if (j == 6) b[i] = v;
            else A[j * 6 + i] = A[i * 6 + j] = v;
        }
}

// What a Gauss-Newton step reads of the tracker state.  k_track_final loads
// it at entry, together with the flags that decide whether the step runs, so
// the serial step starts from registers rather than from dependent loads.
struct GnState {
    double est[12];
    float icp_error, rgb_error;
    int icp_count, rgb_count, active, invalid, iterations;
};

LF_DEV GnState load_gn(const LfTrackState &st) {
    GnState g;
#pragma unroll
    for (int k = 0; k < 12; ++k) g.est[k] = st.est[k];
    g.icp_error = st.icp_error;
    g.rgb_error = st.rgb_error;
    g.icp_count = st.icp_count;
    g.rgb_count = st.rgb_count;
    g.active = st.active;
    g.invalid = st.invalid;
    g.iterations = st.iterations;
    return g;
}

// One Gauss-Newton step of compute_transform (icp_tracker.cpp:50-139), run by
// a single lane after the sums are final: reads g, updates g and writes the
// step's results to st.
LF_DEV void gn_step(LfTrackState &st, GnState &g, const double *sums, int first_iter, bool context_mode) {
    if (first_iter) {
        g.icp_error = FLT_MAX;
        g.rgb_error = FLT_MAX;
        g.icp_count = g.rgb_count = 0;
        g.active = 1;
    }
    double last[12];
    for (int k = 0; k < 12; ++k) last[k] = g.est[k];
    const float last_icp = g.icp_error, last_rgb = g.rgb_error;
    st.last_icp_error = last_icp;
    // the residual terms in FP32, as the reference's (m_estimator.cu:470-480)
    const float icp_r2 = (float)sums[27], icp_n = (float)sums[28];
    const float rgb_r2 = (float)sums[29], rgb_n = (float)sums[30];
    st.stddev = sqrtf(rgb_r2 / (rgb_n - 6.0f));
    double A[36], b[6], x[6];
    unpack27(sums, A, b);
    ldlt6(A, b, x);
    double E[12];
    se3_exp(x, E);
    m34_mul(E, last, g.est);
    g.iterations++;
    bool stop = false, revert = false;
    float icp_error = sqrtf(icp_r2) / icp_n;
    if (isnan(icp_error) && !context_mode) {
        g.invalid = 1;
        stop = true;
    } else if (icp_error > last_icp) {
        if (g.icp_count >= 2) {
            revert = stop = true;
        } else {
            g.icp_count++;
            icp_error = last_icp;
        }
    } else {
        g.icp_count = 0;
    }
    g.icp_error = icp_error;
    if (!stop) {
        float rgb_error = sqrtf(rgb_r2) / rgb_n;
        if (isnan(rgb_error) && !context_mode) {
            g.invalid = 1;
            stop = true;
        } else if (rgb_error > last_rgb) {
            if (g.rgb_count >= 2) {
                revert = stop = true;
            } else {
                g.rgb_count++;
                rgb_error = last_rgb;
            }
        } else {
            g.rgb_count = 0;
        }
        g.rgb_error = rgb_error;
    }
    // Revertable<SE3d>::revert: the estimate goes back to the one before the step
    if (revert)
        for (int k = 0; k < 12; ++k) g.est[k] = last[k];
    g.active = stop ? 0 : 1;
    for (int k = 0; k < 12; ++k) {
        st.prev[k] = last[k];
        st.est[k] = g.est[k];
        st.pose_f[k] = (float)g.est[k];
    }
    st.icp_error = g.icp_error;
    st.rgb_error = g.rgb_error;
    st.icp_count = g.icp_count;
    st.rgb_count = g.rgb_count;
    st.active = g.active;
    st.invalid = g.invalid;
    st.iterations = g.iterations;
}

// Wavefront sum of N values per lane (N a power of two, <= 64) as a
// transpose butterfly: at each exchange a lane keeps half of its values and
// sends the other half to its partner, so N values cost N - 1 + 6 - log2(N)
// shuffles instead of 6 N.  Afterwards lane l holds the wavefront total of
// value l >> (6 - log2 N).  Fixed pairing: deterministic.
// (Compile-time recursion keeps every index constant, so v stays in VGPRs.)
template <int N, int kOff, class T>
LF_DEV void transpose_halve(T *v, int lane) {
    if constexpr (N > 1) {
        constexpr int h = N / 2;
        const bool hi = (lane & kOff) != 0;
#pragma unroll
        for (int i = 0; i < h; ++i) {
            const T keep = hi ? v[h + i] : v[i];
            const T send = hi ? v[i] : v[h + i];
            v[i] = keep + __shfl_xor(send, kOff);
        }
        transpose_halve<h, kOff / 2>(v, lane);
    } else {
#pragma unroll
        for (int o = kOff; o >= 1; o >>= 1) v[0] += __shfl_xor(v[0], o);
    }
}

template <int N, class T>
LF_DEV T transpose_reduce(T (&v)[N]) {
    static_assert(N >= 1 && N <= 64 && (N & (N - 1)) == 0, "N must be a power of two <= 64");
    transpose_halve<N, 32>(v, __lane_id());
    return v[0];
}

// Partial sums of one workgroup (ICP rows into si, RGB rows into sr) into
// value order: transpose butterfly in each wavefront, then the wavefronts meet
// in LDS.  Lanes with tid < kVals return value tid (the others return 0).
template <bool kIcp, bool kRgb>
LF_DEV float block_sums(const FloatSums &sums, float (*s_wave)[kVals]) {
    const float *si = sums.si, *sr = sums.sr;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float v[64];
#pragma unroll
    for (int k = 0; k < 29; ++k) {
        v[k] = kIcp ? si[k] : 0.0f;
        v[29 + k] = kRgb ? sr[k] : 0.0f;
    }
#pragma unroll
    for (int k = kVals; k < 64; ++k) v[k] = 0.0f;
    const float w = transpose_reduce<64>(v);
    if (lane < kVals) s_wave[wave][lane] = w;
    __syncthreads();
    float t = 0.0f;
    if (tid < kVals) {
#pragma unroll
        for (int k = 0; k < kWaves; ++k) t += s_wave[k][tid];
    }
    return t;
}

// Final sum of the G partial rows in a fixed order (deterministic): wavefront w
// reduces values w, w+kWaves, ...; lane l adds partials l, l+64, ... of a value
// (independent loads, in flight together), then a transpose butterfly over the
// wavefront's values.  The kVals totals land in s_sum; the caller
// synchronises before reading them.
LF_DEV void final_sums(const float *partials, int G, float *s_sum) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int kPer = 8;
    static_assert(kPer * kWaves >= kVals, "values per wavefront");
    constexpr int kChunks = kGrid / 64;
    float acc[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int v = wave + kWaves * j;
        const float *src = partials + (size_t)(v < kVals ? v : 0) * G;
        // every load of the value first, then the fixed-order sum.  Indices
        // past G are clamped and their values masked to +0 with a bitwise AND
        // (a select of a loaded value becomes a branch around the load with a
        // wait inside it), so all the loads are in flight together.
        float x[kChunks];
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            const int g = c * 64 + lane;
            const float t = src[g < G ? g : G - 1];
            const unsigned keep = (v < kVals && g < G) ? 0xffffffffu : 0u;
            x[c] = __uint_as_float(__float_as_uint(t) & keep);
        }
        float s = 0.0f;
#pragma unroll
        for (int c = 0; c < kChunks; ++c) s += x[c];
        acc[j] = s;
    }
    const float s = transpose_reduce<kPer>(acc);
    const int v = wave + kWaves * (lane >> 3);
    if ((lane & 7) == 0 && v < kVals) s_sum[v] = s;
}

// The combined FP64 sums of one workgroup: transpose butterfly over the 32
// (padded) values in each wavefront, leaving value l >> 1 in lane l, then
// the wavefronts meet in LDS.  Lanes with tid < kTrackVals return value tid.
LF_DEV double block_sums_d(CombinedSums &sums, double (*s_wave)[32]) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    sums.s[28] = (double)sums.n_icp;
    sums.s[30] = (double)sums.n_rgb;
    const double w = transpose_reduce<32>(sums.s);
    if ((lane & 1) == 0) s_wave[wave][lane >> 1] = w;
    __syncthreads();
    double t = 0.0;
    if (tid < kTrackVals) {
#pragma unroll
        for (int k = 0; k < kWaves; ++k) t += s_wave[k][tid];
    }
    return t;
}

// final_sums of the G combined FP64 partial rows (value v of workgroup g at
// v * G + g), in a fixed order; the kTrackVals totals land in s_sum.  kChunks
// 64-lane chunks of partials per value, enough for G (the chunks past G hold
// only masked zeros, which leave the sums unchanged: a sum started at +0 is
// never -0).
template <int kChunks>
LF_DEV void final_sums_d_chunks(const double *partials, int G, double *s_sum) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int kPer = 4;
    static_assert(kPer * kWaves >= kTrackVals, "values per wavefront");
    double acc[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int v = wave + kWaves * j;
        const double *src = partials + (size_t)(v < kTrackVals ? v : 0) * G;
        double x[kChunks];
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            const int g = c * 64 + lane;
            const double t = src[g < G ? g : G - 1];
            const unsigned long long keep = (v < kTrackVals && g < G) ? ~0ull : 0ull;
            x[c] = __longlong_as_double((long long)((unsigned long long)__double_as_longlong(t) & keep));
        }
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < kChunks; ++c) s += x[c];
        acc[j] = s;
    }
    const double s = transpose_reduce<kPer>(acc);
    const int v = wave + kWaves * (lane >> 4);
    if ((lane & 15) == 0 && v < kTrackVals) s_sum[v] = s;
}
// (a batch of 16 systems: 32 partials per value, one chunk)
LF_DEV void final_sums_d(const double *partials, int G, double *s_sum) {
    if (G <= 64) final_sums_d_chunks<1>(partials, G, s_sum);
    else final_sums_d_chunks<kGrid / 64>(partials, G, s_sum);
}

// Stand-alone reductions (lf_icp_reduce, lf_rgb_step).  kMode 1: ICP rows,
// 2: RGB rows.  Each workgroup writes its kVals partials value-major (value v
// of workgroup g at v * G + g).
template <int kMode>
__global__ __launch_bounds__(kThreads) void k_reduce(ReduceArgs a) {
    __shared__ float s_wave[kWaves][kVals];
    constexpr bool kIcp = kMode == 1, kRgb = kMode == 2;
    const int tid = threadIdx.x;
    float P[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) P[k] = a.pose12[k];
    const float stddev = a.stddev;
    const float th = (float)(1.345 * (double)stddev);
    FloatSums sums;
    sums.clear();
    const int n = a.lv.cols * a.lv.rows;
    for (int i = blockIdx.x * kThreads + tid; i < n; i += gridDim.x * kThreads)
        pixel_rows<kIcp, kRgb, false>(a, load_own<kIcp, kRgb>(a.lv, i), P, stddev, th, sums);
    const float t = block_sums<kIcp, kRgb>(sums, s_wave);
    if (tid < kVals) a.partials[(size_t)tid * gridDim.x + blockIdx.x] = t;
}

template <int kMode>
__global__ __launch_bounds__(kThreads) void k_reduce_final(ReduceArgs a, int G) {
    __shared__ float s_sum[kVals];
    final_sums(a.partials, G, s_sum);
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < 29; ++k) a.out29[k] = s_sum[(kMode == 2 ? 29 : 0) + k];
}

// End of DenseTracking::compute_transform (icp_tracker.cpp:143-155) and the
// odometry chain of DenseOdometry::trackFrame (rgbd_odometry.cpp:39): success
// test, update = est^-1 (or identity), pose = pose * update, pose24 refreshed.
// est and invalid are the frame's final estimate and flag (already in st).
LF_DEV void track_finish(LfTrackState &st, const double *est, int invalid, bool context_mode) {
    double inv_est[12], D[12];
    m34_inv(est, inv_est);
    m34_mul(inv_est, st.init, D);
    // compute_transform(ctx) reports success unconditionally (icp_tracker.cpp:308-313)
    const bool ok = context_mode || (!invalid && !(se3_log_norm(D) > 0.1));
    st.success = ok ? 1 : 0;
    double upd[12];
    if (ok) for (int k = 0; k < 12; ++k) upd[k] = inv_est[k];
    else m34_identity(upd);
    double P[12];
    m34_mul(st.pose, upd, P);
    double Pi[12];
    m34_inv(P, Pi);
    for (int k = 0; k < 12; ++k) {
        st.pose[k] = P[k];
        st.pose24[k] = (float)P[k];
        st.pose24[12 + k] = (float)Pi[k];
    }
}

// One Gauss-Newton iteration of compute_transform, first launch: every
// workgroup sums the ICP and RGB rows of its pixels into one partial row.  The
// pose-independent loads of the first pixel are issued together with the
// tracker-state loads.
// Workgroup wg of the nwg computing the step's partial rows.  kHalf: FP16 maps.
template <bool kHalf>
LF_DEV void track_step_wg(const ReduceArgs &a, int wg, int nwg, double (*s_wave)[32]) {
    const int tid = threadIdx.x;
    const int n = a.lv.cols * a.lv.rows;
    const int stride = nwg * kThreads;
    int i = wg * kThreads + tid;
    // the pixel's own loads and the state loads are in flight together
    Own o = {};
    if (i < n) o = load_own<true, true, kHalf>(a.lv, i);
    float P[12], stddev;
    if (a.first_of_frame) {
#pragma unroll
        for (int k = 0; k < 12; ++k) P[k] = a.init_f[k];
        stddev = 0.0f;
    } else {
        const LfTrackState &st = *a.st;
        const bool run = !st.invalid && (a.first_iter || st.active);
#pragma unroll
        for (int k = 0; k < 12; ++k) P[k] = st.pose_f[k];
        stddev = a.first_iter ? 0.0f : st.stddev;
        // a failed frame, or a level that broke out of its loop: nothing to do
        if (!run) return;
    }
    const float th = (float)(1.345 * (double)stddev);
    CombinedSums sums;
    sums.clear();
    sums.w = (double)a.icp_weight;
    while (i < n) {
        const Own cur = o;
        const int next = i + stride;
        if (next < n) o = load_own<true, true, kHalf>(a.lv, next);
        pixel_rows<true, true, kHalf>(a, cur, P, stddev, th, sums);
        i = next;
    }
    const double t = block_sums_d(sums, s_wave);
    if (tid < kTrackVals) reinterpret_cast<double *>(a.partials)[(size_t)tid * nwg + wg] = t;
}

// Held to 128 VGPRs: 4 waves per SIMD, two 512-lane workgroups per CU.
template <bool kHalf>
__global__ __launch_bounds__(kThreads, 4) void k_track_step(ReduceArgs a) {
    __shared__ double s_wave[kWaves][32];
    track_step_wg<kHalf>(a, blockIdx.x, gridDim.x, s_wave);
}

// Second launch (one workgroup): the partial rows summed in a fixed order,
// then the FP64 step; the first launch of a frame starts the estimate at init
// and the launch that ends the frame runs the finish.  (A single launch with a
// last-workgroup ticket needs an agent-scope release per workgroup, an L2
// write-back on each of the 8 XCDs per workgroup; measured slower than the
// kernel boundary.)
LF_DEV void track_final_wg(const ReduceArgs &a, int G, double *s_sum) {
    const int tid = threadIdx.x;
    GnState g = load_gn(*a.st);
    if (a.first_of_frame) {
#pragma unroll
        for (int k = 0; k < 12; ++k) g.est[k] = a.init[k];
        g.invalid = 0;
        g.iterations = 0;
    } else if (!(!g.invalid && (a.first_iter || g.active))) {
        return;
    }
    final_sums_d(reinterpret_cast<const double *>(a.partials), G, s_sum);
    __syncthreads();
    if (tid < kTrackVals) a.st->sums[tid] = s_sum[tid];
    if (tid == 0) {
        LfTrackState &st = *a.st;
        if (a.first_of_frame)
            for (int k = 0; k < 12; ++k) st.init[k] = a.init[k];
        gn_step(st, g, s_sum, a.first_iter, a.context_mode != 0);
        // the frame ends here when the estimate became invalid, when the
        // finest level broke out of its loop, or after the frame's last step
        if (g.invalid || (a.last_level && !g.active) || a.last_of_frame)
            track_finish(st, g.est, g.invalid, a.context_mode != 0);
    }
}

__global__ __launch_bounds__(kThreads) void k_track_final(ReduceArgs a, int G) {
    __shared__ double s_sum[32];
    track_final_wg(a, G, s_sum);
}

// Launch arguments read from a descriptor table, their pointers asserted global.
LF_DEV ReduceArgs global_args(const ReduceArgs &d) {
    ReduceArgs a = d;
    LfTrackLevel &L = a.lv;
    L.curr_vmap = global_ptr(d.lv.curr_vmap); L.curr_nmap = global_ptr(d.lv.curr_nmap);
    L.last_vmap = global_ptr(d.lv.last_vmap); L.last_nmap = global_ptr(d.lv.last_nmap);
    L.curr_I = global_ptr(d.lv.curr_I); L.last_I = global_ptr(d.lv.last_I);
    L.dIdx = global_ptr(d.lv.dIdx); L.dIdy = global_ptr(d.lv.dIdy);
    a.pose12 = global_ptr(d.pose12); a.partials = global_ptr(d.partials); a.out29 = global_ptr(d.out29);
    a.st = global_ptr(d.st);
    return a;
}

// One system's step of a batch: its launch arguments and its grid (0: the
// system has no step at this index).
struct TrackJob {
    ReduceArgs a;
    int G;
};

// Step s of several systems (each its own tracker state, scratch and
// pyramids): a 2-D grid whose y index is the system; each system's workgroups
// run exactly the single launch's code.
template <bool kHalf>
__global__ __launch_bounds__(kThreads, 4) void k_track_step_batch(const TrackJob *__restrict__ jobs) {
    __shared__ double s_wave[kWaves][32];
    const TrackJob &j = jobs[blockIdx.y];
    if ((int)blockIdx.x >= j.G) return;
    track_step_wg<kHalf>(global_args(j.a), blockIdx.x, j.G, s_wave);
}

// The final sums and Gauss-Newton step of each system: workgroup i, system i.
__global__ __launch_bounds__(kThreads) void k_track_final_batch(const TrackJob *__restrict__ jobs) {
    __shared__ double s_sum[32];
    const TrackJob &j = jobs[blockIdx.x];
    if (j.G <= 0) return;
    track_final_wg(global_args(j.a), j.G, s_sum);
}

__global__ void k_track_finish(LfTrackState *st) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    LfTrackState &s = *st;
    double est[12];
    for (int k = 0; k < 12; ++k) est[k] = s.est[k];
    track_finish(s, est, s.invalid, false);
}

__global__ void k_track_begin(LfTrackState *st, double t0, double t1, double t2, double t3, double t4, double t5,
                              double t6, double t7, double t8, double t9, double t10, double t11) {
    const double T[12] = {t0, t1, t2, t3, t4, t5, t6, t7, t8, t9, t10, t11};
    for (int k = 0; k < 12; ++k) {
        st->est[k] = T[k];
        st->prev[k] = T[k];
        st->init[k] = T[k];
        st->pose_f[k] = (float)T[k];
    }
    st->invalid = 0;
    st->active = 1;
    st->iterations = 0;
}

__global__ void k_track_set_pose(LfTrackState *st, double t0, double t1, double t2, double t3, double t4, double t5,
                                 double t6, double t7, double t8, double t9, double t10, double t11) {
    const double T[12] = {t0, t1, t2, t3, t4, t5, t6, t7, t8, t9, t10, t11};
    double Ti[12];
    m34_inv(T, Ti);
    for (int k = 0; k < 12; ++k) {
        st->pose[k] = T[k];
        st->pose24[k] = (float)T[k];
        st->pose24[12 + k] = (float)Ti[k];
    }
    st->success = 1;
}

int check_level(const LfTrackLevel &lv) {
    if (lv.cols <= 0 || lv.rows <= 0) return lf_set_error(LF_ERR_ARG, "bad level size %dx%d", lv.cols, lv.rows);
    return LF_OK;
}

float icp_angle() { return (float)cos(30 * 3.14 / 180); }

// Workgroups of a reduction: one pixel per lane, at most kGrid workgroups.
int grid_for(const LfTrackLevel &lv) {
    const long n = (long)lv.cols * lv.rows;
    return (int)std::max(1L, std::min((long)kGrid, (n + kThreads - 1) / kThreads));
}

}  // namespace

// The stand-alone reductions' kGrid x kVals FP32 partials, or the tracking
// step's kGrid x kTrackVals FP64 partials, whichever is larger.
size_t lf_reduce_partials(void) {
    return std::max((size_t)kGrid * kVals, (size_t)kGrid * kTrackVals * 2) + 64;
}

extern "C" size_t lf_reduce_scratch_floats(void) { return lf_reduce_partials(); }

int lf_track_begin(LfTrackState *st, const double *T, hipStream_t stream) {
    k_track_begin<<<1, 1, 0, stream>>>(st, T[0], T[1], T[2], T[3], T[4], T[5], T[6], T[7], T[8], T[9], T[10], T[11]);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

static ReduceArgs step_args(const LfTrackLevel &lv, float *scratch, LfTrackState *st, const LfStepFlags &f,
                            float icp_weight) {
    ReduceArgs a = {};
    a.lv = lv;
    a.angle = icp_angle();
    a.dthr = 0.01f;
    a.partials = scratch;
    a.st = st;
    a.first_iter = f.first_of_level;
    a.icp_weight = icp_weight;
    a.first_of_frame = f.first_of_frame;
    a.last_level = f.last_level;
    a.last_of_frame = f.last_of_frame;
    a.context_mode = f.context_mode;
    for (int k = 0; k < 12; ++k) {
        a.init[k] = f.init[k];
        a.init_f[k] = (float)f.init[k];
    }
    return a;
}

int lf_track_iteration(const LfTrackLevel &lv, float *scratch, LfTrackState *st, const LfStepFlags &f,
                       float icp_weight, hipStream_t stream) {
    if (int rc = check_level(lv)) return rc;
    const ReduceArgs a = step_args(lv, scratch, st, f, icp_weight);
    const int G = grid_for(lv);
    if (lv.half_maps) k_track_step<true><<<G, kThreads, 0, stream>>>(a);
    else k_track_step<false><<<G, kThreads, 0, stream>>>(a);
    LF_LAUNCH_CHECK();
    k_track_final<<<1, kThreads, 0, stream>>>(a, G);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

size_t lf_track_job_bytes(void) { return sizeof(TrackJob); }

int lf_track_job_make(const LfTrackLevel &lv, float *scratch, LfTrackState *st, const LfStepFlags &f,
                      float icp_weight, int share, int wgs, void *job, int *grid) {
    if (int rc = check_level(lv)) return rc;
    TrackJob &j = *static_cast<TrackJob *>(job);
    j = TrackJob{};
    j.a = step_args(lv, scratch, st, f, icp_weight);
    // 16 systems at 640x480: 32 workgroups each instead of 512 (the step's
    // lanes sum ~19 level-0 pixels each, with the next pixel's loads in
    // flight, and do 1/16 of the block reductions): 4.1k -> 4.7k frames/s
    // aggregate in mixed precision.  16 and 64 workgroups each measured
    // slower than 32 (5.46k and 5.61k against 5.74k).
    j.G = std::min(grid_for(lv), wgs > 0 ? std::min(wgs, kGrid) : std::max(1, kGrid / std::max(1, share)));
    if (grid) *grid = j.G;
    return LF_OK;
}

void lf_track_job_skip(void *job) {
    TrackJob &j = *static_cast<TrackJob *>(job);
    j = TrackJob{};
    j.G = 0;
}

int lf_track_batch_step(const void *jobs, int n, int gmax, int half_maps, hipStream_t stream) {
    if (!jobs || n <= 0 || n > 65535) return lf_set_error(LF_ERR_ARG, "bad tracking batch of %d", n);
    const TrackJob *J = static_cast<const TrackJob *>(jobs);
    if (gmax > 0) {
        if (half_maps) k_track_step_batch<true><<<dim3(gmax, n), kThreads, 0, stream>>>(J);
        else k_track_step_batch<false><<<dim3(gmax, n), kThreads, 0, stream>>>(J);
        LF_LAUNCH_CHECK();
    }
    k_track_final_batch<<<n, kThreads, 0, stream>>>(J);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

int lf_track_finish(LfTrackState *st, hipStream_t stream) {
    k_track_finish<<<1, 64, 0, stream>>>(st);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

int lf_track_set_pose(LfTrackState *st, const double *T, hipStream_t stream) {
    k_track_set_pose<<<1, 1, 0, stream>>>(st, T[0], T[1], T[2], T[3], T[4], T[5], T[6], T[7], T[8], T[9], T[10], T[11]);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

// icp_reduce of lf_icp_reduce and lf_icp_reduce_noalloc: the partial rows go
// to `partials` (capacity floats), then the final kernel sums them.
static int icp_reduce_into(const float *curr_vmap, const float *curr_nmap, const float *last_vmap,
                           const float *last_nmap, int cols, int rows, const float *pose12, const lf_intrinsics *K,
                           float *partials, size_t capacity, float *sum29, hipStream_t st) {
    ReduceArgs a = {};
    a.lv.curr_vmap = curr_vmap; a.lv.curr_nmap = curr_nmap; a.lv.last_vmap = last_vmap; a.lv.last_nmap = last_nmap;
    a.lv.cols = cols; a.lv.rows = rows; a.lv.K = *K;
    if (int rc = check_level(a.lv)) return rc;
    if ((long long)cols * rows > INT_MAX)
        return lf_set_error(LF_ERR_ARG, "icp_reduce: %dx%d is more than %d pixels", cols, rows, INT_MAX);
    a.pose12 = pose12;
    a.angle = icp_angle();
    a.dthr = 0.01f;
    a.partials = partials;
    a.out29 = sum29;
    const int G = grid_for(a.lv);
    if ((size_t)G * kVals > capacity)
        return lf_set_error(LF_ERR_CAPACITY, "icp_reduce: %d workgroups need %zu scratch floats, %zu available", G,
                            (size_t)G * kVals, capacity);
    k_reduce<1><<<G, kThreads, 0, st>>>(a);
    LF_LAUNCH_CHECK();
    k_reduce_final<1><<<1, kThreads, 0, st>>>(a, G);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" int lf_icp_reduce(const float *curr_vmap, const float *curr_nmap, const float *last_vmap,
                             const float *last_nmap, int cols, int rows, const float *pose12,
                             const lf_intrinsics *K, float *scratch, float *sum29, void *stream) {
    if (!curr_vmap || !curr_nmap || !last_vmap || !last_nmap || !pose12 || !K || !scratch || !sum29)
        return lf_set_error(LF_ERR_ARG, "null argument to icp_reduce");
    return icp_reduce_into(curr_vmap, curr_nmap, last_vmap, last_nmap, cols, rows, pose12, K, scratch,
                           lf_reduce_partials(), sum29, lf_stream(stream));
}

// Scratch of lf_icp_reduce_noalloc: the partial rows of at most kGrid
// workgroups (kGrid * kVals floats, 116 KiB), one copy per GPU, held until the
// process ends.  Host code never passes the array's name to a kernel or a
// copy: on the host it is only a stand-in for the symbol, so its device
// address is looked up per device (hipGetSymbolAddress, with that device
// current) and that address is used everywhere.
constexpr size_t kNoallocScratchFloats = (size_t)kGrid * kVals;
static __device__ __align__(256) float scratch[kNoallocScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float *g_scratch_dev[kMaxDevices] = {};

// The current device's address of `scratch`, resolved on its first use there.
int noalloc_scratch(float **out) {
    int dev = -1;
    LF_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return lf_set_error(LF_ERR_ARG, "device %d is beyond the %d devices of the fixed scratch", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return lf_set_error(LF_ERR_HIP, "hipGetSymbolAddress of the fixed scratch failed on device %d: %s", dev,
                                hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<float *>(p);
    }
    *out = g_scratch_dev[dev];
    return LF_OK;
}
}  // namespace

extern "C" int lf_icp_reduce_noalloc(const float *curr_vmap, const float *curr_nmap, const float *last_vmap,
                                     const float *last_nmap, int cols, int rows, const float *pose12,
                                     const lf_intrinsics *K, float *sum29, void *stream) {
    if (!curr_vmap || !curr_nmap || !last_vmap || !last_nmap || !pose12 || !K || !sum29)
        return lf_set_error(LF_ERR_ARG, "null argument to icp_reduce_noalloc");
    float *partials = nullptr;
    if (int rc = noalloc_scratch(&partials)) return rc;
    // every partial the final kernel reads is written by this call's first
    // kernel, so nothing in the scratch needs clearing
    return icp_reduce_into(curr_vmap, curr_nmap, last_vmap, last_nmap, cols, rows, pose12, K, partials,
                           kNoallocScratchFloats, sum29, lf_stream(stream));
}

extern "C" int lf_rgb_step(const float *curr_intensity, const float *last_intensity, const float *last_vmap,
                           const float *intensity_dx, const float *intensity_dy, int cols, int rows,
                           float stddev_estimate, const float *pose12, const lf_intrinsics *K, float *scratch,
                           float *sum29, void *stream) {
    if (!curr_intensity || !last_intensity || !last_vmap || !intensity_dx || !intensity_dy || !pose12 || !K ||
        !scratch || !sum29)
        return lf_set_error(LF_ERR_ARG, "null argument to rgb_step");
    ReduceArgs a = {};
    a.lv.last_vmap = last_vmap; a.lv.curr_I = curr_intensity; a.lv.last_I = last_intensity;
    a.lv.dIdx = intensity_dx; a.lv.dIdy = intensity_dy;
    a.lv.cols = cols; a.lv.rows = rows; a.lv.K = *K;
    if (int rc = check_level(a.lv)) return rc;
    a.pose12 = pose12;
    a.stddev = stddev_estimate;
    a.partials = scratch;
    a.out29 = sum29;
    const int G = grid_for(a.lv);
    k_reduce<2><<<G, kThreads, 0, lf_stream(stream)>>>(a);
    LF_LAUNCH_CHECK();
    k_reduce_final<2><<<1, kThreads, 0, lf_stream(stream)>>>(a, G);
    LF_LAUNCH_CHECK();
    return LF_OK;
}

extern "C" void lf_unpack_jtjjtr(const float *sum29, float *jtj36, float *jtr6, float *residual2) {
    int k = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j) {
            const float v = sum29[k++];
            if (j == 6) jtr6[i] = v;
            else jtj36[j * 6 + i] = jtj36[i * 6 + j] = v;
        }
    residual2[0] = sum29[27];
    residual2[1] = sum29[28];
}
