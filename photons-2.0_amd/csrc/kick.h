// kick.h -- the reference's kick, drift and wrap (src/photoNs.c:158-196, 242-254) as
// elementwise kernels over the [n][3] FP64 pos/vel/force arrays of one box.
// Included by the translation units that step particles; each keeps its own copy.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

// src/photoNs.c:178-196: add or subtract L until x is in [0, L); far outliers are
// folded first so the loops stay short
__device__ __forceinline__ double wrap_box(double x, double L)
{
    if (!(fabs(x) < 4.0 * L)) x -= floor(x / L) * L;
    while (x < 0.0) x += L;
    while (x >= L) x -= L;
    return x;
}

// the first half of a step: half kicks with the previous forces, long range first
// (src/photoNs.c:158-170), drift, wrap.  acc_pm may be null (no long-range part).
// When step is 0 nothing moves.  out, when set, receives the new positions.
__global__ void k_kick_drift(int64_t n3, double *__restrict__ pos, double *__restrict__ vel,
                             const double *__restrict__ acc, const double *__restrict__ acc_pm, double dkh,
                             double dd, double L, int step, double *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    double x = pos[i];
    if (step) {
        double v = vel[i];
        if (acc_pm) v += acc_pm[i] * dkh;
        v += acc[i] * dkh;
        vel[i] = v;
        x = wrap_box(x + v * dd, L);
        pos[i] = x;
    }
    if (out) out[i] = x;
}

// the second half: half kicks with the new forces, short range first
// (src/photoNs.c:242-254).  in, when set, is the new short-range force computed
// into a batch buffer; it is stored as the box's acc first.
__global__ void k_kick(int64_t n3, double *__restrict__ vel, double *__restrict__ acc, const double *__restrict__ in,
                       const double *__restrict__ acc_pm, double dkh, int step)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    double a = acc[i];
    if (in) {
        a = in[i];
        acc[i] = a;
    }
    if (step) {
        double v = vel[i] + a * dkh;
        if (acc_pm) v += acc_pm[i] * dkh;
        vel[i] = v;
    }
}

inline unsigned kick_blocks(int64_t n3) { return (unsigned)((n3 + 255) / 256); }

}  // namespace
