// sphk_device.hpp -- device-side helpers shared by the gfx950 interaction kernels.
//
// Wave64 throughout: one workgroup = one wavefront = one walk (i-group).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "sphk.h"

namespace sphk {

constexpr int kWave = 64;   // CDNA wavefront width
constexpr int kTile = 64;   // j-entries staged in LDS per tile

// Wendland C2 constants, sph_yamauchi/run/hdr_kernel.hpp:226-227.
constexpr double kCeff0 = 3.342253804929802286;
constexpr double kCeff1 = 13.36901521971920914;

// 1/sqrt(x) in FP64: the v_rsq_f64 seed refined by two Newton-Raphson steps
// y <- y + y (1/2 - x y^2 / 2).  Relative error is at the rounding level.
// x == 0 yields NaN; callers select around it where the reference defines
// the value (self pairs).
__device__ __forceinline__ double rsqrt_f64(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    double e = __builtin_fma(-h * y, y, 0.5);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-h * y, y, 0.5);
    y = __builtin_fma(y, e, y);
    return y;
}

// 1/x in FP64: v_rcp_f64 seed and two Newton-Raphson steps.
__device__ __forceinline__ double rcp_f64(double x)
{
    double y = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-x, y, 1.0);
    y = __builtin_fma(y, e, y);
    return y;
}

// Cube root of a finite x > 0 in FP64.  x = f 2^(3k) with f in [0.5, 4); an
// FP32 seed z ~ f^(-1/3) (v_log_f32, v_exp_f32) is refined by two
// division-free Newton steps z <- z + z (1 - f z^3) / 3, then y = f z^2 gets
// one Newton correction y <- y - (y^3 - f) z^2 / 3.  Within a few ulp of the
// exact cube root.
__device__ __forceinline__ double cbrt_pos_f64(double x)
{
    const int e = __builtin_amdgcn_frexp_exp(x);   // x = m 2^e, m in [0.5, 1)
    const double m = __builtin_amdgcn_frexp_mant(x);
    const int k = e >= 0 ? e / 3 : -((2 - e) / 3);
    const double f = __builtin_amdgcn_ldexp(m, e - 3 * k);
    double z = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)f) * (-1.f / 3.f));
    double t = f * (z * z * z);
    z = __builtin_fma(z * (1. / 3.), 1. - t, z);
    t = f * (z * z * z);
    z = __builtin_fma(z * (1. / 3.), 1. - t, z);
    const double z2 = z * z;
    double y = f * z2;
    const double r = __builtin_fma(y * y, y, -f);
    y = __builtin_fma(-r * (1. / 3.), z2, y);
    return __builtin_amdgcn_ldexp(y, k);
}

// x^(1/3) with pow's special values for x <= 0, inf and NaN (x = 0 -> 0).
__device__ __forceinline__ double cbrt_f64(double x)
{
    if (x > 0. && x < __builtin_inf()) return cbrt_pos_f64(x);
    return x == 0. ? 0. : x;
}

__device__ __forceinline__ float rsqrt_r(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ double rsqrt_r(double x) { return rsqrt_f64(x); }
__device__ __forceinline__ float rcp_r(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float abs_r(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double abs_r(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ double rcp_r(double x) { return rcp_f64(x); }

// W(q) = C0 (1-q)_+^4 (1+4q), hdr_kernel.hpp:246-250.
template <typename R>
__device__ __forceinline__ R kernel0th(R q)
{
    R rmin = R(1) - q;
    rmin = rmin > R(0) ? rmin : R(0);
    const R rmin2 = rmin * rmin;
    return R(kCeff0) * rmin2 * rmin2 * (R(1) + R(4) * q);
}

// W'(q) = C1 [(1-q)_+^4 - (1-q)_+^3 (1+4q)] (hdr_kernel.hpp:251-257) equals
// -5 C1 q (1-q)_+^3; the pair terms use that closed form directly.

// Butterfly sum over the lanes that share one i-particle.  Lane l holds i
// (l & (P-1)) and j-slice (l >> logP); masks P..32 combine the slices.  Every
// lane of the group ends with a bitwise identical total (each step adds the
// same two operands, commutatively).
template <typename T>
__device__ __forceinline__ T slice_sum(T v, int P)
{
    for (int m = P; m < kWave; m <<= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

template <typename T>
__device__ __forceinline__ T slice_max(T v, int P)
{
    for (int m = P; m < kWave; m <<= 1) {
        T o = __shfl_xor(v, m, kWave);
        v = v > o ? v : o;
    }
    return v;
}

// Lane layout of one i-chunk of a walk: ni (1..64) i-particles, P = next
// power of two >= ni, S = 64/P j-slices.
struct LaneMap {
    int ni, P, logP, S, il, slice;
    bool active;
    __device__ __forceinline__ LaneMap(int ni_, int lane)
    {
        ni = ni_;
        logP = (ni <= 1) ? 0 : 32 - __clz(ni - 1);
        P = 1 << logP;
        S = kWave >> logP;
        il = lane & (P - 1);
        slice = lane >> logP;
        active = il < ni;
    }
};

}  // namespace sphk
