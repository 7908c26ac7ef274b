// sphk_adjoint.hip -- vector-Jacobian product of the gravity functor
// (uniform softening), the reverse of gravity_pair_uniform in FP64.
//
// Layout as the forward kernel: one wave per walk, S = floor(64/ni) j-slices,
// 64-record LDS tiles.  The i-side cotangent of a lane stays in registers
// and is combined across slices through LDS at the end of the walk; the
// j-side contribution of a pair goes to a per-tile LDS accumulator of that
// record (LDS atomics), flushed to HBM with one atomic add per record and
// component after the tile.
#include "sphk_adjoint.h"
#include "sphk_device.hpp"
#include "sphk_internal.hpp"

namespace sphk {
namespace {

struct VRec {
    double x, y, z, m;
};

__device__ __forceinline__ int xcd_walk_adj(int b, int n_walk)
{
    // as the forward kernels: chunks of 64 walks dealt to the 8 XCDs in turn
    constexpr int kX = 8, kC = 64;
    const int n_full = n_walk / (kX * kC) * (kX * kC);
    if (b >= n_full) return b;
    const int x = b % kX, k = b / kX;
    return ((k / kC) * kX + x) * kC + k % kC;
}

__device__ void vjp_list(int lane, int S, int slice, int kb, int ke, const int32_t *__restrict__ idx,
                         const sphk_grav_src &src, double px, double py, double pz, double e2, double ax,
                         double ay, double az, double pb, double eb, VRec *sp, double *jb, int32_t *sj,
                         double (&gi)[3],
                         double *outx, double *outy, double *outz)
{
    for (int t0 = kb; t0 < ke; t0 += kTile) {
        const int nt = min(kTile, ke - t0);
        __syncthreads();
        if (lane < nt) {
            const int j = idx ? idx[t0 + lane] : t0 + lane;
            sp[lane] = VRec{src.x[j], src.y[j], src.z[j], src.m[j]};
            sj[lane] = j;
        } else {
            sp[lane] = VRec{1e10, 1e10, 1e10, 0.};
            sj[lane] = -1;
        }
        jb[lane] = 0.;
        jb[kTile + lane] = 0.;
        jb[2 * kTile + lane] = 0.;
        __syncthreads();
        for (int k = slice; k < kTile; k += S) {
            const VRec r = sp[k];
            const double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
            const double s = dx * dx + dy * dy + dz * dz + e2;
            const double w = rsqrt_f64(s);
            const double w3 = w * w * w;
            const double w5 = w3 * w * w;
            const double dot = dx * ax + dy * ay + dz * az;
            // T abar = w3 abar - 3 w5 d (d . abar); then subtract pbar w3 d
            // and add 3 ebar w5 d (the eta sum's gradient is -3 m w5 d)
            const double f = 3. * w5 * (dot - eb) + pb * w3;
            const double cx = r.m * (w3 * ax - f * dx);
            const double cy = r.m * (w3 * ay - f * dy);
            const double cz = r.m * (w3 * az - f * dz);
            gi[0] -= cx;
            gi[1] -= cy;
            gi[2] -= cz;
            atomicAdd(&jb[k], cx);
            atomicAdd(&jb[kTile + k], cy);
            atomicAdd(&jb[2 * kTile + k], cz);
        }
        __syncthreads();
        if (lane < nt) {
            const int j = sj[lane];
            atomicAdd(&outx[j], jb[lane]);
            atomicAdd(&outy[j], jb[kTile + lane]);
            atomicAdd(&outz[j], jb[2 * kTile + lane]);
        }
    }
}

__global__ __launch_bounds__(kWave) void gravity_vjp_kernel(
    int n_walk, const int32_t *__restrict__ i_off, const int32_t *__restrict__ j_off,
    const int32_t *__restrict__ j_idx, sphk_grav_src pj, const int32_t *__restrict__ s_off,
    const int32_t *__restrict__ s_idx, sphk_grav_src ps, sphk_grav_src pi, double e2,
    const double *__restrict__ abx, const double *__restrict__ aby, const double *__restrict__ abz,
    const double *__restrict__ pbar, const double *__restrict__ ebar, double *gx, double *gy, double *gz,
    double *nx, double *ny, double *nz)
{
    __shared__ VRec sp[kTile];
    __shared__ double jb[3 * kTile];
    __shared__ int32_t sj[kTile];
    __shared__ double red[3 * kWave];
    if ((int)blockIdx.x >= n_walk) return;
    const int w = xcd_walk_adj(blockIdx.x, n_walk);
    const int lane = threadIdx.x;
    const int ib = i_off[w], ie = i_off[w + 1];
    for (int c0 = ib; c0 < ie; c0 += kWave) {
        const int ni = min(kWave, ie - c0);
        const int S = kWave / ni, il = lane % ni;
        const int slice = lane / ni < S ? lane / ni : kTile;   // idle lanes sweep nothing
        const int i = c0 + il;
        const double px = pi.x[i], py = pi.y[i], pz = pi.z[i];
        const double ax = abx[i], ay = aby[i], az = abz[i], pb = pbar[i];
        const double eb = ebar ? ebar[i] : 0.;
        double gi[3] = {0., 0., 0.};
        vjp_list(lane, S, slice, j_off[w], j_off[w + 1], j_idx, pj, px, py, pz, e2, ax, ay, az, pb, eb, sp, jb,
                 sj, gi, gx, gy, gz);
        if (s_off)
            vjp_list(lane, S, slice, s_off[w], s_off[w + 1], s_idx, ps, px, py, pz, e2, ax, ay, az, pb, eb, sp,
                     jb, sj, gi, nx, ny, nz);
        __syncthreads();
        for (int q = 0; q < 3; q++) red[q * kWave + lane] = gi[q];
        __syncthreads();
        if (lane < ni) {
            double s[3] = {0., 0., 0.};
            for (int t = 0; t < S; t++)
                for (int q = 0; q < 3; q++) s[q] += red[q * kWave + lane + ni * t];
            atomicAdd(&gx[i], s[0]);
            atomicAdd(&gy[i], s[1]);
            atomicAdd(&gz[i], s[2]);
        }
    }
}

}  // namespace
}  // namespace sphk

extern "C" int sphk_gravity_vjp_dev(sphk_ctx *ctx, int n_walk, const int32_t *i_off, const int32_t *j_off,
                                    const int32_t *j_idx, sphk_grav_src pj, const int32_t *s_off,
                                    const int32_t *s_idx, sphk_grav_src ps, sphk_grav_src pi, double e2,
                                    const double *abx, const double *aby, const double *abz,
                                    const double *pbar, const double *ebar, double *gx, double *gy,
                                    double *gz, double *nx, double *ny, double *nz, void *stream)
{
    using sphk::set_error;
    if (!ctx) return set_error(SPHK_E_ARG, "gravity_vjp: null context");
    if (n_walk < 0 || (n_walk > 0 && (!i_off || !j_off))) return set_error(SPHK_E_ARG, "gravity_vjp: bad walks");
    if (n_walk == 0) return SPHK_OK;
    if (!pi.x || !pi.y || !pi.z || !pj.m || !pj.x || !pj.y || !pj.z || !abx || !aby || !abz || !pbar || !gx ||
        !gy || !gz)
        return set_error(SPHK_E_ARG, "gravity_vjp: null pointer");
    if (s_off && (!ps.m || !ps.x || !ps.y || !ps.z || !nx || !ny || !nz))
        return set_error(SPHK_E_ARG, "gravity_vjp: null SPJ pointer");
    if (!(e2 > 0.)) return set_error(SPHK_E_ARG, "gravity_vjp: softening must be positive");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(sphk::gravity_vjp_kernel, dim3(n_walk), dim3(sphk::kWave), 0, st, n_walk, i_off, j_off,
                       j_idx, pj, s_off, s_idx, ps, pi, e2, abx, aby, abz, pbar, ebar, gx, gy, gz, nx, ny, nz);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(SPHK_E_DEVICE, "gravity_vjp: %s", hipGetErrorString(e));
    return SPHK_OK;
}
