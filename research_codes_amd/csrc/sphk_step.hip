// sphk_step.hip -- per-particle kernels of one SPH step (include/sphk_step.h).
//
// All are HBM-bound elementwise passes: one thread per particle, 256-thread
// workgroups, FP64 throughout.  The diagnostics reduction is two-stage with
// a fixed tree order (per-block partials, then one block), so repeated runs
// are bitwise identical.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "sphk.h"
#include "sphk_internal.hpp"
#include "sphk_step.h"

namespace sphk {

constexpr int kBlock = 256;

static int blocks(int n) { return (n + kBlock - 1) / kBlock; }

__global__ __launch_bounds__(kBlock) void eos_balsara_kernel(int n, sphk_state s, sphk_phys p)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double dens = s.dens[i];
    const double pres = p.ninv * dens * s.uene[i];               // hdr_hgas.hpp:339
    const double vsnd = sqrt((1. + p.ninv) * pres / dens);       // hdr_hgas.hpp:341
    const double ad = fabs(s.divv[i]);
    s.pres[i] = pres;
    s.vsnd[i] = vsnd;
    s.bswt[i] = ad / (ad + fabs(s.rotv[i]) + 1e-4 * vsnd * p.ksrh / s.ksr[i]);  // hdr_sph.hpp:80-81
}

__global__ __launch_bounds__(kBlock) void softening_kernel(int n, const double *ksr, double kinv,
                                                           double e2c, double *eps2)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    if (e2c > 0.) {
        eps2[i] = e2c;
    } else {
        const double h = ksr[i];
        eps2[i] = h * h * kinv * kinv;                            // hdr_gravity.hpp:136
    }
}

__global__ __launch_bounds__(kBlock) void gravity_writeback_kernel(int n, sphk_state s, sphk_phys p,
                                                                   sphk_grav_out g)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    s.g2x[i] = p.G * g.ax[i];
    s.g2y[i] = p.G * g.ay[i];
    s.g2z[i] = p.G * g.az[i];
    // hdr_gravity.hpp:110-111: eta is dPhi/dH through eps = H/kappa; with a
    // constant softening the potential does not depend on H and eta is 0
    // (then nothing the density pass writes is read here)
    if (p.eps_const > 0.) {
        s.eta[i] = 0.;
    } else {
        const double kinv = 1. / p.ksrh;
        const double h = s.ksr[i];
        s.eta[i] = h * h * kinv * kinv * s.grdh[i] / (3. * s.dens[i]) * g.eta[i];
    }
    s.pot[i] = p.G * g.pot[i];
}

__global__ __launch_bounds__(kBlock) void hydro_prep_kernel(int n, sphk_state s, double *hinv,
                                                            double *presc, double *bswt_half,
                                                            double *thrm)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double d = s.dens[i];
    hinv[i] = 1. / s.ksr[i];                                      // hdr_hydro.hpp:50-60
    presc[i] = s.pres[i] * s.grdh[i] / (d * d);
    bswt_half[i] = 0.5 * s.bswt[i];
    const double u = s.uene[i];
    thrm[i] = u > 0. ? u : 0.;
}

__global__ __launch_bounds__(kBlock) void hydro_writeback_kernel(int n, sphk_state s, sphk_phys p,
                                                                 sphk_hydro_out h)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    // hdr_hydro.hpp:21-25
    const double ahx = h.ax[i], ahy = h.ay[i], ahz = h.az[i];
    const double g1x = h.gx[i] * p.G, g1y = h.gy[i] * p.G, g1z = h.gz[i] * p.G;
    s.achx[i] = ahx;
    s.achy[i] = ahy;
    s.achz[i] = ahz;
    s.g1x[i] = g1x;
    s.g1y[i] = g1y;
    s.g1z[i] = g1z;
    s.udot[i] = h.udot[i];
    s.vsmx[i] = h.vsmx[i];
    const double diffu = h.diffu[i];
    s.diffu[i] = diffu;
    // hdr_sph.hpp:95-97
    s.ax[i] = ahx + g1x + s.g2x[i];
    s.ay[i] = ahy + g1y + s.g2y[i];
    s.az[i] = ahz + g1z + s.g2z[i];
    // hdr_hgas.hpp:750-762
    const double alph = s.alph[i], alphu = s.alphu[i], ksr = s.ksr[i];
    double src = -s.divv[i] * (p.alpha_max - alph);
    src = src > 0. ? src : 0.;
    const double tauinv = (0.25 * p.ksrh * s.vsnd[i]) / ksr;
    s.adot[i] = -(alph - p.alpha_min) * tauinv + src;
    double u = s.uene[i];
    u = u > 0. ? u : 0.;
    double lim = p.alphu_max - alphu;
    lim = lim > 0. ? lim : 0.;
    const double srcu = ksr / p.ksrh * fabs(diffu) / sqrt(u + p.eps_u) * lim;
    s.adotu[i] = -(alphu - p.alphu_min) * tauinv + srcu;
}

__global__ __launch_bounds__(kBlock) void predict_kernel(int n, sphk_state s, double dt)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    // hdr_hgas.hpp:646-657
    const double ax = s.ax[i], ay = s.ay[i], az = s.az[i];
    const double vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
    s.x[i] = s.x[i] + vx * dt + 0.5 * ax * dt * dt;
    s.y[i] = s.y[i] + vy * dt + 0.5 * ay * dt * dt;
    s.z[i] = s.z[i] + vz * dt + 0.5 * az * dt * dt;
    s.vx2[i] = vx + 0.5 * ax * dt;
    s.vy2[i] = vy + 0.5 * ay * dt;
    s.vz2[i] = vz + 0.5 * az * dt;
    s.vx[i] = vx + ax * dt;
    s.vy[i] = vy + ay * dt;
    s.vz[i] = vz + az * dt;
    const double u = s.uene[i], ud = s.udot[i];
    s.uene2[i] = u + 0.5 * ud * dt;
    s.uene[i] = u + ud * dt;
    const double a = s.alph[i], ad = s.adot[i];
    s.alph2[i] = a + 0.5 * ad * dt;
    s.alph[i] = a + ad * dt;
    const double au = s.alphu[i], aud = s.adotu[i];
    s.alphu2[i] = au + 0.5 * aud * dt;
    s.alphu[i] = au + aud * dt;
}

__global__ __launch_bounds__(kBlock) void correct_kernel(int n, sphk_state s, double dt)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    // hdr_hgas.hpp:659-665
    s.vx[i] = s.vx2[i] + 0.5 * s.ax[i] * dt;
    s.vy[i] = s.vy2[i] + 0.5 * s.ay[i] * dt;
    s.vz[i] = s.vz2[i] + 0.5 * s.az[i] * dt;
    s.uene[i] = s.uene2[i] + 0.5 * s.udot[i] * dt;
    s.alph[i] = s.alph2[i] + 0.5 * s.adot[i] * dt;
    s.alphu[i] = s.alphu2[i] + 0.5 * s.adotu[i] * dt;
}

// Per-particle terms of the diagnostics: E, Px, Py, Pz, dt_i, m.
__device__ __forceinline__ void diag_terms(const sphk_state &s, const sphk_phys &p, int i, double v[6])
{
    const double m = s.m[i];
    const double vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
    v[0] = m * (0.5 * (vx * vx + vy * vy + vz * vz) + s.uene[i] + 0.5 * s.pot[i]);
    v[1] = m * vx;
    v[2] = m * vy;
    v[3] = m * vz;
    // hdr_hgas.hpp:477-488
    const double ksr = s.ksr[i];
    const double dth = p.c_dt * ksr / (s.vsmx[i] * p.ksrh);
    const double divv = s.divv[i];
    const double dtc = divv < 0. ? -p.c_dt / divv : p.dt_max;
    const double gx = s.g1x[i] + s.g2x[i], gy = s.g1y[i] + s.g2y[i], gz = s.g1z[i] + s.g2z[i];
    const double dtg = p.c_dt * sqrt(ksr / sqrt(gx * gx + gy * gy + gz * gz));
    double a = dth < p.dt_max ? dth : p.dt_max;
    double b = dtg < dtc ? dtg : dtc;
    v[4] = a < b ? a : b;
    v[5] = m;
}

// Block reduction of six values: sums, except the minimum for the slots in
// min_mask (bit k = slot k).
__device__ __forceinline__ void block_reduce6(double v[6], double *sh, int min_mask = 1 << 4)
{
    const int t = threadIdx.x;
    for (int k = 0; k < 6; k++) sh[k * kBlock + t] = v[k];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (t < w) {
            for (int k = 0; k < 6; k++) {
                const double a = sh[k * kBlock + t], b = sh[k * kBlock + t + w];
                sh[k * kBlock + t] = ((min_mask >> k) & 1) ? (a < b ? a : b) : a + b;
            }
        }
        __syncthreads();
    }
    for (int k = 0; k < 6; k++) v[k] = sh[k * kBlock];
}

__global__ __launch_bounds__(kBlock) void diag_partial_kernel(int n, sphk_state s, sphk_phys p,
                                                              double *part)
{
    __shared__ double sh[6 * kBlock];
    double acc[6] = {0., 0., 0., 0., 1e300, 0.};
    // grid-stride with a fixed assignment: block b handles i = b*kBlock + t + k*stride
    const int stride = gridDim.x * kBlock;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        double v[6];
        diag_terms(s, p, i, v);
        for (int k = 0; k < 6; k++) acc[k] = (k == 4) ? (acc[k] < v[k] ? acc[k] : v[k]) : acc[k] + v[k];
    }
    block_reduce6(acc, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) part[k * gridDim.x + blockIdx.x] = acc[k];
}

__global__ __launch_bounds__(kBlock) void diag_final_kernel(int nb, const double *part, double *out)
{
    __shared__ double sh[6 * kBlock];
    double acc[6] = {0., 0., 0., 0., 1e300, 0.};
    for (int b = threadIdx.x; b < nb; b += kBlock)
        for (int k = 0; k < 6; k++) {
            const double v = part[k * nb + b];
            acc[k] = (k == 4) ? (acc[k] < v ? acc[k] : v) : acc[k] + v;
        }
    block_reduce6(acc, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) out[k] = acc[k];
}

}  // namespace sphk

using namespace sphk;

namespace {

// The caller's stream; NULL is the HIP null stream (see include/sphk.h).
// One round of the smoothing-length retry loop (calcDensityKernel,
// hdr_density.hpp:382-414) for particle i: a particle still retrying
// (rs != 0) whose round converged (itr == 0) takes the round's results
// (copyFromForce, hdr_density.hpp:24-31) and gets rs = 0 (its rs to rs_conv
// when given); one that did not converge retries with rs * expand (with
// jump: at least expand * ksr of the round); flag is set if any retries.
__global__ __launch_bounds__(kBlock) void density_commit_kernel(int n, sphk_state s, sphk_dens_out r,
                                                                int64_t *__restrict__ np, double *__restrict__ rs,
                                                                double *__restrict__ rs_conv, double expand, int jump,
                                                                int32_t *__restrict__ flag)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double rsi = rs[i];
    if (rsi == 0.) return;   // converged in an earlier round
    if (r.itr[i] == 0) {
        s.dens[i] = r.dens[i];
        s.divv[i] = r.divv[i];
        s.rotv[i] = r.rotv[i];
        s.grdh[i] = r.grdh[i];
        s.ksr[i] = r.ksr[i];
        np[i] = r.np[i];
        if (rs_conv) rs_conv[i] = rsi;
        rs[i] = 0.;
        return;
    }
    double grown = rsi * expand;
    if (jump) {
        const double k = expand * r.ksr[i];
        grown = grown > k ? grown : k;
    }
    rs[i] = grown;
    *flag = 1;   // every writer stores 1
}

__global__ void zero_flag_kernel(int32_t *flag) { *flag = 0; }

hipStream_t stream_of(sphk_ctx *, void *stream) { return reinterpret_cast<hipStream_t>(stream); }

int launch_check(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(SPHK_E_DEVICE, "%s: %s", what, hipGetErrorString(e));
    return SPHK_OK;
}

bool state_ok(const sphk_state &s)
{
    const void *const *f = reinterpret_cast<const void *const *>(&s);
    for (size_t k = 0; k < sizeof(sphk_state) / sizeof(void *); k++)
        if (!f[k]) return false;
    return true;
}

int pre(sphk_ctx *ctx, int n, const sphk_state *s)
{
    if (!ctx) return set_error(SPHK_E_ARG, "null context");
    if (n < 0) return set_error(SPHK_E_ARG, "negative particle count");
    if (s && !state_ok(*s)) return set_error(SPHK_E_ARG, "null state field");
    return SPHK_OK;
}

// ---- unbound mass (two fixed-order reductions) ---------------------------
__device__ __forceinline__ bool selected(const int32_t *istar, int star, int i)
{
    return istar == nullptr || star < 0 || istar[i] == star;
}

// pass 1: sum rho, rho v (slots 0..3)
__global__ __launch_bounds__(kBlock) void unbound_vcg_partial(int n, sphk_state s, const int32_t *istar,
                                                              int star, double *part)
{
    __shared__ double sh[6 * kBlock];
    double acc[6] = {0., 0., 0., 0., 0., 0.};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        if (!selected(istar, star, i)) continue;
        const double r = s.dens[i];
        acc[0] += r;
        acc[1] += r * s.vx[i];
        acc[2] += r * s.vy[i];
        acc[3] += r * s.vz[i];
    }
    block_reduce6(acc, sh, 0);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) part[k * gridDim.x + blockIdx.x] = acc[k];
}

__device__ __forceinline__ double sigmoid(double x)
{
    if (x >= 0.) return 1. / (1. + exp(-x));
    const double e = exp(x);
    return e / (1. + e);
}

// pass 2: hard and smoothed unbound mass, selected mass (slots 0..2);
// vcg = the pass-1 totals
__global__ __launch_bounds__(kBlock) void unbound_mass_partial(int n, sphk_state s, const int32_t *istar,
                                                               int star, double k, const double *vcg,
                                                               double *part)
{
    __shared__ double sh[6 * kBlock];
    const double cx = vcg[1] / vcg[0], cy = vcg[2] / vcg[0], cz = vcg[3] / vcg[0];
    double acc[6] = {0., 0., 0., 0., 0., 0.};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        if (!selected(istar, star, i)) continue;
        const double vx = s.vx[i] - cx, vy = s.vy[i] - cy, vz = s.vz[i] - cz;
        const double e = 0.5 * (vx * vx + vy * vy + vz * vz) + s.pot[i] + s.uene[i];
        const double m = s.m[i];
        acc[0] += e > 0. ? m : 0.;
        acc[1] += m * sigmoid(k * e);
        acc[2] += m;
    }
    block_reduce6(acc, sh, 0);
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; q++) part[q * gridDim.x + blockIdx.x] = acc[q];
}

__global__ __launch_bounds__(kBlock) void sum_final_kernel(int nb, const double *part, double *out)
{
    __shared__ double sh[6 * kBlock];
    double acc[6] = {0., 0., 0., 0., 0., 0.};
    for (int b = threadIdx.x; b < nb; b += kBlock)
        for (int k = 0; k < 6; k++) acc[k] += part[k * nb + b];
    block_reduce6(acc, sh, 0);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) out[k] = acc[k];
}

// ---- allocation-free diagnostics ------------------------------------------
// The partials of sphk_diagnostics_static: 6 sums per block of the fixed
// 1024-block grid (48 KiB).  Each GPU holds its own copy until the process
// ends; host code reaches it only through the address hipGetSymbolAddress
// returns for that GPU (kept in g_scratch_ptr), never through the name.
constexpr int kDiagBlocks = 1024;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[6 * kDiagBlocks];
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Address of this GPU's scratch (looked up once per device, with that device
// current).
int scratch_of_current_device(double **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(SPHK_E_DEVICE, "diagnostics_static: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(SPHK_E_DEVICE, "diagnostics_static: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(SPHK_E_DEVICE, "diagnostics_static: hipGetSymbolAddress: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return SPHK_OK;
}

// Scratch for the diagnostics partials: one per process and device ordinal.
struct DiagScratch {
    double *part = nullptr;
    double *out = nullptr;
    int cap = 0;
    int device = -1;
};
thread_local DiagScratch g_diag;
thread_local DiagScratch g_unbound;   // out holds 12: pass-1 and pass-2 totals

int diag_scratch(DiagScratch &d, int dev, int nb, int nout)
{
    if (d.device == dev && d.cap >= nb) return SPHK_OK;
    if (d.part) (void)hipFree(d.part);
    if (d.out) (void)hipFree(d.out);
    d = DiagScratch();
    if (hipMalloc(&d.part, sizeof(double) * 6 * nb) != hipSuccess ||
        hipMalloc(&d.out, sizeof(double) * nout) != hipSuccess)
        return set_error(SPHK_E_DEVICE, "diagnostics: hipMalloc failed");
    d.cap = nb;
    d.device = dev;
    return SPHK_OK;
}

}  // namespace

extern "C" {

int sphk_eos_balsara_dev(sphk_ctx *ctx, int n, sphk_state s, sphk_phys p, void *stream)
{
    int rc = pre(ctx, n, &s);
    if (rc || n == 0) return rc;
    hipLaunchKernelGGL(eos_balsara_kernel, dim3(blocks(n)), dim3(kBlock), 0, stream_of(ctx, stream), n, s, p);
    return launch_check("eos_balsara");
}

int sphk_density_commit_dev(sphk_ctx *ctx, int n, sphk_state s, sphk_dens_out round, int64_t *np, double *rs,
                            double *rs_conv, double expand, int jump, int32_t *retry, void *stream)
{
    int rc = pre(ctx, n, &s);
    if (rc) return rc;
    if (!retry) return set_error(SPHK_E_ARG, "density_commit: null retry flag");
    hipStream_t st = stream_of(ctx, stream);
    hipLaunchKernelGGL(zero_flag_kernel, dim3(1), dim3(1), 0, st, retry);
    if (n > 0) {
        if (!np || !rs || !round.dens || !round.divv || !round.rotv || !round.grdh || !round.ksr || !round.np ||
            !round.itr)
            return set_error(SPHK_E_ARG, "density_commit: null array");
        hipLaunchKernelGGL(density_commit_kernel, dim3(blocks(n)), dim3(kBlock), 0, st, n, s, round, np, rs,
                           rs_conv, expand, jump, retry);
    }
    return launch_check("density_commit");
}

int sphk_softening_dev(sphk_ctx *ctx, int n, const double *ksr, double ksrh, double eps_const,
                       double *eps2, void *stream)
{
    int rc = pre(ctx, n, nullptr);
    if (rc || n == 0) return rc;
    if (!ksr || !eps2 || !(ksrh > 0.)) return set_error(SPHK_E_ARG, "softening: bad argument");
    const double e2c = eps_const > 0. ? eps_const * eps_const : 0.;
    hipLaunchKernelGGL(softening_kernel, dim3(blocks(n)), dim3(kBlock), 0, stream_of(ctx, stream), n, ksr,
                       1. / ksrh, e2c, eps2);
    return launch_check("softening");
}

int sphk_gravity_writeback_dev(sphk_ctx *ctx, int n, sphk_state s, sphk_phys p, sphk_grav_out g,
                               void *stream)
{
    int rc = pre(ctx, n, &s);
    if (rc || n == 0) return rc;
    if (!g.ax || !g.ay || !g.az || !g.pot || (!g.eta && !(p.eps_const > 0.)))
        return set_error(SPHK_E_ARG, "gravity_writeback: null (eta may be NULL only under a constant softening)");
    hipLaunchKernelGGL(gravity_writeback_kernel, dim3(blocks(n)), dim3(kBlock), 0, stream_of(ctx, stream), n,
                       s, p, g);
    return launch_check("gravity_writeback");
}

int sphk_hydro_prep_dev(sphk_ctx *ctx, int n, sphk_state s, double *hinv, double *presc,
                        double *bswt_half, double *thrm, void *stream)
{
    int rc = pre(ctx, n, &s);
    if (rc || n == 0) return rc;
    if (!hinv || !presc || !bswt_half || !thrm) return set_error(SPHK_E_ARG, "hydro_prep: null");
    hipLaunchKernelGGL(hydro_prep_kernel, dim3(blocks(n)), dim3(kBlock), 0, stream_of(ctx, stream), n, s,
                       hinv, presc, bswt_half, thrm);
    return launch_check("hydro_prep");
}

int sphk_hydro_writeback_dev(sphk_ctx *ctx, int n, sphk_state s, sphk_phys p, sphk_hydro_out h,
                             void *stream)
{
    int rc = pre(ctx, n, &s);
    if (rc || n == 0) return rc;
    const void *const *f = reinterpret_cast<const void *const *>(&h);
    for (int k = 0; k < 9; k++)
        if (!f[k]) return set_error(SPHK_E_ARG, "hydro_writeback: null field %d", k);
    hipLaunchKernelGGL(hydro_writeback_kernel, dim3(blocks(n)), dim3(kBlock), 0, stream_of(ctx, stream), n,
                       s, p, h);
    return launch_check("hydro_writeback");
}

int sphk_predict_dev(sphk_ctx *ctx, int n, sphk_state s, double dt, void *stream)
{
    int rc = pre(ctx, n, &s);
    if (rc || n == 0) return rc;
    hipLaunchKernelGGL(predict_kernel, dim3(blocks(n)), dim3(kBlock), 0, stream_of(ctx, stream), n, s, dt);
    return launch_check("predict");
}

int sphk_correct_dev(sphk_ctx *ctx, int n, sphk_state s, double dt, void *stream)
{
    int rc = pre(ctx, n, &s);
    if (rc || n == 0) return rc;
    hipLaunchKernelGGL(correct_kernel, dim3(blocks(n)), dim3(kBlock), 0, stream_of(ctx, stream), n, s, dt);
    return launch_check("correct");
}

int sphk_diagnostics(sphk_ctx *ctx, int n, sphk_state s, sphk_phys p, double *out, void *stream)
{
    int rc = pre(ctx, n, &s);
    if (rc) return rc;
    if (!out) return set_error(SPHK_E_ARG, "diagnostics: null output");
    const int dev = sphk_device(ctx);
    const int nb = kDiagBlocks;
    if (g_diag.device != dev || g_diag.cap < nb) {
        if (g_diag.part) (void)hipFree(g_diag.part);
        if (g_diag.out) (void)hipFree(g_diag.out);
        g_diag = DiagScratch();
        if (hipMalloc(&g_diag.part, sizeof(double) * 6 * nb) != hipSuccess ||
            hipMalloc(&g_diag.out, sizeof(double) * 6) != hipSuccess)
            return set_error(SPHK_E_DEVICE, "diagnostics: hipMalloc failed");
        g_diag.cap = nb;
        g_diag.device = dev;
    }
    hipStream_t st = stream_of(ctx, stream);
    hipLaunchKernelGGL(diag_partial_kernel, dim3(nb), dim3(kBlock), 0, st, n, s, p, g_diag.part);
    if ((rc = launch_check("diagnostics"))) return rc;
    hipLaunchKernelGGL(diag_final_kernel, dim3(1), dim3(kBlock), 0, st, nb, g_diag.part, g_diag.out);
    if ((rc = launch_check("diagnostics"))) return rc;
    if (hipMemcpyAsync(out, g_diag.out, sizeof(double) * 6, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return set_error(SPHK_E_DEVICE, "diagnostics: copy-back failed");
    return SPHK_OK;
}

int sphk_diagnostics_static(sphk_ctx *ctx, int n, sphk_state s, sphk_phys p, double *out_dev, void *stream)
{
    int rc = pre(ctx, n, &s);
    if (rc) return rc;
    if (!out_dev) return set_error(SPHK_E_ARG, "diagnostics_static: null output");
    double *part = nullptr;
    if ((rc = scratch_of_current_device(&part))) return rc;
    hipStream_t st = stream_of(ctx, stream);
    // the same kernels and grid as sphk_diagnostics: bitwise the same sums;
    // every partial is written before it is read, nothing needs clearing
    hipLaunchKernelGGL(diag_partial_kernel, dim3(kDiagBlocks), dim3(kBlock), 0, st, n, s, p, part);
    if ((rc = launch_check("diagnostics_static"))) return rc;
    hipLaunchKernelGGL(diag_final_kernel, dim3(1), dim3(kBlock), 0, st, kDiagBlocks, part, out_dev);
    return launch_check("diagnostics_static");
}

int sphk_unbound_mass_dev(sphk_ctx *ctx, int n, sphk_state s, const int32_t *istar, int star, double k,
                          double *out, void *stream)
{
    int rc = pre(ctx, n, &s);
    if (rc) return rc;
    if (!out) return set_error(SPHK_E_ARG, "unbound_mass: null output");
    const int nb = 1024;
    if ((rc = diag_scratch(g_unbound, sphk_device(ctx), nb, 12))) return rc;
    hipStream_t st = stream_of(ctx, stream);
    hipLaunchKernelGGL(unbound_vcg_partial, dim3(nb), dim3(kBlock), 0, st, n, s, istar, star, g_unbound.part);
    hipLaunchKernelGGL(sum_final_kernel, dim3(1), dim3(kBlock), 0, st, nb, g_unbound.part, g_unbound.out);
    hipLaunchKernelGGL(unbound_mass_partial, dim3(nb), dim3(kBlock), 0, st, n, s, istar, star, k, g_unbound.out,
                       g_unbound.part);
    hipLaunchKernelGGL(sum_final_kernel, dim3(1), dim3(kBlock), 0, st, nb, g_unbound.part, g_unbound.out + 6);
    if ((rc = launch_check("unbound_mass"))) return rc;
    double h[12];
    if (hipMemcpyAsync(h, g_unbound.out, sizeof h, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return set_error(SPHK_E_DEVICE, "unbound_mass: copy-back failed");
    out[0] = h[1] / h[0];
    out[1] = h[2] / h[0];
    out[2] = h[3] / h[0];
    out[3] = h[6];
    out[4] = h[7];
    out[5] = h[8];
    return SPHK_OK;
}

}  // extern "C"
