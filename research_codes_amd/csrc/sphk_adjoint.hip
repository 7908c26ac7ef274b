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

// ---------------------------------------------------------------------------
// Hydro functor VJP (FP64).  The symmetric lists hold every interacting pair
// in both directions (r < max(H_i, H_j) puts j in the list of i's walk and i
// in the list of j's walk; pairs beyond both supports contribute exactly
// zero, with zero derivatives, as the kernel is C2 at q = 1).  So the
// cotangent of a particle's inputs is the sum, over its own list, of the
// i-side derivative of its pair terms and the j-side derivative of its
// neighbours' pair terms: one register accumulator per lane, no atomics,
// and a fixed summation order (bitwise reproducible).
// ---------------------------------------------------------------------------
struct HP {           // one particle's hydro inputs (hinv^5 derived)
    double x, y, z, vx, vy, vz, h, h5, p, pc, b, d, c, al, au, u, et, m;
    int32_t id;
};
struct HCot {         // cotangents of that particle's functor outputs
    double ax, ay, az, ud, df, gx, gy, gz;
};
constexpr int kHG = 16;   // x y z vx vy vz hinv pres presc bswt dens vsnd alph alphu thrm eta

__device__ __forceinline__ HP load_hp(const sphk_hydro_src &s, int k)
{
    const double h = s.hinv[k];
    const double h2 = h * h;
    return HP{s.x[k], s.y[k], s.z[k], s.vx[k], s.vy[k], s.vz[k], h, h2 * h2 * h, s.pres[k], s.presc[k],
              s.bswt[k], s.dens[k], s.vsnd[k], s.alph[k], s.alphu[k], s.thrm[k], s.eta[k], s.m[k], s.id[k]};
}

__device__ __forceinline__ HCot load_cot(const sphk_hydro_out &c, int k)
{
    return HCot{c.ax[k], c.ay[k], c.az[k], c.udot[k], c.diffu[k], c.gx[k], c.gy[k], c.gz[k]};
}

// Reverse of the pair term of particle a (the functor's i) with b (its j),
// hdr_hydro.hpp:352-399 in the closed form of the forward kernel, weighted
// by a's output cotangents ca.  ASIDE: add the derivative with respect to
// a's inputs to g, else with respect to b's.
template <bool ASIDE>
__device__ __forceinline__ void hydro_pair_vjp(const HP &a, const HP &b, const HCot &ca, double (&g)[kHG])
{
    if (a.id == b.id) return;
    constexpr double k5 = -5. * kCeff1;
    const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    const double dvx = a.vx - b.vx, dvy = a.vy - b.vy, dvz = a.vz - b.vz;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double ri = 1. / __builtin_sqrt(r2);
    const double r = r2 * ri;
    const double qa = r * a.h, qb = r * b.h;
    const double ta = qa < 1. ? 1. - qa : 0., tb = qb < 1. ? 1. - qb : 0.;
    if (ta == 0. && tb == 0.) return;   // beyond both supports: value and derivatives vanish
    const double Fa = a.h5 * (ta * ta * ta), Fb = b.h5 * (tb * tb * tb);
    const double mk = k5 * b.m;
    const double kri = mk * (Fa + Fb);
    const double ka = kri * r;
    const double rv = dx * dvx + dy * dvy + dz * dvz;
    const double w = rv * ri;
    const double w0 = w < 0. ? w : 0.;
    const double vs = a.c + b.c - 3. * w0;
    const double rhi = 1. / (a.d + b.d);
    const double bs = a.b + b.b, as = a.al + b.al;
    const double av0 = bs * as * vs * w0;
    const double Pt = a.pc + b.pc - 0.5 * av0 * rhi;
    const double tpair = Pt * kri;
    const double dP = a.p - b.p;
    const double vsu2 = __builtin_fabs(dP) * rhi * 2.;
    const double vsu = vsu2 > 0. ? __builtin_sqrt(vsu2) : 0.;
    const double du = a.u - b.u;
    const double aus = a.au + b.au;
    const double Y = 0.25 * av0 * w0 - aus * vsu * du;
    const double X = a.pc * w - rhi * Y;
    const double E = a.et * Fa + b.et * Fb;
    const double dg = mk * E;
    // outputs of a: ax = -0.5 sum tpair d, udot = 0.5 sum ka X,
    // diffu = 2 sum rhi du kri, g = 0.5 sum dg d
    const double Ad = ca.ax * dx + ca.ay * dy + ca.az * dz;
    const double Gd = ca.gx * dx + ca.gy * dy + ca.gz * dz;
    double d_b[3] = {-0.5 * tpair * ca.ax + 0.5 * dg * ca.gx, -0.5 * tpair * ca.ay + 0.5 * dg * ca.gy,
                     -0.5 * tpair * ca.az + 0.5 * dg * ca.gz};
    const double tp_b = -0.5 * Ad;
    const double ka_b = 0.5 * ca.ud * X;
    const double X_b = 0.5 * ca.ud * ka;
    double kri_b = 2. * ca.df * rhi * du;
    double rhi_b = 2. * ca.df * du * kri;
    double du_b = 2. * ca.df * rhi * kri;
    const double dg_b = 0.5 * Gd;
    // X = a.pc w - rhi Y
    double apc_b = X_b * w;
    double w_b = X_b * a.pc;
    rhi_b -= X_b * Y;
    const double Y_b = -X_b * rhi;
    // Y = av0 w0 / 4 - aus vsu du
    double av0_b = 0.25 * w0 * Y_b;
    double w0_b = 0.25 * av0 * Y_b;
    const double aus_b = -vsu * du * Y_b;
    const double vsu_b = -aus * du * Y_b;
    du_b -= aus * vsu * Y_b;
    // tpair = Pt kri, Pt = a.pc + b.pc - av0 rhi / 2
    const double Pt_b = tp_b * kri;
    kri_b += tp_b * Pt;
    apc_b += Pt_b;
    const double bpc_b = Pt_b;
    av0_b -= 0.5 * rhi * Pt_b;
    rhi_b -= 0.5 * av0 * Pt_b;
    // ka = kri r
    kri_b += ka_b * r;
    double r_b = ka_b * kri;
    // dg = mk (a.et Fa + b.et Fb), kri = mk (Fa + Fb)
    const double E_b = dg_b * mk;
    const double aet_b = E_b * Fa, bet_b = E_b * Fb;
    const double Fa_b = E_b * a.et + kri_b * mk;
    const double Fb_b = E_b * b.et + kri_b * mk;
    // av0 = bs as vs w0
    const double bs_b = av0_b * as * vs * w0;
    const double as_b = av0_b * bs * vs * w0;
    const double vs_b = av0_b * bs * as * w0;
    w0_b += av0_b * bs * as * vs;
    // vs = a.c + b.c - 3 w0; w0 = min(w, 0)
    w0_b -= 3. * vs_b;
    if (w < 0.) w_b += w0_b;
    // vsu = sqrt(|dP| rhi 2)
    const double vsu2_b = vsu2 > 0. ? vsu_b * 0.5 / vsu : 0.;
    const double dP_b = vsu2_b * 2. * rhi * (dP > 0. ? 1. : (dP < 0. ? -1. : 0.));
    rhi_b += vsu2_b * 2. * __builtin_fabs(dP);
    // rhi = 1 / (a.d + b.d)
    const double dd_b = -rhi_b * rhi * rhi;
    // w = rv ri, rv = d . dv
    const double rv_b = w_b * ri;
    double ri_b = w_b * rv;
    double dv_b[3] = {rv_b * dx, rv_b * dy, rv_b * dz};
    d_b[0] += rv_b * dvx;
    d_b[1] += rv_b * dvy;
    d_b[2] += rv_b * dvz;
    // F = h^5 t^3, t = 1 - r h (inside the support)
    const double ta_b = Fa_b * 3. * a.h5 * ta * ta, tb_b = Fb_b * 3. * b.h5 * tb * tb;
    double ah_b = Fa_b * 5. * (a.h * a.h) * (a.h * a.h) * (ta * ta * ta) - ta_b * r;
    double bh_b = Fb_b * 5. * (b.h * b.h) * (b.h * b.h) * (tb * tb * tb) - tb_b * r;
    r_b -= ta_b * a.h + tb_b * b.h;
    // ri = 1 / r, r = |d|
    r_b -= ri_b * ri * ri;
    d_b[0] += r_b * ri * dx;
    d_b[1] += r_b * ri * dy;
    d_b[2] += r_b * ri * dz;
    if (ASIDE) {
        g[0] += d_b[0];
        g[1] += d_b[1];
        g[2] += d_b[2];
        g[3] += dv_b[0];
        g[4] += dv_b[1];
        g[5] += dv_b[2];
        g[6] += ah_b;
        g[7] += dP_b;
        g[8] += apc_b;
        g[9] += bs_b;
        g[10] += dd_b;
        g[11] += vs_b;
        g[12] += as_b;
        g[13] += aus_b;
        g[14] += du_b;
        g[15] += aet_b;
    } else {
        g[0] -= d_b[0];
        g[1] -= d_b[1];
        g[2] -= d_b[2];
        g[3] -= dv_b[0];
        g[4] -= dv_b[1];
        g[5] -= dv_b[2];
        g[6] += bh_b;
        g[7] -= dP_b;
        g[8] += bpc_b;
        g[9] += bs_b;
        g[10] += dd_b;
        g[11] += vs_b;
        g[12] += as_b;
        g[13] += aus_b;
        g[14] -= du_b;
        g[15] += bet_b;
    }
}

struct HRecV {        // a staged j-record: inputs and its output cotangents
    HP p;
    HCot c;
};

__global__ __launch_bounds__(kWave) void hydro_vjp_kernel(int n_walk, const int32_t *__restrict__ i_off,
                                                          const int32_t *__restrict__ j_off,
                                                          const int32_t *__restrict__ j_idx, sphk_hydro_src src,
                                                          sphk_hydro_out cot, sphk_hydro_grad grad)
{
    __shared__ HRecV sp[kTile];
    __shared__ double red[kWave];
    if ((int)blockIdx.x >= n_walk) return;
    const int w = xcd_walk_adj(blockIdx.x, n_walk);
    const int lane = threadIdx.x;
    const int ib = i_off[w], ie = i_off[w + 1];
    const int kb = j_off[w], ke = j_off[w + 1];
    for (int c0 = ib; c0 < ie; c0 += kWave) {
        const int ni = min(kWave, ie - c0);
        const int S = kWave / ni, il = lane % ni;
        const int slice = lane / ni < S ? lane / ni : kTile;   // idle lanes sweep nothing
        const int i = c0 + il;
        const HP P = load_hp(src, i);
        const HCot cp = load_cot(cot, i);
        double g[kHG];
#pragma unroll
        for (int q = 0; q < kHG; q++) g[q] = 0.;
        for (int t0 = kb; t0 < ke; t0 += kTile) {
            const int nt = min(kTile, ke - t0);
            __syncthreads();
            if (lane < nt) {
                const int j = j_idx ? j_idx[t0 + lane] : t0 + lane;
                sp[lane] = HRecV{load_hp(src, j), load_cot(cot, j)};
            } else {   // inert: far away, massless, no cotangent
                sp[lane] = HRecV{HP{1e15, 1e15, 1e15, 0., 0., 0., 1., 1., 0., 0., 0., 1., 0., 0., 0., 0., 0., 0., -1},
                                 HCot{0., 0., 0., 0., 0., 0., 0., 0.}};
            }
            __syncthreads();
            for (int k = slice; k < kTile; k += S) {
                const HRecV &Q = sp[k];
                hydro_pair_vjp<true>(P, Q.p, cp, g);
                hydro_pair_vjp<false>(Q.p, P, Q.c, g);
            }
        }
        // per i: the S slices' partials in slice order, then one plain add
        // per field (each particle is the i of exactly one walk lane group)
        double *outs[kHG] = {grad.x, grad.y, grad.z, grad.vx, grad.vy, grad.vz, grad.hinv, grad.pres,
                             grad.presc, grad.bswt, grad.dens, grad.vsnd, grad.alph, grad.alphu, grad.thrm,
                             grad.eta};
#pragma unroll
        for (int q = 0; q < kHG; q++) {
            __syncthreads();
            red[lane] = g[q];
            __syncthreads();
            if (lane < ni) {
                double s = 0.;
                for (int t = 0; t < S; t++) s += red[lane + ni * t];
                outs[q][i] += s;
            }
        }
    }
}

}  // namespace
}  // namespace sphk

extern "C" int sphk_hydro_vjp_dev(sphk_ctx *ctx, int n_walk, const int32_t *i_off, const int32_t *j_off,
                                  const int32_t *j_idx, sphk_hydro_src src, sphk_hydro_out cot,
                                  sphk_hydro_grad grad, void *stream)
{
    using sphk::set_error;
    if (!ctx) return set_error(SPHK_E_ARG, "hydro_vjp: null context");
    if (n_walk < 0 || (n_walk > 0 && (!i_off || !j_off))) return set_error(SPHK_E_ARG, "hydro_vjp: bad walks");
    if (n_walk == 0) return SPHK_OK;
    const void *in[] = {src.id, src.m, src.x, src.y, src.z, src.vx, src.vy, src.vz, src.hinv, src.pres,
                        src.presc, src.bswt, src.dens, src.vsnd, src.alph, src.alphu, src.thrm, src.eta,
                        cot.ax, cot.ay, cot.az, cot.udot, cot.diffu, cot.gx, cot.gy, cot.gz,
                        grad.x, grad.y, grad.z, grad.vx, grad.vy, grad.vz, grad.hinv, grad.pres, grad.presc,
                        grad.bswt, grad.dens, grad.vsnd, grad.alph, grad.alphu, grad.thrm, grad.eta};
    for (const void *q : in)
        if (!q) return set_error(SPHK_E_ARG, "hydro_vjp: null pointer");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(sphk::hydro_vjp_kernel, dim3(n_walk), dim3(sphk::kWave), 0, st, n_walk, i_off, j_off, j_idx,
                       src, cot, grad);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(SPHK_E_DEVICE, "hydro_vjp: %s", hipGetErrorString(e));
    return SPHK_OK;
}

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
