// sphk_adjoint.hip -- vector-Jacobian products of the three functors in
// FP64: gravity (uniform softening, the reverse of gravity_pair_uniform),
// hydro and density.
//
// Layout as the forward kernels: one wave per walk, S = floor(64/ni)
// j-slices, 64-record LDS tiles.  The i-side cotangent of a lane stays in
// registers and is combined across slices through LDS at the end of the
// walk.  The j-side: gravity sweeps each tile a second time with one record
// per lane; density collects it per tile in LDS (atomic adds); hydro needs
// none (its symmetric lists hold both pair orders).  Either way it leaves
// with one HBM atomic add per record and component.
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

// One i-particle's position and output cotangents, as the j-side sweep reads
// them from LDS.
struct VI {
    double x, y, z, ax, ay, az, pb, eb;
};

// Cotangent of record r's position from the pair (i, r): the reverse of the
// uniform-softening pair term, m_r (T abar - pbar w^3 d + 3 ebar w^5 d) with
// T = w^3 I - 3 w^5 d d^T and d = x_i - x_r; the i-side gets its negative.
__device__ __forceinline__ void vjp_pair(const VI &I, const VRec &r, double e2, double &cx, double &cy,
                                         double &cz)
{
    const double dx = I.x - r.x, dy = I.y - r.y, dz = I.z - r.z;
    const double s = dx * dx + dy * dy + dz * dz + e2;
    const double w = rsqrt_f64(s);
    const double w3 = w * w * w;
    const double w5 = w3 * w * w;
    const double dot = dx * I.ax + dy * I.ay + dz * I.az;
    // T abar = w3 abar - 3 w5 d (d . abar); then subtract pbar w3 d
    // and add 3 ebar w5 d (the eta sum's gradient is -3 m w5 d)
    const double f = 3. * w5 * (dot - I.eb) + I.pb * w3;
    cx = r.m * (w3 * I.ax - f * dx);
    cy = r.m * (w3 * I.ay - f * dy);
    cz = r.m * (w3 * I.az - f * dz);
}

// One list of a walk chunk, tile by tile, in two sweeps over the staged
// records: the i-side (lane = (i, slice), register accumulators) and the
// j-side (lane = record, looping over the chunk's i-particles in LDS, one
// global atomic per record and component).  Every pair is evaluated twice;
// summing the j-side in LDS instead needs atomics that the n_i lanes of a
// slice issue to the same record at once.
__device__ void vjp_list(int lane, int S, int slice, int ni, int kb, int ke, const int32_t *__restrict__ idx,
                         const sphk_grav_src &src, const VI &I, const VI *sI, double e2, VRec *sp, int32_t *sj,
                         double (&gi)[3], double *outx, double *outy, double *outz)
{
    for (int t0 = kb; t0 < ke; t0 += kTile) {
        const int nt = min(kTile, ke - t0);
        __syncthreads();
        if (lane < nt) {
            const int j = idx ? idx[t0 + lane] : t0 + lane;
            sp[lane] = VRec{src.x[j], src.y[j], src.z[j], src.m[j]};
            sj[lane] = j;
        }
        __syncthreads();
        for (int k = slice; k < nt; k += S) {
            double cx, cy, cz;
            vjp_pair(I, sp[k], e2, cx, cy, cz);
            gi[0] -= cx;
            gi[1] -= cy;
            gi[2] -= cz;
        }
        if (lane < nt) {
            const VRec r = sp[lane];
            double jx = 0., jy = 0., jz = 0.;
            for (int q = 0; q < ni; q++) {
                double cx, cy, cz;
                vjp_pair(sI[q], r, e2, cx, cy, cz);
                jx += cx;
                jy += cy;
                jz += cz;
            }
            const int j = sj[lane];
            atomicAdd(&outx[j], jx);
            atomicAdd(&outy[j], jy);
            atomicAdd(&outz[j], jz);
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
    __shared__ VI sI[kWave];
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
        const VI I{pi.x[i], pi.y[i], pi.z[i], abx[i], aby[i], abz[i], pbar[i], ebar ? ebar[i] : 0.};
        __syncthreads();   // sI of the previous chunk no longer read
        if (lane < ni) sI[lane] = I;
        double gi[3] = {0., 0., 0.};
        vjp_list(lane, S, slice, ni, j_off[w], j_off[w + 1], j_idx, pj, I, sI, e2, sp, sj, gi, gx, gy, gz);
        if (s_off)
            vjp_list(lane, S, slice, ni, s_off[w], s_off[w + 1], s_idx, ps, I, sI, e2, sp, sj, gi, nx, ny, nz);
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
            for (int k = slice; k < nt; k += S) {
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

// ---------------------------------------------------------------------------
// Density functor VJP (FP64), hdr_density.hpp:254-363.  Per i-particle the
// forward is recomputed: three h-passes rho_k = h_k^3 sum m W(r h_k),
// H_{k+1} = min(eta kappa (m / rho_k)^(1/3), hmax), with S_k = sum m r^2 t^3
// (so that d rho_k / d h_k = 3 rho_k / h_k - 5 C1 h_k^4 S_k), then the grad-h
// sums at H_3.  The reverse runs the outputs' cotangents back through
// Omega, div v and |curl v| (per particle), through the grad-h pair sums (one
// sweep), down the chain H_3 -> rho_2 -> H_2 -> rho_1 -> H_1 -> rho_0 -> H_0
// (per particle), and through the three h-passes' pair sums (one sweep with
// the three kernel weights).  Gather lists are not symmetric, so the j-side
// position and velocity cotangents collect in LDS per tile (atomic adds) and
// leave with one global FP64 atomic per record and component.
// ---------------------------------------------------------------------------
struct DRec {
    double x, y, z, m, vx, vy, vz;
    int32_t id;
};

// Per-i sums over the S slices (fixed order), visible to every lane of the i.
template <int N>
__device__ __forceinline__ void slice_total(double (&v)[N], double *red, int lane, int il, int ni, int S)
{
#pragma unroll
    for (int q = 0; q < N; q++) {
        __syncthreads();
        red[q * kWave + lane] = v[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < N; q++) {
        double s = 0.;
        for (int t = 0; t < S; t++) s += red[q * kWave + il + ni * t];
        v[q] = s;
    }
}

__global__ __launch_bounds__(kWave) void density_vjp_kernel(int n_walk, const int32_t *__restrict__ i_off,
                                                            const int32_t *__restrict__ j_off,
                                                            const int32_t *__restrict__ j_idx, sphk_dens_src src,
                                                            double eta_ksrh, double hmax, sphk_dens_cot cot,
                                                            sphk_dens_grad grad)
{
    constexpr double C0 = kCeff0, k5 = -5. * kCeff1;
    __shared__ DRec sp[kTile];
    __shared__ int32_t sj[kTile];
    __shared__ double jb[6 * kTile];
    __shared__ double red[6 * kWave];
    if ((int)blockIdx.x >= n_walk) return;
    const int w = xcd_walk_adj(blockIdx.x, n_walk);
    const int lane = threadIdx.x;
    const int ib = i_off[w], ie = i_off[w + 1];
    const int kb = j_off[w], ke = j_off[w + 1];

    // one sweep over the walk's list: f(record, k-th slot) per pair of this
    // lane's slice; with FLUSH the j-side cotangents in jb[0..5][k] go to grad
    auto sweep = [&](auto &&f, bool flush, int slice, int S) {
        for (int t0 = kb; t0 < ke; t0 += kTile) {
            const int nt = min(kTile, ke - t0);
            __syncthreads();
            if (lane < nt) {
                const int j = j_idx ? j_idx[t0 + lane] : t0 + lane;
                sp[lane] = DRec{src.x[j], src.y[j], src.z[j], src.m[j], src.vx[j], src.vy[j], src.vz[j], src.id[j]};
                sj[lane] = j;
            } else {
                sp[lane] = DRec{1e15, 1e15, 1e15, 0., 0., 0., 0., -1};
                sj[lane] = -1;
            }
            if (flush)
                for (int q = 0; q < 6; q++) jb[q * kTile + lane] = 0.;
            __syncthreads();
            for (int k = slice; k < nt; k += S) f(sp[k], k);
            if (flush) {
                __syncthreads();
                if (lane < nt) {
                    const int j = sj[lane];
                    double *dst[6] = {grad.x, grad.y, grad.z, grad.vx, grad.vy, grad.vz};
                    for (int q = 0; q < 6; q++) {
                        const double v = jb[q * kTile + lane];
                        if (v != 0.) atomicAdd(&dst[q][j], v);
                    }
                }
            }
        }
    };

    for (int c0 = ib; c0 < ie; c0 += kWave) {
        const int ni = min(kWave, ie - c0);
        const int S = kWave / ni, il = lane % ni;
        const int slice = lane / ni < S ? lane / ni : kTile;   // idle lanes sweep nothing
        const bool lead = lane < ni;
        const int i = c0 + il;
        const double px = src.x[i], py = src.y[i], pz = src.z[i];
        const double pvx = src.vx[i], pvy = src.vy[i], pvz = src.vz[i];
        const double mi = src.m[i];

        // -- forward: three h-passes (rho_k, S_k) and H_{k+1}
        double H[4], rho[3], Sk[3];
        bool capped[4] = {false, false, false, false};
        H[0] = src.ksr[i];
        for (int k = 0; k < 3; k++) {
            const double h = 1. / H[k];
            double v[2] = {0., 0.};
            sweep([&](const DRec &r, int) {
                const double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double q = __builtin_sqrt(r2) * h;
                const double t = q < 1. ? 1. - q : 0.;
                const double t2 = t * t;
                v[0] += r.m * (C0 * t2 * t2 * (1. + 4. * q));
                v[1] += r.m * r2 * (t2 * t);
            }, false, slice, S);
            slice_total<2>(v, red, lane, il, ni, S);
            rho[k] = v[0] * (h * h * h);
            Sk[k] = v[1];
            const double hn = eta_ksrh * cbrt_f64(mi / rho[k]);
            capped[k + 1] = !(hn < hmax);
            H[k + 1] = hn < hmax ? hn : hmax;
        }
        // -- forward grad-h sums at H_3
        const double h3 = 1. / H[3];
        const double h3_4 = (h3 * h3) * (h3 * h3), h3_5 = h3_4 * h3;
        double fg[5] = {0., 0., 0., 0., 0.};   // gsum, dvsum, rot x, y, z (dw folded in)
        sweep([&](const DRec &r, int) {
            const double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
            const double dvx = pvx - r.vx, dvy = pvy - r.vy, dvz = pvz - r.vz;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double q = __builtin_sqrt(r2) * h3;
            const double t = q < 1. ? 1. - q : 0.;
            const double t3 = t * t * t;
            const double dw = k5 * r.m * h3_5 * t3;
            fg[0] += r.m * t3 * (3. + q * (9. - 32. * q));
            fg[1] += (dvx * dx + dvy * dy + dvz * dz) * dw;
            fg[2] += (dvy * dz - dvz * dy) * dw;
            fg[3] += (dvz * dx - dvx * dz) * dw;
            fg[4] += (dvx * dy - dvy * dx) * dw;
        }, false, slice, S);
        slice_total<5>(fg, red, lane, il, ni, S);
        const double g = -C0 * h3_4 * fg[0];
        const double dvv = -fg[1];
        const double rho2 = rho[2];
        const double B = 1. + H[3] * g / (3. * rho2);
        const double Om = 1. / B;
        const double rotm2 = fg[2] * fg[2] + fg[3] * fg[3] + fg[4] * fg[4];
        const double rotm = rotm2 > 0. ? __builtin_sqrt(rotm2) : 0.;

        // -- reverse through the outputs (per particle)
        const double cD = cot.dens[i], cK = cot.ksr[i], cV = cot.divv[i], cR = cot.rotv[i], cG = cot.grdh[i];
        // a particle whose outputs carry no cotangent contributes nothing;
        // it is skipped outright, so that outputs it does not own (e.g. an
        // empty sum over a list built for other particles: 0 * inf) cannot
        // turn into NaN cotangents of its neighbours
        const bool live = cD != 0. || cK != 0. || cV != 0. || cR != 0. || cG != 0.;
        const double Omb = cG + (cV * dvv + cR * rotm) / rho2;
        double rho2b = cD - (cV * dvv + cR * rotm) * Om / (rho2 * rho2);
        const double A = Om / rho2;
        const double dvsb = -cV * A;   // d L / d dvsum (dv = -dvsum)
        double rb[3] = {0., 0., 0.};
        if (rotm > 0.)
            for (int q = 0; q < 3; q++) rb[q] = cR * A * fg[2 + q] / rotm;
        const double Bb = -Omb * Om * Om;
        double H3b = cK + Bb * g / (3. * rho2);
        const double gb = Bb * H[3] / (3. * rho2);
        rho2b -= Bb * H[3] * g / (3. * rho2 * rho2);
        const double gsb = -C0 * h3_4 * gb;            // d L / d gsum
        double h3b = gb * (-4. * C0 * h3 * h3 * h3 * fg[0]);

        // -- reverse grad-h sweep
        double gi[6] = {0., 0., 0., 0., 0., 0.};   // x, y, z, vx, vy, vz of i
        double hsum = 0.;
        sweep([&](const DRec &r, int k) {
            const double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
            const double dvx = pvx - r.vx, dvy = pvy - r.vy, dvz = pvz - r.vz;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double rr = __builtin_sqrt(r2);
            const double q = rr * h3;
            if (!live || !(q < 1.)) return;
            const double t = 1. - q;
            const double t2 = t * t, t3 = t2 * t;
            const double dw = k5 * r.m * h3_5 * t3;
            const double P = 3. + q * (9. - 32. * q), Pp = 9. - 64. * q;
            const double cx = dvy * dz - dvz * dy, cy = dvz * dx - dvx * dz, cz = dvx * dy - dvy * dx;
            const double Sdw = dvsb * (dvx * dx + dvy * dy + dvz * dz) + rb[0] * cx + rb[1] * cy + rb[2] * cz;
            // explicit d and dv terms
            double db[3] = {dw * (dvsb * dvx + (rb[1] * dvz - rb[2] * dvy)),
                            dw * (dvsb * dvy + (rb[2] * dvx - rb[0] * dvz)),
                            dw * (dvsb * dvz + (rb[0] * dvy - rb[1] * dvx))};
            const double vb[3] = {dw * (dvsb * dx + (dy * rb[2] - dz * rb[1])),
                                  dw * (dvsb * dy + (dz * rb[0] - dx * rb[2])),
                                  dw * (dvsb * dz + (dx * rb[1] - dy * rb[0]))};
            // d L / d q: through gsum's m t^3 P(q) and through dw (t = 1 - q)
            const double qb = gsb * r.m * t2 * (t * Pp - 3. * P) - 3. * Sdw * k5 * r.m * h3_5 * t2;
            hsum += Sdw * 5. * k5 * r.m * h3_4 * t3 + qb * rr;
            const double rb_ = qb * h3;
            if (rr > 0.) {
                const double f = rb_ / rr;
                db[0] += f * dx;
                db[1] += f * dy;
                db[2] += f * dz;
            }
            gi[0] += db[0];
            gi[1] += db[1];
            gi[2] += db[2];
            gi[3] += vb[0];
            gi[4] += vb[1];
            gi[5] += vb[2];
            atomicAdd(&jb[0 * kTile + k], -db[0]);
            atomicAdd(&jb[1 * kTile + k], -db[1]);
            atomicAdd(&jb[2 * kTile + k], -db[2]);
            atomicAdd(&jb[3 * kTile + k], -vb[0]);
            atomicAdd(&jb[4 * kTile + k], -vb[1]);
            atomicAdd(&jb[5 * kTile + k], -vb[2]);
        }, true, slice, S);
        double hs[1] = {hsum};
        slice_total<1>(hs, red, lane, il, ni, S);
        h3b += hs[0];
        H3b += h3b * (-h3 * h3);

        // -- reverse down the smoothing-length chain (per particle)
        double rhob[3];
        rhob[2] = rho2b + (capped[3] ? 0. : H3b * (-H[3] / (3. * rho2)));
        double Hb = 0.;
        for (int k = 2; k >= 0; k--) {
            const double h = 1. / H[k];
            const double drho_dh = 3. * rho[k] / h + k5 * (h * h) * (h * h) * Sk[k];
            Hb = rhob[k] * drho_dh * (-h * h);
            if (k > 0) rhob[k - 1] = capped[k] ? 0. : Hb * (-H[k] / (3. * rho[k - 1]));
        }
        // Hb is now d L / d H_0

        // -- reverse of the three h-passes' pair sums (positions only)
        double wk[3], hk[3];
        for (int k = 0; k < 3; k++) {
            hk[k] = 1. / H[k];
            const double h2 = hk[k] * hk[k];
            wk[k] = rhob[k] * k5 * (h2 * h2 * hk[k]);   // rho_bar h^5 (-5 C1)
        }
        sweep([&](const DRec &r, int k) {
            if (!live) return;
            const double dx = px - r.x, dy = py - r.y, dz = pz - r.z;
            const double rr = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
            double c = 0.;
            for (int q = 0; q < 3; q++) {
                const double qq = rr * hk[q];
                if (qq < 1.) {
                    const double t = 1. - qq;
                    c += wk[q] * (t * t * t);
                }
            }
            if (c == 0.) return;
            c *= r.m;
            gi[0] += c * dx;
            gi[1] += c * dy;
            gi[2] += c * dz;
            atomicAdd(&jb[0 * kTile + k], -c * dx);
            atomicAdd(&jb[1 * kTile + k], -c * dy);
            atomicAdd(&jb[2 * kTile + k], -c * dz);
        }, true, slice, S);

        // -- i-side totals
        slice_total<6>(gi, red, lane, il, ni, S);
        if (lead && live) {
            atomicAdd(&grad.x[i], gi[0]);
            atomicAdd(&grad.y[i], gi[1]);
            atomicAdd(&grad.z[i], gi[2]);
            atomicAdd(&grad.vx[i], gi[3]);
            atomicAdd(&grad.vy[i], gi[4]);
            atomicAdd(&grad.vz[i], gi[5]);
            grad.ksr[i] += Hb;
        }
    }
}

}  // namespace
}  // namespace sphk

extern "C" int sphk_density_vjp_dev(sphk_ctx *ctx, int n_walk, const int32_t *i_off, const int32_t *j_off,
                                    const int32_t *j_idx, sphk_dens_src src, sphk_dens_cot cot,
                                    sphk_dens_grad grad, void *stream)
{
    using sphk::set_error;
    if (!ctx) return set_error(SPHK_E_ARG, "density_vjp: null context");
    if (n_walk < 0 || (n_walk > 0 && (!i_off || !j_off))) return set_error(SPHK_E_ARG, "density_vjp: bad walks");
    if (n_walk == 0) return SPHK_OK;
    const void *in[] = {src.id, src.m, src.x, src.y, src.z, src.vx, src.vy, src.vz, src.ksr,
                        cot.dens, cot.ksr, cot.divv, cot.rotv, cot.grdh,
                        grad.x, grad.y, grad.z, grad.vx, grad.vy, grad.vz, grad.ksr};
    for (const void *q : in)
        if (!q) return set_error(SPHK_E_ARG, "density_vjp: null pointer");
    sphk_params p;
    const int rc = sphk_get_params(ctx, &p);
    if (rc) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(sphk::density_vjp_kernel, dim3(n_walk), dim3(sphk::kWave), 0, st, n_walk, i_off, j_off,
                       j_idx, src, p.eta * p.ksrh, p.hmax, cot, grad);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(SPHK_E_DEVICE, "density_vjp: %s", hipGetErrorString(e));
    return SPHK_OK;
}

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
