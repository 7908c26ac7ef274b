// sphk_kernels.hip -- gfx950 kernels for the SPH interaction functors.
//
// Launch shape (all three kernels): one 64-lane workgroup (one wavefront)
// per walk.  A walk's i-particles are processed in chunks of <= 64; inside a
// chunk of ni particles, P = pow2ceil(ni) lanes hold distinct i's and the
// 64/P lane groups split the walk's j-list between them (entry k of a tile
// goes to slice k mod S).  j-entries are staged through LDS in tiles of 64:
// lane k loads entry k (SoA gathers through j_idx, coalesced when the list is
// a run of consecutive particles), then every lane sweeps its slice of the
// tile.  Per-i partial sums are combined across slices with a butterfly of
// cross-lane shuffles.
//
// Precision.  Positions are FP64 everywhere; each pair displacement is
// formed in FP64.  MIXED=false evaluates the pair term in FP64; MIXED=true
// casts the displacement to FP32, evaluates the pair term in FP32, sums a
// tile's partials in FP32 and folds them into FP64 per-i totals.  The three
// smoothing-length passes of the density functor are FP64 in both modes.

#include "sphk_device.hpp"
#include "sphk_internal.hpp"

namespace sphk {

// ---------------------------------------------------------------------------
// Density: scalar calcDensity, sph_yamauchi/run/hdr_density.hpp:254-363.
// ---------------------------------------------------------------------------

// q = sqrt(r2) * hinv evaluated exactly as the reference's FP64 code does
// (separate IEEE roundings, no contraction) -- used to decide q < 1 for pairs
// that lie within a few ulps of the kernel edge.
__device__ __attribute__((noinline)) double exact_q(double dx, double dy, double dz, double hinv)
{
    const double r2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
    return __dmul_rn(__builtin_sqrt(r2), hinv);
}

template <bool MIXED>
__global__ __launch_bounds__(kWave) void density_kernel(
    int n_walk, const int32_t *__restrict__ i_off, const int32_t *__restrict__ j_off,
    const int32_t *__restrict__ j_idx, sphk_dens_src pi, sphk_dens_src pj, sphk_dens_out out,
    double eta_ksrh, double hmax)
{
    using R = typename std::conditional<MIXED, float, double>::type;
    __shared__ double sx[kTile], sy[kTile], sz[kTile], sm[kTile];
    __shared__ R svx[kTile], svy[kTile], svz[kTile], smr[kTile];
    __shared__ int32_t sid[kTile];

    const int w = blockIdx.x;
    if (w >= n_walk) return;
    const int lane = threadIdx.x;
    const int ib = i_off[w], ie = i_off[w + 1];
    const int jb = j_off[w], je = j_off[w + 1];

    for (int c0 = ib; c0 < ie; c0 += kWave) {
        const LaneMap L(min(kWave, ie - c0), lane);
        const int i = c0 + (L.active ? L.il : 0);
        const int32_t id_i = pi.id[i];
        const double px = pi.x[i], py = pi.y[i], pz = pi.z[i];
        const double m_i = pi.m[i];
        double hs = pi.ksr[i];

        double dens = 0.;
        long long np = 0;

        // -- three fixed-point passes: rho(H_k), n(H_k), H_{k+1} (hdr_density.hpp:264-302)
        for (int rep = 0; rep < 3; rep++) {
            const double hi = 1. / hs;
            const double hi3 = hi * hi * hi;
            double rh = 0.;
            int nj = 0;
            for (int t0 = jb; t0 < je; t0 += kTile) {
                const int nt = min(kTile, je - t0);
                __syncthreads();
                if (lane < nt) {
                    const int j = j_idx ? j_idx[t0 + lane] : t0 + lane;
                    sx[lane] = pj.x[j];
                    sy[lane] = pj.y[j];
                    sz[lane] = pj.z[j];
                    sm[lane] = pj.m[j];
                }
                __syncthreads();
                for (int k = L.slice; k < nt; k += L.S) {
                    const double dx = px - sx[k];
                    const double dy = py - sy[k];
                    const double dz = pz - sz[k];
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    const double r1 = r2 > 0. ? r2 * rsqrt_f64(r2) : 0.;
                    double q = r1 * hi;
                    if (__builtin_fabs(q - 1.) < 1e-13) q = exact_q(dx, dy, dz, hi);
                    rh += sm[k] * hi3 * kernel0th(q);
                    nj += (q < 1.) ? 1 : 0;
                }
            }
            rh = slice_sum(rh, L.P);
            nj = slice_sum(nj, L.P);
            double hnew = eta_ksrh * pow(m_i / rh, 1. / 3.);
            hnew = hnew < hmax ? hnew : hmax;
            dens = rh;
            np = nj;
            hs = hnew;
        }

        // -- grad-h, divergence, curl with H_3 (hdr_density.hpp:304-349)
        const double vx = pi.vx[i], vy = pi.vy[i], vz = pi.vz[i];
        const double hi_d = 1. / hs;
        const R hi = (R)hi_d;
        const R hi4 = (R)(hi_d * (hi_d * hi_d * hi_d));
        const R rvx = (R)vx, rvy = (R)vy, rvz = (R)vz;
        double grdh = 0., divv = 0., rotx = 0., roty = 0., rotz = 0.;
        for (int t0 = jb; t0 < je; t0 += kTile) {
            const int nt = min(kTile, je - t0);
            __syncthreads();
            if (lane < nt) {
                const int j = j_idx ? j_idx[t0 + lane] : t0 + lane;
                sx[lane] = pj.x[j];
                sy[lane] = pj.y[j];
                sz[lane] = pj.z[j];
                smr[lane] = (R)pj.m[j];
                svx[lane] = (R)pj.vx[j];
                svy[lane] = (R)pj.vy[j];
                svz[lane] = (R)pj.vz[j];
                sid[lane] = pj.id[j];
            }
            __syncthreads();
            R g = 0, dv = 0, rx = 0, ry = 0, rz = 0;
            for (int k = L.slice; k < nt; k += L.S) {
                const R dpx = (R)(px - sx[k]);
                const R dpy = (R)(py - sy[k]);
                const R dpz = (R)(pz - sz[k]);
                const R dvx = rvx - svx[k];
                const R dvy = rvy - svy[k];
                const R dvz = rvz - svz[k];
                const R r2 = dpx * dpx + dpy * dpy + dpz * dpz;
                const R ri = (id_i != sid[k]) ? rsqrt_r(r2) : R(0);
                const R r1 = r2 * ri;
                const R q = r1 * hi;
                const R kw0 = kernel0th(q);
                const R kw1 = kernel1st(q);
                const R m_j = smr[k];
                g -= (R(3) * kw0 + q * kw1) * hi4 * m_j;
                const R dw = m_j * hi4 * kw1 * ri;
                const R dwx = dw * dpx, dwy = dw * dpy, dwz = dw * dpz;
                dv -= dvx * dwx + dvy * dwy + dvz * dwz;
                rx += dvy * dwz - dvz * dwy;
                ry += dvz * dwx - dvx * dwz;
                rz += dvx * dwy - dvy * dwx;
            }
            grdh += (double)g;
            divv += (double)dv;
            rotx += (double)rx;
            roty += (double)ry;
            rotz += (double)rz;
        }
        grdh = slice_sum(grdh, L.P);
        divv = slice_sum(divv, L.P);
        rotx = slice_sum(rotx, L.P);
        roty = slice_sum(roty, L.P);
        rotz = slice_sum(rotz, L.P);

        // hdr_density.hpp:351-362
        if (L.active && L.slice == 0) {
            const double deni = 1. / dens;
            const double omgi = 1. / (1. + hs * deni * grdh / 3.);
            const double rot2 = rotx * rotx + roty * roty + rotz * rotz;
            double rotv = rot2 != 0. ? __builtin_sqrt(rot2) : 0.;
            rotv *= deni * omgi;
            divv *= deni * omgi;
            out.dens[i] = dens;
            out.np[i] = np;
            out.ksr[i] = hs;
            out.itr[i] = (hs > pi.rs[i]) ? 1 : 0;
            out.divv[i] = divv;
            out.rotv[i] = rotv;
            out.grdh[i] = omgi;
        }
    }
}

// ---------------------------------------------------------------------------
// Hydro: scalar calcHydro, sph_yamauchi/run/hdr_hydro.hpp:320-421.
// ---------------------------------------------------------------------------
template <bool MIXED>
__global__ __launch_bounds__(kWave) void hydro_kernel(
    int n_walk, const int32_t *__restrict__ i_off, const int32_t *__restrict__ j_off,
    const int32_t *__restrict__ j_idx, sphk_hydro_src pi, sphk_hydro_src pj, sphk_hydro_out out)
{
    using R = typename std::conditional<MIXED, float, double>::type;
    __shared__ double sx[kTile], sy[kTile], sz[kTile];
    __shared__ R svx[kTile], svy[kTile], svz[kTile], sm[kTile], shinv[kTile], spres[kTile],
        spresc[kTile], sbswt[kTile], sdens[kTile], svsnd[kTile], salph[kTile], salphu[kTile],
        sthrm[kTile], seta[kTile];
    __shared__ int32_t sid[kTile];

    const int w = blockIdx.x;
    if (w >= n_walk) return;
    const int lane = threadIdx.x;
    const int ib = i_off[w], ie = i_off[w + 1];
    const int jb = j_off[w], je = j_off[w + 1];

    for (int c0 = ib; c0 < ie; c0 += kWave) {
        const LaneMap L(min(kWave, ie - c0), lane);
        const int i = c0 + (L.active ? L.il : 0);
        const int32_t id_i = pi.id[i];
        const double px = pi.x[i], py = pi.y[i], pz = pi.z[i];
        const R vx = (R)pi.vx[i], vy = (R)pi.vy[i], vz = (R)pi.vz[i];
        const double hi_d = pi.hinv[i];
        const R hi = (R)hi_d;
        const R hi4 = (R)(hi_d * (hi_d * hi_d * hi_d));
        const R pp_i = (R)pi.pres[i], pc_i = (R)pi.presc[i], bs_i = (R)pi.bswt[i];
        const R dn_i = (R)pi.dens[i], cs_i = (R)pi.vsnd[i], al_i = (R)pi.alph[i];
        const R au_i = (R)pi.alphu[i], eg_i = (R)pi.thrm[i], et_i = (R)pi.eta[i];

        double ax = 0., ay = 0., az = 0., udot = 0., difu = 0., gx = 0., gy = 0., gz = 0.;
        R vsmx = 0;
        for (int t0 = jb; t0 < je; t0 += kTile) {
            const int nt = min(kTile, je - t0);
            __syncthreads();
            if (lane < nt) {
                const int j = j_idx ? j_idx[t0 + lane] : t0 + lane;
                sx[lane] = pj.x[j];
                sy[lane] = pj.y[j];
                sz[lane] = pj.z[j];
                svx[lane] = (R)pj.vx[j];
                svy[lane] = (R)pj.vy[j];
                svz[lane] = (R)pj.vz[j];
                sm[lane] = (R)pj.m[j];
                shinv[lane] = (R)pj.hinv[j];
                spres[lane] = (R)pj.pres[j];
                spresc[lane] = (R)pj.presc[j];
                sbswt[lane] = (R)pj.bswt[j];
                sdens[lane] = (R)pj.dens[j];
                svsnd[lane] = (R)pj.vsnd[j];
                salph[lane] = (R)pj.alph[j];
                salphu[lane] = (R)pj.alphu[j];
                sthrm[lane] = (R)pj.thrm[j];
                seta[lane] = (R)pj.eta[j];
                sid[lane] = pj.id[j];
            }
            __syncthreads();
            R tax = 0, tay = 0, taz = 0, tud = 0, tdf = 0, tgx = 0, tgy = 0, tgz = 0;
            for (int k = L.slice; k < nt; k += L.S) {
                const R dpx = (R)(px - sx[k]);
                const R dpy = (R)(py - sy[k]);
                const R dpz = (R)(pz - sz[k]);
                const R dvx = vx - svx[k];
                const R dvy = vy - svy[k];
                const R dvz = vz - svz[k];
                const R r2 = dpx * dpx + dpy * dpy + dpz * dpz;
                const R ri = (id_i != sid[k]) ? rsqrt_r(r2) : R(0);
                const R r1 = r2 * ri;
                const R q_i = r1 * hi;
                const R hj = shinv[k];
                const R q_j = r1 * hj;
                const R hi4_j = hj * (hj * hj * hj);
                const R dw_i = hi4 * kernel1st(q_i);
                const R dw_j = hi4_j * kernel1st(q_j);
                const R m_j = sm[k];
                const R ka = (dw_i + dw_j) * m_j;

                const R rv = dpx * dvx + dpy * dvy + dpz * dvz;
                const R w_ij = rv * ri;
                const R w0 = w_ij < R(0) ? w_ij : R(0);
                const R vs = cs_i + svsnd[k] - R(3) * w0;
                const R rhi = rcp_r(dn_i + sdens[k]);
                const R av0 = (bs_i + sbswt[k]) * (al_i + salph[k]) * vs * w0;
                vsmx = vsmx > vs ? vsmx : vs;

                const R ta = (pc_i + spresc[k] - R(0.5) * av0 * rhi) * ka * ri;
                tax -= ta * dpx;
                tay -= ta * dpy;
                taz -= ta * dpz;

                const R vsu2 = __builtin_fabs(pp_i - spres[k]) * rhi * R(2);
                const R vsu = vsu2 != R(0) ? vsu2 * rsqrt_r(vsu2) : R(0);
                const R du = eg_i - sthrm[k];
                tud += ka * (pc_i * w_ij - rhi * (R(0.25) * av0 * w0 - (au_i + salphu[k]) * vsu * du));
                tdf += rhi * du * ka * ri;

                const R dg = m_j * ri * (et_i * dw_i + seta[k] * dw_j);
                tgx += dg * dpx;
                tgy += dg * dpy;
                tgz += dg * dpz;
            }
            ax += (double)tax;
            ay += (double)tay;
            az += (double)taz;
            udot += (double)tud;
            difu += (double)tdf;
            gx += (double)tgx;
            gy += (double)tgy;
            gz += (double)tgz;
        }
        ax = slice_sum(ax, L.P);
        ay = slice_sum(ay, L.P);
        az = slice_sum(az, L.P);
        udot = slice_sum(udot, L.P);
        difu = slice_sum(difu, L.P);
        gx = slice_sum(gx, L.P);
        gy = slice_sum(gy, L.P);
        gz = slice_sum(gz, L.P);
        vsmx = slice_max(vsmx, L.P);
        // hdr_hydro.hpp:402-419
        if (L.active && L.slice == 0) {
            out.ax[i] = ax * 0.5;
            out.ay[i] = ay * 0.5;
            out.az[i] = az * 0.5;
            out.udot[i] = udot * 0.5;
            out.vsmx[i] = (double)vsmx;
            out.diffu[i] = difu * 2.;
            out.gx[i] = gx * 0.5;
            out.gy[i] = gy * 0.5;
            out.gz[i] = gz * 0.5;
        }
    }
}

// ---------------------------------------------------------------------------
// Gravity: scalar calcGravity<TJ>, sph_yamauchi/run/hdr_gravity.hpp:295-338,
// over the EPJ list and then the SPJ list of each walk.
// ---------------------------------------------------------------------------
template <typename R>
struct GravAcc {
    double ax = 0., ay = 0., az = 0., pot = 0., eta = 0.;
};

template <bool MIXED>
__device__ __forceinline__ void gravity_list(
    int lane, const LaneMap &L, int kb, int ke, const int32_t *__restrict__ idx, sphk_grav_src src,
    double px, double py, double pz, double e2_i, double *sx, double *sy, double *sz,
    typename std::conditional<MIXED, float, double>::type *sm,
    typename std::conditional<MIXED, float, double>::type *se, GravAcc<double> &acc)
{
    using R = typename std::conditional<MIXED, float, double>::type;
    const R e2 = (R)e2_i;
    for (int t0 = kb; t0 < ke; t0 += kTile) {
        const int nt = min(kTile, ke - t0);
        __syncthreads();
        if (lane < nt) {
            const int j = idx ? idx[t0 + lane] : t0 + lane;
            sx[lane] = src.x[j];
            sy[lane] = src.y[j];
            sz[lane] = src.z[j];
            sm[lane] = (R)src.m[j];
            se[lane] = (R)src.eps2[j];
        }
        __syncthreads();
        R tax = 0, tay = 0, taz = 0, tpt = 0, tet = 0;
        for (int k = L.slice; k < nt; k += L.S) {
            const R dpx = (R)(px - sx[k]);
            const R dpy = (R)(py - sy[k]);
            const R dpz = (R)(pz - sz[k]);
            const R r2 = dpx * dpx + dpy * dpy + dpz * dpz;
            const R rei_i = rsqrt_r(r2 + e2);
            const R rei3_i = rei_i * rei_i * rei_i;
            const R rei_j = rsqrt_r(r2 + se[k]);
            const R rei3_j = rei_j * rei_j * rei_j;
            const R m_j = sm[k];
            const R dg2 = m_j * (rei3_i + rei3_j);
            tpt -= m_j * (rei_i + rei_j);
            tax -= dpx * dg2;
            tay -= dpy * dg2;
            taz -= dpz * dg2;
            tet += m_j * rei3_i;
        }
        acc.ax += (double)tax;
        acc.ay += (double)tay;
        acc.az += (double)taz;
        acc.pot += (double)tpt;
        acc.eta += (double)tet;
    }
}

template <bool MIXED>
__global__ __launch_bounds__(kWave) void gravity_kernel(
    int n_walk, const int32_t *__restrict__ i_off, const int32_t *__restrict__ j_off,
    const int32_t *__restrict__ j_idx, sphk_grav_src pj, const int32_t *__restrict__ s_off,
    const int32_t *__restrict__ s_idx, sphk_grav_src ps, sphk_grav_src pi, sphk_grav_out out)
{
    using R = typename std::conditional<MIXED, float, double>::type;
    __shared__ double sx[kTile], sy[kTile], sz[kTile];
    __shared__ R sm[kTile], se[kTile];

    const int w = blockIdx.x;
    if (w >= n_walk) return;
    const int lane = threadIdx.x;
    const int ib = i_off[w], ie = i_off[w + 1];

    for (int c0 = ib; c0 < ie; c0 += kWave) {
        const LaneMap L(min(kWave, ie - c0), lane);
        const int i = c0 + (L.active ? L.il : 0);
        const double px = pi.x[i], py = pi.y[i], pz = pi.z[i], e2 = pi.eps2[i];
        GravAcc<double> acc;
        gravity_list<MIXED>(lane, L, j_off[w], j_off[w + 1], j_idx, pj, px, py, pz, e2, sx, sy, sz,
                            sm, se, acc);
        if (s_off)
            gravity_list<MIXED>(lane, L, s_off[w], s_off[w + 1], s_idx, ps, px, py, pz, e2, sx, sy,
                                sz, sm, se, acc);
        acc.ax = slice_sum(acc.ax, L.P);
        acc.ay = slice_sum(acc.ay, L.P);
        acc.az = slice_sum(acc.az, L.P);
        acc.pot = slice_sum(acc.pot, L.P);
        acc.eta = slice_sum(acc.eta, L.P);
        if (L.active && L.slice == 0) {
            out.ax[i] = acc.ax * 0.5;
            out.ay[i] = acc.ay * 0.5;
            out.az[i] = acc.az * 0.5;
            out.pot[i] = acc.pot * 0.5;
            out.eta[i] = acc.eta;
        }
    }
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
hipError_t launch_density(hipStream_t st, int n_walk, const int32_t *i_off, const int32_t *j_off,
                          const int32_t *j_idx, const sphk_dens_src &pi, const sphk_dens_src &pj,
                          const sphk_dens_out &out, double eta_ksrh, double hmax, bool mixed)
{
    if (n_walk <= 0) return hipSuccess;
    if (mixed)
        hipLaunchKernelGGL(density_kernel<true>, dim3(n_walk), dim3(kWave), 0, st, n_walk, i_off,
                           j_off, j_idx, pi, pj, out, eta_ksrh, hmax);
    else
        hipLaunchKernelGGL(density_kernel<false>, dim3(n_walk), dim3(kWave), 0, st, n_walk, i_off,
                           j_off, j_idx, pi, pj, out, eta_ksrh, hmax);
    return hipGetLastError();
}

hipError_t launch_hydro(hipStream_t st, int n_walk, const int32_t *i_off, const int32_t *j_off,
                        const int32_t *j_idx, const sphk_hydro_src &pi, const sphk_hydro_src &pj,
                        const sphk_hydro_out &out, bool mixed)
{
    if (n_walk <= 0) return hipSuccess;
    if (mixed)
        hipLaunchKernelGGL(hydro_kernel<true>, dim3(n_walk), dim3(kWave), 0, st, n_walk, i_off,
                           j_off, j_idx, pi, pj, out);
    else
        hipLaunchKernelGGL(hydro_kernel<false>, dim3(n_walk), dim3(kWave), 0, st, n_walk, i_off,
                           j_off, j_idx, pi, pj, out);
    return hipGetLastError();
}

hipError_t launch_gravity(hipStream_t st, int n_walk, const int32_t *i_off, const int32_t *j_off,
                          const int32_t *j_idx, const sphk_grav_src &pj, const int32_t *s_off,
                          const int32_t *s_idx, const sphk_grav_src &ps, const sphk_grav_src &pi,
                          const sphk_grav_out &out, bool mixed)
{
    if (n_walk <= 0) return hipSuccess;
    if (mixed)
        hipLaunchKernelGGL(gravity_kernel<true>, dim3(n_walk), dim3(kWave), 0, st, n_walk, i_off,
                           j_off, j_idx, pj, s_off, s_idx, ps, pi, out);
    else
        hipLaunchKernelGGL(gravity_kernel<false>, dim3(n_walk), dim3(kWave), 0, st, n_walk, i_off,
                           j_off, j_idx, pj, s_off, s_idx, ps, pi, out);
    return hipGetLastError();
}

}  // namespace sphk
