/*
 * sphk_oracle_simd.c -- the vector CPU functors of the CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY (bench.py's CPU baseline and tests/; nothing in
 * the product links it).  The reference's shipped CPU build (-DUSE_INTRINSICS,
 * -march=core-avx2) takes the vector branches of its functors: four
 * i-particles per 256-bit vector, every j broadcast to the four lanes, 1/x and
 * 1/sqrt(x) by the 4th-order refined estimates of vector_x86.hpp:501-534
 * (calcDensity hdr_density.hpp:82-242, calcHydro hdr_hydro.hpp:128-311,
 * calcGravity hdr_gravity.hpp:194-291).  These are those computations written
 * with AVX2 intrinsics, lane for lane the same IEEE operations in the same
 * order as the oracle's scalar functors with the refined reciprocals in place
 * of 1/x and 1/sqrt(x): orc_density_simd equals orc_density_avx2 bit for bit;
 * hydro and gravity agree with the scalar functors to the reciprocals'
 * rounding (tests/test_oracle_avx2.py).  They exist so that the CPU baseline
 * is timed on the reference's vector path, not on the scalar one.
 *
 * Build: oracle/Makefile (this file with -mavx2, no FMA contraction).
 */
#include <immintrin.h>
#include <math.h>
#include <stdint.h>

#include "sphk_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#define JROW(k) (j_idx ? j_idx[(k)] : (k))

/* vector_x86.hpp:501-511: x0 = rcp estimate, h = 1 - x x0, (1 + h)(1 + h^2) x0 */
static inline __m256d rcp4(__m256d x)
{
    const __m256d one = _mm256_set1_pd(1.);
    const __m256d x0 = _mm256_cvtps_pd(_mm_rcp_ps(_mm256_cvtpd_ps(x)));
    const __m256d h = _mm256_sub_pd(one, _mm256_mul_pd(x, x0));
    return _mm256_mul_pd(_mm256_mul_pd(_mm256_add_pd(one, h), _mm256_add_pd(one, _mm256_mul_pd(h, h))), x0);
}

/* vector_x86.hpp:520-534: x0 = rsqrt estimate, h = 1 - x x0^2,
 * x0 + (1/2 + (3/8 + (5/16 + 35/128 h) h) h) h x0 */
static inline __m256d rsqrt4(__m256d x)
{
    const __m256d x0 = _mm256_cvtps_pd(_mm_rsqrt_ps(_mm256_cvtpd_ps(x)));
    const __m256d h = _mm256_sub_pd(_mm256_set1_pd(1.), _mm256_mul_pd(_mm256_mul_pd(x, x0), x0));
    __m256d p = _mm256_add_pd(_mm256_set1_pd(0.3125), _mm256_mul_pd(_mm256_set1_pd(0.2734375), h));
    p = _mm256_add_pd(_mm256_set1_pd(0.375), _mm256_mul_pd(p, h));
    p = _mm256_add_pd(_mm256_set1_pd(0.5), _mm256_mul_pd(p, h));
    return _mm256_add_pd(x0, _mm256_mul_pd(_mm256_mul_pd(p, h), x0));
}

/* (1 - q)_+ as the scalar functor forms it: 1 - q where 1 - q > 0, else 0 */
static inline __m256d one_minus_q_pos(__m256d q)
{
    const __m256d t = _mm256_sub_pd(_mm256_set1_pd(1.), q);
    return _mm256_and_pd(t, _mm256_cmp_pd(t, _mm256_setzero_pd(), _CMP_GT_OQ));
}

/* W(q) = C0 (1-q)_+^4 (1 + 4q), hdr_kernel.hpp:246-250 */
static inline __m256d kernel0(__m256d q)
{
    const __m256d r = one_minus_q_pos(q);
    const __m256d r2 = _mm256_mul_pd(r, r);
    const __m256d opq = _mm256_add_pd(_mm256_set1_pd(1.), _mm256_mul_pd(_mm256_set1_pd(4.), q));
    return _mm256_mul_pd(_mm256_mul_pd(_mm256_mul_pd(_mm256_set1_pd(ORC_CEFF0), r2), r2), opq);
}

/* W'(q) = C1 [(1-q)_+^4 - (1-q)_+^3 (1 + 4q)], hdr_kernel.hpp:251-257 */
static inline __m256d kernel1(__m256d q)
{
    const __m256d r = one_minus_q_pos(q);
    const __m256d r2 = _mm256_mul_pd(r, r);
    const __m256d r3 = _mm256_mul_pd(r, r2);
    const __m256d r4 = _mm256_mul_pd(r2, r2);
    const __m256d opq = _mm256_add_pd(_mm256_set1_pd(1.), _mm256_mul_pd(_mm256_set1_pd(4.), q));
    return _mm256_mul_pd(_mm256_set1_pd(ORC_CEFF1), _mm256_sub_pd(r4, _mm256_mul_pd(r3, opq)));
}

int orc_density_simd_supported(void)
{
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx2") ? 1 : 0;
}

void orc_density_simd(int n_walk, const int32_t *i_off, const int32_t *j_off, const int32_t *j_idx,
                      orc_dens_src pi, orc_dens_src pj, double eta, double ksrh, double hmax,
                      orc_dens_out out, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int w = 0; w < n_walk; w++) {
        const int jb = j_off[w], je = j_off[w + 1];
        for (int i0 = i_off[w]; i0 < i_off[w + 1]; i0 += 4) {
            const int ni = (i_off[w + 1] - i0 < 4) ? i_off[w + 1] - i0 : 4;
            /* lanes past the walk's last i repeat it; their results are not written */
            double b_id[4], b_x[4], b_y[4], b_z[4], b_vx[4], b_vy[4], b_vz[4], b_hs[4];
            for (int l = 0; l < 4; l++) {
                const int i = i0 + (l < ni ? l : ni - 1);
                b_id[l] = (double)pi.id[i];
                b_x[l] = pi.x[i];
                b_y[l] = pi.y[i];
                b_z[l] = pi.z[i];
                b_vx[l] = pi.vx[i];
                b_vy[l] = pi.vy[i];
                b_vz[l] = pi.vz[i];
                b_hs[l] = pi.ksr[i];
            }
            const __m256d id_i = _mm256_loadu_pd(b_id);
            const __m256d px = _mm256_loadu_pd(b_x), py = _mm256_loadu_pd(b_y), pz = _mm256_loadu_pd(b_z);
            const __m256d vx = _mm256_loadu_pd(b_vx), vy = _mm256_loadu_pd(b_vy), vz = _mm256_loadu_pd(b_vz);
            __m256d hs = _mm256_loadu_pd(b_hs);
            double rh[4], hsv[4];

            /* three fixed-point passes, hdr_density.hpp:264-302 */
            for (int repeat = 0; repeat < 3; repeat++) {
                const __m256d hi = rcp4(hs);
                const __m256d hi3 = _mm256_mul_pd(_mm256_mul_pd(hi, hi), hi);
                __m256d rh_v = _mm256_setzero_pd(), nj_v = _mm256_setzero_pd();
                for (int k = jb; k < je; k++) {
                    const int j = JROW(k);
                    const __m256d dx = _mm256_sub_pd(px, _mm256_set1_pd(pj.x[j]));
                    const __m256d dy = _mm256_sub_pd(py, _mm256_set1_pd(pj.y[j]));
                    const __m256d dz = _mm256_sub_pd(pz, _mm256_set1_pd(pj.z[j]));
                    const __m256d r2 = _mm256_add_pd(_mm256_add_pd(_mm256_mul_pd(dx, dx), _mm256_mul_pd(dy, dy)),
                                                     _mm256_mul_pd(dz, dz));
                    const __m256d q = _mm256_mul_pd(_mm256_sqrt_pd(r2), hi);
                    const __m256d rhj = _mm256_mul_pd(_mm256_mul_pd(_mm256_set1_pd(pj.m[j]), hi3), kernel0(q));
                    rh_v = _mm256_add_pd(rh_v, rhj);
                    nj_v = _mm256_add_pd(nj_v, _mm256_and_pd(_mm256_cmp_pd(q, _mm256_set1_pd(1.), _CMP_LT_OQ),
                                                             _mm256_set1_pd(1.)));
                }
                double njv[4];
                _mm256_storeu_pd(rh, rh_v);
                _mm256_storeu_pd(njv, nj_v);
                /* hdr_density.hpp:288-299 */
                for (int l = 0; l < 4; l++) {
                    const int i = i0 + (l < ni ? l : ni - 1);
                    double h = eta * ksrh * pow(pi.m[i] / rh[l], 1. / 3.);
                    h = (h < hmax) ? h : hmax;
                    hsv[l] = h;
                    if (l < ni) {
                        out.dens[i] = rh[l];
                        out.np[i] = (int64_t)njv[l];
                        out.ksr[i] = h;
                        out.itr[i] = (h > pi.rs[i]) ? 1 : 0;
                    }
                }
                hs = _mm256_loadu_pd(hsv);
            }

            /* grad-h / divergence / curl pass, hdr_density.hpp:304-349 */
            const __m256d hi = rcp4(hs);
            const __m256d hi4 = _mm256_mul_pd(hi, _mm256_mul_pd(_mm256_mul_pd(hi, hi), hi));
            __m256d grdh = _mm256_setzero_pd(), divv = _mm256_setzero_pd();
            __m256d rotx = _mm256_setzero_pd(), roty = _mm256_setzero_pd(), rotz = _mm256_setzero_pd();
            for (int k = jb; k < je; k++) {
                const int j = JROW(k);
                const __m256d dpx = _mm256_sub_pd(px, _mm256_set1_pd(pj.x[j]));
                const __m256d dpy = _mm256_sub_pd(py, _mm256_set1_pd(pj.y[j]));
                const __m256d dpz = _mm256_sub_pd(pz, _mm256_set1_pd(pj.z[j]));
                const __m256d dvx = _mm256_sub_pd(vx, _mm256_set1_pd(pj.vx[j]));
                const __m256d dvy = _mm256_sub_pd(vy, _mm256_set1_pd(pj.vy[j]));
                const __m256d dvz = _mm256_sub_pd(vz, _mm256_set1_pd(pj.vz[j]));
                const __m256d r2 = _mm256_add_pd(_mm256_add_pd(_mm256_mul_pd(dpx, dpx), _mm256_mul_pd(dpy, dpy)),
                                                 _mm256_mul_pd(dpz, dpz));
                const __m256d self = _mm256_cmp_pd(id_i, _mm256_set1_pd((double)pj.id[j]), _CMP_EQ_OQ);
                const __m256d ri = _mm256_andnot_pd(self, rsqrt4(r2));
                const __m256d q = _mm256_mul_pd(_mm256_mul_pd(r2, ri), hi);
                const __m256d kw0 = kernel0(q), kw1 = kernel1(q);
                const __m256d mj = _mm256_set1_pd(pj.m[j]);
                const __m256d ghj = _mm256_add_pd(_mm256_mul_pd(_mm256_set1_pd(3.), kw0), _mm256_mul_pd(q, kw1));
                grdh = _mm256_sub_pd(grdh, _mm256_mul_pd(_mm256_mul_pd(ghj, hi4), mj));
                const __m256d dw = _mm256_mul_pd(_mm256_mul_pd(_mm256_mul_pd(mj, hi4), kw1), ri);
                const __m256d dwx = _mm256_mul_pd(dw, dpx), dwy = _mm256_mul_pd(dw, dpy);
                const __m256d dwz = _mm256_mul_pd(dw, dpz);
                divv = _mm256_sub_pd(divv, _mm256_mul_pd(dvx, dwx));
                divv = _mm256_sub_pd(divv, _mm256_mul_pd(dvy, dwy));
                divv = _mm256_sub_pd(divv, _mm256_mul_pd(dvz, dwz));
                rotx = _mm256_add_pd(rotx, _mm256_mul_pd(dvy, dwz));
                roty = _mm256_add_pd(roty, _mm256_mul_pd(dvz, dwx));
                rotz = _mm256_add_pd(rotz, _mm256_mul_pd(dvx, dwy));
                rotx = _mm256_sub_pd(rotx, _mm256_mul_pd(dvz, dwy));
                roty = _mm256_sub_pd(roty, _mm256_mul_pd(dvx, dwz));
                rotz = _mm256_sub_pd(rotz, _mm256_mul_pd(dvy, dwx));
            }
            /* hdr_density.hpp:351-362 with the refined reciprocals */
            double g[4], dv[4], rx[4], ry[4], rz[4];
            _mm256_storeu_pd(g, grdh);
            _mm256_storeu_pd(dv, divv);
            _mm256_storeu_pd(rx, rotx);
            _mm256_storeu_pd(ry, roty);
            _mm256_storeu_pd(rz, rotz);
            for (int l = 0; l < ni; l++) {
                const int i = i0 + l;
                const double dens = rh[l];
                const double deni = orc_rcp_4th(dens);
                const double omgi = orc_rcp_4th(1. + hsv[l] * deni * g[l] * orc_rcp_4th(3.));
                const double rot2 = rx[l] * rx[l] + ry[l] * ry[l] + rz[l] * rz[l];
                double rotv = rot2 * ((rot2 != 0.) ? orc_rsqrt_4th(rot2) : 0.);
                rotv *= deni * omgi;
                out.divv[i] = dv[l] * (deni * omgi);
                out.rotv[i] = rotv;
                out.grdh[i] = omgi;
            }
        }
    }
}

/* orc_hydro's pair sum (hdr_hydro.hpp:320-421) as the USE_INTRINSICS build
 * runs it (hdr_hydro.hpp:128-311): four i-particles per vector, 1/sqrt(x)
 * and 1/x by rsqrt_4th and rcp_4th; every other operation as orc_hydro's,
 * in its order. */
void orc_hydro_simd(int n_walk, const int32_t *i_off, const int32_t *j_off, const int32_t *j_idx,
                    orc_hydro_src pi, orc_hydro_src pj, orc_hydro_out out, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int w = 0; w < n_walk; w++) {
        const int jb = j_off[w], je = j_off[w + 1];
        for (int i0 = i_off[w]; i0 < i_off[w + 1]; i0 += 4) {
            const int ni = (i_off[w + 1] - i0 < 4) ? i_off[w + 1] - i0 : 4;
            double b[17][4];
            for (int l = 0; l < 4; l++) {
                const int i = i0 + (l < ni ? l : ni - 1);
                const double v[17] = {(double)pi.id[i], pi.x[i], pi.y[i], pi.z[i], pi.vx[i], pi.vy[i],
                                      pi.vz[i], pi.hinv[i], pi.pres[i], pi.presc[i], pi.bswt[i],
                                      pi.dens[i], pi.vsnd[i], pi.alph[i], pi.alphu[i], pi.thrm[i],
                                      pi.eta[i]};
                for (int k = 0; k < 17; k++) b[k][l] = v[k];
            }
            const __m256d id_i = _mm256_loadu_pd(b[0]);
            const __m256d px = _mm256_loadu_pd(b[1]), py = _mm256_loadu_pd(b[2]), pz = _mm256_loadu_pd(b[3]);
            const __m256d vx = _mm256_loadu_pd(b[4]), vy = _mm256_loadu_pd(b[5]), vz = _mm256_loadu_pd(b[6]);
            const __m256d hi_i = _mm256_loadu_pd(b[7]), pp_i = _mm256_loadu_pd(b[8]);
            const __m256d pc_i = _mm256_loadu_pd(b[9]), bs_i = _mm256_loadu_pd(b[10]);
            const __m256d dn_i = _mm256_loadu_pd(b[11]), cs_i = _mm256_loadu_pd(b[12]);
            const __m256d al_i = _mm256_loadu_pd(b[13]), au_i = _mm256_loadu_pd(b[14]);
            const __m256d eg_i = _mm256_loadu_pd(b[15]), et_i = _mm256_loadu_pd(b[16]);
            const __m256d hi4_i = _mm256_mul_pd(hi_i, _mm256_mul_pd(_mm256_mul_pd(hi_i, hi_i), hi_i));
            const __m256d zero = _mm256_setzero_pd();
            const __m256d sign = _mm256_set1_pd(-0.);
            __m256d achx = zero, achy = zero, achz = zero, udot = zero, vsmx = zero, difu = zero;
            __m256d g1x = zero, g1y = zero, g1z = zero;
            for (int k = jb; k < je; k++) {
                const int j = JROW(k);
                const __m256d dpx = _mm256_sub_pd(px, _mm256_set1_pd(pj.x[j]));
                const __m256d dpy = _mm256_sub_pd(py, _mm256_set1_pd(pj.y[j]));
                const __m256d dpz = _mm256_sub_pd(pz, _mm256_set1_pd(pj.z[j]));
                const __m256d dvx = _mm256_sub_pd(vx, _mm256_set1_pd(pj.vx[j]));
                const __m256d dvy = _mm256_sub_pd(vy, _mm256_set1_pd(pj.vy[j]));
                const __m256d dvz = _mm256_sub_pd(vz, _mm256_set1_pd(pj.vz[j]));
                const __m256d r2 = _mm256_add_pd(_mm256_add_pd(_mm256_mul_pd(dpx, dpx), _mm256_mul_pd(dpy, dpy)),
                                                 _mm256_mul_pd(dpz, dpz));
                const __m256d self = _mm256_cmp_pd(id_i, _mm256_set1_pd((double)pj.id[j]), _CMP_EQ_OQ);
                const __m256d ri = _mm256_andnot_pd(self, rsqrt4(r2));
                const __m256d r1 = _mm256_mul_pd(r2, ri);
                const double hj = pj.hinv[j];
                const __m256d q_i = _mm256_mul_pd(r1, hi_i);
                const __m256d q_j = _mm256_mul_pd(r1, _mm256_set1_pd(hj));
                const __m256d dw_i = _mm256_mul_pd(hi4_i, kernel1(q_i));
                const __m256d dw_j = _mm256_mul_pd(_mm256_set1_pd(hj * (hj * hj * hj)), kernel1(q_j));
                const __m256d ka = _mm256_mul_pd(_mm256_add_pd(dw_i, dw_j), _mm256_set1_pd(pj.m[j]));

                const __m256d rv = _mm256_add_pd(_mm256_add_pd(_mm256_mul_pd(dpx, dvx), _mm256_mul_pd(dpy, dvy)),
                                                 _mm256_mul_pd(dpz, dvz));
                const __m256d w_ij = _mm256_mul_pd(rv, ri);
                const __m256d w0 = _mm256_and_pd(w_ij, _mm256_cmp_pd(w_ij, zero, _CMP_LT_OQ));
                const __m256d vs = _mm256_sub_pd(_mm256_add_pd(cs_i, _mm256_set1_pd(pj.vsnd[j])),
                                                 _mm256_mul_pd(_mm256_set1_pd(3.), w0));
                const __m256d rhi = rcp4(_mm256_add_pd(dn_i, _mm256_set1_pd(pj.dens[j])));
                const __m256d av0 = _mm256_mul_pd(
                    _mm256_mul_pd(_mm256_mul_pd(_mm256_add_pd(bs_i, _mm256_set1_pd(pj.bswt[j])),
                                                _mm256_add_pd(al_i, _mm256_set1_pd(pj.alph[j]))),
                                  vs),
                    w0);
                vsmx = _mm256_blendv_pd(vs, vsmx, _mm256_cmp_pd(vsmx, vs, _CMP_GT_OQ));

                const __m256d ta = _mm256_mul_pd(
                    _mm256_mul_pd(_mm256_sub_pd(_mm256_add_pd(pc_i, _mm256_set1_pd(pj.presc[j])),
                                                _mm256_mul_pd(_mm256_mul_pd(_mm256_set1_pd(0.5), av0), rhi)),
                                  ka),
                    ri);
                achx = _mm256_sub_pd(achx, _mm256_mul_pd(ta, dpx));
                achy = _mm256_sub_pd(achy, _mm256_mul_pd(ta, dpy));
                achz = _mm256_sub_pd(achz, _mm256_mul_pd(ta, dpz));

                const __m256d dp = _mm256_sub_pd(pp_i, _mm256_set1_pd(pj.pres[j]));
                const __m256d vsu2 = _mm256_mul_pd(_mm256_mul_pd(_mm256_andnot_pd(sign, dp), rhi),
                                                   _mm256_set1_pd(2.));
                const __m256d vsui = _mm256_andnot_pd(_mm256_cmp_pd(vsu2, zero, _CMP_EQ_OQ), rsqrt4(vsu2));
                const __m256d vsu = _mm256_mul_pd(vsu2, vsui);
                const __m256d du = _mm256_sub_pd(eg_i, _mm256_set1_pd(pj.thrm[j]));
                const __m256d au = _mm256_add_pd(au_i, _mm256_set1_pd(pj.alphu[j]));
                const __m256d inner = _mm256_sub_pd(
                    _mm256_mul_pd(_mm256_mul_pd(_mm256_set1_pd(0.25), av0), w0),
                    _mm256_mul_pd(_mm256_mul_pd(au, vsu), du));
                udot = _mm256_add_pd(udot, _mm256_mul_pd(ka, _mm256_sub_pd(_mm256_mul_pd(pc_i, w_ij),
                                                                           _mm256_mul_pd(rhi, inner))));
                difu = _mm256_add_pd(difu, _mm256_mul_pd(_mm256_mul_pd(_mm256_mul_pd(rhi, du), ka), ri));

                const __m256d dg = _mm256_mul_pd(
                    _mm256_mul_pd(_mm256_set1_pd(pj.m[j]), ri),
                    _mm256_add_pd(_mm256_mul_pd(et_i, dw_i), _mm256_mul_pd(_mm256_set1_pd(pj.eta[j]), dw_j)));
                g1x = _mm256_add_pd(g1x, _mm256_mul_pd(dg, dpx));
                g1y = _mm256_add_pd(g1y, _mm256_mul_pd(dg, dpy));
                g1z = _mm256_add_pd(g1z, _mm256_mul_pd(dg, dpz));
            }
            double o[9][4];
            _mm256_storeu_pd(o[0], achx);
            _mm256_storeu_pd(o[1], achy);
            _mm256_storeu_pd(o[2], achz);
            _mm256_storeu_pd(o[3], udot);
            _mm256_storeu_pd(o[4], vsmx);
            _mm256_storeu_pd(o[5], difu);
            _mm256_storeu_pd(o[6], g1x);
            _mm256_storeu_pd(o[7], g1y);
            _mm256_storeu_pd(o[8], g1z);
            for (int l = 0; l < ni; l++) {
                const int i = i0 + l;
                out.ax[i] = o[0][l] * 0.5;
                out.ay[i] = o[1][l] * 0.5;
                out.az[i] = o[2][l] * 0.5;
                out.udot[i] = o[3][l] * 0.5;
                out.vsmx[i] = o[4][l];
                out.diffu[i] = o[5][l] * 2.;
                out.gx[i] = o[6][l] * 0.5;
                out.gy[i] = o[7][l] * 0.5;
                out.gz[i] = o[8][l] * 0.5;
            }
        }
    }
}

/* One list of orc_gravity's pair sum for four i-particles (hdr_gravity.hpp:
 * 194-291, the USE_INTRINSICS build: 1/sqrt by rsqrt_4th), accumulated. */
static void grav_list4(const int *ii, int ni, orc_grav_src pi, const int32_t *idx, int kb, int ke,
                       orc_grav_src pj, orc_grav_out out)
{
    double b[4][4];
    for (int l = 0; l < 4; l++) {
        const int i = ii[l < ni ? l : ni - 1];
        b[0][l] = pi.x[i];
        b[1][l] = pi.y[i];
        b[2][l] = pi.z[i];
        b[3][l] = pi.eps2[i];
    }
    const __m256d px = _mm256_loadu_pd(b[0]), py = _mm256_loadu_pd(b[1]), pz = _mm256_loadu_pd(b[2]);
    const __m256d e2_i = _mm256_loadu_pd(b[3]);
    __m256d ax = _mm256_setzero_pd(), ay = ax, az = ax, pt = ax, et = ax;
    for (int k = kb; k < ke; k++) {
        const int j = idx ? idx[k] : k;
        const __m256d dpx = _mm256_sub_pd(px, _mm256_set1_pd(pj.x[j]));
        const __m256d dpy = _mm256_sub_pd(py, _mm256_set1_pd(pj.y[j]));
        const __m256d dpz = _mm256_sub_pd(pz, _mm256_set1_pd(pj.z[j]));
        const __m256d r2 = _mm256_add_pd(_mm256_add_pd(_mm256_mul_pd(dpx, dpx), _mm256_mul_pd(dpy, dpy)),
                                         _mm256_mul_pd(dpz, dpz));
        const __m256d rei_i = rsqrt4(_mm256_add_pd(r2, e2_i));
        const __m256d rei3_i = _mm256_mul_pd(_mm256_mul_pd(rei_i, rei_i), rei_i);
        const __m256d rei_j = rsqrt4(_mm256_add_pd(r2, _mm256_set1_pd(pj.eps2[j])));
        const __m256d rei3_j = _mm256_mul_pd(_mm256_mul_pd(rei_j, rei_j), rei_j);
        const __m256d m_j = _mm256_set1_pd(pj.m[j]);
        const __m256d dg2 = _mm256_mul_pd(m_j, _mm256_add_pd(rei3_i, rei3_j));
        pt = _mm256_sub_pd(pt, _mm256_mul_pd(m_j, _mm256_add_pd(rei_i, rei_j)));
        ax = _mm256_sub_pd(ax, _mm256_mul_pd(dpx, dg2));
        ay = _mm256_sub_pd(ay, _mm256_mul_pd(dpy, dg2));
        az = _mm256_sub_pd(az, _mm256_mul_pd(dpz, dg2));
        et = _mm256_add_pd(et, _mm256_mul_pd(m_j, rei3_i));
    }
    double o[5][4];
    _mm256_storeu_pd(o[0], ax);
    _mm256_storeu_pd(o[1], ay);
    _mm256_storeu_pd(o[2], az);
    _mm256_storeu_pd(o[3], pt);
    _mm256_storeu_pd(o[4], et);
    for (int l = 0; l < ni; l++) {
        const int i = ii[l];
        out.ax[i] += o[0][l] * 0.5;
        out.ay[i] += o[1][l] * 0.5;
        out.az[i] += o[2][l] * 0.5;
        out.pot[i] += o[3][l] * 0.5;
        out.eta[i] += o[4][l];
    }
}

/* orc_gravity with four i-particles per vector (the EPJ list, then the SPJ
 * list, as orc_gravity). */
void orc_gravity_simd(int n_walk, const int32_t *i_off,
                      const int32_t *j_off, const int32_t *j_idx, orc_grav_src pj,
                      const int32_t *s_off, const int32_t *s_idx, orc_grav_src ps,
                      orc_grav_src pi, orc_grav_out out, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int w = 0; w < n_walk; w++) {
        for (int i0 = i_off[w]; i0 < i_off[w + 1]; i0 += 4) {
            const int ni = (i_off[w + 1] - i0 < 4) ? i_off[w + 1] - i0 : 4;
            int ii[4];
            for (int l = 0; l < 4; l++) ii[l] = i0 + l;
            for (int l = 0; l < ni; l++) {
                const int i = ii[l];
                out.ax[i] = 0.; out.ay[i] = 0.; out.az[i] = 0.; out.pot[i] = 0.; out.eta[i] = 0.;
            }
            grav_list4(ii, ni, pi, j_idx, j_off[w], j_off[w + 1], pj, out);
            if (s_off) grav_list4(ii, ni, pi, s_idx, s_off[w], s_off[w + 1], ps, out);
        }
    }
}
