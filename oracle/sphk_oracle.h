/*
 * sphk_oracle.h -- CPU oracle for the SPH interaction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a clean-room, scalar, strict-IEEE FP64
 * restatement of the reference's CPU interaction functors (the scalar
 * branches, i.e. the build without USE_INTRINSICS) and of the per-particle
 * glue between them.  It exists to check the HIP product path; nothing in the
 * product links it.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.
 *
 * Build: oracle/Makefile (gcc -O2 -fno-fast-math -ffp-contract=off -fopenmp).
 *
 * Data model (shared with the HIP C-ABI, include/sphk.h):
 *   A "walk" is one i-group with one shared j-list, as handed to the
 *   reference's multiwalk Dispatch functions (sph_yamauchi/run/use_gpu.hpp:28-64).
 *   i-particles of walk w are rows i_off[w] .. i_off[w+1]-1 of the i-arrays.
 *   j-entries of walk w are rows j_idx[j_off[w] .. j_off[w+1]-1] of the
 *   j-arrays (j_idx == NULL means the identity: rows j_off[w] .. j_off[w+1]-1).
 *   All arrays are structure-of-arrays.
 *
 * Pinning status of each function is recorded in its comment; the tests that
 * pin it live in tests/test_oracle_*.py.
 */
#ifndef SPHK_ORACLE_H
#define SPHK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Wendland C2 (3-D) constants, sph_yamauchi/run/hdr_kernel.hpp:218-227. */
#define ORC_ETA_DEFAULT 1.6
#define ORC_KSRH        1.936492
#define ORC_CEFF0       3.342253804929802286
#define ORC_CEFF1       13.36901521971920914

typedef struct {
    const int32_t *id;
    const double *m, *x, *y, *z, *vx, *vy, *vz;
    const double *ksr; /* i-side only: current kernel support radius H */
    const double *rs;  /* i-side only: search radius (itr threshold)   */
} orc_dens_src;

typedef struct {
    double *dens, *divv, *rotv, *grdh, *ksr;
    int64_t *np;
    uint8_t *itr;
} orc_dens_out;

typedef struct {
    const int32_t *id;
    const double *m, *x, *y, *z, *vx, *vy, *vz;
    const double *hinv, *pres, *presc, *bswt, *dens, *vsnd, *alph, *alphu, *thrm, *eta;
} orc_hydro_src;

typedef struct {
    double *ax, *ay, *az;    /* acch */
    double *gx, *gy, *gz;    /* accg (grad-h gravity correction) */
    double *udot, *vsmx, *diffu;
} orc_hydro_out;

typedef struct {
    const double *m, *x, *y, *z, *eps2;
} orc_grav_src;

typedef struct {
    double *ax, *ay, *az, *pot, *eta;
} orc_grav_out;

/* Kernel and its derivative, hdr_kernel.hpp:246-257. */
double orc_kernel0th(double q);
double orc_kernel1st(double q);

/* Scalar calcDensity, hdr_density.hpp:246-365.  eta/ksrh/hmax are the
 * compile-time SK::eta, SK::ksrh and RP::KernelSupportRadiusMaximum. */
void orc_density(int n_walk, const int32_t *i_off, const int32_t *j_off, const int32_t *j_idx,
                 orc_dens_src pi, orc_dens_src pj, double eta, double ksrh, double hmax,
                 orc_dens_out out, int nthreads);
/* The AVX2 branch's arithmetic (hdr_density.hpp:82-242): orc_density with
 * the approximate reciprocals below in place of 1/x and 1/sqrt(x).  Its
 * seeds come from the host CPU's estimate instructions, so its last bits are
 * the host's; it quantifies the shipped CPU path's difference from the
 * scalar functor.  The tail read past nip of that branch is not reproduced. */
void orc_density_avx2(int n_walk, const int32_t *i_off, const int32_t *j_off, const int32_t *j_idx,
                      orc_dens_src pi, orc_dens_src pj, double eta, double ksrh, double hmax,
                      orc_dens_out out, int nthreads);
/* rcp_4th / rsqrt_4th of vector_x86.hpp:501-534, one lane. */
double orc_rcp_4th(double x);
double orc_rsqrt_4th(double x);
/* orc_density_avx2 computed four i-particles per AVX2 vector, as the
 * reference's USE_INTRINSICS calcDensity runs (hdr_density.hpp:82-242):
 * bitwise the same results (sphk_oracle_simd.c).  For timing the CPU
 * baseline on the vector path.  Requires AVX2 (orc_density_simd_supported). */
void orc_density_simd(int n_walk, const int32_t *i_off, const int32_t *j_off, const int32_t *j_idx,
                      orc_dens_src pi, orc_dens_src pj, double eta, double ksrh, double hmax,
                      orc_dens_out out, int nthreads);
int orc_density_simd_supported(void);
/* orc_hydro and orc_gravity four i-particles per AVX2 vector with the
 * USE_INTRINSICS builds' refined reciprocals (hdr_hydro.hpp:128-311,
 * hdr_gravity.hpp:194-291): the CPU baseline's vector path. */
void orc_hydro_simd(int n_walk, const int32_t *i_off, const int32_t *j_off, const int32_t *j_idx,
                    orc_hydro_src pi, orc_hydro_src pj, orc_hydro_out out, int nthreads);
void orc_gravity_simd(int n_walk, const int32_t *i_off,
                      const int32_t *j_off, const int32_t *j_idx, orc_grav_src pj,
                      const int32_t *s_off, const int32_t *s_idx, orc_grav_src ps,
                      orc_grav_src pi, orc_grav_out out, int nthreads);

/* Scalar calcHydro, hdr_hydro.hpp:312-424.  Outputs are assigned. */
void orc_hydro(int n_walk, const int32_t *i_off, const int32_t *j_off, const int32_t *j_idx,
               orc_hydro_src pi, orc_hydro_src pj, orc_hydro_out out, int nthreads);

/* Scalar calcGravity<TJ>, hdr_gravity.hpp:286-342.  Accumulates into out over
 * the EPJ list then (when nsp_off != NULL) the SPJ list of each walk.
 * pi uses only x,y,z,eps2. */
void orc_gravity(int n_walk, const int32_t *i_off,
                 const int32_t *j_off, const int32_t *j_idx, orc_grav_src pj,
                 const int32_t *s_off, const int32_t *s_idx, orc_grav_src ps,
                 orc_grav_src pi, orc_grav_out out, int nthreads);

/* ---- per-particle glue (hdr_hgas.hpp, hdr_sph.hpp, main.cpp) ---------- */

typedef struct {
    double *m, *x, *y, *z, *vx, *vy, *vz;
    double *ax, *ay, *az;           /* acc = acch + accg1 + accg2 */
    double *vx2, *vy2, *vz2;        /* half-step velocity */
    double *uene, *uene2, *udot;
    double *alph, *alph2, *adot;
    double *alphu, *alphu2, *adotu;
    double *diffu, *ksr, *dens, *pres, *vsnd, *divv, *rotv, *bswt, *grdh, *vsmx;
    double *pot, *eta;
    double *achx, *achy, *achz;
    double *g1x, *g1y, *g1z;        /* accg1 */
    double *g2x, *g2y, *g2z;        /* accg2 */
} orc_state;

typedef struct {
    double ninv;        /* 1/n, ideal gas P = ninv * rho * u   (hdr_hgas.hpp:339-341) */
    double ksrh;        /* kappa */
    double G;           /* gravitational constant in code units */
    double alpha_max, alpha_min, alphu_max, alphu_min;
    double eps_u;       /* RP::EpsilonOfInternalEnergy */
    double c_dt;        /* RP::CoefficientOfTimestep */
    double dt_max;      /* RP::MaximumTimestep */
    double eps_const;   /* > 0: constant gravitational softening length (this
                           build's mode; the reference only has eps = H/kappa,
                           hdr_gravity.hpp:136).  The potential then does not
                           depend on H, so the grad-h softening term eta is 0. */
} orc_params;

/* EOS + Balsara: referEquationOfState (hdr_hgas.hpp:318-344) then
 * calcBalsaraSwitch (hdr_sph.hpp:79-89). */
void orc_eos_balsara(int n, orc_state s, orc_params p);
/* HydroEPI/EPJ::copyFromFP derived fields (hdr_hydro.hpp:45-62). */
void orc_hydro_prep(int n, const double *ksr, const double *pres, const double *grdh,
                    const double *dens, const double *bswt, const double *uene,
                    double *hinv, double *presc, double *bswt_half, double *thrm);
/* Gravity write-back SPH::copyFromForce(Gravity) (hdr_gravity.hpp:107-113);
 * eta = 0 when p.eps_const > 0. */
void orc_gravity_writeback(int n, orc_state s, orc_params p,
                           const double *gax, const double *gay, const double *gaz,
                           const double *gpot, const double *geta);
/* Hydro write-back (hdr_hydro.hpp:19-26), sumAcceleration (hdr_sph.hpp:95-97),
 * calcAlphaDot (hdr_hgas.hpp:750-762). */
void orc_hydro_writeback_tail(int n, orc_state s, orc_params p,
                              const double *hax, const double *hay, const double *haz,
                              const double *hgx, const double *hgy, const double *hgz,
                              const double *udot, const double *vsmx, const double *diffu);
/* Per-particle dt (hdr_hgas.hpp:477-488) reduced with min, then the global
 * power-of-two adjustment of main.cpp:241-273.  Returns the new dt, or a
 * negative value when dt would fall to dt_min (the reference aborts). */
double orc_timestep(int n, orc_state s, orc_params p, double dt_prev, double time, double dt_min);
/* predict / correct, hdr_hgas.hpp:646-665 (dnuc = gcor3 = 0). */
void orc_predict(int n, orc_state s, double dt);
void orc_correct(int n, orc_state s, double dt);
/* Total energy (hdr_hgas.hpp:630-634 with gcor3 = 0) and momentum
 * (hdr_hgas.hpp:641-643). */
double orc_energy(int n, orc_state s);
void orc_momentum(int n, orc_state s, double out[3]);
/* calcKernelSupportRadiusMaximum for FlagBinary==0 (hdr_util.hpp:489-505):
 * density-weighted RMS radius about the centre of mass of star `istar_sel`
 * (hdr_util.hpp:7-34). */
double orc_ksr_max(int n, const double *m, const double *x, const double *y, const double *z,
                   const double *dens, const int32_t *istar, int istar_sel);

/* Unbound mass of one snapshot, analysis/unbound_mass/unbound_mass.py:10-28:
 * v_cg = sum rho v / sum rho over the selected particles
 * (analysis/common.py:71-90), e_i = |v_i - v_cg|^2 / 2 + pot_i + u_i, and
 * the mass with e_i > 0.  Also the smoothed mass sum m sigma(k e_i),
 * sigma(x) = 1 / (1 + exp(-x)).  Selection: istar == NULL or star < 0 takes
 * every particle, else those with istar[i] == star.
 * out = {v_cg x, y, z, unbound (hard), unbound (smoothed), selected mass}. */
void orc_unbound_mass(int n, orc_state s, const int32_t *istar, int star, double k, double out[6]);

#ifdef __cplusplus
}
#endif
#endif
