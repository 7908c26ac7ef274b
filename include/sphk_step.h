/*
 * sphk_step.h -- per-particle step kernels around the interaction functors
 * (C-ABI of libsphk.so; device pointers, asynchronous on `stream`, a
 * hipStream_t as void*, NULL being the HIP null stream as in sphk.h).
 *
 * These are the reference's O(N) host loops between and after the three
 * force passes of calcSPHKernel (sph_yamauchi/run/main.cpp:293-358) and its
 * integrator, moved to the device so that a whole step stays in HBM:
 *   referEquationOfState        hdr_hgas.hpp:318-344 (ideal gas P = rho u / n)
 *   calcBalsaraSwitch           hdr_sph.hpp:79-89
 *   GravityEPI/EPJ eps2         hdr_gravity.hpp:134-137,162-166
 *   SPH::copyFromForce(Gravity) hdr_gravity.hpp:107-113
 *   HydroEPI::copyFromFP        hdr_hydro.hpp:45-62
 *   SPH::copyFromForce(Hydro)   hdr_hydro.hpp:19-26
 *   sumAcceleration             hdr_sph.hpp:95-97
 *   calcAlphaDot                hdr_hgas.hpp:750-762
 *   calcTimestep (per particle) hdr_hgas.hpp:477-488
 *   predict / correct           hdr_hgas.hpp:646-665
 *   calcEnergy / calcMomentum   hdr_hgas.hpp:630-644, main.cpp:197-239
 * The ideal-gas minimum energy umin is 0 (hdr_hgas.hpp:342).
 */
#ifndef SPHK_STEP_H
#define SPHK_STEP_H

#include "sphk.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Particle state, structure of arrays (FP64, device pointers).  Field
 * meanings follow class SPH (hdr_sph.hpp:12-50); ax..az is acc, vx2.. the
 * half-step velocity, g1* accg1 (hydro grad-h gravity), g2* accg2 (tree
 * gravity). */
typedef struct {
    double *m, *x, *y, *z, *vx, *vy, *vz;
    double *ax, *ay, *az;
    double *vx2, *vy2, *vz2;
    double *uene, *uene2, *udot;
    double *alph, *alph2, *adot;
    double *alphu, *alphu2, *adotu;
    double *diffu, *ksr, *dens, *pres, *vsnd, *divv, *rotv, *bswt, *grdh, *vsmx;
    double *pot, *eta;
    double *achx, *achy, *achz;
    double *g1x, *g1y, *g1z;
    double *g2x, *g2y, *g2z;
} sphk_state;

/* Physical parameters of the step (RP:: globals and EOS constants). */
typedef struct {
    double ninv;        /* 1/n of P = ninv rho u (n = 1.5 <=> gamma = 5/3) */
    double ksrh;        /* kappa */
    double G;           /* gravitational constant */
    double alpha_max, alpha_min, alphu_max, alphu_min;
    double eps_u;       /* RP::EpsilonOfInternalEnergy */
    double c_dt;        /* RP::CoefficientOfTimestep */
    double dt_max;      /* RP::MaximumTimestep */
    double eps_const;   /* > 0: the gravitational softening is this constant
                           length (not the reference's eps = H/kappa,
                           hdr_gravity.hpp:136), so the potential does not
                           depend on H and the gravity write-back sets the
                           grad-h softening term eta to 0; <= 0: eps = H/kappa */
} sphk_phys;

/* EOS and Balsara switch.  The density functor's outputs are expected in
 * the state already (SPH::copyFromForce, hdr_density.hpp:24-31): the
 * device-resident step points sphk_dens_out at the state arrays. */
int sphk_eos_balsara_dev(sphk_ctx *ctx, int n, sphk_state s, sphk_phys p, void *stream);

/* One round of the smoothing-length retry loop (calcDensityKernel,
 * hdr_density.hpp:382-414), after a density pass into `round` with the
 * gather lists of radius rs: each particle still retrying (rs != 0) whose
 * round converged (itr == 0) takes the round's dens, divv, rotv, grdh, ksr
 * and np (copyFromForce, hdr_density.hpp:24-31) and gets rs = 0 (its rs is
 * kept in rs_conv unless rs_conv is NULL); one that did not retries with
 * rs * expand (jump != 0: at least expand * ksr of the round).  *retry
 * (device int32) = 1 if any particle retries, else 0.  Asynchronous. */
int sphk_density_commit_dev(sphk_ctx *ctx, int n, sphk_state s, sphk_dens_out round, int64_t *np, double *rs,
                            double *rs_conv, double expand, int jump, int32_t *retry, void *stream);

/* eps2 = (ksr / kappa)^2 when eps_const <= 0, else eps_const^2. */
int sphk_softening_dev(sphk_ctx *ctx, int n, const double *ksr, double ksrh, double eps_const,
                       double *eps2, void *stream);

/* accg2, pot and eta from the gravity functor's outputs (eta = 0 when
 * p.eps_const > 0; the functor's eta output may then be NULL). */
int sphk_gravity_writeback_dev(sphk_ctx *ctx, int n, sphk_state s, sphk_phys p,
                               sphk_grav_out g, void *stream);

/* hinv, presc, 0.5*bswt, thrm for the hydro EPI/EPJ. */
int sphk_hydro_prep_dev(sphk_ctx *ctx, int n, sphk_state s, double *hinv, double *presc,
                        double *bswt_half, double *thrm, void *stream);

/* Hydro write-back, acc = acch + accg1 + accg2, alpha-dot. */
int sphk_hydro_writeback_dev(sphk_ctx *ctx, int n, sphk_state s, sphk_phys p,
                             sphk_hydro_out h, void *stream);

int sphk_predict_dev(sphk_ctx *ctx, int n, sphk_state s, double dt, void *stream);
int sphk_correct_dev(sphk_ctx *ctx, int n, sphk_state s, double dt, void *stream);

/* FP64 reductions with a fixed (deterministic) order, written to host
 * memory out[0..5] after a stream synchronisation:
 *   out[0] total energy  sum m (v^2/2 + u + pot/2)
 *   out[1..3] momentum   sum m v
 *   out[4] min over particles of the per-particle timestep
 *   out[5] total mass */
int sphk_diagnostics(sphk_ctx *ctx, int n, sphk_state s, sphk_phys p, double *out, void *stream);

/* Allocation-free variant of sphk_diagnostics.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same kernels, grid and reduction order as sphk_diagnostics (bitwise the same
 * six values), but the results go to DEVICE memory out_dev[0..5],
 * asynchronously on `stream` with no synchronisation, and the 1024 per-block
 * partials live in a file-scope device array of 6 x 1024 doubles (48 KiB) that
 * each GPU holds until the process ends.  No hipMalloc, hipFree or allocating
 * library call is made.  Any n >= 0 is accepted (the kernels stride over the
 * particles with the fixed grid, so the scratch does not grow with n).  Calls
 * that overlap on one GPU share that array and race: serialise them. */
int sphk_diagnostics_static(sphk_ctx *ctx, int n, sphk_state s, sphk_phys p, double *out_dev,
                            void *stream);

/* Unbound mass of the current state (analysis/unbound_mass/unbound_mass.py:10-28,
 * analysis/common.py:71-90): v_cg = sum rho v / sum rho over the selected
 * particles, e_i = |v_i - v_cg|^2 / 2 + pot_i + u_i; hard = sum of m_i with
 * e_i > 0, smoothed = sum m_i / (1 + exp(-k e_i)) (a differentiable
 * surrogate; k in units of 1 / specific energy).  istar (device, may be NULL)
 * with star >= 0 selects one star.  FP64 fixed-order reductions; out (host)
 * = {v_cg x, y, z, hard, smoothed, selected mass} after a stream sync. */
int sphk_unbound_mass_dev(sphk_ctx *ctx, int n, sphk_state s, const int32_t *istar, int star, double k,
                          double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
