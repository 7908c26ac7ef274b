/*
 * sphk_adjoint.h -- vector-Jacobian product of the gravity functor (C-ABI).
 *
 * Forward (sphk_gravity_dev with uniform softening, hdr_gravity.hpp:295-338
 * with every eps2 equal): for particle i, over its EPJ list (particles j)
 * and SPJ list (node monopoles: mass M, centre of mass X),
 *     a_i   = - sum m_j d (r^2 + e2)^(-3/2),   d = x_i - x_j,
 *     pot_i = - sum m_j (r^2 + e2)^(-1/2)      (self term included),
 *     eta_i =   sum m_j (r^2 + e2)^(-3/2)      (the eta output, before the
 *                                              write-back's H/Omega factor).
 * Given cotangents abar (3 x N), pbar (N) and ebar (N), sphk_gravity_vjp_dev
 * adds
 *     d/dx_i:  sum_j m_j ( -T_ij abar_i + pbar_i w3 d - 3 ebar_i w5 d )
 *     d/dx_j:        m_j (  T_ij abar_i - pbar_i w3 d + 3 ebar_i w5 d )
 * with T = w3 I - 3 w5 d d^T, w = (r^2 + e2)^(-1/2), into gx/gy/gz for the
 * particles of EPJ entries and into nx/ny/nz for the nodes of SPJ entries
 * (FP64 throughout; atomic accumulation, so the summation order is not
 * fixed).  sphk_dtree_spread_dev then carries the node cotangents to the
 * particles through X_n = sum_{k in n} m_k x_k / M_n (d X_n / d x_k =
 * m_k / M_n, every node containing k).  The lists and the tree are held
 * fixed (piecewise-constant in x), as in any tree code's derivative.
 * The masses are not differentiated.
 *
 * This operation has no counterpart in the reference; it serves callers that
 * differentiate collision outcomes with respect to initial conditions.
 */
#ifndef SPHK_ADJOINT_H
#define SPHK_ADJOINT_H

#include "sphk.h"
#include "sphk_dtree.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Accumulates (does not clear) gx, gy, gz (particles, N) and nx, ny, nz
 * (nodes; may be NULL when there is no SPJ list).  pj and pi give particle
 * m, x, y, z (eps2 unused: e2 is the uniform softening squared); ps the node
 * monopoles m, x, y, z.  ebar may be NULL (no eta cotangent).  Device
 * pointers; stream as in sphk.h. */
int sphk_gravity_vjp_dev(sphk_ctx *ctx, int n_walk, const int32_t *i_off, const int32_t *j_off,
                         const int32_t *j_idx, sphk_grav_src pj, const int32_t *s_off,
                         const int32_t *s_idx, sphk_grav_src ps, sphk_grav_src pi, double e2,
                         const double *abx, const double *aby, const double *abz, const double *pbar,
                         const double *ebar, double *gx, double *gy, double *gz, double *nx, double *ny,
                         double *nz, void *stream);

/* Cotangents of the density functor's outputs, and of its differentiable
 * inputs (positions, velocities, the starting smoothing length ksr). */
typedef struct {
    const double *dens, *ksr, *divv, *rotv, *grdh;
} sphk_dens_cot;
typedef struct {
    double *x, *y, *z, *vx, *vy, *vz, *ksr;
} sphk_dens_grad;

/* Vector-Jacobian product of the FP64 density functor (sphk_density_dev,
 * hdr_density.hpp:254-363: three smoothing-length passes from ksr, then the
 * grad-h, divergence and curl pass) over gather lists holding every j within
 * the largest H of the passes (as the functor's retry loop leaves them), with
 * the context's eta, kappa and H_max (a capped H has zero derivative): given
 * cotangents of dens, ksr (H_3), divv, rotv and grdh, adds to grad the
 * cotangents of every particle's x, y, z, vx, vy, vz and of the i-particles'
 * starting ksr.  np and itr are not differentiated.  One set of particle
 * arrays serves as both sides.  j-side terms use FP64 atomics (the summation
 * order is not fixed).  No counterpart in the reference; for callers that
 * differentiate collision outcomes. */
int sphk_density_vjp_dev(sphk_ctx *ctx, int n_walk, const int32_t *i_off, const int32_t *j_off,
                         const int32_t *j_idx, sphk_dens_src src, sphk_dens_cot cot, sphk_dens_grad grad,
                         void *stream);

/* Cotangents of the hydro functor's differentiable inputs (sphk_hydro_src
 * fields other than id and m), accumulated by sphk_hydro_vjp_dev. */
typedef struct {
    double *x, *y, *z, *vx, *vy, *vz;
    double *hinv, *pres, *presc, *bswt, *dens, *vsnd, *alph, *alphu, *thrm, *eta;
} sphk_hydro_grad;

/* Vector-Jacobian product of the FP64 hydro functor (sphk_hydro_dev,
 * hdr_hydro.hpp:320-421) over symmetric lists (each walk's list holds every
 * j with r_ij < max(H_i, H_j) for some i of the walk, as the symmetry list
 * builders make them): given cotangents of the outputs ax, ay, az, udot,
 * diffu, gx, gy, gz (vsmx is not differentiated; its pointer is ignored),
 * adds to grad the cotangents of every particle's inputs.  One set of
 * particle arrays serves as both the i- and the j-side.  Each particle's
 * cotangent is summed by its own lanes in a fixed order: no atomics,
 * bitwise reproducible.  This operation has no counterpart in the
 * reference; it serves callers that differentiate collision outcomes. */
int sphk_hydro_vjp_dev(sphk_ctx *ctx, int n_walk, const int32_t *i_off, const int32_t *j_off,
                       const int32_t *j_idx, sphk_hydro_src src, sphk_hydro_out cot,
                       sphk_hydro_grad grad, void *stream);

/* gx[k] += m_k * sum over the nodes n containing k of nx[n] / M_n (same for
 * y, z); m = the particle masses the tree's monopoles were built from. */
int sphk_dtree_spread_dev(const sphk_dtree *t, const double *m, const double *nx, const double *ny,
                          const double *nz, double *gx, double *gy, double *gz);

#ifdef __cplusplus
}
#endif
#endif
