/*
 * sphk.h -- C-ABI of the MI355X (gfx950) SPH interaction library, libsphk.so.
 *
 * This is the drop-in boundary under the FDPS multiwalk functor API of
 * yamauchi1132/Research_Codes.  The reference hands batches of "walks" (an
 * i-group plus the j-list FDPS built for it) to six C++ functions declared in
 * sph_yamauchi/run/use_gpu.hpp:28-64 and implemented in
 * sph_yamauchi/run/hdr_{density,hydro,gravity}GPU.cu.  include/sphk_fdps.hpp
 * keeps those six C++ signatures and the EPI/EPJ/SPJ/Force class layouts;
 * it is a thin adapter over the plain-C entry points below.
 *
 * Conventions
 *  - Plain C: pointers, sizes, int32 offsets.  No C++ or torch types.
 *  - Every entry point returns 0 on success or a negative SPHK_E* code;
 *    sphk_last_error() gives a message.  (The reference asserts or exits;
 *    see hdr_densityGPU.cu:503,521-522 and cuda_pointer.h:9.)
 *  - Walk layout.  Walk w owns i-rows i_off[w] .. i_off[w+1]-1 of the i-arrays
 *    and the j-entries j_idx[j_off[w] .. j_off[w+1]-1], which are rows of the
 *    j-arrays; j_idx == NULL means rows j_off[w] .. j_off[w+1]-1 (a packed
 *    list, as the reference's Dispatch builds, hdr_densityGPU.cu:513-566).
 *  - Arrays are structure-of-arrays, FP64, ids int32.
 *  - "_dev" entry points take DEVICE pointers (inputs resident in HBM) and
 *    launch asynchronously on `stream`, a hipStream_t passed as void*; NULL
 *    is the HIP null stream (the legacy default stream, which is where a
 *    caller's default stream -- torch's included -- runs).  Pass
 *    sphk_stream(ctx) to use the context's own (non-blocking) stream.  "_dispatch"/"_retrieve" take HOST
 *    pointers and mirror the reference's Dispatch/Retrieve split: dispatch
 *    copies in and launches, retrieve waits and copies out.
 *  - precision: SPHK_FP64 evaluates every pair term in FP64; SPHK_MIXED
 *    forms each pair displacement in FP64, casts it to FP32, evaluates the
 *    pair term in FP32 and accumulates per-particle totals in FP64.  The
 *    smoothing-length iteration (neighbour counts, H, itr, density) is
 *    always evaluated in FP64, so np/ksr/itr are identical in both modes.
 *    Any other value is rejected with SPHK_E_ARG.
 */
#ifndef SPHK_H
#define SPHK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SPHK_OK = 0,
    SPHK_E_ARG = -1,       /* invalid argument (null pointer, negative size) */
    SPHK_E_CAPACITY = -2,  /* batch exceeds the context's limits */
    SPHK_E_DEVICE = -3,    /* HIP runtime error */
    SPHK_E_STATE = -4,     /* retrieve without dispatch, tag out of range ... */
    SPHK_E_NONFINITE = -5  /* non-finite input detected by a validating call */
};

enum { SPHK_FP64 = 0, SPHK_MIXED = 1 };

/* Limits of one host dispatch, use_gpu.hpp:20-24 (N_WALK_LIMIT, NI_LIMIT, NJ_LIMIT). */
enum {
    SPHK_WALK_LIMIT = 1000,
    SPHK_NI_LIMIT = 1000 * 1000,
    SPHK_NJ_LIMIT = 1000 * 10000,
    SPHK_TAG_MAX = 4,
    SPHK_MAX_GROUP = 64  /* i-particles per device task; larger walks are split */
};

typedef struct sphk_ctx sphk_ctx;

/* Kernel-side constants.  Replaces the hard-coded SK:: constants and the
 * hidden global RP::KernelSupportRadiusMaximum that the reference's density
 * Dispatch reads (hdr_densityGPU.cu:492-494,576-585; hdr_kernel.hpp:218-227). */
typedef struct {
    double eta;   /* SK::eta, 1.6 in the reference */
    double ksrh;  /* SK::ksrh (kappa), 1.936492 */
    double hmax;  /* RP::KernelSupportRadiusMaximum */
    /* Nonzero: the caller guarantees that every eps2 the gravity functor sees
     * (EPI, EPJ and SPJ) is the same value, e.g. a constant gravitational
     * softening.  The symmetrised pair term then needs one rsqrt instead of
     * two; results are those of the general formula.  0 (default): general. */
    int uniform_softening;
    /* Density functor: capacity (records) of a walk's j-list kept in LDS for
     * its four sweeps, 192 or 256; longer lists stream through 64-record
     * tiles.  192 needs less LDS (more waves per CU), 256 keeps more walks
     * resident.  0 (default): 256. */
    int list_cap;
    /* Launch of the three functors.  0 (default): one workgroup per walk.
     * 1: persistent -- as many workgroups as the chip holds at once, each
     * taking walks from a device queue of its XCD until none is left (one
     * launch for a whole ensemble of runs, any imbalance between them).
     * Results are bitwise those of 0.  Each launch uses the next of 64 walk
     * queues of the context: at most 64 persistent launches of one context
     * may be in flight at once. */
    int persistent;
    /* Host dispatch (sphk_*_dispatch): 1 declares that the caller's source
     * arrays are pinned host memory (hipHostMalloc) left unchanged until the
     * matching sphk_*_retrieve returns; they are then copied to the device
     * directly instead of through the tag's pinned staging block (one host
     * copy of every field less).  0 (default): staged. */
    int pinned_src;
} sphk_params;

/* ---- inputs / outputs (all FP64 unless stated) ------------------------ */

typedef struct {  /* DensityEPI / DensityEPJ, hdr_density.hpp:33-78 */
    const int32_t *id;
    const double *m, *x, *y, *z, *vx, *vy, *vz;
    const double *ksr;  /* i-side only */
    const double *rs;   /* i-side only */
} sphk_dens_src;

typedef struct {  /* Density, hdr_density.hpp:4-22 */
    double *dens, *divv, *rotv, *grdh, *ksr;
    int64_t *np;
    uint8_t *itr;
} sphk_dens_out;

typedef struct {  /* HydroEPI / HydroEPJ, hdr_hydro.hpp:28-126 (m unused on the i-side) */
    const int32_t *id;
    const double *m, *x, *y, *z, *vx, *vy, *vz;
    const double *hinv, *pres, *presc, *bswt, *dens, *vsnd, *alph, *alphu, *thrm, *eta;
} sphk_hydro_src;

typedef struct {  /* Hydro, hdr_hydro.hpp:3-17 */
    double *ax, *ay, *az;  /* acch */
    double *gx, *gy, *gz;  /* accg */
    double *udot, *vsmx, *diffu;
} sphk_hydro_out;

typedef struct {  /* GravityEPI/EPJ and PS::GravitySPJ, hdr_gravity.hpp:63-192 (m unused on the i-side) */
    const double *m, *x, *y, *z, *eps2;
} sphk_grav_src;

typedef struct {  /* Gravity, hdr_gravity.hpp:95-105 */
    double *ax, *ay, *az, *pot, *eta;
} sphk_grav_out;

/* ---- context ---------------------------------------------------------- */

/* Select the device and create a context (one stream, buffers per tag).
 * Replaces setup_gpu (gpu_info.cu:6-34): device = my_proc % device_count. */
int sphk_init(int my_proc, sphk_ctx **ctx);
/* Release a context.  Replaces reset_gpu (gpu_info.cu:36-43). */
int sphk_finalize(sphk_ctx *ctx);
int sphk_set_params(sphk_ctx *ctx, const sphk_params *p);
int sphk_get_params(const sphk_ctx *ctx, sphk_params *p);
const char *sphk_last_error(void);
/* Device ordinal of the context and the hipStream_t it launches on. */
int sphk_device(const sphk_ctx *ctx);
void *sphk_stream(const sphk_ctx *ctx);
/* Block until all work the context launched has finished. */
int sphk_synchronize(sphk_ctx *ctx);

/* ---- device-resident launches (pointers are device pointers) ---------- */

/* Density functor over n_walk walks; replaces DensKernel
 * (hdr_densityGPU.cu:454-484) launched by DispatchKernel_Dens (:496-592). */
int sphk_density_dev(sphk_ctx *ctx, int n_walk, const int32_t *i_off, const int32_t *j_off,
                     const int32_t *j_idx, sphk_dens_src pi, sphk_dens_src pj,
                     sphk_dens_out out, int precision, void *stream);

/* Hydro functor; replaces HydroForceKernel (hdr_hydroGPU.cu:319-363)
 * launched by DispatchKernel_HydroForce (:371-477). */
int sphk_hydro_dev(sphk_ctx *ctx, int n_walk, const int32_t *i_off, const int32_t *j_off,
                   const int32_t *j_idx, sphk_hydro_src pi, sphk_hydro_src pj,
                   sphk_hydro_out out, int precision, void *stream);

/* Gravity functor over the EPJ list then the SPJ list of each walk; replaces
 * GravityForceKernel (hdr_gravityGPU.cu:330-374) launched by
 * DispatchKernel_Gravity (:383-469).  s_off may be NULL (no SPJ).  Outputs
 * are assigned (as the reference's GPU Retrieve does, :485-489).  out.eta
 * may be NULL: the softening-gradient sum is then not formed (callers that
 * do not use it, e.g. under a constant softening). */
int sphk_gravity_dev(sphk_ctx *ctx, int n_walk, const int32_t *i_off,
                     const int32_t *j_off, const int32_t *j_idx, sphk_grav_src pj,
                     const int32_t *s_off, const int32_t *s_idx, sphk_grav_src ps,
                     sphk_grav_src pi, sphk_grav_out out, int precision, void *stream);

/* ---- host-buffer dispatch / retrieve (pointers are host pointers) ------ */
/* i- and j-arrays are packed per walk (j_idx implicit), offsets are counts
 * per walk like the reference's nip/njp arguments.  At most SPHK_TAG_MAX
 * batches may be in flight (one per tag). */

/* DispatchKernel_Dens, hdr_densityGPU.cu:496-592 / use_gpu.hpp:28-33. */
int sphk_density_dispatch(sphk_ctx *ctx, int tag, int n_walk, const int32_t *nip,
                          const int32_t *njp, sphk_dens_src pi, sphk_dens_src pj, int precision);
/* RetrieveKernel_Dens, hdr_densityGPU.cu:594-621 / use_gpu.hpp:35-38. */
int sphk_density_retrieve(sphk_ctx *ctx, int tag, int n_walk, const int32_t *nip,
                          sphk_dens_out out);
/* DispatchKernel_HydroForce, hdr_hydroGPU.cu:371-477 / use_gpu.hpp:40-45. */
int sphk_hydro_dispatch(sphk_ctx *ctx, int tag, int n_walk, const int32_t *nip,
                        const int32_t *njp, sphk_hydro_src pi, sphk_hydro_src pj, int precision);
/* RetrieveKernel_HydroForce, hdr_hydroGPU.cu:479-507 / use_gpu.hpp:47-50. */
int sphk_hydro_retrieve(sphk_ctx *ctx, int tag, int n_walk, const int32_t *nip,
                        sphk_hydro_out out);
/* DispatchKernel_Gravity, hdr_gravityGPU.cu:383-469 / use_gpu.hpp:52-59;
 * nsp may be NULL.  The reference packs EPJ then SPJ into one list per walk. */
int sphk_gravity_dispatch(sphk_ctx *ctx, int tag, int n_walk, const int32_t *nip,
                          const int32_t *njp, const int32_t *nsp, sphk_grav_src pi,
                          sphk_grav_src pj, sphk_grav_src ps, int precision);
/* RetrieveKernel_Gravity, hdr_gravityGPU.cu:471-495 / use_gpu.hpp:61-64. */
int sphk_gravity_retrieve(sphk_ctx *ctx, int tag, int n_walk, const int32_t *nip,
                          sphk_grav_out out);

/* ---- timing ----------------------------------------------------------- */
/* Average duration (ms) of the last `count` launches of kernel `which`
 * (0 density, 1 hydro, 2 gravity) measured with HIP events recorded on the
 * stream each launch used; enable with sphk_timing(ctx, 1). */
int sphk_timing(sphk_ctx *ctx, int enable);
int sphk_kernel_time(sphk_ctx *ctx, int which, double *total_ms, int *count);
int sphk_reset_kernel_time(sphk_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
