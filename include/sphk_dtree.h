/*
 * sphk_dtree.h -- device-resident octree, walk groups and interaction lists
 * (C-ABI).
 *
 * The same tree and list policies as the host builder in sphk_tree.h (the
 * roles FDPS's TreeForForceShort / TreeForForceLong play for the reference,
 * sph_yamauchi/run/main.cpp:377-398), built on the GPU from device arrays:
 * Morton keys (21 bits per dimension over the bounding cube), a breadth-first
 * octree with contiguous children (leaves hold <= n_leaf particles), FP64
 * node monopoles, groups = the largest nodes holding <= n_group particles,
 * and gather / symmetry / long lists.  The node numbering, the walks, the
 * moments and the lists are identical to the host builder's for the same
 * input (every sum and comparison is evaluated in the same order with the
 * same rounding).
 *
 * All array arguments and results are device pointers.  Work is queued on
 * `stream` (NULL = the HIP null stream); calls return after the small
 * device-to-host reads they need (level sizes, list totals), so results are
 * ready for work queued later on the same stream.  Returned arrays are owned
 * by the handle and stay valid until sphk_dtree_free or the next build of
 * that list kind.  Return values: SPHK_OK or a negative SPHK_E_* code with a
 * message in sphk_last_error().
 */
#ifndef SPHK_DTREE_H
#define SPHK_DTREE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sphk_dtree sphk_dtree;

/* perm[k] = index of the particle that goes to sorted position k (ties in
 * key keep index order), as sphk_morton_order on the host.  Asynchronous:
 * perm is written by work queued on `stream`.  A per-device workspace is kept
 * between calls (grown as needed, held until the process ends). */
int sphk_morton_order_dev(int n, const double *x, const double *y, const double *z, int32_t *perm,
                          void *stream);

/* Octree over particles in Morton order; eps2 may be NULL (zero).  Fails
 * with SPHK_E_ARG if the particles are not in Morton order. */
int sphk_dtree_build(int n, const double *x, const double *y, const double *z, const double *m,
                     const double *eps2, int n_leaf, int n_group, void *stream, sphk_dtree **out);
/* The same build into an existing handle, reusing its device buffers (they
 * only grow), so that a tree rebuilt every step allocates nothing once its
 * buffers have reached the run's sizes.  Arrays returned by the handle
 * before are invalid afterwards.  On failure the handle holds no tree
 * (n_walk 0) and may only be rebuilt or freed. */
int sphk_dtree_rebuild(sphk_dtree *t, int n, const double *x, const double *y, const double *z,
                       const double *m, const double *eps2, int n_leaf, int n_group, void *stream);
void sphk_dtree_free(sphk_dtree *t);

int sphk_dtree_n_walk(const sphk_dtree *t);
int sphk_dtree_n_node(const sphk_dtree *t);
/* Walk table: n_walk + 1 offsets into the particle order (device). */
const int32_t *sphk_dtree_i_off(const sphk_dtree *t);

/* kind 0: gather (radius rs), 1: symmetry (radius h), 2: long (theta).
 * Each build walks the tree once into a fixed room of entries per walk in
 * the handle's scratch (2048 candidates per walk for kinds 0 and 1; 8192
 * particles and 2048 nodes for kind 2) and packs the result; if a walk
 * overflows its room (or the walk's small stack), the build runs again as a
 * count walk, a scan and a fill walk, and the handle doubles that room for
 * later builds.  The scratch (n_walk x room entries, at most 8 GiB) is kept
 * with the handle's other buffers.  One host read per build either way. */
int sphk_dtree_list_gather(sphk_dtree *t, const double *rs);
int sphk_dtree_list_symmetry(sphk_dtree *t, const double *h);
int sphk_dtree_list_long(sphk_dtree *t, double theta);
/* Device offsets and indices of a built list, held by the handle; n_j / n_s
 * receive the totals (s_* and n_s are NULL / 0 for kinds 0 and 1).  A list
 * still in its rooms is packed into the handle's arrays first. */
int sphk_dtree_list(sphk_dtree *t, int kind, const int32_t **j_off, const int32_t **j_idx,
                    const int32_t **s_off, const int32_t **s_idx, int64_t *n_j, int64_t *n_s);
/* The totals of a built list (n_s 0 for kinds 0 and 1). */
int sphk_dtree_list_size(const sphk_dtree *t, int kind, int64_t *n_j, int64_t *n_s);
/* A built list into the caller's device arrays (j_off and s_off n_walk + 1
 * entries, j_idx n_j, s_idx n_s; s_* only for kind 2), on the handle's
 * stream: from the rooms straight to the caller's arrays when the list was
 * built in one walk (one pass over the entries instead of a pack and a
 * copy), else a copy of the handle's arrays. */
int sphk_dtree_list_copy(sphk_dtree *t, int kind, int32_t *j_off, int32_t *j_idx, int32_t *s_off,
                         int32_t *s_idx);
/* Node monopoles (device, n_node entries each). */
int sphk_dtree_moments(const sphk_dtree *t, const double **m, const double **x, const double **y,
                       const double **z, const double **eps2);
/* Node structure (device, n_node entries each): particle range, first child
 * (-1 for a leaf), number of children, level, cell side. */
int sphk_dtree_nodes(const sphk_dtree *t, const int32_t **begin, const int32_t **end,
                     const int32_t **child, const int32_t **nchild, const int32_t **level,
                     const double **len);

/* Device-to-device copy of `bytes` on `stream` (for bindings that keep the
 * results in device buffers of their own). */
int sphk_copy_dev(void *dst, const void *src, size_t bytes, void *stream);

/* Reorder particle data in place after a Morton sort: for each of `count`
 * (<= 64) columns, col_c[k] = old col_c[perm[k]], k < n, where perm is a
 * permutation of 0..n-1 (as sphk_morton_order_dev writes) and column c holds
 * n elements of elem[c] = 1, 2, 4 or 8 bytes.  `work` is device scratch of at
 * least sphk_permute_work_bytes(n, count, elem) bytes, 256-byte aligned.
 * Asynchronous on `stream` (two launches for all columns). */
size_t sphk_permute_work_bytes(int n, int count, const int *elem);
int sphk_permute_dev(int n, const int32_t *perm, int count, void *const *cols, const int *elem, void *work,
                     size_t work_bytes, void *stream);

/* *flag (device int32) = 1 if any of the `count` (<= 8) pairs of n-element
 * FP64 device arrays a[c], b[c] differ in any bit, else 0 (the list rebuilds'
 * test for unchanged inputs).  Asynchronous on `stream`. */
int sphk_equal_dev(int n, int count, const double *const *a, const double *const *b, int32_t *flag,
                   void *stream);

#ifdef __cplusplus
}
#endif
#endif
