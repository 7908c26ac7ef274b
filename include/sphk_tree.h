/*
 * sphk_tree.h -- host tree, walk groups and interaction lists (C-ABI).
 *
 * The reference obtains its walks and interaction lists from FDPS (not
 * vendored): TreeForForceShort<...>::Gather for density, ::Symmetry for
 * hydro, TreeForForceLong with a monopole SPJ for gravity, set up with
 * n_leaf = 8, n_group = 256 and theta = 0.5 / 0.7 (sph_yamauchi/run/main.cpp:377-398)
 * and walked by calcForceAllMultiWalk / calcForceMakingTreeMultiWalk
 * (hdr_density.hpp:389, main.cpp:339, hdr_gravity.hpp:393).  This library
 * provides its own octree with the same roles; the list policies are:
 *
 *  - gather (density): j is listed for a group iff |x_j - x_i| < rs_i for at
 *    least one i-particle of the group (the union of the i search spheres;
 *    the tree walk prunes with the bounding box of those spheres).
 *  - symmetry (hydro): j is listed iff r_ij < max(H_i, H_j) for at least one
 *    i of the group (pruned with the box of the i-spheres and the boxes
 *    x_j +- H_j).
 *  - long (gravity): a node is accepted as one SPJ when the squared distance
 *    from its centre of mass to the group's bounding box exceeds
 *    (L_node / theta)^2, L_node the node's cubic cell side; otherwise it is
 *    opened; particles of opened leaves are listed as EPJ (the group itself
 *    is therefore always in its own EPJ list).  SPJ = FP64 monopole
 *    {mass, centre of mass, mass-weighted eps2} (hdr_gravity.hpp:6-91 keeps
 *    the moment in FP32; this tree keeps it in FP64).
 *
 * Particles must be supplied in Morton order (sphk_morton_order gives the
 * permutation); walks are then contiguous index ranges of that order.
 * All pointers are host pointers.  Returned arrays are owned by the handle
 * and stay valid until sphk_tree_free / the next list build of that kind.
 */
#ifndef SPHK_TREE_H
#define SPHK_TREE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sphk_tree sphk_tree;

/* perm[k] = index of the particle that goes to sorted position k. */
int sphk_morton_order(int n, const double *x, const double *y, const double *z, int32_t *perm);

/* Octree over sorted particles (leaves hold <= n_leaf), monopoles from m and
 * eps2 (eps2 may be NULL: zero).  Groups are the largest nodes holding
 * <= n_group particles. */
int sphk_tree_build(int n, const double *x, const double *y, const double *z, const double *m,
                    const double *eps2, int n_leaf, int n_group, sphk_tree **out);
void sphk_tree_free(sphk_tree *t);
/* Replace the walks by fixed chunks of `size` consecutive (Morton-ordered)
 * particles, the last one possibly shorter.  Invalidates built lists. */
int sphk_tree_group_fixed(sphk_tree *t, int size);

int sphk_tree_n_walk(const sphk_tree *t);
int sphk_tree_n_node(const sphk_tree *t);
const int32_t *sphk_tree_i_off(const sphk_tree *t);

/* Lists.  kind: 0 gather (radius = rs), 1 symmetry (radius = H), 2 long
 * (theta).  After a build, sphk_tree_list() returns the offsets/indices. */
int sphk_tree_list_gather(sphk_tree *t, const double *rs);
int sphk_tree_list_symmetry(sphk_tree *t, const double *h);
int sphk_tree_list_long(sphk_tree *t, double theta);
/* j_off has n_walk+1 entries, j_idx has j_off[n_walk]; for kind 2, s_off /
 * s_idx index the node moments (s_* are NULL for kinds 0 and 1). */
int sphk_tree_list(const sphk_tree *t, int kind, const int32_t **j_off, const int32_t **j_idx,
                   const int32_t **s_off, const int32_t **s_idx);
/* Node monopoles: mass, centre of mass, eps2 (n_node entries each). */
int sphk_tree_moments(const sphk_tree *t, const double **m, const double **x, const double **y,
                      const double **z, const double **eps2);

#ifdef __cplusplus
}
#endif
#endif
