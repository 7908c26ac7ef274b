"""Host octree, walk groups and interaction lists (wraps include/sphk_tree.h).

These play the role FDPS plays for the reference (sph_yamauchi/run/main.cpp:377-398):
walks are Morton-contiguous particle ranges of <= n_group particles and each
walk gets a gather (density), symmetry (hydro) or long-range (gravity, EPJ +
monopole SPJ, opening angle theta) list.  The list policies are documented in
include/sphk_tree.h.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import _DP, _I32P, _P, check, lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


def morton_order(x, y, z) -> np.ndarray:
    x, y, z = _f64(x), _f64(y), _f64(z)
    perm = np.empty(len(x), np.int32)
    check(lib().sphk_morton_order(len(x), _p(x), _p(y), _p(z), _p(perm)), "sphk_morton_order")
    return perm


def sort_particles(p: dict) -> tuple[dict, np.ndarray]:
    """Return the particle dict permuted into Morton order and the permutation."""
    perm = morton_order(p["x"], p["y"], p["z"])
    return {k: np.ascontiguousarray(v[perm]) for k, v in p.items()}, perm


class HostTree:
    def __init__(self, x, y, z, m, eps2=None, n_leaf: int = 8, n_group: int = 64):
        self._arrs = [_f64(a) for a in (x, y, z, m)] + [None if eps2 is None else _f64(eps2)]
        h = _P()
        n = len(self._arrs[0])
        check(lib().sphk_tree_build(n, *(_p(a) for a in self._arrs), int(n_leaf), int(n_group),
                                    C.byref(h)), "sphk_tree_build")
        self._h = h
        self.n = n
        self.n_walk = lib().sphk_tree_n_walk(h)
        self.n_node = lib().sphk_tree_n_node(h)
        ioff = lib().sphk_tree_i_off(h)
        self.i_off = np.ctypeslib.as_array(ioff, shape=(self.n_walk + 1,)).copy()

    def group_fixed(self, size: int):
        """Walks = consecutive Morton chunks of `size` particles."""
        check(lib().sphk_tree_group_fixed(self._h, int(size)), "sphk_tree_group_fixed")
        self.n_walk = lib().sphk_tree_n_walk(self._h)
        ioff = lib().sphk_tree_i_off(self._h)
        self.i_off = np.ctypeslib.as_array(ioff, shape=(self.n_walk + 1,)).copy()

    def close(self):
        if self._h:
            lib().sphk_tree_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _get(self, kind: int):
        jo, ji, so, si = _I32P(), _I32P(), _I32P(), _I32P()
        check(lib().sphk_tree_list(self._h, kind, C.byref(jo), C.byref(ji), C.byref(so),
                                   C.byref(si)), "sphk_tree_list")
        j_off = np.ctypeslib.as_array(jo, shape=(self.n_walk + 1,)).copy()
        nj = int(j_off[-1])
        j_idx = np.ctypeslib.as_array(ji, shape=(nj,)).copy() if nj else np.zeros(0, np.int32)
        if kind != 2:
            return j_off, j_idx
        s_off = np.ctypeslib.as_array(so, shape=(self.n_walk + 1,)).copy()
        ns = int(s_off[-1])
        s_idx = np.ctypeslib.as_array(si, shape=(ns,)).copy() if ns else np.zeros(0, np.int32)
        return j_off, j_idx, s_off, s_idx

    def gather(self, rs):
        rs = _f64(rs)
        check(lib().sphk_tree_list_gather(self._h, _p(rs)), "sphk_tree_list_gather")
        return self._get(0)

    def symmetry(self, h):
        h = _f64(h)
        check(lib().sphk_tree_list_symmetry(self._h, _p(h)), "sphk_tree_list_symmetry")
        return self._get(1)

    def long(self, theta: float = 0.5):
        check(lib().sphk_tree_list_long(self._h, float(theta)), "sphk_tree_list_long")
        return self._get(2)

    def moments(self) -> dict:
        ps = [_DP() for _ in range(5)]
        check(lib().sphk_tree_moments(self._h, *(C.byref(p) for p in ps)), "sphk_tree_moments")
        return {k: np.ctypeslib.as_array(p, shape=(self.n_node,)).copy()
                for k, p in zip(("m", "x", "y", "z", "eps2"), ps)}


def brute_force_lists(n: int, x, y, z, group: int = 16):
    """All-pairs lists (every walk lists every particle) over fixed-size groups.

    For small-N parity tests: with complete lists the functors' sums are the
    exact all-pairs sums and no list policy is involved.
    """
    i_off = np.arange(0, n + group, group, dtype=np.int32)
    i_off[-1] = n
    i_off = np.unique(i_off)
    n_walk = len(i_off) - 1
    j_off = (np.arange(n_walk + 1, dtype=np.int64) * n).astype(np.int32)
    j_idx = np.tile(np.arange(n, dtype=np.int32), n_walk)
    return i_off, j_off, j_idx
