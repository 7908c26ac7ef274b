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


# --------------------------------------------------------------------------
# Device-resident tree (include/sphk_dtree.h): same tree, walks and lists as
# HostTree, built on the GPU from torch tensors.
# --------------------------------------------------------------------------
def _dptr(t):
    import torch
    if t is None:
        return None
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()):
        raise ValueError("expected a contiguous CUDA tensor")
    return t.data_ptr()


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def _copy_out(ptr, n: int, dtype):
    """Copy n elements at device address ptr into a new tensor (on the current stream)."""
    import torch
    out = torch.empty(n, dtype=dtype, device="cuda")
    if n:
        check(lib().sphk_copy_dev(out.data_ptr(), ptr, n * out.element_size(), _stream()),
              "sphk_copy_dev")
    return out


def permute_dev(perm, cols, work=None):
    """Reorder CUDA tensors in place, col[k] = old col[perm[k]], with perm an
    int32 permutation tensor (sphk_permute_dev: two launches for all columns).
    `work` is reused scratch (a uint8 CUDA tensor, grown when too small);
    returns the scratch for the next call."""
    import torch
    n, cnt = perm.numel(), len(cols)
    if any(c.numel() != n for c in cols) or perm.dtype != torch.int32:
        raise ValueError("permute_dev: columns of the permutation's length and an int32 permutation expected")
    elem = (C.c_int * cnt)(*[c.element_size() for c in cols])
    need = int(lib().sphk_permute_work_bytes(n, cnt, elem))
    if work is None or work.numel() < need:
        work = torch.empty(need, dtype=torch.uint8, device=perm.device)
    ptrs = (_P * cnt)(*[_dptr(c) for c in cols])
    check(lib().sphk_permute_dev(n, _dptr(perm), cnt, ptrs, elem, work.data_ptr(), work.numel(), _stream()),
          "sphk_permute_dev")
    return work


def morton_order_dev(x, y, z):
    """Device permutation into Morton order (int64 tensor), as morton_order."""
    import torch
    n = x.numel()
    perm = torch.empty(n, dtype=torch.int32, device=x.device)
    check(lib().sphk_morton_order_dev(n, _dptr(x), _dptr(y), _dptr(z), _dptr(perm), _stream()),
          "sphk_morton_order_dev")
    return perm.long()


class DeviceTree:
    """The octree, walks and lists of HostTree, built and held on the GPU.

    Inputs are float64 CUDA tensors in Morton order; outputs are tensors on
    the same device (copies owned by the caller)."""

    def __init__(self, x, y, z, m, eps2=None, n_leaf: int = 8, n_group: int = 64):
        h = _P()
        n = x.numel()
        self.n = n
        check(lib().sphk_dtree_build(n, _dptr(x), _dptr(y), _dptr(z), _dptr(m), _dptr(eps2),
                                     int(n_leaf), int(n_group), _stream(), C.byref(h)),
              "sphk_dtree_build")
        self._h = h
        self._read_walks()
        self._src = tuple(None if a is None else a.clone() for a in (x, y, z, m, eps2)) + (int(n_leaf),
                                                                                          int(n_group))

    def _read_walks(self):
        import torch
        h = self._h
        self.n_walk = lib().sphk_dtree_n_walk(h)
        self.n_node = lib().sphk_dtree_n_node(h)
        self.i_off = _copy_out(lib().sphk_dtree_i_off(h), self.n_walk + 1, torch.int32)

    def rebuild(self, x, y, z, m, eps2=None, n_leaf: int = 8, n_group: int = 64):
        """Build again for new particles into this handle's device buffers
        (no allocation once they have reached the run's sizes).  Inputs equal
        to the last build's (kept as copies) leave the tree as it is."""
        import torch
        if not self._h:
            raise RuntimeError("tree is closed")
        src = (x, y, z, m, eps2, int(n_leaf), int(n_group))
        last = getattr(self, "_src", None)
        if last is not None and last[5:] == src[5:] and all(
                (a is None) == (b is None) and (a is None or a.shape == b.shape)
                for a, b in zip(last[:5], src[:5])):
            # one device-side bitwise comparison of all inputs, one host read
            pairs = [(a, b) for a, b in zip(last[:5], src[:5]) if a is not None]
            if any(b.dtype != torch.float64 for _, b in pairs):
                raise ValueError("DeviceTree inputs must be float64")
            if getattr(self, "_flag", None) is None:
                self._flag = torch.zeros(1, dtype=torch.int32, device=x.device)
            A = (_P * len(pairs))(*[a.data_ptr() for a, _ in pairs])
            B = (_P * len(pairs))(*[_dptr(b) for _, b in pairs])
            check(lib().sphk_equal_dev(x.numel(), len(pairs), A, B, self._flag.data_ptr(), _stream()),
                  "sphk_equal_dev")
            if int(self._flag.item()) == 0:
                return self
        self.n = x.numel()
        self._src = None
        check(lib().sphk_dtree_rebuild(self._h, self.n, _dptr(x), _dptr(y), _dptr(z), _dptr(m), _dptr(eps2),
                                       int(n_leaf), int(n_group), _stream()), "sphk_dtree_rebuild")
        self._read_walks()
        self._src = tuple(None if a is None else a.clone() for a in src[:5]) + src[5:]
        return self

    def close(self):
        if self._h:
            lib().sphk_dtree_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _get(self, kind: int):
        """The built list in new tensors (packed straight from the build's
        rooms when it was built in one walk)."""
        import torch
        nj, ns = C.c_int64(), C.c_int64()
        check(lib().sphk_dtree_list_size(self._h, kind, C.byref(nj), C.byref(ns)), "sphk_dtree_list_size")
        t32 = lambda n: torch.empty(max(n, 1), dtype=torch.int32, device="cuda")   # noqa: E731
        jo, ji = t32(self.n_walk + 1), t32(nj.value)
        so, si = (t32(self.n_walk + 1), t32(ns.value)) if kind == 2 else (None, None)
        check(lib().sphk_dtree_list_copy(self._h, kind, _dptr(jo), _dptr(ji), _dptr(so), _dptr(si)),
              "sphk_dtree_list_copy")
        out = [jo, ji[:nj.value]]
        if kind == 2:
            out += [so, si[:ns.value]]
        return tuple(out)

    def gather(self, rs):
        check(lib().sphk_dtree_list_gather(self._h, _dptr(rs)), "sphk_dtree_list_gather")
        return self._get(0)

    def symmetry(self, h):
        check(lib().sphk_dtree_list_symmetry(self._h, _dptr(h)), "sphk_dtree_list_symmetry")
        return self._get(1)

    def long(self, theta: float):
        check(lib().sphk_dtree_list_long(self._h, float(theta)), "sphk_dtree_list_long")
        return self._get(2)

    def moments(self) -> dict:
        import torch
        ps = [_P() for _ in range(5)]
        check(lib().sphk_dtree_moments(self._h, *(C.byref(p) for p in ps)), "sphk_dtree_moments")
        return {k: _copy_out(p.value, self.n_node, torch.float64)
                for k, p in zip(("m", "x", "y", "z", "eps2"), ps)}

    def nodes(self) -> dict:
        import torch
        ps = [_P() for _ in range(6)]
        check(lib().sphk_dtree_nodes(self._h, *(C.byref(p) for p in ps)), "sphk_dtree_nodes")
        names = ("begin", "end", "child", "nchild", "level", "len")
        return {k: _copy_out(p.value, self.n_node, torch.float64 if k == "len" else torch.int32)
                for k, p in zip(names, ps)}
