"""Device-resident launches of the SPH interaction functors (torch tensors in HBM).

torch provides device memory and the stream; the arithmetic is the HIP code in
libsphk.so.  Every launch goes on torch's current stream so it is ordered with
the surrounding tensor work.

The functor names and argument meaning follow the reference's multiwalk API
(sph_yamauchi/run/use_gpu.hpp:28-64): a launch takes a walk table (i-ranges),
a j-list per walk, the EPI and EPJ fields, and writes the Force fields.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from ._lib import check, lib

FP64 = L.SPHK_FP64
MIXED = L.SPHK_MIXED

_DT = {"id": torch.int32, "np": torch.int64, "itr": torch.uint8}


def _ptr(t: torch.Tensor | None, name: str, dtype=None):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError(f"{name}: tensor must be on the GPU")
    want = dtype or _DT.get(name, torch.float64)
    if t.dtype != want:
        raise TypeError(f"{name}: dtype {t.dtype}, expected {want}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor must be contiguous")
    return t.data_ptr()


def _mk(cls, fields, d: dict | None):
    d = d or {}
    return cls(**{f: _ptr(d.get(f), f) for f in fields})


def to_device(a: np.ndarray | torch.Tensor, device="cuda") -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.to(device).contiguous()
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


class Walks:
    """A walk table plus one list per walk, resident on the device."""

    def __init__(self, i_off, j_off, j_idx=None, s_off=None, s_idx=None, device="cuda"):
        if isinstance(i_off, torch.Tensor):
            self._from_tensors(i_off, j_off, j_idx, s_off, s_idx)
            return
        self.n_walk = len(i_off) - 1
        self.i_off = to_device(np.asarray(i_off, np.int32), device)
        self.j_off = to_device(np.asarray(j_off, np.int32), device)
        self.j_idx = None if j_idx is None else to_device(np.asarray(j_idx, np.int32), device)
        self.s_off = None if s_off is None else to_device(np.asarray(s_off, np.int32), device)
        self.s_idx = None if s_idx is None else to_device(np.asarray(s_idx, np.int32), device)
        self.n_i = int(i_off[-1])
        self.n_j = int(j_off[-1])
        self.n_s = 0 if s_off is None else int(s_off[-1])
        io = np.asarray(i_off, np.int64)
        jo = np.asarray(j_off, np.int64)
        ni = np.diff(io)
        # interaction-list entries evaluated: sum over walks of n_i * n_j
        self.entries = int((ni * np.diff(jo)).sum())
        if s_off is not None:
            self.entries_spj = int((ni * np.diff(np.asarray(s_off, np.int64))).sum())
        else:
            self.entries_spj = 0


    def _from_tensors(self, i_off, j_off, j_idx, s_off, s_idx):
        """Lists already on the device (int32 tensors, e.g. from DeviceTree)."""
        for t in (i_off, j_off, j_idx, s_off, s_idx):
            if t is not None and (t.dtype != torch.int32 or not t.is_cuda):
                raise TypeError("device walk tables must be int32 CUDA tensors")
        self.n_walk = i_off.numel() - 1
        self.i_off, self.j_off, self.j_idx, self.s_off, self.s_idx = i_off, j_off, j_idx, s_off, s_idx
        self._sizes = None   # read from the device on first use (a host synchronisation)

    def _read_sizes(self):
        if self._sizes is None:
            i_off, j_off, s_off = self.i_off, self.j_off, self.s_off
            ni = (i_off[1:] - i_off[:-1]).long()
            v = [i_off[-1].long(), j_off[-1].long(), (ni * (j_off[1:] - j_off[:-1]).long()).sum()]
            if s_off is not None:
                v += [s_off[-1].long(), (ni * (s_off[1:] - s_off[:-1]).long()).sum()]
            h = torch.stack(v).cpu().tolist()
            self._sizes = dict(n_i=h[0], n_j=h[1], entries=h[2], n_s=h[3] if len(h) > 3 else 0,
                               entries_spj=h[4] if len(h) > 3 else 0)
        return self._sizes

    def __getattr__(self, name):
        # sizes of device-built tables (interaction-list entries = sum over
        # walks of n_i * n_j), computed lazily: list builds need not wait
        if name in ("n_i", "n_j", "n_s", "entries", "entries_spj") and "_sizes" in self.__dict__:
            return self._read_sizes()[name]
        raise AttributeError(name)


class Context:
    """One device context: sphk_init(my_proc) ... sphk_finalize."""

    def __init__(self, my_proc: int = 0, eta: float = 1.6, ksrh: float = 1.936492,
                 hmax: float = 1e30):
        h = C.c_void_p()
        check(lib().sphk_init(int(my_proc), C.byref(h)), "sphk_init")
        self._h = h
        self.set_params(eta, ksrh, hmax)

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().sphk_finalize(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, eta=None, ksrh=None, hmax=None, uniform_softening=None, list_cap=None,
                   persistent=None):
        p = L.Params()
        if getattr(self, "_params", None) is not None:
            p = self._params
        if eta is not None:
            p.eta = float(eta)
        if ksrh is not None:
            p.ksrh = float(ksrh)
        if hmax is not None:
            p.hmax = float(hmax)
        if uniform_softening is not None:
            p.uniform_softening = int(bool(uniform_softening))
        if list_cap is not None:
            p.list_cap = int(list_cap)
        if persistent is not None:
            p.persistent = int(bool(persistent))
        check(lib().sphk_set_params(self._h, C.byref(p)), "sphk_set_params")
        self._params = p

    @property
    def params(self):
        return dict(eta=self._params.eta, ksrh=self._params.ksrh, hmax=self._params.hmax,
                    uniform_softening=bool(self._params.uniform_softening), list_cap=self._params.list_cap,
                    persistent=bool(self._params.persistent))

    @staticmethod
    def _stream():
        return torch.cuda.current_stream().cuda_stream

    def density(self, w: Walks, pi: dict, pj: dict, out: dict, precision: int = FP64):
        check(lib().sphk_density_dev(self._h, w.n_walk, w.i_off.data_ptr(), w.j_off.data_ptr(),
                                     None if w.j_idx is None else w.j_idx.data_ptr(),
                                     _mk(L.DensSrc, L.DENS_SRC, pi), _mk(L.DensSrc, L.DENS_SRC, pj),
                                     _mk(L.DensOut, L.DENS_OUT, out), int(precision),
                                     self._stream()), "sphk_density_dev")

    def hydro(self, w: Walks, pi: dict, pj: dict, out: dict, precision: int = FP64):
        check(lib().sphk_hydro_dev(self._h, w.n_walk, w.i_off.data_ptr(), w.j_off.data_ptr(),
                                   None if w.j_idx is None else w.j_idx.data_ptr(),
                                   _mk(L.HydroSrc, L.HYDRO_SRC, pi), _mk(L.HydroSrc, L.HYDRO_SRC, pj),
                                   _mk(L.HydroOut, L.HYDRO_OUT, out), int(precision),
                                   self._stream()), "sphk_hydro_dev")

    def hydro_vjp(self, w: Walks, src: dict, cot: dict, grad: dict):
        """Add to grad (HYDRO_GRAD fields) the vector-Jacobian product of the
        FP64 hydro functor over symmetric lists w, for output cotangents cot
        (HYDRO_OUT fields; vsmx unused)."""
        check(lib().sphk_hydro_vjp_dev(self._h, w.n_walk, w.i_off.data_ptr(), w.j_off.data_ptr(),
                                       None if w.j_idx is None else w.j_idx.data_ptr(),
                                       _mk(L.HydroSrc, L.HYDRO_SRC, src),
                                       _mk(L.HydroOut, L.HYDRO_OUT, {k: v for k, v in cot.items() if k != "vsmx"}),
                                       _mk(L.HydroGrad, L.HYDRO_GRAD, grad), self._stream()),
              "sphk_hydro_vjp_dev")

    def density_vjp(self, w: Walks, src: dict, cot: dict, grad: dict):
        """Add to grad (DENS_GRAD fields) the vector-Jacobian product of the
        FP64 density functor over gather lists w (one particle set for both
        sides), for output cotangents cot (DENS_COT fields), with this
        context's eta, kappa and H_max."""
        check(lib().sphk_density_vjp_dev(self._h, w.n_walk, w.i_off.data_ptr(), w.j_off.data_ptr(),
                                         None if w.j_idx is None else w.j_idx.data_ptr(),
                                         _mk(L.DensSrc, L.DENS_SRC, src), _mk(L.DensCot, L.DENS_COT, cot),
                                         _mk(L.DensGrad, L.DENS_GRAD, grad), self._stream()),
              "sphk_density_vjp_dev")

    def gravity(self, w: Walks, pj: dict, pi: dict, out: dict, ps: dict | None = None,
                precision: int = FP64):
        check(lib().sphk_gravity_dev(self._h, w.n_walk, w.i_off.data_ptr(), w.j_off.data_ptr(),
                                     None if w.j_idx is None else w.j_idx.data_ptr(),
                                     _mk(L.GravSrc, L.GRAV_SRC, pj),
                                     None if w.s_off is None else w.s_off.data_ptr(),
                                     None if w.s_idx is None else w.s_idx.data_ptr(),
                                     _mk(L.GravSrc, L.GRAV_SRC, ps), _mk(L.GravSrc, L.GRAV_SRC, pi),
                                     _mk(L.GravOut, L.GRAV_OUT, out), int(precision),
                                     self._stream()), "sphk_gravity_dev")

    # --- timing (HIP events recorded around each launch on its stream) ---
    def timing(self, enable: bool = True):
        check(lib().sphk_timing(self._h, int(bool(enable))), "sphk_timing")
        self._timing = bool(enable)

    def kernel_time(self, which: int) -> tuple[float, int]:
        ms = C.c_double()
        cnt = C.c_int()
        check(lib().sphk_kernel_time(self._h, int(which), C.byref(ms), C.byref(cnt)),
              "sphk_kernel_time")
        return ms.value, cnt.value

    def reset_kernel_time(self):
        check(lib().sphk_reset_kernel_time(self._h), "sphk_reset_kernel_time")


def empty_dens_out(n: int, device="cuda") -> dict:
    out = {k: torch.empty(n, dtype=torch.float64, device=device)
           for k in ("dens", "divv", "rotv", "grdh", "ksr")}
    out["np"] = torch.empty(n, dtype=torch.int64, device=device)
    out["itr"] = torch.empty(n, dtype=torch.uint8, device=device)
    return out


def empty_out(fields, n: int, device="cuda") -> dict:
    return {k: torch.empty(n, dtype=torch.float64, device=device) for k in fields}
