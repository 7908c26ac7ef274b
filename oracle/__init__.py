"""CPU oracle for the SPH interaction hot path -- TEST INFRASTRUCTURE ONLY.

A clean-room scalar FP64 restatement (sphk_oracle.c) of the reference's CPU
interaction functors and per-particle glue:

  kernel0th/kernel1st   sph_yamauchi/run/hdr_kernel.hpp:246-257
  density               sph_yamauchi/run/hdr_density.hpp:246-365 (scalar calcDensity)
  hydro                 sph_yamauchi/run/hdr_hydro.hpp:312-424  (scalar calcHydro)
  gravity               sph_yamauchi/run/hdr_gravity.hpp:286-342 (scalar calcGravity<TJ>)
  eos_balsara           hdr_hgas.hpp:318-344, hdr_sph.hpp:79-89
  hydro_prep            hdr_hydro.hpp:45-62
  gravity_writeback     hdr_gravity.hpp:107-113
  hydro_writeback_tail  hdr_hydro.hpp:19-26, hdr_sph.hpp:95-97, hdr_hgas.hpp:750-762
  timestep              hdr_hgas.hpp:477-488, main.cpp:241-273
  predict / correct     hdr_hgas.hpp:646-665
  energy / momentum     hdr_hgas.hpp:630-643
  ksr_max               hdr_util.hpp:7-34,489-505

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this module, and only as the checker.  The product never imports it.

Pinning: see the module docstrings of tests/test_oracle_*.py; every function
above is checked there against closed forms, invariants or brute force.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_P = C.c_void_p


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def build_planted(k: int, out_dir) -> Path:
    """A variant of the oracle with planted mistake k (sphk_oracle.c,
    ORC_PLANT), for tests that show a pinning check catches it."""
    out = Path(out_dir) / f"liboracle_plant{k}.so"
    subprocess.run(["gcc", "-O2", "-fno-fast-math", "-ffp-contract=off", "-fPIC", "-fopenmp",
                    "-std=c11", f"-DORC_PLANT={int(k)}", "-shared", "-o", str(out),
                    str(HERE / "sphk_oracle.c"), "-lm"], check=True)
    return out


class use:
    """Context manager: route the module's functions through another build
    of the oracle (e.g. build_planted's) while it is active."""

    def __init__(self, path):
        self.path = Path(path)

    def __enter__(self):
        global _lib
        self._saved = _lib
        _lib = _configure(C.CDLL(str(self.path)))
        return self

    def __exit__(self, *exc):
        global _lib
        _lib = self._saved
        return False


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = _configure(C.CDLL(str(LIB_PATH)))
    return _lib


def _configure(d):
    d.orc_kernel0th.restype = C.c_double
    d.orc_kernel0th.argtypes = [C.c_double]
    d.orc_kernel1st.restype = C.c_double
    d.orc_kernel1st.argtypes = [C.c_double]
    for f in (d.orc_rcp_4th, d.orc_rsqrt_4th):
        f.restype = C.c_double
        f.argtypes = [C.c_double]
    d.orc_timestep.restype = C.c_double
    d.orc_energy.restype = C.c_double
    d.orc_ksr_max.restype = C.c_double
    d.orc_unbound_mass.restype = None
    d.orc_unbound_mass.argtypes = [C.c_int, State, _P, C.c_int, C.c_double, C.POINTER(C.c_double)]
    return d


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class DensSrc(C.Structure):
    _fields_ = [(n, _P) for n in ("id", "m", "x", "y", "z", "vx", "vy", "vz", "ksr", "rs")]


class DensOut(C.Structure):
    _fields_ = [(n, _P) for n in ("dens", "divv", "rotv", "grdh", "ksr", "np", "itr")]


HYDRO_FIELDS = ("id", "m", "x", "y", "z", "vx", "vy", "vz", "hinv", "pres", "presc", "bswt",
                "dens", "vsnd", "alph", "alphu", "thrm", "eta")


class HydroSrc(C.Structure):
    _fields_ = [(n, _P) for n in HYDRO_FIELDS]


class HydroOut(C.Structure):
    _fields_ = [(n, _P) for n in ("ax", "ay", "az", "gx", "gy", "gz", "udot", "vsmx", "diffu")]


class GravSrc(C.Structure):
    _fields_ = [(n, _P) for n in ("m", "x", "y", "z", "eps2")]


class GravOut(C.Structure):
    _fields_ = [(n, _P) for n in ("ax", "ay", "az", "pot", "eta")]


STATE_FIELDS = ("m", "x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "vx2", "vy2", "vz2",
                "uene", "uene2", "udot", "alph", "alph2", "adot", "alphu", "alphu2", "adotu",
                "diffu", "ksr", "dens", "pres", "vsnd", "divv", "rotv", "bswt", "grdh", "vsmx",
                "pot", "eta", "achx", "achy", "achz", "g1x", "g1y", "g1z", "g2x", "g2y", "g2z")


class State(C.Structure):
    _fields_ = [(n, _P) for n in STATE_FIELDS]


class Params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("ninv", "ksrh", "G", "alpha_max", "alpha_min",
                                          "alphu_max", "alphu_min", "eps_u", "c_dt", "dt_max",
                                          "eps_const")]


def kernel0th(q: float) -> float:
    return lib().orc_kernel0th(float(q))


def kernel1st(q: float) -> float:
    return lib().orc_kernel1st(float(q))


def _walk_args(i_off, j_off, j_idx):
    i_off = _i32(i_off)
    j_off = _i32(j_off)
    j_idx = None if j_idx is None else _i32(j_idx)
    return i_off, j_off, j_idx


def density(i_off, j_off, j_idx, pi: dict, pj: dict, eta=1.6, ksrh=1.936492, hmax=1e30,
            nthreads=0, avx2=False, simd=False) -> dict:
    """Scalar calcDensity over walks.  pi/pj: dicts of SoA arrays.  avx2=True:
    the AVX2 branch's arithmetic (approximate reciprocals) instead; simd=True:
    the same arithmetic four i-particles per AVX2 vector (bitwise equal)."""
    i_off, j_off, j_idx = _walk_args(i_off, j_off, j_idx)
    keep = []

    def src(p, iside):
        d = {k: (_i32(p[k]) if k == "id" else _f64(p[k])) for k in
             ("id", "m", "x", "y", "z", "vx", "vy", "vz") + (("ksr", "rs") if iside else ())}
        keep.append(d)
        return DensSrc(**{k: _ptr(v) for k, v in d.items()})

    ni = int(i_off[-1])
    out = {k: np.zeros(ni) for k in ("dens", "divv", "rotv", "grdh", "ksr")}
    out["np"] = np.zeros(ni, np.int64)
    out["itr"] = np.zeros(ni, np.uint8)
    o = DensOut(**{k: _ptr(v) for k, v in out.items()})
    if simd:
        fn = _simd().orc_density_simd
    else:
        fn = lib().orc_density_avx2 if avx2 else lib().orc_density
    fn(len(i_off) - 1, _ptr(i_off), _ptr(j_off), _ptr(j_idx), src(pi, True),
                      src(pj, False), C.c_double(eta), C.c_double(ksrh), C.c_double(hmax), o,
                      int(nthreads))
    return out


def _simd():
    if not lib().orc_density_simd_supported():
        raise RuntimeError("the vector functors need a CPU with AVX2")
    return lib()


def hydro(i_off, j_off, j_idx, pi: dict, pj: dict, nthreads=0, simd=False) -> dict:
    """Scalar calcHydro over walks; simd=True: four i-particles per AVX2
    vector with the refined reciprocals (the reference's vector build)."""
    i_off, j_off, j_idx = _walk_args(i_off, j_off, j_idx)
    keep = []

    def src(p):
        d = {k: (_i32(p[k]) if k == "id" else _f64(p[k])) for k in HYDRO_FIELDS if k in p}
        keep.append(d)
        return HydroSrc(**{k: _ptr(v) for k, v in d.items()})

    ni = int(i_off[-1])
    out = {k: np.zeros(ni) for k in ("ax", "ay", "az", "gx", "gy", "gz", "udot", "vsmx", "diffu")}
    o = HydroOut(**{k: _ptr(v) for k, v in out.items()})
    fn = _simd().orc_hydro_simd if simd else lib().orc_hydro
    fn(len(i_off) - 1, _ptr(i_off), _ptr(j_off), _ptr(j_idx), src(pi), src(pj), o, int(nthreads))
    return out


def gravity(i_off, j_off, j_idx, pj: dict, pi: dict, s_off=None, s_idx=None, ps: dict | None = None,
            nthreads=0, simd=False) -> dict:
    """Scalar calcGravity over walks (EPJ then SPJ list); simd=True: four
    i-particles per AVX2 vector with rsqrt_4th (the reference's vector build)."""
    i_off, j_off, j_idx = _walk_args(i_off, j_off, j_idx)
    keep = []

    def src(p):
        if p is None:
            return GravSrc()
        d = {k: _f64(p[k]) for k in ("m", "x", "y", "z", "eps2") if k in p}
        keep.append(d)
        return GravSrc(**{k: _ptr(v) for k, v in d.items()})

    s_off = None if s_off is None else _i32(s_off)
    s_idx = None if s_idx is None else _i32(s_idx)
    ni = int(i_off[-1])
    out = {k: np.zeros(ni) for k in ("ax", "ay", "az", "pot", "eta")}
    o = GravOut(**{k: _ptr(v) for k, v in out.items()})
    fn = _simd().orc_gravity_simd if simd else lib().orc_gravity
    fn(len(i_off) - 1, _ptr(i_off), _ptr(j_off), _ptr(j_idx), src(pj),
       _ptr(s_off), _ptr(s_idx), src(ps), src(pi), o, int(nthreads))
    return out


def make_state(n: int, **init) -> dict:
    s = {k: np.zeros(n) for k in STATE_FIELDS}
    for k, v in init.items():
        s[k] = _f64(v).copy()
    return s


def _state(s: dict):
    for k in STATE_FIELDS:
        s[k] = _f64(s[k])
    return State(**{k: _ptr(s[k]) for k in STATE_FIELDS})


def params(**kw) -> Params:
    d = dict(ninv=1.0 / 1.5, ksrh=1.936492, G=1.0, alpha_max=2.0, alpha_min=0.1, alphu_max=0.0,
             alphu_min=0.0, eps_u=0.0, c_dt=0.1, dt_max=10.0, eps_const=0.0)
    d.update(kw)
    return Params(**d)


def eos_balsara(s: dict, p: Params) -> None:
    lib().orc_eos_balsara(len(s["m"]), _state(s), p)


def hydro_prep(s: dict) -> dict:
    n = len(s["m"])
    out = {k: np.zeros(n) for k in ("hinv", "presc", "bswt_half", "thrm")}
    lib().orc_hydro_prep(n, *(_ptr(_f64(s[k])) for k in ("ksr", "pres", "grdh", "dens", "bswt", "uene")),
                         *(_ptr(out[k]) for k in ("hinv", "presc", "bswt_half", "thrm")))
    return out


def gravity_writeback(s: dict, p: Params, g: dict) -> None:
    lib().orc_gravity_writeback(len(s["m"]), _state(s), p,
                                *(_ptr(_f64(g[k])) for k in ("ax", "ay", "az", "pot", "eta")))


def hydro_writeback_tail(s: dict, p: Params, h: dict) -> None:
    lib().orc_hydro_writeback_tail(len(s["m"]), _state(s), p,
                                   *(_ptr(_f64(h[k])) for k in ("ax", "ay", "az", "gx", "gy", "gz",
                                                                 "udot", "vsmx", "diffu")))


def timestep(s: dict, p: Params, dt_prev: float, time: float, dt_min: float = 1e-16) -> float:
    return lib().orc_timestep(len(s["m"]), _state(s), p, C.c_double(dt_prev), C.c_double(time),
                              C.c_double(dt_min))


def predict(s: dict, dt: float) -> None:
    lib().orc_predict(len(s["m"]), _state(s), C.c_double(dt))


def correct(s: dict, dt: float) -> None:
    lib().orc_correct(len(s["m"]), _state(s), C.c_double(dt))


def energy(s: dict) -> float:
    return lib().orc_energy(len(s["m"]), _state(s))


def momentum(s: dict) -> np.ndarray:
    out = (C.c_double * 3)()
    lib().orc_momentum(len(s["m"]), _state(s), out)
    return np.array(out[:])


def ksr_max(m, x, y, z, dens, istar=None, istar_sel=0) -> float:
    arrs = [_f64(a) for a in (m, x, y, z, dens)]
    ist = None if istar is None else _i32(istar)
    return lib().orc_ksr_max(len(arrs[0]), *(_ptr(a) for a in arrs), _ptr(ist), int(istar_sel))


def unbound_mass(s: dict, istar=None, star: int = -1, k: float = 20.0) -> np.ndarray:
    """{v_cg (3), unbound mass, smoothed unbound mass, selected mass}."""
    out = (C.c_double * 6)()
    ist = None if istar is None else _i32(istar)
    lib().orc_unbound_mass(len(s["m"]), _state(s), _ptr(ist), int(star), float(k), out)
    return np.array(out[:])
