"""On-disk formats of the reference's run directory, read and written in numpy.

* run header `.head` (RunParameter::readAscii, sph_yamauchi/run/hdr_hgas.hpp:775-795):
  ``Time TimeEnd dtAscii dtHexa / FlagDamping FlagNuclear / alpha_max alpha_min /
  alphu_max alphu_min / C_dt / N``;
* initial particle data `.data` (HelmholtzGas::readAscii, hdr_hgas.hpp:115-134):
  ``id istar m x y z vx vy vz u alpha alphu ksr cmps[13]`` in cgs;
* ASCII snapshot (HelmholtzGas::writeAscii, hdr_hgas.hpp:136-178; read by
  analysis/common.py:35-69): 49 columns, cgs;
* hexadecimal checkpoint (HelmholtzGas::writeHexa, hdr_hgas.hpp:286-316): every
  FP64 field as the hex of its bit pattern, so a restart is bit-exact;
* `time.log` lines (hdr_hgas.hpp:1070-1075):
  ``time: step t dt E-E_nuc-E_abs E E_nuc E_abs H_max wall``.

Units: the reference's code units (hdr_heos.hpp:54-104) are L = 1e-3 R_sun,
M = 1e-6 M_sun, T = 1 s; `CodeUnits` converts between them and cgs.  Runs of
this package use G = M = R = 1 units; `CodeUnits.nbody(M, R)` gives the
matching unit system for a star of mass M [g] and radius R [cm].
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

SOLAR_RADIUS = 6.9599e10        # cm     (hdr_heos.hpp:54)
SOLAR_MASS = 1.9891e33          # g      (hdr_heos.hpp:55)
GRAVITY_CONSTANT = 6.6738480e-8  # cgs   (hdr_heos.hpp:56)
N_CMPS = 13                     # NumberOfNucleon (hdr_heos.hpp:115)


@dataclass(frozen=True)
class CodeUnits:
    length: float = SOLAR_RADIUS * 1e-3
    mass: float = SOLAR_MASS * 1e-6
    time: float = 1.0

    @property
    def velocity(self):
        return self.length / self.time

    @property
    def acceleration(self):
        return self.velocity / self.time

    @property
    def energy(self):          # specific energy
        return self.velocity * self.velocity

    @property
    def density(self):
        return self.mass / self.length ** 3

    @property
    def pressure(self):
        return self.mass * self.length / (self.time * self.time * self.length * self.length)

    @property
    def gravity_constant(self):
        """G in these units (GravityConstantInThisUnit, hdr_heos.hpp:102-104)."""
        return GRAVITY_CONSTANT / (self.length ** 3 / self.mass / self.time ** 2)

    @staticmethod
    def nbody(mass: float, radius: float) -> "CodeUnits":
        """Units with G = 1 for mass [g] and length [cm]."""
        return CodeUnits(radius, mass, math.sqrt(radius ** 3 / (GRAVITY_CONSTANT * mass)))


REFERENCE_UNITS = CodeUnits()

# --------------------------------------------------------------------- .head
HEAD_KEYS = ("time", "time_end", "dt_ascii", "dt_hexa", "flag_damping", "flag_nuclear",
             "alpha_max", "alpha_min", "alphu_max", "alphu_min", "c_dt", "n")


def read_head(path, units: CodeUnits = REFERENCE_UNITS) -> dict:
    tok = open(path).read().split()
    if len(tok) < 12:
        raise ValueError(f"{path}: expected 12 header values, found {len(tok)}")
    h = {}
    for k, t in zip(HEAD_KEYS, tok[:12]):
        h[k] = int(t) if k in ("flag_damping", "flag_nuclear", "n") else float(t)
    for k in ("time", "time_end", "dt_ascii", "dt_hexa"):
        h[k] /= units.time
    return h


def write_head(path, h: dict, units: CodeUnits = REFERENCE_UNITS):
    t = {k: h[k] * units.time for k in ("time", "time_end", "dt_ascii", "dt_hexa")}
    with open(path, "w") as f:
        f.write(f"{t['time']:+.16e} {t['time_end']:+.16e} {t['dt_ascii']:+.16e} {t['dt_hexa']:+.16e}\n")
        f.write(f"{int(h['flag_damping'])} {int(h['flag_nuclear'])}\n")
        f.write(f"{h['alpha_max']:+.16e} {h['alpha_min']:+.16e}\n")
        f.write(f"{h['alphu_max']:+.16e} {h['alphu_min']:+.16e}\n")
        f.write(f"{h['c_dt']:+.16e}\n")
        f.write(f"{int(h['n'])}\n")


# --------------------------------------------------------------------- .data
DATA_COLUMNS = ("id", "istar", "m", "x", "y", "z", "vx", "vy", "vz", "uene", "alph", "alphu", "ksr")


def _data_scale(units: CodeUnits) -> dict:
    return {"m": units.mass, "x": units.length, "y": units.length, "z": units.length,
            "vx": units.velocity, "vy": units.velocity, "vz": units.velocity,
            "uene": units.energy, "ksr": units.length}


def read_data(path, units: CodeUnits = REFERENCE_UNITS, n_cmps: int = N_CMPS) -> dict:
    """Particles of a `.data` file, converted from cgs to `units`."""
    a = np.loadtxt(path, ndmin=2)
    if a.shape[1] < len(DATA_COLUMNS) + n_cmps:
        raise ValueError(f"{path}: {a.shape[1]} columns, expected {len(DATA_COLUMNS) + n_cmps}")
    p = {k: a[:, c].copy() for c, k in enumerate(DATA_COLUMNS)}
    p["id"] = p["id"].astype(np.int64)
    p["istar"] = p["istar"].astype(np.int32)
    for k, s in _data_scale(units).items():
        p[k] = p[k] / s
    p["cmps"] = a[:, len(DATA_COLUMNS):len(DATA_COLUMNS) + n_cmps].copy()
    return p


def write_data(path, p: dict, units: CodeUnits = REFERENCE_UNITS, n_cmps: int = N_CMPS):
    n = len(p["m"])
    sc = _data_scale(units)
    cmps = p.get("cmps", np.zeros((n, n_cmps)))
    with open(path, "w") as f:
        for i in range(n):
            vals = [f"{int(p['id'][i])}", f"{int(p['istar'][i])}"]
            vals += [f"{p[k][i] * sc.get(k, 1.0):+.16e}" for k in DATA_COLUMNS[2:]]
            vals += [f"{c:+.16e}" for c in cmps[i]]
            f.write(" ".join(vals) + "\n")


# ------------------------------------------------------------- ASCII snapshot
SNAPSHOT_COLUMNS = (
    "id", "istar", "m", "x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az",
    "uene", "alph", "alphu", "dens", "ksr", "np", "vsnd", "pres", "temp",
    "divv", "rotv", "bswt", "pot", "abar", "zbar", "enuc", "vsmx", "udot", "dnuc",
) + tuple(f"cmps{k}" for k in range(N_CMPS)) + ("pot3", "tempmax0", "tempmax1", "tempmax2", "entr")
assert len(SNAPSHOT_COLUMNS) == 49


def _snap_scale(units: CodeUnits) -> dict:
    L, V = units.length, units.velocity
    s = {"m": units.mass, "x": L, "y": L, "z": L, "vx": V, "vy": V, "vz": V,
         "ax": units.acceleration, "ay": units.acceleration, "az": units.acceleration,
         "uene": units.energy, "dens": units.density, "ksr": L, "vsnd": V, "pres": units.pressure,
         "divv": 1.0 / units.time, "rotv": 1.0 / units.time, "pot": units.energy, "enuc": units.energy,
         "vsmx": V, "udot": units.energy / units.time, "dnuc": units.energy, "pot3": units.energy,
         "tempmax1": units.density, "entr": units.energy}
    return s


def write_snapshot(path, p: dict, units: CodeUnits = REFERENCE_UNITS):
    """The 49-column ASCII snapshot (HelmholtzGas::writeAscii formats:
    %6d %2d then %+e, np as %6d, composition as %+.3e).  Fields missing from
    p (temperature, nuclear and composition quantities of the Helmholtz gas)
    are written as zero."""
    n = len(p["m"])
    sc = _snap_scale(units)

    def col(k):
        v = p.get(k)
        if v is None:
            return np.zeros(n)
        return np.asarray(v, np.float64) * sc.get(k, 1.0)

    cols = {k: col(k) for k in SNAPSHOT_COLUMNS}
    ids = np.asarray(p["id"], np.int64)
    ist = np.asarray(p.get("istar", np.zeros(n)), np.int64)
    nps = np.asarray(p.get("np", np.zeros(n)), np.int64)
    with open(path, "w") as f:
        for i in range(n):
            out = [f"{ids[i]:6d}", f"{ist[i]:2d}"]
            for k in SNAPSHOT_COLUMNS[2:]:
                if k == "np":
                    out.append(f"{nps[i]:6d}")
                elif k.startswith("cmps"):
                    out.append(f"{cols[k][i]:+.3e}")
                else:
                    out.append(f"{cols[k][i]:+e}")
            f.write(" ".join(out) + "\n")


def read_snapshot(path, units: CodeUnits = REFERENCE_UNITS) -> dict:
    """Columns of an ASCII snapshot by name (analysis/common.py:35-69), in `units`."""
    a = np.loadtxt(path, ndmin=2)
    sc = _snap_scale(units)
    p = {k: a[:, c] / sc.get(k, 1.0) for c, k in enumerate(SNAPSHOT_COLUMNS) if c < a.shape[1]}
    for k in ("id", "istar", "np"):
        if k in p:
            p[k] = p[k].astype(np.int64)
    return p


# --------------------------------------------------------- hexadecimal dump
HEXA_COLUMNS = (
    "m", "x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "achx", "achy", "achz",
    "g1x", "g1y", "g1z", "g2x", "g2y", "g2z", "uene", "udot", "alph", "adot", "alphu", "adotu",
    "dens", "pres", "vsnd", "temp", "divv", "rotv", "bswt", "ksr", "grdh", "vsmx", "pot", "enuc",
    "dnuc", "abar", "zbar", "gcor3",
) + tuple(f"cmps{k}" for k in range(N_CMPS)) + ("tempmax0", "tempmax1", "tempmax2")


def _hex(v: np.ndarray) -> list[str]:
    return [format(int(u), "x") for u in np.asarray(v, np.float64).view(np.uint64)]


def write_hexa(path, p: dict):
    """Checkpoint with every FP64 field as the hex of its bit pattern (code
    units, as the reference writes it); id and istar in decimal."""
    n = len(p["m"])
    cols = [_hex(p.get(k, np.zeros(n))) for k in HEXA_COLUMNS]
    ids = np.asarray(p["id"], np.int64)
    ist = np.asarray(p.get("istar", np.zeros(n)), np.int64)
    with open(path, "w") as f:
        for i in range(n):
            f.write(f"{ids[i]:6d} {ist[i]:2d} " + " ".join(c[i] for c in cols) + "\n")


def read_hexa(path) -> dict:
    rows = [line.split() for line in open(path) if line.strip()]
    n = len(rows)
    p = {"id": np.array([int(r[0]) for r in rows], np.int64),
         "istar": np.array([int(r[1]) for r in rows], np.int64)}
    for c, k in enumerate(HEXA_COLUMNS):
        u = np.array([int(r[2 + c], 16) for r in rows], np.uint64) if n else np.zeros(0, np.uint64)
        p[k] = u.view(np.float64)
    return p


# ------------------------------------------------------------------ time.log
def time_log_line(step: int, t: float, dt: float, energy: float, hmax: float, wall: float,
                  e_nuc: float = 0.0, e_abs: float = 0.0, units: CodeUnits = REFERENCE_UNITS,
                  energy_scale: float | None = None) -> str:
    """One `time:` line; energies in code units are multiplied by
    energy_scale (default: specific energy unit x mass unit)."""
    es = units.energy * units.mass if energy_scale is None else energy_scale
    return (f"time: {step:8d} {t * units.time:16.10f} {dt * units.time:+e} "
            f"{(energy - e_nuc - e_abs) * es:+e} {energy * es:+e} {e_nuc * es:+e} {e_abs * es:+e} "
            f"{hmax * units.length:+e} {wall:+e}\n")


def read_time_log(path) -> np.ndarray:
    """Rows of the `time:` lines: step, t, dt, E-E_nuc-E_abs, E, E_nuc, E_abs, H_max, wall."""
    rows = [line.split()[1:] for line in open(path) if line.startswith("time:")]
    return np.array(rows, np.float64).reshape(-1, 9)
