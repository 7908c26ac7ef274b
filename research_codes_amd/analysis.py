"""Post-processing of runs: energy-error report and two-body orbital energy.

Host-side numpy restatements of the reference's analysis scripts (its Python
is post-processing only): energy_check/energy_check.py:27-110.  The unbound
mass, which needs every particle, runs on the GPU
(SPHSystem.unbound_mass, sphk_unbound_mass_dev).
"""
from __future__ import annotations

import numpy as np


def energy_error_report(time_data: np.ndarray, max_time: float | None = None) -> dict:
    """check_energy_error (energy_check/energy_check.py:92-110).

    time_data rows are (t, E) (columns 2 and 4 of time.log).  Errors are
    E - E(first row).  Scanning rows in order up to the first with
    t >= max_time (default: the last row): `max_error` is the error of the
    largest magnitude seen (the first of equal magnitudes; the end row wins
    ties), `end_error` the error at that end row."""
    td = np.asarray(time_data, np.float64)
    if td.ndim != 2 or td.shape[0] == 0 or td.shape[1] < 2:
        raise ValueError("time_data must have rows of (t, E)")
    if max_time is None:
        max_time = td[-1, 0]
    reached = np.nonzero(td[:, 0] >= max_time)[0]
    if reached.size == 0:
        raise ValueError("max_time is later than the last row")
    end = int(reached[0])
    err = td[:, 1] - td[0, 1]
    best, t_best = 0.0, td[0, 0]
    for i in range(end):
        if abs(err[i]) > abs(best):
            best, t_best = err[i], td[i, 0]
    if abs(err[end]) >= abs(best):
        best, t_best = err[end], td[end, 0]
    return {"max_error": best, "max_time": t_best, "end_error": err[end], "end_time": td[end, 0],
            "e0": td[0, 1]}


def _density_centre(p: dict, sel) -> tuple[np.ndarray, np.ndarray]:
    w = p["dens"][sel]
    pos = np.stack([p[c][sel] for c in ("x", "y", "z")], 1)
    vel = np.stack([p[c][sel] for c in ("vx", "vy", "vz")], 1)
    return (w[:, None] * pos).sum(0) / w.sum(), (w[:, None] * vel).sum(0) / w.sum()


def orbital_energy(p: dict, r1: float, r2: float, G: float = 1.0) -> float:
    """calc_energy (energy_check/energy_check.py:27-84): the particles sorted
    by id are split in halves (star 1, star 2); each star is a point mass at
    its density-weighted centre, moving with the density-weighted velocity,
    holding the mass within 2 R of that centre; returns the two-body energy
    K1 + K2 - G m1 m2 / d."""
    order = np.argsort(p["id"], kind="stable")
    q = {k: np.asarray(v)[order] for k, v in p.items()}
    n = len(order)
    halves = (np.arange(n) < n // 2, np.arange(n) >= n // 2)
    cen, mass, ke = [], [], 0.0
    for sel, r in zip(halves, (r1, r2)):
        c, v = _density_centre(q, sel)
        d2 = sum((q[k][sel] - c[a]) ** 2 for a, k in enumerate(("x", "y", "z")))
        m = q["m"][sel][np.sqrt(d2) < 2 * r].sum()
        cen.append(c)
        mass.append(m)
        ke += 0.5 * m * float(v @ v)
    d = float(np.linalg.norm(cen[0] - cen[1]))
    return ke - G * mass[0] * mass[1] / d
