"""Seeded polytrope and two-star collision initial conditions.

Units G = 1; each star is given by its mass M, radius R and polytropic index
n.  The reference builds its stars with a stretched cubic lattice and a
separate Lane-Emden RK4 solve (polytropic_initial_condition/lane_emden.hpp:14-53,
main.cpp:12-227) and a hyperbolic two-body tool
(two_body_initial_condition/main.cpp:87-199).  This module instead samples
particles at random from the Lane-Emden mass profile, so that any N and seed
are possible, and sets up parabolic encounters directly:

  * Lane-Emden theta'' = -theta^n - 2 theta'/xi, RK4 with d(xi) = 1e-4 from a
    series start, surface xi_1 by linear interpolation of the sign change.
  * m_k ~ U(0,1) -> r_k = R xi(mu^-1(m_k)) / xi_1, mu(xi) = -xi^2 theta'(xi)
    normalised; cos(theta) ~ U(-1,1); phi ~ U(0, 2 pi).  Equal masses M/N.
  * rho_c = M xi_1 / (4 pi R^3 |theta'(xi_1)|),
    K = 4 pi G (R/xi_1)^2 rho_c^(1-1/n) / (n+1),  P = K rho^(1+1/n),
    u = P / ((gamma-1) rho), gamma = 5/3 (the reference's ideal gas,
    P = rho u / 1.5, sph_yamauchi/run/hdr_hgas.hpp:339-341, 767-768).
  * H_0 = min(eta kappa (m / rho_LE(r))^(1/3), R/4); alpha = 1, alpha_u = 0
    (polytropic_initial_condition/make_and_output_data.hpp:12-13).
  * Random numbers: SplitMix64 -> 53-bit doubles, seed = 1000*config + star,
    three draws per particle in the order (m, cos theta, phi).
"""
from __future__ import annotations

import math

import numpy as np

GAMMA = 5.0 / 3.0
KSRH = 1.936492
_M64 = (1 << 64) - 1


def splitmix64(seed: int, count: int) -> np.ndarray:
    """First `count` outputs of SplitMix64 started at `seed` (uint64)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        s = np.uint64(seed & _M64) + k * np.uint64(0x9E3779B97F4A7C15)
        z = s
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform53(seed: int, count: int) -> np.ndarray:
    """Doubles in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


class LaneEmden:
    """Tabulated Lane-Emden solution theta(xi), theta'(xi) for index n."""

    def __init__(self, n: float, dxi: float = 1e-4):
        self.n = float(n)
        xi0 = 1e-4
        th = 1.0 - xi0 ** 2 / 6.0 + n * xi0 ** 4 / 120.0
        dth = -xi0 / 3.0 + n * xi0 ** 3 / 30.0

        def f(xi, t, dt):
            return dt, -(max(t, 0.0) ** n) - 2.0 * dt / xi

        xs, ts, ds = [0.0, xi0], [1.0, th], [0.0, dth]
        xi = xi0
        h = dxi
        while True:
            k1 = f(xi, th, dth)
            k2 = f(xi + h / 2, th + h / 2 * k1[0], dth + h / 2 * k1[1])
            k3 = f(xi + h / 2, th + h / 2 * k2[0], dth + h / 2 * k2[1])
            k4 = f(xi + h, th + h * k3[0], dth + h * k3[1])
            th_n = th + h / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
            dth_n = dth + h / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
            if th_n <= 0.0:
                frac = th / (th - th_n)
                xi1 = xi + frac * h
                d1 = dth + frac * (dth_n - dth)
                xs.append(xi1)
                ts.append(0.0)
                ds.append(d1)
                break
            xi += h
            th, dth = th_n, dth_n
            xs.append(xi)
            ts.append(th)
            ds.append(dth)
            if xi > 100.0:
                raise ValueError(f"Lane-Emden n={n} has no finite surface below xi=100")
        self.xi = np.array(xs)
        self.theta = np.array(ts)
        self.dtheta = np.array(ds)
        self.xi1 = xs[-1]
        self.dtheta1 = ds[-1]
        self.mu = -self.xi ** 2 * self.dtheta          # enclosed-mass coordinate
        self.mu[0] = 0.0
        self.mu1 = self.mu[-1]                         # = -xi1^2 theta'(xi1)


def structure(le: LaneEmden, M: float, R: float, G: float = 1.0) -> tuple[float, float]:
    """Central density and polytropic constant of the n-polytrope of mass M,
    radius R: rho_c = M xi_1 / (4 pi R^3 |theta'(xi_1)|),
    K = 4 pi G (R / xi_1)^2 rho_c^(1 - 1/n) / (n + 1)."""
    n = le.n
    rho_c = M * le.xi1 / (4.0 * math.pi * R ** 3 * (-le.dtheta1))
    K = 4.0 * math.pi * G * (R / le.xi1) ** 2 * rho_c ** (1.0 - 1.0 / n) / (n + 1.0)
    return rho_c, K


def polytrope(n_index: float, N: int, M: float = 1.0, R: float = 1.0, seed: int = 0,
              eta: float = 1.6, id0: int = 0, istar: int = 0, le: LaneEmden | None = None) -> dict:
    """N equal-mass particles sampled from an n-polytrope of mass M, radius R."""
    le = le or LaneEmden(n_index)
    u = uniform53(seed, 3 * N).reshape(N, 3)
    mu = np.maximum.accumulate(le.mu)                  # guard monotonicity at round-off
    xi = np.interp(u[:, 0] * le.mu1, mu, le.xi)
    r = R * xi / le.xi1
    cth = 2.0 * u[:, 1] - 1.0
    sth = np.sqrt(np.maximum(0.0, 1.0 - cth * cth))
    phi = 2.0 * math.pi * u[:, 2]
    x = r * sth * np.cos(phi)
    y = r * sth * np.sin(phi)
    z = r * cth
    rho_c, K = structure(le, M, R)
    th = np.interp(xi, le.xi, le.theta)
    rho = rho_c * np.maximum(th, 0.0) ** n_index
    rho = np.maximum(rho, rho_c * 1e-4)               # surface particles: avoid rho = 0
    P = K * rho ** (1.0 + 1.0 / n_index)
    uene = P / ((GAMMA - 1.0) * rho)
    m = np.full(N, M / N)
    # initial H; capped at R/4 where the Lane-Emden density vanishes at the surface
    ksr = np.minimum(eta * KSRH * (m / rho) ** (1.0 / 3.0), 0.25 * R)
    return dict(
        id=np.arange(id0, id0 + N, dtype=np.int32), istar=np.full(N, istar, np.int32),
        m=m, x=x, y=y, z=z, vx=np.zeros(N), vy=np.zeros(N), vz=np.zeros(N),
        uene=uene, ksr=ksr, alph=np.ones(N), alphu=np.zeros(N), rho_le=rho,
    )


def concat(*stars: dict) -> dict:
    return {k: np.concatenate([s[k] for s in stars]) for k in stars[0]}


def collision(star0: dict, star1: dict, separation: float = 3.0, impact: float = 0.0,
              G: float = 1.0) -> dict:
    """Place two stars on a parabolic encounter in their centre-of-mass frame.

    The relative position (star1 - star0) is (sqrt(d^2 - b^2), b, 0) with
    d = separation and b = impact (the lateral offset at that separation);
    the relative velocity is (-v, 0, 0) with v = sqrt(2 G (M0+M1) / d), the
    parabolic (escape) speed at distance d.
    """
    M0 = float(star0["m"].sum())
    M1 = float(star1["m"].sum())
    M = M0 + M1
    d = np.array([math.sqrt(max(separation ** 2 - impact ** 2, 0.0)), impact, 0.0])
    v = np.array([-math.sqrt(2.0 * G * M / separation), 0.0, 0.0])
    s0 = dict(star0)
    s1 = dict(star1)
    # each star's own (sampled) centre of mass and bulk velocity are removed
    # first, so the separation and speed are those of the two centres
    for s, star in ((s0, star0), (s1, star1)):
        for c in ("x", "y", "z", "vx", "vy", "vz"):
            s[c] = star[c] - np.average(star[c], weights=star["m"])
    for k, c in enumerate("xyz"):
        s0[c] = s0[c] - M1 / M * d[k]
        s1[c] = s1[c] + M0 / M * d[k]
        s0["v" + c] = s0["v" + c] - M1 / M * v[k]
        s1["v" + c] = s1["v" + c] + M0 / M * v[k]
    s1["id"] = star1["id"] + len(star0["id"])
    s1["istar"] = np.ones_like(star1["istar"])
    return concat(s0, s1)


# --- the five workloads the benchmark and tests use ---------------------------

# eta for ~50 neighbours: the realised mean of np (particles within H_2, self
# included) on these sampled stars after the smoothing length has relaxed is
# 49.1 (n=1.5) and 50.8 (n=3) at eta = 1.25, for N = 1e4 and 1e5 alike.  The
# uniform-medium count 4 pi/3 (eta kappa)^3 = 50 (eta = 1.18) ignores that the
# density sum includes the particle itself (C0 m / H^3, a quarter of the sum
# at 50 neighbours), which shrinks H: eta = 1.18 realises ~40.
ETA_50 = 1.25


def single_polytrope(N: int = 10_000, seed: int = 1000, eta: float = ETA_50) -> dict:
    """n=1.5, M=R=1, v=0, ~50 neighbours."""
    return polytrope(1.5, N, 1.0, 1.0, seed=seed, eta=eta)


def head_on(N_each: int = 100_000, eta: float = ETA_50, separation: float = 3.0) -> dict:
    le = LaneEmden(1.5)
    a = polytrope(1.5, N_each, 1.0, 1.0, seed=2000, eta=eta, le=le)
    b = polytrope(1.5, N_each, 1.0, 1.0, seed=2001, eta=eta, le=le)
    return collision(a, b, separation, 0.0)


def off_axis(N_total: int = 1_000_000, eta: float = ETA_50) -> dict:
    """n=3 (M=1, R=1) + n=1.5 (M=0.5, R=0.6), equal particle mass, b = 0.25."""
    N0 = int(round(N_total * 1.0 / 1.5))
    N1 = N_total - N0
    a = polytrope(3.0, N0, 1.0, 1.0, seed=3000, eta=eta)
    b = polytrope(1.5, N1, 0.5, 0.6, seed=3001, eta=eta)
    return collision(a, b, 3.0, 0.25)


def ensemble(n_runs: int = 64, n_min: float = 1e4, n_max: float = 2e5, seed: int = 4000,
             eta: float = ETA_50) -> tuple[dict, list[int], list[dict]]:
    """Independent collisions for batched evaluation.

    Run r: N_r = exp(U(ln n_min, ln n_max)) particles, impactor mass ratio
    q ~ U(0.3, 1) with radius q^(-1/3) (the n=1.5 mass-radius relation
    R ~ M^(-1/3)), impact parameter b ~ U(0, 1) at separation 3, both stars
    n=1.5 and equal particle mass.  Returns the particles of all runs
    concatenated (ids unique across runs), the particle count of each run,
    and the per-run parameters."""
    u = uniform53(seed, 3 * n_runs).reshape(n_runs, 3)
    le = LaneEmden(1.5)
    runs, counts, meta, id0 = [], [], [], 0
    for r in range(n_runs):
        n = int(round(math.exp(math.log(n_min) + u[r, 0] * (math.log(n_max) - math.log(n_min)))))
        q = 0.3 + 0.7 * u[r, 1]
        b = u[r, 2]
        n0 = max(2, int(round(n / (1.0 + q))))
        n1 = max(2, n - n0)
        a = polytrope(1.5, n0, 1.0, 1.0, seed=seed + 2 * r + 1, eta=eta, le=le)
        c = polytrope(1.5, n1, q, q ** (-1.0 / 3.0), seed=seed + 2 * r + 2, eta=eta, le=le)
        p = collision(a, c, 3.0, b)
        p["id"] = p["id"] + id0
        id0 += len(p["id"])
        runs.append(p)
        counts.append(len(p["id"]))
        meta.append({"n": len(p["id"]), "q": q, "b": b})
    return concat(*runs), counts, meta
