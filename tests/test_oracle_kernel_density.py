"""Pins the oracle's kernel and density functor (hdr_kernel.hpp:216-257,
hdr_density.hpp:246-365) against mathematics it must satisfy -- none of these
retypes the oracle's formula:

  * W: integral of W(|x|/H) H^-3 over space is 1 (fixes C0 = 21/(2 pi));
    W(q) = 0 for q >= 1; W'(q) equals the numerical derivative of W.
  * isolated particle: the self term alone gives rho = m C0 / H^3, np = 1, and
    the 3-pass fixed point has a closed form -- this fixes which H the
    returned dens/np/ksr belong to (H_2, H_2, H_3) and the H_max cap.
  * uniform lattice: rho -> m/a^3, Omega^-1 -> 1 (3W + qW' integrates to 0).
  * Omega: the grad-h sum equals d rho / dH measured by finite differences of
    brute-force sums (closes the 1/3 and H factors and the sign).
  * linear velocity fields v = A x: divv -> tr A, rotv -> |curl A|.
  * permutation and translation invariance of every output.
"""
import math

import numpy as np
import pytest

import oracle

KSRH = 1.936492
C0 = 21.0 / (2.0 * math.pi)


def test_kernel_normalisation_and_support():
    # 4 pi int_0^1 W(q) q^2 dq = 1  (trapezoid rule on a fine grid)
    q = np.linspace(0.0, 1.0, 20001)
    w = np.array([oracle.kernel0th(v) for v in q])
    integral = 4 * math.pi * np.trapezoid(w * q * q, q)
    assert abs(integral - 1.0) < 1e-8
    assert abs(oracle.kernel0th(0.0) - C0) < 1e-14
    for v in (1.0, 1.0 + 1e-12, 1.5, 7.0):
        assert oracle.kernel0th(v) == 0.0 and oracle.kernel1st(v) == 0.0


def test_kernel_derivative_matches_finite_difference():
    h = 1e-6
    for q in np.linspace(0.01, 0.99, 37):
        fd = (oracle.kernel0th(q + h) - oracle.kernel0th(q - h)) / (2 * h)
        assert abs(oracle.kernel1st(q) - fd) < 1e-7 * max(1.0, abs(fd))


def _one(x, y, z, m=None, v=None, ksr=None, ids=None):
    n = len(x)
    m = np.full(n, 1.0 / n) if m is None else m
    v = np.zeros((3, n)) if v is None else v
    p = dict(id=np.arange(n) if ids is None else ids, m=m, x=x, y=y, z=z, vx=v[0], vy=v[1],
             vz=v[2], ksr=ksr, rs=1.1 * ksr)
    return p


def _all_pairs(n):
    return [0, n], [0, n], None


@pytest.mark.parametrize("H0", [0.3, 1.0, 2.5])
def test_isolated_particle_closed_form(H0):
    m = 0.7
    p = _one(np.zeros(1), np.zeros(1), np.zeros(1), m=np.array([m]), ksr=np.array([H0]))
    eta = 1.6
    d = oracle.density(*_all_pairs(1), p, p, eta=eta, ksrh=KSRH, hmax=1e30)
    # rho_k = m C0 / H_k^3  and  H_{k+1} = eta kappa (m / rho_k)^(1/3) = c H_k
    c = eta * KSRH / C0 ** (1.0 / 3.0)
    H2, H3 = c * c * H0, c ** 3 * H0
    assert d["np"][0] == 1                       # self counted (hdr_density.hpp:285)
    assert abs(d["dens"][0] / (m * C0 / H2 ** 3) - 1) < 1e-13   # dens from H_2
    assert abs(d["ksr"][0] / H3 - 1) < 1e-13                     # ksr is H_3
    assert d["itr"][0] == (H3 > 1.1 * H0)
    assert d["divv"][0] == 0.0 and d["rotv"][0] == 0.0
    # H_max caps every iterate (hdr_density.hpp:291-292)
    cap = 0.5 * H0
    d2 = oracle.density(*_all_pairs(1), p, p, eta=eta, ksrh=KSRH, hmax=cap)
    assert d2["ksr"][0] == cap
    assert abs(d2["dens"][0] / (m * C0 / cap ** 3) - 1) < 1e-13


def _lattice(nside, a=1.0):
    g = (np.arange(nside) - (nside - 1) / 2.0) * a
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


def test_uniform_lattice_density_and_omega():
    a = 0.1
    x, y, z = _lattice(17, a)
    n = len(x)
    m = np.full(n, 2.0)
    centre = np.argmin(x * x + y * y + z * z)
    H = 3.0 * a   # below eta*kappa*a = 3.1a, so hmax = H holds H fixed
    p = _one(x, y, z, m=m, ksr=np.full(n, H))
    pi = {k: (v[[centre]] if hasattr(v, "__len__") else v) for k, v in p.items()}
    # hmax = H freezes H so the sums are taken at a known smoothing length
    d = oracle.density([0, 1], [0, n], None, pi, p, eta=1.6, ksrh=KSRH, hmax=H)
    rho0 = 2.0 / a ** 3
    assert abs(d["dens"][0] / rho0 - 1) < 5e-3   # lattice discretisation at H = 3a
    assert abs(d["grdh"][0] - 1.0) < 5e-3
    assert d["ksr"][0] == H


def _brute_rho(x, y, z, m, i, H):
    r = np.sqrt((x - x[i]) ** 2 + (y - y[i]) ** 2 + (z - z[i]) ** 2)
    q = r / H
    w = np.where(q < 1, C0 * np.clip(1 - q, 0, None) ** 4 * (1 + 4 * q), 0.0)
    return float((m * w).sum() / H ** 3)


def test_omega_equals_finite_difference_of_rho():
    rng = np.random.default_rng(3)
    n = 400
    x, y, z = rng.uniform(-1, 1, (3, n))
    m = rng.uniform(0.5, 1.5, n)
    p = _one(x, y, z, m=m, ksr=np.full(n, 0.6))
    d = oracle.density(*_all_pairs(n), p, p, eta=1.6, ksrh=KSRH, hmax=1e30)
    for i in (0, 17, 123):
        H3 = d["ksr"][i]
        e = 1e-6 * H3
        g = (_brute_rho(x, y, z, m, i, H3 + e) - _brute_rho(x, y, z, m, i, H3 - e)) / (2 * e)
        omega_inv = 1.0 / (1.0 + H3 * g / (3.0 * d["dens"][i]))
        assert abs(d["grdh"][i] - omega_inv) < 1e-6 * abs(omega_inv)


@pytest.mark.parametrize("A", [
    np.diag([1.0, 1.0, 1.0]),                       # expansion: div 3, curl 0
    np.array([[0, -1, 0], [1, 0, 0], [0, 0, 0.0]]),  # rotation about z: div 0, |curl| 2
    np.array([[0.2, 0.5, 0.0], [-0.3, -0.1, 0.4], [0.0, 0.7, 0.3]]),
])
def test_linear_velocity_field(A):
    a = 0.1
    x, y, z = _lattice(15, a)
    n = len(x)
    X = np.vstack([x, y, z])
    v = A @ X
    centre = np.argmin(x * x + y * y + z * z)
    H = 3.0 * a
    p = _one(x, y, z, m=np.ones(n), v=v, ksr=np.full(n, H))
    pi = {k: (val[[centre]] if hasattr(val, "__len__") else val) for k, val in p.items()}
    d = oracle.density([0, 1], [0, n], None, pi, p, eta=1.6, ksrh=KSRH, hmax=H)
    curl = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    assert abs(d["divv"][0] - np.trace(A)) < 2e-3 * max(1.0, abs(np.trace(A)))
    assert abs(d["rotv"][0] - np.linalg.norm(curl)) < 2e-3 * max(1.0, np.linalg.norm(curl))


def test_permutation_and_translation_invariance():
    rng = np.random.default_rng(11)
    n = 300
    x, y, z = rng.uniform(-1, 1, (3, n))
    v = rng.normal(size=(3, n))
    p = _one(x, y, z, m=rng.uniform(0.5, 1.5, n), v=v, ksr=np.full(n, 0.5))
    d = oracle.density(*_all_pairs(n), p, p, eta=1.6, ksrh=KSRH, hmax=1e30)
    perm = rng.permutation(n)
    pj = {k: val[perm] for k, val in p.items()}
    d2 = oracle.density([0, n], [0, n], None, p, pj, eta=1.6, ksrh=KSRH, hmax=1e30)
    assert np.array_equal(d["np"], d2["np"])
    for k in ("dens", "ksr", "divv", "rotv", "grdh"):
        assert np.allclose(d[k], d2[k], rtol=1e-12, atol=1e-12 * np.abs(d[k]).max())
    shift = {**p, "x": p["x"] + 3.0, "y": p["y"] - 2.0}
    d3 = oracle.density(*_all_pairs(n), shift, shift, eta=1.6, ksrh=KSRH, hmax=1e30)
    assert np.array_equal(d["np"], d3["np"])
    for k in ("dens", "ksr", "grdh"):
        assert np.allclose(d[k], d3[k], rtol=1e-10)


def test_np_counts_match_brute_force():
    rng = np.random.default_rng(5)
    n = 500
    x, y, z = rng.uniform(-1, 1, (3, n))
    p = _one(x, y, z, ksr=np.full(n, 0.35))
    # hmax = 0.35 freezes H = 0.35 for all passes
    d = oracle.density(*_all_pairs(n), p, p, eta=1.6, ksrh=KSRH, hmax=0.35)
    r = np.sqrt((x[:, None] - x) ** 2 + (y[:, None] - y) ** 2 + (z[:, None] - z) ** 2)
    assert np.array_equal(d["np"], (r / 0.35 < 1).sum(axis=1))
