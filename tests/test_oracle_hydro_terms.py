"""Pins the magnitudes of the oracle's dissipative hydro terms and of the
per-particle rates and timestep branches that the conservation tests leave
open (hdr_hydro.hpp:366-391, hdr_hgas.hpp:477-488,750-762):

  * signal velocity: two particles approaching at closing speed w give
    vsmx = c_i + c_j + 3 w;
  * artificial viscosity, a converging flow v = -eps x on a cubic lattice
    (spacing H/8): the heating of the central particle, and the force from a
    linear gradient of alpha, against the continuum limit of the pair sum
    (n times a radial quadrature of the kernel derivative, scipy.integrate);
  * thermal conduction with alpha_u != 0 across a linear pressure field and
    a quadratic u: the heating against its continuum limit
    n (8 pi / 7) ... int r^(9/2) dW dr (the |x|^(1/2) of v_u integrated
    over angles);
  * alpha_u-dot: relaxation at rate kappa c / (4 H) towards alpha_u,min, and
    the source (H / kappa) |diffu| / sqrt(u + eps_u) (alpha_u,max - alpha_u),
    zero at alpha_u,max;
  * timestep: the compression branch -C / div v and the gravity branch
    C sqrt(H / |g|) bind when they are the smallest.

The kernel derivative in the quadratures is differentiated here from
W(q) = 21/(2 pi) (1-q)^4 (1+4q) (numpy polynomials), not taken from the
oracle.  Each check is also run against builds of the oracle with one
planted coefficient mistake (oracle.build_planted): every plant must make
the check that covers it fail.
"""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import oracle

KSRH = 1.936492
H = 0.4
A = H / 8.0

_W = np.polynomial.Polynomial([1.0, -1.0]) ** 4 * np.polynomial.Polynomial([1.0, 4.0]) * (21.0 / (2 * np.pi))
_DW = _W.deriv()


def dW(r):
    """dW/dr-like kernel factor of the functor: H^-4 W'(r / H), zero beyond H."""
    q = r / H
    return np.where(q < 1.0, _DW(q), 0.0) / H ** 4


def _hsrc(n, **kw):
    d = dict(id=np.arange(n, dtype=np.int32), m=np.ones(n), x=np.zeros(n), y=np.zeros(n),
             z=np.zeros(n), vx=np.zeros(n), vy=np.zeros(n), vz=np.zeros(n), hinv=np.full(n, 1 / H),
             pres=np.ones(n), presc=np.zeros(n), bswt=np.zeros(n), dens=np.ones(n),
             vsnd=np.ones(n), alph=np.zeros(n), alphu=np.zeros(n), thrm=np.ones(n),
             eta=np.zeros(n))
    for k, v in kw.items():
        d[k] = np.asarray(v, dtype=np.int32 if k == "id" else np.float64) * np.ones(n)
    return d


def _lattice_center(**fields):
    """Central particle of a cubic lattice (spacing A, unit masses) with the
    given fields (callables of x, y, z or constants); returns its hydro output
    and the number density."""
    g = (np.arange(-12, 13)) * A
    x, y, z = (c.ravel() for c in np.meshgrid(g, g, g, indexing="ij"))
    n = len(x)
    rho = 1.0 / A ** 3
    base = dict(x=x, y=y, z=z, dens=rho)
    base.update({k: (v(x, y, z) if callable(v) else v) for k, v in fields.items()})
    p = _hsrc(n, **base)
    c = int(np.argmin(x * x + y * y + z * z))
    pi = {k: v[[c]] for k, v in p.items()}
    return oracle.hydro([0, 1], [0, n], None, pi, p), rho


def _radial(f):
    """int_0^H f(r) dr."""
    return quad(f, 0.0, H, limit=200, epsabs=0.0, epsrel=1e-12)[0]


# ------------------------------------------------------------------ checks

def check_signal_velocity():
    ci, cj, w = 0.7, 1.3, 0.45
    p = _hsrc(2, x=[0.0, 0.1], vx=[w, 0.0], vsnd=[ci, cj])
    h = oracle.hydro([0, 2], [0, 2], None, p, p)
    assert abs(h["vsmx"][0] - (ci + cj + 3 * w)) < 1e-14, h["vsmx"][0]
    # receding: the sound speeds only (the self entry gives 2 c_i)
    p = _hsrc(2, x=[0.0, 0.1], vx=[-w, 0.0], vsnd=[ci, cj])
    h = oracle.hydro([0, 2], [0, 2], None, p, p)
    assert abs(h["vsmx"][0] - (ci + cj)) < 1e-14


def check_viscous_heating():
    eps, c, b, al = 0.8, 0.3, 0.35, 1.2     # bswt input is 0.5 B: b = 0.5 B
    h, rho = _lattice_center(vx=lambda x, y, z: -eps * x, vy=lambda x, y, z: -eps * y,
                             vz=lambda x, y, z: -eps * z, vsnd=c, bswt=b, alph=al)
    # pair (i, j): w = -eps r, v_sig = 2c + 3 eps r, Pi = (2b)(2al) v_sig w,
    # udot_i = 1/2 sum ka (-rhobar^-1 Pi w / 4), ka = 2 m dW, rhobar^-1 = 1 / (2 rho)
    integ = _radial(lambda r: 4 * np.pi * r * r * dW(r) * (2 * c + 3 * eps * r) * eps ** 2 * r * r)
    expect = -(b * al / (2 * rho)) * rho * integ * 1.0      # n = rho for unit masses
    assert abs(h["udot"][0] / expect - 1) < 2e-3, (h["udot"][0], expect)
    assert h["udot"][0] > 0


def check_viscous_force():
    eps, c, b, a0, a1 = 0.8, 0.3, 0.35, 1.0, 0.6
    h, rho = _lattice_center(vx=lambda x, y, z: -eps * x, vy=lambda x, y, z: -eps * y,
                             vz=lambda x, y, z: -eps * z, vsnd=c, bswt=b,
                             alph=lambda x, y, z: a0 + a1 * x)
    # the odd part of alpha_j = a0 + a1 x_j survives:
    # a_x = (b a1 m eps / (2 rho)) sum x_j^2 (2c + 3 eps r) dW, sum x^2 f(r) -> n int r^2/3 f 4 pi r^2 dr
    integ = _radial(lambda r: 4 * np.pi * r ** 4 / 3 * (2 * c + 3 * eps * r) * dW(r))
    expect = (b * a1 * eps / (2 * rho)) * rho * integ
    assert abs(h["ax"][0] / expect - 1) < 2e-3, (h["ax"][0], expect)
    assert abs(h["ay"][0]) < 1e-9 * abs(expect) and abs(h["az"][0]) < 1e-9 * abs(expect)


def check_conduction():
    p1, u2, au = 0.9, 0.7, 0.4
    h, rho = _lattice_center(pres=lambda x, y, z: 5.0 + p1 * x, thrm=lambda x, y, z: 2.0 + u2 * x * x,
                             alphu=au)
    # v_u = sqrt(2 |dP| rhobar^-1) = sqrt(p1 |x_j| / rho), du = -u2 x_j^2:
    # udot_i = -(2 au u2 / (2 rho)) sum dW sqrt(p1 |x_j| / rho) x_j^2
    #        -> -(au u2 sqrt(p1 / rho) / rho) n int dW r^(5/2) r^2 dr * int |mu|^(5/2) dOmega (= 8 pi / 7)
    integ = _radial(lambda r: dW(r) * r ** 4.5) * 8 * np.pi / 7
    expect = -(au * u2 * math.sqrt(p1 / rho) / rho) * rho * integ
    assert abs(h["udot"][0] / expect - 1) < 1e-2, (h["udot"][0], expect)
    assert h["udot"][0] > 0      # conduction heats where u has a minimum


def _tail(n=1, **kw):
    s = oracle.make_state(n, **{k: np.broadcast_to(np.asarray(v, float), (n,)) for k, v in kw.items()})
    h = dict(ax=[0.0] * n, ay=[0.0] * n, az=[0.0] * n, gx=[0.0] * n, gy=[0.0] * n, gz=[0.0] * n,
             udot=[0.0] * n, vsmx=[1.0] * n, diffu=list(s["diffu"]))
    return s, h


def check_alpha_u_dot():
    c, ksr = 1.5, 0.2
    # relaxation only (no diffu): rate kappa c / (4 H)
    p = oracle.params(alphu_max=2.0, alphu_min=0.1, eps_u=0.01)
    s, h = _tail(alphu=0.6, vsnd=c, ksr=ksr, uene=1.0, diffu=0.0, divv=-3.0, alph=1.0)
    oracle.hydro_writeback_tail(s, p, h)
    assert abs(s["adotu"][0] - (-(0.6 - 0.1) * KSRH * c / (4 * ksr))) < 1e-13
    # source only (alpha_u = alpha_u,min): (H / kappa) |diffu| / sqrt(u + eps_u) (max - alpha_u)
    p = oracle.params(alphu_max=2.0, alphu_min=0.5, eps_u=0.01)
    for du, u in ((-3.0, 1.0), (6.0, 1.0), (-3.0, 3.99)):
        s, h = _tail(alphu=0.5, vsnd=c, ksr=ksr, uene=u, diffu=du, divv=-3.0, alph=1.0)
        oracle.hydro_writeback_tail(s, p, h)
        want = ksr / KSRH * abs(du) / math.sqrt(u + 0.01) * (2.0 - 0.5)
        assert abs(s["adotu"][0] / want - 1) < 1e-13, (du, u)
    # no source at (or above) alpha_u,max
    p = oracle.params(alphu_max=0.5, alphu_min=0.5, eps_u=0.01)
    s, h = _tail(alphu=0.5, vsnd=c, ksr=ksr, uene=1.0, diffu=-3.0, divv=-3.0, alph=1.0)
    oracle.hydro_writeback_tail(s, p, h)
    assert s["adotu"][0] == 0.0


def _dt(**kw):
    p = oracle.params(c_dt=0.1, dt_max=1.0)
    s, _ = _tail(**kw)
    dt = oracle.timestep(s, p, dt_prev=1.0, time=0.0)
    assert dt > 0 and math.log2(dt) == round(math.log2(dt))
    return dt


def check_dt_branches():
    # Courant branch alone: C H / (vsmx kappa)
    base = dict(ksr=1.0, vsmx=1e-3, divv=0.0, g2x=1e-12)
    # compression: -C / div v = 0.1 / 300
    dtc = 0.1 / 300.0
    dt = _dt(**{**base, "divv": -300.0})
    assert dt <= dtc < 2 * dt, (dt, dtc)
    assert _dt(**{**base, "divv": 300.0}) > 0.25       # expansion does not limit
    # gravity: C sqrt(H / |g|), g = accg1 + accg2
    dtg = 0.1 * math.sqrt(1.0 / 5.0e4)
    dt = _dt(**{**base, "g2x": 3.0e4, "g1y": 4.0e4})
    assert dt <= dtg < 2 * dt, (dt, dtg)


CHECKS = {
    "signal_velocity": check_signal_velocity,
    "viscous_heating": check_viscous_heating,
    "viscous_force": check_viscous_force,
    "conduction": check_conduction,
    "alpha_u_dot": check_alpha_u_dot,
    "dt_branches": check_dt_branches,
}

# planted mistake (oracle/sphk_oracle.c ORC_PLANT) -> the check that must catch it
PLANTS = {
    1: "signal_velocity",   # v_sig = c_i + c_j - 4 w0
    2: "viscous_force",     # Pi rhobar^-1 weighted 1 instead of 1/2 in the force
    3: "conduction",        # v_u^2 = |dP| rhobar^-1 (factor 2 dropped)
    4: "viscous_heating",   # Pi w0 weighted 1/2 instead of 1/4 in the heating
    5: "alpha_u_dot",       # alpha_u source taken from alpha's compression source
    6: "dt_branches",       # -C * div v instead of -C / div v
    7: "dt_branches",       # C H / sqrt|g| instead of C sqrt(H / |g|)
    8: "conduction",        # alpha_u,i only instead of alpha_u,i + alpha_u,j
}


@pytest.mark.parametrize("name", list(CHECKS))
def test_hydro_term(name):
    CHECKS[name]()


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    d = tmp_path_factory.mktemp("plant")
    return {k: oracle.build_planted(k, d) for k in PLANTS}


@pytest.mark.parametrize("k", list(PLANTS))
def test_planted_mistake_is_caught(planted, k):
    with oracle.use(planted[k]):
        with pytest.raises(AssertionError):
            CHECKS[PLANTS[k]]()
