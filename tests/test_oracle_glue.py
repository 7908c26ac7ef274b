"""Pins the oracle's per-particle glue (hdr_hgas.hpp, hdr_sph.hpp, main.cpp):

  * EOS: P = (gamma-1) rho u and c^2 = gamma P / rho with gamma = 5/3, which is
    the reference's P = rho u / n, n = 1.5 (hdr_hgas.hpp:339-341,767-768);
  * Balsara switch limits (0 without compression, -> 1 for pure compression);
  * gravity write-back: G scaling; eta = 0 under a constant softening (eta
    with eps = H/kappa is pinned by energy conservation,
    tests/test_oracle_conservation.py);
  * alpha-dot: relaxes to alpha_min at rate c kappa / (4 H), is pushed up by
    compression only;
  * timestep: power-of-two halving to below the particle minimum; doubling only
    when time is a multiple of 2 dt (main.cpp:252-270);
  * KDK: exact for constant acceleration; bounded energy error for a harmonic
    oscillator (symplectic);
  * energy, momentum, density-weighted RMS radius: closed forms.
"""
import math

import numpy as np

import oracle

KSRH = 1.936492


def _state(n, **kw):
    return oracle.make_state(n, **{k: np.broadcast_to(np.asarray(v, float), (n,)) for k, v in kw.items()})


def test_eos_is_gamma_five_thirds():
    s = _state(3, dens=[1.0, 2.0, 0.5], uene=[3.0, 1.0, 0.2], ksr=1.0)
    oracle.eos_balsara(s, oracle.params())
    g = 5.0 / 3.0
    assert np.allclose(s["pres"], (g - 1) * s["dens"] * s["uene"], rtol=1e-15)
    assert np.allclose(s["vsnd"], np.sqrt(g * s["pres"] / s["dens"]), rtol=1e-15)


def test_balsara_limits():
    s = _state(3, dens=1.0, uene=1.0, ksr=0.1, divv=[0.0, -50.0, -2.0], rotv=[3.0, 0.0, 2.0])
    oracle.eos_balsara(s, oracle.params())
    assert s["bswt"][0] == 0.0
    assert 0.9999 < s["bswt"][1] < 1.0   # 1 - B = O(1e-4 c kappa / (H |divv|))
    assert abs(s["bswt"][2] - 0.5) < 1e-3


def test_gravity_writeback_scales_by_G_and_zeroes_eta_for_constant_softening():
    """accg2 and pot are G times the functor's sums.  eta itself is pinned by
    energy conservation (tests/test_oracle_conservation.py); here only that a
    constant softening, whose potential does not depend on H, gives eta = 0."""
    s = _state(1, ksr=0.2, grdh=1.1, dens=3.0)
    g = dict(ax=[1.0], ay=[2.0], az=[3.0], pot=[-4.0], eta=[5.0])
    oracle.gravity_writeback(s, oracle.params(G=2.0), g)
    assert s["g2y"][0] == 4.0 and s["pot"][0] == -8.0 and s["eta"][0] > 0.0
    oracle.gravity_writeback(s, oracle.params(G=2.0, eps_const=0.01), g)
    assert s["eta"][0] == 0.0 and s["g2z"][0] == 6.0


def _tail(**kw):
    s = _state(1, **kw)
    h = dict(ax=[0.0], ay=[0.0], az=[0.0], gx=[0.0], gy=[0.0], gz=[0.0], udot=[0.0], vsmx=[1.0],
             diffu=[kw.get("diffu", 0.0)])
    return s, h


def test_alpha_dot():
    p = oracle.params(alpha_max=2.0, alpha_min=0.1)
    s, h = _tail(alph=1.0, vsnd=2.0, ksr=0.5, divv=0.3, uene=1.0)
    oracle.hydro_writeback_tail(s, p, h)
    tauinv = 0.25 * KSRH * 2.0 / 0.5
    assert abs(s["adot"][0] - (-(1.0 - 0.1) * tauinv)) < 1e-14
    s2, h2 = _tail(alph=1.0, vsnd=2.0, ksr=0.5, divv=-0.3, uene=1.0)
    oracle.hydro_writeback_tail(s2, p, h2)
    assert abs(s2["adot"][0] - (s["adot"][0] + 0.3 * (2.0 - 1.0))) < 1e-14


def test_acceleration_sum():
    s = _state(1, g2x=1.5)
    h = dict(ax=[0.25], ay=[0.0], az=[0.0], gx=[2.0], gy=[0.0], gz=[0.0], udot=[0.0], vsmx=[1.0],
             diffu=[0.0])
    s["vsnd"][:] = 1.0
    s["ksr"][:] = 1.0
    oracle.hydro_writeback_tail(s, oracle.params(G=0.5), h)
    assert s["ax"][0] == 0.25 + 1.0 + 1.5


def test_timestep_power_of_two():
    p = oracle.params(c_dt=0.1, dt_max=10.0)
    s = _state(2, ksr=[1.0, 1.0], vsmx=[1.0, 100.0], divv=[0.0, 0.0], g2x=[1e-12, 1e-12])
    # particle dt: 0.1 * 1 / (100 kappa) = 5.16e-4 -> halve 1.0 down to 2^-11
    dt = oracle.timestep(s, p, dt_prev=1.0, time=0.0)
    dtc = 0.1 / (100 * KSRH)
    assert dt <= dtc < 2 * dt and math.log2(dt) == round(math.log2(dt))
    # doubling only on a 2 dt boundary
    s2 = _state(1, ksr=1.0, vsmx=1e-6, divv=0.0, g2x=1e-12)
    assert oracle.timestep(s2, p, dt_prev=0.25, time=1.0) == 0.5
    assert oracle.timestep(s2, p, dt_prev=0.25, time=0.25) == 0.25
    # underflow is reported
    s3 = _state(1, ksr=1.0, vsmx=1e30, divv=0.0, g2x=1e-12)
    assert oracle.timestep(s3, p, dt_prev=1.0, time=0.0, dt_min=1e-16) < 0


def test_kdk_exact_for_constant_acceleration_and_symplectic():
    s = _state(1, m=1.0, x=0.3, vx=-0.2, ax=0.8)
    dt = 0.125
    for _ in range(8):
        oracle.predict(s, dt)
        oracle.correct(s, dt)     # acceleration unchanged
    t = 1.0
    assert abs(s["x"][0] - (0.3 - 0.2 * t + 0.4 * t * t)) < 1e-14
    assert abs(s["vx"][0] - (-0.2 + 0.8 * t)) < 1e-14
    # harmonic oscillator a = -x: energy error stays bounded and O(dt^2)
    s = _state(1, m=1.0, x=1.0, vx=0.0, ax=-1.0)
    dt, emax = 0.05, 0.0
    for _ in range(2000):
        oracle.predict(s, dt)
        s["ax"][:] = -s["x"]
        oracle.correct(s, dt)
        emax = max(emax, abs(0.5 * s["vx"][0] ** 2 + 0.5 * s["x"][0] ** 2 - 0.5))
    assert emax < dt * dt


def test_energy_momentum_ksrmax():
    s = _state(2, m=[1.0, 3.0], vx=[2.0, 0.0], vy=[0.0, 1.0], uene=[0.5, 0.25], pot=[-1.0, -2.0])
    e = oracle.energy(s)
    assert abs(e - (1.0 * (2.0 + 0.5 - 0.5) + 3.0 * (0.5 + 0.25 - 1.0))) < 1e-15
    assert np.allclose(oracle.momentum(s), [2.0, 3.0, 0.0])
    # two equal-density particles at +-1 around the CoM: rms radius 1
    r = oracle.ksr_max([1.0, 1.0], [-1.0, 1.0], [0.0, 0.0], [0.0, 0.0], [2.0, 2.0])
    assert abs(r - 1.0) < 1e-15
    r = oracle.ksr_max([1.0, 1.0, 5.0], [-1.0, 1.0, 9.0], [0, 0, 0], [0, 0, 0], [2.0, 2.0, 7.0],
                       istar=np.array([0, 0, 1], np.int32), istar_sel=0)
    # CoM of star 0 at the origin; rms radius weighted by density over ALL particles
    assert abs(r - math.sqrt((2 + 2 + 7 * 81) / 11.0)) < 1e-14
