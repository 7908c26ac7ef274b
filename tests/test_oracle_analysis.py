"""Pins of the oracle's analysis functions (unbound mass) by closed forms
and invariants (analysis/unbound_mass/unbound_mass.py:10-28,
analysis/common.py:71-90)."""
import math

import numpy as np

import oracle


def _state(n, **kw):
    s = oracle.make_state(n)
    for k, v in kw.items():
        s[k][:] = v
    return s


def test_bound_particles_at_rest():
    # e = pot + u = -0.5 for every particle: nothing unbound; sigma(-10) each
    s = _state(7, m=0.25, dens=1.0, pot=-1.0, uene=0.5)
    out = oracle.unbound_mass(s, k=20.0)
    assert out[3] == 0.0
    assert math.isclose(out[4], 7 * 0.25 / (1 + math.exp(10.0)), rel_tol=1e-14)
    assert out[5] == 7 * 0.25


def test_two_particles_by_hand():
    # v_cg is density weighted: rho = (3, 1), v = (0, 4) -> v_cg = 1
    s = _state(2, m=1.0, pot=-3.0, uene=0.0)
    s["dens"][:] = [3.0, 1.0]
    s["vx"][:] = [0.0, 4.0]
    out = oracle.unbound_mass(s, k=1.0)
    assert out[0] == 1.0 and out[1] == 0.0 and out[2] == 0.0
    # e_0 = 0.5 * 1 - 3 = -2.5 (bound); e_1 = 0.5 * 9 - 3 = 1.5 (unbound)
    assert out[3] == 1.0
    sig = lambda x: 1.0 / (1.0 + math.exp(-x))   # noqa: E731
    assert math.isclose(out[4], sig(-2.5) + sig(1.5), rel_tol=1e-15)
    # mass weighting would give v_cg = 2 and e_0 = -1, e_1 = -1: nothing unbound


def test_galilean_boost_invariance():
    rng = np.random.default_rng(3)
    n = 500
    s = _state(n)
    s["m"][:] = rng.uniform(0.5, 1.5, n)
    s["dens"][:] = rng.uniform(0.1, 2.0, n)
    for c in ("vx", "vy", "vz"):
        s[c][:] = rng.normal(0, 1, n)
    s["pot"][:] = rng.uniform(-2.0, 0.0, n)
    s["uene"][:] = rng.uniform(0.0, 0.5, n)
    a = oracle.unbound_mass(s, k=20.0)
    for c, dv in (("vx", 3.0), ("vy", -1.5), ("vz", 0.25)):
        s[c] += dv
    b = oracle.unbound_mass(s, k=20.0)
    assert np.allclose(b[:3] - a[:3], [3.0, -1.5, 0.25], rtol=0, atol=1e-12)
    assert a[3] == b[3]
    assert math.isclose(a[4], b[4], rel_tol=1e-12)
    # the smoothed mass tends to the hard one for a steep sigmoid
    c = oracle.unbound_mass(s, k=1e6)
    assert abs(c[4] - c[3]) < 1e-3 * c[5]


def test_star_selection():
    s = _state(4, m=1.0, dens=1.0, pot=-1.0, uene=0.0)
    s["vx"][:] = [0.0, 0.0, 10.0, -10.0]
    istar = np.array([0, 0, 1, 1], np.int32)
    a = oracle.unbound_mass(s, istar=istar, star=0)
    b = oracle.unbound_mass(s, istar=istar, star=1)
    assert a[3] == 0.0 and a[5] == 2.0
    assert b[0] == 0.0 and b[3] == 2.0 and b[5] == 2.0   # +-10 about v_cg = 0: e = 49
