"""Quantifies the reference's shipped CPU path, the AVX2 calcDensity
(hdr_density.hpp:82-242, built with USE_INTRINSICS), against the scalar
functor the oracle restates: the only arithmetic difference is that 1/x and
1/sqrt(x) are the single-precision hardware estimates refined by 4th-order
series (rcp_4th, rsqrt_4th, vector_x86.hpp:501-534).

  * over 40 decades, rsqrt_4th is within 2 ulp of 1/sqrt(x) (its series
    truncation, 35/128 h^5, is below FP64 rounding) and rcp_4th within
    2e-14 of 1/x (truncation h^4, h up to 1.5 * 2^-12 from the estimate);
  * on a seeded n=1.5 star with group lists, the AVX2 arithmetic gives the
    scalar functor's density and H to 1e-14 and Omega^-1, div v and |curl v|
    to 1e-12 of their scale, with identical neighbour counts: the shipped
    CPU path and the scalar oracle agree to FP64 rounding, so the scalar
    restatement stands for both;
  * the vectorised form of that arithmetic (four i-particles per AVX2
    vector, as the shipped branch runs; the CPU baseline's density) equals
    the one-lane restatement bit for bit, also for walks whose particle
    count is not a multiple of four;
  * the vector hydro and gravity functors of the CPU baseline (the shipped
    USE_INTRINSICS branches, hdr_hydro.hpp:128-311, hdr_gravity.hpp:194-291)
    agree with the scalar functors to 1e-11 of each field's scale (rcp_4th
    is within 2e-14 of 1/x per call; diffu sums terms of both signs), with
    conduction (alpha_u > 0) and the SPJ lists live.
"""
import numpy as np
import pytest

import oracle
from tests._cases import polytrope_case, scaled_err


def test_refined_reciprocals():
    x = np.logspace(-20, 20, 20001)
    lib = oracle.lib()
    r = np.array([lib.orc_rcp_4th(float(v)) for v in x])
    s = np.array([lib.orc_rsqrt_4th(float(v)) for v in x])
    assert np.abs(r * x - 1).max() < 2e-14
    assert np.abs(s * np.sqrt(x) - 1).max() < 5e-16


def test_avx2_density_matches_scalar():
    case = polytrope_case(3000, seed=1002, n_group=32)
    p = case["p"]
    a = oracle.density(case["i_off"], *case["dens_lists"], p, p, eta=case["eta"], hmax=case["hmax"])
    b = oracle.density(case["i_off"], *case["dens_lists"], p, p, eta=case["eta"], hmax=case["hmax"],
                       avx2=True)
    assert np.array_equal(a["np"], b["np"])
    assert np.array_equal(a["itr"], b["itr"])
    for k in ("dens", "ksr"):
        assert scaled_err(b[k], a[k]) < 1e-14, k
    for k in ("grdh", "divv", "rotv"):
        assert scaled_err(b[k], a[k], 1e-1) < 1e-12, k


@pytest.mark.skipif(not oracle.lib().orc_density_simd_supported(), reason="host CPU without AVX2")
def test_vector_density_equals_one_lane_restatement():
    case = polytrope_case(3000, seed=1003, n_group=32)
    p = case["p"]
    ni = np.diff(case["i_off"])
    assert (ni % 4 != 0).any() and (ni % 4 == 0).any()
    a = oracle.density(case["i_off"], *case["dens_lists"], p, p, eta=case["eta"], hmax=case["hmax"],
                       avx2=True)
    b = oracle.density(case["i_off"], *case["dens_lists"], p, p, eta=case["eta"], hmax=case["hmax"],
                       simd=True)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.skipif(not oracle.lib().orc_density_simd_supported(), reason="host CPU without AVX2")
def test_vector_hydro_and_gravity_match_scalar():
    case = polytrope_case(3000, seed=1004, n_group=32)
    hs = dict(case["hydro_src"])
    hs["alphu"] = np.random.default_rng(4).uniform(0.1, 1.0, len(hs["m"]))
    io, (jo, ji) = case["hydro_i_off"], case["hydro_lists"]
    a = oracle.hydro(io, jo, ji, hs, hs)
    b = oracle.hydro(io, jo, ji, hs, hs, simd=True)
    assert np.abs(a["diffu"]).max() > 0 and np.abs(a["udot"]).max() > 0
    for k in a:
        assert scaled_err(b[k], a[k]) < 1e-11, k
    jl, g = case["grav_lists"], case["grav_src"]
    assert len(jl[3]) > 0
    a = oracle.gravity(case["grav_i_off"], jl[0], jl[1], g, g, s_off=jl[2], s_idx=jl[3], ps=case["grav_mom"])
    b = oracle.gravity(case["grav_i_off"], jl[0], jl[1], g, g, s_off=jl[2], s_idx=jl[3], ps=case["grav_mom"],
                       simd=True)
    for k in a:
        assert scaled_err(b[k], a[k]) < 1e-12, k
