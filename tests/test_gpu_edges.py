"""Edge cases of the three functors through the C-ABI, against the oracle.

An empty walk table launches nothing and leaves the outputs alone.  Tiny
hand-made walks cover what the polytrope cases do not reach: an isolated
particle whose gather list is itself (its H runs up to H_max, itr set), a
walk whose list holds neighbours beyond every support, ragged walks of one
and two i-particles, a gravity walk with empty EPJ and SPJ lists (zero
force, no self term), and an SPJ-only contribution.  Bars as in the parity
tests: np and itr bit-exact, dens and ksr to 1e-12, the FP64 path's other
outputs to 1e-10 of their scale, the mixed path's to 1e-4."""
import numpy as np
import pytest
import torch

import oracle
from research_codes_amd import device as D
from tests._cases import scaled_err

pytestmark = pytest.mark.gpu

KSRH = 1.936492


def _particles():
    rng = np.random.default_rng(11)
    x = np.array([[5.0, 5.0, 5.0], [0.0, 0.0, 0.0], [0.05, 0.0, 0.0], [0.3, 0.0, 0.0],
                  [0.1, 0.1, 0.0], [0.1, 0.12, 0.01]])
    n = len(x)
    p = {"x": x[:, 0].copy(), "y": x[:, 1].copy(), "z": x[:, 2].copy(), "m": np.full(n, 1e-3),
         "vx": 0.1 * rng.normal(size=n), "vy": 0.1 * rng.normal(size=n), "vz": 0.1 * rng.normal(size=n),
         "id": np.arange(n, dtype=np.int32), "ksr": np.array([0.2, 0.1, 0.1, 0.1, 0.08, 0.08])}
    p["rs"] = 1.1 * p["ksr"]
    return p, rng


# density / hydro walks: [0] <- {0}; [1, 2] <- {1, 2, 4}; [3] <- {1, 2, 3}; [4, 5] <- {4, 5, 1, 2}
I_OFF = np.array([0, 1, 3, 4, 6], np.int32)
J_OFF = np.array([0, 1, 4, 7, 11], np.int32)
J_IDX = np.array([0, 1, 2, 4, 1, 2, 3, 4, 5, 1, 2], np.int32)


def _dev(d):
    return {k: D.to_device(v) for k, v in d.items()}


def test_empty_walk_table_is_a_no_op(ctx):
    p, _ = _particles()
    w = D.Walks(np.array([0], np.int32), np.array([0], np.int32), np.zeros(0, np.int32))
    assert w.n_walk == 0
    src = _dev({k: p[k] for k in D.L.DENS_SRC})
    out = D.empty_dens_out(len(p["m"]))
    for v in out.values():
        v.fill_(7)
    ctx.density(w, src, src, out, D.FP64)
    g = {"m": p["m"], "x": p["x"], "y": p["y"], "z": p["z"], "eps2": (p["ksr"] / KSRH) ** 2}
    gout = D.empty_out(D.L.GRAV_OUT, len(p["m"]))
    for v in gout.values():
        v.fill_(7)
    ctx.gravity(w, _dev(g), _dev(g), gout, precision=D.MIXED)
    torch.cuda.synchronize()
    assert all(bool((v == 7).all()) for v in list(out.values()) + list(gout.values()))


@pytest.mark.parametrize("prec", [D.FP64, D.MIXED], ids=["fp64", "mixed"])
def test_tiny_density_walks_match_oracle(ctx, prec):
    p, _ = _particles()
    hmax = 10.0
    ref = oracle.density(I_OFF, J_OFF, J_IDX, p, p, eta=1.6, ksrh=KSRH, hmax=hmax)
    ctx.set_params(eta=1.6, ksrh=KSRH, hmax=hmax)
    src = _dev({k: p[k] for k in D.L.DENS_SRC})
    out = D.empty_dens_out(len(p["m"]))
    ctx.density(D.Walks(I_OFF, J_OFF, J_IDX), src, src, out, prec)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    assert ref["itr"][0] == 1 and ref["ksr"][0] > 2 * p["rs"][0]   # the isolated particle's H ran away
    assert np.array_equal(got["np"], ref["np"])
    assert np.array_equal(got["itr"], ref["itr"])
    assert scaled_err(got["ksr"], ref["ksr"]) < 1e-12
    assert scaled_err(got["dens"], ref["dens"]) < 1e-12
    tol = 1e-10 if prec == D.FP64 else 1e-4
    for k in ("divv", "rotv", "grdh"):
        e = scaled_err(got[k], ref[k], floor_frac=1e-3 if prec == D.FP64 else 1e-1)
        print(f"[edges] density[{prec}] {k}: {e:.2e}")
        assert e < tol, k


@pytest.mark.parametrize("prec", [D.FP64, D.MIXED], ids=["fp64", "mixed"])
def test_tiny_hydro_walks_match_oracle(ctx, prec):
    p, rng = _particles()
    n = len(p["m"])
    h = {k: p[k] for k in ("id", "m", "x", "y", "z", "vx", "vy", "vz")}
    h["hinv"] = 1.0 / p["ksr"]
    for k in ("pres", "presc", "dens", "vsnd", "alph", "thrm", "eta"):
        h[k] = rng.uniform(0.5, 2.0, n)
    h["bswt"] = rng.uniform(0.0, 0.5, n)
    h["alphu"] = rng.uniform(0.0, 1.0, n)
    ref = oracle.hydro(I_OFF, J_OFF, J_IDX, h, h)
    out = D.empty_out(D.L.HYDRO_OUT, n)
    src = _dev(h)
    ctx.hydro(D.Walks(I_OFF, J_OFF, J_IDX), src, src, out, prec)
    torch.cuda.synchronize()
    tol = 1e-10 if prec == D.FP64 else 1e-4
    for k in D.L.HYDRO_OUT:
        e = scaled_err(out[k].cpu().numpy(), ref[k], floor_frac=1e-3 if prec == D.FP64 else 1e-1)
        print(f"[edges] hydro[{prec}] {k}: {e:.2e}")
        assert e < tol, k


@pytest.mark.parametrize("prec", [D.FP64, D.MIXED], ids=["fp64", "mixed"])
def test_tiny_gravity_walks_match_oracle(ctx, prec):
    p, _ = _particles()
    n = len(p["m"])
    g = {"m": p["m"], "x": p["x"], "y": p["y"], "z": p["z"], "eps2": (p["ksr"] / KSRH) ** 2}
    # [0] <- nothing; [1, 2] <- {0, 1, 2}; [3] <- {3}; [4, 5] <- all six and one far SPJ
    j_off = np.array([0, 0, 3, 4, 10], np.int32)
    j_idx = np.array([0, 1, 2, 3, 0, 1, 2, 3, 4, 5], np.int32)
    s_off = np.array([0, 0, 0, 0, 1], np.int32)
    s_idx = np.array([0], np.int32)
    ps = {"m": np.array([0.5]), "x": np.array([-3.0]), "y": np.array([1.0]), "z": np.array([2.0]),
          "eps2": np.array([1e-4])}
    ref = oracle.gravity(I_OFF, j_off, j_idx, g, g, s_off, s_idx, ps)
    ctx.set_params(uniform_softening=False)
    out = D.empty_out(D.L.GRAV_OUT, n)
    w = D.Walks(I_OFF, j_off, j_idx, s_off, s_idx)
    ctx.gravity(w, _dev(g), _dev(g), out, ps=_dev(ps), precision=prec)
    torch.cuda.synchronize()
    got = {k: v.cpu().numpy() for k, v in out.items()}
    assert all(got[k][0] == 0.0 for k in D.L.GRAV_OUT)        # empty lists: no force, no self term
    tol = 1e-10 if prec == D.FP64 else 1e-4
    for k in D.L.GRAV_OUT:
        e = scaled_err(got[k], ref[k], floor_frac=1e-1 if k in ("ax", "ay", "az") else 1e-3)
        print(f"[edges] gravity[{prec}] {k}: {e:.2e}")
        assert e < tol, k


def test_unknown_precision_is_rejected(ctx):
    p, _ = _particles()
    src = _dev({k: p[k] for k in D.L.DENS_SRC})
    out = D.empty_dens_out(len(p["m"]))
    with pytest.raises(D.L.SphkError, match="precision"):
        ctx.density(D.Walks(I_OFF, J_OFF, J_IDX), src, src, out, 2)
    g = _dev({"m": p["m"], "x": p["x"], "y": p["y"], "z": p["z"], "eps2": (p["ksr"] / KSRH) ** 2})
    with pytest.raises(D.L.SphkError, match="precision"):
        ctx.gravity(D.Walks(I_OFF, J_OFF, J_IDX), g, g, D.empty_out(D.L.GRAV_OUT, len(p["m"])), precision=-1)
