"""Device step kernels (include/sphk_step.h) against the oracle's glue on the
same random states, and a short device-resident KDK run.

  * EOS/Balsara, softening, gravity and hydro write-backs (incl. acceleration
    sum and alpha-dot), predict/correct: within 1e-13 relative (FP64 kernels,
    same formulas; only contraction/rounding order differs);
  * diagnostics: energy, momentum, min particle timestep and mass vs the
    oracle's reductions, to rounding;
  * 10 KDK steps of a 2x4000-particle head-on collision: the mixed-precision
    run's energy drift stays within 2x the FP64 run's (+ 1e-7 absolute slack)
    and positions agree to 1e-6 of the system size.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from research_codes_amd import _lib as L
from research_codes_amd import ic
from research_codes_amd.device import FP64, MIXED, Context
from research_codes_amd.step import SPHSystem, phys

pytestmark = pytest.mark.gpu


def _random_state(n, seed=0):
    rng = np.random.default_rng(seed)
    s = {k: rng.uniform(0.5, 1.5, n) for k in L.STATE_FIELDS}
    for k in ("vx", "vy", "vz", "ax", "ay", "az", "divv", "g1x", "g2x", "diffu", "pot"):
        s[k] = rng.normal(size=n)
    return s


def _dev_state(s):
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in s.items()}
    return t, L.State(**{k: t[k].data_ptr() for k in L.STATE_FIELDS})


def _cmp(t, s, keys, tol=1e-13):
    for k in keys:
        got = t[k].cpu().numpy()
        assert np.allclose(got, s[k], rtol=tol, atol=tol * np.abs(s[k]).max()), k


def test_step_kernels_match_oracle(ctx):
    n = 5000
    lib = L.lib()
    ph = phys(G=0.7, alphu_max=1.0, eps_u=1e-3)
    op = oracle.params(G=0.7, alphu_max=1.0, eps_u=1e-3)
    st = torch.cuda.current_stream().cuda_stream
    h = ctx.handle

    s = _random_state(n, 1)
    t, ss = _dev_state(s)
    L.check(lib.sphk_eos_balsara_dev(h, n, ss, ph, st), "eos")
    oracle.eos_balsara(s, op)
    torch.cuda.synchronize()
    _cmp(t, s, ("pres", "vsnd", "bswt"))

    g = {k: np.random.default_rng(2).normal(size=n) for k in L.GRAV_OUT}
    gt = {k: torch.from_numpy(v).cuda() for k, v in g.items()}
    go = L.GravOut(**{k: gt[k].data_ptr() for k in L.GRAV_OUT})
    L.check(lib.sphk_gravity_writeback_dev(h, n, ss, ph, go, st), "gwb")
    oracle.gravity_writeback(s, op, g)
    torch.cuda.synchronize()
    _cmp(t, s, ("g2x", "g2y", "g2z", "pot", "eta"))

    hp = {k: torch.zeros(n, dtype=torch.float64, device="cuda") for k in ("hinv", "presc", "bswt_half", "thrm")}
    L.check(lib.sphk_hydro_prep_dev(h, n, ss, *(hp[k].data_ptr() for k in ("hinv", "presc", "bswt_half", "thrm")), st), "prep")
    ref = oracle.hydro_prep(s)
    torch.cuda.synchronize()
    for k in ref:
        assert np.allclose(hp[k].cpu().numpy(), ref[k], rtol=1e-14), k

    hh = {k: np.random.default_rng(3).normal(size=n) for k in L.HYDRO_OUT}
    hh["vsmx"] = np.abs(hh["vsmx"]) + 0.1
    ht = {k: torch.from_numpy(v).cuda() for k, v in hh.items()}
    ho = L.HydroOut(**{k: ht[k].data_ptr() for k in L.HYDRO_OUT})
    L.check(lib.sphk_hydro_writeback_dev(h, n, ss, ph, ho, st), "hwb")
    oracle.hydro_writeback_tail(s, op, hh)
    torch.cuda.synchronize()
    _cmp(t, s, ("ax", "ay", "az", "adot", "adotu", "udot", "vsmx", "g1x", "diffu"), 1e-12)

    out = (C.c_double * 6)()
    L.check(lib.sphk_diagnostics(h, n, ss, ph, out, st), "diag")
    assert abs(out[0] / oracle.energy(s) - 1) < 1e-12
    assert np.allclose(out[1:4], oracle.momentum(s), rtol=1e-12)
    assert abs(out[5] - s["m"].sum()) < 1e-12 * s["m"].sum()
    # min particle dt: the oracle's global step uses it through a power-of-two rule
    dt = oracle.timestep(s, op, dt_prev=2.0 ** 10, time=0.0)
    assert dt <= out[4] < 2 * dt

    L.check(lib.sphk_predict_dev(h, n, ss, 0.01, st), "predict")
    oracle.predict(s, 0.01)
    L.check(lib.sphk_correct_dev(h, n, ss, 0.01, st), "correct")
    oracle.correct(s, 0.01)
    torch.cuda.synchronize()
    _cmp(t, s, ("x", "y", "z", "vx", "vy", "vz", "uene", "alph", "alphu"), 1e-13)


def _run(prec, steps=10, n_each=4000):
    p = ic.head_on(n_each, eta=1.18)
    with Context(0) as c:
        s = SPHSystem(p, c, eta=1.18, eps=0.01, precision=prec, n_group=16)
        s.initialize()
        e0 = s.diagnostics()["energy"]
        es = []
        for _ in range(steps):
            es.append(s.kdk()["energy"])
        e1 = s.diagnostics()["energy"]
        order = torch.argsort(s.id.long())
        pos = torch.stack([s.s[k][order] for k in ("x", "y", "z")]).cpu().numpy()
        return e0, e1, pos, s.time


def test_kdk_energy_drift_mixed_vs_fp64():
    e0d, e1d, pd, td = _run(FP64)
    e0m, e1m, pm, tm = _run(MIXED)
    assert td == tm
    drift64 = abs(e1d / e0d - 1)
    driftm = abs(e1m / e0m - 1)
    print(f"[kdk] t={td:.4g} drift fp64 {drift64:.3e} mixed {driftm:.3e}")
    assert driftm <= 2 * drift64 + 1e-7
    assert np.max(np.abs(pm - pd)) < 1e-6 * 3.0


def test_density_retry_loop_is_repeatable():
    """Regression: the kernels run on the caller's stream (NULL = HIP null
    stream), so torch work between launches is ordered with them; the retry
    loop from one snapshot must give bitwise identical H, rho and rs."""
    p = ic.head_on(4000, eta=1.18)
    with Context(0) as c:
        s = SPHSystem(p, c, eta=1.18, eps=0.01, precision=FP64, n_group=16)
        snap = {k: v.clone() for k, v in s.s.items()}
        res = []
        for _ in range(3):
            for k, v in snap.items():
                s.s[k].copy_(v)
            s.converge_density()
            res.append([s.s["dens"].clone(), s.s["ksr"].clone(), s.rs.clone()])
        for r in res[1:]:
            for a, b in zip(res[0], r):
                assert torch.equal(a, b)


def test_per_kernel_group_sizes_agree():
    """Each functor may use its own walk grouping.  On one state with a search
    radius well above H (so no intermediate h-pass is cut off by the list),
    the FP64 force step must not depend on it beyond summation-order rounding
    (theta = 0 makes gravity a direct sum, which does not depend on the groups)."""
    p = ic.head_on(3000, eta=1.18)
    with Context(0) as c:
        s = SPHSystem(p, c, eta=1.18, eps=0.01, theta=0.0, precision=FP64, n_group=16)
        s.initialize()
        snap = {k: v.clone() for k, v in s.s.items()}
        res = []
        for groups in ((16, 16, 16), (8, 32, 64), (64, 8, 32)):
            for k, v in snap.items():
                s.s[k].copy_(v)
            s.n_group.update(zip(("dens", "hydro", "grav"), groups))
            s.rs.copy_(1.5 * s.s["ksr"])
            s.build_lists(("dens",))
            s.density()
            assert int(s.itr.sum()) == 0          # no particle outgrew its search radius
            s.eos_balsara()
            s.build_lists(("hydro", "grav"))      # symmetry lists from the new H
            s.gravity()
            s.hydro()
            res.append({k: s.s[k].clone() for k in ("dens", "ksr", "ax", "ay", "az", "udot", "pot")})
    errs = {}
    for b in res[1:]:
        for k, a in res[0].items():
            scale = a.abs().max().clamp_min(1e-300)
            errs[k] = max(errs.get(k, 0.0), float(((a - b[k]).abs() / scale).max()))
    print("[groups] max scaled diff", {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) < 1e-12, errs


def test_unbound_mass_matches_oracle():
    rng = np.random.default_rng(17)
    n = 100_000
    s = oracle.make_state(n)
    s["m"][:] = rng.uniform(0.5, 1.5, n) / n
    s["dens"][:] = rng.uniform(0.1, 2.0, n)
    for c in ("vx", "vy", "vz"):
        s[c][:] = rng.normal(0.3, 1.0, n)
    s["pot"][:] = rng.uniform(-2.0, 0.0, n)
    s["uene"][:] = rng.uniform(0.0, 0.5, n)
    istar = rng.integers(0, 2, n).astype(np.int32)
    t = {k: torch.from_numpy(v).cuda() for k, v in s.items()}
    ist = torch.from_numpy(istar).cuda()
    lib = L.lib()
    with Context(0) as c:
        ss = L.State(**{k: t[k].data_ptr() for k in L.STATE_FIELDS})
        for star in (-1, 0, 1):
            out = (C.c_double * 6)()
            L.check(lib.sphk_unbound_mass_dev(c.handle, n, ss, ist.data_ptr(), star, 20.0, out, None),
                    "unbound")
            ref = oracle.unbound_mass(s, istar=istar, star=star, k=20.0)
            got = np.array(out[:])
            print(f"[unbound] star {star}: {got} vs {ref}")
            assert np.allclose(got, ref, rtol=1e-12, atol=1e-15)


def test_run_driver_round_trip(tmp_path):
    """A run directory in the reference's formats (cgs .data, .head) runs
    through the KDK driver and writes time.log, ASCII snapshots and a hexa dump."""
    from research_codes_amd import formats
    from research_codes_amd.run import run

    p = ic.head_on(1000, eta=1.6)
    nb = formats.CodeUnits.nbody(formats.SOLAR_MASS, formats.SOLAR_RADIUS)
    formats.write_data(tmp_path / "two.data", p, units=nb)
    formats.write_head(tmp_path / "two.head", dict(time=0.0, time_end=1e5, dt_ascii=1e3, dt_hexa=1e4,
                                                   flag_damping=0, flag_nuclear=0, alpha_max=2.0,
                                                   alpha_min=0.1, alphu_max=0.0, alphu_min=0.0,
                                                   c_dt=0.1, n=len(p["m"])))
    res = run(tmp_path / "two.head", tmp_path / "two.data", tmp_path / "out", max_steps=3)
    rows = formats.read_time_log(tmp_path / "out" / "time.log")
    assert rows.shape[0] == 3 and list(rows[:, 0]) == [0, 1, 2]
    # the first step relaxes H from the sampled initial condition (eps = H/kappa
    # changes with it, and the reference drops the d eps / dH term): energy
    # moves by ~1e-3 there and by far less afterwards
    e = rows[:, 4]
    assert np.all(np.abs(e / e[0] - 1) < 1e-2)
    assert abs(e[2] / e[1] - 1) < 1e-4
    snap0 = formats.read_snapshot(tmp_path / "out" / "sph_t000000.dat")
    assert np.array_equal(snap0["id"], np.sort(p["id"]))
    # positions on disk are cgs; read back in reference code units
    order = np.argsort(p["id"])
    assert np.allclose(snap0["x"], p["x"][order] * nb.length / formats.REFERENCE_UNITS.length, rtol=1e-5,
                       atol=1e-6 * nb.length / formats.REFERENCE_UNITS.length)
    hexa = formats.read_hexa(tmp_path / "out" / "sph_t000000.hex")
    assert np.array_equal(hexa["id"], snap0["id"])
    assert res["steps"] == 3


@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
def test_ensemble_batched_equals_per_run(prec):
    """Independent runs concatenated into one state (one tree per run, one
    walk table per functor) give bitwise the results of running each alone."""
    p, counts, _ = ic.ensemble(3, 1500, 3000, seed=77)

    def force(q, segments):
        with Context(0) as c:
            s = SPHSystem(q, c, eta=1.18, eps=0.01, precision=prec, n_group=16, hmax=1e30,
                          segments=segments)
            s.converge_density()
            s.eos_balsara()
            s.build_lists(("hydro", "grav"))
            s.gravity()
            s.hydro()
            snap = s.snapshot()
            return {k: snap[k] for k in ("id", "dens", "ksr", "ax", "ay", "az", "udot", "pot")}

    whole = force(p, counts)
    parts, a = [], 0
    for n in counts:
        parts.append(force({k: v[a:a + n] for k, v in p.items()}, None))
        a += n
    for k in whole:
        got = np.concatenate([q[k] for q in parts])
        assert np.array_equal(whole[k], got), k


@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
def test_captured_force_step_replays_bitwise(prec):
    """The force step captured in a HIP graph gives the eager step's results
    on every replay, with eager work and synchronisations between replays."""
    p = ic.head_on(3000, eta=1.18)
    with Context(0) as c:
        s = SPHSystem(p, c, eta=1.18, eps=0.01, precision=prec, n_group=16)
        s.initialize()
        s.rs.copy_(1.1 * s.s["ksr"])
        s.build_lists()
        snap = {k: v.clone() for k, v in s.s.items()}

        def restore():
            for k, v in snap.items():
                s.s[k].copy_(v)

        restore()
        s.force(prec)
        ref = {k: s.s[k].clone() for k in ("dens", "ksr", "ax", "ay", "az", "udot", "pot", "vsmx")}
        g = s.capture_force(prec)
        for rep in range(4):
            restore()
            _ = (s.s["x"] * 2.0).sum().item()     # eager work between replays
            g.replay()
            torch.cuda.synchronize()
            for k, v in ref.items():
                assert torch.equal(s.s[k], v), (rep, k)


def test_collision_loss_finite_difference_sensitivity():
    """The smoothed unbound mass of a short fixed-step collision is smooth in
    the impactor's offset and velocity: central differences with h and h/2
    agree (the finite-difference side of the sensitivity check)."""
    from research_codes_amd.sensitivity import fd_gradient, head_on_loss

    def f(dx0, dv0):
        return head_on_loss(dx0, dv0, n_each=1500, steps=4, dt=2e-3, precision=FP64)

    base = f(0.0, 0.0)
    assert 0.0 < base < 2.0
    g1 = fd_gradient(f, [0.0, 0.0], h=1e-4)
    g2 = fd_gradient(f, [0.0, 0.0], h=5e-5)
    print(f"[sens] loss {base:.9e} grad(h) {g1} grad(h/2) {g2}")
    assert np.all(np.isfinite(g1)) and np.all(np.abs(g1 - g2) <= 1e-2 * np.abs(g1).max() + 1e-9)


@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64", "mixed"])
def test_persistent_launches_equal_ordinary(prec):
    """Persistent launches (grid-resident workgroups taking walks from the
    per-XCD queues) give bitwise the ordinary launches' results: an ensemble
    of three runs in one launch per functor, again (each launch leaves its
    queue zeroed for the next), and replayed from a HIP graph; a run with
    fewer walks than the grid holds workgroups (walk count not a multiple
    of 8) too: the grid's surplus workgroups claim nothing."""
    p, counts, _ = ic.ensemble(3, 1500, 3000, seed=78)
    with Context(0) as c:
        s = SPHSystem(p, c, eta=1.18, eps=0.01, precision=prec, n_group=16, hmax=1e30, segments=counts)
        s.initialize()
        s.rs.copy_(1.1 * s.s["ksr"])
        s.build_lists()
        snap = {k: v.clone() for k, v in s.s.items()}
        keys = ("dens", "ksr", "np", "ax", "ay", "az", "udot", "pot", "vsmx")

        def step(persistent):
            c.set_params(persistent=persistent)
            for k, v in snap.items():
                s.s[k].copy_(v)
            s.force(prec)
            torch.cuda.synchronize()
            return {k: (s.np_ if k == "np" else s.s[k]).clone() for k in keys}

        ref = step(False)
        assert all(bool(torch.isfinite(v.double()).all()) for v in ref.values())
        for rep in range(3):
            got = step(True)
            for k in keys:
                assert torch.equal(got[k], ref[k]), (rep, k)
        g = s.capture_force(prec)            # persistent is still set
        for rep in range(3):
            for k, v in snap.items():
                s.s[k].copy_(v)
            _ = (s.s["x"] * 2.0).sum().item()
            g.replay()
            torch.cuda.synchronize()
            for k in keys:
                assert torch.equal((s.np_ if k == "np" else s.s[k]), ref[k]), ("graph", rep, k)
        c.set_params(persistent=False)

    small = ic.head_on(200, eta=1.18)
    out = []
    for persistent in (False, True):
        with Context(0) as c:
            c.set_params(persistent=persistent)
            s = SPHSystem(small, c, eta=1.18, eps=0.01, precision=prec, n_group=16)
            s.initialize()
            print(f"[persistent] small run: walks {s.lists['dens'].n_walk} / {s.lists['hydro'].n_walk} / "
                  f"{s.lists['grav'].n_walk}", flush=True)
            out.append({k: s.s[k].clone() for k in ("dens", "ax", "ay", "az", "udot")})
    assert all(bool(torch.isfinite(v).all()) for v in out[0].values())
    for k in out[0]:
        assert torch.equal(out[0][k], out[1][k]), k
