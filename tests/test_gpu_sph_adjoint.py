"""Reverse-mode gradient of an SPH collision run (research_codes_amd.sph_adjoint).

(1) The differentiable restatement of the initial force evaluation and of a
KDK step reproduces the HIP step (FP64) on the recorded lists, field by
field; (2) the gradient of the smoothed unbound mass after four fixed steps
with respect to the impactor's offset and velocity change matches central
differences of the HIP run itself (theta = 0: direct-sum gravity, so the
outcome is smooth in the initial conditions)."""
import numpy as np
import pytest
import torch

from research_codes_amd import ic
from research_codes_amd.device import FP64, Context
from research_codes_amd.sensitivity import fd_gradient, head_on_loss
from research_codes_amd.sph_adjoint import STATE, _state, head_on_gradient, init_t, step_t
from research_codes_amd.step import SPHSystem

pytestmark = pytest.mark.gpu

N_EACH = 1500
DT = 2e-3


@pytest.fixture(scope="module")
def ctx():
    with Context(0) as c:
        yield c


def _compare(out, s, tol, tag):
    worst = 0.0
    for k in STATE:
        a, b = out[k].detach(), s.s[k]
        err = float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))
        worst = max(worst, err)
        assert err <= tol, f"{tag}: {k} differs by {err:.3e} (relative to its largest value)"
    print(f"[sph-adjoint] {tag}: restated state within {worst:.2e} of the HIP state")


def test_restated_steps_match_hip(ctx):
    p = ic.head_on(N_EACH, eta=1.18)
    s = SPHSystem(p, ctx, eta=1.18, eps=0.01, theta=0.5, precision=FP64, n_group=32)
    st, m, rec = _state(s), s.s["m"].clone(), {}
    s.initialize(record=rec)
    with torch.no_grad():
        out = init_t(st, m, rec, s)
    _compare(out, s, 1e-10, "initial force")
    for n in range(2):
        st, m, rec = _state(s), s.s["m"].clone(), {}
        s.kdk(dt=DT, record=rec)
        with torch.no_grad():
            out = step_t(st, m, rec, s, DT)
        _compare(out, s, 1e-10, f"step {n}")


@pytest.mark.parametrize("k", [20.0, 2.0])
def test_collision_gradient_matches_finite_differences(k):
    """k = 20 is the loss's slope in the sensitivity study; k = 2 gives
    gradients four orders of magnitude larger at this short run."""
    steps = 4
    loss, g, _ = head_on_gradient(n_each=N_EACH, steps=steps, dt=DT, k=k, theta=0.0, checkpoint_every=3)

    def f(dx0, dv0):
        return head_on_loss(dx0, dv0, n_each=N_EACH, steps=steps, dt=DT, k=k, precision=FP64, theta=0.0)

    base = f(0.0, 0.0)
    fd = fd_gradient(f, [0.0, 0.0], h=1e-4)
    print(f"[sph-adjoint] k={k}: loss {loss:.12e} (FD run {base:.12e}); adjoint {g} vs FD {fd}")
    assert abs(loss - base) <= 1e-12 * abs(base)
    assert np.all(np.isfinite(g)) and np.all(g != 0.0)
    assert np.all(np.abs(g - fd) <= 1e-5 * np.abs(fd).max())


def test_long_rollout_gradient_matches_finite_differences():
    """200 steps with the power-of-two timestep (held fixed by the gradient
    and replayed by the differences), stars in contact at t=0, direct-sum
    gravity: the contact shock makes the loss O(0.1)."""
    n_each, steps = 3000, 200
    loss, g, dts = head_on_gradient(n_each=n_each, steps=steps, dt=None, k=20.0, theta=0.0,
                                    separation=2.0)

    def f(dx0, dv0):
        return head_on_loss(dx0, dv0, n_each=n_each, k=20.0, precision=FP64, theta=0.0,
                            separation=2.0, dts=dts)

    fd = fd_gradient(f, [0.0, 0.0], h=1e-5)
    rel = np.abs(g - fd) / np.abs(fd).max()
    print(f"[sph-adjoint] 200 steps to t={sum(dts):.4g}: loss {loss:.6e}, adjoint {g} vs FD {fd}, rel {rel}")
    assert loss > 1e-3 and np.all(rel <= 1e-6)


def test_collision_run_rejects_unsupported_systems(ctx):
    from research_codes_amd.device import MIXED
    from research_codes_amd.sph_adjoint import CollisionRun
    p = ic.head_on(300, eta=1.18)
    with pytest.raises(ValueError, match="FP64"):
        CollisionRun(SPHSystem(p, ctx, eta=1.18, eps=0.01, precision=MIXED), DT)
    with pytest.raises(ValueError, match="softening"):
        CollisionRun(SPHSystem(p, ctx, eta=1.18, precision=FP64), DT)
    n = len(p["x"])
    with pytest.raises(ValueError, match="ensemble"):
        CollisionRun(SPHSystem(p, ctx, eta=1.18, eps=0.01, precision=FP64, segments=[n // 2, n - n // 2]), DT)
