"""Adjoint of the gravity functor (include/sphk_adjoint.h, research_codes_amd.adjoint).

Checked against (1) the identity d(1/2 sum m pot)/dx = -m a for direct
summation, (2) torch autograd of an all-pairs FP64 reference, (3) central
finite differences of the tree forward (theta = 0.5) at fixed tree topology,
and (4) central finite differences (eps = 1e-4) of a checkpointed
kick-drift-kick rollout's smoothed unbound mass."""
import numpy as np
import pytest
import torch

from research_codes_amd.adjoint import (TreeGravity, gravity_rollout, potential_energy,
                                        smoothed_unbound_mass)
from research_codes_amd.device import FP64, MIXED, Context
from research_codes_amd.tree import DeviceTree, morton_order_dev

pytestmark = pytest.mark.gpu


def _clusters(n, seed, sep=4.0):
    g = torch.Generator().manual_seed(seed)
    pos = torch.randn(n, 3, generator=g, dtype=torch.float64)
    pos[n // 2:, 0] += sep
    m = torch.rand(n, generator=g, dtype=torch.float64) * 0.5 + 0.5
    return [pos[:, k].cuda().contiguous() for k in range(3)], (m / m.sum()).cuda()


@pytest.fixture(scope="module")
def ctx():
    with Context(0) as c:
        yield c


def test_potential_gradient_is_minus_m_a(ctx):
    (x, y, z), m = _clusters(600, 1)
    grav = TreeGravity(ctx, m, eps=0.05, theta=0.0, n_group=16)
    xs = [t.clone().requires_grad_(True) for t in (x, y, z)]
    pe = potential_energy(grav, m, *xs)
    g = torch.autograd.grad(pe, xs)
    ax, ay, az, _ = grav(x, y, z)
    scale = float((m * torch.sqrt(ax ** 2 + ay ** 2 + az ** 2)).max())
    err = max(float((gi + m * ai).abs().max()) for gi, ai in zip(g, (ax, ay, az))) / scale
    print(f"[adjoint] |dPE/dx + m a| / max|m a| = {err:.2e}")
    assert err < 1e-12


def _torch_direct(m, e2, x, y, z):
    dx = x[:, None] - x[None, :]
    dy = y[:, None] - y[None, :]
    dz = z[:, None] - z[None, :]
    w = (dx * dx + dy * dy + dz * dz + e2) ** -0.5
    w3 = w ** 3
    return (-(m[None, :] * dx * w3).sum(1), -(m[None, :] * dy * w3).sum(1),
            -(m[None, :] * dz * w3).sum(1), -(m[None, :] * w).sum(1))


def test_vjp_matches_torch_all_pairs(ctx):
    (x, y, z), m = _clusters(300, 2)
    eps = 0.1
    g = torch.Generator(device="cuda").manual_seed(3)
    cot = [torch.randn(300, generator=g, dtype=torch.float64, device="cuda") for _ in range(4)]
    grav = TreeGravity(ctx, m, eps=eps, theta=0.0, n_group=16)
    xs = [t.clone().requires_grad_(True) for t in (x, y, z)]
    out = grav(*xs)
    ours = torch.autograd.grad(sum((o * c).sum() for o, c in zip(out, cot)), xs)
    xr = [t.clone().requires_grad_(True) for t in (x, y, z)]
    ref_out = _torch_direct(m, eps * eps, *xr)
    for o, r in zip(out, ref_out):
        assert torch.allclose(o.detach(), r.detach(), rtol=1e-12, atol=1e-12 * float(r.detach().abs().max()))
    ref = torch.autograd.grad(sum((o * c).sum() for o, c in zip(ref_out, cot)), xr)
    for a, b in zip(ours, ref):
        err = float((a - b).abs().max() / b.abs().max())
        print(f"[adjoint] VJP vs torch all pairs: {err:.2e}")
        assert err < 1e-11


def _lists(x, y, z, m, eps, theta):
    perm = morton_order_dev(x, y, z)
    xs, ys, zs = x[perm].contiguous(), y[perm].contiguous(), z[perm].contiguous()
    t = DeviceTree(xs, ys, zs, m[perm].contiguous(), torch.full_like(xs, eps * eps), 8, 32)
    out = [perm.cpu()] + [a.cpu() for a in t.long(theta)]
    t.close()
    return out


@pytest.mark.parametrize("prec", [FP64, MIXED], ids=["fp64fwd", "mixedfwd"])
def test_tree_vjp_finite_difference(ctx, prec):
    (x, y, z), m = _clusters(800, 4, sep=6.0)
    eps, theta, h = 0.05, 0.5, 1e-7
    g = torch.Generator(device="cuda").manual_seed(5)
    v = [torch.randn(800, generator=g, dtype=torch.float64, device="cuda") for _ in range(3)]
    cot = [torch.randn(800, generator=g, dtype=torch.float64, device="cuda") for _ in range(4)]
    grav = TreeGravity(ctx, m, eps=eps, theta=theta, n_group=32, precision=FP64)
    plus = [a + h * b for a, b in zip((x, y, z), v)]
    minus = [a - h * b for a, b in zip((x, y, z), v)]
    # the derivative holds the tree fixed: the perturbation must not change it
    for a, b in zip(_lists(*plus, m, eps, theta), _lists(*minus, m, eps, theta)):
        assert torch.equal(a, b)
    fd = (sum((o * c).sum() for o, c in zip(grav(*plus), cot))
          - sum((o * c).sum() for o, c in zip(grav(*minus), cot))) / (2 * h)
    grav.precision = prec        # the backward is FP64 in both modes
    xs = [t.clone().requires_grad_(True) for t in (x, y, z)]
    gr = torch.autograd.grad(sum((o * c).sum() for o, c in zip(grav(*xs), cot)), xs)
    jv = sum((a * b).sum() for a, b in zip(gr, v))
    rel = abs(float(jv - fd)) / abs(float(fd))
    print(f"[adjoint] tree VJP.v {float(jv):.9e} vs FD {float(fd):.9e}: {rel:.2e}")
    assert rel < 1e-6


def test_rollout_gradient_finite_difference(ctx):
    """Sensitivity of a collision outcome to the impactor's initial offset and
    velocity: 12 checkpointed KDK steps of two clusters, loss = smoothed
    unbound mass; autograd vs central differences with eps = 1e-4."""
    (x, y, z), m = _clusters(128, 6, sep=3.0)
    vx = torch.zeros_like(x)
    vx[64:] = -0.6
    vy, vz = torch.zeros_like(x), torch.zeros_like(x)
    imp = torch.zeros_like(x)
    imp[64:] = 1.0
    grav = TreeGravity(ctx, m, eps=0.1, theta=0.0, n_group=16)

    def loss(p):   # p = (dx0, dvx0): impactor offset and velocity change
        xx = x + p[0] * imp
        vv = vx + p[1] * imp
        st = gravity_rollout(grav, xx, y, z, vv, vy, vz, dt=0.05, steps=12, checkpoint_every=5)
        return smoothed_unbound_mass(grav, m, *st, k=20.0)

    p = torch.zeros(2, dtype=torch.float64, device="cuda", requires_grad=True)
    gr = torch.autograd.grad(loss(p), p)[0]
    h = 1e-4
    fd = []
    for k in range(2):
        e = torch.zeros(2, dtype=torch.float64, device="cuda")
        e[k] = h
        with torch.no_grad():
            fd.append(float((loss(e) - loss(-e)) / (2 * h)))
    print(f"[adjoint] rollout gradient {gr.tolist()} vs FD {fd}")
    for a, b in zip(gr.tolist(), fd):
        assert abs(a - b) <= 1e-5 * max(abs(b), 1e-3)
