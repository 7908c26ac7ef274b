"""Differentiable tree gravity (torch.autograd.Function over the HIP functor).

Forward: the gravity functor (sphk_gravity_dev, uniform softening) on lists
from the device tree.  Backward: the hand-written vector-Jacobian product
(sphk_gravity_vjp_dev, include/sphk_adjoint.h) followed by the tree spread of
the node cotangents to the particles (sphk_dtree_spread_dev).  The tree and
lists are held fixed between forward and backward, as a tree code's
derivative is (piecewise constant in the positions).

`gravity_rollout` integrates a pure-gravity system with kick-drift-kick
steps through this Function, recomputing each segment of `checkpoint_every`
steps in the backward pass (torch.utils.checkpoint), so memory holds one
state per segment instead of one per step.
"""
from __future__ import annotations

import torch
from torch.utils.checkpoint import checkpoint

from . import _lib as L
from ._lib import check, lib
from .device import FP64, Context, Walks
from .tree import DeviceTree, morton_order_dev


def _stream():
    return torch.cuda.current_stream().cuda_stream


class TreeGravity:
    """a, pot of N particles with masses m and softening eps (uniform).

    Call with float64 CUDA position tensors (any order); returns
    (ax, ay, az, pot) times G, differentiable with respect to the positions.
    `with_eta` also returns the functor's eta sum (sum m_j w^3, no G), the
    input of the SPH grad-h gravity term."""

    def __init__(self, ctx: Context, m: torch.Tensor, eps: float, theta: float = 0.5, n_group: int = 32,
                 n_leaf: int = 8, precision: int = FP64, G: float = 1.0):
        if not eps > 0:
            raise ValueError("the adjoint needs a positive uniform softening")
        self.ctx = ctx
        self.m = m.contiguous()
        self.e2 = float(eps) ** 2
        self.theta, self.n_group, self.n_leaf = theta, n_group, n_leaf
        self.precision, self.G = precision, G

    def __call__(self, x, y, z):
        return self.with_eta(x, y, z)[:4]

    def with_eta(self, x, y, z):
        perm = morton_order_dev(x.detach(), y.detach(), z.detach())
        if bool((perm == torch.arange(perm.numel(), device=perm.device)).all()):
            out = _GravityFn.apply(x, y, z, perm, self)
            return tuple(o * self.G for o in out[:4]) + (out[4],)
        inv = torch.empty_like(perm)
        inv[perm] = torch.arange(perm.numel(), device=perm.device)
        out = _GravityFn.apply(x[perm], y[perm], z[perm], perm, self)
        return tuple(o[inv] * self.G for o in out[:4]) + (out[4][inv],)

    # forward on sorted particles; returns the outputs and what the backward
    # pass needs (the tree, its lists and monopoles)
    def _forward(self, x, y, z, perm):
        ms = self.m[perm].contiguous()
        e2 = torch.full_like(ms, self.e2)
        t = DeviceTree(x, y, z, ms, e2, self.n_leaf, self.n_group)
        st = _Saved(t, Walks(t.i_off, *t.long(self.theta)), t.moments(), ms, e2)
        self.ctx.set_params(uniform_softening=True)
        src = dict(m=ms, x=x, y=y, z=z, eps2=e2)
        out = {k: torch.empty_like(ms) for k in L.GRAV_OUT}
        self.ctx.gravity(st.walks, src, src, out, ps=st.mom, precision=self.precision)
        return (out["ax"], out["ay"], out["az"], out["pot"], out["eta"]), st

    def _backward(self, st, x, y, z, gax, gay, gaz, gpot, geta):
        w, mom, t = st.walks, st.mom, st.tree
        n = x.numel()
        g = [torch.zeros(n, dtype=torch.float64, device=x.device) for _ in range(3)]
        nb = [torch.zeros(t.n_node, dtype=torch.float64, device=x.device) for _ in range(3)]
        src = L.GravSrc(m=st.ms.data_ptr(), x=x.data_ptr(), y=y.data_ptr(), z=z.data_ptr(),
                        eps2=st.e2.data_ptr())
        ps = L.GravSrc(m=mom["m"].data_ptr(), x=mom["x"].data_ptr(), y=mom["y"].data_ptr(),
                       z=mom["z"].data_ptr(), eps2=mom["eps2"].data_ptr())
        cot = [c.contiguous() for c in (gax, gay, gaz, gpot, geta)]
        check(lib().sphk_gravity_vjp_dev(self.ctx.handle, w.n_walk, w.i_off.data_ptr(), w.j_off.data_ptr(),
                                         w.j_idx.data_ptr(), src, w.s_off.data_ptr(), w.s_idx.data_ptr(), ps,
                                         src, self.e2, *(c.data_ptr() for c in cot), *(a.data_ptr() for a in g),
                                         *(a.data_ptr() for a in nb), _stream()), "sphk_gravity_vjp_dev")
        check(lib().sphk_dtree_spread_dev(t._h, st.ms.data_ptr(), *(a.data_ptr() for a in nb),
                                          *(a.data_ptr() for a in g)), "sphk_dtree_spread_dev")
        return g


class _Saved:
    """Tree, lists and monopoles of one forward call (freed with it)."""

    def __init__(self, tree, walks, mom, ms, e2):
        self.tree, self.walks, self.mom, self.ms, self.e2 = tree, walks, mom, ms, e2


class _GravityFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, z, perm, op):
        x, y, z = x.contiguous(), y.contiguous(), z.contiguous()
        (ax, ay, az, pot, eta), st = op._forward(x, y, z, perm)
        ctx.op, ctx.st = op, st
        ctx.save_for_backward(x, y, z)
        return ax, ay, az, pot, eta

    @staticmethod
    def backward(ctx, gax, gay, gaz, gpot, geta):
        x, y, z = ctx.saved_tensors
        z0 = torch.zeros_like(x)
        gx, gy, gz = ctx.op._backward(ctx.st, x, y, z,
                                      *(z0 if c is None else c for c in (gax, gay, gaz, gpot, geta)))
        return gx, gy, gz, None, None


def potential_energy(grav: TreeGravity, m, x, y, z):
    """1/2 sum m_i pot_i (self terms included, as the functor's pot is)."""
    return 0.5 * (m * grav(x, y, z)[3]).sum()


def gravity_rollout(grav: TreeGravity, x, y, z, vx, vy, vz, dt: float, steps: int,
                    checkpoint_every: int = 20):
    """`steps` kick-drift-kick steps of pure gravity, differentiable with
    respect to the initial positions and velocities, recomputed per segment
    of `checkpoint_every` steps in the backward pass."""
    def segment(x, y, z, vx, vy, vz, n):
        ax, ay, az, _ = grav(x, y, z)
        for _ in range(n):
            vx, vy, vz = vx + 0.5 * dt * ax, vy + 0.5 * dt * ay, vz + 0.5 * dt * az
            x, y, z = x + dt * vx, y + dt * vy, z + dt * vz
            ax, ay, az, _ = grav(x, y, z)
            vx, vy, vz = vx + 0.5 * dt * ax, vy + 0.5 * dt * ay, vz + 0.5 * dt * az
        return x, y, z, vx, vy, vz

    state = (x, y, z, vx, vy, vz)
    done = 0
    while done < steps:
        n = min(checkpoint_every, steps - done)
        state = checkpoint(segment, *state, n, use_reentrant=False)
        done += n
    return state


def smoothed_unbound_mass(grav: TreeGravity, m, x, y, z, vx, vy, vz, k: float = 20.0, u=None):
    """sum m sigma(k e), e = |v - v_cg|^2 / 2 + pot (+ u): the differentiable
    surrogate of the unbound mass, with v_cg mass-weighted (no density in a
    pure-gravity system)."""
    M = m.sum()
    cx, cy, cz = ((m * v).sum() / M for v in (vx, vy, vz))
    pot = grav(x, y, z)[3]
    e = 0.5 * ((vx - cx) ** 2 + (vy - cy) ** 2 + (vz - cz) ** 2) + pot
    if u is not None:
        e = e + u
    return (m * torch.sigmoid(k * e)).sum()
