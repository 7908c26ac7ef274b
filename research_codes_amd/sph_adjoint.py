"""Reverse-mode gradient of a fixed-step SPH collision run.

The forward run is the HIP path: SPHSystem.initialize, then KDK steps with a
fixed dt, with the device state checkpointed every `checkpoint_every` steps.
The backward pass re-runs each segment's HIP steps from its checkpoint
(bitwise the same steps), recording per step the Morton reorder, the H_max
and the density and hydro walk lists (SPHSystem.kdk(record=...)).  It then
walks the segment backwards: each step is restated as a differentiable FP64
function of its input state over exactly those lists and linearised at the
recorded state, and the cotangent is pulled back with torch.autograd.grad.
The restatement follows the functors and glue the HIP step runs:

* predict / correct (hdr_hgas.hpp:646-665);
* the density functor's three smoothing-length passes and the grad-h,
  divergence and curl pass (hdr_density.hpp:254-363) on the device, through
  its hand-written vector-Jacobian product (sphk_density_vjp_dev);
* EOS and Balsara switch (hdr_hgas.hpp:318-344, hdr_sph.hpp:79-89);
* tree gravity and its write-back (hdr_gravity.hpp:107-113, 295-338),
  through the hand-written vector-Jacobian product (adjoint.TreeGravity);
* the hydro functor (hdr_hydro.hpp:320-421) on the device, through its
  hand-written vector-Jacobian product (sphk_hydro_vjp_dev), its EPI
  preparation (hdr_hydro.hpp:50-60) and write-back with the acceleration sum
  and the viscosity switches' rates (hdr_hydro.hpp:21-25, hdr_sph.hpp:95-97,
  hdr_hgas.hpp:750-762).

As in any tree code's derivative, the tree, the lists and the reorder are
held fixed (piecewise constant in the state).  The SPH kernel vanishes with
continuous derivatives at the list edges, so list changes do not make the
outcome jump; gravity's multipole acceptance does (theta = 0 removes it).
The smoothing-length cap H_max is held fixed (it binds only where H reaches
it), and the timestep is fixed (the power-of-two rule makes the outcome
piecewise constant in dt).  Gravity needs a uniform softening.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib as L
from . import ic
from .adjoint import TreeGravity
from .device import FP64, Context, Walks
from .step import SPHSystem

C0 = 21.0 / (2.0 * math.pi)      # Wendland C2 normalisation, hdr_kernel.hpp:226
C1 = 4.0 * C0                    # hdr_kernel.hpp:227

# per-particle state carried from step to step (SPHSystem.s field names)
STATE = ("x", "y", "z", "vx", "vy", "vz", "uene", "alph", "alphu", "ax", "ay", "az", "udot", "adot",
         "adotu", "ksr", "pot", "dens")


# ----------------------------------------------------------------- kernels
def _w0(q):
    """W(q) = C0 (1-q)_+^4 (1+4q), hdr_kernel.hpp:246-250."""
    t = torch.clamp(1.0 - q, min=0.0)
    t2 = t * t
    return C0 * t2 * t2 * (1.0 + 4.0 * q)


def _w1(q):
    """W'(q) = C1 [(1-q)_+^4 - (1-q)_+^3 (1+4q)], hdr_kernel.hpp:251-257."""
    t = torch.clamp(1.0 - q, min=0.0)
    t2 = t * t
    return C1 * (t2 * t2 - t2 * t * (1.0 + 4.0 * q))


def _sqrt0(s):
    """sqrt with a zero derivative at s = 0 (the functors' 0 cases: self
    pairs, zero curl, equal pressures)."""
    pos = s > 0
    return torch.where(pos, torch.sqrt(torch.where(pos, s, torch.ones_like(s))), torch.zeros_like(s))


class _SegSum(torch.autograd.Function):
    """Per-particle sums of pair values whose particle indices are sorted
    (a segmented reduction: no atomics, unlike index_add over pairs that
    share their i); the gradient is the gather of the sums' cotangent."""

    @staticmethod
    def forward(ctx, v, idx, lengths):
        ctx.save_for_backward(idx)
        return torch.segment_reduce(v, "sum", lengths=lengths, unsafe=True, initial=0.0)

    @staticmethod
    def backward(ctx, g):
        (idx,) = ctx.saved_tensors
        return g[idx], None, None


class _GatherSorted(torch.autograd.Function):
    """x[idx] for sorted idx; the gradient is a segmented sum."""

    @staticmethod
    def forward(ctx, x, idx, lengths):
        ctx.save_for_backward(lengths)
        return x[idx]

    @staticmethod
    def backward(ctx, g):
        (lengths,) = ctx.saved_tensors
        return torch.segment_reduce(g, "sum", lengths=lengths, unsafe=True, initial=0.0), None, None


class _Gather(torch.autograd.Function):
    """x[idx]; the gradient is scattered with atomic adds (no sort)."""

    @staticmethod
    def forward(ctx, x, idx):
        ctx.save_for_backward(idx)
        ctx.n = x.numel()
        return x[idx]

    @staticmethod
    def backward(ctx, g):
        (idx,) = ctx.saved_tensors
        return torch.zeros(ctx.n, dtype=g.dtype, device=g.device).index_add_(0, idx, g), None


class _Seg:
    """A pair set (i, j) sorted by i: per-i sums and the gathers of the two
    sides, with gradients that avoid sorting the indices."""

    def __init__(self, i, j, n: int):
        if i.numel() > 1 and not bool((i[1:] >= i[:-1]).all()):
            raise ValueError("pairs must be sorted by their i index")
        self.i, self.j = i, j
        # pairs per i from the sorted i (cheaper than a histogram)
        bounds = torch.searchsorted(i, torch.arange(n + 1, device=i.device, dtype=i.dtype))
        self.lengths = bounds[1:] - bounds[:-1]

    def sum(self, v):
        return _SegSum.apply(v, self.i, self.lengths)

    def gi(self, x):
        return _GatherSorted.apply(x, self.i, self.lengths) if x.requires_grad else x[self.i]

    def gj(self, x):
        return _Gather.apply(x, self.j) if x.requires_grad else x[self.j]


def expand_pairs(w: Walks) -> tuple[torch.Tensor, torch.Tensor]:
    """(i, j) of every entry of every walk's list: i over the walk's
    particles, j over its list (int64 device tensors, walk order)."""
    io = w.i_off.long()
    jo = w.j_off.long()
    ni = io[1:] - io[:-1]
    nj = jo[1:] - jo[:-1]
    cnt = ni * nj
    total = int(cnt.sum())
    dev = io.device
    if total == 0:
        e = torch.zeros(0, dtype=torch.int64, device=dev)
        return e, e
    wk = torch.repeat_interleave(torch.arange(w.n_walk, device=dev), cnt)
    start = torch.cumsum(cnt, 0) - cnt
    loc = torch.arange(total, device=dev) - start[wk]
    njw = nj[wk]
    i = io[wk] + torch.div(loc, njw, rounding_mode="floor")
    pos = jo[wk] + loc % njw
    j = pos if w.j_idx is None else w.j_idx.long()[pos]
    return i, j


# --------------------------------------------------------- frozen gravity
class FrozenGravity:
    """Tree gravity over a recorded tree and long lists (one KDK step's, from
    SPHSystem.kdk(record=...)): the node membership, the EPJ/SPJ lists and
    the node masses stay those of the record; the nodes' centres of mass are
    recomputed from the positions it is called with, X_n = sum_{k in n} m_k
    x_k / M_n.  This is the tree held fixed, the piecewise-constant branch of
    tree gravity the adjoint differentiates, as a function that can also be
    evaluated away from the recorded state (finite differences at fixed
    topology).  Uniform softening eps2; outputs a and pot times G."""

    def __init__(self, ctx: Context, m, eps2: float, walks: Walks, begin, end, G: float):
        self.ctx, self.m, self.e2, self.walks, self.G = ctx, m.contiguous(), float(eps2), walks, G
        lengths = (end - begin).long()
        n_node = lengths.numel()
        node_of = torch.repeat_interleave(torch.arange(n_node, device=m.device), lengths)
        start = torch.cumsum(lengths, 0) - lengths
        self.idx = begin.long()[node_of] + (torch.arange(node_of.numel(), device=m.device) - start[node_of])
        self.node_of, self.lengths = node_of, lengths
        self.M = torch.segment_reduce(self.m[self.idx], "sum", lengths=lengths, unsafe=True, initial=0.0)
        self.e2n = torch.full_like(self.M, self.e2)
        self.e2p = torch.full_like(self.m, self.e2)

    def centres(self, x, y, z):
        return [_SegSum.apply(_Gather.apply(self.m * c, self.idx), self.node_of, self.lengths) / self.M
                for c in (x, y, z)]

    def __call__(self, x, y, z):
        X, Y, Z = self.centres(x, y, z)
        out = _FrozenGravityFn.apply(x, y, z, X, Y, Z, self)
        return tuple(o * self.G for o in out)


class _FrozenGravityFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, z, X, Y, Z, op):
        x, y, z, X, Y, Z = (t.contiguous() for t in (x, y, z, X, Y, Z))
        src = dict(m=op.m, x=x, y=y, z=z, eps2=op.e2p)
        ps = dict(m=op.M, x=X, y=Y, z=Z, eps2=op.e2n)
        out = {k: torch.empty_like(op.m) for k in L.GRAV_OUT}
        keep = op.ctx.params["uniform_softening"]
        op.ctx.set_params(uniform_softening=True)
        op.ctx.gravity(op.walks, src, src, out, ps=ps, precision=FP64)
        op.ctx.set_params(uniform_softening=keep)
        ctx.op = op
        ctx.save_for_backward(x, y, z, X, Y, Z)
        return out["ax"], out["ay"], out["az"], out["pot"]

    @staticmethod
    def backward(ctx, gax, gay, gaz, gpot):
        op = ctx.op
        x, y, z, X, Y, Z = ctx.saved_tensors
        w = op.walks
        g = [torch.zeros_like(x) for _ in range(3)]
        nb = [torch.zeros_like(X) for _ in range(3)]
        src = L.GravSrc(m=op.m.data_ptr(), x=x.data_ptr(), y=y.data_ptr(), z=z.data_ptr(),
                        eps2=op.e2p.data_ptr())
        ps = L.GravSrc(m=op.M.data_ptr(), x=X.data_ptr(), y=Y.data_ptr(), z=Z.data_ptr(),
                       eps2=op.e2n.data_ptr())
        z0 = torch.zeros_like(x)
        cot = [(z0 if c is None else c).contiguous() for c in (gax, gay, gaz, gpot)]
        L.check(L.lib().sphk_gravity_vjp_dev(op.ctx.handle, w.n_walk, w.i_off.data_ptr(), w.j_off.data_ptr(),
                                             w.j_idx.data_ptr(), src, w.s_off.data_ptr(), w.s_idx.data_ptr(),
                                             ps, src, op.e2, *(c.data_ptr() for c in cot), None,
                                             *(a.data_ptr() for a in g), *(a.data_ptr() for a in nb),
                                             torch.cuda.current_stream().cuda_stream), "sphk_gravity_vjp_dev")
        return (*g, *nb, None)


# ------------------------------------------------------------------ force
def _density_t(P, m, pairs, hmax, eta, ksrh):
    """The density functor on pairs (i, j): returns dens, ksr (= H_3), divv,
    rotv and grdh (= Omega^-1), differentiable in the positions, velocities
    and the starting H (P["ksr"])."""
    i, j = pairs
    seg = _Seg(i, j, m.numel())
    dx, dy, dz = (seg.gi(P[c]) - seg.gj(P[c]) for c in ("x", "y", "z"))
    r2 = dx * dx + dy * dy + dz * dz
    r = _sqrt0(r2)
    mj = m[j]
    H = P["ksr"]
    dens = None
    for _ in range(3):                                  # hdr_density.hpp:264-302
        hi = 1.0 / H
        dens = seg.sum(mj * _w0(r * seg.gi(hi))) * (hi * hi * hi)
        H = torch.clamp(eta * ksrh * (m / dens) ** (1.0 / 3.0), max=hmax)
    hi = 1.0 / H
    hi4 = hi * hi * hi * hi
    self_ = i == j                                      # ids are unique: self = same index
    ri = torch.where(self_, torch.zeros_like(r), 1.0 / torch.where(self_, torch.ones_like(r), r))
    q = r2 * ri * seg.gi(hi)
    kw0, kw1 = _w0(q), _w1(q)
    grdh = -seg.sum((3.0 * kw0 + q * kw1) * mj) * hi4
    dw = mj * seg.gi(hi4) * kw1 * ri
    dvx, dvy, dvz = (seg.gi(P[c]) - seg.gj(P[c]) for c in ("vx", "vy", "vz"))
    divv = -seg.sum((dvx * dx + dvy * dy + dvz * dz) * dw)
    rx = seg.sum((dvy * dz - dvz * dy) * dw)
    ry = seg.sum((dvz * dx - dvx * dz) * dw)
    rz = seg.sum((dvx * dy - dvy * dx) * dw)
    omg = 1.0 / (1.0 + H / dens * grdh / 3.0)           # hdr_density.hpp:351-362
    rot = _sqrt0(rx * rx + ry * ry + rz * rz)
    return dens, H, divv / dens * omg, rot / dens * omg, omg


def _hydro_t(P, m, pairs, d, G):
    """The hydro functor on non-self pairs (i, j) with r < max(H_i, H_j), in
    torch (the autograd reference the device VJP is tested against)."""
    i, j = pairs
    seg = _Seg(i, j, m.numel())
    gi, gj = seg.gi, seg.gj
    dx, dy, dz = (gi(P[c]) - gj(P[c]) for c in ("x", "y", "z"))
    dvx, dvy, dvz = (gi(P[c]) - gj(P[c]) for c in ("vx", "vy", "vz"))
    r = torch.sqrt(dx * dx + dy * dy + dz * dz)
    ri = 1.0 / r
    hi = d["hinv"]
    hi4 = hi * hi * hi * hi
    dwi = gi(hi4) * _w1(r * gi(hi))
    dwj = gj(hi4) * _w1(r * gj(hi))
    mj = m[j]
    ka = (dwi + dwj) * mj
    wij = (dx * dvx + dy * dvy + dz * dvz) * ri
    w0 = torch.clamp(wij, max=0.0)
    vs = gi(d["vsnd"]) + gj(d["vsnd"]) - 3.0 * w0
    rhi = 1.0 / (gi(d["dens"]) + gj(d["dens"]))
    av0 = (gi(d["bswt_half"]) + gj(d["bswt_half"])) * (gi(P["alph"]) + gj(P["alph"])) * vs * w0
    pci = gi(d["presc"])
    ta = (pci + gj(d["presc"]) - 0.5 * av0 * rhi) * ka * ri
    vsu = _sqrt0(torch.abs(gi(d["pres"]) - gj(d["pres"])) * rhi * 2.0)
    du = gi(d["thrm"]) - gj(d["thrm"])
    udot = seg.sum(ka * (pci * wij - rhi * (0.25 * av0 * w0 - (gi(P["alphu"]) + gj(P["alphu"])) * vsu * du)))
    dg = mj * ri * (gi(d["eta"]) * dwi + gj(d["eta"]) * dwj)
    out = {"achx": -0.5 * seg.sum(ta * dx), "achy": -0.5 * seg.sum(ta * dy),
           "achz": -0.5 * seg.sum(ta * dz), "udot": 0.5 * udot,
           "diffu": 2.0 * seg.sum(rhi * du * ka * ri)}
    for c, dc in (("x", dx), ("y", dy), ("z", dz)):
        out["g1" + c] = G * 0.5 * seg.sum(dg * dc)
    return out


_HYDRO_OUTS = ("ax", "ay", "az", "udot", "diffu", "gx", "gy", "gz")


class _HydroFn(torch.autograd.Function):
    """The FP64 hydro functor on the device (sphk_hydro_dev) over symmetric
    lists, differentiated by its hand-written vector-Jacobian product
    (sphk_hydro_vjp_dev).  Inputs: the HYDRO_GRAD fields in that order."""

    @staticmethod
    def forward(ctx, walks, dctx, m, ids, *vals):
        src = dict(zip(L.HYDRO_GRAD, (v.contiguous() for v in vals)))
        src.update(id=ids, m=m)
        out = {k: torch.empty_like(m) for k in L.HYDRO_OUT}
        dctx.hydro(walks, src, src, out, FP64)
        ctx.save_for_backward(*(src[k] for k in L.HYDRO_GRAD))
        ctx.walks, ctx.dctx, ctx.m, ctx.ids = walks, dctx, m, ids
        return tuple(out[k] for k in _HYDRO_OUTS)

    @staticmethod
    def backward(ctx, *gouts):
        src = dict(zip(L.HYDRO_GRAD, ctx.saved_tensors))
        src.update(id=ctx.ids, m=ctx.m)
        cot = {k: torch.zeros_like(ctx.m) if g is None else g.contiguous() for k, g in zip(_HYDRO_OUTS, gouts)}
        grad = {k: torch.zeros_like(ctx.m) for k in L.HYDRO_GRAD}
        ctx.dctx.hydro_vjp(ctx.walks, src, cot, grad)
        return (None, None, None, None) + tuple(grad[k] for k in L.HYDRO_GRAD)


def hydro_dev(walks, dctx, m, ids, inputs: dict) -> dict:
    """Differentiable hydro functor outputs (ax, ay, az, udot, diffu, gx,
    gy, gz; g without the gravitational constant) from the HYDRO_GRAD inputs."""
    out = _HydroFn.apply(walks, dctx, m.detach(), ids, *(inputs[k] for k in L.HYDRO_GRAD))
    return dict(zip(_HYDRO_OUTS, out))


_DENS_IN = ("x", "y", "z", "vx", "vy", "vz", "ksr")


class _DensFn(torch.autograd.Function):
    """The FP64 density functor on the device (sphk_density_dev) over gather
    lists, at the smoothing-length cap hmax, differentiated by its
    hand-written vector-Jacobian product (sphk_density_vjp_dev).  Inputs:
    positions, velocities and the starting H (_DENS_IN order); outputs dens,
    ksr (= H_3), divv, rotv, grdh."""

    @staticmethod
    def _run(dctx, hmax, f):
        keep = dctx.params["hmax"]
        dctx.set_params(hmax=hmax)
        try:
            f()
        finally:
            dctx.set_params(hmax=keep)

    @staticmethod
    def forward(ctx, walks, dctx, hmax, m, ids, *vals):
        src = dict(zip(_DENS_IN, (v.contiguous() for v in vals)))
        src.update(id=ids, m=m, rs=src["ksr"])
        out = {k: torch.empty_like(m) for k in ("dens", "divv", "rotv", "grdh", "ksr")}
        out["np"] = torch.empty(m.numel(), dtype=torch.int64, device=m.device)
        out["itr"] = torch.empty(m.numel(), dtype=torch.uint8, device=m.device)
        _DensFn._run(dctx, hmax, lambda: dctx.density(walks, src, src, out, FP64))
        ctx.save_for_backward(*(src[k] for k in _DENS_IN))
        ctx.walks, ctx.dctx, ctx.hmax, ctx.m, ctx.ids = walks, dctx, hmax, m, ids
        return tuple(out[k] for k in L.DENS_COT)

    @staticmethod
    def backward(ctx, *gouts):
        src = dict(zip(_DENS_IN, ctx.saved_tensors))
        src.update(id=ctx.ids, m=ctx.m, rs=src["ksr"])
        cot = {k: torch.zeros_like(ctx.m) if g is None else g.contiguous() for k, g in zip(L.DENS_COT, gouts)}
        grad = {k: torch.zeros_like(ctx.m) for k in L.DENS_GRAD}
        _DensFn._run(ctx.dctx, ctx.hmax, lambda: ctx.dctx.density_vjp(ctx.walks, src, cot, grad))
        return (None,) * 5 + tuple(grad[k] for k in _DENS_IN)


def density_dev(walks, dctx, hmax, m, ids, inputs: dict):
    """Differentiable density functor outputs (dens, ksr, divv, rotv, grdh)
    from positions, velocities and the starting H (inputs["ksr"])."""
    return _DensFn.apply(walks, dctx, float(hmax), m.detach(), ids, *(inputs[k] for k in _DENS_IN))


def force_t(P: dict, m, rec: dict, sysm: SPHSystem) -> dict:
    """One force evaluation (density, EOS + Balsara, gravity + write-back,
    hydro + write-back + rates) on the recorded lists, differentiable in
    the positions, velocities, u, alpha, alpha_u and the starting H of P."""
    ph = sysm.phys
    ksrh = ph.ksrh
    # density: the device functor and its hand-written VJP over the recorded
    # gather lists, at the step's smoothing-length cap; each particle takes
    # the results of the retry round in which it converged (over that
    # round's lists, as calcDensityKernel copies them)
    rounds = rec.get("dens_rounds") or [(rec["dens"], None)]
    res = None
    for walks, took in rounds:
        r = density_dev(walks, sysm.ctx, rec["hmax"], m, rec["id"], P)
        res = r if res is None else tuple(torch.where(took, a, b) for a, b in zip(r, res))
    dens, H, divv, rotv, omg = res
    # EOS and Balsara switch
    u = P["uene"]
    pres = ph.ninv * dens * u
    vsnd = _sqrt0((1.0 + ph.ninv) * pres / dens)
    ad = torch.abs(divv)
    bswt = ad / (ad + torch.abs(rotv) + 1e-4 * vsnd * ksrh / H)
    # gravity (uniform softening) and its write-back: over the recorded tree
    # and long lists when the record holds them, else a tree of its own
    if "grav" in rec:
        fg = FrozenGravity(sysm.ctx, m.detach(), sysm.eps ** 2, rec["grav"], rec["gnode_begin"],
                           rec["gnode_end"], ph.G)
        g2x, g2y, g2z, pot = fg(P["x"], P["y"], P["z"])
    else:
        tg = TreeGravity(sysm.ctx, m.detach(), eps=sysm.eps, theta=sysm.theta, n_group=sysm.n_group["grav"],
                         n_leaf=sysm.n_leaf, precision=FP64, G=ph.G)
        g2x, g2y, g2z, pot, _ = tg.with_eta(P["x"], P["y"], P["z"])
    # constant softening: the potential does not depend on H, so the grad-h
    # softening term eta of the write-back (hdr_gravity.hpp:110-111) is 0
    eta = torch.zeros_like(dens)
    # hydro: the device functor and its hand-written VJP, on the EPI
    # preparation of hdr_hydro.hpp:50-60
    hin = {"x": P["x"], "y": P["y"], "z": P["z"], "vx": P["vx"], "vy": P["vy"], "vz": P["vz"],
           "hinv": 1.0 / H, "pres": pres, "presc": pres * omg / (dens * dens), "bswt": 0.5 * bswt,
           "dens": dens, "vsnd": vsnd, "alph": P["alph"], "alphu": P["alphu"],
           "thrm": torch.clamp(u, min=0.0), "eta": eta}
    h = hydro_dev(rec["hydro"], sysm.ctx, m, rec["id"], hin)
    # write-back: acceleration sum and the viscosity switches' rates
    out = {"dens": dens, "ksr": H, "pot": pot, "udot": h["udot"]}
    out["ax"] = h["ax"] + ph.G * h["gx"] + g2x
    out["ay"] = h["ay"] + ph.G * h["gy"] + g2y
    out["az"] = h["az"] + ph.G * h["gz"] + g2z
    alph, alphu = P["alph"], P["alphu"]
    src = torch.clamp(-divv * (ph.alpha_max - alph), min=0.0)
    tauinv = 0.25 * ksrh * vsnd / H
    out["adot"] = -(alph - ph.alpha_min) * tauinv + src
    lim = torch.clamp(ph.alphu_max - alphu, min=0.0)
    on = lim > 0
    srcu = torch.where(on, H / ksrh * torch.abs(h["diffu"])
                       / torch.sqrt(torch.clamp(u, min=0.0) + ph.eps_u + (~on).double()) * lim,
                       torch.zeros_like(lim))
    out["adotu"] = -(alphu - ph.alphu_min) * tauinv + srcu
    return out


def init_t(S: dict, m, rec: dict, sysm: SPHSystem) -> dict:
    """The initial force evaluation (SPHSystem.initialize) as a function of
    the initial state."""
    f = force_t(S, m, rec, sysm)
    out = {k: S[k] for k in ("x", "y", "z", "vx", "vy", "vz", "uene", "alph", "alphu")}
    out.update(f)
    return out


def step_t(S: dict, m, rec: dict, sysm: SPHSystem, dt: float) -> dict:
    """One KDK step (SPHSystem.kdk with a fixed dt) as a function of the
    state at its start; m is in that state's order."""
    P = {}
    for c, v in (("x", "vx"), ("y", "vy"), ("z", "vz")):
        a = "a" + c
        P[c] = S[c] + S[v] * dt + 0.5 * S[a] * dt * dt          # hdr_hgas.hpp:646-657
        P[v] = S[v] + S[a] * dt
        P[v + "2"] = S[v] + 0.5 * S[a] * dt
    for k, kd in (("uene", "udot"), ("alph", "adot"), ("alphu", "adotu")):
        P[k] = S[k] + S[kd] * dt
        P[k + "2"] = S[k] + 0.5 * S[kd] * dt
    P["ksr"] = S["ksr"]
    perm = rec["perm"]
    if perm is not None:
        P = {k: v[perm] for k, v in P.items()}
        m = m[perm]
    f = force_t(P, m, rec, sysm)
    out = {"x": P["x"], "y": P["y"], "z": P["z"]}
    for v, a in (("vx", "ax"), ("vy", "ay"), ("vz", "az")):
        out[v] = P[v + "2"] + 0.5 * f[a] * dt                   # hdr_hgas.hpp:659-665
    for k, kd in (("uene", "udot"), ("alph", "adot"), ("alphu", "adotu")):
        out[k] = P[k + "2"] + 0.5 * f[kd] * dt
    out.update(f)
    return out


def smoothed_unbound_t(S: dict, m, k: float):
    """sum m sigma(k e), e = |v - v_cg|^2 / 2 + pot + u, v_cg density-weighted
    (analysis/unbound_mass/unbound_mass.py:10-24, analysis/common.py:71-90,
    the smoothed form of sphk_unbound_mass_dev over all particles)."""
    w = S["dens"]
    W = w.sum()
    c = [(w * S[v]).sum() / W for v in ("vx", "vy", "vz")]
    e = 0.5 * sum((S[v] - cv) ** 2 for v, cv in zip(("vx", "vy", "vz"), c)) + S["pot"] + S["uene"]
    return (m * torch.sigmoid(k * e)).sum()


# ------------------------------------------------------------ checkpoints
_EXTRA = ("id", "istar", "rs", "np_", "itr", "eps2")
# the initial state's differentiable fields
_INIT = ("x", "y", "z", "vx", "vy", "vz", "uene", "alph", "alphu", "ksr")


def _save(sysm: SPHSystem) -> dict:
    ck = {"s": {k: v.clone() for k, v in sysm.s.items()},
          "extra": {k: getattr(sysm, k).clone() for k in _EXTRA},
          "hmax": sysm.hmax, "time": sysm.time, "dt": sysm.dt}
    return ck


def _restore(sysm: SPHSystem, ck: dict):
    for k, v in ck["s"].items():
        sysm.s[k].copy_(v)
    for k, v in ck["extra"].items():
        getattr(sysm, k).copy_(v)
    sysm.hmax = ck["hmax"]
    sysm.ctx.set_params(hmax=ck["hmax"])
    sysm.time, sysm.dt = ck["time"], ck["dt"]


def _state(sysm: SPHSystem) -> dict:
    return {k: sysm.s[k].clone() for k in STATE}


def _pull(out: dict, inp: dict, cot: dict) -> dict:
    keys = [k for k in out if k in cot and out[k].requires_grad]
    g = torch.autograd.grad([out[k] for k in keys], list(inp.values()), [cot[k] for k in keys],
                            allow_unused=True)
    return {k: (torch.zeros_like(v) if gk is None else gk) for (k, v), gk in zip(inp.items(), g)}


class CollisionRun:
    """An SPH run on the HIP path and its reverse-mode gradient.

    `forward(steps)` runs SPHSystem.initialize and `steps` KDK steps,
    checkpointing the state every `checkpoint_every` steps, and returns the
    smoothed unbound mass of the final state.  The steps have length dt, or
    with dt=None the power-of-two rule's lengths, which the gradient then
    holds fixed (kept in `dts`).  `backward()` returns the gradient of the
    loss with respect to the initial state (a dict of per-particle tensors
    per field, in the order the system holds its particles at construction).
    The system needs a uniform gravitational softening.  Its precision is
    that of the forward run and of the backward pass's replay (SPHSystem
    precision; MIXED: FP32 pair terms, FP64 accumulation); the restated step
    the gradient is taken through is FP64, linearised at the run's states.
    `frozen_losses(deltas)` evaluates the loss with every discrete choice of
    this run held (tree, lists, reorders, H_max, step lengths) for perturbed
    initial states: the function whose derivative `backward()` is."""

    def __init__(self, sysm: SPHSystem, dt: float | None, checkpoint_every: int = 20, k: float = 20.0):
        if sysm.eps is None:
            raise ValueError("the gravity adjoint needs a uniform softening (eps)")
        if len(sysm.seg) != 2:
            raise ValueError("one run per system: differentiate the runs of an ensemble one by one")
        if sysm.tree != "device":
            raise ValueError("the gravity adjoint walks the device tree")
        self.sysm, self.every, self.k = sysm, int(checkpoint_every), float(k)
        self.dt = None if dt is None else float(dt)

    def forward(self, steps: int) -> float:
        s = self.sysm
        self.steps = int(steps)
        self.m0 = s.s["m"].clone()
        self.istar0 = s.istar.clone()
        self.init_state = _state(s)
        self.init_rec = {}
        s.initialize(record=self.init_rec)
        self.cks = {}
        self.dts = []
        for n in range(self.steps):
            if n % self.every == 0:
                self.cks[n] = _save(s)
            s.kdk(dt=self.dt)
            self.dts.append(s.dt)
        self.final = _state(s)
        self.m_final = s.s["m"].clone()
        return float(smoothed_unbound_t(self.final, self.m_final, self.k))

    def backward(self) -> dict:
        s = self.sysm
        S = {k: v.clone().requires_grad_(True) for k, v in self.final.items()}
        loss = smoothed_unbound_t(S, self.m_final, self.k)
        cot = dict(zip(S, torch.autograd.grad(loss, list(S.values()), allow_unused=True)))
        cot = {k: torch.zeros_like(S[k]) if v is None else v for k, v in cot.items()}
        for n0 in sorted(self.cks, reverse=True):
            n1 = min(n0 + self.every, self.steps)
            _restore(s, self.cks[n0])
            seg = []
            for n in range(n0, n1):
                st, m, rec = _state(s), s.s["m"].clone(), {}
                s.kdk(dt=self.dts[n], record=rec)
                seg.append((st, m, rec, self.dts[n]))
            for st, m, rec, dt in reversed(seg):
                inp = {k: v.clone().requires_grad_(True) for k, v in st.items()}
                out = step_t(inp, m, rec, s, dt)
                cot = _pull(out, inp, cot)
                del out, inp
        inp = {k: self.init_state[k].clone().requires_grad_(True) for k in _INIT}
        out = init_t(inp, self.m0, self.init_rec, s)
        return _pull(out, inp, cot)

    @torch.no_grad()
    def frozen_losses(self, deltas: list) -> list:
        """The loss after the run's steps, with this run's tree, lists,
        reorders, H_max and step lengths held (re-recorded by replaying its
        HIP steps segment by segment), for initial states perturbed by each
        delta (a dict of per-particle tensors over the _INIT fields, in the
        construction order).  The perturbed states go through the restated
        FP64 step."""
        s = self.sysm
        states = []
        for d in deltas:
            S0 = {k: self.init_state[k] + d[k] if k in d else self.init_state[k] for k in _INIT}
            f = init_t(S0, self.m0, self.init_rec, s)
            states.append({k: f[k] for k in STATE})
        for n0 in sorted(self.cks):
            n1 = min(n0 + self.every, self.steps)
            _restore(s, self.cks[n0])
            for n in range(n0, n1):
                m, rec = s.s["m"].clone(), {}
                s.kdk(dt=self.dts[n], record=rec)
                states = [{k: v for k, v in step_t(S, m, rec, s, self.dts[n]).items() if k in STATE}
                          for S in states]
        return [float(smoothed_unbound_t(S, self.m_final, self.k)) for S in states]


def head_on_gradient(n_each: int = 100_000, steps: int = 200, dt: float | None = 1e-3, k: float = 20.0,
                     eta: float = ic.ETA_50, eps: float = 0.01, theta: float = 0.5, checkpoint_every: int = 20,
                     n_group: int = 32, separation: float = 3.0) -> tuple[float, np.ndarray, list]:
    """Smoothed unbound mass after `steps` KDK steps of the head-on
    collision of two n=1.5 stars (research_codes_amd.sensitivity.head_on_loss
    at dx0 = dv0 = 0, FP64), its gradient with respect to the impactor's
    offset dx0 and velocity change dv0 along the collision axis, and the
    step lengths (dt, or the power-of-two rule's with dt=None)."""
    p = ic.head_on(n_each, eta=eta, separation=separation)
    with Context(0) as c:
        s = SPHSystem(p, c, eta=eta, eps=eps, theta=theta, precision=FP64, n_group=n_group)
        run = CollisionRun(s, dt, checkpoint_every, k)
        loss = run.forward(steps)
        g = run.backward()
        imp = (run.istar0 == 1).double()
        grad = np.array([float((g["x"] * imp).sum()), float((g["vx"] * imp).sum())])
    return loss, grad, run.dts
