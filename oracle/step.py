"""Whole-step restatement of the reference's SPH step on the CPU oracle.

TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and bench.py's
CPU-baseline leg may use it, and only as the checker.

  density_loop  calcDensityKernel, sph_yamauchi/run/hdr_density.hpp:369-415:
                rs = 1.1 ksr; repeat the density walk; a particle with
                rs != 0 whose H_3 > rs retries with rs *= 1.1, otherwise its
                results are copied (copyFromForce, hdr_density.hpp:24-31) and
                rs = 0; stop when no particle retried.
  force         calcSPHKernel, sph_yamauchi/run/main.cpp:293-358: density
                loop, EOS, Balsara, gravity + write-back, hydro + write-back,
                sumAcceleration, calcAlphaDot, then H_max =
                calcKernelSupportRadiusMaximum (hdr_util.hpp:489-505).
  kdk           loopSimulation's predict / force / correct,
                sph_yamauchi/run/hdr_hgas.hpp:1499-1539, with a given dt.

Lists.  FDPS builds the reference's lists and is not available, so this
module builds exact ones with scipy's k-d tree (independent of the product's
tree): one walk per particle; gather = j with |x_i - x_j| < rs_i; symmetry =
j with |x_i - x_j| < max(H_i, H_j); gravity = every particle (direct
summation, the tree's theta -> 0 limit).  A caller may pass its own gather
builder so that results that depend on list composition (the density passes
whose H exceeds rs) are compared on identical lists.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
from scipy.spatial import cKDTree

from . import (density, energy, eos_balsara, gravity, gravity_writeback, hydro, hydro_prep,
               hydro_writeback_tail, ksr_max, make_state, params, predict, correct)

KSRH = 1.936492
EXPAND = 1.1          # hdr_density.hpp:374


def _rows_to_walks(rows):
    """Per-particle j lists (a list of int arrays) as one walk per particle."""
    n = len(rows)
    cnt = np.fromiter((len(r) for r in rows), np.int64, n)
    j_off = np.zeros(n + 1, np.int64)
    np.cumsum(cnt, out=j_off[1:])
    j_idx = np.concatenate([np.asarray(r, np.int32) for r in rows]) if n else np.zeros(0, np.int32)
    return np.arange(n + 1, dtype=np.int32), j_off.astype(np.int32), j_idx.astype(np.int32)


def _pos(s):
    return np.stack([s["x"], s["y"], s["z"]], axis=1)


def exact_gather(s, rs):
    """One walk per particle, j with |x_i - x_j| < rs_i (ascending j)."""
    pos = _pos(s)
    tree = cKDTree(pos)
    rows = []
    for i, cand in enumerate(tree.query_ball_point(pos, r=rs, return_sorted=True)):
        c = np.asarray(cand, np.int64)
        d2 = ((pos[c] - pos[i]) ** 2).sum(axis=1)
        rows.append(c[d2 < rs[i] * rs[i]])
    return _rows_to_walks(rows)


def exact_symmetry(s, h):
    """One walk per particle, j with |x_i - x_j| < max(H_i, H_j) (ascending j)."""
    pos = _pos(s)
    n = len(pos)
    tree = cKDTree(pos)
    ii, jj = [], []
    for i, cand in enumerate(tree.query_ball_point(pos, r=h)):
        c = np.asarray(cand, np.int64)
        c = c[((pos[c] - pos[i]) ** 2).sum(axis=1) < h[i] * h[i]]
        ii.append(np.full(len(c), i))
        jj.append(c)
    ii, jj = np.concatenate(ii), np.concatenate(jj)
    a = sp.coo_matrix((np.ones(len(ii), np.int8), (ii, jj)), shape=(n, n)).tocsr()
    a = (a + a.T).tocsr()
    a.sort_indices()
    return (np.arange(n + 1, dtype=np.int32), a.indptr.astype(np.int32), a.indices.astype(np.int32))


def all_pairs(n, group=64):
    """Walks of `group` particles, each listing every particle."""
    i_off = np.append(np.arange(0, n, group), n).astype(np.int32)
    nw = len(i_off) - 1
    j_off = (np.arange(nw + 1) * n).astype(np.int32)
    return i_off, j_off, np.tile(np.arange(n, dtype=np.int32), nw)


def density_loop(s, ids, eta, hmax, gather=exact_gather, max_rounds=1000):
    """calcDensityKernel (hdr_density.hpp:369-415) on state s; returns
    (rounds, np) and writes dens, divv, rotv, grdh, ksr of each particle from
    the pass in which it converged.  gather(s, rs) -> (i_off, j_off, j_idx)
    over all particles (rs = 0 for converged ones)."""
    n = len(s["m"])
    rs = EXPAND * s["ksr"]
    npart = np.zeros(n, np.int64)
    for rnd in range(1, max_rounds + 1):
        walks = gather(s, rs)
        src = dict(id=ids, m=s["m"], x=s["x"], y=s["y"], z=s["z"], vx=s["vx"], vy=s["vy"],
                   vz=s["vz"], ksr=s["ksr"], rs=rs)
        out = density(*walks, src, src, eta=eta, ksrh=KSRH, hmax=hmax)
        active = rs != 0.0
        retry = active & (out["itr"] != 0)
        done = active & ~retry
        rs = np.where(retry, rs * EXPAND, np.where(done, 0.0, rs))
        for k in ("dens", "divv", "rotv", "grdh", "ksr"):
            s[k] = np.where(done, out[k], s[k])
        npart = np.where(done, out["np"], npart)
        if not retry.any():
            return rnd, npart
    raise RuntimeError(f"density loop did not converge in {max_rounds} rounds")


class OracleRun:
    """The reference's step on the oracle: state s (oracle.make_state
    fields), particle ids, eta, constant softening eps (None: eps = H/kappa,
    hdr_gravity.hpp:136), H_max (initially the caller's, then updated after
    each force evaluation as the reference does)."""

    def __init__(self, p, eta=1.18, eps=None, hmax=1e30, phys=None, gather=exact_gather):
        n = len(p["m"])
        self.s = make_state(n, **{k: p[k] for k in ("m", "x", "y", "z", "vx", "vy", "vz", "uene",
                                                    "ksr", "alph", "alphu")})
        self.ids = np.ascontiguousarray(p["id"], np.int32)
        self.istar = np.ascontiguousarray(p.get("istar", np.zeros(n)), np.int32)
        self.eta, self.eps, self.hmax = eta, eps, hmax
        self.prm = phys or params(eps_const=0.0 if eps is None else float(eps))
        self.gather = gather
        self.np = np.zeros(n, np.int64)
        self.rounds = 0

    def force(self):
        s, n = self.s, len(self.s["m"])
        self.rounds, self.np = density_loop(s, self.ids, self.eta, self.hmax, self.gather)
        eos_balsara(s, self.prm)
        eps2 = (s["ksr"] / KSRH) ** 2 if self.eps is None else np.full(n, float(self.eps) ** 2)
        gsrc = dict(m=s["m"], x=s["x"], y=s["y"], z=s["z"], eps2=eps2)
        g = gravity(*all_pairs(n), gsrc, gsrc)
        gravity_writeback(s, self.prm, g)
        hp = hydro_prep(s)
        hsrc = dict(id=self.ids, m=s["m"], x=s["x"], y=s["y"], z=s["z"], vx=s["vx"], vy=s["vy"],
                    vz=s["vz"], hinv=hp["hinv"], pres=s["pres"], presc=hp["presc"],
                    bswt=hp["bswt_half"], dens=s["dens"], vsnd=s["vsnd"], alph=s["alph"],
                    alphu=s["alphu"], thrm=hp["thrm"], eta=s["eta"])
        h = hydro(*exact_symmetry(s, s["ksr"]), hsrc, hsrc)
        hydro_writeback_tail(s, self.prm, h)
        self.hmax = ksr_max(s["m"], s["x"], s["y"], s["z"], s["dens"], self.istar, 0)

    def kdk(self, dt):
        predict(self.s, dt)
        self.force()
        correct(self.s, dt)

    def energy(self):
        return energy(self.s)
