"""The reference's `run` program on the device path: read a run directory's
header and particle data, integrate with KDK to TimeEnd, and write time.log,
ASCII snapshots every dtAscii and hexadecimal dumps every dtHexa
(sph_yamauchi/run/main.cpp:400-460, hdr_hgas.hpp:1335-1539).

Units are the reference's code units (L = 1e-3 R_sun, M = 1e-6 M_sun,
T = 1 s, G = GravityConstantInThisUnit); the softening follows H (eps = H/kappa)
as in the reference.  Usage:

  python -m research_codes_amd.run <file.head> <file.data> <out_dir> [--steps K]
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

from . import formats
from .device import FP64, MIXED, Context
from .step import SPHSystem, phys


def run(head: str, data: str, out_dir: str, eta: float = 1.6, theta: float = 0.5, n_group: int = 32,
        precision: int = MIXED, max_steps: int | None = None, units=formats.REFERENCE_UNITS,
        log=print) -> dict:
    h = formats.read_head(head, units)
    p = formats.read_data(data, units)
    if len(p["m"]) != h["n"]:
        raise ValueError(f"{data}: {len(p['m'])} particles, header says {h['n']}")
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    ph = phys(G=units.gravity_constant, alpha_max=h["alpha_max"], alpha_min=h["alpha_min"],
              alphu_max=h["alphu_max"], alphu_min=h["alphu_min"], c_dt=h["c_dt"],
              dt_max=max(h["dt_ascii"], 1e-300))
    t_wall = time.perf_counter()
    with Context(0) as ctx:
        s = SPHSystem(p, ctx, eta=eta, eps=None, theta=theta, n_group=n_group, precision=precision,
                      phys_params=ph)
        s.time = h["time"]
        s.initialize()
        next_ascii, next_hexa = h["time"], h["time"]
        step = 0
        with open(out / "time.log", "w") as tl:
            tl.write(f"# Header Data:  {head}\n# AlphaMax:   {h['alpha_max']:+e}\n"
                     f"# AlphaMin:   {h['alpha_min']:+e}\n# C for Time: {h['c_dt']:+e}\n")
            while s.time < h["time_end"] and (max_steps is None or step < max_steps):
                if s.time >= next_ascii:
                    formats.write_snapshot(out / f"sph_t{step:06d}.dat", s.snapshot(), units)
                    next_ascii += h["dt_ascii"]
                if s.time >= next_hexa:
                    formats.write_hexa(out / f"sph_t{step:06d}.hex", s.snapshot())
                    next_hexa += h["dt_hexa"]
                t0 = s.time
                d = s.kdk()
                tl.write(formats.time_log_line(step, t0, s.dt, d["energy"], s.ctx.params["hmax"],
                                               time.perf_counter() - t_wall, units=units))
                tl.flush()
                step += 1
        final = s.diagnostics()
        formats.write_snapshot(out / f"sph_t{step:06d}.dat", s.snapshot(), units)
        log(f"[run] {step} steps to t={s.time:.6g}, E={final['energy']:+.9e}")
        return {"steps": step, "time": s.time, "energy": final["energy"]}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("head")
    ap.add_argument("data")
    ap.add_argument("out_dir")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--eta", type=float, default=1.6)
    ap.add_argument("--n-group", type=int, default=32)
    ap.add_argument("--fp64", action="store_true")
    a = ap.parse_args(argv)
    run(a.head, a.data, a.out_dir, eta=a.eta, n_group=a.n_group, precision=FP64 if a.fp64 else MIXED,
        max_steps=a.steps)
    return 0


if __name__ == "__main__":
    sys.exit(main())
