"""Writes tests/golden/rho_profile_n15_sun.json: the radial density profile
of the reference's relaxed n=1.5, 1 M_sun / 1 R_sun SPH star, binned from
the (r [cm], rho [g cm^-3]) pairs it ships as data
(plot_snap/making_slide/dens/hoge, 33 377 particles): per radial bin of
width 0.05 R_sun the particle count and the median radius and density."""
import json
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/plot_snap/making_slide/dens/hoge")
OUT = Path(__file__).with_name("rho_profile_n15_sun.json")
R_SUN = 6.9599e10   # cm, the reference's unit (hdr_heos.hpp:54)

d = np.loadtxt(SRC)
r = d[:, 0] / R_SUN
edges = np.linspace(0.0, 1.0, 21)
rows = []
for a, b in zip(edges[:-1], edges[1:]):
    s = (r >= a) & (r < b)
    if s.any():
        rows.append({"r_lo": round(float(a), 4), "r_hi": round(float(b), 4), "count": int(s.sum()),
                     "r_median": float(np.median(r[s])), "rho_median": float(np.median(d[s, 1]))})
OUT.write_text(json.dumps({"source": "plot_snap/making_slide/dens/hoge", "particles": int(len(d)),
                           "r_unit": "R_sun = 6.9599e10 cm", "rho_unit": "g cm^-3", "bins": rows},
                          indent=1) + "\n")
print(f"wrote {OUT} ({len(rows)} bins)")
