"""Writes ylm_scipy.json: spherical harmonics from scipy.special.sph_harm_y
(an implementation independent of both the oracle and the HIP kernels) at
seeded random directions, l <= 20 (the largest l of the gtinv table), m <= 0
(Condon-Shortley phase)."""
import json
import os

import numpy as np
from scipy.special import sph_harm_y

rng = np.random.default_rng(11)
dirs = rng.normal(size=(24, 3))
dirs /= np.linalg.norm(dirs, axis=1)[:, None]
LMAX = 20
out = {"lmax": LMAX, "dirs": dirs.tolist(), "ylm_re": [], "ylm_im": []}
for d in dirs:
    th = float(np.arccos(d[2]))
    ph = float(np.arctan2(d[1], d[0]))
    re, im = [], []
    for l in range(LMAX + 1):
        for m in range(-l, 1):
            v = complex(sph_harm_y(l, m, th, ph))
            re.append(v.real)
            im.append(v.imag)
    out["ylm_re"].append(re)
    out["ylm_im"].append(im)
json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ylm_scipy.json"), "w"))
