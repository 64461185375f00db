"""Writes tests/golden/flat_lcdm_distances.json: comoving distances of flat
LCDM (Om=0.3, OL=0.7, h=0.7) from scipy.integrate.quad at 1e-13 relative
tolerance, an independent check of both distance integrators."""
import json
import os

import numpy as np
from scipy.integrate import quad

Om, OL, h = 0.3, 0.7, 0.7
DH = 2997.92458 / h
pairs = [(0.0, z) for z in (0.05, 0.1, 0.3, 0.5, 0.7, 1.0, 1.5, 2.0, 2.5, 3.0, 4.0)]
pairs += [(0.5, 2.0), (1.0, 3.0), (0.1, 2.0), (2.0, 3.0), (0.2340134382247925, 1.123)]
rows = []
for z1, z2 in pairs:
    v, err = quad(lambda z: 1.0 / np.sqrt(Om * (1 + z) ** 3 + OL), z1, z2,
                  epsabs=0, epsrel=1e-13, limit=200)
    rows.append({"z1": z1, "z2": z2, "Dc": DH * v})
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "flat_lcdm_distances.json")
with open(out, "w") as f:
    json.dump({"Om": Om, "OL": OL, "h": h, "rows": rows}, f, indent=1)
print(out)
