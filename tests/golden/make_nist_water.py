"""Write nist_water_23.5mpa.npy: the thermal-property table the reference
ships for its supercritical-water channel example
(example/THERMAL_PROPERTIES/NIST_WATER_23.5MP.DAT; columns P, H, T, D, M, K,
CP, B as read by buildup_property_relations_from_table,
src/input_thermo.f90:617-693), every 4th row plus the last one, as a
float64 array of shape (rows, 8).

    python tests/golden/make_nist_water.py    # needs the reference checkout
"""
import os

import numpy as np

SRC = "/root/reference/example/THERMAL_PROPERTIES/NIST_WATER_23.5MP.DAT"

if __name__ == "__main__":
    rows = []
    with open(SRC) as f:
        next(f)  # header line
        for line in f:
            parts = line.split()
            if len(parts) == 8:
                rows.append([float(v) for v in parts])
    tab = np.array(rows)
    keep = list(range(0, len(tab), 4))
    if keep[-1] != len(tab) - 1:
        keep.append(len(tab) - 1)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nist_water_23.5mpa.npy")
    np.save(out, tab[keep])
    print(out, tab[keep].shape)
