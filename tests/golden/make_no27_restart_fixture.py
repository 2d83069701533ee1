"""Extract a fixture from the reference's sample/no27.chemkin_reaction
restart.bin (the state the reference wrote after 40000 steps of its
reacting boundary-layer sample, air at 300 K and 1 bar): every 14th cell of
every 11th row, ghost layer included, with step and tt.  Run from the
repository root with the reference checked out at /root/reference:

    python tests/golden/make_no27_restart_fixture.py

The grid dimensions come from the sample's Plot3D file; the species count
(8) from its chem.inp (tests/golden/no27_chem.inp)."""
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from shared_cfd_amd import fileio as F  # noqa: E402

SAMPLE = "/root/reference/sample/no27.chemkin_reaction"
xs, _ = F.read_plot3d(os.path.join(SAMPLE, "ideal_bl.x"))
ni, nj = xs[0].shape[0] - 1, xs[0].shape[1] - 1
step, tt, blocks = F.read_restart(os.path.join(SAMPLE, "restart.bin"), [ni], [nj], 11)
a = blocks[0]
cells = [[j, i, [float(v) for v in a[j, i]]] for j in range(0, nj + 2, 11) for i in range(0, ni + 2, 14)]
out = {"source": "sample/no27.chemkin_reaction/restart.bin", "ni": ni, "nj": nj, "dimq": 11, "step": step,
       "tt": tt, "layout": "q(1:8 species, rho u, rho v, E), R, gamma, p at (j, i), i = 0..ni+1, j = 0..nj+1",
       "cells": cells}
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "no27_restart_cells.json"), "w") as f:
    json.dump(out, f, indent=0)
print(len(cells), "cells")
