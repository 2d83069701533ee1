"""Fixtures of the reference's sample/no27.chemkin_reaction: its grid, its
converged cold-flow restart state and the transport-table entries of its
species.  Run from the repository root with the reference checked out at
/root/reference:

    python tests/golden/make_no27_state_fixture.py

Writes tests/golden/no27_state.npz:
  x, y      node coordinates of the one block (ideal_bl.x, 224 x 100 nodes)
  q         q(1:11) of the interior cells, shape (nj, ni, 11): the 8 species
            densities in the mechanism's order, rho u, rho v, E
  rgas, p   R and p of the interior cells from the same file (the T guess of
            calc_T is p / (rho R))
  step, tt  the header of restart.bin
and tests/golden/no27_trans.inp: the header line and the entries of
store/therm_lib/chemkin/core/trans.inp for the sample's species."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from shared_cfd_amd import fileio as F  # noqa: E402

REF = "/root/reference"
SAMPLE = os.path.join(REF, "sample", "no27.chemkin_reaction")
xs, ys = F.read_plot3d(os.path.join(SAMPLE, "ideal_bl.x"))
ni, nj = xs[0].shape[0] - 1, xs[0].shape[1] - 1
step, tt, blocks = F.read_restart(os.path.join(SAMPLE, "restart.bin"), [ni], [nj], 11)
a = blocks[0]
np.savez_compressed(os.path.join(HERE, "no27_state.npz"), x=xs[0], y=ys[0], q=a[1:-1, 1:-1, :11],
                    rgas=a[1:-1, 1:-1, 11], p=a[1:-1, 1:-1, 13], step=np.array(step), tt=np.array(tt))

species = ["N2", "O2", "H2", "OH", "O", "H", "HO2", "H2O"]
lines = open(os.path.join(REF, "store", "therm_lib", "chemkin", "core", "trans.inp")).read().splitlines()
out = [lines[0]]
k = 1
while k < len(lines):
    head = lines[k]
    body = []
    k += 1
    while k < len(lines) and lines[k][:2] in (" V", " C"):
        body.append(lines[k])
        k += 1
    name, partner = head[:16].strip(), head[16:34].strip()
    if name in species and not partner:
        out += [head] + body
with open(os.path.join(HERE, "no27_trans.inp"), "w") as f:
    f.write("\n".join(out) + "\n")
print(ni, nj, step, tt, len(out), "trans lines")
