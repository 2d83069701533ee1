"""On-disk formats of shared_cfd and block-interface detection.

* Plot3D grids as the reference reads and writes them (ASCII: the number of
  planes, then ``ni+1 nj+1 1`` per plane, then per plane the x, y and z node
  values, four per line, i fastest): store/dim/2d/geometry.f90:27-78
  (init_xrh), store/cond/core/read_plot3d.py:7-70.
* Fortran sequential unformatted files in the reference's byte order
  (big-endian, ``-fconvert=big-endian`` of store/architecture/PC/
  Makefile.head:4), with the gfortran record markers (4-byte lengths; a
  negative length marks a record continued in the next sub-record).
* ``restart.bin`` / ``result*.bin`` (step, tt, then q(1:dimq), R, gamma, p
  for i = 0..ni+1, j = 0..nj+1 of every plane): store/core/inout.f90:1-142.
* ``geometry.bin`` (per plane ni, nj and the node pairs (xh, rh) for
  i = 0..ni, j = 0..nj): store/dim/2d/geometry.f90:182-193.
* Block-interface detection: boundary node lines of all planes that share
  consecutive coincident nodes become rows of the 11-integer cut table the
  library takes (``shcfd_set_cut_table``), in the meaning of
  store/cond/core/mod_cond.py:28-50 and :86-91 (geo, width, order, then
  plane, constant-index cell, first cell, outward sign for each side).
  Boundary stretches not covered by a cut are returned as the segments the
  user's boundary conditions must cover (toSetManually of
  store/cond/core/set_cond_var.py:46-88).
"""
from __future__ import annotations

import dataclasses
import struct
from typing import List, Sequence, Tuple

import numpy as np


# ---------------------------------------------------------------------------
# Plot3D
# ---------------------------------------------------------------------------
def read_plot3d(path: str) -> Tuple[List[np.ndarray], List[np.ndarray]]:
    """Node coordinates of every plane: lists of x and y arrays of shape
    (ni+1, nj+1), indexed [i, j] (z is read and dropped)."""
    with open(path) as f:
        tokens = f.read().split()
    pos = 0

    def take(n):
        nonlocal pos
        out = tokens[pos:pos + n]
        pos += n
        return out

    nplane = int(float(take(1)[0]))
    dims = []
    for _ in range(nplane):
        a, b, c = (int(float(t)) for t in take(3))
        if c != 1:
            raise ValueError(f"{path}: third dimension {c} is not 1")
        dims.append((a, b))
    xs, ys = [], []
    for (a, b) in dims:
        n = a * b
        x = np.array(take(n), dtype=np.float64).reshape(b, a).T.copy()
        y = np.array(take(n), dtype=np.float64).reshape(b, a).T.copy()
        take(n)
        xs.append(x)
        ys.append(y)
    if pos != len(tokens):
        raise ValueError(f"{path}: {len(tokens) - pos} values left over")
    return xs, ys


def write_plot3d(path: str, xs: Sequence[np.ndarray], ys: Sequence[np.ndarray]) -> None:
    """Write planes of node coordinates ((ni+1, nj+1) arrays) with z = 0."""
    with open(path, "w") as f:
        f.write(f"{len(xs):8d}\n")
        for x in xs:
            f.write(f"{x.shape[0]:8d} {x.shape[1]:8d} {1:8d}\n")
        for x, y in zip(xs, ys):
            for arr in (x, y, np.zeros_like(x)):
                flat = np.asarray(arr, dtype=np.float64).T.reshape(-1)
                for k in range(0, flat.size, 4):
                    f.write(" ".join(f"{v:24.16e}" for v in flat[k:k + 4]) + "\n")  # 17 digits: exact


# ---------------------------------------------------------------------------
# Fortran sequential unformatted records
# ---------------------------------------------------------------------------
def read_records(path: str, endian: str = ">") -> List[bytes]:
    """All records of a sequential unformatted file (gfortran markers)."""
    with open(path, "rb") as f:
        data = f.read()
    recs, pos, cur = [], 0, b""
    while pos < len(data):
        (n,) = struct.unpack(endian + "i", data[pos:pos + 4])
        m = abs(n)
        cur += data[pos + 4:pos + 4 + m]
        (tail,) = struct.unpack(endian + "i", data[pos + 4 + m:pos + 8 + m])
        if abs(tail) != m:
            raise ValueError(f"{path}: record markers {n} / {tail} differ at byte {pos}")
        pos += 8 + m
        if n >= 0:
            recs.append(cur)
            cur = b""
    if cur:
        raise ValueError(f"{path}: unterminated record")
    return recs


def write_records(path: str, records: Sequence[bytes], endian: str = ">", limit: int = 2 ** 31 - 9) -> None:
    """Sequential unformatted records as gfortran writes them: a record
    longer than `limit` bytes is split into subrecords; a subrecord's head
    marker is negative when another subrecord follows it and its tail marker
    is negative when another one precedes it."""
    with open(path, "wb") as f:
        for r in records:
            parts = [r[k:k + limit] for k in range(0, len(r), limit)] or [b""]
            last = len(parts) - 1
            for k, p in enumerate(parts):
                head = -len(p) if k < last else len(p)
                tail = -len(p) if k > 0 else len(p)
                f.write(struct.pack(endian + "i", head) + p + struct.pack(endian + "i", tail))


# ---------------------------------------------------------------------------
# restart.bin / result.bin and geometry.bin
# ---------------------------------------------------------------------------
def read_restart(path: str, ni: Sequence[int], nj: Sequence[int], dimq: int):
    """(step, tt, [array (nj+2, ni+2, dimq+3)]) per plane: q(1:dimq), R,
    gamma, p at i = 0..ni+1 (axis 1), j = 0..nj+1 (axis 0)."""
    recs = read_records(path)
    if len(recs) != 3:
        raise ValueError(f"{path}: {len(recs)} records, expected 3")
    step = struct.unpack(">i", recs[0])[0]
    tt = struct.unpack(">d", recs[1])[0]
    vals = np.frombuffer(recs[2], dtype=">f8").astype(np.float64)
    need = sum((b + 2) * (a + 2) * (dimq + 3) for a, b in zip(ni, nj))
    if vals.size != need:
        raise ValueError(f"{path}: {vals.size} values, the planes need {need}")
    out, k = [], 0
    for a, b in zip(ni, nj):
        n = (b + 2) * (a + 2) * (dimq + 3)
        out.append(vals[k:k + n].reshape(b + 2, a + 2, dimq + 3).copy())
        k += n
    return step, tt, out


def write_restart(path: str, step: int, tt: float, blocks: Sequence[np.ndarray]) -> None:
    body = b"".join(np.ascontiguousarray(b, dtype=">f8").tobytes() for b in blocks)
    write_records(path, [struct.pack(">i", step), struct.pack(">d", tt), body])


def read_geometry(path: str):
    """[(ni, nj, xh (ni+1, nj+1), rh (ni+1, nj+1))] per plane."""
    recs = read_records(path)
    out = []
    for k in range(0, len(recs), 3):
        ni = struct.unpack(">i", recs[k])[0]
        nj = struct.unpack(">i", recs[k + 1])[0]
        v = np.frombuffer(recs[k + 2], dtype=">f8").astype(np.float64).reshape(nj + 1, ni + 1, 2)
        out.append((ni, nj, v[..., 0].T.copy(), v[..., 1].T.copy()))
    return out


def write_geometry(path: str, xs: Sequence[np.ndarray], ys: Sequence[np.ndarray]) -> None:
    recs = []
    for x, y in zip(xs, ys):
        ni, nj = x.shape[0] - 1, x.shape[1] - 1
        pairs = np.stack([x.T, y.T], axis=-1)  # (nj+1, ni+1, 2): i fastest, (x, y) innermost
        recs += [struct.pack(">i", ni), struct.pack(">i", nj), np.ascontiguousarray(pairs, dtype=">f8").tobytes()]
    write_records(path, recs)


# ---------------------------------------------------------------------------
# block-interface detection
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Interfaces:
    cuts: np.ndarray          # (n, 11) int32 rows of the cut table
    uncovered: List[Tuple[int, int, int, int]]  # (plane 1-based, side, first cell, last cell)


def _lines(xs, ys):
    """Boundary node lines in the reference's order: per plane the i-lower,
    i-upper, j-lower and j-upper sides (cind = 0 for i = const, lu = 0 for
    the lower side), node index along the line from 0."""
    out = []
    for p, (x, y) in enumerate(zip(xs, ys)):
        n_i, n_j = x.shape
        for cind, lu in ((0, 0), (0, 1), (1, 0), (1, 1)):
            if cind == 0:
                k = 0 if lu == 0 else n_i - 1
                pts = np.stack([x[k, :], y[k, :]], axis=-1)
            else:
                k = 0 if lu == 0 else n_j - 1
                pts = np.stack([x[:, k], y[:, k]], axis=-1)
            out.append((p, cind, lu, pts))
    return out


def detect_interfaces(xs: Sequence[np.ndarray], ys: Sequence[np.ndarray], rtol: float = 1e-10) -> Interfaces:
    """Cut-table rows for every stretch of two boundary lines (of the same or
    different planes) whose consecutive nodes coincide, in the same or the
    opposite direction.  Nodes coincide when both coordinates agree to
    rtol times the extent of the grid (the reference compares exactly, which
    is the same for grids written with the same printed values; generated
    grids whose blocks round their shared nodes differently also match).  The earlier line in the reference's order is
    side 1; s1 is the first cell (1-based) of the stretch on it, s2 the cell
    of side 2 that meets it, order = +1 / -1 for the same / opposite
    direction.  A stretch found from its first node is not searched again
    from its later nodes."""
    lines = _lines(xs, ys)
    ext = max(max(float(np.ptp(x)), float(np.ptp(y))) for x, y in zip(xs, ys))
    atol = rtol * (ext if ext > 0.0 else 1.0)

    def same(a, b):
        return abs(a[0] - b[0]) <= atol and abs(a[1] - b[1]) <= atol

    rows = []
    covered = [np.zeros(len(l[3]) - 1, dtype=bool) for l in lines]
    for a, (p1, c1, lu1, P) in enumerate(lines):
        na = len(P)
        i = 0
        while i < na - 1:
            found = False
            for b in range(a, len(lines)):
                p2, c2, lu2, Q = lines[b]
                nb = len(Q)
                # coincident nodes with a node index later in the reference's list
                cand = np.nonzero((np.abs(Q[:, 0] - P[i, 0]) <= atol) & (np.abs(Q[:, 1] - P[i, 1]) <= atol))[0]
                for jq in cand:
                    if b == a and jq <= i:
                        continue
                    for pm in (1, -1):
                        jn = jq + pm
                        if not (0 <= jn < nb) or not same(Q[jn], P[i + 1]):
                            continue
                        w = 1
                        while (i + w + 1 < na and 0 <= jq + (w + 1) * pm < nb
                               and same(Q[jq + (w + 1) * pm], P[i + w + 1])):
                            w += 1
                        # w coincident edges from node i (side 1) and node jq (side 2)
                        s1 = i + 1
                        s2 = jq + 1 if pm > 0 else jq
                        pm1, pm2 = -1 + 2 * lu1, -1 + 2 * lu2
                        ncell1 = xs[p1].shape[c1] - 1
                        ncell2 = xs[p2].shape[c2] - 1
                        cc1 = 1 if pm1 < 0 else ncell1
                        cc2 = 1 if pm2 < 0 else ncell2
                        rows.append([c1 * 2 + c2 + 1, w, pm, p1 + 1, cc1, s1, pm1, p2 + 1, cc2, s2, pm2])
                        covered[a][i:i + w] = True
                        lo = min(s2, s2 + pm * (w - 1))
                        covered[b][lo - 1:lo - 1 + w] = True
                        i += w - 1
                        found = True
                        break
                    if found:
                        break
                if found:
                    break
            i += 1
    uncovered = []
    for a, (p, cind, lu, P) in enumerate(lines):
        side = cind * 2 + lu  # 0 i-min, 1 i-max, 2 j-min, 3 j-max (the BC table's sides)
        cov = covered[a]
        k = 0
        while k < cov.size:
            if cov[k]:
                k += 1
                continue
            e = k
            while e + 1 < cov.size and not cov[e + 1]:
                e += 1
            uncovered.append((p + 1, side, k + 1, e + 1))
            k = e + 1
    return Interfaces(np.array(rows, dtype=np.int32).reshape(-1, 11), uncovered)
