"""Python restatement of the reference's Chemkin input readers.

TEST INFRASTRUCTURE ONLY: pins the library's C++ parser (shcfd_mech_parse)
and feeds mechanism tables to the C oracle.  Follows
store/therm_lib/chemkin/core/sub_chem.f90: read_cheminp (:191-598) with its
helpers next_line (:519-539), read_tail_number(_arr) (:504-559),
read_head_char (:560-573), search_SYM (:574-587), fetch_slash (:588-597);
set_trans_data (:788-889).  Fills a ctypes mirror of shcfd_mech
(include/shcfd_types.h) defined here, independently of the product.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

MAX_NE, MAX_NS, MAX_NR, MAX_TSEC = 8, 16, 48, 5
_d, _i = ctypes.c_double, ctypes.c_int


class OrMech(ctypes.Structure):
    _fields_ = [
        ("ne", _i), ("ns", _i), ("nr", _i), ("nt", _i),
        ("mw", _d * MAX_NS), ("inv_mw", _d * MAX_NS), ("tthre", (_d * 3) * MAX_NS),
        ("coeff", ((_d * 7) * 2) * MAX_NS), ("es", (_d * MAX_NE) * MAX_NS),
        ("reverse_off", _i * MAX_NR), ("kind", _i * MAX_NR), ("num_nu", (_i * 2) * MAX_NR),
        ("ind_nu", ((_i * 3) * 2) * MAX_NR), ("snu", _i * MAX_NR), ("exist_m", _i * MAX_NR),
        ("num_m", _i * MAX_NR), ("ind_m", (_i * MAX_NS) * MAX_NR), ("men", (_d * MAX_NS) * MAX_NR),
        ("abe", (_d * 3) * MAX_NR), ("cabe", (_d * 3) * MAX_NR), ("troe", (_d * 4) * MAX_NR),
        ("tr2th", _i * MAX_NS), ("nsec_trans", _i * MAX_NS),
        ("trange", ((_d * 2) * MAX_TSEC) * MAX_NS), ("trans", ((_d * 4) * MAX_TSEC) * MAX_NS),
        ("rtol", _d), ("atol", _d),
    ]


# invRc = 1d0/(1.987e0): the literal is default (single) precision
INV_RC = 1.0 / float(np.float32(1.987))


def _fline(s, width=100):
    s = s.rstrip("\n").rstrip("\r").replace("\t", " ")
    return s[:width]


class _Reader:
    def __init__(self, path):
        with open(path) as f:
            self.lines = f.readlines()
        self.k = 0

    def next_line(self):
        while self.k < len(self.lines):
            s = _fline(self.lines[self.k])
            self.k += 1
            ind = s.find("!")
            if ind == 0:
                continue
            if ind > 0:
                s = s[:ind]
            s = s.rstrip()
            if s:
                return s
        return ""


def _real(tok):
    return float(tok.strip().replace("D", "E").replace("d", "E"))


def _fixed(line, c0, w):
    f = line[c0:c0 + w].replace(" ", "").replace("D", "E").replace("d", "E")
    if not f:
        return 0.0
    for p in range(1, len(f)):
        if f[p] in "+-" and f[p - 1] not in "Ee":
            f = f[:p] + "E" + f[p:]
            break
    return float(f)


def _int(line, c0, w):
    f = line[c0:c0 + w].strip()
    return int(f) if f else 0


def _tail_number(line):
    line = line.rstrip()
    ind = line.rfind(" ")
    v = _real(line[ind + 1:])
    return v, (line[:ind + 1].rstrip() if ind >= 0 else "")


def _tail_numbers(line, num):
    arr = [0.0] * num
    for i in range(num - 1, -1, -1):
        arr[i], line = _tail_number(line)
    if num == 3:
        arr[0] = math.log(arr[0])
        arr[2] = arr[2] * INV_RC
    else:
        arr[1] = 1.0 / arr[1]
        arr[2] = 1.0 / arr[2]
    return arr, line


def _head(line):
    line = line.lstrip()
    ind = line.find(" ")
    if ind < 0:
        return line, ""
    return line[:ind], line[ind:].strip()


def _search(s, sym):
    ind = s.find("+")
    if ind < 0:
        buf, rest = s.rstrip(), ""
    else:
        buf, rest = s[:ind].rstrip(), s[ind + 1:].rstrip()
    if buf not in sym:
        raise ValueError(f"unknown species {buf!r}")
    return sym.index(buf), rest


def _slash(line):
    a = line.index("/")
    b = line.index("/", a + 1)
    return line[a + 1:b], line[b + 1:].strip()


def read_cheminp(chem_inp, mw_inp, trans_inp=None, rtol=1e-6, atol=1e-6) -> OrMech:
    m = OrMech()
    R = _Reader(chem_inp)
    assert R.next_line() == "elements"
    elm = []
    while True:
        s = R.next_line().strip()
        if s in ("end", "END"):
            break
        while s:
            t, s = _head(s)
            elm.append(t)
    m.ne = len(elm)
    mwe = {}
    M = _Reader(mw_inp)
    while True:
        s = M.next_line()
        if not s:
            break
        t, s = _head(s)
        if t in elm and t not in mwe:
            mwe[t], _ = _tail_number(s)
    assert len(mwe) == len(elm), "element missing from MW.inp"
    assert R.next_line() == "species"
    spc = []
    while True:
        s = R.next_line().strip()
        if s in ("end", "END"):
            break
        while s:
            t, s = _head(s)
            spc.append(t)
    m.ns = len(spc)
    assert R.next_line() == "thermo all"
    s = R.next_line()
    tdef = [_fixed(s, 0, 10), _fixed(s, 10, 10), _fixed(s, 20, 10)]
    while True:
        s = R.next_line()
        if s in ("end", "END"):
            break
        name = s[:24].split(" ")[0]
        if name not in spc:
            R.next_line(); R.next_line(); R.next_line()
            continue
        k = spc.index(name)
        assert s[44] == "G"
        pairs = [(s[24 + 5 * i:26 + 5 * i], _int(s, 26 + 5 * i, 3)) for i in range(4)]
        pairs.append((s[73:75], _int(s, 75, 3)))
        tb = [_fixed(s, 45, 10), _fixed(s, 55, 10), _fixed(s, 65, 8)]
        for sym, cnt in pairs:
            if sym.strip():
                m.es[k][elm.index(sym.strip())] = cnt
        mw = 0.0
        for e in range(len(elm)):
            mw = mw + m.es[k][e] * mwe[elm[e]]
        m.mw[k] = mw
        for i in range(3):
            m.tthre[k][i] = tb[i] if tb[0] != 0.0 else tdef[i]
        l2, l3, l4 = R.next_line(), R.next_line(), R.next_line()
        hi = [_fixed(l2, 15 * i, 15) for i in range(5)] + [_fixed(l3, 0, 15), _fixed(l3, 15, 15)]
        lo = [_fixed(l3, 30 + 15 * i, 15) for i in range(3)] + [_fixed(l4, 15 * i, 15) for i in range(4)]
        for i in range(7):
            m.coeff[k][1][i] = hi[i]
            m.coeff[k][0][i] = lo[i]
    assert R.next_line() == "reactions      cal/mole  moles"
    nr = 0
    while True:
        s = R.next_line().strip()
        if s in ("end", "END"):
            break
        if s in ("duplicate", "DUPLICATE"):
            continue
        ind = s.find("/")
        if ind < 0:
            r = nr
            nr += 1
            abe, s = _tail_numbers(s, 3)
            for i in range(3):
                m.abe[r][i] = abe[i]
            ind = s.find("<=>")
            if ind >= 0:
                lr, lp = s[:ind], s[ind + 3:]
            else:
                ind = s.find("=")
                lr, lp = s[:ind], s[ind + 1:]
            lr, lp = lr.rstrip(), lp.rstrip()
            if lp.startswith(">"):
                m.reverse_off[r] = 1
                lp = lp[1:]
            if "(+M)" in lr:
                assert lr.find("(+M)") + 4 == len(lr)
                lr, lp = lr[:-4], lp[:-4]
                m.exist_m[r] = 1
            elif "+M" in lr:
                assert lr.find("+M") + 2 == len(lr)
                lr, lp = lr[:-2], lp[:-2]
                m.exist_m[r] = 1
                m.kind[r] = 4
            for side, txt in ((0, lr), (1, lp)):
                i = 0
                while True:
                    j, txt = _search(txt, spc)
                    m.ind_nu[r][side][i] = j
                    i += 1
                    if not txt:
                        break
                m.num_nu[r][side] = i
        else:
            r = nr - 1
            key, rest = s[:ind].rstrip(), s[ind:]
            if key in ("REV", "rev"):
                m.kind[r] = m.kind[r] + 1
                txt, rest = _slash(rest)
                v, _ = _tail_numbers(txt, 3)
                for i in range(3):
                    m.cabe[r][i] = v[i]
            elif key in ("LOW", "low"):
                m.kind[r] = 2
                txt, rest = _slash(rest)
                v, _ = _tail_numbers(txt, 3)
                for i in range(3):
                    m.cabe[r][i] = v[i]
            elif key in ("TROE", "troe"):
                m.kind[r] = 3
                txt, rest = _slash(rest)
                v, _ = _tail_numbers(txt, 4)
                for i in range(4):
                    m.troe[r][i] = v[i]
            else:
                assert m.exist_m[r]
                j = 0
                sp = key
                while True:
                    m.ind_m[r][j], _ = _search(sp, spc)
                    txt, rest = _slash(rest)
                    v, _ = _tail_number(txt)
                    m.men[r][j] = v - 1.0
                    j += 1
                    if not rest:
                        break
                    ind = rest.find("/")
                    sp, rest = rest[:ind], rest[ind:]
                m.num_m[r] = j
    for i in range(nr):
        m.snu[i] = m.num_nu[i][1] - m.num_nu[i][0]
    for j in range(m.ns):
        m.inv_mw[j] = 1.0 / m.mw[j]
    m.nr = max(nr, 1)
    m.rtol, m.atol = rtol, atol
    if trans_inp:
        with open(trans_inp) as f:
            lines = [_fline(x, 300) for x in f.readlines()]
        k = 1
        while k < len(lines):
            s = lines[k]
            k += 1
            if s.startswith("!"):
                continue
            if s[:3] in ("END", "end"):
                break
            sname = s[:15].rstrip()
            comment = s[16:31].strip()
            iv, ic = _int(s, 35, 1), _int(s, 37, 1)
            if comment or sname not in spc:
                k += iv + ic
                continue
            t = m.nt
            m.nt += 1
            m.tr2th[t] = spc.index(sname)
            m.nsec_trans[t] = iv
            for i in range(iv):
                v = lines[k]
                k += 1
                assert v[1] == "V"
                m.trange[t][i][0] = _fixed(v, 2, 7)
                m.trange[t][i][1] = _fixed(v, 11, 7)
                for j in range(4):
                    m.trans[t][i][j] = _fixed(v, 20 + 15 * j, 15)
            k += ic
    return m
