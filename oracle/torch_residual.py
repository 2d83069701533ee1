"""Float64 torch restatement of the steady ideal-gas residual (inviscid or
laminar), for checking the hand-written discrete-adjoint kernels with torch
autograd.

TEST INFRASTRUCTURE ONLY: imported by tests/ (the checker), never by the
product.  It restates, for one block with nY = 1:

* face metrics from the nodes (dim/2d/geometry.f90:79-169),
* primitive and ideal-gas thermo state (core/scheme.f90:1-29,
  therm_lib/ideal/thermal_model.raw.f90:7-33),
* ghost fills: cut rows (cond/no-nV/condition.tail.f90:2-70, the cut_copro
  row format), side rows (zero gradient, curved slip wall, no-slip wall,
  fixed state, inflow of given T, u, v), then the corners
  (condition.tail.f90:331-356),
* MUSCL face states with the reference limiter (high_order/muscl/muscl.f90:7-73),
* the SLAU flux (core/scheme.f90:71-153),
* when viscous: cell centres with extrapolated ghost centres and the face
  metrics from them (viscosity/viscous/no-nV/2d/sch_viscous.f90:1-54), the
  viscous flux (sch_viscous.f90:106-178) with the species enthalpy
  vhi = kappa/(kappa-1) R T of the ideal gas (ghosts filled like w),
* the flux balance (LU-SGS/2d/time_lusgs_viscous.f90:12-22, Area = 0) of
  TG - TGv,
* the wall-force objective J = sum dsj ((TG - TGv)_u a_x + (TG - TGv)_v a_y)
  over j-min faces.

Arrays are (j, i) ordered like the reference layout without ghosts for q and
(i, j) for the nodes, matching Case.xh/Case.rh.  Its fluxes are pinned
against the C oracle's set_tg in tests/test_adjoint_host.py.
"""
from __future__ import annotations

import torch

BC_EXTRAP, BC_REFLECT, BC_NOSLIP, BC_FIXED, BC_INFLOW_TUV, BC_OUTFLOW_P, BC_SLIP_NORMAL = range(7)
SIDE_IMIN, SIDE_IMAX, SIDE_JMIN, SIDE_JMAX = range(4)


def face_metrics(x, y):
    """x, y: (ni+1, nj+1) nodes -> dsi, vni for i-faces (ni+1, nj) and dsj,
    vnj for j-faces (ni, nj+1)."""
    dxi = x[:, 1:] - x[:, :-1]
    dyi = y[:, 1:] - y[:, :-1]
    dsi = torch.sqrt(dxi * dxi + dyi * dyi)
    vni = torch.stack([dyi / dsi, -dxi / dsi], dim=-1)
    dxj = x[1:, :] - x[:-1, :]
    dyj = y[1:, :] - y[:-1, :]
    dsj = torch.sqrt(dxj * dxj + dyj * dyj)
    vnj = torch.stack([-dyj / dsj, dxj / dsj], dim=-1)
    return dsi, vni, dsj, vnj


def ideal_w(q, kap, rg, nu=0.0):
    """q (..., 4) -> w (..., 10): rho, u, v, p, Y, gamma, H, R, mu and the
    species enthalpy vhi."""
    rho = q[..., 0]
    inv = 1.0 / rho
    u = q[..., 1] * inv
    v = q[..., 2] * inv
    Y = q[..., 0] * inv
    ei = q[..., 3] / rho - 0.5 * (u * u + v * v)
    T = (ei * (kap - 1.0)) / rg
    p = (rho * rg) * T
    H = (q[..., 3] + p) / rho
    one = torch.ones_like(rho)
    vhi = ((kap / (kap - 1.0)) * rg) * T
    return torch.stack([rho, u, v, p, Y, kap * one, H, rg * one, rho * nu, vhi], dim=-1)


def _cut_cells(t, inc, copy):
    geo, order = int(t[0]), int(t[2])
    p1, c1, s1, pm1 = int(t[3]) - 1, int(t[4]), int(t[5]), int(t[6])
    p2, c2, s2, pm2 = int(t[7]) - 1, int(t[8]), int(t[9]), int(t[10])
    a1, a2 = s1 + inc, s2 + inc * order
    to1 = copy < 2
    lay = copy & 1
    g = (c1 + (2 * pm1 if lay else pm1)) if to1 else (c2 + (2 * pm2 if lay else pm2))
    src = ((c2 - pm2) if lay else c2) if to1 else ((c1 - pm1) if lay else c1)
    if geo == 1:
        return ((g, a1), (src, a2)) if to1 else ((g, a2), (src, a1))
    if geo == 2:
        return ((g, a1), (a2, src)) if to1 else ((a2, g), (src, a1))
    if geo == 3:
        return ((a1, g), (src, a2)) if to1 else ((g, a2), (a1, src))
    if geo == 4:
        return ((a1, g), (a2, src)) if to1 else ((a2, g), (a1, src))
    raise ValueError(geo)


def ghosted_w(w_int, ni, nj, cut_rows, bc_rows, pool, vni, vnj, kap=1.4, rg=1.0, nu=0.0):
    """Interior w (nj, ni, 10) -> w with 2 ghost layers (nj+4, ni+4, 10).
    Single block; cut rows first, then side rows (a later fill of the same
    ghost replaces an earlier one).  Built out of place for autograd."""
    W0 = torch.nn.functional.pad(w_int.movedim(-1, 0), (2, 2, 2, 2)).movedim(0, -1)

    def at(i, j):
        return (j + 1, i + 1)

    fills = {}
    for t in cut_rows:
        for copy in range(4):
            for inc in range(int(t[1])):
                (di, dj), (si, sj) = _cut_cells(t, inc, copy)
                fills[at(di, dj)] = W0[at(si, sj)]
    for row in bc_rows:
        side, s, e, kind, off = int(row[1]), int(row[2]), int(row[3]), int(row[4]), int(row[5])
        n = nj if side <= SIDE_IMAX else ni
        s, e = max(s, 1), min(e, n)
        for tpos in range(s, e + 1):
            if side == SIDE_IMIN:
                ii, jj, g1, g2 = 1, tpos, (0, tpos), (-1, tpos)
            elif side == SIDE_IMAX:
                ii, jj, g1, g2 = ni, tpos, (ni + 1, tpos), (ni + 2, tpos)
            elif side == SIDE_JMIN:
                ii, jj, g1, g2 = tpos, 1, (tpos, 0), (tpos, -1)
            else:
                ii, jj, g1, g2 = tpos, nj, (tpos, nj + 1), (tpos, nj + 2)
            if kind == BC_FIXED:
                wv = pool[off:off + 10]
            elif kind == BC_INFLOW_TUV:
                # T, u, v given, p from the cell (ideal gas)
                T, u, v = pool[off], pool[off + 1], pool[off + 2]
                p_in = W0[at(ii, jj)][3]
                rho = p_in / (rg * T)
                cp = (kap / (kap - 1.0)) * rg
                one = torch.ones_like(rho)
                wv = torch.stack([rho, u * one, v * one, p_in, one, kap * one, cp * T + 0.5 * (u * u + v * v),
                                  rg * one, nu * rho, cp * T * one])
            else:
                wv = W0[at(ii, jj)]
                if kind == BC_NOSLIP:
                    wv = torch.cat([wv[:1], -wv[1:3], wv[3:]])
                elif kind == BC_SLIP_NORMAL:
                    if side == SIDE_JMIN:
                        nx, ny = vnj[ii - 1, 0, 0], vnj[ii - 1, 0, 1]
                    elif side == SIDE_JMAX:
                        nx, ny = vnj[ii - 1, nj, 0], vnj[ii - 1, nj, 1]
                    elif side == SIDE_IMIN:
                        nx, ny = vni[0, jj - 1, 0], vni[0, jj - 1, 1]
                    else:
                        nx, ny = vni[ni, jj - 1, 0], vni[ni, jj - 1, 1]
                    u, v = wv[1], wv[2]
                    un = nx * u + ny * v
                    wv = torch.cat([wv[:1], (u - 2.0 * un * nx)[None], (v - 2.0 * un * ny)[None], wv[3:]])
                elif kind != BC_EXTRAP:
                    raise NotImplementedError(kind)
            fills[at(*g1)] = wv
            fills[at(*g2)] = wv
    keys = list(fills.keys())
    rows = torch.tensor([k[0] for k in keys])
    cols = torch.tensor([k[1] for k in keys])
    vals = torch.stack([fills[k] for k in keys])
    W1 = W0.index_put((rows, cols), vals)
    # corners: the mean of the corner cell and its two side ghosts
    cr, cc, cv = [], [], []
    for pi in (-1, 1):
        i = 1 if pi < 0 else ni
        for pj in (-1, 1):
            j = 1 if pj < 0 else nj
            cr.append(at(i + pi, j + pj)[0])
            cc.append(at(i + pi, j + pj)[1])
            cv.append((1.0 / 3.0) * ((W1[at(i, j)] + W1[at(i, j + pj)]) + W1[at(i + pi, j)]))
    return W1.index_put((torch.tensor(cr), torch.tensor(cc)), torch.stack(cv))


def centres(x, y):
    """Cell centres (nj+2, ni+2) with one layer of ghost centres: the means
    of the four nodes, then x(0) = 2 x(1) - x(2) on the i-sides and after
    them the j-sides (dim/2d/geometry.f90:79-169)."""
    def one(z):
        c = 0.25 * (((z[:-1, :-1] + z[1:, :-1]) + z[:-1, 1:]) + z[1:, 1:])
        c = c.transpose(0, 1)  # (nj, ni)
        c = torch.cat([(2.0 * c[:, 0] - c[:, 1])[:, None], c, (2.0 * c[:, -1] - c[:, -2])[:, None]], dim=1)
        return torch.cat([(2.0 * c[0] - c[1])[None], c, (2.0 * c[-1] - c[-2])[None]], dim=0)
    return one(x), one(y)


def geojac(xs, rs):
    """Face metrics (g11, g21, g12, g22) from the centre coordinates of the
    stencil (c, c1, a, b, e, f), DIR 0 layout (xi along c -> c1)."""
    xc, x1, xa, xb, xe, xf = xs
    rc, r1, ra, rb, re, rf = rs
    dxdxi, dydxi = x1 - xc, r1 - rc
    dxdeta = (((xa + xb) - xe) - xf) * 0.25
    dydeta = (((ra + rb) - re) - rf) * 0.25
    det = (dxdxi * dydeta) - (dydxi * dxdeta)
    return dydeta / det, -dydxi / det, -dxdeta / det, dxdxi / det


def viscous_flux(ws, g, n1, n2, dir_):
    """Viscous flux (..., 4) from the w of the stencil cells (c, c1, a, b, e,
    f), the face metrics and unit normal.  dir_ 0: c1 - c is the xi
    difference; 1: the eta difference."""
    wc, w1, wa, wb, we, wf = ws
    T = lambda w: (w[..., 3] / w[..., 0]) / w[..., 7]
    avg = lambda k: (wc[..., k] + w1[..., k]) * 0.5
    rho, u, v, gamm, Rg, mu, lhi = avg(0), avg(1), avg(2), avg(5), avg(7), avg(8), avg(9)
    dn = [w1[..., 1] - wc[..., 1], w1[..., 2] - wc[..., 2], T(w1) - T(wc), w1[..., 4] - wc[..., 4]]
    cd = lambda a_, b_, e_, f_: (((a_ + b_) - e_) - f_) * 0.25
    dc = [cd(wa[..., 1], wb[..., 1], we[..., 1], wf[..., 1]), cd(wa[..., 2], wb[..., 2], we[..., 2], wf[..., 2]),
          cd(T(wa), T(wb), T(we), T(wf)), cd(wa[..., 4], wb[..., 4], we[..., 4], wf[..., 4])]
    dxi, deta = (dn, dc) if dir_ == 0 else (dc, dn)
    g11, g21, g12, g22 = g
    D = mu / rho
    k_heat = ((mu * gamm) / (gamm - 1.0)) * Rg
    dudx = dxi[0] * g11 + deta[0] * g21
    dvdx = dxi[1] * g11 + deta[1] * g21
    dTdx = dxi[2] * g11 + deta[2] * g21
    dudr = dxi[0] * g12 + deta[0] * g22
    dvdr = dxi[1] * g12 + deta[1] * g22
    dTdr = dxi[2] * g12 + deta[2] * g22
    dYdx = dxi[3] * g11 + deta[3] * g21
    dYdr = dxi[3] * g12 + deta[3] * g22
    divu = dudx + dvdr
    t_xx = -((2.0 / 3.0) * mu) * divu + (2.0 * mu) * dudx
    t_rr = -((2.0 / 3.0) * mu) * divu + (2.0 * mu) * dvdr
    t_xr = mu * (dvdx + dudr)
    qx = -(k_heat * dTdx) - ((rho * D) * (dYdx * lhi))
    qr = -(k_heat * dTdr) - ((rho * D) * (dYdr * lhi))
    f0 = ((rho * D) * dYdx) * n1 + ((rho * D) * dYdr) * n2
    f1 = t_xx * n1 + t_xr * n2
    f2 = t_xr * n1 + t_rr * n2
    f3 = (((u * t_xx) + (v * t_xr)) - qx) * n1 + (((u * t_xr) + (v * t_rr)) - qr) * n2
    return torch.stack([f0, f1, f2, f3], dim=-1)


def muscl_delta(Du, Dl, k_mus):
    s = ((2.0 * Du) * Dl) / (((Du * Du) + (Dl * Dl)) + 1e-300)
    du = (0.5 * s) * (((1.0 + k_mus * s) * Du) + ((1.0 - k_mus * s) * Dl))
    dl = (0.5 * s) * (((1.0 + k_mus * s) * Dl) + ((1.0 - k_mus * s) * Du))
    return du, dl


def muscl_faces(W, axis, hp, k_mus, high_order):
    """Face states along `axis` (1: i, 0: j) of the ghosted W.  Returns the
    left/right states of faces 0..n (n interior cells along the axis)."""
    Wm = W.movedim(axis, 0)
    n = Wm.shape[0] - 4
    w_m, w_0, w_1, w_2 = Wm[0:n + 1], Wm[1:n + 2], Wm[2:n + 3], Wm[3:n + 4]
    if not high_order:
        return w_0.movedim(0, axis), w_1.movedim(0, axis)
    du0, _ = muscl_delta(w_1 - w_0, w_0 - w_m, k_mus)
    _, dl1 = muscl_delta(w_2 - w_1, w_1 - w_0, k_mus)
    return (w_0 + hp * du0).movedim(0, axis), (w_1 - hp * dl1).movedim(0, axis)


def slau(wl, wr, n1, n2):
    """SLAU flux (..., 4) of face states wl/wr (..., 9), unit normal (n1, n2)."""
    rhol, ul, vl, pl, gl = wl[..., 0], wl[..., 1], wl[..., 2], wl[..., 3], wl[..., 5]
    rhor, ur, vr, pr, gr = wr[..., 0], wr[..., 1], wr[..., 2], wr[..., 3], wr[..., 5]
    al = torch.sqrt((gl * pl) / rhol)
    unl = (n1 * ul) + (n2 * vl)
    ar = torch.sqrt((gr * pr) / rhor)
    unr = (n1 * ur) + (n2 * vr)
    dp = pr - pl
    a_bar = 0.5 * (ar + al)
    p_bar = 0.5 * (pr + pl)
    ml = unl / a_bar
    mr = unr / a_bar
    mt = (1.0 / a_bar) * torch.sqrt(0.5 * ((((ul * ul) + (vl * vl)) + (ur * ur)) + (vr * vr)))
    m_tilde = torch.where(mt > 1.0, torch.ones_like(mt), mt)
    chi = (1.0 - m_tilde) * (1.0 - m_tilde)
    tt = torch.where(ml < 0.0, ml, torch.zeros_like(ml))
    gp = -torch.where(tt > -1.0, tt, -torch.ones_like(tt))
    tt = torch.where(mr > 0.0, mr, torch.zeros_like(mr))
    gm = torch.where(tt < 1.0, tt, torch.ones_like(tt))
    g = gp * gm
    un_bar_abs = ((torch.abs(unr) * rhor) + (torch.abs(unl) * rhol)) / (rhor + rhol)
    unp = ((1.0 - g) * un_bar_abs) + (g * torch.abs(unl))
    unm = ((1.0 - g) * un_bar_abs) + (g * torch.abs(unr))
    dm = 0.5 * (((rhol * (unl + unp)) + (rhor * (unr - unm))) - ((chi / a_bar) * dp))
    sgn = lambda x: torch.where(x >= 0.0, torch.ones_like(x), -torch.ones_like(x))
    betar = torch.where(torch.abs(mr) < 1.0, (0.25 * (2.0 + mr)) * ((mr - 1.0) * (mr - 1.0)), 0.5 * (1.0 - sgn(mr)))
    betal = torch.where(torch.abs(ml) < 1.0, (0.25 * (2.0 - ml)) * ((ml + 1.0) * (ml + 1.0)), 0.5 * (1.0 + sgn(ml)))
    p_tilde = (p_bar + ((0.5 * (betal - betar)) * (pl - pr))) + (((1.0 - chi) * ((betar + betal) - 1.0)) * p_bar)
    cp = 0.5 * (dm + torch.abs(dm))
    cm = 0.5 * (dm - torch.abs(dm))
    f0 = (cp * wl[..., 4]) + (cm * wr[..., 4])
    f1 = ((cp * ul) + (cm * ur)) + (p_tilde * n1)
    f2 = ((cp * vl) + (cm * vr)) + (p_tilde * n2)
    f3 = (cp * wl[..., 6]) + (cm * wr[..., 6])
    return torch.stack([f0, f1, f2, f3], dim=-1)


def residual(q, x, y, pool, case, obj=None):
    """R (nj, ni, 4), the net face fluxes TGi - TGvi (nj, ni+1, 4) and
    TGj - TGvj (nj+1, ni, 4), and the objective of one block.  q: (nj, ni, 4); x, y: (ni+1, nj+1); pool: bc pool tensor.
    obj = (i0, i1, ax, ay) or None."""
    p = case.params
    ni, nj = case.ni[0], case.nj[0]
    dsi, vni, dsj, vnj = face_metrics(x, y)
    w = ideal_w(q, p.kappa, p.rgas, p.nu)
    W = ghosted_w(w, ni, nj, case.cut_rows, case.bc_rows, pool, vni, vnj, p.kappa, p.rgas, p.nu)
    hp = p.phi_mus * 0.5
    # i-faces: rows j = 1..nj, faces i = 0..ni
    Wi = W[2:nj + 2]
    li, ri = muscl_faces(Wi, 1, hp, p.k_mus, p.high_order)
    ni_ = vni.transpose(0, 1)  # (nj, ni+1, 2)
    TGi = slau(li, ri, ni_[..., 0], ni_[..., 1])
    Wj = W[:, 2:ni + 2]
    lj, rj = muscl_faces(Wj, 0, hp, p.k_mus, p.high_order)
    nj_ = vnj.transpose(0, 1)  # (nj+1, ni, 2)
    TGj = slau(lj, rj, nj_[..., 0], nj_[..., 1])
    dsi_t = dsi.transpose(0, 1)[..., None]  # (nj, ni+1, 1)
    dsj_t = dsj.transpose(0, 1)[..., None]  # (nj+1, ni, 1)
    if p.viscous:
        XC, RC = centres(x, y)
        # i-faces (fi, fj), fi = 0..ni, fj = 1..nj: cell (i, j) at W[j+1, i+1], centre at XC[j, i]
        wsi = [W[2:nj + 2, 1:ni + 2], W[2:nj + 2, 2:ni + 3], W[3:nj + 3, 2:ni + 3], W[3:nj + 3, 1:ni + 2],
               W[1:nj + 1, 2:ni + 3], W[1:nj + 1, 1:ni + 2]]
        sl_i = [(slice(1, nj + 1), slice(0, ni + 1)), (slice(1, nj + 1), slice(1, ni + 2)),
                (slice(2, nj + 2), slice(1, ni + 2)), (slice(2, nj + 2), slice(0, ni + 1)),
                (slice(0, nj), slice(1, ni + 2)), (slice(0, nj), slice(0, ni + 1))]
        gi = geojac([XC[s] for s in sl_i], [RC[s] for s in sl_i])
        TGi = TGi - viscous_flux(wsi, gi, ni_[..., 0], ni_[..., 1], 0)
        # j-faces (fi, fj), fi = 1..ni, fj = 0..nj; stencil (c, c1, a, b, e, f) with xi along a/b - e/f
        wsj = [W[1:nj + 2, 2:ni + 2], W[2:nj + 3, 2:ni + 2], W[2:nj + 3, 3:ni + 3], W[1:nj + 2, 3:ni + 3],
               W[2:nj + 3, 1:ni + 1], W[1:nj + 2, 1:ni + 1]]
        sl_j = [(slice(0, nj + 1), slice(1, ni + 1)), (slice(1, nj + 2), slice(1, ni + 1)),
                (slice(1, nj + 2), slice(2, ni + 2)), (slice(0, nj + 1), slice(2, ni + 2)),
                (slice(1, nj + 2), slice(0, ni)), (slice(0, nj + 1), slice(0, ni))]
        # geojac of a j-face: xi from (a + b - e - f)/4, eta from c1 - c
        xs, rs = [XC[s] for s in sl_j], [RC[s] for s in sl_j]
        dxdxi = (((xs[2] + xs[3]) - xs[4]) - xs[5]) * 0.25
        dydxi = (((rs[2] + rs[3]) - rs[4]) - rs[5]) * 0.25
        dxdeta, dydeta = xs[1] - xs[0], rs[1] - rs[0]
        det = (dxdxi * dydeta) - (dydxi * dxdeta)
        gj = (dydeta / det, -dydxi / det, -dxdeta / det, dxdxi / det)
        TGj = TGj - viscous_flux(wsj, gj, nj_[..., 0], nj_[..., 1], 1)
    Fi = dsi_t * TGi
    Fj = dsj_t * TGj
    R = (Fi[:, :-1] - Fi[:, 1:]) + (Fj[:-1, :] - Fj[1:, :])
    J = None
    if obj is not None:
        i0, i1, ax, ay = obj
        sl = slice(i0 - 1, i1)
        J = (dsj_t[0, sl, 0] * (TGj[0, sl, 1] * ax + TGj[0, sl, 2] * ay)).sum()
    return R, TGi, TGj, J
