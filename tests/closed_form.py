"""Closed-form references and synthetic geometries for pinning the oracle.

- exact Riemann solution of the Euler equations (Toro, "Riemann Solvers and
  Numerical Methods for Fluid Dynamics", 3rd ed., ch. 4: pressure function,
  Newton iteration for p*, sampling of the self-similar solution)
- a closed rectangular box on the solver grid in the reference's geometry
  encoding (geoA / geoG / geoM / mask as cutcell.js:144-184 lays them out)
"""
import numpy as np


# ---------------------------------------------------------------- exact Riemann solver
def _fK(p, rK, pK, cK, g):
    if p > pK:  # shock
        A = 2.0 / ((g + 1) * rK)
        B = (g - 1) / (g + 1) * pK
        f = (p - pK) * np.sqrt(A / (p + B))
        df = np.sqrt(A / (B + p)) * (1 - (p - pK) / (2 * (B + p)))
    else:  # rarefaction
        f = 2 * cK / (g - 1) * ((p / pK) ** ((g - 1) / (2 * g)) - 1)
        df = 1 / (rK * cK) * (p / pK) ** (-(g + 1) / (2 * g))
    return f, df


def star_state(WL, WR, g):
    rL, uL, pL = WL
    rR, uR, pR = WR
    cL, cR = np.sqrt(g * pL / rL), np.sqrt(g * pR / rR)
    p = max(1e-8, 0.5 * (pL + pR) - 0.125 * (uR - uL) * (rL + rR) * (cL + cR))
    for _ in range(100):
        fL, dL = _fK(p, rL, pL, cL, g)
        fR, dR = _fK(p, rR, pR, cR, g)
        pn = p - (fL + fR + uR - uL) / (dL + dR)
        pn = max(pn, 1e-10)
        if abs(pn - p) / (0.5 * (pn + p)) < 1e-14:
            p = pn
            break
        p = pn
    fL, _ = _fK(p, rL, pL, cL, g)
    fR, _ = _fK(p, rR, pR, cR, g)
    u = 0.5 * (uL + uR) + 0.5 * (fR - fL)
    return p, u


def sample(WL, WR, g, s):
    """Exact (rho, u, p) at similarity coordinates s = x/t."""
    rL, uL, pL = WL
    rR, uR, pR = WR
    cL, cR = np.sqrt(g * pL / rL), np.sqrt(g * pR / rR)
    ps, us = star_state(WL, WR, g)
    out = np.zeros((len(s), 3))
    for k, x in enumerate(s):
        if x <= us:  # left of contact
            if ps > pL:
                SL = uL - cL * np.sqrt((g + 1) / (2 * g) * ps / pL + (g - 1) / (2 * g))
                if x <= SL:
                    out[k] = (rL, uL, pL)
                else:
                    r = rL * ((ps / pL + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pL + 1))
                    out[k] = (r, us, ps)
            else:
                SHL = uL - cL
                cs = cL * (ps / pL) ** ((g - 1) / (2 * g))
                STL = us - cs
                if x <= SHL:
                    out[k] = (rL, uL, pL)
                elif x > STL:
                    out[k] = (rL * (ps / pL) ** (1 / g), us, ps)
                else:
                    f = 2 / (g + 1) + (g - 1) / ((g + 1) * cL) * (uL - x)
                    out[k] = (rL * f ** (2 / (g - 1)), 2 / (g + 1) * (cL + (g - 1) / 2 * uL + x), pL * f ** (2 * g / (g - 1)))
        else:
            if ps > pR:
                SR = uR + cR * np.sqrt((g + 1) / (2 * g) * ps / pR + (g - 1) / (2 * g))
                if x >= SR:
                    out[k] = (rR, uR, pR)
                else:
                    r = rR * ((ps / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pR + 1))
                    out[k] = (r, us, ps)
            else:
                SHR = uR + cR
                cs = cR * (ps / pR) ** ((g - 1) / (2 * g))
                STR = us + cs
                if x >= SHR:
                    out[k] = (rR, uR, pR)
                elif x < STR:
                    out[k] = (rR * (ps / pR) ** (1 / g), us, ps)
                else:
                    f = 2 / (g + 1) - (g - 1) / ((g + 1) * cR) * (uR - x)
                    out[k] = (rR * f ** (2 / (g - 1)), 2 / (g + 1) * (-cR + (g - 1) / 2 * uR + x), pR * f ** (2 * g / (g - 1)))
    return out


# ---------------------------------------------------------------- box geometry
def box_geometry(Nx, Nr, solid_cols=(0,), solid_right=True, solid_top=True):
    """Closed channel of full cells: solid left/right columns and top row, axis at j=0.
    Returns mask, geoA, geoG, geoM in the reference's layout (cutcell.js:144-184)."""
    solid = np.zeros((Nr, Nx), bool)
    for c in solid_cols:
        solid[:, c] = True
    if solid_right:
        solid[:, Nx - 1] = True
    if solid_top:
        solid[Nr - 1, :] = True
    return cells_geometry(solid)


def merge_cells(geoG, geoM, Nx, Nr, slave, host):
    """Merge full cell `slave` = (i, j) into its 4-neighbour `host` (cutcell.js:121-142,
    164-184 encoding): both members carry the group sums (V, Am, Wx, Wr) in geoG, the slave
    its host offset in geoM.xy, the host the slave's direction bit in geoM.z (1: west of the
    host, 2: east, 4: south, 8: north)."""
    G = geoG.reshape(Nr, Nx, 4)
    M = geoM.reshape(Nr, Nx, 4)
    (si, sj), (hi, hj) = slave, host
    assert abs(si - hi) + abs(sj - hj) == 1
    group = G[sj, si].astype(np.float64) + G[hj, hi].astype(np.float64)
    G[sj, si] = group
    G[hj, hi] = group
    M[sj, si, 0] = hi - si
    M[sj, si, 1] = hj - sj
    bit = 1 if si < hi else 2 if si > hi else 4 if sj < hj else 8
    M[hj, hi, 2] = float(int(M[hj, hi, 2]) | bit)


def cells_geometry(solid):
    """Geometry of full square cells (no cut cells) with the given solid mask (Nr x Nx):
    volumes j + 1/2, r-weighted face areas, wall vectors by closure, in the reference's
    layout (cutcell.js:144-184). Returns mask, geoA, geoG, geoM (flat float32 arrays)."""
    solid = np.asarray(solid, bool)
    Nr, Nx = solid.shape
    j = np.arange(Nr)[:, None] * np.ones((1, Nx))
    V = np.where(solid, 0.0, j + 0.5)
    Am = np.where(solid, 0.0, 1.0)
    W1 = Nx + 1
    Ax = np.zeros((Nr + 1, W1))
    Ar = np.zeros((Nr + 1, W1))
    for jj in range(Nr):
        for ii in range(Nx + 1):
            left = solid[jj, ii - 1] if ii > 0 else False
            right = solid[jj, ii] if ii < Nx else False
            Ax[jj, ii] = 0.0 if (left or right) else jj + 0.5
    for jj in range(Nr + 1):
        for ii in range(Nx):
            lo = solid[jj - 1, ii] if jj > 0 else False
            hi = solid[jj, ii] if jj < Nr else False
            Ar[jj, ii] = 0.0 if (lo or hi) else float(jj)
    geoA = np.zeros(((Nr + 1) * W1, 4), np.float32)
    g = geoA.reshape(Nr + 1, W1, 4)
    g[:Nr, :Nx, 0] = V
    g[:Nr, :Nx, 1] = Am
    g[:Nr, :, 2] = Ax[:Nr]
    g[:, :Nx, 3] = Ar[:, :Nx]
    Wx = np.where(solid, 0.0, Ax[:Nr, :Nx] - Ax[:Nr, 1:Nx + 1])
    Wr = np.where(solid, 0.0, Am - (Ar[1:Nr + 1, :Nx] - Ar[:Nr, :Nx]))
    Wx[np.abs(Wx) < 1e-6] = 0
    Wr[np.abs(Wr) < 1e-6] = 0
    geoG = np.zeros((Nr, Nx, 4), np.float32)
    geoG[..., 0] = np.where(solid, 1.0, V)
    geoG[..., 1] = Am
    geoG[..., 2] = Wx
    geoG[..., 3] = Wr
    geoM = np.zeros((Nr, Nx, 4), np.float32)
    return solid.astype(np.uint8).ravel(), geoA.ravel(), geoG.ravel(), geoM.ravel()


def uniform_params(rho=1.0, p=1.0, gamma=1.4):
    """A parameter block with no inlet, no physics and free stream = (rho, 0, 0, p)."""
    from rnozzle import abi
    P = abi.NzParams()
    G = 1.0 / (gamma - 1)
    P.amb[:] = [rho, 0, 0, p]
    P.res[:] = [p, 1.0, 1.0]
    P.inlet[:] = [0.0, 1.0, G, G]
    P.pStart = p
    P.ramp = 0
    P.RcCells = 0
    P.inj = 0
    P.order2 = 1
    P.hllBlend = 1.0
    P.ambCons[:] = [rho, 0, 0, p * G]
    P.ambB[:] = [0, 0, G, 0]
    P.gamR[:] = [G, G]
    P.uInlet0 = 0
    P.vMax = 1e3
    P.cfl = 0.5
    P.dtMax = 10.0
    P.wFloor = 0.0
    P.visc = 0
    P.chem = 0
    P.gas[:] = [1.0, 1.0, 288.15, 1.0]
    P.Rs[:] = [1.0, 1.0, 1.0]
    P.cap[:] = [2400, 2800]
    P.chemK[:] = [0, 1, 0, p]
    P.Tw = 1.0
    return P


def state_from_prims(rho, u, v, p, gamma=1.4):
    G = 1.0 / (gamma - 1)
    A = np.stack([rho, rho * u, rho * v, p * G + 0.5 * rho * (u * u + v * v)], -1).astype(np.float32)
    B = np.stack([np.zeros_like(rho), np.zeros_like(rho), np.full_like(rho, G), np.zeros_like(rho)], -1).astype(np.float32)
    return A.ravel(), B.ravel()
