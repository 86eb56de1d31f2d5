"""Pins for the oracle functions that the whole-flow checks leave open, each against a
reference that does not share the oracle's code path:

- the spray's PCG hash and its 24-bit uniform (spray.js:33-34) against a literal table of
  the published pcg_hash (Jarzynski & Olano 2020: PCG's 32-bit LCG step, O'Neill 2014,
  followed by the RXS-M-XS output permutation)
- the face flux (solver.js:177-255) against HLLC in Toro's second form (Toro, "Riemann
  Solvers and Numerical Methods for Fluid Dynamics", 3rd ed., sec. 10.4.2) and the HLLE
  flux with Einfeldt's wave-speed estimates, on random state pairs; the HLL blend weight
  (solver.js:182-190) on constructed tangential pressure jumps
- the slip-wall pressure (solver.js:301-306) against the exact reflected-wave pressure of
  the symmetric Riemann problem (exact solver of tests/closed_form.py)
- the inlet, far-field and outflow ghost rules (solver.js:76-118) through the flux of the
  boundary face, with the ghost built from the isentropic stagnation relations
- the spray deposit and apply (spray.js:125-179): a parcel in a merged slave cell gives its
  mass, momentum and energy to the group, which gas + liquid conserve
- the respawn draws (spray.js:43-56): element frequencies against the 256-slot table and
  every drawn angle, radius and diameter against the literal hash values
"""
import numpy as np
import pytest

import oracle as orc
from closed_form import cells_geometry, merge_cells, star_state, state_from_prims, uniform_params

M32 = 0xFFFFFFFF


# ---------------------------------------------------------------- PCG hash
def lcg32(state):
    """PCG's 32-bit LCG step with its default multiplier and increment."""
    return (state * 747796405 + 2891336453) & M32


def rxs_m_xs32(state):
    """PCG's RXS M XS 32/32 output permutation: random xorshift, multiply, fixed xorshift."""
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M32
    return ((word >> 22) ^ word) & M32


def pcg_hash(v):
    return rxs_m_xs32(lcg32(v & M32))


# pcg_hash(v) for these inputs (pcg_hash(0) = 129708002 is the value quoted for the
# published shader function)
PCG_TABLE = {0: 129708002, 1: 2831084092, 2: 2055130248, 3: 2131687100, 7919: 4282043348,
             65535: 131527928, 65536: 688544357, 123456789: 4272394698, 2654435761: 1883031434,
             2147483647: 3919530446, 2147483648: 566699590, 3735928559: 1730779506,
             4294967294: 2155513543, 4294967295: 3861530882}


def spray_rnd(a, n):
    """rnd(a, n) of spray.js:34: the low 24 bits of pcg(a + 7919 n) over 2^24."""
    return (pcg_hash((a + n * 7919) & M32) & 0xFFFFFF) / 16777216.0


def test_pcg_known_answers():
    L = orc.lib()
    for v, want in PCG_TABLE.items():
        assert pcg_hash(v) == want, v
        assert (L.oracle_pcg(v) & M32) == want, v
    rng = np.random.default_rng(7)
    for v in rng.integers(0, 1 << 32, 4096, dtype=np.uint64):
        assert (L.oracle_pcg(int(v)) & M32) == pcg_hash(int(v))
    # the 24-bit uniform: exact in float32, in [0, 1)
    for a in (0, 5, 2654435761, 4294967295):
        for n in range(1, 6):
            got = float(L.oracle_rnd(a & M32, n))
            assert got == spray_rnd(a, n)
            assert 0.0 <= got < 1.0


# ---------------------------------------------------------------- Riemann fluxes
def sound(W):
    r, un, ut, p, G = W[:5]
    return np.sqrt((1 + 1 / G) * p / r)


def cons(W):
    """Conserved (rho, rho un, rho ut, E, rho Z, rho F, rho O) and the physical normal flux."""
    r, un, ut, p, G, Z, F, O = W
    E = p * G + 0.5 * r * (un * un + ut * ut)
    U = np.array([r, r * un, r * ut, E, r * Z, r * F, r * O])
    Fl = np.array([r * un, r * un * un + p, r * un * ut, un * (E + p), r * un * Z, r * un * F, r * un * O])
    return U, Fl


def einfeldt_speeds(WL, WR):
    """HLLE wave-speed estimates (Einfeldt 1988): min / max of the outer characteristic
    speeds of each state and of the Roe-averaged state; G = 1/(gamma - 1) is Roe-averaged
    like the velocities, c^2 = (H - q^2 / 2) / G."""
    rL, uL, vL, pL, GL = WL[:5]
    rR, uR, vR, pR, GR = WR[:5]
    HL = (cons(WL)[0][3] + pL) / rL
    HR = (cons(WR)[0][3] + pR) / rR
    wl, wr = np.sqrt(rL), np.sqrt(rR)
    avg = lambda a, b: (wl * a + wr * b) / (wl + wr)
    u, v, H, G = avg(uL, uR), avg(vL, vR), avg(HL, HR), avg(GL, GR)
    c = np.sqrt(max((H - 0.5 * (u * u + v * v)) / G, 1e-12))
    return min(uL - sound(WL), u - c), max(uR + sound(WR), u + c)


def textbook_flux(WL, WR, w):
    """(1 - w) HLLC + w HLLE in the face-normal frame; returns the 7 conserved fluxes, the
    face speed uf (mass flux over the upwind density) and the upwind G."""
    UL, FL = cons(WL)
    UR, FR = cons(WR)
    SL, SR = einfeldt_speeds(WL, WR)
    if SL >= 0:
        return FL, WL[1], WL[4]
    if SR <= 0:
        return FR, WR[1], WR[4]
    rL, uL, pL = WL[0], WL[1], WL[3]
    rR, uR, pR = WR[0], WR[1], WR[3]
    Ss = (pR - pL + rL * uL * (SL - uL) - rR * uR * (SR - uR)) / (rL * (SL - uL) - rR * (SR - uR))
    if Ss >= 0:
        WK, UK, FK, SK = WL, UL, FL, SL
    else:
        WK, UK, FK, SK = WR, UR, FR, SR
    rK, uK, pK = WK[0], WK[1], WK[3]
    # Toro (10.44): F*K = [S* (SK UK - FK) + SK p* D*] / (SK - S*), D* = (0, 1, 0, S*, 0, 0, 0),
    # with p* = pK + rhoK (SK - uK)(S* - uK) the same from either side
    pstar = pK + rK * (SK - uK) * (Ss - uK)
    D = np.array([0.0, 1.0, 0.0, Ss, 0.0, 0.0, 0.0])
    Fc = (Ss * (SK * UK - FK) + SK * pstar * D) / (SK - Ss)
    # HLLE: the state averaged over the wave fan, then the left Rankine-Hugoniot jump
    Uh = (SR * UR - SL * UL - (FR - FL)) / (SR - SL)
    Fh = FL + SL * (Uh - UL)
    uf = Ss * (SK - uK) / (SK - Ss)
    return (1 - w) * Fc + w * Fh, uf, WK[4]


def oracle_face(F12, normal_r):
    """Oracle face texture (F, FY, FO) -> the 7 fluxes in the normal frame, uf, uf Gf."""
    F = F12[:4].copy()
    if normal_r:
        F[[1, 2]] = F[[2, 1]]
    return np.concatenate([F, [F12[4], F12[5], F12[8]]]), F12[7], F12[6]


def pair_grid(WL, WR, normal_r, ptan=None, solid=None, Nx=6, Nr=6):
    """A fluid grid with state WL on the low side and WR on the high side of the face of
    interest: the x-face (i=3, j=2) between cells (2,2) and (3,2), or the r-face (i=2, j=3)
    between cells (2,2) and (2,3). `ptan` = pressures of the face's four tangential
    neighbours (low side -, low side +, high side -, high side +). First order, so the face
    states are the cells' own."""
    if solid is None:
        solid = np.zeros((Nr, Nx), bool)
    mask, geoA, geoG, geoM = cells_geometry(solid)
    r = np.zeros((Nr, Nx)); un = np.zeros((Nr, Nx)); ut = np.zeros((Nr, Nx)); p = np.zeros((Nr, Nx))
    G = np.zeros((Nr, Nx)); Y = np.zeros((3, Nr, Nx))
    ii, jj = np.meshgrid(np.arange(Nx), np.arange(Nr))
    high = (jj >= 3) if normal_r else (ii >= 3)
    for sel, W in ((~high, WL), (high, WR)):
        r[sel], un[sel], ut[sel], p[sel], G[sel] = W[0], W[1], W[2], W[3], W[4]
        for q in range(3):
            Y[q][sel] = W[5 + q]
    if ptan is not None:
        cells = ([(2, 1), (2, 3), (3, 1), (3, 3)] if not normal_r else [(1, 2), (3, 2), (1, 3), (3, 3)])
        for (ci, cj), pt in zip(cells, ptan):
            p[cj, ci] = pt
    u, v = (ut, un) if normal_r else (un, ut)
    E = p * G + 0.5 * r * (u * u + v * v)
    A = np.stack([r, r * u, r * v, E], -1).astype(np.float32)
    B = np.stack([r * Y[0], r * Y[1], G, r * Y[2]], -1).astype(np.float32)
    return (mask, geoA, geoG, geoM), A, B


def probe_face(WL, WR, normal_r, hll, ptan=None, solid=None):
    geo, A, B = pair_grid(WL, WR, normal_r, ptan, solid)
    Nr, Nx = A.shape[:2]
    P = uniform_params()
    P.order2 = 0
    P.hllBlend = hll
    o = orc.Oracle(Nx, Nr, 64)
    o.set_geometry(*geo)
    o.set_state(A.ravel(), B.ravel())
    _, _, Fx, Fr, _ = o.stage1_probe(P)
    return oracle_face(Fr[3, 2] if normal_r else Fx[2, 3], normal_r), A, B


def f32_state(W):
    """The state as the grid stores it (float32 conserved, then the primitive recovery)."""
    r, un, ut, p, G, Z, F, O = W
    E = np.float32(p * G + 0.5 * r * (un * un + ut * ut))
    r32, m1, m2, G32 = np.float32(r), np.float32(r * un), np.float32(r * ut), np.float32(G)
    rd = float(r32)
    und, utd = float(m1) / rd, float(m2) / rd
    pd = (float(E) - 0.5 * rd * (und * und + utd * utd)) / float(G32)
    Ys = [float(np.float32(r * y)) / rd for y in (Z, F, O)]
    return (rd, und, utd, pd, float(G32), *Ys)


def random_pairs(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        pair = []
        for _ in range(2):
            pair.append((rng.uniform(0.2, 5.0), rng.uniform(-2.5, 2.5), rng.uniform(-1.5, 1.5), rng.uniform(0.1, 5.0),
                         rng.choice([2.5, 5.0, 3.0]), rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0, 1)))
        out.append(pair)
    return out


@pytest.mark.parametrize("normal_r", [False, True])
@pytest.mark.parametrize("hll", [0, 1])
def test_hllc_hlle_random_pairs(normal_r, hll):
    """hllw = 0: the oracle's flux is Toro's HLLC; hllw = 1 (every tangential jump 1 %,
    blend 1e3, so the weight saturates): the HLLE flux, for mass, both momenta, energy and
    the three advected fractions. The face speed and its G flux stay HLLC's."""
    cases = random_pairs(40, 11 + hll + 2 * normal_r)
    seen = set()
    for WL, WR in cases:
        WLs, WRs = f32_state(WL), f32_state(WR)
        ptan = None
        if hll:
            ptan = (WLs[3] * 1.01, WLs[3] * 1.01, WRs[3] * 1.01, WRs[3] * 1.01)
        (F7, uf, ufG), _, _ = probe_face(WL, WR, normal_r, 1e3 if hll else 0.0, ptan)
        want, wuf, wG = textbook_flux(WLs, WRs, 1.0 if hll else 0.0)
        scale = 1.0 + np.max(np.abs(want))
        np.testing.assert_allclose(F7, want, rtol=1e-11, atol=1e-11 * scale)
        assert uf == pytest.approx(wuf, rel=1e-11, abs=1e-11 * scale)
        assert ufG == pytest.approx(wuf * wG, rel=1e-11, abs=1e-11 * scale)
        SL, SR = einfeldt_speeds(WLs, WRs)
        seen.add("left" if SL >= 0 else "right" if SR <= 0 else "fan")
    assert "fan" in seen and len(seen) >= 2      # subsonic fans and a supersonic face


def test_hllc_resolves_a_contact_exactly():
    """An isolated contact (equal u and p, different density and fractions) is an exact
    solution of HLLC: mass flux rho_upwind u, no diffusion; HLLE smears it."""
    WL = (2.0, 0.5, 0.25, 1.0, 2.5, 0.75, 0.5, 0.0)
    WR = (0.5, 0.5, -0.25, 1.0, 2.5, 0.0, 0.0, 0.5)
    (F7, uf, _), _, _ = probe_face(WL, WR, False, 0.0)
    assert F7[0] == pytest.approx(2.0 * 0.5, rel=1e-13)
    assert F7[1] == pytest.approx(2.0 * 0.25 + 1.0, rel=1e-13)
    assert F7[2] == pytest.approx(2.0 * 0.5 * 0.25, rel=1e-13)
    assert F7[4] == pytest.approx(2.0 * 0.5 * 0.75, rel=1e-13)
    assert F7[6] == pytest.approx(0.0, abs=1e-15)
    assert uf == pytest.approx(0.5, rel=1e-13)
    ptan = (1.01, 1.01, 1.01, 1.01)
    (Fh, _, _), _, _ = probe_face(WL, WR, False, 1e3, ptan)
    assert abs(Fh[0] - 1.0) > 0.05            # HLLE diffuses the contact


def smoothstep(e0, e1, x):
    t = min(max((x - e0) / (e1 - e0), 0.0), 1.0)
    return t * t * (3 - 2 * t)


@pytest.mark.parametrize("normal_r", [False, True])
def test_hll_blend_weight(normal_r):
    """The blend weight is smoothstep(0.15, 0.6, blend * jm), jm the largest relative jump
    |p_t - p| / min(p_t, p) between each face cell and its two tangential neighbours; a
    wall neighbour counts as no jump. Constructed jumps put the weight on the ramp, at both
    ends and at a wall; the oracle's flux is then (1 - w) HLLC + w HLLE."""
    WL = (1.0, 0.3, 0.1, 1.0, 2.5, 0.0, 0.0, 0.0)
    WR = (0.6, 0.2, -0.1, 0.8, 2.5, 0.0, 0.0, 0.0)
    WLs, WRs = f32_state(WL), f32_state(WR)
    pL, pR = WLs[3], WRs[3]
    cases = [
        # tangential pressures, blend, expected jm
        ((pL * 1.10, pL, pR, pR), 3.0, 0.10),                 # one jump on the low side
        ((pL, pL / 1.2, pR, pR * 1.05), 2.0, 0.20),            # a drop counts relative to the smaller
        ((pL * 1.04, pL * 1.02, pR * 1.06, pR / 1.05), 5.0, 0.06),
        ((pL, pL, pR, pR), 1.0, 0.0),                          # no jump: pure HLLC
        ((pL * 2, pL, pR, pR), 1.0, 1.0),                      # saturated: pure HLLE
    ]
    for ptan, blend, jm in cases:
        # jm from the pressures the grid's float32 conserved states give back
        sides = (WL, WL, WR, WR)
        ptan32 = [f32_state(tuple(W[:3]) + (pt,) + tuple(W[4:]))[3] for W, pt in zip(sides, ptan)]
        jm_got = max(abs(ptan32[0] - pL) / min(ptan32[0], pL), abs(ptan32[1] - pL) / min(ptan32[1], pL),
                     abs(ptan32[2] - pR) / min(ptan32[2], pR), abs(ptan32[3] - pR) / min(ptan32[3], pR))
        assert jm_got == pytest.approx(jm, rel=1e-5, abs=1e-6)
        w = smoothstep(0.15, 0.6, jm_got * float(np.float32(blend)))
        (F7, _, _), _, _ = probe_face(WL, WR, normal_r, blend, ptan)
        want, _, _ = textbook_flux(WLs, WRs, w)
        np.testing.assert_allclose(F7, want, rtol=1e-10, atol=1e-12)
        if 0 < w < 1:
            # the weight itself, from the mass and energy fluxes
            Fc, _, _ = textbook_flux(WLs, WRs, 0.0)
            Fh, _, _ = textbook_flux(WLs, WRs, 1.0)
            for q in (0, 3):
                assert (F7[q] - Fc[q]) / (Fh[q] - Fc[q]) == pytest.approx(w, rel=1e-8)
    # a solid tangential neighbour (wall texel p = 1) is replaced by the cell's own pressure
    Nx, Nr = 6, 6
    solid = np.zeros((Nr, Nx), bool)
    wall = (2, 3) if not normal_r else (3, 2)
    solid[wall[1], wall[0]] = True
    WL3 = (1.0, 0.3, 0.1, 3.0, 2.5, 0.0, 0.0, 0.0)
    W3s = f32_state(WL3)
    (F7, _, _), _, _ = probe_face(WL3, WR, normal_r, 1.0, None, solid)
    want, _, _ = textbook_flux(W3s, WRs, 0.0)          # every other tangential neighbour equal
    np.testing.assert_allclose(F7, want, rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------- slip-wall pressure
@pytest.mark.parametrize("gamma", [1.4, 1.2])
def test_wall_pressure_exact_reflection(gamma):
    """A uniform stream meets the closed east wall. In stage 1 the wall cell's x momentum
    changes by dt/V (Ax_w F_w - Wx p_w); p_w must be the exact pressure of the symmetric
    Riemann problem (rho, u, p | rho, -u, p): the reflected-shock value for flow into the
    wall, the double-rarefaction value for flow away from it, 0 once the gas would cavitate."""
    Nx, Nr = 8, 5
    solid = np.zeros((Nr, Nx), bool)
    solid[:, Nx - 1] = True
    mask, geoA, geoG, geoM = cells_geometry(solid)
    G = 1.0 / (gamma - 1)
    for u in (0.9, 0.35, 0.05, -0.05, -0.4, -1.2, -8.0):
        rho, p = 0.8, 1.3
        A, B = state_from_prims(np.full((Nr, Nx), rho), np.full((Nr, Nx), u), np.zeros((Nr, Nx)), np.full((Nr, Nx), p),
                                gamma=gamma)
        P = uniform_params(rho=rho, p=p, gamma=gamma)
        P.amb[1] = u
        P.order2 = 0
        o = orc.Oracle(Nx, Nr, 64)
        o.set_geometry(mask, geoA, geoG, geoM)
        o.set_state(A, B)
        A1, _, Fx, _, info = o.stage1_probe(P)
        dt = info[0]
        q = A.reshape(Nr, Nx, 4).astype(np.float64)
        r32, m32, E32 = q[2, Nx - 2, 0], q[2, Nx - 2, 1], q[2, Nx - 2, 3]
        uu = m32 / r32
        pp = (E32 - 0.5 * r32 * uu * uu) / G
        gA = geoA.reshape(Nr + 1, Nx + 1, 4)
        gG = geoG.reshape(Nr, Nx, 4)
        j, i = 2, Nx - 2
        V, Wx, Aw = gG[j, i, 0], gG[j, i, 2], abs(gA[j, i, 2])
        pw = (Aw * Fx[j, i, 1] - (A1[j, i, 1] - m32) * V / dt) / Wx
        c = np.sqrt(gamma * pp / r32)
        if 1 + 0.5 * (gamma - 1) * uu / c <= 0:
            want = 0.0
        else:
            want, us = star_state((r32, uu, pp), (r32, -uu, pp), gamma)
            assert abs(us) < 1e-9
        assert pw == pytest.approx(want, rel=1e-10, abs=1e-12), (u, pw, want)


# ---------------------------------------------------------------- boundary ghosts
def boundary_probe(Q, amb, P, Nx=8, Nr=6, inj=0, B_w=None):
    """Open grid (no solid cells), uniform interior state Q = (rho, u, v, p, G); returns the
    oracle's stage-1 face fluxes."""
    mask, geoA, geoG, geoM = cells_geometry(np.zeros((Nr, Nx), bool))
    r, u, v, p, G = Q
    E = p * G + 0.5 * r * (u * u + v * v)
    A = np.tile(np.array([r, r * u, r * v, E], np.float32), (Nr, Nx, 1))
    B = np.tile(np.array([0.0, 0.0, G, 0.0 if B_w is None else B_w], np.float32), (Nr, Nx, 1))
    P.amb[:] = amb
    P.order2 = 0
    P.hllBlend = 0.0
    o = orc.Oracle(Nx, Nr, 64)
    o.set_geometry(mask, geoA, geoG, geoM)
    o.set_state(A.ravel(), B.ravel())
    return o.stage1_probe(P), A


def stored(Q):
    """Primitive state as the float32 conserved storage returns it."""
    r, u, v, p, G = Q
    W = f32_state((r, u, v, p, G, 0.0, 0.0, 0.0))
    return W


@pytest.mark.parametrize("side", ["top", "right"])
def test_far_field_and_outflow_ghosts(side):
    """Top boundary (far field, solver.js:100-108) and right boundary (outflow, 109-118):
    an outflowing supersonic state leaves through a copy of itself, a subsonic outflow
    through itself at the ambient pressure, an inflow takes the ambient state (at the right
    boundary with its u clamped to >= 0). Each branch is pinned through the boundary face's
    flux = HLLC(interior, ghost); a uniform supersonic outflow is preserved."""
    G = 2.5
    amb = (0.7, 0.15, -0.05, 0.9)
    P = uniform_params(rho=amb[0], p=amb[3])
    Nx, Nr = 8, 6
    amb = tuple(float(np.float32(x)) for x in amb)   # the parameter block is float32
    ambW = (amb[0], amb[1], amb[2], amb[3], G)        # ghost G = inlet[3] = G of the free stream
    for name, normal in (("supersonic", 2.4), ("subsonic", 0.4), ("inflow", -0.3), ("at rest", 0.0)):
        if side == "top":
            Q = (1.2, 0.35, normal, 1.1, G)
        else:
            Q = (1.2, normal, 0.2, 1.1, G)
        (A1, _, Fx, Fr, _), A = boundary_probe(Q, amb, P)
        Qs = stored(Q)
        c = np.sqrt((1 + 1 / G) * Qs[3] / Qs[0])
        un = Qs[2] if side == "top" else Qs[1]
        if un >= c:
            ghost = Qs[:5]
        elif un > 0:
            ghost = (Qs[0], Qs[1], Qs[2], amb[3], G)
        else:
            ghost = ambW if side == "top" else (amb[0], max(amb[1], 0.0), amb[2], amb[3], G)
        if side == "top":
            F12 = Fr[Nr, 3]
            WLn = (Qs[0], Qs[2], Qs[1], Qs[3], G, 0, 0, 0)
            WRn = (ghost[0], ghost[2], ghost[1], ghost[3], ghost[4], 0, 0, 0)
        else:
            F12 = Fx[2, Nx]
            WLn = (Qs[0], Qs[1], Qs[2], Qs[3], G, 0, 0, 0)
            WRn = (ghost[0], ghost[1], ghost[2], ghost[3], ghost[4], 0, 0, 0)
        F7, _, _ = oracle_face(F12, side == "top")
        want, _, _ = textbook_flux(WLn, WRn, 0.0)
        np.testing.assert_allclose(F7, want, rtol=1e-11, atol=1e-12, err_msg=f"{side} {name}")
    # a uniform supersonic stream along the axis (v = 0: a uniform radial velocity is no
    # axisymmetric solution) enters through the free stream and leaves through both ghosts
    # unchanged
    mask, geoA, geoG, geoM = cells_geometry(np.zeros((Nr, Nx), bool))
    P.amb[:] = (1.25, 2.5, 0.0, 1.125)               # exact in float32, as the stored state
    P.order2 = 1
    P.hllBlend = 1.0
    A, B = state_from_prims(np.full((Nr, Nx), 1.25), np.full((Nr, Nx), 2.5), np.zeros((Nr, Nx)), np.full((Nr, Nx), 1.125))
    o = orc.Oracle(Nx, Nr, 64)
    o.set_geometry(mask, geoA, geoG, geoM)
    o.set_state(A, B)
    o.step(5, P)
    S = o.state_f64()[0]
    np.testing.assert_allclose(S, A.astype(np.float64), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("absorb", [0.0, 1.0])
@pytest.mark.parametrize("ramp", [0.0, 5.0])
def test_inlet_ghost_isentropic(absorb, ramp):
    """Chamber inflow (solver.js:81-99): the ghost is the stagnation state (P0, T0) expanded
    isentropically to the inflow speed u, h0 = cp T + u^2 / 2 with cp = gamma R / (gamma - 1),
    p = P0 (T / T0)^(gamma / (gamma - 1)), rho = p / (R T), carrying pure exhaust. u is the
    interior speed clamped to [0, c0 / 2] (absorption 0) or the slow inflow average
    ubar * smoothstep(1, 2, P0 / pStart) (absorption 1); P0 ramps from pStart at t = 0."""
    gamma, Rg, T0, P0, pStart = 1.25, 1.3, 3.0, 6.0, 2.5
    G = 1 / (gamma - 1)
    P = uniform_params(rho=1.0, p=1.0)
    P.res[:] = [P0, T0, Rg]
    P.inlet[:] = [0.0, absorb, G, 2.5]
    P.pStart = pStart
    P.ramp = ramp
    P.RcCells = 4.0
    ubar = 0.6
    Nx, Nr = 8, 6
    Q = (1.5, 0.45, 0.0, 1.2, G)
    (A1, B1, Fx, _, _), A = boundary_probe(Q, (1.0, 0.0, 0.0, 1.0), P, B_w=ubar)
    Qs = stored(Q)
    P0n = float(np.float32(pStart)) if ramp > 0 else float(np.float32(P0))
    T0f, Rf, Gf = (float(np.float32(x)) for x in (T0, Rg, G))
    gf = 1 + 1 / Gf
    c0 = np.sqrt(gf * Rf * T0f)
    if absorb == 0:
        u = min(max(Qs[1], 0.0), 0.5 * c0)
    else:
        u = float(np.float32(ubar)) * smoothstep(1.0, 2.0, P0n / float(np.float32(pStart)))
    cp = gf * Rf / (gf - 1)
    T = T0f - u * u / (2 * cp)
    p = P0n * (T / T0f) ** (gf / (gf - 1))
    ghost = (p / (Rf * T), u, 0.0, p, Gf, 1.0, 1.0, 0.0)
    for j in range(Nr):
        F7, uf, _ = oracle_face(Fx[j, 0], False)
        WR = (Qs[0], Qs[1], Qs[2], Qs[3], Qs[4], 0.0, 0.0, ubar / Qs[0])
        if j + 0.5 < 4.0:
            want, _, _ = textbook_flux(ghost, WR, 0.0)
        else:   # above the chamber radius: the free stream
            want, _, _ = textbook_flux((1.0, 0.0, 0.0, 1.0, 2.5, 0, 0, 0), WR, 0.0)
        np.testing.assert_allclose(F7[:6], want[:6], rtol=1e-10, atol=1e-12, err_msg=f"row {j}")


# ---------------------------------------------------------------- spray deposit / apply
def spray_box(parcel, gas_u=0.0, gas_v=0.0, NP=64):
    """Closed box of hot gas with cell (4, 2) merged into its east neighbour (5, 2); one
    parcel in ring slot NP - 1 (the ring spawns massless parcels elsewhere). Returns the
    oracle after one step, and the same run without the parcel."""
    Nx, Nr = 12, 6
    solid = np.zeros((Nr, Nx), bool)
    solid[:, 0] = solid[:, Nx - 1] = True
    solid[Nr - 1, :] = True
    mask, geoA, geoG, geoM = cells_geometry(solid)
    merge_cells(geoG, geoM, Nx, Nr, (4, 2), (5, 2))
    rho, p = 0.25, 1.0          # T = 1152.6 K: the drop evaporates
    A, B = state_from_prims(np.full((Nr, Nx), rho), np.full((Nr, Nx), gas_u), np.full((Nr, Nx), gas_v),
                            np.full((Nr, Nx), p))
    P = uniform_params(rho=rho, p=p)
    P.inj = 1
    P.sprayMhat = 0.0
    P.sprayTScale = 0.1 / 287.6
    P.fuelVec[:] = (800.0, 400.0, 3e5, 0.02)
    P.oxVec[:] = (1100.0, 300.0, 2e5, 0.015)
    P.eF, P.eO, P.Ge = 0.7, 0.3, 4.0
    runs = []
    for with_parcel in (True, False):
        o = orc.Oracle(Nx, Nr, 64)
        o.set_geometry(mask, geoA, geoG, geoM)
        o.set_parcels(NP)
        o.set_injector(np.zeros(256 * 3 * 4, np.float32))
        o.set_state(A, B)
        if with_parcel:
            o.set_parcel(NP - 1, *parcel)
        runs.append(o)
    return runs, (Nx, Nr, geoG.reshape(Nr, Nx, 4), mask.reshape(Nr, Nx)), P


@pytest.mark.parametrize("ptype", [0.0, 1.0])
@pytest.mark.parametrize("gas", [(0.0, 0.0), (0.2, 0.05)])
def test_spray_deposit_into_merged_group(ptype, gas):
    """A fuel (type 0) or oxidiser (type 1) drop in the merged slave cell (4, 2): what it
    gives up goes to its group (host (5, 2), group volume V) and both members take the same
    state. Against the same step without the drop: rho += dm/V; momentum += dp/V with
    dp = m_old u_old - m_new u_new; E += (dm_f eF + dm_o eO + u_gas . dp) / V, u_gas the
    group's velocity before the apply; G -> mix(G, Ge, min(dm / (V rho_new), 1));
    (rho Z, rho F, rho O) += (dm, dm_f, dm_o) / V. Gas + liquid mass and momentum are
    conserved, every other cell is unchanged."""
    NP = 64
    a0 = np.array([4.3, 2.6, 0.45, -0.12], np.float32)
    b0 = np.array([20e-6, 1e-3, ptype, 0.0], np.float32)
    (o, o0), (Nx, Nr, gG, solid), P = spray_box((a0, b0), *gas, NP=NP)
    Pa, Pb = o.parcels(NP)
    for x in (o, o0):
        x.step(1, P)
    SA, SB = (s.reshape(Nr, Nx, 4) for s in o.state_f64())
    RA, RB = (s.reshape(Nr, Nx, 4) for s in o0.state_f64())
    PA, PB = (x.reshape(-1, 4)[NP - 1] for x in o.parcels_f64(NP))
    m0, m1 = float(b0[1]), PB[1]
    u0 = a0[2:].astype(np.float64)
    dm = m0 - m1
    assert dm > 1e-6 * m0                               # it evaporated
    dp = m1 * (u0 - PA[2:]) + dm * u0
    assert np.all(np.abs(dp) > 0)
    V = float(gG[2, 5, 0])
    assert V == 5.0 and float(gG[2, 4, 0]) == 5.0     # the group volume, on both members
    A, B = RA[2, 5], RB[2, 5]
    vel = A[1:3] / A[0]
    dmf, dmo = dm * (1 - ptype), dm * ptype
    rn = A[0] + dm / V
    want_A = np.array([rn, A[1] + dp[0] / V, A[2] + dp[1] / V,
                       A[3] + (dmf * P.eF + dmo * P.eO + vel @ dp) / V])
    phi = min(dm / V / rn, 1.0)
    want_B = np.array([B[0] + dm / V, B[1] + dmf / V, B[2] * (1 - phi) + float(np.float32(P.Ge)) * phi, B[3] + dmo / V])
    for cell in ((5, 2), (4, 2)):
        np.testing.assert_allclose(SA[cell[1], cell[0]], want_A, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(SB[cell[1], cell[0]], want_B, rtol=1e-12, atol=1e-15)
    other = np.ones((Nr, Nx), bool)
    other[2, 4] = other[2, 5] = False
    assert np.array_equal(SA[other], RA[other]) and np.array_equal(SB[other], RB[other])
    # gas + liquid conserve mass and momentum (the group counted once, at its host)
    d_gas = (SA[2, 5] - RA[2, 5]) * V
    assert d_gas[0] == pytest.approx(dm, rel=1e-12)
    np.testing.assert_allclose(d_gas[1:3], m0 * u0 - m1 * PA[2:], rtol=1e-12)


# ---------------------------------------------------------------- respawn draws
def test_respawn_draws():
    """The ring spawns 4 parcels per step at its head (seed 1, 2, ... per step). Each draws
    its element from the 256-slot table with t = floor(rnd(id, 1) 255.999), id = idx xor
    (seed 2654435761), then angle e1.z + (2 rnd(id, 2) - 1) e2.z, radius
    max(e1.y + (rnd(id, 3) - 1/2) e2.w, 0.05), diameter e2.y exp(1.4 (rnd(id, 4) - 1/2)).
    Every drawn parcel matches these with the literal hash; the element frequencies follow
    the table's slot counts and every draw stays inside its jitter bounds."""
    NP, steps = 1024, 600
    Nx, Nr = 16, 8
    solid = np.zeros((Nr, Nx), bool)
    solid[:, 0] = solid[:, Nx - 1] = True
    solid[Nr - 1, :] = True
    mask, geoA, geoG, geoM = cells_geometry(solid)
    A, B = state_from_prims(np.full((Nr, Nx), 1.0), np.zeros((Nr, Nx)), np.zeros((Nr, Nx)), np.full((Nr, Nx), 1.0))
    P = uniform_params()
    P.inj = 1
    P.sprayMhat = 1e-4
    P.sprayTScale = 1e-6
    P.fuelVec[:] = (800.0, 2000.0, 3e5, 0.02)     # boiling point above the gas temperature
    P.oxVec[:] = (1100.0, 2000.0, 2e5, 0.015)
    # three elements over 256 slots: 40, 100 and 116 slots (a largest-remainder table)
    counts = (40, 100, 116)
    elem = np.repeat(np.arange(3), counts)
    tab = np.zeros((3, 256, 4), np.float32)
    tab[0, :, 0] = 2.0 + elem                      # x (identifies the element)
    tab[0, :, 1] = np.array([1.5, 3.0, 0.03])[elem]  # r (the third hits the 0.05 floor)
    tab[0, :, 2] = np.array([0.2, -0.3, 0.6])[elem]  # angle
    tab[0, :, 3] = np.array([0.5, 0.8, 0.3])[elem]   # speed
    tab[1, :, 0] = np.array([0.0, 1.0, 0.0])[elem]   # type
    tab[1, :, 1] = np.array([1e-5, 2e-5, 5e-5])[elem]
    tab[1, :, 2] = np.array([0.1, 0.05, 0.2])[elem]  # angle jitter
    tab[1, :, 3] = np.array([0.4, 1.0, 0.2])[elem]   # radius jitter
    o = orc.Oracle(Nx, Nr, 64)
    o.set_geometry(mask, geoA, geoG, geoM)
    o.set_parcels(NP)
    o.set_injector(tab.ravel())
    o.set_state(A, B)
    drawn = np.zeros(3, int)
    for s in range(steps):
        o.step(1, P)
        dt = float(o.step_info()[0])
        PA, PB = (x.reshape(-1, 4) for x in o.parcels_f64(NP))
        seed = 1 + s
        for k in range(4):
            idx = (4 * s + k) % NP
            pid = (idx ^ ((seed * 2654435761) & M32)) & M32
            t = int(np.float32(spray_rnd(pid, 1)) * 255.999)
            e = elem[t]
            drawn[e] += 1
            ang = float(tab[0, t, 2]) + (spray_rnd(pid, 2) - 0.5) * 2 * float(tab[1, t, 2])
            rr = max(float(tab[0, t, 1]) + (spray_rnd(pid, 3) - 0.5) * float(tab[1, t, 3]), 0.05)
            d = float(tab[1, t, 1]) * np.exp((spray_rnd(pid, 4) - 0.5) * 1.4)
            a, b = PA[idx], PB[idx]
            assert a[0] == float(tab[0, t, 0]), (s, k)
            assert a[1] == pytest.approx(rr, rel=1e-7)
            assert a[2] == pytest.approx(float(tab[0, t, 3]) * np.cos(ang), rel=1e-6, abs=1e-7)
            assert a[3] == pytest.approx(float(tab[0, t, 3]) * np.sin(ang), rel=1e-6, abs=1e-7)
            assert b[0] == pytest.approx(d, rel=1e-6)
            assert b[1] == pytest.approx(float(np.float32(P.sprayMhat)) * dt / 4, rel=1e-6)
            assert b[2] == float(tab[1, t, 0])
            # jitter bounds
            assert abs(ang - float(tab[0, t, 2])) <= float(tab[1, t, 2]) * (1 + 1e-6)
            assert abs(a[1] - float(tab[0, t, 1])) <= 0.5 * float(tab[1, t, 3]) * (1 + 1e-6) or a[1] == pytest.approx(0.05)
            assert np.exp(-0.7) * (1 - 1e-6) <= b[0] / float(tab[1, t, 1]) <= np.exp(0.7) * (1 + 1e-6)
    n = 4 * steps
    freq = drawn / n
    want = np.array(counts) / 256.0
    sigma = np.sqrt(want * (1 - want) / n)
    assert np.all(np.abs(freq - want) < 4 * sigma), (freq, want)


def lognormal_diameter(d_med, sigma, pid):
    """The lognormal option's draw: Box-Muller on rnd 4 and 6 of the literal hash."""
    u1, u2 = 1.0 - spray_rnd(pid, 4), spray_rnd(pid, 6)
    return d_med * np.exp(sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * np.pi * u2))


def test_lognormal_drop_size_option():
    """The drop-size option of the injector table (third row y = sigma > 0): each spawned
    diameter is d_med exp(sigma N), N the Box-Muller normal deviate of the literal hash's
    draws 4 and 6; over the sample, ln(d / d_med) has mean 0 and standard deviation sigma.
    The setup writes the option: median at dropD0Frac x the orifice, sigma in the table."""
    from rnozzle import configs
    from rnozzle.sim_setup import SimSetup
    s = SimSetup(configs.kerolox_injector(Nx=160, dropD0Frac=0.1, dropSigma=0.3))
    tab = s.injector.reshape(3, 256, 4)
    Do = 1.5e-3
    film = tab[1, :, 1] == np.float32(80e-6)
    assert np.all(tab[1, ~film, 1] == np.float32(0.1 * Do)) and np.all(tab[2, :, 1] == np.float32(0.3))
    ref = SimSetup(configs.kerolox_injector(Nx=160)).injector.reshape(3, 256, 4)
    assert np.all(ref[2, :, 1] == 0)                      # the reference's tables leave it 0
    NP, steps, sigma = 1024, 500, 0.3
    Nx, Nr = 16, 8
    solid = np.zeros((Nr, Nx), bool)
    solid[:, 0] = solid[:, Nx - 1] = True
    solid[Nr - 1, :] = True
    mask, geoA, geoG, geoM = cells_geometry(solid)
    A, B = state_from_prims(np.full((Nr, Nx), 1.0), np.zeros((Nr, Nx)), np.zeros((Nr, Nx)), np.full((Nr, Nx), 1.0))
    P = uniform_params()
    P.inj = 1
    P.sprayMhat = 1e-4
    P.sprayTScale = 1e-6
    P.fuelVec[:] = (800.0, 2000.0, 3e5, 0.02)
    P.oxVec[:] = (1100.0, 2000.0, 2e5, 0.015)
    t = np.zeros((3, 256, 4), np.float32)
    t[0, :, 0], t[0, :, 1], t[0, :, 3] = 3.0, 2.0, 0.4
    t[1, :, 1] = 2e-5
    t[2, :, 1] = sigma
    o = orc.Oracle(Nx, Nr, 64)
    o.set_geometry(mask, geoA, geoG, geoM)
    o.set_parcels(NP)
    o.set_injector(t.ravel())
    o.set_state(A, B)
    logs = []
    for s_ in range(steps):
        o.step(1, P)
        _, PB = (x.reshape(-1, 4) for x in o.parcels_f64(NP))
        for k in range(4):
            idx = (4 * s_ + k) % NP
            pid = (idx ^ (((1 + s_) * 2654435761) & M32)) & M32
            want = lognormal_diameter(float(np.float32(2e-5)), float(np.float32(sigma)), pid)
            assert PB[idx, 0] == pytest.approx(want, rel=1e-6)
            logs.append(np.log(PB[idx, 0] / float(np.float32(2e-5))))
    logs = np.array(logs)
    n = len(logs)
    assert abs(np.mean(logs)) < 4 * sigma / np.sqrt(n)
    assert abs(np.std(logs) - sigma) < 4 * sigma / np.sqrt(2 * n)
