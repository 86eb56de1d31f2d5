// oracle.cpp — CPU restatement of the reference solver's per-step GPU passes.
//
// TEST INFRASTRUCTURE ONLY. This is the checker the HIP kernels are compared
// against (parity tests, smoke check, and the CPU baseline timing in bench.py).
// The product path never links, loads or calls it.
//
// It follows the reference pass by pass, with the same intermediate "textures"
// (padded primitives, face-flux arrays, T1/T2, three rotating state copies),
// so a reader can hold each loop next to the GLSL it restates:
//   toPrim / soundSpeed         solver.js:27-40
//   FS_PRIM                     solver.js:56-126   -> prim_pass
//   FS_FLUX_X / FS_FLUX_R       solver.js:133-259  -> flux_pass
//   FS_UPDATE                   solver.js:264-363  -> update_pass
//   FS_REDUCE1/2, FS_STEPINFO   solver.js:368-411  -> cfl
//   FS_RESCALE                  solver.js:418-437  -> rescale
//   GLSL_GAS + FS_VCELL         solver.js:445-489  -> vcell_pass
//   FS_VFLUX_X / FS_VFLUX_R     solver.js:490-539  -> vflux_pass
//   FS_SOURCE                   solver.js:540-663  -> source_pass
//   FS_WALLROW                  solver.js:665-674  -> wall_row
//   FS_DERIVED                  render.js:60-134   -> derived
//   GPUSolver.step schedule     solver.js:849-891  -> step
//   FS_SPRAY_UPDATE             spray.js:19-122    -> spray_update
//   VS/FS_SPRAY_DEP             spray.js:125-149   -> spray_deposit (serial, in parcel order)
//   FS_SPRAY_APPLY              spray.js:150-179   -> spray_apply
//   GPUSpray.step / reset       spray.js:242-297
// Templated on the arithmetic type: float reproduces the reference's FP32
// arithmetic, double is the FP64 error reference. OpenMP over rows.
//
// How each pass is pinned (tests/test_oracle_closed_form.py, tests/test_oracle_pins.py,
// tests/test_setup_golden.py):
//   prim / flux / update    exact Riemann solutions (Sod, Toro test 2), closed-box
//                           conservation, free-stream preservation on cut cells
//   face flux               Toro's HLLC (second form) and the Einfeldt HLLE flux on random
//                           state pairs at blend weight 0 and 1; the weight on constructed
//                           tangential pressure jumps (ramp, both ends, wall neighbour);
//                           an isolated contact kept exactly
//   wall pressure           the exact reflected-shock / double-rarefaction pressure of the
//                           symmetric Riemann problem, and the cavitation floor
//   ghost cells             inlet from the isentropic stagnation relations (ramp, absorbing
//                           speed); far field and outflow, every branch, through the
//                           boundary face's flux; a supersonic stream kept uniform
//   CFL / step info         brute force over the definition
//   inlet + whole step      choked mass flow against Pc At / c*
//   vcell / vflux           axisymmetric Couette: momentum mu a / r and energy 2 mu a^2
//                           exactly, with Sutherland's air viscosity (ISA table value)
//   source: wall model      Werner-Wengle law at the Eckert reference temperature, both
//                           branches
//   source: afterburning    heat = fuel burnt x heat of combustion; mixing limit = shear
//                           rate; Arrhenius limit linear in p with Ta = 9000 K;
//                           fuel- / air-limited availability
//   source: separation      the Zukoski plateau rise 1 + M/2 as the onset, the 1.2x
//                           ramp, the 4- and 6-cell probe reach, momentum removed at the
//                           rate rho |u| u; the subsonic band damping
//   spray deposit / apply   a drop in a merged slave cell: mass, both momenta, energy
//                           (latent and drag work), G mixing and fractions into its group;
//                           gas + liquid mass and momentum conserved
//   spray respawn           every draw against a literal table of the published pcg_hash;
//                           element frequencies against the slot counts; jitter bounds
//   spray update            PCG hash (permutation, uniformity), Stokes drag relaxation,
//                           d2-law evaporation with the Spalding number, mass balance;
//                           break-up to the Weber-12 diameter with its time scale
//                           (d / U) sqrt(rho_l / rho_g); impingement fan (speed kept,
//                           +-25 degrees, d / 4); axis mirror and 0.3 restitution
//   rescale                 ambient cells take p fp, T fT at the same gas constant and
//                           u + du; exhaust cells (Z >= 0.03) unchanged; monotone between;
//                           mass fractions and G kept
//   wall row                (u, v, p, sqrt(gamma p / rho)) of the named near-wall cell
//   derived                 closed forms on constructed states; field 9 against Prandtl's
//                           mixing length mu_t = rho l^2 |du/dr|; field 13 against the
//                           Arrhenius burn rate and the fuel the step burns
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "nozzle.h"

namespace {

template <class T>
struct v4 {
  T x, y, z, w;
};
template <class T>
inline v4<T> mk(T a, T b, T c, T d) { return v4<T>{a, b, c, d}; }
template <class T>
inline v4<T> load4(const std::vector<T>& a, size_t k) { return v4<T>{a[k * 4], a[k * 4 + 1], a[k * 4 + 2], a[k * 4 + 3]}; }
template <class T>
inline void store4(std::vector<T>& a, size_t k, const v4<T>& v) { a[k * 4] = v.x; a[k * 4 + 1] = v.y; a[k * 4 + 2] = v.z; a[k * 4 + 3] = v.w; }

template <class T>
inline T clampT(T x, T lo, T hi) { return std::min(std::max(x, lo), hi); }
template <class T>
inline T smoothstepT(T e0, T e1, T x) {
  const T t = clampT<T>((x - e0) / (e1 - e0), T(0), T(1));
  return t * t * (T(3) - T(2) * t);
}
template <class T>
inline T mixT(T a, T b, T t) { return a * (T(1) - t) + b * t; }

template <class T>
struct Prim {
  v4<T> w;  // rho, u, v, p
  v4<T> s;  // Z, F, G, O
};

// toPrim (solver.js:33-39)
template <class T>
inline Prim<T> toPrim(const v4<T>& a, const v4<T>& b) {
  const T RHO_MIN = T(1e-7), P_MIN = T(1e-7);
  const T rho = std::max(a.x, RHO_MIN);
  const T u = a.y / rho, v = a.z / rho;
  const T G = std::max(b.z, T(0.5));
  const T p = std::max((a.w - T(0.5) * rho * (u * u + v * v)) / G, P_MIN);
  Prim<T> q;
  q.w = mk(rho, u, v, p);
  q.s = mk(clampT(b.x / rho, T(0), T(1)), clampT(b.y / rho, T(0), T(1)), G, clampT(b.w / rho, T(0), T(1)));
  return q;
}
template <class T>
inline T soundSpeed(const Prim<T>& q) { return std::sqrt((T(1) + T(1) / q.s.z) * q.w.w / q.w.x); }

template <class T>
inline T Tp(float v) { return (T)v; }

struct Injector {
  std::vector<float> tab;  // 256 x 3 RGBA
};

template <class T>
struct Oracle {
  int Nx = 0, Nr = 0;
  // textures
  std::vector<uint8_t> mask;
  std::vector<T> geoA, geoG, geoM, jWall;
  std::vector<T> SA[3], SB[3];
  std::vector<T> W, Wb;                        // padded (Nx+4) x (Nr+4)
  std::vector<T> FxA, FxB, FxC, FrA, FrB, FrC;  // (Nx+1) x (Nr+1)
  std::vector<T> T1, T2;
  std::vector<T> wallRow;
  T step_[2][4];
  int stepIdx = 0, cur = 0;
  int64_t stepCount = 0;
  // spray
  int NP = 65536;
  std::vector<T> PA[2], PB[2], PC[2], src;
  std::vector<T> tab;
  int spCur = 0;
  uint32_t head = 0, seed = 1;

  void resize(int nx, int nr) {
    Nx = nx;
    Nr = nr;
    const size_t N = (size_t)Nx * Nr, NL = (size_t)(Nx + 1) * (Nr + 1), NPd = (size_t)(Nx + 4) * (Nr + 4);
    mask.assign(N, 0);
    geoA.assign(NL * 4, T(0));
    geoG.assign(N * 4, T(0));
    geoM.assign(N * 4, T(0));
    jWall.assign((size_t)Nx * 4, T(0));
    for (int k = 0; k < 3; ++k) { SA[k].assign(N * 4, T(0)); SB[k].assign(N * 4, T(0)); }
    W.assign(NPd * 4, T(0));
    Wb.assign(NPd * 4, T(0));
    for (auto* f : {&FxA, &FxB, &FxC, &FrA, &FrB, &FrC}) f->assign(NL * 4, T(0));
    T1.assign(N * 4, T(0));
    T2.assign(N * 4, T(0));
    wallRow.assign((size_t)Nx * 4, T(0));
    std::memset(step_, 0, sizeof(step_));
    stepIdx = 0;
    cur = 0;
    src.assign(N * 4, T(0));
    tab.assign(256 * 3 * 4, T(0));
    spray_alloc();
  }
  void spray_alloc() {
    for (int k = 0; k < 2; ++k) {
      PA[k].assign((size_t)NP * 4, T(0));
      PB[k].assign((size_t)NP * 4, T(0));
      PC[k].assign((size_t)NP * 4, T(0));
    }
    spray_reset();
  }
  // GPUSpray.reset (spray.js:242-254)
  void spray_reset() {
    for (int k = 0; k < 2; ++k) {
      for (int p = 0; p < NP; ++p) {
        PA[k][(size_t)p * 4] = T(-1);
        PA[k][(size_t)p * 4 + 1] = T(-1);
        PA[k][(size_t)p * 4 + 2] = T(0);
        PA[k][(size_t)p * 4 + 3] = T(0);
      }
      std::fill(PB[k].begin(), PB[k].end(), T(0));
      std::fill(PC[k].begin(), PC[k].end(), T(0));
    }
    head = 0;
  }
  bool solidAt(int i, int j) const { return mask[(size_t)j * Nx + i] != 0; }

  // ---------------- FS_PRIM (solver.js:56-126)
  void prim_pass(const std::vector<T>& A, const std::vector<T>& B, int timeSel, const nz_params& P) {
    const int PW = Nx + 4, PH = Nr + 4;
    const T* st = step_[stepIdx];
#pragma omp parallel for schedule(static)
    for (int pj = 0; pj < PH; ++pj)
      for (int pi = 0; pi < PW; ++pi) {
        int i = pi - 2, j = pj - 2;
        const bool mir = j < 0;
        if (mir) j = -j - 1;
        Prim<T> q;
        q.w = mk(Tp<T>(P.amb[0]), Tp<T>(P.amb[1]), Tp<T>(P.amb[2]), Tp<T>(P.amb[3]));
        q.s = mk(T(0), T(0), Tp<T>(P.inlet[3]), T(0));
        bool wall = false;
        if (i < 0) {
          if (j < Nr && solidAt(0, j)) wall = true;
          else if (T(j) + T(0.5) < Tp<T>(P.RcCells) && P.inj) wall = true;
          else if (T(j) + T(0.5) < Tp<T>(P.RcCells)) {
            const T t = timeSel == 1 ? st[1] : st[2];
            const T f = Tp<T>(P.ramp) > T(0) ? smoothstepT(T(0), Tp<T>(P.ramp), t) : T(1);
            const T P0 = mixT(Tp<T>(P.pStart), Tp<T>(P.res[0]), f);
            const T T0 = Tp<T>(P.res[1]), R = Tp<T>(P.res[2]), G = Tp<T>(P.inlet[2]), g = T(1) + T(1) / G;
            const v4<T> a = load4(A, (size_t)j * Nx + 0);
            const T ubar = std::max(B[((size_t)j * Nx + 0) * 4 + 3], T(0));
            const T c0 = std::sqrt(g * R * T0);
            const T ui = clampT(a.y / std::max(a.x, T(1e-7)), T(0), T(0.5) * c0);
            const T ut = ubar * smoothstepT(T(1), T(2), P0 / std::max(Tp<T>(P.pStart), T(1e-6)));
            const T u = mixT(ui, ut, Tp<T>(P.inlet[1]));
            const T TT = T0 - u * u * (g - T(1)) / (T(2) * g * R);
            const T p = P0 * std::pow(TT / T0, g * G);
            q.w = mk(p / (R * TT), u, T(0), p);
            q.s = mk(T(1), T(1), G, T(0));
          }
        } else if (j >= Nr) {
          const int ci = std::min(i, Nx - 1), cj = Nr - 1;
          if (!solidAt(ci, cj)) {
            const Prim<T> t = toPrim(load4(A, (size_t)cj * Nx + ci), load4(B, (size_t)cj * Nx + ci));
            const T cs = soundSpeed(t);
            if (t.w.z >= cs) q = t;
            else if (t.w.z > T(0)) { q = t; q.w.w = Tp<T>(P.amb[3]); }
          }
        } else if (i >= Nx) {
          const int ci = Nx - 1, cj = j;
          if (!solidAt(ci, cj)) {
            const Prim<T> t = toPrim(load4(A, (size_t)cj * Nx + ci), load4(B, (size_t)cj * Nx + ci));
            const T cs = soundSpeed(t);
            if (t.w.y >= cs) q = t;
            else if (t.w.y > T(0)) { q = t; q.w.w = Tp<T>(P.amb[3]); }
            else q.w.y = std::max(q.w.y, T(0));
          }
        } else {
          if (solidAt(i, j)) wall = true;
          else q = toPrim(load4(A, (size_t)j * Nx + i), load4(B, (size_t)j * Nx + i));
        }
        const size_t o = (size_t)pj * PW + pi;
        if (wall) {
          store4(W, o, mk(T(-1), T(0), T(0), T(1)));
          store4(Wb, o, mk(T(0), T(0), T(0), T(0)));
          continue;
        }
        if (mir) q.w.z = -q.w.z;
        store4(W, o, q.w);
        store4(Wb, o, q.s);
      }
  }

  // padded fetch helpers (cell coordinates, -2..N+1)
  v4<T> Wat(int i, int j) const { return load4(W, (size_t)(j + 2) * (Nx + 4) + (i + 2)); }
  v4<T> Wbat(int i, int j) const { return load4(Wb, (size_t)(j + 2) * (Nx + 4) + (i + 2)); }

  // vanLeer limiter, one component (solver.js:156-160)
  static T vanLeer(T a, T b) {
    const T ab = a * b;
    T den = a + b;
    if (std::fabs(den) < T(1e-30)) den = T(1);
    return T(2) * std::max(ab, T(0)) / den;
  }
  static T jumpRatio(T a, T b) { return std::fabs(a - b) / std::min(a, b); }

  // ---------------- FS_FLUX_X / FS_FLUX_R (solver.js:133-259)
  void flux_pass(const nz_params& P) {
    for (int dirR = 0; dirR < 2; ++dirR) {
      const int FW = dirR ? Nx : Nx + 1, FH = dirR ? Nr + 1 : Nr;
      std::vector<T>& oF = dirR ? FrA : FxA;
      std::vector<T>& oFY = dirR ? FrB : FxB;
      std::vector<T>& oFO = dirR ? FrC : FxC;
#pragma omp parallel for schedule(static)
      for (int j = 0; j < FH; ++j)
        for (int i = 0; i < FW; ++i) {
          const size_t fo = (size_t)j * (Nx + 1) + i;
          store4(oF, fo, mk(T(0), T(0), T(0), T(0)));
          store4(oFY, fo, mk(T(0), T(0), T(0), T(0)));
          store4(oFO, fo, mk(T(0), T(0), T(0), T(0)));
          const v4<T> geo = load4(geoA, fo);
          T area;
          int dx_, dy_, tx_, ty_;
          if (dirR) {
            area = geo.w;
            if (j == 0 || area == T(0)) continue;
            dx_ = 0; dy_ = 1; tx_ = 1; ty_ = 0;
          } else {
            area = geo.z;
            if (area == T(0)) continue;
            dx_ = 1; dy_ = 0; tx_ = 0; ty_ = 1;
          }
          auto cellP = [&](int ci, int cj, bool& s) {
            Prim<T> q;
            q.w = Wat(ci, cj);
            q.s = Wbat(ci, cj);
            s = q.w.x < T(0);
            if (dirR) std::swap(q.w.y, q.w.z);
            return q;
          };
          auto pAt = [&](int ci, int cj, T fb) {
            const v4<T> w = Wat(ci, cj);
            return w.x < T(0) ? fb : w.w;
          };
          auto mirrorN = [](Prim<T> q) { q.w.y = -q.w.y; return q; };
          bool s0, s1, s2, s3;
          const int k0x = i - 2 * dx_, k0y = j - 2 * dy_, k1x = i - dx_, k1y = j - dy_, k3x = i + dx_, k3y = j + dy_;
          Prim<T> c0 = cellP(k0x, k0y, s0), c1 = cellP(k1x, k1y, s1);
          Prim<T> c2 = cellP(i, j, s2), c3 = cellP(k3x, k3y, s3);
          if (s1 && s2) continue;
          T hllw = T(0);
          if (Tp<T>(P.hllBlend) > T(0) && !(s1 || s2)) {
            const T p1 = c1.w.w, p2 = c2.w.w;
            const T jm = std::max(std::max(jumpRatio(pAt(k1x + tx_, k1y + ty_, p1), p1), jumpRatio(pAt(k1x - tx_, k1y - ty_, p1), p1)),
                                  std::max(jumpRatio(pAt(i + tx_, j + ty_, p2), p2), jumpRatio(pAt(i - tx_, j - ty_, p2), p2)));
            hllw = smoothstepT(T(0.15), T(0.6), jm * Tp<T>(P.hllBlend));
          }
          Prim<T> WL, WR;
          if (s1) { WR = c2; WL = mirrorN(c2); }
          else if (s2) { WL = c1; WR = mirrorN(c1); }
          else {
            if (s0) c0 = mirrorN(c1);
            if (s3) c3 = mirrorN(c2);
            WL = c1;
            WR = c2;
            if (P.order2 && area > T(0)) {
              const T* a0 = &c0.w.x; const T* a1 = &c1.w.x; const T* a2 = &c2.w.x; const T* a3 = &c3.w.x;
              const T* b0 = &c0.s.x; const T* b1 = &c1.s.x; const T* b2 = &c2.s.x; const T* b3 = &c3.s.x;
              T tL[4], tR[4], yL[4], yR[4];
              for (int c = 0; c < 4; ++c) {
                const T dL = vanLeer(a1[c] - a0[c], a2[c] - a1[c]);
                const T dR = vanLeer(a2[c] - a1[c], a3[c] - a2[c]);
                tL[c] = a1[c] + T(0.5) * dL;
                tR[c] = a2[c] - T(0.5) * dR;
                const T eL = vanLeer(b1[c] - b0[c], b2[c] - b1[c]);
                const T eR = vanLeer(b2[c] - b1[c], b3[c] - b2[c]);
                yL[c] = b1[c] + T(0.5) * eL;
                yR[c] = b2[c] - T(0.5) * eR;
              }
              if (tL[0] > T(0) && tL[3] > T(0) && tR[0] > T(0) && tR[3] > T(0)) {
                const T lo[4] = {T(0), T(0), T(0.5), T(0)}, hi[4] = {T(1), T(1), T(100), T(1)};
                WL.w = mk(tL[0], tL[1], tL[2], tL[3]);
                WR.w = mk(tR[0], tR[1], tR[2], tR[3]);
                WL.s = mk(clampT(yL[0], lo[0], hi[0]), clampT(yL[1], lo[1], hi[1]), clampT(yL[2], lo[2], hi[2]), clampT(yL[3], lo[3], hi[3]));
                WR.s = mk(clampT(yR[0], lo[0], hi[0]), clampT(yR[1], lo[1], hi[1]), clampT(yR[2], lo[2], hi[2]), clampT(yR[3], lo[3], hi[3]));
              }
            }
          }
          // HLLC blended to HLL (solver.js:211-251)
          const T GL = WL.s.z, GR = WR.s.z;
          const T rL = WL.w.x, uL = WL.w.y, vL = WL.w.z, pL = WL.w.w;
          const T rR = WR.w.x, uR = WR.w.y, vR = WR.w.z, pR = WR.w.w;
          const T cL = std::sqrt((T(1) + T(1) / GL) * pL / rL), cR = std::sqrt((T(1) + T(1) / GR) * pR / rR);
          const T EL = pL * GL + T(0.5) * rL * (uL * uL + vL * vL);
          const T ER = pR * GR + T(0.5) * rR * (uR * uR + vR * vR);
          const T sL = std::sqrt(rL), sR = std::sqrt(rR);
          const T uT = (sL * uL + sR * uR) / (sL + sR);
          const T HT = (sL * (EL + pL) / rL + sR * (ER + pR) / rR) / (sL + sR);
          const T vT = (sL * vL + sR * vR) / (sL + sR);
          const T GT = (sL * GL + sR * GR) / (sL + sR);
          const T cT = std::sqrt(std::max((HT - T(0.5) * (uT * uT + vT * vT)) / GT, T(1e-12)));
          const T SL = std::min(uL - cL, uT - cT);
          const T SR = std::max(uR + cR, uT + cT);
          const T FL[4] = {rL * uL, rL * uL * uL + pL, rL * uL * vL, uL * (EL + pL)};
          const T FR[4] = {rR * uR, rR * uR * uR + pR, rR * uR * vR, uR * (ER + pR)};
          const T YL[3] = {WL.s.x, WL.s.y, WL.s.w}, YR[3] = {WR.s.x, WR.s.y, WR.s.w};
          T FYL[3], FYR[3];
          for (int c = 0; c < 3; ++c) { FYL[c] = rL * uL * YL[c]; FYR[c] = rR * uR * YR[c]; }
          T F[4], FY[3], uf, Gf;
          if (SL >= T(0)) {
            for (int c = 0; c < 4; ++c) F[c] = FL[c];
            for (int c = 0; c < 3; ++c) FY[c] = FYL[c];
            uf = uL; Gf = GL;
          } else if (SR <= T(0)) {
            for (int c = 0; c < 4; ++c) F[c] = FR[c];
            for (int c = 0; c < 3; ++c) FY[c] = FYR[c];
            uf = uR; Gf = GR;
          } else {
            const T UL[4] = {rL, rL * uL, rL * vL, EL};
            const T UR[4] = {rR, rR * uR, rR * vR, ER};
            const T inv = T(1) / (SR - SL);
            T Fhll[4], FYhll[3];
            for (int c = 0; c < 4; ++c) Fhll[c] = (SR * FL[c] - SL * FR[c] + SL * SR * (UR[c] - UL[c])) * inv;
            for (int c = 0; c < 3; ++c) FYhll[c] = (SR * FYL[c] - SL * FYR[c] + SL * SR * (rR * YR[c] - rL * YL[c])) * inv;
            const T Ss = (pR - pL + rL * uL * (SL - uL) - rR * uR * (SR - uR)) / (rL * (SL - uL) - rR * (SR - uR));
            const bool left = Ss >= T(0);
            const T rK = left ? rL : rR, uK = left ? uL : uR, vK = left ? vL : vR, pK = left ? pL : pR;
            const T EK = left ? EL : ER, SK = left ? SL : SR;
            const T* YK = left ? YL : YR;
            const T* FYK = left ? FYL : FYR;
            const T* FK = left ? FL : FR;
            const T* UK = left ? UL : UR;
            const T kk = rK * (SK - uK) / (SK - Ss);
            const T Us[4] = {kk * T(1), kk * Ss, kk * vK, kk * (EK / rK + (Ss - uK) * (Ss + pK / (rK * (SK - uK))))};
            for (int c = 0; c < 4; ++c) {
              const T Fc = FK[c] + SK * (Us[c] - UK[c]);
              F[c] = mixT(Fc, Fhll[c], hllw);
            }
            for (int c = 0; c < 3; ++c) {
              const T FYc = FYK[c] + SK * (kk * YK[c] - rK * YK[c]);
              FY[c] = mixT(FYc, FYhll[c], hllw);
            }
            uf = uK + SK * ((SK - uK) / (SK - Ss) - T(1));
            Gf = left ? GL : GR;
          }
          if (dirR) std::swap(F[1], F[2]);
          store4(oF, fo, mk(F[0], F[1], F[2], F[3]));
          store4(oFY, fo, mk(FY[0], FY[1], uf * Gf, uf));
          store4(oFO, fo, mk(FY[2], T(0), T(0), T(0)));
        }
    }
  }

  // ---------------- FS_UPDATE (solver.js:264-363)
  void update_pass(const std::vector<T>& A, const std::vector<T>& B, const std::vector<T>& A0, const std::vector<T>& B0,
                   int stage, std::vector<T>& oA, std::vector<T>& oB, const nz_params& P) {
    const T* st = step_[stepIdx];
    const int W1 = Nx + 1;
    auto memberFlux = [&](int ci, int cj, T* L, T* LY, T& LO) {
      const v4<T> g = load4(geoA, (size_t)cj * W1 + ci);
      const T aw = std::fabs(g.z), ae = std::fabs(geoA[((size_t)cj * W1 + ci + 1) * 4 + 2]);
      const T as_ = std::fabs(g.w), an = std::fabs(geoA[((size_t)(cj + 1) * W1 + ci) * 4 + 3]);
      const size_t kw = (size_t)cj * W1 + ci, ke = kw + 1, ks = kw, kn = (size_t)(cj + 1) * W1 + ci;
      for (int c = 0; c < 4; ++c) {
        L[c] -= ae * FxA[ke * 4 + c] - aw * FxA[kw * 4 + c];
        L[c] -= an * FrA[kn * 4 + c] - as_ * FrA[ks * 4 + c];
        LY[c] -= ae * FxB[ke * 4 + c] - aw * FxB[kw * 4 + c];
        LY[c] -= an * FrB[kn * 4 + c] - as_ * FrB[ks * 4 + c];
      }
      if (P.inj) {
        LO -= ae * FxC[ke * 4] - aw * FxC[kw * 4];
        LO -= an * FrC[kn * 4] - as_ * FrC[ks * 4];
      }
    };
    auto wallPressure = [](const Prim<T>& q, T un) {
      const T g = T(1) + T(1) / q.s.z, p = q.w.w, rho = q.w.x, c = std::sqrt(g * p / rho);
      if (un > T(0)) {
        const T a = T(0.25) * (g + T(1)) * un;
        return p + rho * un * (a + std::sqrt(a * a + c * c));
      }
      const T f = std::max(T(1) + T(0.5) * (g - T(1)) * un / c, T(0));
      return p * std::pow(f, T(2) * g / (g - T(1)));
    };
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Nr; ++j)
      for (int i = 0; i < Nx; ++i) {
        const size_t k = (size_t)j * Nx + i;
        if (solidAt(i, j)) {
          store4(oA, k, mk(Tp<T>(P.ambCons[0]), Tp<T>(P.ambCons[1]), Tp<T>(P.ambCons[2]), Tp<T>(P.ambCons[3])));
          store4(oB, k, mk(Tp<T>(P.ambB[0]), Tp<T>(P.ambB[1]), Tp<T>(P.ambB[2]), Tp<T>(P.ambB[3])));
          continue;
        }
        const v4<T> gm = load4(geoM, k);
        const int Hi = i + (int)std::lround((double)gm.x), Hj = j + (int)std::lround((double)gm.y);
        const size_t Hk = (size_t)Hj * Nx + Hi;
        const v4<T> gg = load4(geoG, k);
        const v4<T> a = load4(A, Hk), b = load4(B, Hk);
        const Prim<T> q = toPrim(a, b);
        T L[4] = {0, 0, 0, 0}, LY[4] = {0, 0, 0, 0}, LO = 0;
        memberFlux(Hi, Hj, L, LY, LO);
        const int sm = (int)std::lround((double)geoM[Hk * 4 + 2]);
        if (sm & 1) memberFlux(Hi - 1, Hj, L, LY, LO);
        if (sm & 2) memberFlux(Hi + 1, Hj, L, LY, LO);
        if (sm & 4) memberFlux(Hi, Hj - 1, L, LY, LO);
        if (sm & 8) memberFlux(Hi, Hj + 1, L, LY, LO);
        const T Wx = gg.z, Wr = gg.w;
        const T wl = std::sqrt(Wx * Wx + Wr * Wr);
        if (wl > T(1e-6)) {
          const T nx = Wx / wl, nr = Wr / wl;
          const T pw = wallPressure(q, q.w.y * nx + q.w.z * nr);
          L[1] -= Wx * pw;
          L[2] -= Wr * pw;
        }
        L[2] += q.w.w * gg.y;
        const T invV = T(1) / std::max(gg.x, T(1e-6));
        const T dt = st[0];
        const T LB[4] = {LY[0], LY[1], LY[2] - q.s.z * LY[3], LO};
        const T av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
        T U[4], UY[4];
        if (stage == 1) {
          for (int c = 0; c < 4; ++c) { U[c] = av[c] + dt * invV * L[c]; UY[c] = bv[c] + dt * invV * LB[c]; }
        } else {
          const v4<T> a0 = load4(A0, Hk), b0 = load4(B0, Hk);
          const T a0v[4] = {a0.x, a0.y, a0.z, a0.w}, b0v[4] = {b0.x, b0.y, b0.z, b0.w};
          for (int c = 0; c < 4; ++c) {
            U[c] = T(0.5) * (a0v[c] + av[c] + dt * invV * L[c]);
            UY[c] = T(0.5) * (b0v[c] + bv[c] + dt * invV * LB[c]);
          }
        }
        const T rho = std::max(U[0], T(1e-7));
        T vx = U[1] / std::max(U[0], T(1e-7)), vy = U[2] / std::max(U[0], T(1e-7));
        const T sp = std::sqrt(vx * vx + vy * vy);
        if (sp > Tp<T>(P.vMax)) { vx *= Tp<T>(P.vMax) / sp; vy *= Tp<T>(P.vMax) / sp; }
        if (!(rho < T(1e30)) || !(vx < T(1e30) && vx > T(-1e30))) {
          store4(oA, k, mk(Tp<T>(P.ambCons[0]), Tp<T>(P.ambCons[1]), Tp<T>(P.ambCons[2]), Tp<T>(P.ambCons[3])));
          store4(oB, k, mk(Tp<T>(P.ambB[0]), Tp<T>(P.ambB[1]), Tp<T>(P.ambB[2]), Tp<T>(P.ambB[3])));
          continue;
        }
        const T ke_ = T(0.5) * rho * (vx * vx + vy * vy);
        const T Gn = clampT(UY[2], Tp<T>(P.gamR[0]), Tp<T>(P.gamR[1]));
        const T E = std::max(U[3], T(1e-7) * Gn + ke_);
        const T Zn = clampT(UY[0], T(0), rho);
        store4(oA, k, mk(rho, rho * vx, rho * vy, E));
        T ubar = T(0);
        if (P.inj) ubar = clampT(UY[3], T(0), rho);
        else if (i == 0) {
          const T u0 = B[k * 4 + 3];
          const T uBar0 = Tp<T>(P.inlet[0]), uBar1 = T((float)(1.6 * (double)P.uInlet0)), uBar2 = Tp<T>(std::max(P.ramp, 0.f));
          const bool on = stage == 2 && st[2] > uBar2;
          ubar = on ? std::min(u0 + std::min(dt * uBar0, T(1)) * (vx - u0), uBar1) : u0;
        }
        store4(oB, k, mk(Zn, clampT(UY[1], T(0), Zn), Gn, ubar));
      }
  }

  // ---------------- CFL: FS_REDUCE1/2 + FS_STEPINFO (solver.js:368-411)
  void cfl(const std::vector<T>& A, const std::vector<T>& B, const nz_params& P) {
    T wmax = T(0);
#pragma omp parallel for reduction(max : wmax) schedule(static)
    for (int j = 0; j < Nr; ++j)
      for (int i = 0; i < Nx; ++i) {
        if (solidAt(i, j)) continue;
        const size_t k = (size_t)j * Nx + i;
        const Prim<T> q = toPrim(load4(A, k), load4(B, k));
        wmax = std::max(wmax, std::fabs(q.w.y) + std::fabs(q.w.z) + T(2) * soundSpeed(q));
      }
    const T w = std::max(Tp<T>(P.wFloor), wmax);
    const T* prev = step_[stepIdx];
    T dt = std::min(Tp<T>(P.cfl) / w, Tp<T>(P.dtMax));
    if (prev[0] > T(0)) dt = std::min(dt, prev[0] * T(1.2));
    T* nx = step_[1 - stepIdx];
    nx[0] = dt;
    nx[1] = prev[2];
    nx[2] = prev[2] + dt;
    nx[3] = w;
    stepIdx = 1 - stepIdx;
  }

  // ---------------- GLSL_GAS (solver.js:445-450)
  static T sutherland(T TK, const nz_params& P) { return Tp<T>(P.gas[3]) * T(1.458e-6) * TK * std::sqrt(TK) / (TK + T(110.4)); }
  static T Rmix(const v4<T>& s, const nz_params& P) {
    return (T(1) - s.x) * Tp<T>(P.gas[0]) + s.y * Tp<T>(P.Rs[0]) + s.w * Tp<T>(P.Rs[1]) + std::max(s.x - s.y - s.w, T(0)) * Tp<T>(P.Rs[2]);
  }

  // ---------------- FS_VCELL (solver.js:451-489)
  void vcell_pass(const nz_params& P) {
    const T dt = step_[stepIdx][0];
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Nr; ++j)
      for (int i = 0; i < Nx; ++i) {
        const size_t k = (size_t)j * Nx + i;
        const v4<T> w0 = Wat(i, j);
        store4(T1, k, mk(T(0), T(0), T(0), T(0)));
        store4(T2, k, mk(T(0), T(0), T(0), T(0)));
        if (w0.x < T(0)) continue;
        const v4<T> wE = Wat(i + 1, j), wW = Wat(i - 1, j), wN = Wat(i, j + 1), wS = Wat(i, j - 1);
        const T fE = wE.x < T(0) ? T(0) : T(1), fW = wW.x < T(0) ? T(0) : T(1);
        const T fN = wN.x < T(0) ? T(0) : T(1), fS = wS.x < T(0) ? T(0) : T(1);
        const T dx_ = std::max(fE + fW, T(1)), dr_ = std::max(fN + fS, T(1));
        const T ux = (fE * (wE.y - w0.y) + fW * (w0.y - wW.y)) / dx_;
        const T vx = (fE * (wE.z - w0.z) + fW * (w0.z - wW.z)) / dx_;
        const T ur = (fN * (wN.y - w0.y) + fS * (w0.y - wS.y)) / dr_;
        const T vr = (fN * (wN.z - w0.z) + fS * (w0.z - wS.z)) / dr_;
        const T r = T(j) + T(0.5), vor = w0.z / r;
        const T div = ux + vr + vor;
        const T S2 = T(2) * (ux * ux + vr * vr + vor * vor) + (ur + vx) * (ur + vx) - T(0.6666667f) * div * div;
        const T om = vx - ur;
        const T sens = om * om / (om * om + div * div + T(1e-8));
        const T Smag = std::sqrt(std::max(S2, T(0))) * sens;
        const v4<T> sc = Wbat(i, j);
        const T rho = w0.x, p = w0.w;
        const T Rm = Rmix(sc, P);
        const T mu = sutherland(p / (rho * Rm) * Tp<T>(P.gas[2]), P);
        const T l = geoM[k * 4 + 3];
        T mut = rho * l * l * Smag;
        if (T(i) < Tp<T>(P.chamberNu[1])) mut = std::max(mut, rho * Tp<T>(P.chamberNu[0]));
        store4(T1, k, mk(ux, ur, vx, vr));
        store4(T2, k, mk(std::min(mu + mut, T(0.08) * rho / std::max(dt, T(1e-9))), (sc.z + T(1)) * p / rho, Smag, mu));
      }
  }

  // ---------------- FS_VFLUX_X / FS_VFLUX_R (solver.js:490-539)
  void vflux_pass() {
    const T PR_INV = T(1.25);
    for (int dirR = 0; dirR < 2; ++dirR) {
      const int FW = dirR ? Nx : Nx + 1, FH = dirR ? Nr + 1 : Nr;
      std::vector<T>& oF = dirR ? FrA : FxA;
      std::vector<T>& oFY = dirR ? FrB : FxB;
      std::vector<T>& oFO = dirR ? FrC : FxC;
#pragma omp parallel for schedule(static)
      for (int j = 0; j < FH; ++j)
        for (int i = 0; i < FW; ++i) {
          const size_t fo = (size_t)j * (Nx + 1) + i;
          store4(oF, fo, mk(T(0), T(0), T(0), T(0)));
          store4(oFY, fo, mk(T(0), T(0), T(0), T(0)));
          store4(oFO, fo, mk(T(0), T(0), T(0), T(0)));
          const v4<T> geo = load4(geoA, fo);
          int k1x, k1y;
          if (dirR) {
            if (geo.w == T(0) || j == 0 || j >= Nr) continue;
            k1x = i; k1y = j - 1;
          } else {
            if (geo.z == T(0) || i == 0 || i >= Nx) continue;
            k1x = i - 1; k1y = j;
          }
          const v4<T> w1 = Wat(k1x, k1y), w2 = Wat(i, j);
          if (w1.x < T(0) || w2.x < T(0)) continue;
          const size_t c1 = (size_t)k1y * Nx + k1x, c2 = (size_t)j * Nx + i;
          const v4<T> t1 = load4(T1, c1), t2 = load4(T1, c2);
          const v4<T> m1 = load4(T2, c1), m2 = load4(T2, c2);
          const v4<T> s1 = Wbat(k1x, k1y), s2 = Wbat(i, j);
          const T mu = T(0.5) * (m1.x + m2.x);
          const T uf = T(0.5) * (w1.y + w2.y), vf = T(0.5) * (w1.z + w2.z);
          const v4<T> tt = mk(T(0.5) * (t1.x + t2.x), T(0.5) * (t1.y + t2.y), T(0.5) * (t1.z + t2.z), T(0.5) * (t1.w + t2.w));
          T r, ux, vx, ur, vr;
          if (dirR) {
            r = std::max(T(j), T(0.5));
            ux = tt.x; vx = tt.z; ur = w2.y - w1.y; vr = w2.z - w1.z;
          } else {
            r = T(j) + T(0.5);
            ux = w2.y - w1.y; vx = w2.z - w1.z; ur = tt.y; vr = tt.w;
          }
          const T div = ux + vr + vf / r;
          const T txr = mu * (ur + vx);
          const T kk = mu * PR_INV;
          v4<T> F;
          if (dirR) {
            const T trr = mu * (T(2) * vr - T(0.6666667f) * div);
            F = mk(-kk * (s2.w - s1.w), -txr, -trr, -(uf * txr + vf * trr) - kk * (m2.y - m1.y));
          } else {
            const T txx = mu * (T(2) * ux - T(0.6666667f) * div);
            F = mk(-kk * (s2.w - s1.w), -txx, -txr, -(uf * txx + vf * txr) - kk * (m2.y - m1.y));
          }
          store4(oF, fo, F);
          store4(oFY, fo, mk(-kk * (s2.x - s1.x), -kk * (s2.y - s1.y), -kk * (s2.z - s1.z), T(0)));
        }
    }
  }

  // ---------------- FS_SOURCE (solver.js:540-663)
  void source_pass(const std::vector<T>& A, const std::vector<T>& B, std::vector<T>& oA, std::vector<T>& oB, const nz_params& P) {
    const T dt = step_[stepIdx][0];
    const int W1 = Nx + 1;
    auto memberFlux = [&](int ci, int cj, T* L, T* LY) {
      const v4<T> g = load4(geoA, (size_t)cj * W1 + ci);
      const T aw = std::fabs(g.z), ae = std::fabs(geoA[((size_t)cj * W1 + ci + 1) * 4 + 2]);
      const T as_ = std::fabs(g.w), an = std::fabs(geoA[((size_t)(cj + 1) * W1 + ci) * 4 + 3]);
      const size_t kw = (size_t)cj * W1 + ci, ke = kw + 1, ks = kw, kn = (size_t)(cj + 1) * W1 + ci;
      for (int c = 0; c < 4; ++c) {
        L[c] -= ae * FxA[ke * 4 + c] - aw * FxA[kw * 4 + c];
        L[c] -= an * FrA[kn * 4 + c] - as_ * FrA[ks * 4 + c];
        LY[c] -= ae * FxB[ke * 4 + c] - aw * FxB[kw * 4 + c];
        LY[c] -= an * FrB[kn * 4 + c] - as_ * FrB[ks * 4 + c];
      }
    };
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Nr; ++j)
      for (int i = 0; i < Nx; ++i) {
        const size_t k = (size_t)j * Nx + i;
        if (solidAt(i, j)) {
          store4(oA, k, mk(Tp<T>(P.ambCons[0]), Tp<T>(P.ambCons[1]), Tp<T>(P.ambCons[2]), Tp<T>(P.ambCons[3])));
          store4(oB, k, mk(Tp<T>(P.ambB[0]), Tp<T>(P.ambB[1]), Tp<T>(P.ambB[2]), Tp<T>(P.ambB[3])));
          continue;
        }
        const int Hi = i + (int)std::lround((double)geoM[k * 4]), Hj = j + (int)std::lround((double)geoM[k * 4 + 1]);
        const size_t Hk = (size_t)Hj * Nx + Hi;
        const v4<T> gg = load4(geoG, k);
        const v4<T> a = load4(A, Hk), b = load4(B, Hk);
        const Prim<T> q = toPrim(a, b);
        const T rho = q.w.x, p = q.w.w, G = q.s.z;
        const T velx = q.w.y, vely = q.w.z;
        const T V = std::max(gg.x, T(1e-6));
        const T Rm = Rmix(q.s, P), Tg = p / (rho * Rm);
        T L[4] = {0, 0, 0, 0}, LY[4] = {0, 0, 0, 0};
        const v4<T> t1 = load4(T1, Hk), t2 = load4(T2, Hk);
        if (P.visc) {
          memberFlux(Hi, Hj, L, LY);
          const int sm = (int)std::lround((double)geoM[Hk * 4 + 2]);
          if (sm & 1) memberFlux(Hi - 1, Hj, L, LY);
          if (sm & 2) memberFlux(Hi + 1, Hj, L, LY);
          if (sm & 4) memberFlux(Hi, Hj - 1, L, LY);
          if (sm & 8) memberFlux(Hi, Hj + 1, L, LY);
          const T r = V / std::max(gg.y, T(1e-6)), vor = vely / r;
          L[2] -= t2.x * (T(2) * vor - T(0.6666667f) * (t1.x + t1.w + vor)) * gg.y;
          const T wl = std::sqrt(gg.z * gg.z + gg.w * gg.w);
          if (wl > T(1e-6)) {
            const T nx = gg.z / wl, nr = gg.w / wl;
            const T vn = velx * nx + vely * nr;
            const T utx = velx - vn * nx, utr = vely - vn * nr;
            const T up = std::sqrt(utx * utx + utr * utr);
            const T cp = Rm * (G + T(1));
            const T Taw = Tg + T(0.89) * up * up / (T(2) * cp);
            const T Ts = std::max(Tg + T(0.5) * (Tp<T>(P.Tw) - Tg) + T(0.22) * (Taw - Tg), T(1e-3));
            const T rs = p / (Rm * Ts), mus = sutherland(Ts * Tp<T>(P.gas[2]), P);
            const T dz = T(2) * clampT(T(0.5) * V / wl, T(0.25), T(1));
            const T nz_ = mus / (rs * dz);
            T tw = up <= T(69.7) * nz_ ? T(2) * mus * up / dz
                                       : rs * std::pow(T(7.187) * std::pow(nz_, T(1.142857)) + T(0.13769) * std::pow(nz_, T(0.142857)) * up, T(1.75));
            tw = std::min(tw, T(0.3) * rho * up * V / std::max(wl * dt, T(1e-9)));
            if (up > T(1e-6)) { L[1] -= tw * wl * utx / up; L[2] -= tw * wl * utr / up; }
            const T qw = tw / std::max(up, T(1e-4)) * cp * (Taw - Tp<T>(P.Tw)) * T(1.24);
            const T qmax = T(0.3) * p * G * V / std::max(wl * dt, T(1e-9));
            L[3] -= clampT(qw, -qmax, qmax) * wl;
          }
          if (sm & 16) {
            const T up = std::sqrt(velx * velx + vely * vely), Mn = up / std::sqrt((T(1) + T(1) / G) * p / rho);
            if (Mn > T(1)) {
              const T shx = velx / up, shy = vely / up;
              T pd = T(0);
              for (int n = 4; n <= 6; n += 2)
                for (int m = 0; m <= 1; ++m) {
                  int Dx = Hi + (int)std::lround((double)(T(n) * shx));
                  // GLSL round() of an exact .5 is implementation-defined; lround rounds away from zero
                  int Dy = Hj + (int)std::floor(T(n) * shy) + m;
                  Dx = std::min(std::max(Dx, 0), Nx - 1);
                  Dy = std::min(std::max(Dy, 0), Nr - 1);
                  const v4<T> wd = Wat(Dx, Dy);
                  if (wd.x > T(0)) pd = std::max(pd, wd.w);
                }
              const T crit = T(1) + T(0.5) * Mn;
              const T f = smoothstepT(crit, T(1.2) * crit, pd / p) * V * rho * up;
              L[1] -= f * velx;
              L[2] -= f * vely;
            } else if (velx > T(0) && p < T(1.3) * Tp<T>(P.chemK[3])) {
              L[1] -= T(0.5) * V * rho * up * velx;
              L[2] -= T(0.5) * V * rho * up * vely;
            }
          }
        }
        const T invV = dt / V;
        const T dO = invV * L[0];
        L[0] = T(0);
        T U[4] = {a.x + invV * L[0], a.y + invV * L[1], a.z + invV * L[2], a.w + invV * L[3]};
        T UY[4] = {b.x + invV * LY[0] / T(1), b.y + invV * LY[1] / T(1), b.z + invV * LY[2] / rho, b.w + invV * T(0) / T(1)};
        if (P.inj) UY[3] += dO;
        if (P.chem) {
          const T Z = q.s.x, F = q.s.y, O = q.s.w, s = Tp<T>(P.chemK[1]);
          const T avail = P.inj ? std::min(F, (O + T(0.233) * std::max(T(1) - Z, T(0))) / s)
                                : std::min(F, std::max(T(1) - Z - s * (Z - F), T(0)) / s);
          if (avail > T(1e-5)) {
            const T TK = Tg * Tp<T>(P.gas[2]);
            const T kin = Tp<T>(P.chemK[2]) * p * std::exp(T(-9000) / TK) * (T(1) - smoothstepT(Tp<T>(P.cap[0]), Tp<T>(P.cap[1]), TK));
            const T mixr = P.inj ? std::max(t2.z, T(4) * t2.x / rho) : t2.z;
            const T burn = avail * std::min(std::min(mixr, kin) * dt, T(0.5));
            UY[1] -= rho * burn;
            if (P.inj) UY[3] = std::max(UY[3] - rho * burn * s, T(0));
            U[3] += rho * burn * Tp<T>(P.chemK[0]);
          }
        }
        const T rn = std::max(U[0], T(1e-7));
        const T vnx = U[1] / rn, vny = U[2] / rn;
        if (!(rn < T(1e30)) || !(vnx < T(1e30) && vnx > T(-1e30))) {
          store4(oA, k, mk(Tp<T>(P.ambCons[0]), Tp<T>(P.ambCons[1]), Tp<T>(P.ambCons[2]), Tp<T>(P.ambCons[3])));
          store4(oB, k, mk(Tp<T>(P.ambB[0]), Tp<T>(P.ambB[1]), Tp<T>(P.ambB[2]), Tp<T>(P.ambB[3])));
          continue;
        }
        const T Gn = clampT(UY[2], Tp<T>(P.gamR[0]), Tp<T>(P.gamR[1]));
        const T Zn = clampT(UY[0], T(0), rn);
        store4(oA, k, mk(rn, rn * vnx, rn * vny, std::max(U[3], T(1e-7) * Gn + T(0.5) * rn * (vnx * vnx + vny * vny))));
        store4(oB, k, mk(Zn, clampT(UY[1], T(0), Zn), Gn, P.inj ? clampT(UY[3], T(0), rn) : B[k * 4 + 3]));
      }
  }

  // ---------------- FS_RESCALE (solver.js:418-437), applied through GPUSolver.rescaleAmbient (998-1010)
  void rescale(T fp, T fT, T du, const nz_params& P) {
    std::vector<T>& A = SA[cur];
    std::vector<T>& B = SB[cur];
    std::vector<T>& oA = SA[(cur + 1) % 3];
    std::vector<T>& oB = SB[(cur + 1) % 3];
    const T thr = T(0.03);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Nr; ++j)
      for (int i = 0; i < Nx; ++i) {
        const size_t k = (size_t)j * Nx + i;
        const v4<T> a = load4(A, k), b = load4(B, k);
        if (solidAt(i, j)) {
          store4(oA, k, mk(Tp<T>(P.ambCons[0]), Tp<T>(P.ambCons[1]), Tp<T>(P.ambCons[2]), Tp<T>(P.ambCons[3])));
          store4(oB, k, mk(Tp<T>(P.ambB[0]), Tp<T>(P.ambB[1]), Tp<T>(P.ambB[2]), Tp<T>(P.ambB[3])));
          continue;
        }
        const Prim<T> q = toPrim(a, b);
        const T w = T(1) - smoothstepT(T(0), thr, q.s.x);
        const T rho = q.w.x * mixT(T(1), fp / fT, w);
        const T p = q.w.w * mixT(T(1), fp, w);
        const T vx = q.w.y + du * w, vy = q.w.z;
        store4(oA, k, mk(rho, rho * vx, rho * vy, p * q.s.z + T(0.5) * rho * (vx * vx + vy * vy)));
        store4(oB, k, mk(rho * q.s.x, rho * q.s.y, q.s.z, b.w));
      }
    cur = (cur + 1) % 3;
  }

  // ---------------- FS_WALLROW (solver.js:665-674)
  void wall_row() {
    const std::vector<T>& A = SA[cur];
    const std::vector<T>& B = SB[cur];
    for (int i = 0; i < Nx; ++i) {
      const int j = (int)jWall[(size_t)i * 4];
      const size_t k = (size_t)j * Nx + i;
      const Prim<T> q = toPrim(load4(A, k), load4(B, k));
      store4(wallRow, i, mk(q.w.y, q.w.z, q.w.w, soundSpeed(q)));
    }
  }

  // ---------------- FS_DERIVED (render.js:60-134), the display fields of render.js derive()
  void derived(const nz_derived_params& d, double velScale, float* out) const {
    const std::vector<T>& A = SA[cur];
    const std::vector<T>& B = SB[cur];
    auto idx = [&](int a, int b) { return (size_t)std::min(std::max(b, 0), Nr - 1) * Nx + std::min(std::max(a, 0), Nx - 1); };
    auto solid = [&](int a, int b) { return mask[idx(a, b)] != 0; };
    auto P = [&](int a, int b) { const size_t k = idx(a, b); return toPrim(load4(A, k), load4(B, k)); };
    auto lrho = [&](int a, int b) { return std::log(std::max(A[idx(a, b < 0 ? -b - 1 : b) * 4], T(1e-7))); };
    const int f = d.field == 8 ? 0 : d.field;
    const T vel = Tp<T>((float)velScale);
    for (int j = 0; j < Nr; ++j)
      for (int i = 0; i < Nx; ++i) {
        float* o = out + ((size_t)j * Nx + i) * 4;
        if (solid(i, j)) { o[0] = 0.f; o[1] = 0.f; o[2] = 0.f; o[3] = -1.f; continue; }
        const Prim<T> q = P(i, j);
        const T rho = q.w.x, p = q.w.w, Z = q.s.x, F = q.s.y, O = q.s.w;
        const T c = soundSpeed(q);
        const T spd = std::sqrt(q.w.y * q.w.y + q.w.z * q.w.z);
        const T M = spd / c;
        const T Rm = (T(1) - Z) * Tp<T>(d.RaN) + F * Tp<T>(d.Rs[0]) + O * Tp<T>(d.Rs[1]) +
                     std::max(Z - F - O, T(0)) * Tp<T>(d.Rs[2]);
        const T TT = p / (rho * Rm) * Tp<T>(288.15f);
        T v = TT;
        switch (f) {
          case 1: v = M; break;
          case 2: {
            const T l0 = lrho(i, j);
            const T le = solid(i + 1, j) ? l0 : lrho(i + 1, j), lw = solid(i - 1, j) ? l0 : lrho(i - 1, j);
            const T ln = solid(i, j + 1) ? l0 : lrho(i, j + 1), ls = (j > 0 && solid(i, j - 1)) ? l0 : lrho(i, j - 1);
            v = T(1) - std::exp(-Tp<T>(d.schlGain) * std::sqrt((le - lw) * (le - lw) + (ln - ls) * (ln - ls)) * T(0.5));
            break;
          }
          case 3: v = std::log(p * T(101325)) / std::log(T(10)); break;
          case 4: v = std::log(rho * T(1.225f)) / std::log(T(10)); break;
          case 5: v = spd * vel; break;
          case 6: {
            const Prim<T> e = P(i + 1, j), w = P(i - 1, j), n = P(i, j + 1);
            const T su = j > 0 ? P(i, j - 1).w.y : q.w.y;
            v = std::fabs(T(0.5) * (e.w.z - w.w.z) - T(0.5) * (n.w.y - su)) * vel / Tp<T>(d.dx);
            break;
          }
          case 7: v = Z; break;
          case 9: {
            // vis.T2 is an RGBA32F texture (render.js:426): its texels hold float values
            const T m0 = T((float)T2[((size_t)j * Nx + i) * 4]), m3 = T((float)T2[((size_t)j * Nx + i) * 4 + 3]);
            v = m3 > T(0) ? std::log(std::max(m0 / m3, T(1))) / std::log(T(10)) : T(0);
            break;
          }
          case 10: v = Z > T(0.003f) ? (Z - F) / Z : T(0); break;
          case 11: v = F; break;
          case 12: v = O; break;
          case 13: {
            const T m0 = T((float)T2[((size_t)j * Nx + i) * 4]), m2 = T((float)T2[((size_t)j * Nx + i) * 4 + 2]);
            const T avail = std::min(F, (O + T(0.233f) * std::max(T(1) - Z, T(0))) / Tp<T>(d.chemK[1]));
            const T kin = Tp<T>(d.chemK[2]) * p * std::exp(T(-9000) / std::max(TT, T(200)));
            v = std::log(T(1) + T(1e5f) * avail * std::min(std::max(m2, T(4) * m0 / rho), kin)) / std::log(T(10));
            break;
          }
          default: break;
        }
        T t = (v - Tp<T>(d.range[0])) / (Tp<T>(d.range[1]) - Tp<T>(d.range[0]));
        if (f == 2) t = v;
        if (f == 6) t = std::log(T(1) + T(20) * std::max(v, T(0)) / Tp<T>(d.range[1])) / std::log(T(21));
        o[0] = (float)t;
        o[1] = (float)Z;
        o[2] = (float)((TT - Tp<T>(d.Ta)) / std::max(Tp<T>(d.Tc) - Tp<T>(d.Ta), T(1)));
        o[3] = (float)M;
      }
  }

  // ---------------- spray (spray.js:19-179, 256-297)
  static uint32_t pcg(uint32_t v) {
    const uint32_t s = v * 747796405u + 2891336453u;
    const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
  }
  static T rnd(uint32_t a, uint32_t n) { return T((float)(pcg(a + n * 7919u) & 0xffffffu) / 16777216.0f); }
  bool wallAt(T x, T r) const {
    const int ci = (int)std::floor(x), cj = (int)std::floor(r);
    if (ci < 0 || cj < 0 || ci >= Nx || cj >= Nr) return false;
    return Wat(ci, cj).x < T(0);
  }
  void spray_update(const nz_params& P) {
    const std::vector<T>& fA = PA[spCur];
    const std::vector<T>& fB = PB[spCur];
    std::vector<T>& oA = PA[1 - spCur];
    std::vector<T>& oB = PB[1 - spCur];
    std::vector<T>& oC = PC[1 - spCur];
    const T dt = step_[stepIdx][0];
    const int perStep = 4;
    const T fan = T(25 * M_PI / 180);
#pragma omp parallel for schedule(static)
    for (int idx = 0; idx < NP; ++idx) {
      const v4<T> a = load4(fA, idx), b = load4(fB, idx);
      int rel = idx - (int)head;
      if (rel < 0) rel += NP;
      if (rel < perStep) {
        const uint32_t id = (uint32_t)idx ^ (seed * 2654435761u);
        const int t = (int)(rnd(id, 1u) * T(255.999));
        const v4<T> e1 = load4(tab, t), e2 = load4(tab, 256 + t), e3 = load4(tab, 512 + t);
        const T ang = e1.z + (rnd(id, 2u) - T(0.5)) * T(2) * e2.z;
        const T r = std::max(e1.y + (rnd(id, 3u) - T(0.5)) * e2.w, T(0.05));
        T d;
        if (e3.y > T(0)) {
          // lognormal drop size (injector-table option of this library, not the reference's):
          // median e2.y, spread e3.y, Box-Muller normal deviate from draws 4 and 6
          const T u1 = T(1) - rnd(id, 4u), u2 = rnd(id, 6u);
          const T nrm = std::sqrt(T(-2) * std::log(u1)) * std::cos(T(6.2831853071795865) * u2);
          d = e2.y * std::exp(e3.y * nrm);
        } else {
          d = e2.y * std::exp((rnd(id, 4u) - T(0.5)) * T(1.4));
        }
        store4(oA, idx, mk(e1.x, r, e1.w * std::cos(ang), e1.w * std::sin(ang)));
        store4(oB, idx, mk(d, Tp<T>(P.sprayMhat) * dt / T(perStep), e2.x, e3.x));
        store4(oC, idx, b.y > T(0) ? mk(b.y, b.y * a.z, b.y * a.w, T(0)) : mk(T(0), T(0), T(0), T(0)));
        continue;
      }
      store4(oA, idx, a);
      store4(oB, idx, b);
      store4(oC, idx, mk(T(0), T(0), T(0), T(0)));
      if (b.y <= T(0)) continue;
      const int gi = std::min(std::max((int)std::floor(a.x), 0), Nx - 1);
      const int gj = std::min(std::max((int)std::floor(a.y), 0), Nr - 1);
      const v4<T> w = Wat(gi, gj);
      if (w.x < T(0)) {
        store4(oA, idx, mk(a.x + T(1), std::max(a.y - T(0.5), T(0.05)), std::max(a.z, T(0)), a.w));
        continue;
      }
      const float* pr = b.z < T(0.5) ? P.fuelVec : P.oxVec;
      const T prx = pr[0], pry = pr[1], prz = pr[2], prw = pr[3];
      T vx = a.z, vy = a.w;
      const T m = b.y, d = b.x, dts = dt * Tp<T>(P.sprayTScale);
      T dn = d, mn = m;
      if (w.x > T(0)) {
        const v4<T> s = Wbat(gi, gj);
        const T rm = (T(1) - s.x) * Tp<T>(P.gas[0]) + s.y * Tp<T>(P.Rs[0]) + s.w * Tp<T>(P.Rs[1]) + std::max(s.x - s.y - s.w, T(0)) * Tp<T>(P.Rs[2]);
        const T TT = w.w / (w.x * rm) * Tp<T>(P.gas[2]);
        const T rhoG = w.x * T(1.225);
        const T rvx = w.y - vx, rvy = w.z - vy;
        const T Urel = std::sqrt(rvx * rvx + rvy * rvy) * T(287.6) + T(1e-3);
        const T mu = T(1.458e-6) * TT * std::sqrt(TT) / (TT + T(110.4));
        const T Re = rhoG * Urel * d / mu;
        const T invTau = T(18) * mu * (T(1) + T(0.15) * std::pow(std::min(Re, T(1000)), T(0.687))) / (prx * d * d);
        const T ex = std::exp(-dts * invTau);
        vx = w.y + (vx - w.y) * ex;
        vy = w.z + (vy - w.z) * ex;
        const T cp = rm * T(287.05) * (s.z + T(1));
        const T Bn = std::max(cp * (TT - pry), T(0)) / prz;
        const T K = T(8) * (mu / T(0.7)) / prx * std::log(T(1) + Bn) * (T(1) + T(0.27) * std::sqrt(Re));
        const T d2 = d * d - K * dts;
        dn = d2 > T(1e-12) ? std::sqrt(d2) : T(0);
        const T f = dn / d;
        mn = dn > T(0) ? m * f * f * f : T(0);
        const T we = rhoG * Urel * Urel * dn / prw;
        if (we > T(12) && dn > T(0)) {
          const T dst = T(12) * prw / (rhoG * Urel * Urel);
          const T tb = T(3) * dn / Urel * std::sqrt(prx / rhoG);
          dn = mixT(dn, std::max(dst, T(2e-6)), T(1) - std::exp(-dts / tb));
        }
      }
      T xi = b.w;
      if (xi > T(0) && a.x >= xi) {
        const uint32_t id = (uint32_t)idx ^ (seed * 2654435761u);
        const T sp = std::sqrt(vx * vx + vy * vy), phi = (rnd(id, 5u) - T(0.5)) * T(2) * fan;
        vx = sp * std::cos(phi);
        vy = sp * std::sin(phi);
        dn *= T(0.25);
        xi = T(0);
      }
      T x = a.x + vx * dt, r = a.y + vy * dt;
      if (r < T(0)) { r = -r; vy = -vy; }
      if (x < T(0)) { x = -x; vx = std::fabs(vx) * T(0.3); }
      if (wallAt(x, r)) {
        const bool hx = wallAt(x, a.y), hr = wallAt(a.x, r);
        if (hx) { vx *= T(-0.3); x = a.x; }
        if (hr) { vy *= T(-0.3); r = a.y; }
        if (!hx && !hr) { x = a.x; r = a.y; vx *= T(-0.3); vy *= T(-0.3); }
      }
      const T dm = m - mn;
      const T dpx = mn * (a.z - vx) + dm * a.z, dpy = mn * (a.w - vy) + dm * a.w;
      store4(oC, idx, mk(dm, dpx, dpy, T(0)));
      if (x >= T(Nx) || r >= T(Nr) || mn <= T(0)) {
        store4(oA, idx, mk(T(-1), T(-1), T(0), T(0)));
        store4(oB, idx, mk(T(0), T(0), T(0), T(0)));
        continue;
      }
      store4(oA, idx, mk(x, r, vx, vy));
      store4(oB, idx, mk(dn, mn, b.z, xi));
    }
  }
  // additive point deposit into the host cell of each parcel's old cell (spray.js:125-149, 274-284)
  void spray_deposit() {
    std::fill(src.begin(), src.end(), T(0));
    const std::vector<T>& fA = PA[spCur];
    const std::vector<T>& fB = PB[spCur];
    const std::vector<T>& tC = PC[1 - spCur];
    for (int idx = 0; idx < NP; ++idx) {
      const v4<T> a = load4(fA, idx), c = load4(tC, idx);
      if (a.x < T(0) || (c.x == T(0) && c.y == T(0) && c.z == T(0))) continue;
      const T type = fB[(size_t)idx * 4 + 2];
      int gi = std::min(std::max((int)std::floor(a.x), 0), Nx - 1);
      int gj = std::min(std::max((int)std::floor(a.y), 0), Nr - 1);
      const size_t g = (size_t)gj * Nx + gi;
      gi += (int)std::lround((double)geoM[g * 4]);
      gj += (int)std::lround((double)geoM[g * 4 + 1]);
      const size_t h = (size_t)gj * Nx + gi;
      src[h * 4] += c.x * (T(1) - type);
      src[h * 4 + 1] += c.x * type;
      src[h * 4 + 2] += c.y;
      src[h * 4 + 3] += c.z;
    }
  }
  void spray_apply(const std::vector<T>& A, const std::vector<T>& B, std::vector<T>& oA, std::vector<T>& oB, const nz_params& P) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Nr; ++j)
      for (int i = 0; i < Nx; ++i) {
        const size_t k = (size_t)j * Nx + i;
        if (solidAt(i, j)) {
          store4(oA, k, mk(Tp<T>(P.ambCons[0]), Tp<T>(P.ambCons[1]), Tp<T>(P.ambCons[2]), Tp<T>(P.ambCons[3])));
          store4(oB, k, mk(Tp<T>(P.ambB[0]), Tp<T>(P.ambB[1]), Tp<T>(P.ambB[2]), Tp<T>(P.ambB[3])));
          continue;
        }
        const int Hi = i + (int)std::lround((double)geoM[k * 4]), Hj = j + (int)std::lround((double)geoM[k * 4 + 1]);
        const size_t Hk = (size_t)Hj * Nx + Hi;
        const v4<T> s = load4(src, Hk);
        const v4<T> a = load4(A, Hk), b = load4(B, Hk);
        const T invV = T(1) / std::max(geoG[k * 4], T(1e-6));
        const T dm = s.x + s.y;
        if (dm == T(0) && s.z == T(0) && s.w == T(0)) { store4(oA, k, a); store4(oB, k, b); continue; }
        const T velx = a.y / std::max(a.x, T(1e-7)), vely = a.z / std::max(a.x, T(1e-7));
        const T rn = a.x + dm * invV;
        const T mx = a.y + s.z * invV, my = a.z + s.w * invV;
        T E = a.w + (s.x * Tp<T>(P.eF) + s.y * Tp<T>(P.eO) + (velx * s.z + vely * s.w)) * invV;
        const T phi = std::min(dm * invV / rn, T(1));
        const T Gn = mixT(b.z, Tp<T>(P.Ge), phi);
        const T vnx = mx / rn, vny = my / rn;
        E = std::max(E, T(0.02) * a.w + T(0.5) * rn * (vnx * vnx + vny * vny) + T(1e-7) * Gn);
        store4(oA, k, mk(rn, mx, my, E));
        store4(oB, k, mk(b.x + dm * invV, b.y + s.x * invV, Gn, b.w + s.y * invV));
      }
  }

  // ---------------- GPUSolver.step (solver.js:849-891)
  void step(int n, const nz_params& P, bool sprayOn) {
    for (int k = 0; k < n; ++k) {
      const int c0 = cur, c1 = (cur + 1) % 3, c2 = (cur + 2) % 3;
      cfl(SA[c0], SB[c0], P);
      prim_pass(SA[c0], SB[c0], 1, P);
      flux_pass(P);
      update_pass(SA[c0], SB[c0], SA[c0], SB[c0], 1, SA[c1], SB[c1], P);
      prim_pass(SA[c1], SB[c1], 2, P);
      flux_pass(P);
      update_pass(SA[c1], SB[c1], SA[c0], SB[c0], 2, SA[c2], SB[c2], P);
      cur = (cur + 2) % 3;
      if (P.visc || P.chem) {
        const int nx = (cur + 1) % 3;
        vcell_pass(P);
        if (P.visc) vflux_pass();
        source_pass(SA[cur], SB[cur], SA[nx], SB[nx], P);
        cur = nx;
      }
      if (P.inj && sprayOn) {
        const int nx = (cur + 1) % 3;
        spray_update(P);
        spray_deposit();
        spray_apply(SA[cur], SB[cur], SA[nx], SB[nx], P);
        cur = nx;
        spCur = 1 - spCur;
        head = (head + 4) % (uint32_t)NP;
        seed = seed + 1;
      }
      stepCount++;
    }
  }
};

struct Handle {
  int prec = 32;
  Oracle<float> f;
  Oracle<double> d;
  bool sprayOn = false;
};

template <class T>
void copy_in(std::vector<T>& dst, const float* src, size_t n) {
  for (size_t k = 0; k < n; ++k) dst[k] = (T)src[k];
}
template <class T>
void copy_out(float* dst, const std::vector<T>& src, size_t n) {
  for (size_t k = 0; k < n; ++k) dst[k] = (float)src[k];
}

template <class T>
void set_geo(Oracle<T>& o, const uint8_t* mask, const float* geoA, const float* geoG, const float* geoM, const float* jWall) {
  const size_t N = (size_t)o.Nx * o.Nr, NL = (size_t)(o.Nx + 1) * (o.Nr + 1);
  std::memcpy(o.mask.data(), mask, N);
  copy_in(o.geoA, geoA, NL * 4);
  copy_in(o.geoG, geoG, N * 4);
  copy_in(o.geoM, geoM, N * 4);
  if (jWall) copy_in(o.jWall, jWall, (size_t)o.Nx * 4);
}

template <class T>
void set_state(Oracle<T>& o, const float* A, const float* B) {
  const size_t N = (size_t)o.Nx * o.Nr * 4;
  for (int k = 0; k < 3; ++k) { copy_in(o.SA[k], A, N); copy_in(o.SB[k], B, N); }
  std::memset(o.step_, 0, sizeof(o.step_));
  o.cur = 0;
  o.stepIdx = 0;
  o.stepCount = 0;
  o.spray_reset();
}

template <class T>
void set_inj(Oracle<T>& o, const float* tab) { copy_in(o.tab, tab, 256 * 3 * 4); }

// The first stage of the next step (CFL, FS_PRIM, FS_FLUX_X/R, FS_UPDATE stage 1) from the
// current state, leaving the oracle as it was: the stage-1 state, the face-flux textures
// (per face F = (mass, x mom, r mom, energy), FY = (rho Z u, rho F u, uf Gf, uf),
// FO = (rho O u, 0, 0, 0), unweighted, as solver.js:246-255 writes them) and the step info.
template <class T>
void stage1_probe(Oracle<T>& o, const nz_params& P, double* A1, double* B1, double* Fx, double* Fr, double* info) {
  T saved[2][4];
  std::memcpy(saved, o.step_, sizeof(saved));
  const int savedIdx = o.stepIdx;
  const int c0 = o.cur;
  std::vector<T> tA(o.SA[c0].size()), tB(o.SB[c0].size());
  o.cfl(o.SA[c0], o.SB[c0], P);
  for (int k = 0; k < 4; ++k) info[k] = (double)o.step_[o.stepIdx][k];
  o.prim_pass(o.SA[c0], o.SB[c0], 1, P);
  o.flux_pass(P);
  o.update_pass(o.SA[c0], o.SB[c0], o.SA[c0], o.SB[c0], 1, tA, tB, P);
  for (size_t k = 0; k < tA.size(); ++k) { A1[k] = (double)tA[k]; B1[k] = (double)tB[k]; }
  const size_t NL = (size_t)(o.Nx + 1) * (o.Nr + 1);
  for (size_t f = 0; f < NL; ++f)
    for (int q = 0; q < 4; ++q) {
      Fx[f * 12 + q] = (double)o.FxA[f * 4 + q];
      Fx[f * 12 + 4 + q] = (double)o.FxB[f * 4 + q];
      Fx[f * 12 + 8 + q] = (double)o.FxC[f * 4 + q];
      Fr[f * 12 + q] = (double)o.FrA[f * 4 + q];
      Fr[f * 12 + 4 + q] = (double)o.FrB[f * 4 + q];
      Fr[f * 12 + 8 + q] = (double)o.FrC[f * 4 + q];
    }
  std::memcpy(o.step_, saved, sizeof(saved));
  o.stepIdx = savedIdx;
}

}  // namespace

#define DISPATCH(h, CALL) ((h)->prec == 64 ? ((h)->d.CALL) : ((h)->f.CALL))

extern "C" {

void* oracle_create(int Nx, int Nr, int prec) {
  Handle* h = new Handle();
  h->prec = prec == 64 ? 64 : 32;
  if (h->prec == 64) h->d.resize(Nx, Nr); else h->f.resize(Nx, Nr);
  return h;
}
void oracle_destroy(void* p) { delete static_cast<Handle*>(p); }
int oracle_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

void oracle_set_geometry(void* p, const uint8_t* mask, const float* geoA, const float* geoG, const float* geoM, const float* jWall) {
  Handle* h = static_cast<Handle*>(p);
  if (h->prec == 64) set_geo(h->d, mask, geoA, geoG, geoM, jWall); else set_geo(h->f, mask, geoA, geoG, geoM, jWall);
}

// GPUSolver.setState (solver.js:816-828)
void oracle_set_state(void* p, const float* A, const float* B) {
  Handle* h = static_cast<Handle*>(p);
  if (h->prec == 64) set_state(h->d, A, B); else set_state(h->f, A, B);
}
void oracle_set_injector(void* p, const float* tab) {
  Handle* h = static_cast<Handle*>(p);
  h->sprayOn = tab != nullptr;
  if (!tab) return;
  if (h->prec == 64) set_inj(h->d, tab); else set_inj(h->f, tab);
}
void oracle_set_parcels(void* p, int n) {
  Handle* h = static_cast<Handle*>(p);
  if (h->prec == 64) { h->d.NP = n; h->d.spray_alloc(); } else { h->f.NP = n; h->f.spray_alloc(); }
}
void oracle_step(void* p, int n, const nz_params* P) {
  Handle* h = static_cast<Handle*>(p);
  DISPATCH(h, step(n, *P, h->sprayOn));
}
void oracle_rescale(void* p, float fp, float fT, float du, const nz_params* P) {
  Handle* h = static_cast<Handle*>(p);
  if (h->prec == 64) h->d.rescale(fp, fT, du, *P); else h->f.rescale(fp, fT, du, *P);
}
void oracle_read_state(void* p, float* A, float* B) {
  Handle* h = static_cast<Handle*>(p);
  if (h->prec == 64) {
    const size_t N = (size_t)h->d.Nx * h->d.Nr * 4;
    copy_out(A, h->d.SA[h->d.cur], N); copy_out(B, h->d.SB[h->d.cur], N);
  } else {
    const size_t N = (size_t)h->f.Nx * h->f.Nr * 4;
    copy_out(A, h->f.SA[h->f.cur], N); copy_out(B, h->f.SB[h->f.cur], N);
  }
}
void oracle_read_state_f64(void* p, double* A, double* B) {
  Handle* h = static_cast<Handle*>(p);
  if (h->prec == 64) {
    const size_t N = (size_t)h->d.Nx * h->d.Nr * 4;
    std::memcpy(A, h->d.SA[h->d.cur].data(), N * 8); std::memcpy(B, h->d.SB[h->d.cur].data(), N * 8);
  } else {
    const size_t N = (size_t)h->f.Nx * h->f.Nr * 4;
    for (size_t k = 0; k < N; ++k) { A[k] = h->f.SA[h->f.cur][k]; B[k] = h->f.SB[h->f.cur][k]; }
  }
}
void oracle_read_step_info(void* p, float* out) {
  Handle* h = static_cast<Handle*>(p);
  for (int k = 0; k < 4; ++k) out[k] = h->prec == 64 ? (float)h->d.step_[h->d.stepIdx][k] : h->f.step_[h->f.stepIdx][k];
}
void oracle_read_t2(void* p, float* out) {
  Handle* h = static_cast<Handle*>(p);
  if (h->prec == 64) copy_out(out, h->d.T2, h->d.T2.size()); else copy_out(out, h->f.T2, h->f.T2.size());
}
void oracle_wall_row(void* p, float* out) {
  Handle* h = static_cast<Handle*>(p);
  if (h->prec == 64) { h->d.wall_row(); copy_out(out, h->d.wallRow, h->d.wallRow.size()); }
  else { h->f.wall_row(); copy_out(out, h->f.wallRow, h->f.wallRow.size()); }
}
void oracle_read_parcels(void* p, float* PA, float* PB) {
  Handle* h = static_cast<Handle*>(p);
  if (h->prec == 64) { copy_out(PA, h->d.PA[h->d.spCur], h->d.PA[0].size()); copy_out(PB, h->d.PB[h->d.spCur], h->d.PB[0].size()); }
  else { copy_out(PA, h->f.PA[h->f.spCur], h->f.PA[0].size()); copy_out(PB, h->f.PB[h->f.spCur], h->f.PB[0].size()); }
}
void oracle_read_parcels_f64(void* p, double* PA, double* PB) {
  Handle* h = static_cast<Handle*>(p);
  auto get = [&](auto& o) {
    for (size_t k = 0; k < o.PA[0].size(); ++k) { PA[k] = (double)o.PA[o.spCur][k]; PB[k] = (double)o.PB[o.spCur][k]; }
  };
  if (h->prec == 64) get(h->d); else get(h->f);
}
// FS_DERIVED display fields of the current state (velScale: U_REF)
void oracle_derived(void* p, const nz_derived_params* d, double velScale, float* out) {
  Handle* h = static_cast<Handle*>(p);
  if (h->prec == 64) h->d.derived(*d, velScale, out); else h->f.derived(*d, velScale, out);
}
// place parcel k (a = x, r, u, v; b = d, m, type, x_impinge), for single-droplet tests
void oracle_set_parcel(void* p, int k, const float* a, const float* b) {
  Handle* h = static_cast<Handle*>(p);
  auto put = [&](auto& o) {
    if (k < 0 || k >= o.NP) return;
    for (int q = 0; q < 4; ++q) {
      o.PA[o.spCur][(size_t)k * 4 + q] = a[q];
      o.PB[o.spCur][(size_t)k * 4 + q] = b[q];
    }
  };
  if (h->prec == 64) put(h->d); else put(h->f);
}
// pieces of one pass, for closed-form tests of the restatement itself
int oracle_pcg(uint32_t v) { return (int)Oracle<float>::pcg(v); }
float oracle_rnd(uint32_t a, uint32_t n) { return Oracle<float>::rnd(a, n); }
// stage 1 of the next step without committing it (A1, B1: Nx*Nr*4; Fx, Fr: (Nx+1)*(Nr+1)*12;
// info: the step info of that step)
void oracle_stage1_probe(void* p, const nz_params* P, double* A1, double* B1, double* Fx, double* Fr, double* info) {
  Handle* h = static_cast<Handle*>(p);
  if (h->prec == 64) stage1_probe(h->d, *P, A1, B1, Fx, Fr, info);
  else stage1_probe(h->f, *P, A1, B1, Fx, Fr, info);
}
}
