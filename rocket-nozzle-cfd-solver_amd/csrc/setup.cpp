// setup.cpp — host-side solver setup: grid, cut-cell geometry, parameter block,
// initial state and injector table for the nozzle solver.
//
// Restates, in C++ with the same double-precision arithmetic and float32
// storage, what the reference computes in JavaScript before any GPU work:
//   physics.js:7-199    constants, ISA atmosphere, 1-D theory, Rao/cone contour
//   cutcell.js:14-186   r-weighted cut-cell volumes, apertures, sliver merging
//   simsetup.js:7-311   grid snapping, turbulence length, initial state, P block,
//                       injector thermo and the 256-slot spray element table
// Float32 rounding points follow the reference's Float32Array stores so the
// outputs can be compared bit for bit with the reference's own.
#include "nozzle.h"
#include "nz_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace nz {

// ---- constants (physics.js:7-19)
const double G0 = 9.80665;
const double RU = 8314.462618;
const double R_AIR = 287.053;
const double P_REF = 101325.0;
const double T_REF = 288.15;
const double R_REF = R_AIR;
const double RHO_REF = P_REF / (R_REF * T_REF);
const double U_REF = std::sqrt(P_REF / RHO_REF);
const double GAMMA_AIR = 1.4;
const double CHEM_TAU0 = 5.5e-8;

static inline float f32(double v) { return (float)v; }
// JavaScript Math.round: nearest, halves toward +infinity
static inline double js_round(double v) { return std::floor(v + 0.5); }

// ---- ISA 1976 (physics.js:24-50)
static const double ISA[8][4] = {
    {0, 288.15, -0.0065, 101325.0},   {11000, 216.65, 0.0, 22632.06},
    {20000, 216.65, 0.001, 5474.889}, {32000, 228.65, 0.0028, 868.0187},
    {47000, 270.65, 0.0, 110.9063},   {51000, 270.65, -0.0028, 66.93887},
    {71000, 214.65, -0.002, 3.956420}, {86000, 186.87, 0.0, 0.3734}};

Atm atmosphere(double h) {
  h = std::max(0.0, std::min(h, 120000.0));
  const double* L = ISA[0];
  for (int k = 0; k < 8; ++k)
    if (h >= ISA[k][0]) L = ISA[k];
  const double hb = L[0], Tb = L[1], lapse = L[2], pb = L[3];
  const double k = G0 / R_AIR;
  Atm a;
  if (std::fabs(lapse) < 1e-12) {
    a.T = Tb;
    a.p = pb * std::exp(-k * (h - hb) / Tb);
  } else {
    a.T = Tb + lapse * (h - hb);
    a.p = pb * std::pow(Tb / a.T, k / lapse);
  }
  a.rho = a.p / (R_AIR * a.T);
  a.a = std::sqrt(1.4 * R_AIR * a.T);
  return a;
}

// ---- isentropic relations (physics.js:55-72)
double areaRatioFromMach(double M, double g) {
  const double t = (2 / (g + 1)) * (1 + 0.5 * (g - 1) * M * M);
  return std::pow(t, (g + 1) / (2 * (g - 1))) / M;
}
double machFromAreaRatio(double eps, double g, bool supersonic) {
  if (eps <= 1.0) return 1.0;
  double lo = supersonic ? 1.0 : 1e-6, hi = supersonic ? 100.0 : 1.0;
  for (int k = 0; k < 200; ++k) {
    const double mid = 0.5 * (lo + hi);
    const double f = areaRatioFromMach(mid, g) - eps;
    if (supersonic) {
      if (f > 0) hi = mid; else lo = mid;
    } else {
      if (f > 0) lo = mid; else hi = mid;
    }
  }
  return 0.5 * (lo + hi);
}
static double machFromPressureRatio(double p0p, double g) {
  return std::sqrt(std::max(0.0, (std::pow(p0p, (g - 1) / g) - 1) * 2 / (g - 1)));
}

// ---- 1-D ideal nozzle (physics.js:75-106)
Theory nozzleTheory(double Pc, double Tc, double g, double Mw, double Dt, double De, double pa) {
  Theory t;
  const double R = RU / Mw;
  t.At = M_PI * 0.25 * Dt * Dt;
  t.Ae = M_PI * 0.25 * De * De;
  t.eps = t.Ae / t.At;
  t.Me = machFromAreaRatio(t.eps, g, true);
  const double fT = 1 + 0.5 * (g - 1) * t.Me * t.Me;
  t.pe = Pc * std::pow(fT, -g / (g - 1));
  t.Te = Tc / fT;
  t.ue = t.Me * std::sqrt(g * R * t.Te);
  const double Gam = std::sqrt(g) * std::pow(2 / (g + 1), (g + 1) / (2 * (g - 1)));
  t.cstar = std::sqrt(R * Tc) / Gam;
  t.mdot = Pc * t.At / t.cstar;
  t.F = t.mdot * t.ue + (t.pe - pa) * t.Ae;
  t.Fvac = t.mdot * t.ue + t.pe * t.Ae;
  t.Isp = t.F / (t.mdot * G0);
  t.c0 = std::sqrt(g * R * Tc);
  t.uLim = std::sqrt(2 * g / (g - 1) * R * Tc);
  (void)machFromPressureRatio;
  return t;
}

// ---- contour (physics.js:112-199)
static void raoAngles(double eps, double pct, double* thn_out, double* the_out) {
  const double le = std::log10(std::max(eps, 2.0));
  double thn = 14.5 + 11.0 * le;
  double the = 16.5 - 5.3 * le;
  const double d = 80 - pct;
  thn += d * 0.30;
  the += d * 0.22;
  thn = std::min(std::max(thn, 12.0), 48.0);
  the = std::min(std::max(the, 2.0), thn - 4);
  *thn_out = thn * M_PI / 180;
  *the_out = the * M_PI / 180;
}

double Contour::radiusAt(double x) const {
  double r;
  if (x <= 0) r = Rc;
  else if (x >= xe) r = Re;
  else {
    size_t lo = 0, hi = xs.size() - 1;
    while (hi - lo > 1) {
      const size_t m = (lo + hi) >> 1;
      if (xs[m] <= x) lo = m; else hi = m;
    }
    const double t = (x - xs[lo]) / (xs[hi] - xs[lo]);
    r = rs[lo] + t * (rs[hi] - rs[lo]);
  }
  // acoustic-cavity pocket (chamberHardwareContour, simsetup.js:155-159)
  if (cavD > 0 && x < cavW) r = r + cavD;
  return r;
}

Contour buildContour(const nz_config& o) {
  Contour c;
  const double bellPct = o.bellPct != 0 ? o.bellPct : 80;
  c.Rt = 0.5 * o.Dt;
  c.Re = std::max(0.5 * o.De, c.Rt * 1.001);
  c.Rc = c.Rt * std::sqrt(std::max(o.CR, 1.5));
  c.Lc = std::max(o.chamberLen * c.Rc, 1.5 * c.Rt);
  const double convDeg = o.convAngle > 0 ? o.convAngle : 35.0;
  const double thc = convDeg * M_PI / 180;
  const double Ru = 1.5 * c.Rt;
  const double Rf = std::min(0.6 * c.Rc, 1.0 * c.Rt * 1.5);
  auto push = [&](double x, double r) {
    if (c.xs.empty() || x > c.xs.back() + 1e-12) { c.xs.push_back(x); c.rs.push_back(r); }
  };
  push(0, c.Rc);
  push(c.Lc, c.Rc);
  const int nA = 24;
  for (int k = 1; k <= nA; ++k) {
    const double ph = thc * k / nA;
    push(c.Lc + Rf * std::sin(ph), c.Rc - Rf * (1 - std::cos(ph)));
  }
  double x1 = c.xs.back(), r1 = c.rs.back();
  const double r2 = c.Rt + Ru * (1 - std::cos(thc));
  if (r1 > r2) {
    const double x2 = x1 + (r1 - r2) / std::tan(thc);
    push(x2, r2);
    x1 = x2;
    r1 = r2;
  }
  c.xt = x1 + Ru * std::sin(thc);
  for (int k = 1; k <= nA; ++k) {
    const double ph = thc * (1 - (double)k / nA);
    push(c.xt - Ru * std::sin(ph), c.Rt + Ru * (1 - std::cos(ph)));
  }
  const double Rd = 0.382 * c.Rt;
  c.eps = (c.Re * c.Re) / (c.Rt * c.Rt);
  double thn, the, xe;
  if (o.type == NZ_CONE) {
    thn = (o.coneAngle != 0 ? o.coneAngle : 15) * M_PI / 180;
    the = thn;
  } else {
    raoAngles(c.eps, bellPct, &thn, &the);
  }
  for (int k = 1; k <= nA; ++k) {
    const double ph = thn * k / nA;
    push(c.xt + Rd * std::sin(ph), c.Rt + Rd * (1 - std::cos(ph)));
  }
  const double xN = c.xs.back(), rN = c.rs.back();
  if (o.type == NZ_CONE) {
    xe = xN + (c.Re - rN) / std::tan(thn);
    const int n = 40;
    for (int k = 1; k <= n; ++k) push(xN + (xe - xN) * k / n, rN + (c.Re - rN) * k / n);
  } else {
    const double L15 = (c.Re - c.Rt) / std::tan(15 * M_PI / 180);
    xe = c.xt + bellPct / 100 * L15;
    xe = std::max(xe, xN + (c.Re - rN) / std::tan(thn) * 1.02);
    const double tn = std::tan(thn), te = std::tan(the);
    double qx = (c.Re - rN + xN * tn - xe * te) / (tn - te);
    qx = std::min(std::max(qx, xN + 1e-6), xe - 1e-6);
    double qr = rN + tn * (qx - xN);
    qr = std::min(std::max(qr, rN), c.Re * 1.2);
    const int n = 80;
    for (int k = 1; k <= n; ++k) {
      const double s = (double)k / n, a = (1 - s) * (1 - s), b = 2 * s * (1 - s), cc = s * s;
      push(a * xN + b * qx + cc * xe, a * rN + b * qr + cc * c.Re);
    }
  }
  c.xe = c.xs.back();
  c.thn = thn;
  c.the = the;
  return c;
}

// ---- propellant tables (physics.js:215-230)
struct Fluid { double rho, Tb, Tinj, L, cl, sigma, Mw; };
static const Fluid FL_RP1 = {810, 650, 300, 250e3, 2000, 0.02, 100};
static const Fluid FL_CH4 = {420, 190, 112, 300e3, 3500, 0.005, 16};
static const Fluid FL_H2 = {40, 60, 40, 20e3, 14000, 0.001, 2};
static const Fluid FL_LOX = {1140, 155, 90, 150e3, 1700, 0.005, 32};
struct Prop { double Tc, gamma, Mw, qAb, sAir; bool spray; double OF, s; const Fluid* fuel; const Fluid* ox; };
static const Prop PROPS[8] = {
    {3670, 1.22, 22.0, 6.8, 1.8, true, 2.3, 3.4, &FL_RP1, &FL_LOX},   // kerolox
    {3550, 1.21, 20.5, 3.0, 0.8, true, 3.0, 4.0, &FL_CH4, &FL_LOX},   // methalox
    {3400, 1.20, 13.5, 4.2, 1.2, true, 6.0, 8.0, &FL_H2, &FL_LOX},    // hydrolox
    {3150, 1.23, 26.0, 2.0, 0.5, false, 0, 0, nullptr, nullptr},     // hybrid
    {3100, 1.20, 25.0, 4.6, 1.2, false, 0, 0, nullptr, nullptr},     // solid
    {1150, 1.28, 13.0, 14.0, 4.2, false, 0, 0, nullptr, nullptr},    // hydrazine
    {300, 1.40, 28.0, 0, 1, false, 0, 0, nullptr, nullptr},          // coldgas
    {3000, 1.22, 22.0, 3.0, 1, false, 0, 0, nullptr, nullptr},       // custom
};

static bool injectorActive(const nz_config& c) {
  return c.injMode && c.prop >= 0 && c.prop < 8 && PROPS[c.prop].spray;
}

// ---- injector thermo (simsetup.js:138-153)
static InjData injectorThermo(const nz_config& cfg, double ReN, double Ge) {
  const Prop& sp = PROPS[cfg.prop];
  const Fluid& f = *sp.fuel;
  const Fluid& o = *sp.ox;
  auto Rn = [](double Mw) { return (RU / Mw) / R_REF; };
  auto Leff = [](const Fluid& x) { return x.L + x.cl * (x.Tb - x.Tinj); };
  InjData d;
  d.Rf = Rn(f.Mw);
  d.Ro = Rn(o.Mw);
  d.eF = d.Rf * Ge * f.Tb / T_REF - Leff(f) / (U_REF * U_REF);
  d.eO = d.Ro * Ge * o.Tb / T_REF - Leff(o) / (U_REF * U_REF);
  const double Yo = sp.OF / (1 + sp.OF), Yf = 1 - Yo;
  const double burnt = std::min(Yf, Yo / sp.s);
  d.Fres = Yf - burnt;
  d.Rp = (ReN - d.Fres * d.Rf) / std::max(1 - d.Fres, 1e-3);
  d.qHat = (ReN * (Ge + 1) * cfg.Tc / T_REF - (Yf * d.eF + Yo * d.eO)) / burnt;
  d.s = sp.s;
  d.Ge = Ge;
  d.fuelVec[0] = f.rho; d.fuelVec[1] = f.Tb; d.fuelVec[2] = Leff(f); d.fuelVec[3] = f.sigma;
  d.oxVec[0] = o.rho; d.oxVec[1] = o.Tb; d.oxVec[2] = Leff(o); d.oxVec[3] = o.sigma;
  d.fuelRho = f.rho;
  d.oxRho = o.rho;
  return d;
}

// ---- chamber hardware solids (simsetup.js:161-175)
static std::vector<Solid> chamberSolids(const nz_config& cfg, const Contour& c, double dx) {
  std::vector<Solid> out;
  auto snap = [dx](double v) { return js_round(v / dx) * dx; };
  if (cfg.injType == NZ_INJ_PINTLE) out.push_back({-dx, snap(0.5 * c.Rc), -dx, snap(0.22 * c.Rc)});
  const int n = (int)js_round(cfg.baffles);
  const double h = snap(cfg.baffleH / 100 * c.Rc);
  const double t = std::max(2 * dx, snap(0.04 * c.Rc));
  for (int k = 1; k <= n && h > 0; ++k) {
    const double r = snap((double)k / (n + 1) * c.Rc);
    out.push_back({-dx, h, r - t / 2, r + t / 2});
  }
  if (cfg.cavD > 0) {
    const double w = snap(cfg.cavW / 100 * c.Rc), d = (cfg.cavD / 100) * c.Rc;
    const double e = std::max(3 * dx, 0.12 * c.Rt);
    out.push_back({w, w + e, c.Rc, c.Rc + d + e});
  }
  return out;
}

// ---- cut-cell geometry (cutcell.js:14-186)
std::vector<float> contourTable(const Contour& contour, int Nx, double dx) {
  const int SUB = 8, nTab = Nx * SUB + 1;
  std::vector<float> Rtab(nTab);
  for (int k = 0; k < nTab; ++k) {
    const double X = ((double)k / SUB) * dx;
    Rtab[k] = f32(X <= contour.xe ? contour.radiusAt(X) / dx : -1e9);
  }
  return Rtab;
}

void buildCutGeometry(const CutInput& o, CutGeometry& G) {
  const int Nx = o.Nx, Nr = o.Nr;
  const double dx = o.dx;
  const Contour& contour = *o.contour;
  const bool cut = !o.stair;
  const double xeC = contour.xe / dx, twC = o.tw / dx, xBaseC = o.xBase / dx, bodyC = o.bodyR / dx;
  const int SUB = 8;
  const int nTab = Nx * SUB + 1;
  const std::vector<float> Rtab = contourTable(contour, Nx, dx);
  auto RinAt = [&](double xc) -> double {
    if (xc > xeC) return -1e9;
    const double f = std::min(std::max(xc * SUB, 0.0), (double)(nTab - 1));
    const int k = (int)std::min(std::floor(f), (double)(nTab - 2));
    const double t = f - k;
    const double a = Rtab[k], b = Rtab[k + 1];
    if (a < -1e8 || b < -1e8) return contour.radiusAt(std::min(xc * dx, contour.xe)) / dx;
    return a + (b - a) * t;
  };
  std::vector<Solid> solids;
  for (const Solid& s : o.solids) solids.push_back({s.x0 / dx, s.x1 / dx, s.r0 / dx, s.r1 / dx});
  auto fluid = [&](double xc, double rc) -> bool {
    for (const Solid& s : solids)
      if (xc >= s.x0 && xc <= s.x1 && rc >= s.r0 && rc <= s.r1) return false;
    if (xc <= xeC) {
      const double ri = RinAt(xc);
      if (rc >= ri && rc <= ri + twC) return false;
      if (o.base && xc < xBaseC && rc >= ri && rc < bodyC) return false;
    }
    return true;
  };
  std::vector<float> colMin(Nx), colMax(Nx);
  for (int i = 0; i < Nx; ++i) {
    double mn = 1e9, mx = -1e9;
    for (int a = 0; a <= SUB; ++a) {
      const double r = Rtab[i * SUB + a];
      if (r > -1e8) { mn = std::min(mn, r); mx = std::max(mx, r); }
    }
    colMin[i] = f32(mn);
    colMax[i] = f32(mx);
  }
  auto overlaps = [](double a0, double a1, double b0, double b1) { return a0 < b1 && a1 > b0; };
  auto maybeCut = [&](int i, int j) -> bool {
    if (!cut) return false;
    const double x0 = i, x1 = i + 1, r0 = j, r1 = j + 1;
    for (const Solid& s : solids)
      if (overlaps(x0, x1, s.x0 - 1e-6, s.x1 + 1e-6) && overlaps(r0, r1, s.r0 - 1e-6, s.r1 + 1e-6)) return true;
    if (colMax[i] < -1e8) return false;
    const double mn = colMin[i], mx = colMax[i];
    if (overlaps(r0, r1, mn - 1e-6, mx + 1e-6)) return true;
    if (overlaps(r0, r1, mn + twC - 1e-6, mx + twC + 1e-6)) return true;
    if (xeC >= x0 && xeC <= x1 && overlaps(r0, r1, mn, mx + twC)) return true;
    if (o.base) {
      if (xBaseC >= x0 && xBaseC <= x1 && overlaps(r0, r1, mn, bodyC)) return true;
      if (x0 < xBaseC && bodyC >= r0 && bodyC <= r1) return true;
    }
    return false;
  };

  const int NC = 8, NF = 12;
  const size_t NN = (size_t)Nx * Nr;
  std::vector<float> V(NN, 0.f), Am(NN, 0.f);
  std::vector<uint8_t> isCut(NN, 0);
  for (int j = 0; j < Nr; ++j)
    for (int i = 0; i < Nx; ++i) {
      const size_t k = (size_t)j * Nx + i;
      if (maybeCut(i, j)) {
        double am = 0, v = 0;
        for (int b = 0; b < NC; ++b) {
          const double rc = j + (b + 0.5) / NC;
          for (int a = 0; a < NC; ++a)
            if (fluid(i + (a + 0.5) / NC, rc)) { am += 1; v += rc; }
        }
        Am[k] = f32(am / (NC * NC));
        V[k] = f32(v / (NC * NC));
        isCut[k] = (Am[k] > 0 && Am[k] < 1) ? 1 : 0;
      } else if (fluid(i + 0.5, j + 0.5)) {
        Am[k] = 1;
        V[k] = f32(j + 0.5);
      }
    }
  const int W1 = Nx + 1;
  const size_t NL = (size_t)W1 * (Nr + 1);
  std::vector<float> Ax(NL, 0.f), Ar(NL, 0.f);
  auto cutAt = [&](int i, int j) -> int {
    return (i >= 0 && i < Nx && j >= 0 && j < Nr) ? isCut[(size_t)j * Nx + i] : 0;
  };
  for (int j = 0; j < Nr; ++j)
    for (int i = 0; i <= Nx; ++i) {
      double a;
      if (cut && (cutAt(i - 1, j) || cutAt(i, j) || maybeCut(std::min(i, Nx - 1), j) || (i > 0 && maybeCut(i - 1, j)))) {
        a = 0;
        for (int b = 0; b < NF; ++b) {
          const double rc = j + (b + 0.5) / NF;
          if (fluid(i, rc)) a += rc;
        }
        a /= NF;
      } else {
        a = fluid(i, j + 0.5) ? j + 0.5 : 0;
      }
      Ax[(size_t)j * W1 + i] = f32(a);
    }
  for (int j = 0; j <= Nr; ++j)
    for (int i = 0; i < Nx; ++i) {
      double a;
      if (cut && (cutAt(i, j - 1) || cutAt(i, j) || (j < Nr && maybeCut(i, j)) || (j > 0 && maybeCut(i, j - 1)))) {
        double f = 0;
        for (int b = 0; b < NF; ++b)
          if (fluid(i + (b + 0.5) / NF, j)) f++;
        a = j * f / NF;
      } else {
        a = fluid(i + 0.5, j) ? (double)j : 0;
      }
      Ar[(size_t)j * W1 + i] = f32(a);
    }

  // classification (cutcell.js:106-120)
  std::vector<uint8_t> solid(NN, 0);
  std::vector<float> kap(NN);
  for (int j = 0; j < Nr; ++j)
    for (int i = 0; i < Nx; ++i) {
      const size_t k = (size_t)j * Nx + i;
      kap[k] = f32((double)V[k] / (j + 0.5));
      if (kap[k] < 0.02) solid[k] = 1;
    }
  auto zeroSolidFaces = [&]() {
    for (int j = 0; j < Nr; ++j)
      for (int i = 0; i < Nx; ++i) {
        if (!solid[(size_t)j * Nx + i]) continue;
        Ax[(size_t)j * W1 + i] = 0;
        Ax[(size_t)j * W1 + i + 1] = 0;
        Ar[(size_t)j * W1 + i] = 0;
        Ar[(size_t)(j + 1) * W1 + i] = 0;
      }
  };
  zeroSolidFaces();
  // sliver merging (cutcell.js:121-142); eligible hosts have kappa >= 0.5 so
  // the result does not depend on the visiting order
  std::vector<int64_t> host(NN, -1);
  const double KMERGE = 0.5;
  for (int pass = 0; pass < 2; ++pass)
    for (int j = 0; j < Nr; ++j)
      for (int i = 0; i < Nx; ++i) {
        const size_t k = (size_t)j * Nx + i;
        if (solid[k] || kap[k] >= KMERGE || host[k] >= 0) continue;
        int64_t best = -1;
        double bestA = 0;
        const int ci[4] = {i - 1, i + 1, i, i};
        const int cj[4] = {j, j, j - 1, j + 1};
        const double ca[4] = {(double)Ax[(size_t)j * W1 + i] / (j + 0.5),
                              (double)Ax[(size_t)j * W1 + i + 1] / (j + 0.5),
                              j > 0 ? (double)Ar[(size_t)j * W1 + i] / j : 0.0,
                              (double)Ar[(size_t)(j + 1) * W1 + i] / (j + 1)};
        for (int c = 0; c < 4; ++c) {
          const int ii = ci[c], jj = cj[c];
          if (ii < 0 || ii >= Nx || jj < 0 || jj >= Nr) continue;
          const size_t kk = (size_t)jj * Nx + ii;
          if (solid[kk] || kap[kk] < KMERGE || host[kk] >= 0) continue;
          if (ca[c] > bestA + 1e-9) { bestA = ca[c]; best = (int64_t)kk; }
        }
        if (best >= 0 && bestA > 0.02) host[k] = best;
        else if (pass == 1) solid[k] = 1;
      }
  zeroSolidFaces();
  // first-order flags (cutcell.js:144-153)
  auto partial = [&](int i, int j) -> bool {
    if (!(i >= 0 && i < Nx && j >= 0 && j < Nr)) return false;
    const size_t k = (size_t)j * Nx + i;
    return !solid[k] && (kap[k] < 0.98 || host[k] >= 0);
  };
  G.Nx = Nx;
  G.Nr = Nr;
  G.geoA.assign(NL * 4, 0.f);
  for (int j = 0; j <= Nr; ++j)
    for (int i = 0; i <= Nx; ++i) {
      const size_t q = ((size_t)j * W1 + i) * 4;
      if (i < Nx && j < Nr) {
        const size_t k = (size_t)j * Nx + i;
        G.geoA[q] = V[k];
        G.geoA[q + 1] = Am[k];
      }
      if (j < Nr) {
        const float a = Ax[(size_t)j * W1 + i];
        G.geoA[q + 2] = (partial(i - 1, j) || partial(i, j)) ? -a : a;
      }
      if (i < Nx) {
        const float a = Ar[(size_t)j * W1 + i];
        G.geoA[q + 3] = (partial(i, j - 1) || partial(i, j)) ? -a : a;
      }
    }
  // wall vectors by closure and group sums (cutcell.js:154-184)
  std::vector<float> Wx(NN, 0.f), Wr(NN, 0.f);
  for (int j = 0; j < Nr; ++j)
    for (int i = 0; i < Nx; ++i) {
      const size_t k = (size_t)j * Nx + i;
      if (solid[k]) continue;
      // evaluated in double and rounded once, as the Float32Array store does
      Wx[k] = f32((double)Ax[(size_t)j * W1 + i] - (double)Ax[(size_t)j * W1 + i + 1]);
      Wr[k] = f32((double)Am[k] - ((double)Ar[(size_t)(j + 1) * W1 + i] - (double)Ar[(size_t)j * W1 + i]));
      if (std::fabs(Wx[k]) < 1e-6) Wx[k] = 0;
      if (std::fabs(Wr[k]) < 1e-6) Wr[k] = 0;
    }
  std::vector<float> gV(NN, 0.f), gA(NN, 0.f), gX(NN, 0.f), gR(NN, 0.f);
  std::vector<uint8_t> slaveMask(NN, 0);
  for (size_t k = 0; k < NN; ++k)
    if (!solid[k]) {
      const size_t h = host[k] >= 0 ? (size_t)host[k] : k;
      gV[h] += V[k];
      gA[h] += Am[k];
      gX[h] += Wx[k];
      gR[h] += Wr[k];
    }
  for (int j = 0; j < Nr; ++j)
    for (int i = 0; i < Nx; ++i) {
      const size_t k = (size_t)j * Nx + i;
      if (solid[k] || host[k] < 0) continue;
      const int64_t h = host[k];
      const int hi = (int)(h % Nx), hj = (int)((h - hi) / Nx);
      const int bit = i < hi ? 1 : i > hi ? 2 : j < hj ? 4 : 8;
      slaveMask[h] |= bit;
    }
  G.geoG.assign(NN * 4, 0.f);
  G.geoM.assign(NN * 4, 0.f);
  G.nCut = G.nMerged = 0;
  for (int j = 0; j < Nr; ++j)
    for (int i = 0; i < Nx; ++i) {
      const size_t k = (size_t)j * Nx + i, q = k * 4;
      if (solid[k]) { G.geoG[q] = 1; continue; }
      const size_t h = host[k] >= 0 ? (size_t)host[k] : k;
      const int hi = (int)(h % Nx), hj = (int)((h - hi) / Nx);
      G.geoG[q] = gV[h]; G.geoG[q + 1] = gA[h]; G.geoG[q + 2] = gX[h]; G.geoG[q + 3] = gR[h];
      G.geoM[q] = (float)(hi - i); G.geoM[q + 1] = (float)(hj - j);
      G.geoM[q + 2] = slaveMask[k]; G.geoM[q + 3] = kap[k];
      if (kap[k] < 0.98) G.nCut++;
      if (host[k] >= 0) G.nMerged++;
    }
  G.mask = solid;
  G.kappa = kap;
  G.Ax = Ax;
  G.Ar = Ar;
}

// 1-D inflow speed estimate (simsetup.js:251-254)
static double chamberInflowSpeed(const nz_config& cfg, double c0) {
  const double M = machFromAreaRatio(std::max(cfg.CR, 1.05), cfg.gamma, false);
  return M * c0 / std::sqrt(1 + 0.5 * (cfg.gamma - 1) * M * M);
}

Atm ambientOf(const nz_config& cfg) {
  Atm a = atmosphere(cfg.alt);
  if (cfg.ambP > 0 || cfg.ambT > 0) {
    if (cfg.ambP > 0) a.p = cfg.ambP;
    if (cfg.ambT > 0) a.T = cfg.ambT;
    a.rho = a.p / (R_AIR * a.T);
    a.a = std::sqrt(1.4 * R_AIR * a.T);
  }
  return a;
}

// ---- computeParams (simsetup.js:257-311)
ParamsD computeParams(const nz_config& cfg, const Contour& contour, double dx, double PcSi, double pStartSi) {
  ParamsD P;
  const double g = cfg.gamma, Ge = 1 / (g - 1), Ga = 1 / (GAMMA_AIR - 1);
  const Atm atm = ambientOf(cfg);
  const double pa = atm.p / P_REF;
  const double rhoa = pa / (atm.T / T_REF);
  const double ca = std::sqrt(GAMMA_AIR * pa / rhoa);
  const double uinf = cfg.mach * ca;
  const double ReN = (RU / cfg.Mw) / R_REF;
  const double TcN = cfg.Tc / T_REF;
  const double PcN = PcSi / P_REF;
  const double c0 = std::sqrt(g * ReN * TcN);
  const double uLim = std::sqrt(2 * g / (g - 1) * ReN * TcN);
  const double tScale = dx / U_REF;
  const bool injOn = injectorActive(cfg);
  P.inj = injOn;
  if (injOn) P.injData = injectorThermo(cfg, ReN, Ge);
  const bool chem = injOn || (cfg.afterburn && cfg.qAb > 0);
  double thMdot = 0;
  if (injOn) thMdot = nozzleTheory(PcSi, cfg.Tc, g, cfg.Mw, cfg.Dt, cfg.De, atm.p).mdot;
  if (injOn) { P.Rs[0] = P.injData.Rf; P.Rs[1] = P.injData.Ro; P.Rs[2] = P.injData.Rp; }
  else { P.Rs[0] = P.Rs[1] = P.Rs[2] = ReN; }
  if (injOn) { P.chamberNu[0] = 0.01 * chamberInflowSpeed(cfg, c0) * contour.Rc / dx; P.chamberNu[1] = contour.xt / dx; }
  else { P.chamberNu[0] = 0; P.chamberNu[1] = 0; }
  if (injOn) { P.cap[0] = 1.02 * cfg.Tc; P.cap[1] = 1.25 * cfg.Tc; }
  else { P.cap[0] = 2400; P.cap[1] = 2800; }
  if (injOn) {
    P.sprayMhat = thMdot / (2 * M_PI * RHO_REF * U_REF * dx * dx);
    P.sprayMdot = thMdot;
    P.sprayTScale = dx / U_REF;
  }
  P.gamma = g;
  P.amb[0] = rhoa; P.amb[1] = uinf; P.amb[2] = 0; P.amb[3] = pa;
  P.ambCons[0] = rhoa; P.ambCons[1] = rhoa * uinf; P.ambCons[2] = 0; P.ambCons[3] = pa * Ga + 0.5 * rhoa * uinf * uinf;
  P.ambB[0] = 0; P.ambB[1] = 0; P.ambB[2] = Ga; P.ambB[3] = 0;
  P.gamR[0] = std::min(Ge, Ga); P.gamR[1] = std::max(Ge, Ga);
  P.inlet[0] = c0 / (6 * contour.xt / dx);
  P.inlet[1] = cfg.inletAbsorb;
  P.inlet[2] = Ge;
  P.inlet[3] = Ga;
  P.uInlet0 = chamberInflowSpeed(cfg, c0);
  P.visc = cfg.viscous != 0;
  P.chem = chem;
  P.gas[0] = 1.0; P.gas[1] = ReN; P.gas[2] = T_REF; P.gas[3] = U_REF / (dx * P_REF);
  if (injOn) {
    P.chemK[0] = P.injData.qHat; P.chemK[1] = P.injData.s; P.chemK[2] = (dx / U_REF) / CHEM_TAU0; P.chemK[3] = pa;
  } else {
    P.chemK[0] = cfg.qAb * 1e6 / (U_REF * U_REF);
    P.chemK[1] = std::max(cfg.sAir != 0 ? cfg.sAir : 1.0, 0.05);
    P.chemK[2] = (dx / U_REF) / CHEM_TAU0;
    P.chemK[3] = pa;
  }
  P.Tw = std::min(cfg.Tw != 0 ? cfg.Tw : 800.0, std::max(cfg.Tc, 1.0)) / T_REF;
  P.res[0] = PcN; P.res[1] = TcN; P.res[2] = ReN;
  P.pStart = std::min(pStartSi / P_REF, PcN);
  P.ramp = cfg.startMode == NZ_START_BURST ? 0 : (cfg.rampMs * 1e-3) / tScale;
  P.RcCells = (contour.Rc + (injOn ? cfg.cavD / 100 * contour.Rc : 0)) / dx + 0.5;
  P.cfl = cfg.cfl != 0 ? cfg.cfl : 0.6;
  P.dtMax = 10.0;
  P.wFloor = 2.0 * c0 + std::fabs(uinf);
  P.vMax = 3.0 * (uLim + c0 + std::fabs(uinf) + ca);
  P.hllBlend = cfg.hllBlend;
  P.order2 = cfg.order2 != 0;
  P.ReN = ReN;
  P.Ta = atm.T;
  P.pa = atm.p;
  P.uinf = uinf;
  return P;
}

void paramsToFloat(const ParamsD& d, nz_params& P) {
  std::memset(&P, 0, sizeof(P));
  for (int k = 0; k < 4; ++k) {
    P.amb[k] = f32(d.amb[k]); P.inlet[k] = f32(d.inlet[k]); P.ambCons[k] = f32(d.ambCons[k]);
    P.ambB[k] = f32(d.ambB[k]); P.gas[k] = f32(d.gas[k]); P.chemK[k] = f32(d.chemK[k]);
    P.fuelVec[k] = f32(d.injData.fuelVec[k]); P.oxVec[k] = f32(d.injData.oxVec[k]);
  }
  for (int k = 0; k < 3; ++k) { P.res[k] = f32(d.res[k]); P.Rs[k] = f32(d.Rs[k]); }
  for (int k = 0; k < 2; ++k) { P.gamR[k] = f32(d.gamR[k]); P.chamberNu[k] = f32(d.chamberNu[k]); P.cap[k] = f32(d.cap[k]); }
  P.pStart = f32(d.pStart); P.ramp = f32(d.ramp); P.RcCells = f32(d.RcCells);
  P.inj = d.inj ? 1 : 0;
  P.order2 = d.order2 ? 1 : 0;
  P.hllBlend = f32(d.hllBlend);
  P.uInlet0 = f32(d.uInlet0); P.vMax = f32(d.vMax);
  P.cfl = f32(d.cfl); P.dtMax = f32(d.dtMax); P.wFloor = f32(d.wFloor);
  P.visc = d.visc ? 1 : 0; P.chem = d.chem ? 1 : 0;
  P.Tw = f32(d.Tw);
  P.sprayMhat = f32(d.sprayMhat); P.sprayTScale = f32(d.sprayTScale);
  P.eF = f32(d.injData.eF); P.eO = f32(d.injData.eO); P.Ge = f32(d.injData.Ge);
}

// double-precision values in nz_params field order
int paramsToDoubles(const ParamsD& d, double* out, int n) {
  std::vector<double> v;
  for (int k = 0; k < 4; ++k) v.push_back(d.amb[k]);
  for (int k = 0; k < 3; ++k) v.push_back(d.res[k]);
  for (int k = 0; k < 4; ++k) v.push_back(d.inlet[k]);
  v.push_back(d.pStart); v.push_back(d.ramp); v.push_back(d.RcCells); v.push_back(d.inj ? 1 : 0);
  v.push_back(d.order2 ? 1 : 0); v.push_back(d.hllBlend);
  for (int k = 0; k < 4; ++k) v.push_back(d.ambCons[k]);
  for (int k = 0; k < 4; ++k) v.push_back(d.ambB[k]);
  for (int k = 0; k < 2; ++k) v.push_back(d.gamR[k]);
  v.push_back(d.uInlet0); v.push_back(d.vMax);
  v.push_back(d.cfl); v.push_back(d.dtMax); v.push_back(d.wFloor);
  v.push_back(d.visc ? 1 : 0); v.push_back(d.chem ? 1 : 0);
  for (int k = 0; k < 4; ++k) v.push_back(d.gas[k]);
  for (int k = 0; k < 3; ++k) v.push_back(d.Rs[k]);
  for (int k = 0; k < 2; ++k) v.push_back(d.chamberNu[k]);
  for (int k = 0; k < 2; ++k) v.push_back(d.cap[k]);
  for (int k = 0; k < 4; ++k) v.push_back(d.chemK[k]);
  v.push_back(d.Tw);
  v.push_back(d.sprayMhat); v.push_back(d.sprayTScale);
  for (int k = 0; k < 4; ++k) v.push_back(d.injData.fuelVec[k]);
  for (int k = 0; k < 4; ++k) v.push_back(d.injData.oxVec[k]);
  v.push_back(d.injData.eF); v.push_back(d.injData.eO); v.push_back(d.injData.Ge);
  const int m = std::min(n, (int)v.size());
  for (int k = 0; k < m; ++k) out[k] = v[k];
  return (int)v.size();
}

// ---- injector element table (simsetup.js:181-237)
static std::vector<float> buildInjectorTable(const nz_config& cfg, const Contour& c, double dx, const ParamsD& P) {
  const InjData& d = P.injData;
  const double Pc = cfg.Pc, dp = (cfg.injDP != 0 ? cfg.injDP : 20) / 100 * Pc;
  const double OF = cfg.OF > 0 ? cfg.OF : PROPS[cfg.prop].OF;
  const double Vf = 0.8 * std::sqrt(2 * dp / d.fuelRho) / U_REF, Vo = 0.8 * std::sqrt(2 * dp / d.oxRho) / U_REF;
  const double Do = (cfg.injD != 0 ? cfg.injD : 1.5) * 1e-3;
  const int n = std::max(1, (int)js_round(cfg.injRings != 0 ? cfg.injRings : 4));
  const double mf = 1 / (1 + OF), mo = OF / (1 + OF), film = std::min(cfg.film / 100, 0.5);
  const double Ri = 0.9 * c.Rc, deg = M_PI / 180, C = 1 / dx;
  struct El { double type, weight, x, r, angle, speed, drop, cone, width, xi; };
  std::vector<El> el;
  auto add = [&](double type, double weight, double x, double r, double angle, double speed, double drop,
                 double cone, double width, double xi) {
    el.push_back({type, weight, x * C, r * C, angle, speed, drop, cone, width * C, xi * C});
  };
  const double x0 = 0.5 * dx;
  auto ringR = [&](int k, int m) { return (k + 0.5) / m * Ri; };
  auto shareOf = [](const std::vector<double>& rs) {
    double s = 0;
    for (double r : rs) s += r;
    if (s == 0) s = 1;
    std::vector<double> o;
    for (double r : rs) o.push_back(r / s);
    return o;
  };
  if (cfg.injType == NZ_INJ_COAX) {
    std::vector<double> rs;
    for (int k = 0; k < n; ++k) rs.push_back(ringR(k, n));
    const std::vector<double> w = shareOf(rs);
    const double g = 1.2 * Do;
    for (int k = 0; k < n; ++k) {
      const double r = rs[k];
      add(1, mo * w[k], x0, r, 0, Vo, 0.035 * Do, 3 * deg, 0.5 * Do, 0);
      add(0, 0.5 * mf * (1 - film) * w[k], x0, r + g, 0, 3 * Vf, 0.015 * Do, 8 * deg, 0.4 * Do, 0);
      add(0, 0.5 * mf * (1 - film) * w[k], x0, std::max(r - g, 0.2 * Do), 0, 3 * Vf, 0.015 * Do, 8 * deg, 0.4 * Do, 0);
    }
  } else if (cfg.injType == NZ_INJ_IMPINGE) {
    const double th = (cfg.injAngle != 0 ? cfg.injAngle : 30) * deg, xi = 6 * Do, del = xi * std::tan(th);
    std::vector<double> rs;
    for (int k = 0; k < n; ++k) rs.push_back(std::max(ringR(k, n), del + Do));
    const std::vector<double> w = shareOf(rs);
    for (int k = 0; k < n; ++k) {
      const double r = rs[k];
      add(0, mf * (1 - film) * w[k], x0, r + del, -th, Vf, 0.12 * Do, 1 * deg, 0.5 * Do, xi);
      add(1, mo * w[k], x0, r - del, th, Vo, 0.12 * Do, 1 * deg, 0.5 * Do, xi);
    }
  } else if (cfg.injType == NZ_INJ_PINTLE) {
    const double Rp = 0.22 * c.Rc, Lp = 0.5 * c.Rc;
    const double th = std::atan(std::sqrt(std::max(cfg.injTMR != 0 ? cfg.injTMR : 1.0, 0.01)));
    const double xt = Lp + 1.5 * dx;
    add(1, mo, xt, Rp, th, Vo, 0.04 * Do, 6 * deg, 0.6 * Do, 0);
    add(0, mf * (1 - film), xt, Rp + Do, th, Vf, 0.04 * Do, 6 * deg, 0.6 * Do, 0);
  } else {
    const int m = std::max(2, n);
    std::vector<double> rs, rf, ro;
    for (int k = 0; k < m; ++k) rs.push_back(ringR(k, m));
    for (int k = 0; k < m; ++k) (k % 2 == 0 ? rf : ro).push_back(rs[k]);
    const std::vector<double> f = shareOf(rf), o = shareOf(ro);
    int a = 0, b = 0;
    for (int k = 0; k < m; ++k) {
      if (k % 2 == 0) add(0, mf * (1 - film) * f[a++], x0, rs[k], 0, Vf, 0.03 * Do, 4 * deg, 0.6 * Do, 0);
      else add(1, mo * o[b++], x0, rs[k], 0, Vo, 0.03 * Do, 4 * deg, 0.6 * Do, 0);
    }
  }
  // drop-size law (an extension; 0 keeps the reference's): every element's median drop at
  // dropD0Frac x the orifice, lognormal spread dropSigma (the film stays as the reference has it)
  if (cfg.dropD0Frac > 0)
    for (El& e : el) e.drop = cfg.dropD0Frac * Do;
  if (film > 0) add(0, mf * film, x0, 0.97 * c.Rc, 2 * deg, 0.5 * Vf, 80e-6, 1 * deg, 0.015 * c.Rc, 0);
  // largest-remainder fill of 256 slots; the reference sorts with
  // Array.prototype.sort, which is stable, so ties keep element order
  double total = 0;
  for (const El& e : el) total += e.weight;
  const int N = 256;
  struct Slot { size_t e; double q; int n; };
  std::vector<Slot> slots;
  for (size_t k = 0; k < el.size(); ++k) slots.push_back({k, el[k].weight / total * N, 0});
  int used = 0;
  for (Slot& s : slots) { s.n = (int)std::floor(s.q); used += s.n; }
  std::vector<size_t> order(slots.size());
  for (size_t k = 0; k < order.size(); ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return (slots[b].q - slots[b].n) < (slots[a].q - slots[a].n);
  });
  for (size_t k : order)
    if (used < N) { slots[k].n++; used++; }
  std::vector<float> tab((size_t)N * 3 * 4, 0.f);
  int t = 0;
  for (const Slot& s : slots)
    for (int q = 0; q < s.n; ++q, ++t) {
      const El& e = el[s.e];
      float* r0 = &tab[(size_t)t * 4];
      r0[0] = f32(e.x); r0[1] = f32(e.r); r0[2] = f32(e.angle); r0[3] = f32(e.speed);
      float* r1 = &tab[(size_t)(N + t) * 4];
      r1[0] = f32(e.type); r1[1] = f32(e.drop); r1[2] = f32(e.cone); r1[3] = f32(e.width);
      float* r2 = &tab[(size_t)(2 * N + t) * 4];
      r2[0] = f32(e.xi);
      r2[1] = cfg.dropSigma > 0 ? f32(cfg.dropSigma) : 0.f;   // > 0: lognormal drop-size draw
    }
  return tab;
}

// the cut-cell builder's input for this setup (pointers into the setup)
void cutcellInput(const Setup& S, nz_cutcell_in& in) {
  std::memset(&in, 0, sizeof(in));
  in.Nx = S.Nx; in.Nr = S.Nr; in.dx = S.dx; in.tw = S.tw; in.xBase = S.xBase; in.bodyR = S.bodyR;
  in.base = S.cfg.base ? 1 : 0;
  in.stair = S.cfg.wallModel == NZ_WALL_STAIR ? 1 : 0;
  in.Rtab = S.Rtab.data();
  in.npts = (int32_t)S.contour.xs.size();
  in.xs = S.contour.xs.data();
  in.rs = S.contour.rs.data();
  in.Rc = S.contour.Rc; in.Re = S.contour.Re; in.xe = S.contour.xe;
  in.cavD = S.contour.cavD; in.cavW = S.contour.cavW;
  in.nsolids = (int32_t)(S.solidsFlat.size() / 4);
  in.solids = S.solidsFlat.empty() ? nullptr : S.solidsFlat.data();
}

// ---- buildSimSetup (simsetup.js:7-126)
Setup* buildSimSetup(const nz_config& cfg_in) {
  Setup* S = new Setup();
  nz_config cfg = cfg_in;
  const int Nx = cfg.Nx;
  const int Nr = std::max(16, (int)js_round(Nx * cfg.aspect));
  const bool injOn = injectorActive(cfg);
  Contour contour = buildContour(cfg);
  if (injOn && cfg.cavD > 0) {
    contour.cavW = cfg.cavW / 100 * contour.Rc;
    contour.cavD = cfg.cavD / 100 * contour.Rc;
  }
  if (injOn && !(cfg.dx > 0))
    cfg.domainL = std::max(1.3 * contour.xe,
                           1.25 * (contour.Re + std::max(contour.Rc, contour.Re) * 0.12) / cfg.aspect);
  double dx = cfg.domainL / Nx;
  int nt = 0;
  bool gridLimited = false;
  if (cfg.dx > 0) dx = cfg.dx;
  else {
    const int minCells = cfg.minCells != 0 ? cfg.minCells : 8;
    const int want = std::max(minCells, (int)js_round(contour.Rt / dx));
    const int fit = (int)std::floor(contour.Rt / (contour.xe / (0.75 * Nx)));
    nt = std::max(3, std::min(want, std::max(fit, std::min((int)js_round(contour.Rt / dx), want))));
    gridLimited = nt < minCells;
    dx = contour.Rt / nt;
  }
  (void)gridLimited;
  const Atm atm = ambientOf(cfg);
  const double g = cfg.gamma, Ge = 1 / (g - 1), Ga = 1 / (GAMMA_AIR - 1);
  S->theory = nozzleTheory(cfg.Pc, cfg.Tc, g, cfg.Mw, cfg.Dt, cfg.De, atm.p);

  const double tw = std::max(2.6 * dx, 0.12 * contour.Rt);
  const double bodyR = cfg.base ? std::max(contour.Rc, contour.Re) * (cfg.bodyRatio != 0 ? cfg.bodyRatio : 1.5) + tw : 0;
  const double xBase = contour.xt;
  std::vector<float> RinCol(Nx);
  for (int i = 0; i < Nx; ++i) {
    const double x = (i + 0.5) * dx;
    RinCol[i] = x <= contour.xe ? f32(contour.radiusAt(x)) : std::numeric_limits<float>::infinity();
  }
  const std::vector<Solid> hw = injOn ? chamberSolids(cfg, contour, dx) : std::vector<Solid>();
  CutInput ci;
  ci.Nx = Nx; ci.Nr = Nr; ci.dx = dx; ci.contour = &contour; ci.tw = tw; ci.base = cfg.base != 0;
  ci.xBase = xBase; ci.bodyR = bodyR; ci.stair = cfg.wallModel == NZ_WALL_STAIR; ci.solids = hw;
  S->cfg = cfg; S->Nx = Nx; S->Nr = Nr; S->dx = dx; S->tw = tw; S->bodyR = bodyR; S->xBase = xBase;
  S->contour = contour;
  S->Rtab = contourTable(contour, Nx, dx);
  S->solidsFlat.clear();
  for (const Solid& q : hw) S->solidsFlat.insert(S->solidsFlat.end(), {q.x0, q.x1, q.r0, q.r1});
  if (cfg.deviceGeometry) {
    // the same geometry, built by the GPU kernels (cutcell.hip)
    CutGeometry& G = S->geo;
    const size_t N = (size_t)Nx * Nr, NL = (size_t)(Nx + 1) * (Nr + 1);
    G.Nx = Nx; G.Nr = Nr;
    G.mask.assign(N, 0); G.geoA.assign(NL * 4, 0.f); G.geoG.assign(N * 4, 0.f); G.geoM.assign(N * 4, 0.f);
    G.kappa.assign(N, 0.f); G.Ax.assign(NL, 0.f); G.Ar.assign(NL, 0.f);
    nz_cutcell_in in;
    cutcellInput(*S, in);
    int32_t counts[2] = {0, 0};
    if (nz_cutcell_build(&in, G.mask.data(), G.geoA.data(), G.geoG.data(), G.geoM.data(), G.kappa.data(), G.Ax.data(),
                         G.Ar.data(), counts) != 0) {
      delete S;
      return nullptr;
    }
    G.nCut = counts[0];
    G.nMerged = counts[1];
  } else {
    buildCutGeometry(ci, S->geo);
  }
  CutGeometry& geo = S->geo;
  const std::vector<uint8_t>& mask = geo.mask;
  // turbulence length scale and boundary-layer band flag (simsetup.js:51-63)
  const double CS = 0.17, KAPPA = 0.41;
  for (int j = 0; j < Nr; ++j)
    for (int i = 0; i < Nx; ++i) {
      const double x = (i + 0.5) * dx, r = (j + 0.5) * dx;
      double l = CS;
      if (x <= contour.xe && r < RinCol[i]) {
        const double s = x > contour.xt ? x - contour.xt : x;
        const double delta = 0.025 * s + (x > contour.xt ? 0.05 * contour.Rt : 0);
        l = std::max(CS, std::min(KAPPA * (RinCol[i] - r), 0.09 * delta) / dx);
        if (x > contour.xt + 0.5 * contour.Rt && RinCol[i] - r < std::max(2 * dx, delta))
          geo.geoM[((size_t)j * Nx + i) * 4 + 2] += 16;
      }
      geo.geoM[((size_t)j * Nx + i) * 4 + 3] = cfg.turbulence ? f32(l) : 0.f;
    }
  S->cfg = cfg;
  S->Nx = Nx;
  S->Nr = Nr;
  S->dx = dx;
  S->tw = tw;
  S->bodyR = bodyR;
  S->xBase = xBase;
  S->contour = contour;
  S->atm = atm;
  S->RinCol = RinCol;
  S->solids = hw;
  S->tScale = dx / U_REF;
  S->P = computeParams(cfg, contour, dx, cfg.Pc, atm.p);
  S->injOn = injOn;
  // probes (simsetup.js:104-118)
  S->iExit = std::max(1, std::min(Nx - 1, (int)std::floor(contour.xe / dx) - 1));
  S->jExit = 0;
  while (S->jExit < Nr && !mask[(size_t)S->jExit * Nx + S->iExit]) S->jExit++;
  S->iChamber = std::max(1, (int)std::floor(0.5 * contour.Lc / dx));
  for (const Solid& s : hw)
    if (s.r0 <= 0) S->iChamber = std::max(S->iChamber, (int)std::ceil(s.x1 / dx) + 2);
  S->iThroat = (int)std::floor(contour.xt / dx);
  S->jWall.assign((size_t)Nx * 4, 0.f);
  for (int i = 0; i <= S->iExit; ++i) {
    int j = std::min(Nr - 1, (int)std::floor(RinCol[i] / dx));
    while (j > 0 && (mask[(size_t)j * Nx + i] || geo.kappa[(size_t)j * Nx + i] < 0.5)) j--;
    S->jWall[(size_t)i * 4] = (float)j;
  }
  if (injOn) S->injTab = buildInjectorTable(cfg, contour, dx, S->P);
  (void)Ge;
  (void)Ga;
  return S;
}

// ---- initialState (simsetup.js:72-102)
void initialState(const Setup& S, float* A, float* B) {
  const int Nx = S.Nx, Nr = S.Nr;
  const double dx = S.dx;
  const ParamsD& P = S.P;
  const double g = S.cfg.gamma, Ge = 1 / (g - 1), Ga = 1 / (GAMMA_AIR - 1);
  const double pa = P.amb[3], rhoa = P.amb[0];
  const double ReN = P.ReN, TcN = P.res[1], PcN = P.res[0];
  const double rho0 = PcN / (ReN * TcN);
  const double E0 = PcN * Ge;
  const Contour& c = S.contour;
  std::memset(A, 0, sizeof(float) * (size_t)Nx * Nr * 4);
  std::memset(B, 0, sizeof(float) * (size_t)Nx * Nr * 4);
  for (int j = 0; j < Nr; ++j) {
    const double r = (j + 0.5) * dx;
    for (int i = 0; i < Nx; ++i) {
      const double x = (i + 0.5) * dx;
      const size_t k = ((size_t)j * Nx + i) * 4;
      const bool inNozzle = x <= c.xe && r < S.RinCol[i];
      if (S.injOn && inNozzle && x < c.xt) {
        A[k] = f32(rho0); A[k + 1] = 0; A[k + 2] = 0; A[k + 3] = f32(E0);
        B[k] = f32(rho0); B[k + 1] = f32(rho0 * P.injData.Fres); B[k + 2] = f32(Ge);
      } else if (S.cfg.startMode == NZ_START_BURST && inNozzle && x < c.xt) {
        A[k] = f32(rho0); A[k + 1] = 0; A[k + 2] = 0; A[k + 3] = f32(E0);
        B[k] = f32(rho0); B[k + 1] = f32(rho0); B[k + 2] = f32(Ge);
      } else if (inNozzle) {
        A[k] = f32(rhoa); A[k + 1] = 0; A[k + 2] = 0; A[k + 3] = f32(pa * Ga); B[k + 2] = f32(Ga);
      } else {
        A[k] = f32(P.ambCons[0]); A[k + 1] = f32(P.ambCons[1]); A[k + 2] = 0; A[k + 3] = f32(P.ambCons[3]);
        B[k + 2] = f32(Ga);
      }
    }
  }
  if (!S.injOn)
    for (int j = 0; j < Nr; ++j) B[(size_t)j * Nx * 4 + 3] = f32(P.uInlet0);
  const std::vector<float>& gm = S.geo.geoM;
  for (size_t k = 0; k < (size_t)Nx * Nr; ++k) {
    const size_t q = k * 4;
    const float hx = gm[q], hy = gm[q + 1];
    if (S.geo.mask[k] || (hx == 0 && hy == 0)) continue;
    const size_t h = (size_t)((int64_t)k + (int64_t)hy * Nx + (int64_t)hx) * 4;
    for (int cc = 0; cc < 4; ++cc) { A[q + cc] = A[h + cc]; B[q + cc] = B[h + cc]; }
  }
}

}  // namespace nz

// ======================================================================
// C ABI
// ======================================================================
using namespace nz;

extern "C" {

void nz_config_default(nz_config* c) {
  std::memset(c, 0, sizeof(*c));
  // S defaults (app.js:12-26) passed through cfgFromS (app.js:201-211)
  c->Dt = 23 / 1000.0;
  c->De = std::max(119.0, 23 * 1.02) / 1000;
  c->type = NZ_BELL;
  c->bellPct = 80;
  c->coneAngle = 15;
  c->CR = 6;
  c->chamberLen = 1.6;
  c->Pc = 196 * 1e5 * 100 / 100;
  c->Tc = 3670;
  c->gamma = 1.22;
  c->Mw = 22;
  c->alt = 0;
  c->mach = 0;
  c->base = 0;
  c->bodyRatio = 1.5;
  c->domainL = 1.2;
  c->Nx = 1024;
  c->aspect = 0.28;
  c->startMode = NZ_START_RAMP;
  c->rampMs = 0.12;
  c->cfl = 0.6;
  c->order2 = 1;
  c->wallModel = NZ_WALL_CUT;
  c->minCells = 8;
  c->viscous = 1;
  c->afterburn = 1;
  c->qAb = 6.8;
  c->sAir = 1.8;
  c->Tw = 800;
  c->inletAbsorb = 100 / 100.0;
  c->prop = NZ_PROP_KEROLOX;
  c->injMode = 0;
  c->injType = NZ_INJ_SHOWER;
  c->injRings = 6;
  c->injDP = 20;
  c->injD = 1.5;
  c->injTMR = 1;
  c->injAngle = 30;
  c->OF = 2.3;
  c->film = 0;
  c->baffles = 0;
  c->baffleH = 30;
  c->cavD = 0;
  c->cavW = 40;
  c->dx = 0;
  c->hllBlend = 1.0;
  c->convAngle = 35;
  c->ambP = 0;
  c->ambT = 0;
  c->turbulence = 1;
}

int nz_config_preset(nz_config* c, const char* name) {
  // ENGINE_PRESETS (physics.js:233-309) applied as applyPreset does (app.js:1108-1121)
  struct Pre { const char* key; int prop; double Pc, Dt, De; int type; double bellPct, coneAngle, CR, chamberLen, alt, mach; int base; double domain;
               int inj; int injType; double injRings, injDP, injD, injAngle, injTMR, film, baffles, baffleH, cavD, cavW; };
  static const Pre pre[] = {
      {"injShower", NZ_PROP_KEROLOX, 70, 30, 85, NZ_BELL, 80, 15, 5, 6, 0, 0, 0, 0.5, 1, NZ_INJ_SHOWER, 6, 20, 1.5, 0, 0, 0, 0, 30, 0, 40},
      {"injCoax", NZ_PROP_METHALOX, 70, 30, 85, NZ_BELL, 80, 15, 5, 6, 0, 0, 0, 0.5, 1, NZ_INJ_COAX, 4, 15, 2, 0, 0, 0, 0, 30, 0, 40},
      {"injImpinge", NZ_PROP_KEROLOX, 70, 30, 85, NZ_BELL, 80, 15, 5, 6, 0, 0, 0, 0.5, 1, NZ_INJ_IMPINGE, 4, 20, 1.2, 30, 0, 0, 0, 30, 0, 40},
      {"injPintle", NZ_PROP_KEROLOX, 70, 30, 85, NZ_BELL, 80, 15, 5, 6, 0, 0, 0, 0.5, 1, NZ_INJ_PINTLE, 1, 20, 1, 0, 0.4, 0, 0, 30, 0, 40},
      {"injHardware", NZ_PROP_KEROLOX, 70, 30, 85, NZ_BELL, 80, 15, 5, 6, 0, 0, 0, 0.5, 1, NZ_INJ_IMPINGE, 4, 20, 1.2, 30, 0, 10, 1, 30, 12, 40},
      {"demo", NZ_PROP_KEROLOX, 196, 23, 119, NZ_BELL, 80, 15, 6, 1.6, 0, 0, 0, 1.2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
      {"overexp", NZ_PROP_HYDROLOX, 120, 20, 160, NZ_BELL, 80, 15, 5, 1.4, 0, 0, 0, 1.4, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
      {"diamonds", NZ_PROP_METHALOX, 90, 30, 125, NZ_BELL, 80, 15, 5, 1.5, 0, 0, 0, 1.2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
      {"ideal", NZ_PROP_KEROLOX, 70, 34, 99, NZ_BELL, 80, 15, 5, 1.5, 0, 0, 0, 1.2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
      {"underexp", NZ_PROP_KEROLOX, 80, 30, 70, NZ_BELL, 80, 15, 5, 1.5, 22, 0, 0, 1.4, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
      {"coldgas", NZ_PROP_COLDGAS, 20, 20, 30, NZ_CONE, 80, 15, 4, 1.4, 0, 0, 0, 0.6, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
      {"hybrid", NZ_PROP_HYBRID, 30, 30, 64, NZ_CONE, 80, 12, 4, 2.0, 0, 0, 0, 1.0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
      {"upper", NZ_PROP_HYDROLOX, 100, 16, 100, NZ_BELL, 80, 15, 5, 1.5, 30, 0, 0, 1.4, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
      {"flight", NZ_PROP_KEROLOX, 100, 26, 90, NZ_BELL, 80, 15, 5, 1.5, 12, 2.0, 1, 1.2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
  };
  for (const Pre& p : pre) {
    if (std::strcmp(p.key, name) != 0) continue;
    const Prop& pr = PROPS[p.prop];
    c->prop = p.prop;
    c->Tc = pr.Tc; c->gamma = pr.gamma; c->Mw = pr.Mw; c->qAb = pr.qAb; c->sAir = pr.sAir;
    c->Pc = p.Pc * 1e5;
    c->Dt = p.Dt / 1000;
    c->De = std::max(p.De, p.Dt * 1.02) / 1000;
    c->type = p.type; c->bellPct = p.bellPct; c->coneAngle = p.coneAngle; c->CR = p.CR; c->chamberLen = p.chamberLen;
    c->alt = p.alt * 1000; c->mach = p.mach; c->base = p.base; c->domainL = p.domain;
    c->injMode = 0;
    if (pr.spray) c->OF = pr.OF;
    if (p.inj) {
      c->injMode = 1; c->injType = p.injType; c->injRings = (int)p.injRings; c->injDP = p.injDP; c->injD = p.injD;
      if (p.injAngle != 0) c->injAngle = p.injAngle;
      if (p.injTMR != 0) c->injTMR = p.injTMR;
      c->film = p.film; c->baffles = p.baffles; c->baffleH = p.baffleH; c->cavD = p.cavD; c->cavW = p.cavW;
    }
    return 0;
  }
  return nz_fail(-2, std::string("unknown preset: ") + name);
}

int nz_setup_build(const nz_config* cfg, nz_setup** out) {
  if (!cfg || !out) return nz_fail(-1, "nz_setup_build: null argument");
  if (cfg->Nx < 4 || cfg->Nx > 65536) return nz_fail(-1, "nz_setup_build: Nx out of range");
  if (!(cfg->Dt > 0) || !(cfg->gamma > 1) || !(cfg->Mw > 0) || !(cfg->domainL > 0 || cfg->dx > 0))
    return nz_fail(-1, "nz_setup_build: invalid configuration");
  try {
    *out = reinterpret_cast<nz_setup*>(buildSimSetup(*cfg));
    if (!*out) return -10;   // the device geometry build failed; nz_last_error() says why
  } catch (const std::exception& e) {
    return nz_fail(-3, std::string("nz_setup_build: ") + e.what());
  }
  return 0;
}

void nz_setup_free(nz_setup* s) { delete reinterpret_cast<Setup*>(s); }

int nz_setup_get_info(const nz_setup* sp, nz_setup_info* o) {
  if (!sp || !o) return nz_fail(-1, "nz_setup_get_info: null argument");
  const Setup& S = *reinterpret_cast<const Setup*>(sp);
  std::memset(o, 0, sizeof(*o));
  o->Nx = S.Nx; o->Nr = S.Nr; o->dx = S.dx; o->tScale = S.tScale; o->tw = S.tw; o->bodyR = S.bodyR; o->xBase = S.xBase;
  o->iExit = S.iExit; o->jExit = S.jExit; o->iChamber = S.iChamber; o->iThroat = S.iThroat;
  o->injector = S.injOn ? 1 : 0;
  o->nCut = S.geo.nCut; o->nMerged = S.geo.nMerged;
  const Contour& c = S.contour;
  o->Rt = c.Rt; o->Re = c.Re; o->Rc = c.Rc; o->Lc = c.Lc; o->xt = c.xt; o->xe = c.xe; o->eps = c.eps; o->thn = c.thn; o->the = c.the;
  o->atmT = S.atm.T; o->atmP = S.atm.p;
  o->thMe = S.theory.Me; o->thPe = S.theory.pe; o->thMdot = S.theory.mdot; o->thF = S.theory.F; o->thIsp = S.theory.Isp;
  o->thCstar = S.theory.cstar; o->thC0 = S.theory.c0; o->thUlim = S.theory.uLim;
  return 0;
}

int nz_setup_get_params(const nz_setup* sp, nz_params* P) {
  if (!sp || !P) return nz_fail(-1, "nz_setup_get_params: null argument");
  paramsToFloat(reinterpret_cast<const Setup*>(sp)->P, *P);
  return 0;
}

int nz_setup_get_params_f64(const nz_setup* sp, double* out, int n) {
  if (!sp || !out) return nz_fail(-1, "nz_setup_get_params_f64: null argument");
  return paramsToDoubles(reinterpret_cast<const Setup*>(sp)->P, out, n);
}

int nz_setup_compute_params(const nz_setup* sp, const nz_config* cfg, double PcSi, double pStartSi, nz_params* P) {
  if (!sp || !cfg || !P) return nz_fail(-1, "nz_setup_compute_params: null argument");
  const Setup& S = *reinterpret_cast<const Setup*>(sp);
  paramsToFloat(computeParams(*cfg, S.contour, S.dx, PcSi, pStartSi), *P);
  return 0;
}

int nz_setup_get_geometry(const nz_setup* sp, uint8_t* mask, float* geoA, float* geoG, float* geoM, float* jWall,
                          float* kappa, float* Ax, float* Ar) {
  if (!sp) return nz_fail(-1, "nz_setup_get_geometry: null setup");
  const Setup& S = *reinterpret_cast<const Setup*>(sp);
  const CutGeometry& g = S.geo;
  if (mask) std::memcpy(mask, g.mask.data(), g.mask.size());
  if (geoA) std::memcpy(geoA, g.geoA.data(), g.geoA.size() * 4);
  if (geoG) std::memcpy(geoG, g.geoG.data(), g.geoG.size() * 4);
  if (geoM) std::memcpy(geoM, g.geoM.data(), g.geoM.size() * 4);
  if (jWall) std::memcpy(jWall, S.jWall.data(), S.jWall.size() * 4);
  if (kappa) std::memcpy(kappa, g.kappa.data(), g.kappa.size() * 4);
  if (Ax) std::memcpy(Ax, g.Ax.data(), g.Ax.size() * 4);
  if (Ar) std::memcpy(Ar, g.Ar.data(), g.Ar.size() * 4);
  return 0;
}

int nz_setup_initial_state(const nz_setup* sp, float* A, float* B) {
  if (!sp || !A || !B) return nz_fail(-1, "nz_setup_initial_state: null argument");
  initialState(*reinterpret_cast<const Setup*>(sp), A, B);
  return 0;
}

int nz_setup_get_injector(const nz_setup* sp, float* tab) {
  if (!sp) return nz_fail(-1, "nz_setup_get_injector: null setup");
  const Setup& S = *reinterpret_cast<const Setup*>(sp);
  if (!S.injOn) return -1;
  if (tab) std::memcpy(tab, S.injTab.data(), S.injTab.size() * 4);
  return 0;
}

int nz_setup_get_cutcell_input(const nz_setup* sp, nz_cutcell_in* out) {
  if (!sp || !out) return nz_fail(-1, "nz_setup_get_cutcell_input: null argument");
  nz::cutcellInput(*reinterpret_cast<const Setup*>(sp), *out);
  return 0;
}

int nz_setup_get_probes(const nz_setup* sp, nz_probes* out) {
  if (!sp || !out) return nz_fail(-1, "nz_setup_get_probes: null argument");
  const Setup& S = *reinterpret_cast<const Setup*>(sp);
  std::memset(out, 0, sizeof(*out));
  out->iExit = S.iExit;
  out->jExit = S.jExit;
  out->iThroat = S.iThroat;
  out->iChamber = S.iChamber;
  out->dx = S.dx;
  out->exact = 1;
  out->pa = S.P.amb[3];
  for (int k = 0; k < 3; ++k) out->Rs[k] = S.P.Rs[k];
  out->RaN = 1.0;   // computeParams' RaN (simsetup.js:309)
  out->sprayMdot = S.P.inj ? S.P.sprayMdot : 0.0;
  return 0;
}

int nz_wall_separation(const nz_setup* sp, const float* row, double paSi, nz_separation* out) {
  if (!sp || !row || !out) return nz_fail(-1, "nz_wall_separation: null argument");
  const Setup& S = *reinterpret_cast<const Setup*>(sp);
  const Contour& C = S.contour;
  const double dx = S.dx;
  // walk upstream from the exit while the near-wall flow runs backwards or slows below
  // 0.6 c; up to two attached cells are bridged (simsetup.js:352-360)
  int i = S.iExit, gaps = 0, last = -1;
  for (; i > S.iThroat + 2 && gaps <= 2; --i) {
    const size_t k = (size_t)i * 4;
    const double x = (i + 0.5) * dx;
    const double slope = (C.radiusAt(x + 0.5 * dx) - C.radiusAt(x - 0.5 * dx)) / dx;
    const double ut = (row[k] + slope * row[k + 1]) / std::hypot(1.0, slope);
    if (ut < 0 || std::hypot((double)row[k], (double)row[k + 1]) < 0.6 * row[k + 3]) { last = i; gaps = 0; }
    else if (last >= 0 || S.iExit - i > 2) gaps++;
  }
  if (last < 0 || S.iExit - last < 3) return 0;
  // wall pressure ahead of the separation shock: the lowest in the 30 cells upstream
  double pSep = INFINITY;
  size_t kS = (size_t)last * 4;
  for (int q = last; q >= std::max(S.iThroat + 1, last - 30); --q)
    if (row[(size_t)q * 4 + 2] < pSep) { pSep = row[(size_t)q * 4 + 2]; kS = (size_t)q * 4; }
  const double Msep = std::hypot((double)row[kS], (double)row[kS + 1]) / std::max((double)row[kS + 3], 1e-6);
  if (Msep < 1.05) return 0;   // nothing supersonic ahead of it (the nozzle has not started)
  const double x = (last + 0.5) * dx, R = C.radiusAt(x);
  out->x = x;
  out->frac = (x - C.xt) / (C.xe - C.xt);
  out->eps = (R * R) / (C.Rt * C.Rt);
  out->pRatio = pSep * P_REF / std::max(paSi, 1e-6);
  out->Msep = Msep;
  return 1;
}

int nz_summerfield_point(const nz_setup* sp, double PcSi, double gamma, double paSi, double* frac) {
  if (!sp || !frac) return nz_fail(-1, "nz_summerfield_point: null argument");
  const Contour& C = reinterpret_cast<const Setup*>(sp)->contour;
  // first station of the divergent section where 1-D isentropic wall pressure drops
  // below 0.4 of ambient (simsetup.js:375-384)
  const int nS = 200;
  for (int k = 0; k <= nS; ++k) {
    const double x = C.xt + (C.xe - C.xt) * k / nS, R = C.radiusAt(x);
    const double M = machFromAreaRatio(std::max((R * R) / (C.Rt * C.Rt), 1.0001), gamma, true);
    if (PcSi * std::pow(1 + 0.5 * (gamma - 1) * M * M, -gamma / (gamma - 1)) < 0.4 * paSi) {
      *frac = (double)k / nS;
      return 1;
    }
  }
  return 0;
}

double nz_setup_radius_at(const nz_setup* sp, double x) {
  if (!sp) return 0;
  return reinterpret_cast<const Setup*>(sp)->contour.radiusAt(x);
}

int nz_nozzle_theory(double Pc, double Tc, double gamma, double Mw, double Dt, double De, double pa, double* o) {
  if (!o) return nz_fail(-1, "nz_nozzle_theory: null output");
  const Theory t = nozzleTheory(Pc, Tc, gamma, Mw, Dt, De, pa);
  const double v[14] = {t.At, t.Ae, t.eps, t.Me, t.pe, t.Te, t.ue, t.cstar, t.mdot, t.F, t.Fvac, t.Isp, t.c0, t.uLim};
  std::memcpy(o, v, sizeof(v));
  return 0;
}

void nz_atmosphere(double h, double* o) {
  const Atm a = atmosphere(h);
  o[0] = a.T; o[1] = a.p; o[2] = a.rho; o[3] = a.a;
}

}  // extern "C"
