// nz_internal.h — host-side types shared by the setup code and the solver runtime.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "nozzle.h"

// records the calling thread's last error and returns `code`
int nz_fail(int code, const std::string& msg);

namespace nz {

extern const double G0, RU, R_AIR, P_REF, T_REF, R_REF, RHO_REF, U_REF, GAMMA_AIR, CHEM_TAU0;

struct Atm { double T = 0, p = 0, rho = 0, a = 0; };
struct Theory {
  double At = 0, Ae = 0, eps = 0, Me = 0, pe = 0, Te = 0, ue = 0, cstar = 0, mdot = 0, F = 0, Fvac = 0, Isp = 0, c0 = 0, uLim = 0;
};
struct Contour {
  std::vector<double> xs, rs;
  double Rt = 0, Re = 0, Rc = 0, Lc = 0, xt = 0, xe = 0, eps = 0, thn = 0, the = 0;
  double cavW = 0, cavD = 0;  // acoustic-cavity pocket (m), injector mode only
  double radiusAt(double x) const;
};
struct Solid { double x0, x1, r0, r1; };
struct InjData {
  double Rf = 0, Ro = 0, Rp = 0, eF = 0, eO = 0, Fres = 0, qHat = 0, s = 0, Ge = 0;
  double fuelVec[4] = {0, 0, 0, 0}, oxVec[4] = {0, 0, 0, 0};
  double fuelRho = 0, oxRho = 0;
};
// the reference's P object at double precision
struct ParamsD {
  bool inj = false;
  InjData injData;
  double Rs[3] = {0, 0, 0}, chamberNu[2] = {0, 0}, cap[2] = {0, 0};
  double sprayMhat = 0, sprayMdot = 0, sprayTScale = 0;
  double gamma = 0;
  double amb[4] = {0, 0, 0, 0}, ambCons[4] = {0, 0, 0, 0}, ambB[4] = {0, 0, 0, 0};
  double gamR[2] = {0, 0}, inlet[4] = {0, 0, 0, 0};
  double uInlet0 = 0;
  bool visc = false, chem = false;
  double gas[4] = {0, 0, 0, 0}, chemK[4] = {0, 0, 0, 0};
  double Tw = 0, res[3] = {0, 0, 0}, pStart = 0, ramp = 0, RcCells = 0;
  double cfl = 0, dtMax = 0, wFloor = 0, vMax = 0, hllBlend = 1;
  bool order2 = true;
  double ReN = 0, Ta = 0, pa = 0, uinf = 0;
};
struct CutInput {
  int Nx = 0, Nr = 0;
  double dx = 0;
  const Contour* contour = nullptr;
  double tw = 0, xBase = 0, bodyR = 0;
  bool base = false, stair = false;
  std::vector<Solid> solids;  // metres
};
struct CutGeometry {
  int Nx = 0, Nr = 0;
  std::vector<uint8_t> mask;
  std::vector<float> geoA, geoG, geoM, kappa, Ax, Ar;
  int nCut = 0, nMerged = 0;
};
struct Setup {
  nz_config cfg;
  std::vector<float> Rtab;         // the cut-cell builder's contour table (nz_cutcell_in)
  std::vector<double> solidsFlat;  // its solids, 4 per solid (m)
  int Nx = 0, Nr = 0;
  double dx = 0, tw = 0, bodyR = 0, xBase = 0, tScale = 0;
  Contour contour;
  Atm atm;
  Theory theory;
  std::vector<float> RinCol;
  std::vector<Solid> solids;
  CutGeometry geo;
  ParamsD P;
  bool injOn = false;
  int iExit = 0, jExit = 0, iChamber = 0, iThroat = 0;
  std::vector<float> jWall;
  std::vector<float> injTab;
};

Atm atmosphere(double h);
Atm ambientOf(const nz_config& cfg);
double areaRatioFromMach(double M, double g);
double machFromAreaRatio(double eps, double g, bool supersonic);
Theory nozzleTheory(double Pc, double Tc, double g, double Mw, double Dt, double De, double pa);
Contour buildContour(const nz_config& o);
void buildCutGeometry(const CutInput& in, CutGeometry& out);
// the inner-wall radius (cells) every 1/8 cell, -1e9 past the exit (cutcell.js:19-27)
std::vector<float> contourTable(const Contour& c, int Nx, double dx);
ParamsD computeParams(const nz_config& cfg, const Contour& contour, double dx, double PcSi, double pStartSi);
void paramsToFloat(const ParamsD& d, nz_params& P);
int paramsToDoubles(const ParamsD& d, double* out, int n);
Setup* buildSimSetup(const nz_config& cfg);
void cutcellInput(const Setup& S, nz_cutcell_in& in);
void initialState(const Setup& S, float* A, float* B);

}  // namespace nz
