// nz_device.h — device-side data layout and per-cell / per-face physics of the
// MI355X nozzle solver. Written for gfx950 (wave64); included by solver.hip only.
//
// Layout in HBM (structure of arrays, one instance):
//   U[8][Ns]      conserved state planes rho, rho*u, rho*v, E, rho*Z, rho*F, G, w
//                 (w = inflow-speed average ubar, or rho*O in injector mode), row-major j*Nx+i
//   ax[Nr][Nx+1]  signed r-weighted x-face areas (geoA.z; negative = first-order face)
//   ar[Nr+1][Nx]  signed r-weighted r-face areas (geoA.w)
//   gg[Nx*Nr]     float4 group (V, Am, Wx, Wr) (geoG)
//   code[Nx*Nr]   packed per-cell flags: solid, host offset, slave bits, boundary-layer band
//   lt[Nx*Nr]     turbulence length scale in cells (geoM.w)
// The physics restates the reference GLSL (solver.js / spray.js line numbers
// at each function); the fusion, tiling and data layout are this library's own.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace nzk {

// ---- per-cell code word
enum : uint32_t {
  C_SOLID = 1u,
  C_SLAVE_SHIFT = 1,   // bits 1..4: slave bits 1,2,4,8 of a host (cutcell.js:168-174)
  C_BL = 1u << 5,      // boundary-layer band flag (geoM.z bit 16, simsetup.js:60)
  C_IS_SLAVE = 1u << 6,
  C_HX_SHIFT = 8,      // host offset dx + 1 (2 bits)
  C_HY_SHIFT = 10,     // host offset dy + 1 (2 bits)
  C_EXS_SHIFT = 12,    // bits 12..19 of a host: 1 + slot of its first slave outside the tile (0: none listed)
};
// per tile of the stage kernels: faces of merged cells outside the tile (3 per such
// slave, listed host by host) and the cells their stencils reach beyond the 2-cell halo
constexpr int MAX_EXTRA_FACES = 192;
constexpr int MAX_FAR_CELLS = 64;
constexpr int FAR_HALO = 4;   // far cells are looked up in a window this many cells around the tile
__device__ __forceinline__ int code_hdx(uint32_t c) { return (int)((c >> C_HX_SHIFT) & 3u) - 1; }
__device__ __forceinline__ int code_hdy(uint32_t c) { return (int)((c >> C_HY_SHIFT) & 3u) - 1; }
__device__ __forceinline__ uint32_t code_slaves(uint32_t c) { return (c >> C_SLAVE_SHIFT) & 15u; }
__device__ __forceinline__ int code_exslot(uint32_t c) { return (int)((c >> C_EXS_SHIFT) & 255u) - 1; }

// ---- parameters of one instance (the reference's P block plus derived uniforms)
struct Consts {
  float amb[4], res[3], inlet[4], pStart, ramp, RcCells;
  int inj;
  int order2;
  float hllBlend;
  float ambCons[4], ambB[4], gamR[2];
  float uBar[3];   // (inlet[0], 1.6*uInlet0, max(ramp,0)) as solver.js:935 passes them
  float vMax;
  float cfl, dtMax, wFloor;
  int visc, chem;
  float gas[4], Rs[3], chamberNu[2], cap[2], chemK[4], Tw;
  float sprayMhat, sprayTScale, fuelVec[4], oxVec[4], eF, eO, Ge;
  float fan;       // impingement fan half-angle (spray.js:271)
  float resScale;  // residual normalisation 1/rho_chamber
  int spray;       // spray stepped this step
  int NP;          // parcels
  // fixed-point spray deposit: a release v is summed as the integer rint(v * 2^depExp / B),
  // B = the largest spawned parcel mass (mass channels) or twice that times depVel, a bound
  // on parcel and gas speeds (momentum channels); depExp leaves NP such terms headroom in
  // 63 bits, so integer sums, whose order does not matter, carry every release
  float depVel;
  int depExp;
  // batched solves: an instance stops (flags bit 0) at t >= tEnd (grid time), after
  // minSteps steps at residual <= resTol, or after maxSteps steps; zero disables a test
  float tEnd, resTol;
  int minSteps, maxSteps;
};

// per-instance mutable scalars kept on the device between launches
struct InstState {
  double cur[4];   // step info of the step in progress (dt, t_old, t_new, w), in the state's precision
  double last[4];  // step info of the last completed step (readStepInfo)
  unsigned long long wacc;   // max wave speed being accumulated (double bits, >= 0)
  unsigned long long wsnap;  // max wave speed of the current state
  uint32_t head;   // spray ring head
  uint32_t seed;   // spray seed
  long long steps; // completed steps since setState
  double resAcc;   // sum of squares of d(rho)/dt over groups, this step
  double resLast;  // RMS of the last completed step
  double nGroups;  // number of fluid groups (residual normalisation)
  uint32_t flags;  // bit 0: instance inactive (skipped); stopped on bit 1: the residual, 2: time, 3: step count
  uint32_t pad0_;
  uint32_t flagsSeen;  // `flags` as the last stage 2 (or close kernel) read them: stage 1, whose first
                       // workgroup writes `flags`, reads this copy instead
  uint32_t pad_;
  double mmax;     // largest parcel mass spawned since the state was set (stage 1 raises it)
};

// The next step's stage 1 folds the tile partials of the last stage 2 into the time step
// itself (every workgroup reads them) when there are at most this many; with more, a close
// kernel folds them after every step. The partials array always holds the current tiling's
// count of valid entries (tilesX * tilesY): stage 2 writes one per tile, and the CFL set-up
// and a tile-shape change leave it folded into entry 0 with zeros after it, which every
// tiling reads alike (a zero changes neither the max nor the sum).
constexpr uint32_t FUSED_CLOSE_MAX = 2048;

// 64-byte aligned, so a descriptor spans exactly sizeof(DevInst) / 64 scalar-cache lines
struct alignas(64) DevInst {
  int Nx, Nr;
  int tilesX, tilesY, tile0;  // tile grid of this instance inside a launch
  int pblk0, npblk;           // parcel blocks of this instance inside a spray launch
  int cblk0, ncblk;           // cell blocks of this instance inside a reduction launch
  long long Ns;               // plane stride in elements
  void* U0;                   // state U^n (updated in place by stage 2)
  void* U1;                   // state U^(1)
  const float* ax;
  const float* ar;
  const float4* gg;
  const uint32_t* code;
  const float* lt;
  const int* exStart;         // per tile (instance-local): first entry in exFaces; tiles + 1 entries
  const uint2* exFaces;       // faces of merged cells outside their host's tile (x | dir << 31, y)
  const int* fcStart;         // per tile: first entry in fcCells; tiles + 1 entries
  const int2* fcCells;        // cells those faces reach beyond the tile's 2-cell halo
  void* exFlux;               // per extra face: 9 inviscid + 7 viscous weighted fluxes (state type)
  unsigned long long* src;    // 4 planes of Ns, fixed point (Consts::depExp): fuel vapour, ox vapour,
                              // x momentum, r momentum
  float* parc;                // 8 planes of NP: x, r, u, v, d, m, type, xi
  const float* tab;           // spray element table 256 x 3 RGBA
  InstState* st;
  double2* part;              // per tile: (max wave speed, residual sum) of the last stage-2 launch
  Consts c;
};

// A workgroup's first touch of its instance descriptor: one scalar load per 64-byte line,
// all issued together and waited for once, so that the field loads that follow (which
// the compiler places next to their uses, each with its own wait) hit the scalar cache
// instead of each paying a cold round trip to memory at kernel start.
__device__ __forceinline__ void touch_desc(const DevInst* __restrict__ d) {
  const unsigned* w = reinterpret_cast<const unsigned*>(d);
  unsigned v[sizeof(DevInst) / 64];
#pragma unroll
  for (int l = 0; l < (int)(sizeof(DevInst) / 64); ++l) v[l] = __builtin_nontemporal_load(w + 16 * l);
#pragma unroll
  for (int l = 0; l < (int)(sizeof(DevInst) / 64); ++l) asm volatile("" ::"s"(v[l]));
}

// Global-memory access: a uniform 64-bit base (SGPRs) plus a 32-bit per-lane byte offset
// in the global address space, so the compiler emits global loads / stores with an SGPR
// base instead of flat accesses through generic pointers with 64-bit lane addresses.
// Every array indexed this way holds fewer than 2^28 elements per base.
#define NZ_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T gld(const T* base, unsigned i) {
  unsigned off = i * (unsigned)sizeof(T);
  return *(const NZ_GLOBAL T*)((const NZ_GLOBAL char*)base + off);
}
// A workgroup-uniform load from memory that no kernel writes while it runs (host-built
// tables): through the constant address space, so it is a scalar load waited on by its
// own counter rather than a vector load plus a lane read.
template <class T>
__device__ __forceinline__ T cld(const T* base, unsigned i) {
  return *((const __attribute__((address_space(4))) T*)base + i);
}
__device__ __forceinline__ float4 gld(const float4* base, unsigned i) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  unsigned off = i * 16u;
  const v4f v = *(const NZ_GLOBAL v4f*)((const NZ_GLOBAL char*)base + off);
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ double2 gld(const double2* base, unsigned i) {
  typedef double v2d __attribute__((ext_vector_type(2)));
  unsigned off = i * 16u;
  const v2d v = *(const NZ_GLOBAL v2d*)((const NZ_GLOBAL char*)base + off);
  return make_double2(v.x, v.y);
}
template <class T>
__device__ __forceinline__ void gst(T* base, unsigned i, T v) {
  unsigned off = i * (unsigned)sizeof(T);
  *(NZ_GLOBAL T*)((NZ_GLOBAL char*)base + off) = v;
}

// probes and parameter values of the integrated quantities (exitPerformance,
// chamberPerformance, cellPrimSI), in double as the reference's P object holds them
struct IntegCtx {
  int iExit, jExit, iThroat, iChamber, inj;
  double pa, Rs[3], RaN;
};
// the integrals' cell pass uses at most this many blocks (one double2 partial each)
constexpr int INTEG_BLOCKS_MAX = 1024;

// view parameters of the FS_DERIVED display fields (render.js:420-435)
struct DerivedC {
  int field;
  float r0, r1, RaN, Rs[3], chemK[4], Tc, Ta, schl, dx, vel;  // vel: U_REF as render.js passes it
};

template <typename R>
struct Prim {
  R r, u, v, p, Z, F, G, O;
};

// Division and square root in the state's precision. FP32 uses the hardware
// reciprocal and square root (v_rcp_f32 / v_sqrt_f32, about 1 ulp; GLSL ES 3.00
// highp division, which the reference runs on, guarantees only 2.5 ulp); FP64 keeps
// IEEE division and square root so that the FP64 run is the error reference.
__device__ __forceinline__ float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double dv(double a, double b) { return a / b; }
__device__ __forceinline__ float rc(float b) { return __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double rc(double b) { return 1.0 / b; }
__device__ __forceinline__ float sq(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double sq(double x) { return sqrt(x); }

// Power and exponential in the state's precision. FP32 uses the hardware base-2
// exponential and logarithm (v_exp_f32 / v_log_f32), which is how GLSL ES 3.00 defines
// pow(x, y) = exp2(y log2 x) (x > 0; x = 0 gives 0 for y > 0); FP64 keeps the library.
__device__ __forceinline__ float powR(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
__device__ __forceinline__ double powR(double x, double y) { return pow(x, y); }
__device__ __forceinline__ float expR(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ double expR(double x) { return exp(x); }
// natural logarithm, sine and cosine likewise (GLSL's log, sin, cos are implementation
// precision; the hardware v_log_f32 / v_sin_f32 / v_cos_f32 serve FP32)
__device__ __forceinline__ float logR(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
__device__ __forceinline__ double logR(double x) { return log(x); }
__device__ __forceinline__ float sinR(float x) { return __sinf(x); }
__device__ __forceinline__ double sinR(double x) { return sin(x); }
__device__ __forceinline__ float cosR(float x) { return __cosf(x); }
__device__ __forceinline__ double cosR(double x) { return cos(x); }

template <typename R>
__device__ __forceinline__ R clampR(R x, R lo, R hi) { return fmin(fmax(x, lo), hi); }
template <typename R>
__device__ __forceinline__ R smoothstepR(R e0, R e1, R x) {
  const R t = clampR<R>(dv(x - e0, e1 - e0), R(0), R(1));
  return t * t * (R(3) - R(2) * t);
}
template <typename R>
__device__ __forceinline__ R mixR(R a, R b, R t) { return a * (R(1) - t) + b * t; }

// toPrim (solver.js:33-39)
template <typename R>
__device__ __forceinline__ Prim<R> to_prim(R a0, R a1, R a2, R a3, R b0, R b1, R b2, R b3) {
  Prim<R> q;
  q.r = fmax(a0, R(1e-7));
  const R ir = rc(q.r);
  q.u = a1 * ir;
  q.v = a2 * ir;
  q.G = fmax(b2, R(0.5));
  q.p = fmax(dv(a3 - R(0.5) * q.r * (q.u * q.u + q.v * q.v), q.G), R(1e-7));
  q.Z = clampR(b0 * ir, R(0), R(1));
  q.F = clampR(b1 * ir, R(0), R(1));
  q.O = clampR(b3 * ir, R(0), R(1));
  return q;
}
template <typename R>
__device__ __forceinline__ R sound_speed(const Prim<R>& q) { return sq(dv((R(1) + rc(q.G)) * q.p, q.r)); }

template <typename R>
__device__ __forceinline__ Prim<R> wall_prim() {
  Prim<R> q;
  q.r = R(-1); q.u = R(0); q.v = R(0); q.p = R(1);
  q.Z = R(0); q.F = R(0); q.G = R(0); q.O = R(0);
  return q;
}

template <typename R>
__device__ __forceinline__ Prim<R> load_prim(const R* __restrict__ U, long long Ns, long long k) {
  const unsigned i = (unsigned)k;
  return to_prim<R>(gld(U, i), gld(U + Ns, i), gld(U + 2 * Ns, i), gld(U + 3 * Ns, i), gld(U + 4 * Ns, i),
                    gld(U + 5 * Ns, i), gld(U + 6 * Ns, i), gld(U + 7 * Ns, i));
}

// Raw state and code word of the one grid cell that the ghost rules of cell (i, j)
// read: (i, j) clamped into the grid after the axis mirror (j < 0 -> -j - 1). Loading
// it unconditionally lets a thread issue the loads of all its cells at once.
template <typename R>
struct CellRaw {
  R v[8];
  uint32_t code;
};
template <typename R>
__device__ __forceinline__ CellRaw<R> cell_src_load(const DevInst& I, const R* __restrict__ U, int i, int j) {
  const int jm = j < 0 ? -j - 1 : j;
  const int ci = min(max(i, 0), I.Nx - 1), cj = min(max(jm, 0), I.Nr - 1);
  const unsigned k = (unsigned)(cj * I.Nx + ci);
  CellRaw<R> s;
  s.code = gld(I.code, k);
  for (int q = 0; q < 8; ++q) s.v[q] = gld(U + q * I.Ns, k);
  return s;
}

// Primitive state with boundary ghost cells at cell (i, j), i in [-2, Nx+1],
// j in [-2, Nr+1], from its source cell: stagnation inlet with ignition ramp and
// absorbing face, far field, outflow, axis mirror; solid cells flagged rho = -1
// (solver.js:70-125).
template <typename R>
__device__ Prim<R> prim_from(const DevInst& I, const CellRaw<R>& s, int i, int j, R t) {
  const Consts& c = I.c;
  const int Nx = I.Nx, Nr = I.Nr;
  const bool mir = j < 0;
  if (mir) j = -j - 1;
  Prim<R> q;
  q.r = R(c.amb[0]); q.u = R(c.amb[1]); q.v = R(c.amb[2]); q.p = R(c.amb[3]);
  q.Z = R(0); q.F = R(0); q.G = R(c.inlet[3]); q.O = R(0);
  if (i < 0) {
    const bool inCh = R(j) + R(0.5) < R(c.RcCells);
    if (j < Nr && (s.code & C_SOLID)) return wall_prim<R>();
    if (inCh && c.inj) return wall_prim<R>();
    if (inCh && j < Nr) {
      const R f = R(c.ramp) > R(0) ? smoothstepR(R(0), R(c.ramp), t) : R(1);
      const R P0 = mixR(R(c.pStart), R(c.res[0]), f);
      const R T0 = R(c.res[1]), Rg = R(c.res[2]), G = R(c.inlet[2]), g = R(1) + R(1) / G;
      const R a0 = s.v[0], a1 = s.v[1];
      const R ubar = fmax(s.v[7], R(0));
      const R c0 = sqrt(g * Rg * T0);
      const R ui = clampR(a1 / fmax(a0, R(1e-7)), R(0), R(0.5) * c0);
      const R ut = ubar * smoothstepR(R(1), R(2), P0 / fmax(R(c.pStart), R(1e-6)));
      const R u = mixR(ui, ut, R(c.inlet[1]));
      const R T = T0 - u * u * (g - R(1)) / (R(2) * g * Rg);
      const R p = P0 * powR(T / T0, g * G);
      q.r = p / (Rg * T); q.u = u; q.v = R(0); q.p = p;
      q.Z = R(1); q.F = R(1); q.G = G; q.O = R(0);
    }
  } else if (j >= Nr) {
    if (!(s.code & C_SOLID)) {
      const Prim<R> tq = to_prim<R>(s.v[0], s.v[1], s.v[2], s.v[3], s.v[4], s.v[5], s.v[6], s.v[7]);
      const R cs = sound_speed(tq);
      if (tq.v >= cs) q = tq;
      else if (tq.v > R(0)) { q = tq; q.p = R(c.amb[3]); }
    }
  } else if (i >= Nx) {
    if (!(s.code & C_SOLID)) {
      const Prim<R> tq = to_prim<R>(s.v[0], s.v[1], s.v[2], s.v[3], s.v[4], s.v[5], s.v[6], s.v[7]);
      const R cs = sound_speed(tq);
      if (tq.u >= cs) q = tq;
      else if (tq.u > R(0)) { q = tq; q.p = R(c.amb[3]); }
      else q.u = fmax(q.u, R(0));
    }
  } else {
    if (s.code & C_SOLID) return wall_prim<R>();
    q = to_prim<R>(s.v[0], s.v[1], s.v[2], s.v[3], s.v[4], s.v[5], s.v[6], s.v[7]);
  }
  if (mir) q.v = -q.v;
  return q;
}

template <typename R>
__device__ __forceinline__ Prim<R> prim_at(const DevInst& I, const R* __restrict__ U, int i, int j, R t) {
  return prim_from<R>(I, cell_src_load<R>(I, U, i, j), i, j, t);
}

// van Leer limited slope of one component (solver.js:156-160)
template <typename R>
__device__ __forceinline__ R van_leer(R a, R b) {
  R den = a + b;
  if (fabs(den) < R(1e-30)) den = R(1);
  return dv(R(2) * fmax(a * b, R(0)), den);
}
// MUSCL face value c +- half the van Leer slope of the differences a, b (solver.js:198-203):
// c + s 0.5 (2 max(ab, 0) / (a + b)) = c + s max(ab, 0) / (a + b). FP32 folds the two
// factors of 2 away and forms it as one multiply-add on the hardware reciprocal, whose
// magnitude is capped at 1e30 in place of the reference's "a + b below 1e-30 -> 1" (where
// max(ab, 0) < 2.5e-61, both add nothing to c); FP64 keeps the reference's form.
template <bool PLUS>
__device__ __forceinline__ float muscl(float c, float a, float b) {
  const float m = fmaxf(a * b, 0.f);
  const float r = __builtin_amdgcn_fmed3f(__builtin_amdgcn_rcpf(a + b), -1e30f, 1e30f);
  return fmaf(PLUS ? m : -m, r, c);
}
template <bool PLUS>
__device__ __forceinline__ double muscl(double c, double a, double b) {
  return PLUS ? c + 0.5 * van_leer(a, b) : c - 0.5 * van_leer(a, b);
}
template <typename R>
__device__ __forceinline__ R jump_ratio(R a, R b) { return dv(fabs(a - b), fmin(a, b)); }

// The MUSCL values of cell c in one direction, from its neighbours cm (behind) and cp
// (ahead), in the face-normal frame (u normal): vp the value at its face ahead (the left
// state there), vm the value at its face behind (the right state there), each formed as
// muscl forms it in face_flux, so a face built from two cells' values is bit for bit the
// face face_flux builds from the 4-cell stencil. A wall neighbour (r < 0) is the cell's
// mirror image, its normal velocity reversed (solver.js:192-197). Order r, u, v, p, Z, F, G, O.
template <typename R>
__device__ __forceinline__ void muscl_values(Prim<R> cm, const Prim<R>& c, Prim<R> cp, R* vp, R* vm) {
  if (cm.r < R(0)) { cm = c; cm.u = -cm.u; }
  if (cp.r < R(0)) { cp = c; cp.u = -cp.u; }
  const R cc[8] = {c.r, c.u, c.v, c.p, c.Z, c.F, c.G, c.O};
  const R cb[8] = {cm.r, cm.u, cm.v, cm.p, cm.Z, cm.F, cm.G, cm.O};
  const R ca[8] = {cp.r, cp.u, cp.v, cp.p, cp.Z, cp.F, cp.G, cp.O};
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    vp[q] = muscl<true>(cc[q], cc[q] - cb[q], ca[q] - cc[q]);
    vm[q] = muscl<false>(cc[q], cc[q] - cb[q], ca[q] - cc[q]);
  }
}

// The larger tangential pressure jump ratio at cell c, from its tangential neighbours a, b
// (a wall neighbour counts as c's own pressure): the cell's part of the HLL blend weight
// (solver.js:183-190).
template <typename R>
__device__ __forceinline__ R tang_jump(const Prim<R>& c, const Prim<R>& a, const Prim<R>& b) {
  return fmax(jump_ratio(a.r < R(0) ? c.p : a.p, c.p), jump_ratio(b.r < R(0) ? c.p : b.p, c.p));
}

// Face flux: MUSCL van Leer reconstruction with positivity fallback, wall mirror
// states, HLLC with Roe-averaged wave speeds blended to HLL by `hllw`
// (solver.js:177-255). Inputs are in the face-normal frame (u normal).
// Outputs F = (mass, normal mom, tangential mom, energy), FY = (rhoZ u, rhoF u,
// uf*Gf, uf), FO = rhoO u.
template <typename R>
__device__ __forceinline__ void riemann_flux(const Prim<R>& WL, const Prim<R>& WR, R hllw, R* F, R* FY, R& FO);
template <typename R>
__device__ __forceinline__ void face_flux(Prim<R> c0, Prim<R> c1, Prim<R> c2, Prim<R> c3, R hllw, bool second,
                                          R* F, R* FY, R& FO) {
  const bool s0 = c0.r < R(0), s1 = c1.r < R(0), s2 = c2.r < R(0), s3 = c3.r < R(0);
  Prim<R> WL, WR;
  if (s1) { WR = c2; WL = c2; WL.u = -WL.u; }
  else if (s2) { WL = c1; WR = c1; WR.u = -WR.u; }
  else {
    if (s0) { c0 = c1; c0.u = -c0.u; }
    if (s3) { c3 = c2; c3.u = -c3.u; }
    WL = c1;
    WR = c2;
    if (second) {
      const R tLr = muscl<true>(c1.r, c1.r - c0.r, c2.r - c1.r);
      const R tLu = muscl<true>(c1.u, c1.u - c0.u, c2.u - c1.u);
      const R tLv = muscl<true>(c1.v, c1.v - c0.v, c2.v - c1.v);
      const R tLp = muscl<true>(c1.p, c1.p - c0.p, c2.p - c1.p);
      const R tRr = muscl<false>(c2.r, c2.r - c1.r, c3.r - c2.r);
      const R tRu = muscl<false>(c2.u, c2.u - c1.u, c3.u - c2.u);
      const R tRv = muscl<false>(c2.v, c2.v - c1.v, c3.v - c2.v);
      const R tRp = muscl<false>(c2.p, c2.p - c1.p, c3.p - c2.p);
      if (tLr > R(0) && tLp > R(0) && tRr > R(0) && tRp > R(0)) {
        WL.r = tLr; WL.u = tLu; WL.v = tLv; WL.p = tLp;
        WR.r = tRr; WR.u = tRu; WR.v = tRv; WR.p = tRp;
        WL.Z = clampR(muscl<true>(c1.Z, c1.Z - c0.Z, c2.Z - c1.Z), R(0), R(1));
        WL.F = clampR(muscl<true>(c1.F, c1.F - c0.F, c2.F - c1.F), R(0), R(1));
        WL.G = clampR(muscl<true>(c1.G, c1.G - c0.G, c2.G - c1.G), R(0.5), R(100));
        WL.O = clampR(muscl<true>(c1.O, c1.O - c0.O, c2.O - c1.O), R(0), R(1));
        WR.Z = clampR(muscl<false>(c2.Z, c2.Z - c1.Z, c3.Z - c2.Z), R(0), R(1));
        WR.F = clampR(muscl<false>(c2.F, c2.F - c1.F, c3.F - c2.F), R(0), R(1));
        WR.G = clampR(muscl<false>(c2.G, c2.G - c1.G, c3.G - c2.G), R(0.5), R(100));
        WR.O = clampR(muscl<false>(c2.O, c2.O - c1.O, c3.O - c2.O), R(0), R(1));
      }
    }
  }
  riemann_flux(WL, WR, hllw, F, FY, FO);
}

// Face states of the face between cells c1 and c2 from their MUSCL values (muscl_values:
// vL the left cell's value ahead, vR the right cell's value behind): as face_flux forms them.
template <typename R>
__device__ __forceinline__ void face_states(const Prim<R>& c1, const Prim<R>& c2, const R* vL, const R* vR, bool second,
                                            Prim<R>& WL, Prim<R>& WR) {
  const bool s1 = c1.r < R(0), s2 = c2.r < R(0);
  if (s1) { WR = c2; WL = c2; WL.u = -WL.u; return; }
  if (s2) { WL = c1; WR = c1; WR.u = -WR.u; return; }
  WL = c1;
  WR = c2;
  if (second && vL[0] > R(0) && vL[3] > R(0) && vR[0] > R(0) && vR[3] > R(0)) {
    WL.r = vL[0]; WL.u = vL[1]; WL.v = vL[2]; WL.p = vL[3];
    WR.r = vR[0]; WR.u = vR[1]; WR.v = vR[2]; WR.p = vR[3];
    WL.Z = clampR(vL[4], R(0), R(1));
    WL.F = clampR(vL[5], R(0), R(1));
    WL.G = clampR(vL[6], R(0.5), R(100));
    WL.O = clampR(vL[7], R(0), R(1));
    WR.Z = clampR(vR[4], R(0), R(1));
    WR.F = clampR(vR[5], R(0), R(1));
    WR.G = clampR(vR[6], R(0.5), R(100));
    WR.O = clampR(vR[7], R(0), R(1));
  }
}

// HLLC flux of the face states WL, WR with Roe-averaged wave speeds, blended to HLL by
// `hllw` (solver.js:206-255), in the face-normal frame.
template <typename R>
__device__ __forceinline__ void riemann_flux(const Prim<R>& WL, const Prim<R>& WR, R hllw, R* F, R* FY, R& FO) {
  const R GL = WL.G, GR = WR.G;
  const R rL = WL.r, uL = WL.u, vL = WL.v, pL = WL.p;
  const R rR = WR.r, uR = WR.u, vR = WR.v, pR = WR.p;
  const R irL = rc(rL), irR = rc(rR);
  const R cL = sq((R(1) + rc(GL)) * pL * irL), cR = sq((R(1) + rc(GR)) * pR * irR);
  const R EL = pL * GL + R(0.5) * rL * (uL * uL + vL * vL);
  const R ER = pR * GR + R(0.5) * rR * (uR * uR + vR * vR);
  const R sL = sq(rL), sR = sq(rR);
  const R isum = rc(sL + sR);
  const R uT = (sL * uL + sR * uR) * isum;
  const R HT = (sL * (EL + pL) * irL + sR * (ER + pR) * irR) * isum;
  const R vT = (sL * vL + sR * vR) * isum;
  const R GT = (sL * GL + sR * GR) * isum;
  const R cT = sq(fmax(dv(HT - R(0.5) * (uT * uT + vT * vT), GT), R(1e-12)));
  const R SL = fmin(uL - cL, uT - cT);
  const R SR = fmax(uR + cR, uT + cT);
  const R mL = rL * uL, mR = rR * uR;
  const R FL0 = mL, FL1 = mL * uL + pL, FL2 = mL * vL, FL3 = uL * (EL + pL);
  const R FR0 = mR, FR1 = mR * uR + pR, FR2 = mR * vR, FR3 = uR * (ER + pR);
  if (SL >= R(0)) {
    F[0] = FL0; F[1] = FL1; F[2] = FL2; F[3] = FL3;
    FY[0] = mL * WL.Z; FY[1] = mL * WL.F; FO = mL * WL.O;
    FY[2] = uL * GL; FY[3] = uL;
    return;
  }
  if (SR <= R(0)) {
    F[0] = FR0; F[1] = FR1; F[2] = FR2; F[3] = FR3;
    FY[0] = mR * WR.Z; FY[1] = mR * WR.F; FO = mR * WR.O;
    FY[2] = uR * GR; FY[3] = uR;
    return;
  }
  const R Ss = dv(pR - pL + mL * (SL - uL) - mR * (SR - uR), rL * (SL - uL) - rR * (SR - uR));
  const bool left = Ss >= R(0);
  const R rK = left ? rL : rR, uK = left ? uL : uR, vK = left ? vL : vR, pK = left ? pL : pR;
  const R EK = left ? EL : ER, SK = left ? SL : SR;
  const R YK0 = left ? WL.Z : WR.Z, YK1 = left ? WL.F : WR.F, YK2 = left ? WL.O : WR.O;
  const R FK0 = left ? FL0 : FR0, FK1 = left ? FL1 : FR1, FK2 = left ? FL2 : FR2, FK3 = left ? FL3 : FR3;
  const R mK = rK * uK;
  const R iSKs = rc(SK - Ss);
  const R kk = rK * (SK - uK) * iSKs;
  const R irK = left ? irL : irR;  // 1 / rK, already formed for the side it takes
  const R Us0 = kk, Us1 = kk * Ss, Us2 = kk * vK, Us3 = kk * (EK * irK + (Ss - uK) * (Ss + dv(pK * irK, SK - uK)));
  F[0] = FK0 + SK * (Us0 - rK);
  F[1] = FK1 + SK * (Us1 - mK);
  F[2] = FK2 + SK * (Us2 - rK * vK);
  F[3] = FK3 + SK * (Us3 - EK);
  FY[0] = mK * YK0 + SK * (kk * YK0 - rK * YK0);
  FY[1] = mK * YK1 + SK * (kk * YK1 - rK * YK1);
  FO = mK * YK2 + SK * (kk * YK2 - rK * YK2);
  // blend towards HLL (mix(hllc, hll, w)); with w = 0 the mix returns the HLLC flux
  // itself, so the HLL flux is formed only where some face of the wave needs it
  if (hllw > R(0)) {
    const R inv = rc(SR - SL);
    const R SLR = SL * SR;
    const R Fh0 = (SR * FL0 - SL * FR0 + SLR * (rR - rL)) * inv;
    const R Fh1 = (SR * FL1 - SL * FR1 + SLR * (rR * uR - rL * uL)) * inv;
    const R Fh2 = (SR * FL2 - SL * FR2 + SLR * (rR * vR - rL * vL)) * inv;
    const R Fh3 = (SR * FL3 - SL * FR3 + SLR * (ER - EL)) * inv;
    const R Yh0 = (SR * mL * WL.Z - SL * mR * WR.Z + SLR * (rR * WR.Z - rL * WL.Z)) * inv;
    const R Yh1 = (SR * mL * WL.F - SL * mR * WR.F + SLR * (rR * WR.F - rL * WL.F)) * inv;
    const R Yh2 = (SR * mL * WL.O - SL * mR * WR.O + SLR * (rR * WR.O - rL * WL.O)) * inv;
    F[0] = mixR(F[0], Fh0, hllw);
    F[1] = mixR(F[1], Fh1, hllw);
    F[2] = mixR(F[2], Fh2, hllw);
    F[3] = mixR(F[3], Fh3, hllw);
    FY[0] = mixR(FY[0], Yh0, hllw);
    FY[1] = mixR(FY[1], Yh1, hllw);
    FO = mixR(FO, Yh2, hllw);
  }
  const R uf = uK + SK * ((SK - uK) * iSKs - R(1));
  FY[2] = uf * (left ? GL : GR);
  FY[3] = uf;
}

// HLL blending weight of a face from its two cells' tangential jump ratios (tang_jump;
// solver.js:183-190)
template <typename R>
__device__ __forceinline__ R hll_weight(R j1, R j2, R blend) {
  return smoothstepR(R(0.15), R(0.6), fmax(j1, j2) * blend);
}

// reflected-Riemann pressure on a slip wall (solver.js:301-306)
template <typename R>
__device__ __forceinline__ R wall_pressure(const Prim<R>& q, R un) {
  const R g = R(1) + rc(q.G), p = q.p, rho = q.r, c = sq(dv(g * p, rho));
  if (un > R(0)) {
    const R a = R(0.25) * (g + R(1)) * un;
    return p + rho * un * (a + sq(a * a + c * c));
  }
  const R f = fmax(R(1) + dv(R(0.5) * (g - R(1)) * un, c), R(0));
  return p * powR(f, dv(R(2) * g, g - R(1)));
}

// Sutherland viscosity and mixture gas constant (solver.js:448-449)
template <typename R>
__device__ __forceinline__ R sutherland(R TK, const Consts& c) { return dv(R(c.gas[3]) * R(1.458e-6) * TK * sq(TK), TK + R(110.4)); }
template <typename R>
__device__ __forceinline__ R rmix(R Z, R F, R O, const Consts& c) {
  return (R(1) - Z) * R(c.gas[0]) + F * R(c.Rs[0]) + O * R(c.Rs[1]) + fmax(Z - F - O, R(0)) * R(c.Rs[2]);
}

// Per-cell viscous quantities (FS_VCELL, solver.js:458-487): T1 = (ux, ur, vx, vr),
// T2 = (mu_eff, h, |S|*sensor, mu_lam). Neighbours wE..wS may be walls (r < 0).
template <typename R>
__device__ __forceinline__ void vcell(const Prim<R>& w0, const Prim<R>& wE, const Prim<R>& wW, const Prim<R>& wN,
                                      const Prim<R>& wS, int i, int j, R l, R dt, const Consts& c, R* T1, R* T2) {
  const R fE = wE.r < R(0) ? R(0) : R(1), fW = wW.r < R(0) ? R(0) : R(1);
  const R fN = wN.r < R(0) ? R(0) : R(1), fS = wS.r < R(0) ? R(0) : R(1);
  const R idx = rc(fmax(fE + fW, R(1))), idr = rc(fmax(fN + fS, R(1)));
  const R ux = (fE * (wE.u - w0.u) + fW * (w0.u - wW.u)) * idx;
  const R vx = (fE * (wE.v - w0.v) + fW * (w0.v - wW.v)) * idx;
  const R ur = (fN * (wN.u - w0.u) + fS * (w0.u - wS.u)) * idr;
  const R vr = (fN * (wN.v - w0.v) + fS * (w0.v - wS.v)) * idr;
  const R r = R(j) + R(0.5), vor = dv(w0.v, r);
  const R div = ux + vr + vor;
  const R S2 = R(2) * (ux * ux + vr * vr + vor * vor) + (ur + vx) * (ur + vx) - R(0.6666667f) * div * div;
  const R om = vx - ur;
  const R sens = dv(om * om, om * om + div * div + R(1e-8));
  const R Smag = sq(fmax(S2, R(0))) * sens;
  const R Rm = rmix(w0.Z, w0.F, w0.O, c);
  const R ir0 = rc(w0.r);
  const R mu = sutherland(dv(w0.p * ir0, Rm) * R(c.gas[2]), c);
  R mut = w0.r * l * l * Smag;
  if (R(i) < R(c.chamberNu[1])) mut = fmax(mut, w0.r * R(c.chamberNu[0]));
  T1[0] = ux; T1[1] = ur; T1[2] = vx; T1[3] = vr;
  T2[0] = fmin(mu + mut, dv(R(0.08) * w0.r, fmax(dt, R(1e-9))));
  T2[1] = (w0.G + R(1)) * w0.p * ir0;
  T2[2] = Smag;
  T2[3] = mu;
}

// Viscous face flux between cells 1 (west/south) and 2 (FS_VFLUX, solver.js:499-536).
// F = (ox-vapour diffusion, x mom, r mom, energy), FY = (Z, F, G diffusion).
template <typename R>
__device__ __forceinline__ void visc_face(bool dirR, int j, const Prim<R>& w1, const Prim<R>& w2, const R* t1, const R* t2,
                                          const R* m1, const R* m2, R* F, R* FY) {
  const R PR_INV = R(1.25);
  const R mu = R(0.5) * (m1[0] + m2[0]);
  const R uf = R(0.5) * (w1.u + w2.u), vf = R(0.5) * (w1.v + w2.v);
  R r, ux, vx, ur, vr;
  if (dirR) {
    r = fmax(R(j), R(0.5));
    ux = R(0.5) * (t1[0] + t2[0]); vx = R(0.5) * (t1[2] + t2[2]);
    ur = w2.u - w1.u; vr = w2.v - w1.v;
  } else {
    r = R(j) + R(0.5);
    ux = w2.u - w1.u; vx = w2.v - w1.v;
    ur = R(0.5) * (t1[1] + t2[1]); vr = R(0.5) * (t1[3] + t2[3]);
  }
  const R div = ux + vr + dv(vf, r);
  const R txr = mu * (ur + vx);
  const R k = mu * PR_INV;
  F[0] = -k * (w2.O - w1.O);
  if (dirR) {
    const R trr = mu * (R(2) * vr - R(0.6666667f) * div);
    F[1] = -txr; F[2] = -trr; F[3] = -(uf * txr + vf * trr) - k * (m2[1] - m1[1]);
  } else {
    const R txx = mu * (R(2) * ux - R(0.6666667f) * div);
    F[1] = -txx; F[2] = -txr; F[3] = -(uf * txx + vf * txr) - k * (m2[1] - m1[1]);
  }
  FY[0] = -k * (w2.Z - w1.Z);
  FY[1] = -k * (w2.F - w1.F);
  FY[2] = -k * (w2.G - w1.G);
}

// PCG hash and uniform in [0,1) (spray.js:33-34)
__device__ __forceinline__ uint32_t pcg(uint32_t v) {
  const uint32_t s = v * 747796405u + 2891336453u;
  const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
  return (w >> 22u) ^ w;
}
__device__ __forceinline__ float rnd(uint32_t a, uint32_t n) { return (float)(pcg(a + n * 7919u) & 0xffffffu) / 16777216.0f; }

// Fixed-point units of the spray deposit this step: s[0] for the mass channels, s[1] for the
// momentum channels (integer units per normalised unit)
__device__ __forceinline__ void dep_scales(const Consts& c, double mmax, double* s) {
  const double one = ldexp(1.0, c.depExp);
  const double bm = fmax(mmax, 1e-300);
  s[0] = one / bm;
  s[1] = one / fmax(2.0 * bm * (double)c.depVel, 1e-300);
}
// one release in fixed point, clamped to twice its bound (a release beyond it would break the
// bound the headroom assumes; it cannot occur for parcels spawned by this library)
__device__ __forceinline__ unsigned long long dep_fix(double v, double scale, int depExp) {
  const double lim = ldexp(2.0, depExp);
  const double r = rint(v * scale);
  const double q = r == r ? fmin(fmax(r, -lim), lim) : 0.0;  // a NaN release adds nothing
  return (unsigned long long)(long long)q;
}

}  // namespace nzk
