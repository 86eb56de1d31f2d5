// kernels.hip — fused per-step kernels of the MI355X nozzle solver (gfx950, wave64).
//
// One solver step is two (three with the spray) launches instead of the
// reference's ~21 draws (solver.js:849-891):
//   k_stage<1>   LDS tile of primitives with a 2-cell ghost halo (FS_PRIM) ->
//                each x/r face flux computed once (MUSCL + HLLC/HLL, FS_FLUX) ->
//                merged-group finite-volume update U1 = U0 + dt L(U0) (FS_UPDATE).
//   k_spray      parcel push, evaporation, break-up; deposits its release with
//                float atomics into the host cell (FS_SPRAY_UPDATE + FS_SPRAY_DEP).
//   k_stage<2>   the same flux/update on U1 giving U^{n+1} = (U0 + U1 + dt L(U1))/2,
//                fused with the viscous / wall / afterburning split step
//                (FS_VCELL, FS_VFLUX, FS_SOURCE), the spray source apply
//                (FS_SPRAY_APPLY) and the CFL wave-speed max of the new state
//                (FS_REDUCE1/2). U^{n+1} overwrites U0 in place: every group's
//                old value is read and rewritten by its host cell's thread only.
// The time step of the next step is taken by the first launch of that step from
// the wave-speed max and the previous step info (FS_STEPINFO), so dt never
// leaves the device. Launches carry a table of instances: one launch advances
// every resident solve (batched serving uses the same kernels).
#include "nz_device.h"

#include <algorithm>

namespace nzk {

// ------------------------------------------------------------------------------
// instance lookup for a flattened tile index (instances sorted by tile0)
__device__ __forceinline__ int find_inst(const DevInst* __restrict__ I, int n, int t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (I[m].tile0 <= t) lo = m; else hi = m - 1;
  }
  return lo;
}

// FS_STEPINFO (solver.js:404-409): dt from the wave-speed max of the current state,
// in the arithmetic type of the state
template <typename R>
__device__ __forceinline__ void step_info(const Consts& c, const InstState& s, R* cur) {
  const R w = fmax(R(c.wFloor), R(__longlong_as_double((long long)s.wsnap)));
  R dt = fmin(R(c.cfl) / w, R(c.dtMax));
  const R lastDt = R(s.last[0]), lastT = R(s.last[2]);
  if (lastDt > R(0)) dt = fmin(dt, lastDt * R(1.2f));
  cur[0] = dt;
  cur[1] = lastT;
  cur[2] = lastT + dt;
  cur[3] = w;
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double m = 0.0;
  if (threadIdx.x == 0) {
    for (int k = 0; k < NT / 64; ++k) m = fmax(m, red[k]);
  }
  return m;  // valid on thread 0
}
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double m = 0.0;
  if (threadIdx.x == 0) {
    for (int k = 0; k < NT / 64; ++k) m += red[k];
  }
  return m;
}

// ------------------------------------------------------------------------------
// One face's inviscid flux times |area| (FS_FLUX_X/R for the face, as memberFlux
// weights it), with primitives from `getP(i, j)`: the LDS tile, or the tile with a
// global fall-back for faces of merged cells that lie outside it.
template <typename R, class GetP>
__device__ __forceinline__ void face_flux_w(const DevInst& I, bool dirR, int fi, int fj, GetP getP, R* F, R* FY, R& FO) {
  for (int q = 0; q < 4; ++q) { F[q] = R(0); FY[q] = R(0); }
  FO = R(0);
  const int Nx = I.Nx, Nr = I.Nr;
  const bool inDom = dirR ? (fi >= 0 && fi < Nx && fj >= 0 && fj <= Nr) : (fi >= 0 && fi <= Nx && fj >= 0 && fj < Nr);
  if (!inDom) return;
  const float area = dirR ? I.ar[(long long)fj * Nx + fi] : I.ax[(long long)fj * (Nx + 1) + fi];
  if (area == 0.f || (dirR && fj == 0)) return;
  const int dx = dirR ? 0 : 1, dy = dirR ? 1 : 0, tx = dy, ty = dx;
  Prim<R> c0 = getP(fi - 2 * dx, fj - 2 * dy), c1 = getP(fi - dx, fj - dy);
  Prim<R> c2 = getP(fi, fj), c3 = getP(fi + dx, fj + dy);
  const bool s1 = c1.r < R(0), s2 = c2.r < R(0);
  if (s1 && s2) return;
  R hllw = R(0);
  if (I.c.hllBlend > 0.f && !(s1 || s2)) {
    const Prim<R> a1 = getP(fi - dx + tx, fj - dy + ty), b1 = getP(fi - dx - tx, fj - dy - ty);
    const Prim<R> a2 = getP(fi + tx, fj + ty), b2 = getP(fi - tx, fj - ty);
    hllw = hll_weight<R>(c1.p, c2.p, a1.r < R(0) ? c1.p : a1.p, b1.r < R(0) ? c1.p : b1.p,
                         a2.r < R(0) ? c2.p : a2.p, b2.r < R(0) ? c2.p : b2.p, R(I.c.hllBlend));
  }
  if (dirR) {
    R t;
    t = c0.u; c0.u = c0.v; c0.v = t;
    t = c1.u; c1.u = c1.v; c1.v = t;
    t = c2.u; c2.u = c2.v; c2.v = t;
    t = c3.u; c3.u = c3.v; c3.v = t;
  }
  face_flux<R>(c0, c1, c2, c3, hllw, I.c.order2 && area > 0.f, F, FY, FO);
  if (dirR) { const R t = F[1]; F[1] = F[2]; F[2] = t; }
  const R a = R(fabsf(area));
  for (int q = 0; q < 4; ++q) { F[q] *= a; FY[q] *= a; }
  FO *= a;
}

// One face's viscous flux times |area| (FS_VFLUX_X/R); getT(i, j, T1, T2) gives the
// per-cell gradients / viscosities of an interior fluid cell.
template <typename R, class GetP, class GetT>
__device__ __forceinline__ void vface_w(const DevInst& I, bool dirR, int fi, int fj, GetP getP, GetT getT, R* F, R* FY) {
  for (int q = 0; q < 4; ++q) F[q] = R(0);
  for (int q = 0; q < 3; ++q) FY[q] = R(0);
  const int Nx = I.Nx, Nr = I.Nr;
  const bool on = dirR ? (fi >= 0 && fi < Nx && fj > 0 && fj < Nr) : (fi > 0 && fi < Nx && fj >= 0 && fj < Nr);
  if (!on) return;
  const float area = dirR ? I.ar[(long long)fj * Nx + fi] : I.ax[(long long)fj * (Nx + 1) + fi];
  if (area == 0.f) return;
  const int k1i = dirR ? fi : fi - 1, k1j = dirR ? fj - 1 : fj;
  const Prim<R> w1 = getP(k1i, k1j), w2 = getP(fi, fj);
  if (w1.r < R(0) || w2.r < R(0)) return;
  R t1[4], m1[4], t2[4], m2[4];
  getT(k1i, k1j, t1, m1);
  getT(fi, fj, t2, m2);
  visc_face<R>(dirR, fj, w1, w2, t1, t2, m1, m2, F, FY);
  const R a = R(fabsf(area));
  for (int q = 0; q < 4; ++q) F[q] *= a;
  for (int q = 0; q < 3; ++q) FY[q] *= a;
}

// T1/T2 of an interior cell from primitives given by getP (zero for walls)
template <typename R, class GetP>
__device__ __forceinline__ void vcell_at(const DevInst& I, GetP getP, R dt, int i, int j, R* T1, R* T2) {
  for (int q = 0; q < 4; ++q) { T1[q] = R(0); T2[q] = R(0); }
  const Prim<R> w0 = getP(i, j);
  if (w0.r < R(0)) return;
  vcell<R>(w0, getP(i + 1, j), getP(i - 1, j), getP(i, j + 1), getP(i, j - 1), i, j, R(I.lt[(long long)j * I.Nx + i]), dt,
           I.c, T1, T2);
}

// T1/T2 of an interior cell from global memory
template <typename R>
__device__ void vcell_global(const DevInst& I, const R* U, R t, R dt, int i, int j, R* T1, R* T2) {
  auto g = [&](int a, int b) { return prim_at<R>(I, U, a, b, t); };
  vcell_at<R>(I, g, dt, i, j, T1, T2);
}

// extra face of a merged cell outside the tile: (face x | dir << 31, face y)
__device__ __forceinline__ bool ex_dir(uint2 e) { return (e.x >> 31) != 0; }
__device__ __forceinline__ int ex_i(uint2 e) { return (int)(e.x & 0x7fffffffu); }
__device__ __forceinline__ int ex_j(uint2 e) { return (int)e.y; }

// ------------------------------------------------------------------------------
template <typename R, int TX, int TY>
struct StageSmem {
  static constexpr int PW = TX + 4, PH = TY + 4;   // primitive tile with 2-cell halo
  static constexpr int NFX = (TX + 1) * TY, NFR = TX * (TY + 1);
  static constexpr int TW = TX + 2, THh = TY + 2;  // T1/T2 tile with 1-cell ring
  R P[8][PH][PW];
  union {
    struct {
      R Fx[9][TY][TX + 1];
      R Fr[9][TY + 1][TX];
    } inv;
    struct {
      R T[8][THh][TW];
      R Vx[7][TY][TX + 1];
      R Vr[7][TY + 1][TX];
    } vis;
  } u;
  static constexpr int MAXE = MAX_EXTRA_FACES;     // faces of merged cells outside the tile
  static constexpr int MAXF = MAX_FAR_CELLS;       // cells those faces reach beyond the halo
  static constexpr int WW = TX + 2 * FAR_HALO, WH = TY + 2 * FAR_HALO;
  uint2 ex[MAXE];
  int2 fc[MAXF];
  R PF[8][MAXF];
  unsigned char fmap[WH][WW];                      // far-cell slot by window position
  int nEx, nF;
  double red[16];
  double redd[16];
  int last;
};

template <typename R, int TX, int TY, int STAGE, bool PHYS>
__global__ void __launch_bounds__(TX* TY) k_stage(const DevInst* __restrict__ insts, int ninst) {
  constexpr int NT = TX * TY;
  using S = StageSmem<R, TX, TY>;
  __shared__ S sm;
  const int tid = threadIdx.x;
  const int t = blockIdx.x;
  const int ii = find_inst(insts, ninst, t);
  const DevInst& I = insts[ii];
  const Consts& c = I.c;
  InstState* st = I.st;
  const bool active = !(st->flags & 1u);
  const int Nx = I.Nx, Nr = I.Nr;
  const long long Ns = I.Ns;
  const int lt_ = t - I.tile0;
  const int i0 = (lt_ % I.tilesX) * TX, j0 = (lt_ / I.tilesX) * TY;
  R* __restrict__ U0 = static_cast<R*>(I.U0);
  R* __restrict__ U1 = static_cast<R*>(I.U1);
  const R* __restrict__ Uin = STAGE == 1 ? U0 : U1;

  // time step of this step: FS_STEPINFO at the start of the step
  R cur[4];
  if (STAGE == 1) {
    step_info<R>(c, *st, cur);
    if (lt_ == 0 && tid == 0 && active) { st->cur[0] = cur[0]; st->cur[1] = cur[1]; st->cur[2] = cur[2]; st->cur[3] = cur[3]; }
  } else {
    cur[0] = R(st->cur[0]); cur[1] = R(st->cur[1]); cur[2] = R(st->cur[2]); cur[3] = R(st->cur[3]);
  }
  const R dt = cur[0];
  const R tsel = STAGE == 1 ? cur[1] : cur[2];

  double wloc = 0.0;
  double rloc = 0.0;
  if (active) {
    // ---- 1. primitive tile with ghost halo (FS_PRIM); this tile's extra faces
    for (int idx = tid; idx < S::PW * S::PH; idx += NT) {
      const int li = idx % S::PW, lj = idx / S::PW;
      const Prim<R> q = prim_at<R>(I, Uin, i0 - 2 + li, j0 - 2 + lj, tsel);
      sm.P[0][lj][li] = q.r; sm.P[1][lj][li] = q.u; sm.P[2][lj][li] = q.v; sm.P[3][lj][li] = q.p;
      sm.P[4][lj][li] = q.Z; sm.P[5][lj][li] = q.F; sm.P[6][lj][li] = q.G; sm.P[7][lj][li] = q.O;
    }
    // extra faces of merged cells outside the tile, and the cells they reach beyond
    // the halo (primitives loaded here in parallel, so no thread waits on them later)
    const int exBase = I.exStart[lt_];
    const int nEx = min(I.exStart[lt_ + 1] - exBase, S::MAXE);
    const int fcBase = I.fcStart[lt_];
    const int nF = min(I.fcStart[lt_ + 1] - fcBase, S::MAXF);
    for (int e = tid; e < nEx; e += NT) sm.ex[e] = I.exFaces[exBase + e];
    for (int e = tid; e < nF; e += NT) {
      const int2 fcell = I.fcCells[fcBase + e];
      sm.fc[e] = fcell;
      const Prim<R> q = prim_at<R>(I, Uin, fcell.x, fcell.y, tsel);
      sm.PF[0][e] = q.r; sm.PF[1][e] = q.u; sm.PF[2][e] = q.v; sm.PF[3][e] = q.p;
      sm.PF[4][e] = q.Z; sm.PF[5][e] = q.F; sm.PF[6][e] = q.G; sm.PF[7][e] = q.O;
    }
    if (nF > 0) {  // uniform over the workgroup
      for (int w = tid; w < S::WW * S::WH; w += NT) sm.fmap[w / S::WW][w % S::WW] = 0xFF;
      __syncthreads();
      for (int e = tid; e < nF; e += NT) {
        const int2 fcell = I.fcCells[fcBase + e];
        const int wi = fcell.x - (i0 - FAR_HALO), wj = fcell.y - (j0 - FAR_HALO);
        if (wi >= 0 && wi < S::WW && wj >= 0 && wj < S::WH) sm.fmap[wj][wi] = (unsigned char)e;
      }
    }
    R* __restrict__ exFlux = static_cast<R*>(I.exFlux) + (long long)exBase * 16;
    auto P_at = [&](int li, int lj) {
      Prim<R> q;
      q.r = sm.P[0][lj][li]; q.u = sm.P[1][lj][li]; q.v = sm.P[2][lj][li]; q.p = sm.P[3][lj][li];
      q.Z = sm.P[4][lj][li]; q.F = sm.P[5][lj][li]; q.G = sm.P[6][lj][li]; q.O = sm.P[7][lj][li];
      return q;
    };
    // primitives by global cell index: the LDS tile, which always covers in-tile faces
    auto tileP = [&](int gi, int gj) { return P_at(gi - (i0 - 2), gj - (j0 - 2)); };
    // ... and, for faces of merged cells outside the tile, the preloaded far cells
    // (slot from the window map), else global memory
    auto anyP = [&](int gi, int gj) {
      const int li = gi - (i0 - 2), lj = gj - (j0 - 2);
      if (li >= 0 && li < S::PW && lj >= 0 && lj < S::PH) return P_at(li, lj);
      const int wi = gi - (i0 - FAR_HALO), wj = gj - (j0 - FAR_HALO);
      if (nF > 0 && wi >= 0 && wi < S::WW && wj >= 0 && wj < S::WH) {
        const int e = sm.fmap[wj][wi];
        if (e < nF) {
          Prim<R> q;
          q.r = sm.PF[0][e]; q.u = sm.PF[1][e]; q.v = sm.PF[2][e]; q.p = sm.PF[3][e];
          q.Z = sm.PF[4][e]; q.F = sm.PF[5][e]; q.G = sm.PF[6][e]; q.O = sm.PF[7][e];
          return q;
        }
      }
      return prim_at<R>(I, Uin, gi, gj, tsel);
    };
    __syncthreads();

    // ---- 2. each face flux once (FS_FLUX_X / FS_FLUX_R), times |area|; extra faces last
    for (int f = tid; f < S::NFX + S::NFR + nEx; f += NT) {
      R F[4], FY[4], FO;
      if (f < S::NFX + S::NFR) {
        const bool dirR = f >= S::NFX;
        const int g = dirR ? f - S::NFX : f;
        const int fx = dirR ? g % TX : g % (TX + 1);
        const int fy = dirR ? g / TX : g / (TX + 1);
        face_flux_w<R>(I, dirR, i0 + fx, j0 + fy, tileP, F, FY, FO);
        if (dirR) {
          for (int q = 0; q < 4; ++q) { sm.u.inv.Fr[q][fy][fx] = F[q]; sm.u.inv.Fr[4 + q][fy][fx] = FY[q]; }
          sm.u.inv.Fr[8][fy][fx] = FO;
        } else {
          for (int q = 0; q < 4; ++q) { sm.u.inv.Fx[q][fy][fx] = F[q]; sm.u.inv.Fx[4 + q][fy][fx] = FY[q]; }
          sm.u.inv.Fx[8][fy][fx] = FO;
        }
      } else {
        const int e = f - S::NFX - S::NFR;
        const uint2 ef = sm.ex[e];
        face_flux_w<R>(I, ex_dir(ef), ex_i(ef), ex_j(ef), anyP, F, FY, FO);
        for (int q = 0; q < 4; ++q) { exFlux[e * 16 + q] = F[q]; exFlux[e * 16 + 4 + q] = FY[q]; }
        exFlux[e * 16 + 8] = FO;
      }
    }
    __syncthreads();

    // weighted inviscid flux of face f (W, E, S, N = 0..3) of member cell (mi, mj):
    // the in-tile array for faces of the tile, the slave's extra-face slots (`slot`,
    // advanced per face taken) for the others, or evaluated here when it has none
    auto invFace = [&](int f, int mi_, int mj_, int& slot, R* F9) {
      const bool dirR = f >= 2;
      const int fi = mi_ + (f == 1 ? 1 : 0), fj = mj_ + (f == 3 ? 1 : 0);
      const int fx = fi - i0, fy = fj - j0;
      if (!dirR && fx >= 0 && fx <= TX && fy >= 0 && fy < TY) {
        for (int q = 0; q < 9; ++q) F9[q] = sm.u.inv.Fx[q][fy][fx];
      } else if (dirR && fx >= 0 && fx < TX && fy >= 0 && fy <= TY) {
        for (int q = 0; q < 9; ++q) F9[q] = sm.u.inv.Fr[q][fy][fx];
      } else if (slot >= 0 && slot < nEx) {
        for (int q = 0; q < 9; ++q) F9[q] = exFlux[slot * 16 + q];
        ++slot;
      } else {
        face_flux_w<R>(I, dirR, fi, fj, anyP, F9, F9 + 4, F9[8]);
      }
    };

    // ---- 3. merged-group finite-volume update (FS_UPDATE)
    const int tx = tid % TX, ty = tid / TX;
    const int ci = i0 + tx, cj = j0 + ty;
    const bool inDom = ci < Nx && cj < Nr;
    const long long k = (long long)cj * Nx + ci;
    const uint32_t code = inDom ? I.code[k] : C_IS_SLAVE;
    const bool solid = (code & C_SOLID) != 0;
    const bool host = inDom && !solid && !(code & C_IS_SLAVE);
    R A[4], B[4];         // group state after the stage
    R ubarM[5];           // member-specific B.w (inflow-speed average in column 0)
    int mi[5], mj[5], nm = 0;
    int mslot[5];         // first extra-face slot of each member (-1: faces in the tile or none)
    bool guard = false;   // NaN guard hit
    R rho_old = R(0);
    float4 gg = make_float4(0.f, 0.f, 0.f, 0.f);
    Prim<R> qS;           // state after the stage, as primitives (physics input)
    if (host) {
      mi[0] = ci; mj[0] = cj; mslot[0] = -1; nm = 1;
      const uint32_t sl = code_slaves(code);
      if (sl & 1u) { mi[nm] = ci - 1; mj[nm] = cj; ++nm; }
      if (sl & 2u) { mi[nm] = ci + 1; mj[nm] = cj; ++nm; }
      if (sl & 4u) { mi[nm] = ci; mj[nm] = cj - 1; ++nm; }
      if (sl & 8u) { mi[nm] = ci; mj[nm] = cj + 1; ++nm; }
      // slaves outside the tile take 3 consecutive extra-face slots each, in member order
      int es = code_exslot(code) * 3;
      for (int m = 1; m < nm; ++m) {
        const bool inT = mi[m] >= i0 && mi[m] < i0 + TX && mj[m] >= j0 && mj[m] < j0 + TY;
        mslot[m] = (inT || es < 0) ? -1 : es;
        if (!inT && es >= 0) es += 3;
      }
      R L[4], LY[4], LO;
      for (int q = 0; q < 4; ++q) {
        L[q] = -(sm.u.inv.Fx[q][ty][tx + 1] - sm.u.inv.Fx[q][ty][tx]) - (sm.u.inv.Fr[q][ty + 1][tx] - sm.u.inv.Fr[q][ty][tx]);
        LY[q] = -(sm.u.inv.Fx[4 + q][ty][tx + 1] - sm.u.inv.Fx[4 + q][ty][tx]) - (sm.u.inv.Fr[4 + q][ty + 1][tx] - sm.u.inv.Fr[4 + q][ty][tx]);
      }
      LO = -(sm.u.inv.Fx[8][ty][tx + 1] - sm.u.inv.Fx[8][ty][tx]) - (sm.u.inv.Fr[8][ty + 1][tx] - sm.u.inv.Fr[8][ty][tx]);
      for (int m = 1; m < nm; ++m) {
        // the slave's own net flux: west/south faces enter, east/north faces leave
        R Fw[9], Fe_[9], Fs[9], Fn[9];
        int slot = mslot[m];
        invFace(0, mi[m], mj[m], slot, Fw);
        invFace(1, mi[m], mj[m], slot, Fe_);
        invFace(2, mi[m], mj[m], slot, Fs);
        invFace(3, mi[m], mj[m], slot, Fn);
        for (int q = 0; q < 4; ++q) {
          L[q] += -(Fe_[q] - Fw[q]) - (Fn[q] - Fs[q]);
          LY[q] += -(Fe_[4 + q] - Fw[4 + q]) - (Fn[4 + q] - Fs[4 + q]);
        }
        LO += -(Fe_[8] - Fw[8]) - (Fn[8] - Fs[8]);
      }
      gg = I.gg[k];
      R a[4], b[4];
      for (int q = 0; q < 4; ++q) { a[q] = Uin[q * Ns + k]; b[q] = Uin[(4 + q) * Ns + k]; }
      const Prim<R> q = to_prim<R>(a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]);
      const R Wx = R(gg.z), Wr = R(gg.w);
      const R wl = sqrt(Wx * Wx + Wr * Wr);
      if (wl > R(1e-6)) {
        const R iwl = rc(wl);
        const R pw = wall_pressure<R>(q, q.u * (Wx * iwl) + q.v * (Wr * iwl));
        L[1] -= Wx * pw;
        L[2] -= Wr * pw;
      }
      L[2] += q.p * R(gg.y);
      const R invV = rc(fmax(R(gg.x), R(1e-6)));
      const R LB[4] = {LY[0], LY[1], LY[2] - q.G * LY[3], LO};
      R Uq[4], UY[4];
      if (STAGE == 1) {
        for (int m = 0; m < 4; ++m) { Uq[m] = a[m] + dt * invV * L[m]; UY[m] = b[m] + dt * invV * LB[m]; }
      } else {
        R a0[4], b0[4];
        for (int m = 0; m < 4; ++m) { a0[m] = U0[m * Ns + k]; b0[m] = U0[(4 + m) * Ns + k]; }
        rho_old = a0[0];
        for (int m = 0; m < 4; ++m) {
          Uq[m] = R(0.5) * (a0[m] + a[m] + dt * invV * L[m]);
          UY[m] = R(0.5) * (b0[m] + b[m] + dt * invV * LB[m]);
        }
      }
      const R rho = fmax(Uq[0], R(1e-7));
      const R iU0 = rc(fmax(Uq[0], R(1e-7)));
      R vx = Uq[1] * iU0, vy = Uq[2] * iU0;
      const R sp = sq(vx * vx + vy * vy);
      if (sp > R(c.vMax)) { const R f = dv(R(c.vMax), sp); vx *= f; vy *= f; }
      if (!(rho < R(1e30)) || !(vx < R(1e30) && vx > R(-1e30))) {
        guard = true;
        for (int m = 0; m < 4; ++m) { A[m] = R(c.ambCons[m]); B[m] = R(c.ambB[m]); }
        for (int m = 0; m < nm; ++m) ubarM[m] = R(c.ambB[3]);
      } else {
        const R ke = R(0.5) * rho * (vx * vx + vy * vy);
        const R Gn = clampR(UY[2], R(c.gamR[0]), R(c.gamR[1]));
        const R E = fmax(Uq[3], R(1e-7) * Gn + ke);
        const R Zn = clampR(UY[0], R(0), rho);
        A[0] = rho; A[1] = rho * vx; A[2] = rho * vy; A[3] = E;
        B[0] = Zn; B[1] = clampR(UY[1], R(0), Zn); B[2] = Gn;
        B[3] = c.inj ? clampR(UY[3], R(0), rho) : R(0);
        for (int m = 0; m < nm; ++m) {
          R ub = B[3];
          if (!c.inj) {
            ub = R(0);
            if (mi[m] == 0) {
              const R u0 = Uin[7 * Ns + (long long)mj[m] * Nx];
              const bool on = STAGE == 2 && cur[2] > R(c.uBar[2]);
              ub = on ? fmin(u0 + fmin(dt * R(c.uBar[0]), R(1)) * (vx - u0), R(c.uBar[1])) : u0;
            }
          }
          ubarM[m] = ub;
        }
      }
      qS = to_prim<R>(A[0], A[1], A[2], A[3], B[0], B[1], B[2], ubarM[0]);
    }

    // ---- 4. viscous / wall / afterburning split step (FS_VCELL, FS_VFLUX, FS_SOURCE)
    if (STAGE == 2 && PHYS && (c.visc || c.chem)) {  // per instance: a batch mixes both kinds
      __syncthreads();  // inviscid fluxes consumed; reuse their LDS
      for (int idx = tid; idx < S::TW * S::THh; idx += NT) {
        const int ri = idx % S::TW, rj = idx / S::TW;
        const int li = ri + 1, lj = rj + 1;
        const int gi = i0 - 2 + li, gj = j0 - 2 + lj;
        R T1[4] = {0, 0, 0, 0}, T2[4] = {0, 0, 0, 0};
        if (gi >= 0 && gi < Nx && gj >= 0 && gj < Nr && sm.P[0][lj][li] >= R(0)) {
          vcell<R>(P_at(li, lj), P_at(li + 1, lj), P_at(li - 1, lj), P_at(li, lj + 1), P_at(li, lj - 1), gi, gj,
                   R(I.lt[(long long)gj * Nx + gi]), dt, c, T1, T2);
        }
        for (int q = 0; q < 4; ++q) { sm.u.vis.T[q][rj][ri] = T1[q]; sm.u.vis.T[4 + q][rj][ri] = T2[q]; }
      }
      __syncthreads();
      // per-cell T1/T2: the LDS ring tile, else computed from primitives
      auto anyT = [&](int gi, int gj, R* T1, R* T2) {
        const int ri = gi - (i0 - 1), rj = gj - (j0 - 1);
        if (ri >= 0 && ri < S::TW && rj >= 0 && rj < S::THh) {
          for (int q = 0; q < 4; ++q) { T1[q] = sm.u.vis.T[q][rj][ri]; T2[q] = sm.u.vis.T[4 + q][rj][ri]; }
          return;
        }
        vcell_at<R>(I, anyP, dt, gi, gj, T1, T2);
      };
      auto tileT = [&](int gi, int gj, R* T1, R* T2) {
        const int ri = gi - (i0 - 1), rj = gj - (j0 - 1);
        for (int q = 0; q < 4; ++q) { T1[q] = sm.u.vis.T[q][rj][ri]; T2[q] = sm.u.vis.T[4 + q][rj][ri]; }
      };
      if (c.visc) {
        for (int f = tid; f < S::NFX + S::NFR + nEx; f += NT) {
          R F[4], FY[3];
          if (f < S::NFX + S::NFR) {
            const bool dirR = f >= S::NFX;
            const int g = dirR ? f - S::NFX : f;
            const int fx = dirR ? g % TX : g % (TX + 1);
            const int fy = dirR ? g / TX : g / (TX + 1);
            vface_w<R>(I, dirR, i0 + fx, j0 + fy, tileP, tileT, F, FY);
            if (dirR) {
              for (int q = 0; q < 4; ++q) sm.u.vis.Vr[q][fy][fx] = F[q];
              for (int q = 0; q < 3; ++q) sm.u.vis.Vr[4 + q][fy][fx] = FY[q];
            } else {
              for (int q = 0; q < 4; ++q) sm.u.vis.Vx[q][fy][fx] = F[q];
              for (int q = 0; q < 3; ++q) sm.u.vis.Vx[4 + q][fy][fx] = FY[q];
            }
          } else {
            const int e = f - S::NFX - S::NFR;
            const uint2 ef = sm.ex[e];
            vface_w<R>(I, ex_dir(ef), ex_i(ef), ex_j(ef), anyP, anyT, F, FY);
            for (int q = 0; q < 4; ++q) exFlux[e * 16 + 9 + q] = F[q];
            for (int q = 0; q < 3; ++q) exFlux[e * 16 + 13 + q] = FY[q];
          }
        }
        __syncthreads();
      }
      auto visFace = [&](int f, int mi_, int mj_, int& slot, R* F7) {
        const bool dirR = f >= 2;
        const int fi = mi_ + (f == 1 ? 1 : 0), fj = mj_ + (f == 3 ? 1 : 0);
        const int fx = fi - i0, fy = fj - j0;
        if (!dirR && fx >= 0 && fx <= TX && fy >= 0 && fy < TY) {
          for (int q = 0; q < 7; ++q) F7[q] = sm.u.vis.Vx[q][fy][fx];
        } else if (dirR && fx >= 0 && fx < TX && fy >= 0 && fy <= TY) {
          for (int q = 0; q < 7; ++q) F7[q] = sm.u.vis.Vr[q][fy][fx];
        } else if (slot >= 0 && slot < nEx) {
          for (int q = 0; q < 7; ++q) F7[q] = exFlux[slot * 16 + 9 + q];
          ++slot;
        } else {
          vface_w<R>(I, dirR, fi, fj, anyP, anyT, F7, F7 + 4);
        }
      };
      if (host) {
        const R rho = qS.r, p = qS.p, G = qS.G, velx = qS.u, vely = qS.v;
        const R V = fmax(R(gg.x), R(1e-6));
        const R Rm = rmix(qS.Z, qS.F, qS.O, c), Tg = dv(p, rho * Rm);
        R L[4] = {0, 0, 0, 0}, LY[3] = {0, 0, 0};
        R T1h[4], T2h[4];
        for (int q = 0; q < 4; ++q) { T1h[q] = sm.u.vis.T[q][ty + 1][tx + 1]; T2h[q] = sm.u.vis.T[4 + q][ty + 1][tx + 1]; }
        if (c.visc) {
          for (int m = 0; m < nm; ++m) {
            R Fw[7], Fe_[7], Fs[7], Fn[7];
            int slot = mslot[m];
            visFace(0, mi[m], mj[m], slot, Fw);
            visFace(1, mi[m], mj[m], slot, Fe_);
            visFace(2, mi[m], mj[m], slot, Fs);
            visFace(3, mi[m], mj[m], slot, Fn);
            for (int q = 0; q < 4; ++q) L[q] += -(Fe_[q] - Fw[q]) - (Fn[q] - Fs[q]);
            for (int q = 0; q < 3; ++q) LY[q] += -(Fe_[4 + q] - Fw[4 + q]) - (Fn[4 + q] - Fs[4 + q]);
          }
          // axisymmetric hoop stress (solver.js:587-588)
          const R r = dv(V, fmax(R(gg.y), R(1e-6))), vor = dv(vely, r);
          L[2] -= T2h[0] * (R(2) * vor - R(0.6666667f) * (T1h[0] + T1h[3] + vor)) * R(gg.y);
          // Werner-Wengle wall shear + Reynolds-analogy heat flux (solver.js:590-607)
          const R gx = R(gg.z), gr = R(gg.w);
          const R wl = sqrt(gx * gx + gr * gr);
          if (wl > R(1e-6)) {
            const R iwl = rc(wl);
            const R nx = gx * iwl, nr = gr * iwl;
            const R vn = velx * nx + vely * nr;
            const R utx = velx - vn * nx, utr = vely - vn * nr;
            const R up = sqrt(utx * utx + utr * utr);
            const R cp = Rm * (G + R(1));
            const R Taw = Tg + dv(R(0.89) * up * up, R(2) * cp);
            const R Ts = fmax(Tg + R(0.5) * (R(c.Tw) - Tg) + R(0.22) * (Taw - Tg), R(1e-3));
            const R rs = dv(p, Rm * Ts), mus = sutherland(Ts * R(c.gas[2]), c);
            const R dz = R(2) * clampR(R(0.5) * V * iwl, R(0.25), R(1));
            const R nzv = dv(mus, rs * dz);
            R tw = up <= R(69.7) * nzv ? dv(R(2) * mus * up, dz)
                                       : rs * pow(R(7.187) * pow(nzv, R(1.142857)) + R(0.13769) * pow(nzv, R(0.142857)) * up, R(1.75));
            const R iwdt = rc(fmax(wl * dt, R(1e-9)));
            tw = fmin(tw, R(0.3) * rho * up * V * iwdt);
            if (up > R(1e-6)) { const R f = dv(tw * wl, up); L[1] -= f * utx; L[2] -= f * utr; }
            const R qw = dv(tw, fmax(up, R(1e-4))) * cp * (Taw - R(c.Tw)) * R(1.24);
            const R qmax = R(0.3) * p * G * V * iwdt;
            L[3] -= clampR(qw, -qmax, qmax) * wl;
          }
          // Zukoski separation in the boundary-layer band (solver.js:613-630)
          if (code & C_BL) {
            const R up = sq(velx * velx + vely * vely), Mn = dv(up, sq(dv((R(1) + rc(G)) * p, rho)));
            if (Mn > R(1)) {
              const R iup = rc(up);
              const R shx = velx * iup, shy = vely * iup;
              R pd = R(0);
              for (int n = 4; n <= 6; n += 2)
                for (int m = 0; m <= 1; ++m) {
                  int Dx = ci + (int)round(R(n) * shx);
                  int Dy = cj + (int)floor(R(n) * shy) + m;
                  Dx = min(max(Dx, 0), Nx - 1);
                  Dy = min(max(Dy, 0), Nr - 1);
                  const long long kd = (long long)Dy * Nx + Dx;
                  if (!(I.code[kd] & C_SOLID)) pd = fmax(pd, load_prim<R>(U1, Ns, kd).p);
                }
              const R crit = R(1) + R(0.5) * Mn;
              const R fs = smoothstepR(crit, R(1.2) * crit, dv(pd, p)) * V * rho * up;
              L[1] -= fs * velx;
              L[2] -= fs * vely;
            } else if (velx > R(0) && p < R(1.3) * R(c.chemK[3])) {
              L[1] -= R(0.5) * V * rho * up * velx;
              L[2] -= R(0.5) * V * rho * up * vely;
            }
          }
        }
        const R invV = dv(dt, V);
        const R dO = invV * L[0];
        R Uq[4] = {A[0], A[1] + invV * L[1], A[2] + invV * L[2], A[3] + invV * L[3]};
        R UY[4] = {B[0] + invV * LY[0], B[1] + invV * LY[1], B[2] + dv(invV * LY[2], rho), B[3]};
        if (c.inj) UY[3] += dO;
        // afterburning / injector combustion (solver.js:639-654)
        if (c.chem) {
          const R Z = qS.Z, Fv = qS.F, O = qS.O, s = R(c.chemK[1]);
          const R avail = c.inj ? fmin(Fv, (O + R(0.233) * fmax(R(1) - Z, R(0))) / s)
                                : fmin(Fv, fmax(R(1) - Z - s * (Z - Fv), R(0)) / s);
          if (avail > R(1e-5)) {
            const R TK = Tg * R(c.gas[2]);
            const R kin = R(c.chemK[2]) * p * exp(R(-9000) / TK) * (R(1) - smoothstepR(R(c.cap[0]), R(c.cap[1]), TK));
            const R mixr = c.inj ? fmax(T2h[2], R(4) * T2h[0] / rho) : T2h[2];
            const R burn = avail * fmin(fmin(mixr, kin) * dt, R(0.5));
            UY[1] -= rho * burn;
            if (c.inj) UY[3] = fmax(UY[3] - rho * burn * s, R(0));
            Uq[3] += rho * burn * R(c.chemK[0]);
          }
        }
        const R rn = fmax(Uq[0], R(1e-7));
        const R irn = rc(rn);
        const R vnx = Uq[1] * irn, vny = Uq[2] * irn;
        if (!(rn < R(1e30)) || !(vnx < R(1e30) && vnx > R(-1e30))) {
          guard = true;
          for (int m = 0; m < 4; ++m) { A[m] = R(c.ambCons[m]); B[m] = R(c.ambB[m]); }
          for (int m = 0; m < nm; ++m) ubarM[m] = R(c.ambB[3]);
        } else {
          const R Gn = clampR(UY[2], R(c.gamR[0]), R(c.gamR[1]));
          const R Zn = clampR(UY[0], R(0), rn);
          A[0] = rn; A[1] = rn * vnx; A[2] = rn * vny;
          A[3] = fmax(Uq[3], R(1e-7) * Gn + R(0.5) * rn * (vnx * vnx + vny * vny));
          B[0] = Zn; B[1] = clampR(UY[1], R(0), Zn); B[2] = Gn;
          if (c.inj) {
            B[3] = clampR(UY[3], R(0), rn);
            for (int m = 0; m < nm; ++m) ubarM[m] = B[3];
          }
        }
      }
    }

    // ---- 5. spray gas sources (FS_SPRAY_APPLY, spray.js:158-178), consumed and cleared
    if (STAGE == 2 && c.spray && host) {
      float* __restrict__ src = I.src;
      const R s0 = R(src[k]), s1 = R(src[Ns + k]), s2 = R(src[2 * Ns + k]), s3 = R(src[3 * Ns + k]);
      const R dm = s0 + s1;
      if (!(dm == R(0) && s2 == R(0) && s3 == R(0))) {
        src[k] = 0.f; src[Ns + k] = 0.f; src[2 * Ns + k] = 0.f; src[3 * Ns + k] = 0.f;
        const R invV = rc(fmax(R(gg.x), R(1e-6)));
        const R iA0 = rc(fmax(A[0], R(1e-7)));
        const R velx = A[1] * iA0, vely = A[2] * iA0;
        const R rn = A[0] + dm * invV;
        const R mx = A[1] + s2 * invV, my = A[2] + s3 * invV;
        R E = A[3] + (s0 * R(c.eF) + s1 * R(c.eO) + (velx * s2 + vely * s3)) * invV;
        const R irn = rc(rn);
        const R phi = fmin(dm * invV * irn, R(1));
        const R Gn = mixR(B[2], R(c.Ge), phi);
        const R vnx = mx * irn, vny = my * irn;
        E = fmax(E, R(0.02) * A[3] + R(0.5) * rn * (vnx * vnx + vny * vny) + R(1e-7) * Gn);
        A[0] = rn; A[1] = mx; A[2] = my; A[3] = E;
        B[0] = B[0] + dm * invV;
        B[1] = B[1] + s0 * invV;
        B[2] = Gn;
        // every member takes the host's B (spray.js:161-177)
        const R bw = ubarM[0] + s1 * invV;
        for (int m = 0; m < nm; ++m) ubarM[m] = bw;
      }
    }

    // ---- 6. write the group, wave speed for the next step's CFL, residual
    R* __restrict__ Uout = STAGE == 1 ? U1 : U0;
    if (inDom && solid) {
      for (int q = 0; q < 4; ++q) { Uout[q * Ns + k] = R(c.ambCons[q]); Uout[(4 + q) * Ns + k] = R(c.ambB[q]); }
    }
    if (host) {
      for (int m = 0; m < nm; ++m) {
        const long long km = (long long)mj[m] * Nx + mi[m];
        for (int q = 0; q < 4; ++q) Uout[q * Ns + km] = A[q];
        for (int q = 0; q < 3; ++q) Uout[(4 + q) * Ns + km] = B[q];
        Uout[7 * Ns + km] = ubarM[m];
      }
      if (STAGE == 2) {
        const Prim<R> qn = to_prim<R>(A[0], A[1], A[2], A[3], B[0], B[1], B[2], ubarM[0]);
        wloc = (double)(fabs(qn.u) + fabs(qn.v) + R(2) * sound_speed(qn));
        // slaves are fluid cells holding the same state
        const double dr = (double)(A[0] - rho_old) / fmax((double)dt, 1e-30);
        rloc = dr * dr;
      }
    }
  }

  if (STAGE == 2) {
    // per-tile partials, reduced by k_finalize (no same-address atomics across workgroups)
    const double wb = block_max<NT>(wloc, sm.red);
    const double rb = block_sum<NT>(rloc, sm.redd);
    if (tid == 0) I.part[lt_] = make_double2(active ? wb : 0.0, active ? rb : 0.0);
  }
}

// Closes the step of every instance (one workgroup per instance): wave-speed max
// and residual over its tiles, step info of the completed step (FS_STEPINFO's
// input for the next one), spray ring head and seed (spray.js:295-296).
__global__ void __launch_bounds__(256) k_finalize(const DevInst* __restrict__ insts, int ninst) {
  __shared__ double red[16], redd[16];
  const int n = blockIdx.x;
  if (n >= ninst) return;
  const DevInst& J = insts[n];
  InstState* s = J.st;
  if (s->flags & 1u) return;
  const int nt = J.tilesX * J.tilesY;
  double w = 0.0, r = 0.0;
  for (int k = threadIdx.x; k < nt; k += 256) {
    const double2 p = J.part[k];
    w = fmax(w, p.x);
    r += p.y;
  }
  const double wb = block_max<256>(w, red);
  const double rb = block_sum<256>(r, redd);
  if (threadIdx.x == 0) {
    s->last[0] = s->cur[0]; s->last[1] = s->cur[1]; s->last[2] = s->cur[2]; s->last[3] = s->cur[3];
    s->wsnap = (unsigned long long)__double_as_longlong(wb);
    s->wacc = 0ull;
    s->steps += 1;
    s->resLast = sqrt(rb / fmax(s->nGroups, 1.0)) * (double)J.c.resScale;
    s->resAcc = 0.0;
    if (J.c.spray) {
      s->head = (s->head + 4u) % (uint32_t)J.c.NP;
      s->seed = s->seed + 1u;
    }
    // batched solves stop on simulated time or on the residual
    if (J.c.tEnd > 0.f && s->last[2] >= (double)J.c.tEnd) s->flags |= 1u | 4u;
    if (J.c.resTol > 0.f && s->steps >= J.c.minSteps && s->resLast <= (double)J.c.resTol) s->flags |= 1u | 2u;
    if (J.c.maxSteps > 0 && s->steps >= J.c.maxSteps) s->flags |= 1u | 8u;
    if (J.c.maxSteps > 0 && s->steps >= J.c.maxSteps) s->flags |= 1u | 8u;
  }
}

// ------------------------------------------------------------------------------
// Lagrangian parcels (FS_SPRAY_UPDATE, spray.js:40-121) with the point deposit
// (VS_SPRAY_DEP, spray.js:131-142) as float atomics into the host cell.
template <typename R>
__global__ void __launch_bounds__(256) k_spray(const DevInst* __restrict__ insts, int ninst) {
  // instance of this block: the last one whose parcel blocks start at or before it
  int lo = 0, hi = ninst - 1;
  while (lo < hi) {
    const int m = (lo + hi + 1) >> 1;
    if (insts[m].pblk0 <= (int)blockIdx.x) lo = m; else hi = m - 1;
  }
  const DevInst& I = insts[lo];
  if ((int)blockIdx.x >= I.pblk0 + I.npblk) return;
  const Consts& c = I.c;
  const InstState* st = I.st;
  if ((st->flags & 1u) || !c.spray) return;
  const int NP = c.NP;
  const int idx = ((int)blockIdx.x - I.pblk0) * 256 + threadIdx.x;
  if (idx >= NP) return;
  const int Nx = I.Nx, Nr = I.Nr;
  const long long Ns = I.Ns;
  const R* __restrict__ U1 = static_cast<const R*>(I.U1);
  float* __restrict__ pc = I.parc;
  const R dt = R(st->cur[0]);
  const uint32_t head = st->head, seed = st->seed;
  const R ax_ = pc[idx], ay = pc[NP + idx], az = pc[2 * NP + idx], aw = pc[3 * NP + idx];
  const R bx = pc[4 * NP + idx], by = pc[5 * NP + idx], bz = pc[6 * NP + idx], bw = pc[7 * NP + idx];
  int rel = idx - (int)head;
  if (rel < 0) rel += NP;
  R o[8] = {ax_, ay, az, aw, bx, by, bz, bw};
  R dmo = 0, dpx = 0, dpy = 0;
  const int perStep = 4;
  if (rel < perStep) {
    // new parcel; the liquid left in the replaced one goes to the gas
    const uint32_t id = (uint32_t)idx ^ (seed * 2654435761u);
    const int tt = (int)(R(rnd(id, 1u)) * R(255.999f));
    const float* e1 = I.tab + tt * 4;
    const float* e2 = I.tab + (256 + tt) * 4;
    const float* e3 = I.tab + (512 + tt) * 4;
    const R ang = R(e1[2]) + (R(rnd(id, 2u)) - R(0.5)) * R(2) * R(e2[2]);
    const R r = fmax(R(e1[1]) + (R(rnd(id, 3u)) - R(0.5)) * R(e2[3]), R(0.05f));
    const R d = R(e2[1]) * exp((R(rnd(id, 4u)) - R(0.5)) * R(1.4f));
    o[0] = e1[0]; o[1] = r; o[2] = R(e1[3]) * cos(ang); o[3] = R(e1[3]) * sin(ang);
    o[4] = d; o[5] = R(c.sprayMhat) * dt / R(perStep); o[6] = e2[0]; o[7] = e3[0];
    if (by > R(0)) { dmo = by; dpx = by * az; dpy = by * aw; }
  } else if (by > R(0)) {
    const int gi = min(max((int)floor(ax_), 0), Nx - 1), gj = min(max((int)floor(ay), 0), Nr - 1);
    const long long kg = (long long)gj * Nx + gi;
    if (I.code[kg] & C_SOLID) {
      // born inside solid hardware: walk out into the flow
      o[0] = ax_ + R(1); o[1] = fmax(ay - R(0.5), R(0.05f)); o[2] = fmax(az, R(0)); o[3] = aw;
    } else {
      const float* pr = bz < R(0.5) ? c.fuelVec : c.oxVec;
      const R pr0 = pr[0], pr1 = pr[1], pr2 = pr[2], pr3 = pr[3];
      R vx = az, vy = aw;
      const R m = by, d = bx, dts = dt * R(c.sprayTScale);
      R dn = d, mn = m;
      const Prim<R> w = load_prim<R>(U1, Ns, kg);
      {
        const R rm = (R(1) - w.Z) * R(c.gas[0]) + w.F * R(c.Rs[0]) + w.O * R(c.Rs[1]) + fmax(w.Z - w.F - w.O, R(0)) * R(c.Rs[2]);
        const R T = w.p / (w.r * rm) * R(c.gas[2]);
        const R rhoG = w.r * R(1.225f);
        const R rvx = w.u - vx, rvy = w.v - vy;
        const R Urel = sqrt(rvx * rvx + rvy * rvy) * R(287.6f) + R(1e-3f);
        const R mu = R(1.458e-6f) * T * sqrt(T) / (T + R(110.4f));
        const R Re = rhoG * Urel * d / mu;
        const R invTau = R(18) * mu * (R(1) + R(0.15f) * pow(fmin(Re, R(1000)), R(0.687f))) / (pr0 * d * d);
        const R ex = exp(-dts * invTau);
        vx = w.u + (vx - w.u) * ex;
        vy = w.v + (vy - w.v) * ex;
        const R cp = rm * R(287.05f) * (w.G + R(1));
        const R Bn = fmax(cp * (T - pr1), R(0)) / pr2;
        const R K = R(8) * (mu / R(0.7f)) / pr0 * log(R(1) + Bn) * (R(1) + R(0.27f) * sqrt(Re));
        const R d2 = d * d - K * dts;
        dn = d2 > R(1e-12f) ? sqrt(d2) : R(0);
        const R f = dn / d;
        mn = dn > R(0) ? m * f * f * f : R(0);
        const R we = rhoG * Urel * Urel * dn / pr3;
        if (we > R(12) && dn > R(0)) {
          const R dst = R(12) * pr3 / (rhoG * Urel * Urel);
          const R tb = R(3) * dn / Urel * sqrt(pr0 / rhoG);
          dn = mixR(dn, fmax(dst, R(2e-6f)), R(1) - exp(-dts / tb));
        }
      }
      R xi = bw;
      if (xi > R(0) && ax_ >= xi) {
        const uint32_t id = (uint32_t)idx ^ (seed * 2654435761u);
        const R sp = sqrt(vx * vx + vy * vy), phi = (R(rnd(id, 5u)) - R(0.5)) * R(2) * R(c.fan);
        vx = sp * cos(phi);
        vy = sp * sin(phi);
        dn *= R(0.25);
        xi = R(0);
      }
      R x = ax_ + vx * dt, r = ay + vy * dt;
      if (r < R(0)) { r = -r; vy = -vy; }
      if (x < R(0)) { x = -x; vx = fabs(vx) * R(0.3f); }
      auto wallAt = [&](R px, R pr_) {
        const int qi = (int)floor(px), qj = (int)floor(pr_);
        if (qi < 0 || qj < 0 || qi >= Nx || qj >= Nr) return false;
        return (I.code[(long long)qj * Nx + qi] & C_SOLID) != 0;
      };
      if (wallAt(x, r)) {
        const bool hx = wallAt(x, ay), hr = wallAt(ax_, r);
        if (hx) { vx *= R(-0.3f); x = ax_; }
        if (hr) { vy *= R(-0.3f); r = ay; }
        if (!hx && !hr) { x = ax_; r = ay; vx *= R(-0.3f); vy *= R(-0.3f); }
      }
      const R dm = m - mn;
      dmo = dm;
      dpx = mn * (az - vx) + dm * az;
      dpy = mn * (aw - vy) + dm * aw;
      if (x >= R(Nx) || r >= R(Nr) || mn <= R(0)) {
        o[0] = R(-1); o[1] = R(-1); o[2] = 0; o[3] = 0; o[4] = 0; o[5] = 0; o[6] = 0; o[7] = 0;
      } else {
        o[0] = x; o[1] = r; o[2] = vx; o[3] = vy; o[4] = dn; o[5] = mn; o[6] = bz; o[7] = xi;
      }
    }
  }
  for (int q = 0; q < 8; ++q) pc[q * NP + idx] = (float)o[q];
  // deposit into the host of the cell the parcel started in
  if (!(ax_ < R(0)) && !(dmo == R(0) && dpx == R(0) && dpy == R(0))) {
    int gi = min(max((int)floor(ax_), 0), Nx - 1), gj = min(max((int)floor(ay), 0), Nr - 1);
    const uint32_t cg = I.code[(long long)gj * Nx + gi];
    gi += code_hdx(cg);
    gj += code_hdy(cg);
    const long long h = (long long)gj * Nx + gi;
    float* __restrict__ src = I.src;
    unsafeAtomicAdd(&src[h], (float)(dmo * (R(1) - bz)));
    unsafeAtomicAdd(&src[Ns + h], (float)(dmo * bz));
    unsafeAtomicAdd(&src[2 * Ns + h], (float)dpx);
    unsafeAtomicAdd(&src[3 * Ns + h], (float)dpy);
  }
}

// ------------------------------------------------------------------------------
// wave-speed max of the current state (after setState / rescale): per-block partials,
// then one workgroup folds them into the instance's snapshot
template <typename R>
__global__ void __launch_bounds__(256) k_cfl_partial(DevInst I) {
  __shared__ double red[16];
  const long long N = (long long)I.Nx * I.Nr;
  const R* U = static_cast<const R*>(I.U0);
  double w = 0.0;
  for (long long k = (long long)blockIdx.x * 256 + threadIdx.x; k < N; k += (long long)gridDim.x * 256) {
    if (I.code[k] & C_SOLID) continue;
    const Prim<R> q = load_prim<R>(U, I.Ns, k);
    w = fmax(w, (double)(fabs(q.u) + fabs(q.v) + R(2) * sound_speed(q)));
  }
  const double wb = block_max<256>(w, red);
  if (threadIdx.x == 0) I.part[blockIdx.x] = make_double2(wb, 0.0);
}
__global__ void __launch_bounds__(256) k_cfl_reduce(DevInst I, int nblk) {
  __shared__ double red[16];
  double w = 0.0;
  for (int k = threadIdx.x; k < nblk; k += 256) w = fmax(w, I.part[k].x);
  const double wb = block_max<256>(w, red);
  if (threadIdx.x == 0) {
    I.st->wsnap = (unsigned long long)__double_as_longlong(wb);
    I.st->wacc = 0ull;
  }
}

// FS_RESCALE (solver.js:425-436), in place
template <typename R>
__global__ void k_rescale(DevInst I, float fp, float fT, float du) {
  const long long N = (long long)I.Nx * I.Nr;
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  R* U = static_cast<R*>(I.U0);
  const long long Ns = I.Ns;
  const Consts& c = I.c;
  if (I.code[k] & C_SOLID) {
    for (int q = 0; q < 4; ++q) { U[q * Ns + k] = R(c.ambCons[q]); U[(4 + q) * Ns + k] = R(c.ambB[q]); }
    return;
  }
  const R bw = U[7 * Ns + k];
  const Prim<R> q = load_prim<R>(U, Ns, k);
  const R w = R(1) - smoothstepR(R(0), R(0.03f), q.Z);
  const R rho = q.r * mixR(R(1), R(fp) / R(fT), w);
  const R p = q.p * mixR(R(1), R(fp), w);
  const R vx = q.u + R(du) * w, vy = q.v;
  U[k] = rho; U[Ns + k] = rho * vx; U[2 * Ns + k] = rho * vy; U[3 * Ns + k] = p * q.G + R(0.5) * rho * (vx * vx + vy * vy);
  U[4 * Ns + k] = rho * q.Z; U[5 * Ns + k] = rho * q.F; U[6 * Ns + k] = q.G; U[7 * Ns + k] = bw;
}

// FS_WALLROW (solver.js:668-673)
template <typename R>
__global__ void k_wall_row(DevInst I, const float* __restrict__ jw, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= I.Nx) return;
  const int j = min(max((int)jw[i * 4], 0), I.Nr - 1);
  const Prim<R> q = load_prim<R>(static_cast<const R*>(I.U0), I.Ns, (long long)j * I.Nx + i);
  out[i * 4] = (float)q.u; out[i * 4 + 1] = (float)q.v; out[i * 4 + 2] = (float)q.p; out[i * 4 + 3] = (float)sound_speed(q);
}

// T2 of the last physics pass, recomputed from U^(1) (the reference's vis.T2, render.js:426)
template <typename R>
__global__ void k_t2(DevInst I, float* __restrict__ out) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)I.Nx * I.Nr) return;
  const int i = (int)(k % I.Nx), j = (int)(k / I.Nx);
  const R* U1 = static_cast<const R*>(I.U1);
  R T1[4], T2[4];
  vcell_global<R>(I, U1, R(I.st->last[2]), R(I.st->last[0]), i, j, T1, T2);
  for (int q = 0; q < 4; ++q) out[k * 4 + q] = (float)T2[q];
}

// interleave a rectangle of planes p0..p0+3 into RGBA rows (the texture layout readPixels returns)
template <typename R>
__global__ void k_pack(const R* __restrict__ U, long long Ns, int Nx, int p0, int x, int y, int w, int h, float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= w * h) return;
  const int i = x + k % w, j = y + k / w;
  const long long c = (long long)j * Nx + i;
  for (int q = 0; q < 4; ++q) out[(long long)k * 4 + q] = (float)U[(p0 + q) * Ns + c];
}
// the last step info as four floats (readStepInfo)
__global__ void k_step_info(const InstState* __restrict__ st, float* __restrict__ out) {
  if (threadIdx.x < 4) out[threadIdx.x] = (float)st->last[threadIdx.x];
}
// interleave the full state in double (FP64 error analysis read-back)
template <typename R>
__global__ void k_pack_f64(const R* __restrict__ U, long long Ns, long long N, double* __restrict__ A, double* __restrict__ B) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  for (int q = 0; q < 4; ++q) { A[k * 4 + q] = (double)U[q * Ns + k]; B[k * 4 + q] = (double)U[(4 + q) * Ns + k]; }
}
// de-interleave RGBA A/B into the state planes (setState upload)
template <typename R>
__global__ void k_unpack(R* __restrict__ U, long long Ns, long long N, const float* __restrict__ A, const float* __restrict__ B) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  for (int q = 0; q < 4; ++q) { U[q * Ns + k] = R(A[k * 4 + q]); U[(4 + q) * Ns + k] = R(B[k * 4 + q]); }
}
// parcels all dead (GPUSpray.reset, spray.js:242-254) and sources cleared
__global__ void k_spray_reset(float* __restrict__ parc, int NP, float* __restrict__ src, long long Ns) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < NP) {
    parc[k] = -1.f; parc[NP + k] = -1.f;
    for (int q = 2; q < 8; ++q) parc[(long long)q * NP + k] = 0.f;
  }
  if (k < Ns)
    for (int q = 0; q < 4; ++q) src[q * Ns + k] = 0.f;
}

// FP64 integrals of the current state: exit plane, throat plane, chamber probe, totals
template <typename R>
__global__ void __launch_bounds__(256) k_integrals(DevInst I, int iExit, int jExit, int iThroat, int iChamber, double* __restrict__ out) {
  __shared__ double red[16];
  const int tid = threadIdx.x;
  const R* U = static_cast<const R*>(I.U0);
  const long long Ns = I.Ns;
  const int Nx = I.Nx, Nr = I.Nr;
  const double pa = (double)I.c.amb[3];
  double F = 0, md = 0, Mw = 0, pw = 0, Aa = 0, mt = 0, mass = 0, en = 0;
  for (int j = tid; j < jExit; j += 256) {
    const long long k = (long long)j * Nx + iExit;
    const double rho = fmax((double)U[k], 1e-9), u = (double)U[Ns + k] / rho, v = (double)U[2 * Ns + k] / rho;
    const double G = (double)U[6 * Ns + k], g = 1.0 + 1.0 / G;
    const double p = fmax(((double)U[3 * Ns + k] - 0.5 * rho * (u * u + v * v)) / G, 1e-9);
    const double dA = 2.0 * M_PI * fabs((double)I.ax[(long long)j * (Nx + 1) + iExit + 1]);
    const double cs = sqrt(g * p / rho);
    const double m = (double)U[4 * Ns + k] * u * dA;
    F += (rho * u * u + p - pa) * dA;
    md += m;
    Mw += m * sqrt(u * u + v * v) / cs;
    pw += p * dA;
    Aa += dA;
  }
  // throat plane: fluid cells from the axis up to the first solid one
  int jT = 0;
  while (jT < Nr && !(I.code[(long long)jT * Nx + iThroat] & C_SOLID)) ++jT;
  for (int j = tid; j < jT; j += 256) {
    const long long k = (long long)j * Nx + iThroat;
    const double rho = fmax((double)U[k], 1e-9), u = (double)U[Ns + k] / rho;
    mt += rho * u * 2.0 * M_PI * fabs((double)I.ax[(long long)j * (Nx + 1) + iThroat]);
  }
  for (long long k = tid; k < (long long)Nx * Nr; k += 256) {
    const uint32_t cd = I.code[k];
    if ((cd & C_SOLID) || (cd & C_IS_SLAVE)) continue;
    const double V = (double)I.gg[k].x;
    mass += (double)U[k] * V;
    en += (double)U[3 * Ns + k] * V;
  }
  double vals[8] = {F, md, Mw, pw, Aa, mt, mass, en};
  for (int q = 0; q < 8; ++q) {
    const double s = block_sum<256>(vals[q], red);
    __syncthreads();
    if (tid == 0) out[q] = s;
  }
  if (tid == 0) {
    const long long k = iChamber;
    const double rho = fmax((double)U[k], 1e-9), u = (double)U[Ns + k] / rho, v = (double)U[2 * Ns + k] / rho;
    const double G = (double)U[6 * Ns + k];
    const double p = fmax(((double)U[3 * Ns + k] - 0.5 * rho * (u * u + v * v)) / G, 1e-9);
    const double Y = fmin(fmax((double)U[4 * Ns + k] / rho, 0.0), 1.0);
    const double Fv = fmax((double)U[5 * Ns + k] / rho, 0.0);
    const double Ov = I.c.inj ? fmax((double)U[7 * Ns + k] / rho, 0.0) : 0.0;
    const double Rm = (1 - Y) * 1.0 + Fv * I.c.Rs[0] + Ov * I.c.Rs[1] + fmax(Y - Fv - Ov, 0.0) * I.c.Rs[2];
    out[8] = p;
    out[9] = p / (rho * Rm);
    out[10] = I.st->resLast;
  }
}

}  // namespace nzk

// ==============================================================================
// host-side launchers (the runtime in runtime.hip calls only these)
#include "nz_launch.h"

namespace nzk {

__global__ void k_set_inst(DevInst* dst, DevInst src) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = src;
}

void tile_shape(bool fp64, int* tx, int* ty) {
  *tx = fp64 ? TX64 : TX32;
  *ty = fp64 ? TY64 : TY32;
}

hipError_t launch_set_inst(DevInst* dst, const DevInst& src, hipStream_t s) {
  hipLaunchKernelGGL(k_set_inst, dim3(1), dim3(64), 0, s, dst, src);
  return hipGetLastError();
}

template <typename R, int TX, int TY>
static hipError_t stage_t(int stage, bool phys, const DevInst* d, int ninst, int ntiles, hipStream_t s) {
  const dim3 g(ntiles), b(TX * TY);
  if (stage == 1) hipLaunchKernelGGL((k_stage<R, TX, TY, 1, false>), g, b, 0, s, d, ninst);
  else if (phys) hipLaunchKernelGGL((k_stage<R, TX, TY, 2, true>), g, b, 0, s, d, ninst);
  else hipLaunchKernelGGL((k_stage<R, TX, TY, 2, false>), g, b, 0, s, d, ninst);
  return hipGetLastError();
}
hipError_t launch_stage(int stage, bool phys, bool fp64, const DevInst* d, int ninst, int ntiles, hipStream_t s) {
  return fp64 ? stage_t<double, TX64, TY64>(stage, phys, d, ninst, ntiles, s)
              : stage_t<float, TX32, TY32>(stage, phys, d, ninst, ntiles, s);
}
hipError_t launch_finalize(const DevInst* d, int ninst, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3(ninst), dim3(256), 0, s, d, ninst);
  return hipGetLastError();
}
hipError_t launch_spray(bool fp64, const DevInst* d, int ninst, int nblk, hipStream_t s) {
  if (fp64) hipLaunchKernelGGL((k_spray<double>), dim3(nblk), dim3(256), 0, s, d, ninst);
  else hipLaunchKernelGGL((k_spray<float>), dim3(nblk), dim3(256), 0, s, d, ninst);
  return hipGetLastError();
}
hipError_t launch_cfl_init(bool fp64, const DevInst& I, hipStream_t s) {
  // as many blocks as the instance has tiles (its partials array holds one per tile)
  const long long N = (long long)I.Nx * I.Nr;
  const int nblk = (int)std::min<long long>(std::min<long long>((N + 255) / 256, 1024), (long long)I.tilesX * I.tilesY);
  if (fp64) hipLaunchKernelGGL((k_cfl_partial<double>), dim3(nblk), dim3(256), 0, s, I);
  else hipLaunchKernelGGL((k_cfl_partial<float>), dim3(nblk), dim3(256), 0, s, I);
  hipLaunchKernelGGL(k_cfl_reduce, dim3(1), dim3(256), 0, s, I, nblk);
  return hipGetLastError();
}
hipError_t launch_rescale(bool fp64, const DevInst& I, float fp, float fT, float du, hipStream_t s) {
  const long long N = (long long)I.Nx * I.Nr;
  const dim3 g((unsigned)((N + 255) / 256)), b(256);
  if (fp64) hipLaunchKernelGGL((k_rescale<double>), g, b, 0, s, I, fp, fT, du);
  else hipLaunchKernelGGL((k_rescale<float>), g, b, 0, s, I, fp, fT, du);
  return hipGetLastError();
}
hipError_t launch_wall_row(bool fp64, const DevInst& I, const float* jw, float* out, hipStream_t s) {
  const dim3 g((I.Nx + 255) / 256), b(256);
  if (fp64) hipLaunchKernelGGL((k_wall_row<double>), g, b, 0, s, I, jw, out);
  else hipLaunchKernelGGL((k_wall_row<float>), g, b, 0, s, I, jw, out);
  return hipGetLastError();
}
hipError_t launch_t2(bool fp64, const DevInst& I, float* out, hipStream_t s) {
  const long long N = (long long)I.Nx * I.Nr;
  const dim3 g((unsigned)((N + 255) / 256)), b(256);
  if (fp64) hipLaunchKernelGGL((k_t2<double>), g, b, 0, s, I, out);
  else hipLaunchKernelGGL((k_t2<float>), g, b, 0, s, I, out);
  return hipGetLastError();
}
hipError_t launch_pack(bool fp64, const void* U, long long Ns, int Nx, int p0, int x, int y, int w, int h, float* out, hipStream_t s) {
  const dim3 g((unsigned)((w * h + 255) / 256)), b(256);
  if (fp64) hipLaunchKernelGGL((k_pack<double>), g, b, 0, s, static_cast<const double*>(U), Ns, Nx, p0, x, y, w, h, out);
  else hipLaunchKernelGGL((k_pack<float>), g, b, 0, s, static_cast<const float*>(U), Ns, Nx, p0, x, y, w, h, out);
  return hipGetLastError();
}
hipError_t launch_unpack(bool fp64, void* U, long long Ns, long long N, const float* A, const float* B, hipStream_t s) {
  const dim3 g((unsigned)((N + 255) / 256)), b(256);
  if (fp64) hipLaunchKernelGGL((k_unpack<double>), g, b, 0, s, static_cast<double*>(U), Ns, N, A, B);
  else hipLaunchKernelGGL((k_unpack<float>), g, b, 0, s, static_cast<float*>(U), Ns, N, A, B);
  return hipGetLastError();
}
hipError_t launch_step_info(const InstState* st, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_step_info, dim3(1), dim3(64), 0, s, st, out);
  return hipGetLastError();
}
hipError_t launch_pack_f64(bool fp64, const void* U, long long Ns, long long N, double* A, double* B, hipStream_t s) {
  const dim3 g((unsigned)((N + 255) / 256)), b(256);
  if (fp64) hipLaunchKernelGGL((k_pack_f64<double>), g, b, 0, s, static_cast<const double*>(U), Ns, N, A, B);
  else hipLaunchKernelGGL((k_pack_f64<float>), g, b, 0, s, static_cast<const float*>(U), Ns, N, A, B);
  return hipGetLastError();
}
hipError_t launch_spray_reset(float* parc, int NP, float* src, long long Ns, hipStream_t s) {
  const long long n = NP > Ns ? NP : Ns;
  hipLaunchKernelGGL(k_spray_reset, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, parc, NP, src, Ns);
  return hipGetLastError();
}
hipError_t launch_integrals(bool fp64, const DevInst& I, int iExit, int jExit, int iThroat, int iChamber, double* out, hipStream_t s) {
  if (fp64) hipLaunchKernelGGL((k_integrals<double>), dim3(1), dim3(256), 0, s, I, iExit, jExit, iThroat, iChamber, out);
  else hipLaunchKernelGGL((k_integrals<float>), dim3(1), dim3(256), 0, s, I, iExit, jExit, iThroat, iChamber, out);
  return hipGetLastError();
}

}  // namespace nzk
