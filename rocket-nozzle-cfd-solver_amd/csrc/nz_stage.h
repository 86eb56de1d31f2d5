// nz_stage.h — the fused stage kernel (k_stage) and the device helpers it shares
// with the other kernels. Included by kernels.hip, and by stage_inst.hip, which is
// compiled once per tile shape listed in stage_shapes.h.
//
// k_stage<R, TX, TY, STAGE, PHYS>: one workgroup of TX*TY threads per TX x TY tile.
//   LDS tile of primitives with a 2-cell ghost halo (FS_PRIM) -> each x/r face flux
//   computed once (MUSCL + HLLC/HLL, FS_FLUX) -> merged-group finite-volume update
//   (FS_UPDATE). STAGE 1 gives U1 = U0 + dt L(U0); STAGE 2 gives
//   U^{n+1} = (U0 + U1 + dt L(U1)) / 2 in place over U0, fused with the viscous /
//   wall / afterburning split step (FS_VCELL, FS_VFLUX, FS_SOURCE), the spray source
//   apply (FS_SPRAY_APPLY) and the per-tile CFL wave-speed and residual partials.
#pragma once
#include "nz_device.h"

namespace nzk {

// ------------------------------------------------------------------------------
// Tile of a workgroup. The dispatcher deals workgroups round-robin over the 8 XCDs
// (workgroup b runs on XCD b % 8), and each XCD has its own L2. The tiles are cut into
// chunks of `chunk` consecutive tiles (whole tile rows), dealt round-robin to the XCDs:
// an XCD's workgroups take its chunks in order, so the halos shared inside a chunk are
// fetched into one L2 once, while the costly tiles (walls, merged cells) of one region
// of the nozzle spread over several XCDs. Tiles past the last whole round of chunks go
// to the XCDs in contiguous runs. Bijective on [0, nblk).
#ifndef NZ_XCD_CHUNK_ROWS
#define NZ_XCD_CHUNK_ROWS 2
#endif
__device__ __forceinline__ int xcd_tile(int b, int nblk, int chunk) {
  constexpr int NXCD = 8;
  const int x = b % NXCD, l = b / NXCD;
  const int per = (nblk / (NXCD * chunk)) * chunk;  // chunked tiles per XCD
  if (l < per) return ((l / chunk) * NXCD + x) * chunk + l % chunk;
  const int rest = nblk - per * NXCD, q = rest / NXCD, r = rest % NXCD;
  return per * NXCD + x * q + min(x, r) + (l - per);
}

// FS_STEPINFO (solver.js:404-409): dt from the wave-speed max of the current state,
// in the arithmetic type of the state
template <typename R>
__device__ __forceinline__ void step_info(const Consts& c, double last0, double last2, double wmax, R* cur) {
  const R w = fmax(R(c.wFloor), R(wmax));
  R dt = fmin(R(c.cfl) / w, R(c.dtMax));
  const R lastDt = R(last0), lastT = R(last2);
  if (lastDt > R(0)) dt = fmin(dt, lastDt * R(1.2f));
  cur[0] = dt;
  cur[1] = lastT;
  cur[2] = lastT + dt;
  cur[3] = w;
}

// A double moved across the lanes of a wave by the DPP control CTRL (both 32-bit halves);
// lanes of rows outside ROW_MASK take `id`, the identity of the reduction step.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double x, double id) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(id), __double2loint(x), CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(id), __double2hiint(x), CTRL, ROW_MASK, 0xF, false);
  return __hiloint2double(hi, lo);
}
// The value of lane L - 1 of the wave (DPP wave_shr:1; lane 0 takes 0). Every lane of the
// wave must be active.
__device__ __forceinline__ float lane_prev(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ double lane_prev(double x) { return dpp_move<0x138, 0xF>(x, 0.0); }
// Max of w (>= 0) and sum of r over the 64 lanes of a wave, left in every lane, in one fixed
// order: swaps within quads (lane ^ 1, lane ^ 2), the half-row and row mirrors (each row of
// 16 lanes then holds its row's result in every lane), row 0 into row 1 and row 2 into
// row 3 (row_bcast15), rows 0-1 into rows 2-3 (row_bcast31); lane 63 then holds the total,
// which every lane reads. Register-to-register moves, where a butterfly of lane shuffles
// is a chain of LDS round trips. Every lane of the wave must be active.
__device__ __forceinline__ void wave_max_sum(double& w, double& r) {
  w = fmax(w, dpp_move<0xB1, 0xF>(w, 0.0)); r += dpp_move<0xB1, 0xF>(r, 0.0);
  w = fmax(w, dpp_move<0x4E, 0xF>(w, 0.0)); r += dpp_move<0x4E, 0xF>(r, 0.0);
  w = fmax(w, dpp_move<0x141, 0xF>(w, 0.0)); r += dpp_move<0x141, 0xF>(r, 0.0);
  w = fmax(w, dpp_move<0x140, 0xF>(w, 0.0)); r += dpp_move<0x140, 0xF>(r, 0.0);
  w = fmax(w, dpp_move<0x142, 0xA>(w, 0.0)); r += dpp_move<0x142, 0xA>(r, 0.0);
  w = fmax(w, dpp_move<0x143, 0xC>(w, 0.0)); r += dpp_move<0x143, 0xC>(r, 0.0);
  w = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(w), 63), __builtin_amdgcn_readlane(__double2loint(w), 63));
  r = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(r), 63), __builtin_amdgcn_readlane(__double2loint(r), 63));
}

// Wave-speed max and residual sum of an instance's tile partials, in one fixed order for
// every caller: 256 virtual lanes v take partials v, v + 256, ... in turn; each group of
// 64 lanes folds by wave_max_sum (which leaves the same bits in all its lanes); the four
// group results combine as max and (R0 + R1) + (R2 + R3). A workgroup of NT threads (NT
// divides 256) runs virtual lanes tid, tid + NT, ...; the loads of four consecutive
// partials of a lane are issued together. fold_groups leaves the group results in W[4],
// R[4] (LDS), which fold_combine reads after a barrier.
// The fold in two parts, so that its first loads can be issued with a tile's own loads and
// consumed later: fold_issue loads each virtual lane's first four partials, fold_finish folds
// them, loads and folds the rest (more than 1024 partials), and leaves the group results.
template <int NT>
struct FoldLoads {
  double2 p[256 / NT][4];
};
template <int NT>
__device__ __forceinline__ void fold_issue(const double2* __restrict__ part, uint32_t n, int tid, FoldLoads<NT>& f) {
  static_assert(256 % NT == 0 && NT % 64 == 0, "fold: the workgroup size must divide 256");
#pragma unroll
  for (int m = 0; m < 256 / NT; ++m) {
    const uint32_t v = (uint32_t)(tid + m * NT);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      // every lane loads, from an index clamped into the n partials; fold_finish drops the
      // values of indices past n (selecting here would wait for the loads here)
      f.p[m][j] = gld(part, min(v + 256u * j, n - 1u));
    }
  }
}
template <int NT>
__device__ __forceinline__ void fold_finish(const double2* __restrict__ part, uint32_t n, int tid, const FoldLoads<NT>& f,
                                            double* W, double* R) {
#pragma unroll
  for (int m = 0; m < 256 / NT; ++m) {
    const uint32_t v = (uint32_t)(tid + m * NT);
    double w = 0.0, r = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool in = v + 256u * j < n;
      w = fmax(w, in ? f.p[m][j].x : 0.0);
      r += in ? f.p[m][j].y : 0.0;
    }
    for (uint32_t k0 = v + 1024; k0 < n; k0 += 1024) {
      double2 p[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t k = k0 + 256u * j;
        p[j] = k < n ? part[k] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        w = fmax(w, p[j].x);
        r += p[j].y;
      }
    }
    wave_max_sum(w, r);
    if ((v & 63u) == 0u) { W[v >> 6] = w; R[v >> 6] = r; }
  }
}
template <int NT>
__device__ __forceinline__ void fold_groups(const double2* __restrict__ part, uint32_t n, int tid, double* W, double* R) {
  FoldLoads<NT> f;
  fold_issue<NT>(part, n, tid, f);
  fold_finish<NT>(part, n, tid, f, W, R);
}
__device__ __forceinline__ void fold_combine(const double* W, const double* R, double& w, double& r) {
  w = fmax(fmax(W[0], W[1]), fmax(W[2], W[3]));
  r = (R[0] + R[1]) + (R[2] + R[3]);
}

// Stop tests of batched solves after the last completed step: simulated time, residual
// (after a minimum step count) and step count; 0 while the solve goes on
__device__ __forceinline__ uint32_t stop_bits(const Consts& c, double last2, long long steps, double resLast) {
  uint32_t f = 0u;
  if (c.tEnd > 0.f && last2 >= (double)c.tEnd) f |= 1u | 4u;
  if (c.resTol > 0.f && steps > 0 && steps >= c.minSteps && resLast <= (double)c.resTol) f |= 1u | 2u;
  if (c.maxSteps > 0 && steps >= c.maxSteps) f |= 1u | 8u;
  return f;
}
__device__ __forceinline__ uint32_t stop_bits(const Consts& c, const InstState& s, double resLast) {
  return stop_bits(c, s.last[2], s.steps, resLast);
}

// The step-state fields every workgroup of a stage launch reads, loaded together at its
// start as global loads. The other stage writes them, so no workgroup of this launch
// changes them. Stage 1: flagsSeen, last, steps, wsnap; stage 2: flags, cur.
struct StepView {
  uint32_t flags;
  double last0, last2;
  long long steps;
  unsigned long long wsnap;
  double cur[4];
};
template <int STAGE>
__device__ __forceinline__ StepView load_step_view(const InstState* st) {
  const NZ_GLOBAL InstState* g = (const NZ_GLOBAL InstState*)st;
  StepView v;
  if (STAGE == 1) {
    v.flags = g->flagsSeen;
    v.last0 = g->last[0];
    v.last2 = g->last[2];
    v.steps = g->steps;
    v.wsnap = g->wsnap;
    v.cur[0] = v.cur[1] = v.cur[2] = v.cur[3] = 0.0;
  } else {
    v.flags = g->flags;
    v.last0 = v.last2 = 0.0;
    v.steps = 0;
    v.wsnap = 0ull;
    for (int q = 0; q < 4; ++q) v.cur[q] = g->cur[q];
  }
  return v;
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
// max of w (>= 0) and sum of r over the workgroup in one pass (one barrier); valid on thread 0
template <int NT>
__device__ __forceinline__ void block_max_sum(double& w, double& r, double* redW, double* redR) {
  wave_max_sum(w, r);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { redW[wv] = w; redR[wv] = r; }
  __syncthreads();
  if (threadIdx.x == 0) {
    w = 0.0;
    r = 0.0;
    for (int k = 0; k < NT / 64; ++k) { w = fmax(w, redW[k]); r += redR[k]; }
  }
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
// global fall-back for faces of merged cells that lie outside it; the signed area from
// `getA(dirR, i, j)` (LDS for the tile's own faces, else global memory). DIR 0 / 1 fixes
// the direction at compile time (x / r faces; the `dirR` argument is then ignored), so
// the cell addresses are constant offsets and the velocity swap costs nothing; DIR -1
// takes it from `dirR`.
template <typename R, int DIR = -1, class GetP, class GetA>
__device__ __forceinline__ void face_flux_w(const DevInst& I, bool dirR_, int fi, int fj, GetP getP, GetA getA, R* F,
                                            R* FY, R& FO) {
  const bool dirR = DIR < 0 ? dirR_ : DIR == 1;
  if constexpr (DIR >= 0) {
    // a face of the LDS tile: every cell it reads is in the tile and its area is 0
    // outside the domain, so the flux is formed unconditionally and then kept or set
    // to zero (no divergent early exits; a dropped face's arithmetic may be non-finite
    // and is discarded by the selects)
    constexpr int dx = DIR == 1 ? 0 : 1, dy = DIR == 1 ? 1 : 0, tx = dy, ty = dx;
    const float area = getA(dirR, fi, fj);
    Prim<R> c0 = getP(fi - 2 * dx, fj - 2 * dy), c1 = getP(fi - dx, fj - dy);
    Prim<R> c2 = getP(fi, fj), c3 = getP(fi + dx, fj + dy);
    const bool s1 = c1.r < R(0), s2 = c2.r < R(0);
    const bool keep = area != 0.f && !(DIR == 1 && fj == 0) && !(s1 && s2);
    R hllw = R(0);
    if (I.c.hllBlend > 0.f) {
      const Prim<R> a1 = getP(fi - dx + tx, fj - dy + ty), b1 = getP(fi - dx - tx, fj - dy - ty);
      const Prim<R> a2 = getP(fi + tx, fj + ty), b2 = getP(fi - tx, fj - ty);
      const R w = hll_weight<R>(tang_jump(c1, a1, b1), tang_jump(c2, a2, b2), R(I.c.hllBlend));
      hllw = (s1 || s2) ? R(0) : w;
    }
    if (DIR == 1) {
      R t;
      t = c0.u; c0.u = c0.v; c0.v = t;
      t = c1.u; c1.u = c1.v; c1.v = t;
      t = c2.u; c2.u = c2.v; c2.v = t;
      t = c3.u; c3.u = c3.v; c3.v = t;
    }
    face_flux<R>(c0, c1, c2, c3, hllw, I.c.order2 && area > 0.f, F, FY, FO);
    if (DIR == 1) { const R t = F[1]; F[1] = F[2]; F[2] = t; }
    const R a = R(fabsf(area));
    for (int q = 0; q < 4; ++q) { F[q] = keep ? F[q] * a : R(0); FY[q] = keep ? FY[q] * a : R(0); }
    FO = keep ? FO * a : R(0);
    return;
  }
  for (int q = 0; q < 4; ++q) { F[q] = R(0); FY[q] = R(0); }
  FO = R(0);
  const int Nx = I.Nx, Nr = I.Nr;
  const bool inDom = dirR ? (fi >= 0 && fi < Nx && fj >= 0 && fj <= Nr) : (fi >= 0 && fi <= Nx && fj >= 0 && fj < Nr);
  if (!inDom) return;
  const float area = getA(dirR, fi, fj);
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
    hllw = hll_weight<R>(tang_jump(c1, a1, b1), tang_jump(c2, a2, b2), R(I.c.hllBlend));
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

// the primitives with the velocity components exchanged (the r-face-normal frame)
template <typename R>
__device__ __forceinline__ Prim<R> swap_uv(Prim<R> q) {
  const R t = q.u; q.u = q.v; q.v = t;
  return q;
}

// One tile face's inviscid flux times |area| from its two cells' primitives c1, c2, MUSCL
// values vL (c1's value ahead) and vR (c2's value behind) and tangential jump ratios j1, j2
// (formed once per cell, in the face-normal frame of direction DIR): the same flux, bit for
// bit, as face_flux_w forms from the 4-cell stencil, and dropped likewise.
template <typename R, int DIR>
__device__ __forceinline__ void face_flux_pc(const DevInst& I, int fj, const Prim<R>& c1, const Prim<R>& c2, const R* vL,
                                             const R* vR, R j1, R j2, float area, R* F, R* FY, R& FO) {
  const bool s1 = c1.r < R(0), s2 = c2.r < R(0);
  const bool keep = area != 0.f && !(DIR == 1 && fj == 0) && !(s1 && s2);
  const R hllw = (I.c.hllBlend > 0.f && !(s1 || s2)) ? hll_weight<R>(j1, j2, R(I.c.hllBlend)) : R(0);
  Prim<R> WL, WR;
  face_states(c1, c2, vL, vR, I.c.order2 && area > 0.f, WL, WR);
  riemann_flux(WL, WR, hllw, F, FY, FO);
  if (DIR == 1) { const R t = F[1]; F[1] = F[2]; F[2] = t; }
  const R a = R(fabsf(area));
  for (int q = 0; q < 4; ++q) { F[q] = keep ? F[q] * a : R(0); FY[q] = keep ? FY[q] * a : R(0); }
  FO = keep ? FO * a : R(0);
}

// One face's viscous flux times |area| (FS_VFLUX_X/R); getT(i, j, T1, T2) gives the
// per-cell gradients / viscosities of an interior fluid cell.
template <typename R, int DIR = -1, class GetP, class GetT, class GetA>
__device__ __forceinline__ void vface_w(const DevInst& I, bool dirR_, int fi, int fj, GetP getP, GetT getT, GetA getA, R* F,
                                        R* FY) {
  const bool dirR = DIR < 0 ? dirR_ : DIR == 1;  // as in face_flux_w
  for (int q = 0; q < 4; ++q) F[q] = R(0);
  for (int q = 0; q < 3; ++q) FY[q] = R(0);
  const int Nx = I.Nx, Nr = I.Nr;
  const bool on = dirR ? (fi >= 0 && fi < Nx && fj > 0 && fj < Nr) : (fi > 0 && fi < Nx && fj >= 0 && fj < Nr);
  if (!on) return;
  const float area = getA(dirR, fi, fj);
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
  vcell<R>(w0, getP(i + 1, j), getP(i - 1, j), getP(i, j + 1), getP(i, j - 1), i, j, R(gld(I.lt, (unsigned)(j * I.Nx + i))), dt,
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

// signed r-weighted face area from global memory (face inside the domain)
__device__ __forceinline__ float face_area(const DevInst& I, bool dirR, int fi, int fj) {
  return dirR ? gld(I.ar, (unsigned)(fj * I.Nx + fi)) : gld(I.ax, (unsigned)(fj * (I.Nx + 1) + fi));
}

// ------------------------------------------------------------------------------
template <typename R, int TX, int TY>
struct StageInvSmem {
  R Fx[9][TY][TX + 1];
  union {
    R Fr[9][TY + 1][TX];
    // per-cell face states, until the r faces are formed: each cell's r-direction MUSCL
    // value at its north face and x-direction tangential jump ratio, for the south face of
    // the cell above it (none for the top row, whose north faces the boundary pass forms)
    struct {
      R TS[8][TY - 1][TX];
      R JS[TY - 1][TX];
    } pc;
  };
};
template <typename R, int TX, int TY>
struct StageVisSmem {
  R T[8][TY + 2][TX + 2];
  R Vx[7][TY][TX + 1];
  R Vr[7][TY + 1][TX];
};
// the viscous arrays share the inviscid flux arrays' LDS, and exist only in the kernels
// that run the split step
template <typename R, int TX, int TY, bool VIS>
union StageFluxSmem {
  StageInvSmem<R, TX, TY> inv;
};
template <typename R, int TX, int TY>
union StageFluxSmem<R, TX, TY, true> {
  StageInvSmem<R, TX, TY> inv;
  StageVisSmem<R, TX, TY> vis;
};

template <typename R, int TX, int TY, bool VIS>
struct StageSmem {
  static constexpr int PW = TX + 4, PH = TY + 4;   // primitive tile with 2-cell halo
  static constexpr int NFX = (TX + 1) * TY, NFR = TX * (TY + 1);
  static constexpr int TW = TX + 2, THh = TY + 2;  // T1/T2 tile with 1-cell ring
  // the halo's cells outside the tile: rows 0, 1, then rows PH - 2, PH - 1 (each left to
  // right), then columns 0, 1, PW - 2, PW - 1 of rows 2 .. PH - 3, row by row
  static constexpr int NRING = PW * PH - TX * TY;
  static __device__ __forceinline__ int2 ring_cell(int q) {
    if (q < 2 * PW) return make_int2(q % PW, q / PW);
    q -= 2 * PW;
    if (q < 2 * PW) return make_int2(q % PW, PH - 2 + q / PW);
    q -= 2 * PW;
    const int c = q & 3;
    return make_int2(c < 2 ? c : PW - 4 + c, 2 + (q >> 2));
  }
  R P[8][PH][PW];
  float AX[TY][TX + 1];                            // signed areas of the tile's x faces
  float AR[TY + 1][TX];                            // ... and r faces (0 outside the domain)
  StageFluxSmem<R, TX, TY, VIS> u;
  static constexpr int MAXE = MAX_EXTRA_FACES;     // faces of merged cells outside the tile
  static constexpr int MAXF = MAX_FAR_CELLS;       // cells those faces reach beyond the halo
  static constexpr int WW = TX + 2 * FAR_HALO, WH = TY + 2 * FAR_HALO;
  uint2 ex[MAXE];
  int2 fc[MAXF];
  R PF[8][MAXF];
  unsigned char fmap[WH][WW];                      // far-cell slot by window position
  int nEx, nF;
  double foldW[4], foldR[4];                       // group results of the last step's partials (stage 1)
  double red[TX * TY / 64];                        // one slot per wave of the block reductions
  double redd[TX * TY / 64];
};

// W: minimum waves per SIMD the compiler must fit for stage 1 and the stage 2 kernel
// without the split step (1: no constraint); the viscous stage 2 is never constrained
// Launch arguments: the instance table, its length, the XCD chunk (tiles) and, for more
// than one instance, the instance of every launch tile. Each comes from the kernel
// arguments (preloaded into SGPRs), so the first memory access of a workgroup is its
// instance's descriptor, or the one table entry that names it.
template <typename R, int TX, int TY, int STAGE, bool PHYS, int W>
__global__ void __launch_bounds__(TX* TY, (STAGE == 2 && PHYS) ? 1 : W)
    k_stage(const DevInst* __restrict__ insts, int ninst, int chunk, const unsigned short* __restrict__ tileInst) {
  constexpr int NT = TX * TY;
  // per-cell face states (shapes whose tile rows fit a wave): each cell's MUSCL values and
  // tangential jump ratios are formed once, by its own thread, and the tile's west column
  // and south row of faces move to the boundary pass. The throughput shapes pack that pass
  // into as few waves as possible: there the moved faces are free only while all four
  // sides still fit one wave (2 (TX + TY) <= 64); otherwise they cost a wave more than the
  // per-cell values save, and those shapes keep the 4-cell stencil per face.
  constexpr bool PC = TX <= 64 && (W < 4 || 2 * (TX + TY) <= 64);
  using S = StageSmem<R, TX, TY, STAGE == 2 && PHYS>;
  __shared__ S sm;
  const int tid = threadIdx.x;
  const int t = xcd_tile(blockIdx.x, gridDim.x, chunk);
  const int ii = ninst == 1 ? 0 : (int)tileInst[t];
  touch_desc(insts + ii);
  const DevInst& I = insts[ii];
  const Consts& c = I.c;
  InstState* st = I.st;
  const int Nx = I.Nx, Nr = I.Nr;
  const long long Ns = I.Ns;
  const int lt_ = t - I.tile0;
  const int i0 = (lt_ % I.tilesX) * TX, j0 = (lt_ / I.tilesX) * TY;
  R* __restrict__ U0 = static_cast<R*>(I.U0);
  R* __restrict__ U1 = static_cast<R*>(I.U1);
  const R* __restrict__ Uin = STAGE == 1 ? U0 : U1;

  // Step state. Stage 1 closes the last step and opens this one: the stop tests of the
  // last completed step (simulated time and step count here; with the residual once the
  // partials are folded), then FS_STEPINFO (solver.js:404-409) from the wave-speed max
  // of the partials the last stage 2 wrote, folded here by wave 0 of every workgroup
  // (`fused`) or by the close kernel that followed it. Stage 2 takes the time step stage
  // 1 wrote; its first workgroup files the step as completed. Each field is written in a
  // launch in which no workgroup reads it: stage 1's first workgroup sets the stop bits in
  // `flags`, which stage 2 reads, while stage 1 reads `flagsSeen`, the copy that stage 2's
  // first workgroup (or the close kernel) writes.
  bool was = false, active = false;
  R cur[4] = {R(0), R(0), R(0), R(0)};
  R dt = R(0), tsel = R(0);
  // the last step's partials: one per tile of this tiling (see FUSED_CLOSE_MAX)
  const uint32_t nPart = (uint32_t)(I.tilesX * I.tilesY);
  const bool fused = nPart <= FUSED_CLOSE_MAX;
  StepView sv;
  auto step_state = [&]() {
    was = !(sv.flags & 1u);
    active = was;
    if (STAGE == 1) {
      tsel = R(sv.last2);
      if (was) {
        const uint32_t f = stop_bits(c, sv.last2, sv.steps, INFINITY);
        if (f) {
          active = false;
          if (lt_ == 0) {
            // the solve stops here; its last step's residual still decides the stop reason
            if (fused) fold_groups<NT>(I.part, nPart, tid, sm.foldW, sm.foldR);
            __syncthreads();
            if (tid == 0) {
              if (fused) {
                double w, r;
                fold_combine(sm.foldW, sm.foldR, w, r);
                st->wsnap = (unsigned long long)__double_as_longlong(w);
                st->resAcc = r;
                st->resLast = sqrt(r / fmax(st->nGroups, 1.0)) * (double)c.resScale;
              }
              st->flags |= stop_bits(c, *st, st->resLast);
            }
          }
        }
      }
    } else {
      cur[0] = R(sv.cur[0]); cur[1] = R(sv.cur[1]); cur[2] = R(sv.cur[2]); cur[3] = R(sv.cur[3]);
      dt = cur[0];
      tsel = cur[2];
      if (lt_ == 0 && tid == 0) st->flagsSeen = st->flags;
      if (active && lt_ == 0 && tid == 0) {
        st->last[0] = st->cur[0]; st->last[1] = st->cur[1]; st->last[2] = st->cur[2]; st->last[3] = st->cur[3];
        st->steps += 1;
        if (c.spray) {
          st->head = (st->head + 4u) % (uint32_t)c.NP;
          st->seed = st->seed + 1u;
        }
      }
    }
  };
  // Throughput-tuned shapes (W = 4, many workgroups per CU in turn) read the step state
  // first; the latency-tuned shapes issue the tile's loads first (below).
  if constexpr (W >= 4) {
    sv = load_step_view<STAGE>(st);
    step_state();
  }

  // ---- 1. primitive tile with ghost halo (FS_PRIM); this tile's extra faces. The tile's
  // loads depend only on the tile; on the latency-tuned shapes they are issued ahead of
  // the step state (which decides whether the solve is still running and gives the time
  // the ghost rules need), so that its round trip overlaps them instead of preceding them.
  // This thread's own cell: its group geometry, for the update phase, loaded here so that
  // its latency overlaps the tile load and the face loop (its code word and state come with
  // the tile's cells below: a thread's first tile cell is its own)
  const int tx = tid % TX, ty = tid / TX;
  const int ci = i0 + tx, cj = j0 + ty;
  const bool inDom = ci < Nx && cj < Nr;
  const long long k = (long long)cj * Nx + ci;
  // Every load below is issued by every lane, from an address clamped into its array, and
  // lanes with nothing to load discard the value: loads under a divergent branch would be
  // waited for inside it, one round trip per branch instead of one for the whole set.
  float4 gg = make_float4(0.f, 0.f, 0.f, 0.f);
  // stage 2: the own cell's state at the start of the step (the second stage averages it
  // in), loaded with the tile rather than at the update, where its round trip was exposed
  R u0v[8];
  {
    const unsigned kc = (unsigned)(min(cj, Nr - 1) * Nx + min(ci, Nx - 1));
    const float4 g4 = gld(I.gg, kc);
    if (inDom) gg = g4;
    if constexpr (STAGE == 2) {
#pragma unroll
      for (int m = 0; m < 8; ++m) u0v[m] = gld(U0 + m * Ns, kc);
    }
  }
  // eddy-viscosity length scales of this thread's cells of the viscous ring tile
  constexpr int LPT = (S::TW * S::THh + NT - 1) / NT;
  float ltv[LPT];
  if constexpr (STAGE == 2 && PHYS) {
#pragma unroll
    for (int r = 0; r < LPT; ++r) {
      const int idx = tid + r * NT;
      const int gi = i0 - 1 + idx % S::TW, gj = j0 - 1 + idx / S::TW;
      ltv[r] = (idx < S::TW * S::THh && gi >= 0 && gi < Nx && gj >= 0 && gj < Nr) ? gld(I.lt, (unsigned)(gj * Nx + gi)) : 0.f;
    }
  }
  // this tile's extra faces (faces of its merged cells outside the tile) and the far
  // cells they reach: workgroup-uniform index loads, issued before the tile's own loads
  // so that their latency is not added after them
  const int exBase = cld(I.exStart, (unsigned)lt_);
  const int nEx = min(cld(I.exStart, (unsigned)lt_ + 1u) - exBase, S::MAXE);
  const int fcBase = cld(I.fcStart, (unsigned)lt_);
  const int nF = min(cld(I.fcStart, (unsigned)lt_ + 1u) - fcBase, S::MAXF);
  // every load of this thread's cells and face areas is issued before any is used: its own
  // cell (tile position (tx, ty) + 2), then ring cells of the 2-cell halo (tile_ring_cell)
  constexpr int CPT = 1 + (S::NRING + NT - 1) / NT, FPT = (S::NFX + S::NFR + NT - 1) / NT;
  CellRaw<R> raw[CPT];
  float area[FPT];
  raw[0] = cell_src_load<R>(I, Uin, ci, cj);
#pragma unroll
  for (int r = 1; r < CPT; ++r) {
    const int2 lc = S::ring_cell(min(tid + (r - 1) * NT, S::NRING - 1));  // lanes past the ring reload its last cell
    raw[r] = cell_src_load<R>(I, Uin, i0 - 2 + lc.x, j0 - 2 + lc.y);
  }
#pragma unroll
  for (int r = 0; r < FPT; ++r) {
    const int f = min(tid + r * NT, S::NFX + S::NFR - 1);
    const bool dirR = f >= S::NFX;
    const int g = dirR ? f - S::NFX : f;
    const int fi = i0 + (dirR ? g % TX : g % (TX + 1)), fj = j0 + (dirR ? g / TX : g / (TX + 1));
    // x faces: fi <= Nx, fj < Nr; r faces: fi < Nx, fj <= Nr (0 outside)
    const bool in = dirR ? (fi < Nx && fj <= Nr) : (fi <= Nx && fj < Nr);
    // one load from a per-lane base (no branch on the direction); both bases are taken as
    // workgroup-uniform values first, so the choice is a register select, not a per-lane
    // load of the descriptor field
    const float* pax = I.ax;
    const float* par = I.ar;
    asm("" : "+s"(pax));
    asm("" : "+s"(par));
    const float* base = dirR ? par : pax;
    const int row = dirR ? Nx : Nx + 1;
    const float ld = gld(base, (unsigned)(min(fj, dirR ? Nr : Nr - 1) * row + min(fi, dirR ? Nx - 1 : Nx)));
    area[r] = (in && tid + r * NT < S::NFX + S::NFR) ? ld : 0.f;
  }
  // the last step's partials (stage 1): loaded with the tile, so that their round trip
  // overlaps the tile's, and folded after the face fluxes
  // (tiles of fewer than 256 threads hold more partials per thread: they load them late)
  constexpr bool EARLY_FOLD = NT >= 256;
  FoldLoads<NT> fl;
  if constexpr (W < 4) sv = load_step_view<STAGE>(st);
  // (issued whether or not this launch folds them: a branch here would take the first
  // use of the loads, and its wait, into it)
  if (STAGE == 1 && EARLY_FOLD) fold_issue<NT>(I.part, nPart, tid, fl);
  // every load above is issued before any of them is waited for: nothing below may be
  // moved above this point, and no load above below it
  asm volatile("" ::: "memory");
  if constexpr (W < 4) step_state();

  double wloc = 0.0;
  double rloc = 0.0;
  if (active) {
#pragma unroll
    for (int r = 0; r < CPT; ++r) {
      const int idx = tid + (r - 1) * NT;
      if (r == 0 || idx < S::NRING) {
        const int2 lc = r == 0 ? make_int2(tx + 2, ty + 2) : S::ring_cell(idx);
        const int li = lc.x, lj = lc.y;
        const Prim<R> q = prim_from<R>(I, raw[r], i0 - 2 + li, j0 - 2 + lj, tsel);
        sm.P[0][lj][li] = q.r; sm.P[1][lj][li] = q.u; sm.P[2][lj][li] = q.v; sm.P[3][lj][li] = q.p;
        sm.P[4][lj][li] = q.Z; sm.P[5][lj][li] = q.F; sm.P[6][lj][li] = q.G; sm.P[7][lj][li] = q.O;
      }
    }
#pragma unroll
    for (int r = 0; r < FPT; ++r) {
      const int f = tid + r * NT;
      if (f < S::NFX + S::NFR) {
        const bool dirR = f >= S::NFX;
        const int g = dirR ? f - S::NFX : f;
        if (dirR) sm.AR[g / TX][g % TX] = area[r]; else sm.AX[g / (TX + 1)][g % (TX + 1)] = area[r];
      }
    }
    // extra faces of merged cells outside the tile, and the cells they reach beyond
    // the halo (primitives loaded here in parallel, so no thread waits on them later)
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
    auto tileA = [&](bool dirR, int fi, int fj) { return dirR ? sm.AR[fj - j0][fi - i0] : sm.AX[fj - j0][fi - i0]; };
    auto globA = [&](bool dirR, int fi, int fj) { return face_area(I, dirR, fi, fj); };
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

    // ---- 2. each face flux once (FS_FLUX_X / FS_FLUX_R), times |area|: the west and
    // the south face of every tile cell (one each per thread, direction fixed at compile
    // time), then the tile's east column and north row of faces and the extra faces
    auto putX = [&](int fx, int fy, const R* F, const R* FY, R FO) {
      for (int q = 0; q < 4; ++q) { sm.u.inv.Fx[q][fy][fx] = F[q]; sm.u.inv.Fx[4 + q][fy][fx] = FY[q]; }
      sm.u.inv.Fx[8][fy][fx] = FO;
    };
    auto putR = [&](int fx, int fy, const R* F, const R* FY, R FO) {
      for (int q = 0; q < 4; ++q) { sm.u.inv.Fr[q][fy][fx] = F[q]; sm.u.inv.Fr[4 + q][fy][fx] = FY[q]; }
      sm.u.inv.Fr[8][fy][fx] = FO;
    };
    if constexpr (PC) {
      // each cell's MUSCL values in both directions and its tangential jump ratios; the
      // x-direction value at its east face reaches the west face of the next cell from the
      // neighbouring lane (the tile's west column of faces is the boundary pass's), the
      // r-direction value at its north face the south face of the cell above through LDS
      // (the south row likewise)
      const int li = tx + 2, lj = ty + 2;
      const Prim<R> cc = P_at(li, lj), cw = P_at(li - 1, lj), ce = P_at(li + 1, lj);
      const Prim<R> cs = P_at(li, lj - 1), cn = P_at(li, lj + 1);
      const bool blend = I.c.hllBlend > 0.f;
      R pX[8], mX[8], pY[8], mY[8];
      muscl_values(cw, cc, ce, pX, mX);
      muscl_values(swap_uv(cs), swap_uv(cc), swap_uv(cn), pY, mY);
      const R jY = blend ? tang_jump(cc, cn, cs) : R(0);  // x faces: tangential neighbours in r
      const R jX = blend ? tang_jump(cc, ce, cw) : R(0);  // r faces: tangential neighbours in x
      if (ty < TY - 1) {
        for (int q = 0; q < 8; ++q) sm.u.inv.pc.TS[q][ty][tx] = pY[q];
        sm.u.inv.pc.JS[ty][tx] = jX;
      }
      {
        R vW[8];
        for (int q = 0; q < 8; ++q) vW[q] = lane_prev(pX[q]);
        const R jW = lane_prev(jY);
        R F[4], FY[4], FO;
        face_flux_pc<R, 0>(I, j0 + ty, cw, cc, vW, mX, jW, jY, sm.AX[ty][tx], F, FY, FO);
        if (tx > 0) putX(tx, ty, F, FY, FO);
      }
      __syncthreads();
      // the south neighbour's values, read before any r-face flux overwrites them
      const int tb = ty > 0 ? ty - 1 : 0;
      R vS[8];
      for (int q = 0; q < 8; ++q) vS[q] = sm.u.inv.pc.TS[q][tb][tx];
      const R jS = sm.u.inv.pc.JS[tb][tx];
      __syncthreads();
      {
        R F[4], FY[4], FO;
        face_flux_pc<R, 1>(I, j0 + ty, swap_uv(cs), swap_uv(cc), vS, mY, jS, jX, sm.AR[ty][tx], F, FY, FO);
        if (ty > 0) putR(tx, ty, F, FY, FO);
      }
    } else {
      R F[4], FY[4], FO;
      face_flux_w<R, 0>(I, false, i0 + tx, j0 + ty, tileP, tileA, F, FY, FO);
      putX(tx, ty, F, FY, FO);
      face_flux_w<R, 1>(I, true, i0 + tx, j0 + ty, tileP, tileA, F, FY, FO);
      putR(tx, ty, F, FY, FO);
    }
    // The tile's north row of r faces and east column of x faces (with per-cell face
    // states also its south row and west column), then the extra faces. Work item v:
    // [0, TX) north faces, [TX, 2 TX) south faces, from B_E the east faces and then the
    // west faces, from B_X the extra faces. Shapes tuned for latency (W < 4: one wave of
    // workgroups fills the GPU) start each kind on a wave of its own, so the kinds run side
    // by side on different SIMDs rather than one after another in one divergent wave, each
    // with the flux routine specialised for its direction; the throughput shapes (W = 4,
    // many workgroups per CU in turn) pack the tile's faces into as few waves as possible
    // and evaluate them with the one routine that takes the direction at run time.
    constexpr bool SPLIT = W < 4;
    constexpr int NB = PC ? 2 : 1;  // with per-cell face states also the south row and west column
    constexpr int B_E = SPLIT ? (NB * TX + 63) / 64 * 64 : NB * TX, B_X = B_E + (SPLIT ? (NB * TY + 63) / 64 * 64 : NB * TY);
    for (int v = tid; v < B_X + nEx; v += NT) {
      R F[4], FY[4], FO;
      if (v < B_X) {
        const bool dirR = v < B_E;
        const int k = dirR ? v : v - B_E;
        if (k < NB * (dirR ? TX : TY)) {
          const bool sw = k >= (dirR ? TX : TY);  // the south row / west column
          const int kk = sw ? k - (dirR ? TX : TY) : k;
          const int fx = dirR ? kk : (sw ? 0 : TX), fy = dirR ? (sw ? 0 : TY) : kk;
          if constexpr (SPLIT) {
            if (dirR) face_flux_w<R, 1>(I, true, i0 + fx, j0 + fy, tileP, tileA, F, FY, FO);
            else face_flux_w<R, 0>(I, false, i0 + fx, j0 + fy, tileP, tileA, F, FY, FO);
          } else {
            face_flux_w<R>(I, dirR, i0 + fx, j0 + fy, tileP, tileA, F, FY, FO);
          }
          if (dirR) putR(fx, fy, F, FY, FO);
          else putX(fx, fy, F, FY, FO);
        }
      } else {
        const int e = v - B_X;
        const uint2 ef = sm.ex[e];
        face_flux_w<R>(I, ex_dir(ef), ex_i(ef), ex_j(ef), anyP, globA, F, FY, FO);
        for (int q = 0; q < 4; ++q) { gst(exFlux, e * 16 + q, F[q]); gst(exFlux, e * 16 + 4 + q, FY[q]); }
        gst(exFlux, e * 16 + 8, FO);
      }
    }
    // the last step's partials (stage 1), folded once this wave's faces are formed: the fold
    // is a chain of cross-lane steps with little arithmetic, whose latency the other waves'
    // face work now hides (the time step it gives is first needed by the update below)
    if (STAGE == 1 && fused) {
      if constexpr (EARLY_FOLD) fold_finish<NT>(I.part, nPart, tid, fl, sm.foldW, sm.foldR);
      else fold_groups<NT>(I.part, nPart, tid, sm.foldW, sm.foldR);
    }
    __syncthreads();

    // this step's time step (stage 1); the first workgroup records it, with the folded
    // wave speed and residual of the last step and that step's residual stop test
    if (STAGE == 1) {
      double wF = 0.0, rF = 0.0;
      if (fused) fold_combine(sm.foldW, sm.foldR, wF, rF);
      const double wmax = fused ? wF : __longlong_as_double((long long)sv.wsnap);
      step_info<R>(c, sv.last0, sv.last2, wmax, cur);
      dt = cur[0];
      if (lt_ == 0 && tid == 0) {
        st->cur[0] = cur[0]; st->cur[1] = cur[1]; st->cur[2] = cur[2]; st->cur[3] = cur[3];
        // the parcels this step spawns weigh sprayMhat dt / 4: the bound of the deposit's
        // fixed-point scale, which this step's spray and stage 2 read
        if (c.spray) st->mmax = fmax(st->mmax, (double)(float)(R(c.sprayMhat) * cur[0] / R(4)));
        if (fused) {
          st->wsnap = (unsigned long long)__double_as_longlong(wF);
          st->resAcc = rF;
          st->resLast = sqrt(rF / fmax(st->nGroups, 1.0)) * (double)c.resScale;
          st->flags |= stop_bits(c, *st, st->resLast);
        }
      }
    }

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
        for (int q = 0; q < 9; ++q) F9[q] = gld(exFlux, slot * 16 + q);
        ++slot;
      } else {
        face_flux_w<R>(I, dirR, fi, fj, anyP, globA, F9, F9 + 4, F9[8]);
      }
    };

    // ---- 3. merged-group finite-volume update (FS_UPDATE); the thread's own cell's
    // state and code word came with the tile (raw[0])
    const uint32_t code = inDom ? raw[0].code : (uint32_t)C_IS_SLAVE;
    R a[4], b[4];
    for (int q = 0; q < 4; ++q) { a[q] = raw[0].v[q]; b[q] = raw[0].v[4 + q]; }
    const bool solid = (code & C_SOLID) != 0;
    const bool host = inDom && !solid && !(code & C_IS_SLAVE);
    R A[4], B[4];         // group state after the stage
    // Group members in fixed slots: 0 the host, 1..4 its west, east, south and north
    // neighbours, present as its slave bits say (cutcell.js:168-174). The short loops over
    // them are unrolled over the five slots and the face loops take a member's cell from its
    // slot number, so no array here is indexed at run time (which would put it in scratch).
    uint32_t mem = 0u;    // bit m: member slot m present
    R ubarM[5];           // member-specific B.w (inflow-speed average in column 0)
    int mi[5], mj[5];
    mi[0] = ci; mj[0] = cj; mi[1] = ci - 1; mj[1] = cj; mi[2] = ci + 1; mj[2] = cj;
    mi[3] = ci; mj[3] = cj - 1; mi[4] = ci; mj[4] = cj + 1;
    auto mem_i = [&](int m) { return ci + (m == 1 ? -1 : m == 2 ? 1 : 0); };
    auto mem_j = [&](int m) { return cj + (m == 3 ? -1 : m == 4 ? 1 : 0); };
    // slaves outside the tile take 3 consecutive extra-face slots each, in member order: the
    // first slot of member m, with es the running slot (from the host's code word)
    auto mem_slot = [&](int m, int& es) {
      const int xi = mem_i(m), yj = mem_j(m);
      const bool inT = xi >= i0 && xi < i0 + TX && yj >= j0 && yj < j0 + TY;
      const int slot = (inT || es < 0) ? -1 : es;
      if (!inT && es >= 0) es += 3;
      return slot;
    };
    bool guard = false;   // NaN guard hit
    R rho_old = R(0);
    Prim<R> qS;           // state after the stage, as primitives (physics input)
    if (host) {
      mem = 1u | (code_slaves(code) << 1);
      R L[4], LY[4], LO;
      for (int q = 0; q < 4; ++q) {
        L[q] = -(sm.u.inv.Fx[q][ty][tx + 1] - sm.u.inv.Fx[q][ty][tx]) - (sm.u.inv.Fr[q][ty + 1][tx] - sm.u.inv.Fr[q][ty][tx]);
        LY[q] = -(sm.u.inv.Fx[4 + q][ty][tx + 1] - sm.u.inv.Fx[4 + q][ty][tx]) - (sm.u.inv.Fr[4 + q][ty + 1][tx] - sm.u.inv.Fr[4 + q][ty][tx]);
      }
      LO = -(sm.u.inv.Fx[8][ty][tx + 1] - sm.u.inv.Fx[8][ty][tx]) - (sm.u.inv.Fr[8][ty + 1][tx] - sm.u.inv.Fr[8][ty][tx]);
      int es = code_exslot(code) * 3;
      for (int m = 1; m < 5; ++m) {
        if (!((mem >> m) & 1u)) continue;
        // the slave's own net flux: west/south faces enter, east/north faces leave
        R Fw[9], Fe_[9], Fs[9], Fn[9];
        int slot = mem_slot(m, es);
        const int xi = mem_i(m), yj = mem_j(m);
        invFace(0, xi, yj, slot, Fw);
        invFace(1, xi, yj, slot, Fe_);
        invFace(2, xi, yj, slot, Fs);
        invFace(3, xi, yj, slot, Fn);
        for (int q = 0; q < 4; ++q) {
          L[q] += -(Fe_[q] - Fw[q]) - (Fn[q] - Fs[q]);
          LY[q] += -(Fe_[4 + q] - Fw[4 + q]) - (Fn[4 + q] - Fs[4 + q]);
        }
        LO += -(Fe_[8] - Fw[8]) - (Fn[8] - Fs[8]);
      }
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
        const R* a0 = u0v;
        const R* b0 = u0v + 4;
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
        for (int m = 0; m < 5; ++m) ubarM[m] = R(c.ambB[3]);
      } else {
        const R ke = R(0.5) * rho * (vx * vx + vy * vy);
        const R Gn = clampR(UY[2], R(c.gamR[0]), R(c.gamR[1]));
        const R E = fmax(Uq[3], R(1e-7) * Gn + ke);
        const R Zn = clampR(UY[0], R(0), rho);
        A[0] = rho; A[1] = rho * vx; A[2] = rho * vy; A[3] = E;
        B[0] = Zn; B[1] = clampR(UY[1], R(0), Zn); B[2] = Gn;
        B[3] = c.inj ? clampR(UY[3], R(0), rho) : R(0);
#pragma unroll
        for (int m = 0; m < 5; ++m) {
          if (!((mem >> m) & 1u)) continue;  // an absent slot's cell may lie outside the grid
          R ub = B[3];
          if (!c.inj) {
            ub = R(0);
            if (mi[m] == 0) {
              const R u0 = gld(Uin + 7 * Ns, (unsigned)(mj[m] * Nx));
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
    if constexpr (STAGE == 2 && PHYS) {
    if (c.visc || c.chem) {  // per instance: a batch mixes both kinds
      // Zukoski separation (below) reads the stage-1 state a few cells downstream of a
      // supersonic boundary-layer host: waves holding a boundary-layer host issue those
      // loads here and fold them after the viscous cells, so their latency overlaps that
      // work. zSup: the host is supersonic; zup, zMn: its speed and Mach number.
      const bool zHost = c.visc && host && (code & C_BL);
      const bool zWave = __ballot(zHost) != 0ull;  // uniform over the wave
      bool zSup = false;
      R zup = R(0), zMn = R(0), zpd = R(0);
      uint32_t zc[4];
      R zq[4][5];
      if (zWave) {
        zup = sq(qS.u * qS.u + qS.v * qS.v);
        zMn = dv(zup, sq(dv((R(1) + rc(qS.G)) * qS.p, qS.r)));
        zSup = zHost && zMn > R(1);
        const R iup = rc(zup);
        const R shx = zSup ? qS.u * iup : R(0), shy = zSup ? qS.v * iup : R(0);
#pragma unroll
        for (int z = 0; z < 4; ++z) {
          const int n = z < 2 ? 4 : 6, m = z & 1;
          const int Dx = min(max(ci + (int)round(R(n) * shx), 0), Nx - 1);
          const int Dy = min(max(cj + (int)floor(R(n) * shy) + m, 0), Nr - 1);
          const unsigned kd = (unsigned)(Dy * Nx + Dx);
          zc[z] = gld(I.code, kd);
          for (int q = 0; q < 4; ++q) zq[z][q] = gld(U1 + q * Ns, kd);
          zq[z][4] = gld(U1 + 6 * Ns, kd);
        }
      }
      __syncthreads();  // inviscid fluxes consumed; reuse their LDS
#pragma unroll
      for (int r = 0; r < LPT; ++r) {
        const int idx = tid + r * NT;
        if (idx >= S::TW * S::THh) break;
        const int ri = idx % S::TW, rj = idx / S::TW;
        const int li = ri + 1, lj = rj + 1;
        const int gi = i0 - 2 + li, gj = j0 - 2 + lj;
        R T1[4] = {0, 0, 0, 0}, T2[4] = {0, 0, 0, 0};
        if (gi >= 0 && gi < Nx && gj >= 0 && gj < Nr && sm.P[0][lj][li] >= R(0)) {
          vcell<R>(P_at(li, lj), P_at(li + 1, lj), P_at(li - 1, lj), P_at(li, lj + 1), P_at(li, lj - 1), gi, gj,
                   R(ltv[r]), dt, c, T1, T2);
        }
        for (int q = 0; q < 4; ++q) { sm.u.vis.T[q][rj][ri] = T1[q]; sm.u.vis.T[4 + q][rj][ri] = T2[q]; }
      }
      __syncthreads();
      // the probes' largest pressure over the cells that are not solid, in probe order
      if (zWave) {
#pragma unroll
        for (int z = 0; z < 4; ++z)
          if (!(zc[z] & C_SOLID))
            zpd = fmax(zpd, to_prim<R>(zq[z][0], zq[z][1], zq[z][2], zq[z][3], R(0), R(0), zq[z][4], R(0)).p);
      }
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
        // as the inviscid faces: west and south face of each cell, then the rest
        auto putVX = [&](int fx, int fy, const R* F, const R* FY) {
          for (int q = 0; q < 4; ++q) sm.u.vis.Vx[q][fy][fx] = F[q];
          for (int q = 0; q < 3; ++q) sm.u.vis.Vx[4 + q][fy][fx] = FY[q];
        };
        auto putVR = [&](int fx, int fy, const R* F, const R* FY) {
          for (int q = 0; q < 4; ++q) sm.u.vis.Vr[q][fy][fx] = F[q];
          for (int q = 0; q < 3; ++q) sm.u.vis.Vr[4 + q][fy][fx] = FY[q];
        };
        {
          R F[4], FY[3];
          vface_w<R, 0>(I, false, i0 + tx, j0 + ty, tileP, tileT, tileA, F, FY);
          putVX(tx, ty, F, FY);
          vface_w<R, 1>(I, true, i0 + tx, j0 + ty, tileP, tileT, tileA, F, FY);
          putVR(tx, ty, F, FY);
        }
        // the north row, east column and extra faces, laid out as in the inviscid pass
        constexpr int VB_E = SPLIT ? (TX + 63) / 64 * 64 : TX, VB_X = VB_E + (SPLIT ? (TY + 63) / 64 * 64 : TY);
        for (int v = tid; v < VB_X + nEx; v += NT) {
          R F[4], FY[3];
          if (v < VB_X) {
            const bool dirR = v < VB_E;
            const int k = dirR ? v : v - VB_E;
            if (k < (dirR ? TX : TY)) {
              const int fx = dirR ? k : TX, fy = dirR ? TY : k;
              if constexpr (SPLIT) {
                if (dirR) vface_w<R, 1>(I, true, i0 + fx, j0 + fy, tileP, tileT, tileA, F, FY);
                else vface_w<R, 0>(I, false, i0 + fx, j0 + fy, tileP, tileT, tileA, F, FY);
              } else {
                vface_w<R>(I, dirR, i0 + fx, j0 + fy, tileP, tileT, tileA, F, FY);
              }
              if (dirR) putVR(fx, fy, F, FY);
              else putVX(fx, fy, F, FY);
            }
          } else {
            const int e = v - VB_X;
            const uint2 ef = sm.ex[e];
            vface_w<R>(I, ex_dir(ef), ex_i(ef), ex_j(ef), anyP, anyT, globA, F, FY);
            for (int q = 0; q < 4; ++q) gst(exFlux, e * 16 + 9 + q, F[q]);
            for (int q = 0; q < 3; ++q) gst(exFlux, e * 16 + 13 + q, FY[q]);
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
          for (int q = 0; q < 7; ++q) F7[q] = gld(exFlux, slot * 16 + 9 + q);
          ++slot;
        } else {
          vface_w<R>(I, dirR, fi, fj, anyP, anyT, globA, F7, F7 + 4);
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
          int es = code_exslot(code) * 3;
          for (int m = 0; m < 5; ++m) {
            if (!((mem >> m) & 1u)) continue;
            R Fw[7], Fe_[7], Fs[7], Fn[7];
            int slot = m == 0 ? -1 : mem_slot(m, es);
            const int xi = mem_i(m), yj = mem_j(m);
            visFace(0, xi, yj, slot, Fw);
            visFace(1, xi, yj, slot, Fe_);
            visFace(2, xi, yj, slot, Fs);
            visFace(3, xi, yj, slot, Fn);
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
                                       : rs * powR(R(7.187) * powR(nzv, R(1.142857)) + R(0.13769) * powR(nzv, R(0.142857)) * up, R(1.75));
            const R iwdt = rc(fmax(wl * dt, R(1e-9)));
            tw = fmin(tw, R(0.3) * rho * up * V * iwdt);
            if (up > R(1e-6)) { const R f = dv(tw * wl, up); L[1] -= f * utx; L[2] -= f * utr; }
            const R qw = dv(tw, fmax(up, R(1e-4))) * cp * (Taw - R(c.Tw)) * R(1.24);
            const R qmax = R(0.3) * p * G * V * iwdt;
            L[3] -= clampR(qw, -qmax, qmax) * wl;
          }
          // Zukoski separation in the boundary-layer band (solver.js:613-630)
          if (code & C_BL) {
            // probes at n = 4, 6 along the flow, rows floor(n shy) + 0, 1: loaded and folded above
            const R up = zup, Mn = zMn;
            if (zSup) {
              const R crit = R(1) + R(0.5) * Mn;
              const R fs = smoothstepR(crit, R(1.2) * crit, dv(zpd, p)) * V * rho * up;
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
            const R kin = R(c.chemK[2]) * p * expR(R(-9000) / TK) * (R(1) - smoothstepR(R(c.cap[0]), R(c.cap[1]), TK));
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
          for (int m = 0; m < 5; ++m) ubarM[m] = R(c.ambB[3]);
        } else {
          const R Gn = clampR(UY[2], R(c.gamR[0]), R(c.gamR[1]));
          const R Zn = clampR(UY[0], R(0), rn);
          A[0] = rn; A[1] = rn * vnx; A[2] = rn * vny;
          A[3] = fmax(Uq[3], R(1e-7) * Gn + R(0.5) * rn * (vnx * vnx + vny * vny));
          B[0] = Zn; B[1] = clampR(UY[1], R(0), Zn); B[2] = Gn;
          if (c.inj) {
            B[3] = clampR(UY[3], R(0), rn);
            for (int m = 0; m < 5; ++m) ubarM[m] = B[3];
          }
        }
      }
    }
    }

    // ---- 5. spray gas sources (FS_SPRAY_APPLY, spray.js:158-178), consumed and cleared
    if (STAGE == 2 && c.spray && host) {
      unsigned long long* __restrict__ src = I.src;
      const unsigned ku = (unsigned)k;
      const unsigned long long q0 = gld(src, ku), q1 = gld(src + Ns, ku), q2 = gld(src + 2 * Ns, ku), q3 = gld(src + 3 * Ns, ku);
      if ((q0 | q1 | q2 | q3) != 0ull) {
        gst(src, ku, 0ull); gst(src + Ns, ku, 0ull); gst(src + 2 * Ns, ku, 0ull); gst(src + 3 * Ns, ku, 0ull);
        double sc[2];
        dep_scales(c, st->mmax, sc);
        const R s0 = R((double)(long long)q0 / sc[0]), s1 = R((double)(long long)q1 / sc[0]);
        const R s2 = R((double)(long long)q2 / sc[1]), s3 = R((double)(long long)q3 / sc[1]);
        const R dm = s0 + s1;
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
        for (int m = 0; m < 5; ++m) ubarM[m] = bw;
      }
    }

    // ---- 6. write the group, wave speed for the next step's CFL, residual
    R* __restrict__ Uout = STAGE == 1 ? U1 : U0;
    if (inDom && solid) {
      for (int q = 0; q < 4; ++q) { gst(Uout + q * Ns, (unsigned)k, R(c.ambCons[q])); gst(Uout + (4 + q) * Ns, (unsigned)k, R(c.ambB[q])); }
    }
    if (host) {
#pragma unroll
      for (int m = 0; m < 5; ++m) {
        if (!((mem >> m) & 1u)) continue;
        const long long km = (long long)mj[m] * Nx + mi[m];
        for (int q = 0; q < 4; ++q) gst(Uout + q * Ns, (unsigned)km, A[q]);
        for (int q = 0; q < 3; ++q) gst(Uout + (4 + q) * Ns, (unsigned)km, B[q]);
        gst(Uout + 7 * Ns, (unsigned)km, ubarM[m]);
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
    // per-tile partials, folded by the next stage 1 or the close kernel (no same-address
    // atomics across workgroups)
    double wb = wloc, rb = rloc;
    block_max_sum<NT>(wb, rb, sm.red, sm.redd);
    if (tid == 0) I.part[lt_] = make_double2(active ? wb : 0.0, active ? rb : 0.0);
  }
}

}  // namespace nzk
