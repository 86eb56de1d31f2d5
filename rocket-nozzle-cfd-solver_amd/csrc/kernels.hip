// kernels.hip — fused per-step kernels of the MI355X nozzle solver (gfx950, wave64).
//
// One solver step is two (three with the spray) launches instead of the
// reference's ~21 draws (solver.js:849-891):
//   k_stage<1>   LDS tile of primitives with a 2-cell ghost halo (FS_PRIM) ->
//                each x/r face flux computed once (MUSCL + HLLC/HLL, FS_FLUX) ->
//                merged-group finite-volume update U1 = U0 + dt L(U0) (FS_UPDATE).
//   k_spray      parcel push, evaporation, break-up; deposits its release into the
//                host cell, summed per cell in LDS, then one float atomic per cell and
//                channel (FS_SPRAY_UPDATE + FS_SPRAY_DEP).
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
#include "nz_stage.h"
#include "stage_shapes.h"

#include <hip/hip_ext.h>

#include <algorithm>

namespace nzk {
// Folds every instance's tile partials of the last stage 2 (one workgroup per instance): the
// wave-speed max for the next time step, the residual of the last step and the stop
// tests of batched solves. It runs after every step when an instance has more partials
// than the next stage 1 folds itself, and once after each run of steps, so that what the
// host reads is current. It changes nothing the next stage 1 would not compute alike.
__global__ void __launch_bounds__(256) k_close(const DevInst* __restrict__ insts, int ninst) {
  __shared__ double W[4], R[4];
  const int n = blockIdx.x;
  if (n >= ninst) return;
  const DevInst& J = insts[n];
  InstState* s = J.st;
  if (s->flags & 1u) {
    if (threadIdx.x == 0) s->flagsSeen = s->flags;
    return;
  }
  fold_groups<256>(J.part, (uint32_t)(J.tilesX * J.tilesY), threadIdx.x, W, R);
  __syncthreads();
  double w, r;
  fold_combine(W, R, w, r);
  if (threadIdx.x == 0) {
    s->wsnap = (unsigned long long)__double_as_longlong(w);
    s->resAcc = r;
    s->resLast = sqrt(r / fmax(s->nGroups, 1.0)) * (double)J.c.resScale;
    s->flags |= stop_bits(J.c, *s, s->resLast);
    s->flagsSeen = s->flags;
  }
}

// ------------------------------------------------------------------------------
// Lagrangian parcels (FS_SPRAY_UPDATE, spray.js:40-121) with the point deposit
// (VS_SPRAY_DEP, spray.js:131-142) into the host cell. Parcels crowd into few cells
// near the injector, where per-parcel atomics on the same addresses serialise; so each
// workgroup first sums its parcels' deposits per host cell in an LDS hash table (open
// addressing, twice as many slots as parcels, so every insert ends) and then adds each
// distinct cell's sums to the source planes with one atomic per channel. Every sum is of
// 64-bit fixed-point integers (dep_scales), which add associatively: the sources, and so
// the run, come out bit for bit the same whatever order the atomics land in, and a solve
// gives the same bits alone and in a batch.
constexpr int DEP_LOG = 9, DEP_SLOTS = 1 << DEP_LOG;
static_assert(DEP_SLOTS >= 2 * 256, "deposit table: at least two slots per parcel of a workgroup");
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
  __shared__ int depKey[DEP_SLOTS];
  __shared__ unsigned long long depVal[4][DEP_SLOTS];
  for (int q = threadIdx.x; q < DEP_SLOTS; q += 256) {
    depKey[q] = -1;
    depVal[0][q] = 0ull; depVal[1][q] = 0ull; depVal[2][q] = 0ull; depVal[3][q] = 0ull;
  }
  int depH = -1;   // host cell this parcel deposits into (-1: none)
  unsigned long long dep[4] = {0ull, 0ull, 0ull, 0ull};
  if (idx < NP) {
    const int Nx = I.Nx, Nr = I.Nr;
    const long long Ns = I.Ns;
    const R* __restrict__ U1 = static_cast<const R*>(I.U1);
    float* __restrict__ pc = I.parc;
    const R dt = R(st->cur[0]);
    const uint32_t head = st->head, seed = st->seed;
    const unsigned ui = (unsigned)idx, un = (unsigned)NP;
    const R ax_ = gld(pc, ui), ay = gld(pc, un + ui), az = gld(pc, 2 * un + ui), aw = gld(pc, 3 * un + ui);
    const R bx = gld(pc, 4 * un + ui), by = gld(pc, 5 * un + ui), bz = gld(pc, 6 * un + ui), bw = gld(pc, 7 * un + ui);
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
      R d;
      if (e3[1] > 0.f) {
        // lognormal drop size (an injector-table option): median e2.y, spread e3.y, the normal
        // deviate by Box-Muller from draws 4 and 6
        const R u1 = R(1) - R(rnd(id, 4u)), u2 = R(rnd(id, 6u));
        const R nrm = sq(R(-2) * logR(u1)) * cosR(R(6.2831853071795865) * u2);
        d = R(e2[1]) * expR(R(e3[1]) * nrm);
      } else {
        d = R(e2[1]) * expR((R(rnd(id, 4u)) - R(0.5)) * R(1.4f));   // spray.js:53
      }
      o[0] = e1[0]; o[1] = r; o[2] = R(e1[3]) * cosR(ang); o[3] = R(e1[3]) * sinR(ang);
      o[4] = d; o[5] = R(c.sprayMhat) * dt / R(perStep); o[6] = e2[0]; o[7] = e3[0];
      if (by > R(0)) { dmo = by; dpx = by * az; dpy = by * aw; }
    } else if (by > R(0)) {
      const int gi = min(max((int)floor(ax_), 0), Nx - 1), gj = min(max((int)floor(ay), 0), Nr - 1);
      const long long kg = (long long)gj * Nx + gi;
      if (gld(I.code, (unsigned)kg) & C_SOLID) {
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
          const R T = dv(w.p, w.r * rm) * R(c.gas[2]);
          const R rhoG = w.r * R(1.225f);
          const R rvx = w.u - vx, rvy = w.v - vy;
          const R Urel = sq(rvx * rvx + rvy * rvy) * R(287.6f) + R(1e-3f);
          const R mu = dv(R(1.458e-6f) * T * sq(T), T + R(110.4f));
          const R Re = dv(rhoG * Urel * d, mu);
          const R invTau = dv(R(18) * mu * (R(1) + R(0.15f) * powR(fmin(Re, R(1000)), R(0.687f))), pr0 * d * d);
          const R ex = expR(-dts * invTau);
          vx = w.u + (vx - w.u) * ex;
          vy = w.v + (vy - w.v) * ex;
          const R cp = rm * R(287.05f) * (w.G + R(1));
          const R Bn = dv(fmax(cp * (T - pr1), R(0)), pr2);
          const R K = dv(R(8) * dv(mu, R(0.7f)), pr0) * logR(R(1) + Bn) * (R(1) + R(0.27f) * sq(Re));
          const R d2 = d * d - K * dts;
          dn = d2 > R(1e-12f) ? sq(d2) : R(0);
          const R f = dv(dn, d);
          mn = dn > R(0) ? m * f * f * f : R(0);
          const R we = dv(rhoG * Urel * Urel * dn, pr3);
          if (we > R(12) && dn > R(0)) {
            const R dst = dv(R(12) * pr3, rhoG * Urel * Urel);
            const R tb = dv(R(3) * dn, Urel) * sq(dv(pr0, rhoG));
            dn = mixR(dn, fmax(dst, R(2e-6f)), R(1) - expR(-dv(dts, tb)));
          }
        }
        R xi = bw;
        if (xi > R(0) && ax_ >= xi) {
          const uint32_t id = (uint32_t)idx ^ (seed * 2654435761u);
          const R sp = sq(vx * vx + vy * vy), phi = (R(rnd(id, 5u)) - R(0.5)) * R(2) * R(c.fan);
          vx = sp * cosR(phi);
          vy = sp * sinR(phi);
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
    for (int q = 0; q < 8; ++q) gst(pc, (unsigned)q * un + ui, (float)o[q]);
    // deposit into the host of the cell the parcel started in
    if (!(ax_ < R(0)) && !(dmo == R(0) && dpx == R(0) && dpy == R(0))) {
      int gi = min(max((int)floor(ax_), 0), Nx - 1), gj = min(max((int)floor(ay), 0), Nr - 1);
      const uint32_t cg = I.code[(long long)gj * Nx + gi];
      gi += code_hdx(cg);
      gj += code_hdy(cg);
      depH = gj * Nx + gi;
      double sc[2];
      dep_scales(c, st->mmax, sc);
      dep[0] = dep_fix((double)(dmo * (R(1) - bz)), sc[0], c.depExp);
      dep[1] = dep_fix((double)(dmo * bz), sc[0], c.depExp);
      dep[2] = dep_fix((double)dpx, sc[1], c.depExp);
      dep[3] = dep_fix((double)dpy, sc[1], c.depExp);
    }
  }
  __syncthreads();
  if (depH >= 0) {
    unsigned slot = ((unsigned)depH * 2654435761u) >> (32 - DEP_LOG);
    for (;;) {
      const int prev = atomicCAS(&depKey[slot], -1, depH);
      if (prev == -1 || prev == depH) break;
      slot = (slot + 1u) & (unsigned)(DEP_SLOTS - 1);
    }
    for (int q = 0; q < 4; ++q) atomicAdd(&depVal[q][slot], dep[q]);
  }
  __syncthreads();
  unsigned long long* __restrict__ src = I.src;
  const long long Ns = I.Ns;
  for (int q = threadIdx.x; q < DEP_SLOTS; q += 256) {
    const int h = depKey[q];
    if (h < 0) continue;
    for (int ch = 0; ch < 4; ++ch)
      if (depVal[ch][q]) atomicAdd(&src[ch * Ns + h], depVal[ch][q]);
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
__global__ void __launch_bounds__(256) k_cfl_reduce(DevInst I, int nblk, int cap) {
  __shared__ double red[16];
  double w = 0.0;
  for (int k = threadIdx.x; k < nblk; k += 256) w = fmax(w, I.part[k].x);
  const double wb = block_max<256>(w, red);
  __syncthreads();
  // the partials folded into entry 0, zeros after it (valid for every tiling)
  for (int k = 1 + threadIdx.x; k < cap; k += 256) I.part[k] = make_double2(0.0, 0.0);
  if (threadIdx.x == 0) {
    I.part[0] = make_double2(wb, 0.0);
    I.st->wsnap = (unsigned long long)__double_as_longlong(wb);
    I.st->wacc = 0ull;
  }
}

// the n partials of a tiling folded (in the stage kernels' fixed order) into entry 0, with
// zeros up to cap: the form every tiling reads alike (before a tile-shape change)
__global__ void __launch_bounds__(256) k_part_canon(double2* __restrict__ part, int n, int cap) {
  __shared__ double W[4], R[4];
  fold_groups<256>(part, (uint32_t)n, threadIdx.x, W, R);
  __syncthreads();
  double w, r;
  fold_combine(W, R, w, r);
  for (int k = 1 + threadIdx.x; k < cap; k += 256) part[k] = make_double2(0.0, 0.0);
  if (threadIdx.x == 0) part[0] = make_double2(w, r);
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
__global__ void k_spray_reset(float* __restrict__ parc, int NP, unsigned long long* __restrict__ src, long long Ns) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < NP) {
    parc[k] = -1.f; parc[NP + k] = -1.f;
    for (int q = 2; q < 8; ++q) parc[(long long)q * NP + k] = 0.f;
  }
  if (k < Ns)
    for (int q = 0; q < 4; ++q) src[q * Ns + k] = 0ull;
}

// FP64 integrals of the current state: exit plane, throat plane, chamber probe, totals
// FS_DERIVED (render.js:60-134): one display field per cell plus (Z, Tn, M); neighbours
// clamped into the grid, the axis mirrored for the density gradient and the vorticity.
// Constants are the shader's float literals and uniforms, also in the FP64 build.
template <typename R>
__global__ void __launch_bounds__(256) k_derived(DevInst I, DerivedC d, const float* __restrict__ T2, float* __restrict__ out) {
  const int Nx = I.Nx, Nr = I.Nr;
  const long long N = (long long)Nx * Nr;
  const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
  if (k >= N) return;
  const int i = (int)(k % Nx), j = (int)(k / Nx);
  const R* __restrict__ U = static_cast<const R*>(I.U0);
  const long long Ns = I.Ns;
  auto cell = [&](int a, int b) { return (long long)min(max(b, 0), Nr - 1) * Nx + min(max(a, 0), Nx - 1); };
  auto solid = [&](int a, int b) { return (gld(I.code, (unsigned)cell(a, b)) & C_SOLID) != 0; };
  auto P = [&](int a, int b) { return load_prim<R>(U, Ns, cell(a, b)); };
  auto lrho = [&](int a, int b) { return log(fmax(gld(U, (unsigned)cell(a, b < 0 ? -b - 1 : b)), R(1e-7))); };
  float* o = out + k * 4;
  if (solid(i, j)) { o[0] = 0.f; o[1] = 0.f; o[2] = 0.f; o[3] = -1.f; return; }
  const Prim<R> q = P(i, j);
  const R rho = q.r, p = q.p;
  const R c = sound_speed(q);
  const R spd = sqrt(q.u * q.u + q.v * q.v);
  const R M = spd / c;
  const R Rm = (R(1) - q.Z) * R(d.RaN) + q.F * R(d.Rs[0]) + q.O * R(d.Rs[1]) + fmax(q.Z - q.F - q.O, R(0)) * R(d.Rs[2]);
  const R T = p / (rho * Rm) * R(288.15f);
  const int f = d.field == 8 ? 0 : d.field;
  R v = T;
  if (f == 1) v = M;
  else if (f == 2) {
    const R l0 = lrho(i, j);
    const R le = solid(i + 1, j) ? l0 : lrho(i + 1, j);
    const R lw = solid(i - 1, j) ? l0 : lrho(i - 1, j);
    const R ln = solid(i, j + 1) ? l0 : lrho(i, j + 1);
    const R ls = (j > 0 && solid(i, j - 1)) ? l0 : lrho(i, j - 1);
    const R g = sqrt((le - lw) * (le - lw) + (ln - ls) * (ln - ls)) * R(0.5);
    v = R(1) - exp(-R(d.schl) * g);
  } else if (f == 3) v = log(p * R(101325)) / log(R(10));
  else if (f == 4) v = log(rho * R(1.225f)) / log(R(10));
  else if (f == 5) v = spd * R(d.vel);
  else if (f == 6) {
    const Prim<R> e = P(i + 1, j), w = P(i - 1, j), n = P(i, j + 1);
    const R su = j > 0 ? P(i, j - 1).u : q.u;
    const R om = R(0.5) * (e.v - w.v) - R(0.5) * (n.u - su);
    v = fabs(om) * R(d.vel) / R(d.dx);
  } else if (f == 7) v = q.Z;
  else if (f == 9) {
    const float m0 = T2[k * 4], m3 = T2[k * 4 + 3];
    v = m3 > 0.f ? log(fmax(R(m0) / R(m3), R(1))) / log(R(10)) : R(0);
  } else if (f == 10) v = q.Z > R(0.003f) ? (q.Z - q.F) / q.Z : R(0);
  else if (f == 11) v = q.F;
  else if (f == 12) v = q.O;
  else if (f == 13) {
    const R m0 = R(T2[k * 4]), m2 = R(T2[k * 4 + 2]);
    const R avail = fmin(q.F, (q.O + R(0.233f) * fmax(R(1) - q.Z, R(0))) / R(d.chemK[1]));
    const R kin = R(d.chemK[2]) * p * exp(R(-9000) / fmax(T, R(200)));
    v = log(R(1) + R(1e5) * avail * fmin(fmax(m2, R(4) * m0 / rho), kin)) / log(R(10));
  }
  R t = (v - R(d.r0)) / (R(d.r1) - R(d.r0));
  if (f == 2) t = v;
  if (f == 6) t = log(R(1) + R(20) * fmax(v, R(0)) / R(d.r1)) / log(R(21));
  o[0] = (float)t;
  o[1] = (float)q.Z;
  o[2] = (float)((T - R(d.Ta)) / fmax(R(d.Tc) - R(d.Ta), R(1)));
  o[3] = (float)M;
}

// Integrated quantities in FP64, summed in a fixed order so that every call gives the same
// bits. k_integrals_cells: block b sums mass and energy over fluid group hosts in its own
// run of cells (thread t takes cells t, t + 256, ... of the run in order; block_sum folds
// the lanes by a butterfly per wave and the four waves in order) into part[b].
// k_integrals_close: one block folds those partials the same way, sums the exit plane
// (exitPerformance, simsetup.js:314-337) and the throat plane (chamberPerformance,
// 239-250), and reads the chamber probe (cellPrimSI, 338-348).
template <typename R>
__global__ void __launch_bounds__(256) k_integrals_cells(DevInst I, long long run, double2* __restrict__ part) {
  __shared__ double red[8];
  const long long N = (long long)I.Nx * I.Nr;
  const long long k0 = (long long)blockIdx.x * run, k1 = min(k0 + run, N);
  const R* __restrict__ U = static_cast<const R*>(I.U0);
  const long long Ns = I.Ns;
  double m = 0.0, e = 0.0;
  for (long long k = k0 + threadIdx.x; k < k1; k += 256) {
    const uint32_t cd = I.code[k];
    if ((cd & C_SOLID) || (cd & C_IS_SLAVE)) continue;
    const double V = (double)I.gg[k].x;
    m += (double)U[k] * V;
    e += (double)U[3 * Ns + k] * V;
  }
  // both sums folded as block_sum folds each (a butterfly per wave, then the waves in order)
  for (int o = 32; o > 0; o >>= 1) {
    m += __shfl_xor(m, o, 64);
    e += __shfl_xor(e, o, 64);
  }
  if ((threadIdx.x & 63) == 0) { red[threadIdx.x >> 6] = m; red[4 + (threadIdx.x >> 6)] = e; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ms = 0.0, es = 0.0;
    for (int k = 0; k < 4; ++k) { ms += red[k]; es += red[4 + k]; }
    part[blockIdx.x] = make_double2(ms, es);
  }
}

template <typename R>
__global__ void __launch_bounds__(256) k_integrals_close(DevInst I, IntegCtx q, const double2* __restrict__ part, int nb,
                                                         double* __restrict__ out) {
  __shared__ double redn[10][4];
  __shared__ int jmin[4];
  const int tid = threadIdx.x;
  const R* __restrict__ U = static_cast<const R*>(I.U0);
  const long long Ns = I.Ns;
  const int Nx = I.Nx, Nr = I.Nr;
  double mass = 0.0, en = 0.0;
  for (int b = tid; b < nb; b += 256) { mass += part[b].x; en += part[b].y; }
  // exit plane: column iExit up to jExit, face areas of the plane iExit + 1
  double F = 0, md = 0, Mw = 0, pw = 0, Aa = 0;
  for (int j = tid; j < q.jExit; j += 256) {
    const long long k = (long long)j * Nx + q.iExit;
    const double rho = fmax((double)U[k], 1e-9), u = (double)U[Ns + k] / rho, v = (double)U[2 * Ns + k] / rho;
    const double G = (double)U[6 * Ns + k], g = 1.0 + 1.0 / G;
    const double p = fmax(((double)U[3 * Ns + k] - 0.5 * rho * (u * u + v * v)) / G, 1e-9);
    const double dA = 2.0 * M_PI * fabs((double)I.ax[(long long)j * (Nx + 1) + q.iExit + 1]);
    const double c = sqrt(g * p / rho);
    const double m = (double)U[4 * Ns + k] * u * dA;
    F += (rho * u * u + p - q.pa) * dA;
    md += m;
    Mw += m * sqrt(u * u + v * v) / c;
    pw += p * dA;
    Aa += dA;
  }
  // throat plane: fluid cells from the axis up to the first solid one
  int js = Nr;
  for (int j = tid; j < Nr; j += 256)
    if (I.code[(long long)j * Nx + q.iThroat] & C_SOLID) js = min(js, j);
  for (int o = 32; o > 0; o >>= 1) js = min(js, __shfl_xor(js, o, 64));
  if ((tid & 63) == 0) jmin[tid >> 6] = js;
  __syncthreads();
  const int jT = min(min(jmin[0], jmin[1]), min(jmin[2], jmin[3]));
  double mt = 0, mO = 0, mF = 0;
  for (int j = tid; j < jT; j += 256) {
    const long long k = (long long)j * Nx + q.iThroat;
    const double rho = fmax((double)U[k], 1e-9), u = (double)U[Ns + k] / rho;
    const double dA = 2.0 * M_PI * fabs((double)I.ax[(long long)j * (Nx + 1) + q.iThroat]);
    mt += rho * u * dA;
    mO += (double)U[7 * Ns + k] * u * dA;
    mF += (double)U[5 * Ns + k] * u * dA;
  }
  // the ten sums together: each is folded exactly as block_sum folds it (a butterfly per
  // wave, then the four waves in order), with one barrier for all of them
  constexpr int NS = 10;
  double vals[NS] = {F, md, Mw, pw, Aa, mt, mass, en, mO, mF};
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int n = 0; n < NS; ++n) vals[n] += __shfl_xor(vals[n], o, 64);
  if ((tid & 63) == 0)
#pragma unroll
    for (int n = 0; n < NS; ++n) redn[n][tid >> 6] = vals[n];
  __syncthreads();
  if (tid < NS) {
    double m = 0.0;
    for (int k = 0; k < 4; ++k) m += redn[tid][k];
    constexpr int slot[NS] = {0, 1, 2, 3, 4, 5, 6, 7, 11, 12};
    out[slot[tid]] = m;
  }
  if (tid == 0) {
    // chamber probe at (iChamber, 0)
    const long long k = q.iChamber;
    const double rho = fmax((double)U[k], 1e-9), u = (double)U[Ns + k] / rho, v = (double)U[2 * Ns + k] / rho;
    const double G = (double)U[6 * Ns + k];
    const double p = fmax(((double)U[3 * Ns + k] - 0.5 * rho * (u * u + v * v)) / G, 1e-9);
    const double Y = fmin(fmax((double)U[4 * Ns + k] / rho, 0.0), 1.0);
    const double Fv = fmax((double)U[5 * Ns + k] / rho, 0.0);
    const double Ov = q.inj ? fmax((double)U[7 * Ns + k] / rho, 0.0) : 0.0;
    const double Rm = (1.0 - Y) * q.RaN + Fv * q.Rs[0] + Ov * q.Rs[1] + fmax(Y - Fv - Ov, 0.0) * q.Rs[2];
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


hipError_t launch_set_inst(DevInst* dst, const DevInst& src, hipStream_t s) {
  hipLaunchKernelGGL(k_set_inst, dim3(1), dim3(64), 0, s, dst, src);
  return hipGetLastError();
}

#define NZ_DECL_STAGE(name, R, TX, TY, W)                                                                       \
  hipError_t stage_launch_##name(int stage, bool phys, const DevInst* d, int ninst, int ntiles, int chunk,        \
                                 const unsigned short* tileInst, hipStream_t s, hipEvent_t e0, hipEvent_t e1);
NZ_STAGE_SHAPES_F32(NZ_DECL_STAGE)
NZ_STAGE_SHAPES_F64(NZ_DECL_STAGE)
#undef NZ_DECL_STAGE

namespace {
struct ShapeEntry {
  int tx, ty, w;
  hipError_t (*launch)(int, bool, const DevInst*, int, int, int, const unsigned short*, hipStream_t, hipEvent_t, hipEvent_t);
};
#define NZ_SHAPE_ENTRY(name, R, TX, TY, W) {TX, TY, W, stage_launch_##name},
const ShapeEntry kShapes32[] = {NZ_STAGE_SHAPES_F32(NZ_SHAPE_ENTRY)};
const ShapeEntry kShapes64[] = {NZ_STAGE_SHAPES_F64(NZ_SHAPE_ENTRY)};
#undef NZ_SHAPE_ENTRY
}  // namespace

int tile_shape_count(bool fp64) {
  return fp64 ? (int)(sizeof(kShapes64) / sizeof(kShapes64[0])) : (int)(sizeof(kShapes32) / sizeof(kShapes32[0]));
}

bool tile_shape(bool fp64, int shape, int* tx, int* ty, int* waves) {
  if (shape < 0 || shape >= tile_shape_count(fp64)) return false;
  const ShapeEntry& e = fp64 ? kShapes64[shape] : kShapes32[shape];
  *tx = e.tx;
  *ty = e.ty;
  if (waves) *waves = e.w;
  return true;
}

hipError_t launch_stage(int stage, bool phys, bool fp64, int shape, const DevInst* d, int ninst, int ntiles,
                        int tilesX0, const unsigned short* tileInst, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  if (shape < 0 || shape >= tile_shape_count(fp64)) return hipErrorInvalidValue;
  if (ninst > 1 && !tileInst) return hipErrorInvalidValue;
  const ShapeEntry& e = fp64 ? kShapes64[shape] : kShapes32[shape];
  const int chunk = tilesX0 * NZ_XCD_CHUNK_ROWS > 1 ? tilesX0 * NZ_XCD_CHUNK_ROWS : 1;
  return e.launch(stage, phys, d, ninst, ntiles, chunk, tileInst, s, e0, e1);
}
hipError_t launch_close(const DevInst* d, int ninst, hipStream_t s) {
  hipLaunchKernelGGL(k_close, dim3(ninst), dim3(256), 0, s, d, ninst);
  return hipGetLastError();
}
hipError_t launch_spray(bool fp64, const DevInst* d, int ninst, int nblk, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  if (e0 && e1) {
    if (fp64) hipExtLaunchKernelGGL((k_spray<double>), dim3(nblk), dim3(256), 0, s, e0, e1, 0, d, ninst);
    else hipExtLaunchKernelGGL((k_spray<float>), dim3(nblk), dim3(256), 0, s, e0, e1, 0, d, ninst);
  } else {
    if (fp64) hipLaunchKernelGGL((k_spray<double>), dim3(nblk), dim3(256), 0, s, d, ninst);
    else hipLaunchKernelGGL((k_spray<float>), dim3(nblk), dim3(256), 0, s, d, ninst);
  }
  return hipGetLastError();
}
hipError_t launch_part_canon(double2* part, int n, int cap, hipStream_t s) {
  hipLaunchKernelGGL(k_part_canon, dim3(1), dim3(256), 0, s, part, n, cap);
  return hipGetLastError();
}
hipError_t launch_cfl_init(bool fp64, const DevInst& I, int partCap, hipStream_t s) {
  // one partial per block: at most ceil(N / 256), which no shape's tile count undercuts
  // (the smallest tile holds 128 cells in FP32, 64 in FP64), so they fit the partials array
  const long long N = (long long)I.Nx * I.Nr;
  const int nblk = (int)std::min<long long>((N + 255) / 256, 1024);
  if (fp64) hipLaunchKernelGGL((k_cfl_partial<double>), dim3(nblk), dim3(256), 0, s, I);
  else hipLaunchKernelGGL((k_cfl_partial<float>), dim3(nblk), dim3(256), 0, s, I);
  hipLaunchKernelGGL(k_cfl_reduce, dim3(1), dim3(256), 0, s, I, nblk, partCap);
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
hipError_t launch_spray_reset(float* parc, int NP, unsigned long long* src, long long Ns, hipStream_t s) {
  const long long n = NP > Ns ? NP : Ns;
  hipLaunchKernelGGL(k_spray_reset, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, parc, NP, src, Ns);
  return hipGetLastError();
}
hipError_t launch_derived(bool fp64, const DevInst& I, const DerivedC& d, const float* T2, float* out, hipStream_t s) {
  const long long N = (long long)I.Nx * I.Nr;
  const dim3 g((unsigned)((N + 255) / 256)), b(256);
  if (fp64) hipLaunchKernelGGL((k_derived<double>), g, b, 0, s, I, d, T2, out);
  else hipLaunchKernelGGL((k_derived<float>), g, b, 0, s, I, d, T2, out);
  return hipGetLastError();
}
hipError_t launch_integrals(bool fp64, const DevInst& I, const IntegCtx& q, double* scratch, double* out, hipStream_t s) {
  // one block per 1024 cells, at most INTEG_BLOCKS_MAX; partials in scratch
  const long long N = (long long)I.Nx * I.Nr;
  const int nb = (int)std::max<long long>(1, std::min<long long>((N + 1023) / 1024, INTEG_BLOCKS_MAX));
  const long long run = (N + nb - 1) / nb;
  double2* part = reinterpret_cast<double2*>(scratch);
  if (fp64) {
    hipLaunchKernelGGL((k_integrals_cells<double>), dim3(nb), dim3(256), 0, s, I, run, part);
    hipLaunchKernelGGL((k_integrals_close<double>), dim3(1), dim3(256), 0, s, I, q, part, nb, out);
  } else {
    hipLaunchKernelGGL((k_integrals_cells<float>), dim3(nb), dim3(256), 0, s, I, run, part);
    hipLaunchKernelGGL((k_integrals_close<float>), dim3(1), dim3(256), 0, s, I, q, part, nb, out);
  }
  return hipGetLastError();
}

}  // namespace nzk
