// cutcell.hip — the cut-cell geometry on the GPU (buildCutGeometry, cutcell.js:14-186).
//
// Every cell's 8 x 8 sub-sample volume and area, every face's 12-sample open area, the
// classification, the sliver merge, the closure wall vectors and the group sums, one thread
// per cell or face. The merge is order-independent (an eligible host has kappa >= 0.5, so it
// is never a slave and never turned solid; a sliver's choice reads only such cells), so every
// sliver decides in parallel what the reference's two sequential passes decide. Arithmetic is
// FP64 with rounding to float32 where the reference stores into a Float32Array, and this file
// is compiled without floating-point contraction, so the output is bit for bit the host
// restatement's (setup.cpp) and the reference's. Group sums are float32 additions in
// ascending cell order, as the reference's accumulation loop makes them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "nozzle.h"
#include "nz_internal.h"

namespace nzc {

struct CutParams {
  int Nx, Nr, nTab, npts, nsol, base, cut;
  double dx, xeC, twC, xBaseC, bodyC;
  double Rc, Re, xe, cavD, cavW;  // contour (m) for the radius look-up past the table's end
};
// device inputs and intermediates of one build
struct CutBufs {
  const float* Rtab;    // nTab: inner-wall radius (cells) every 1/8 cell, -1e9 beyond the exit
  const double* xs;     // npts contour polyline (m)
  const double* rs;
  const double* sol;    // nsol x (x0, x1, r0, r1), cells
  float *colMin, *colMax;
  float *V, *Am, *kap, *Wx, *Wr;
  unsigned char *isCut, *solid0, *solid;
  int* host;
  float *Ax, *Ar;       // (Nx + 1) x (Nr + 1), row stride Nx + 1
  // outputs
  unsigned char* mask;
  float *geoA, *geoG, *geoM;
  unsigned int* counts;  // nCut, nMerged
};

constexpr int SUB = 8, NC = 8, NF = 12;

// radiusAt of the contour polyline (physics.js:190-197, cavity pocket simsetup.js:155-159)
__device__ double radius_at(const CutParams& p, const CutBufs& b, double x) {
  double r;
  if (x <= 0) r = p.Rc;
  else if (x >= p.xe) r = p.Re;
  else {
    int lo = 0, hi = p.npts - 1;
    while (hi - lo > 1) {
      const int m = (lo + hi) >> 1;
      if (b.xs[m] <= x) lo = m; else hi = m;
    }
    const double t = (x - b.xs[lo]) / (b.xs[hi] - b.xs[lo]);
    r = b.rs[lo] + t * (b.rs[hi] - b.rs[lo]);
  }
  if (p.cavD > 0 && x < p.cavW) r = r + p.cavD;
  return r;
}

// RinAt (cutcell.js:28-35): the table interpolated linearly, the contour itself in the last
// interval before the exit
__device__ double rin_at(const CutParams& p, const CutBufs& b, double xc) {
  if (xc > p.xeC) return -1e9;
  const double f = fmin(fmax(xc * SUB, 0.0), (double)(p.nTab - 1));
  const int k = (int)fmin(floor(f), (double)(p.nTab - 2));
  const double t = f - k;
  const double a = b.Rtab[k], c = b.Rtab[k + 1];
  if (a < -1e8 || c < -1e8) return radius_at(p, b, fmin(xc * p.dx, p.xe)) / p.dx;
  return a + (c - a) * t;
}

// fluid(x, r) (cutcell.js:36-44)
__device__ bool fluid(const CutParams& p, const CutBufs& b, double xc, double rc) {
  for (int s = 0; s < p.nsol; ++s) {
    const double* q = b.sol + 4 * s;
    if (xc >= q[0] && xc <= q[1] && rc >= q[2] && rc <= q[3]) return false;
  }
  if (xc <= p.xeC) {
    const double ri = rin_at(p, b, xc);
    if (rc >= ri && rc <= ri + p.twC) return false;
    if (p.base && xc < p.xBaseC && rc >= ri && rc < p.bodyC) return false;
  }
  return true;
}

__device__ __forceinline__ bool overlaps(double a0, double a1, double b0, double b1) { return a0 < b1 && a1 > b0; }

// maybeCut (cutcell.js:53-67)
__device__ bool maybe_cut(const CutParams& p, const CutBufs& b, int i, int j) {
  if (!p.cut) return false;
  const double x0 = i, x1 = i + 1, r0 = j, r1 = j + 1;
  for (int s = 0; s < p.nsol; ++s) {
    const double* q = b.sol + 4 * s;
    if (overlaps(x0, x1, q[0] - 1e-6, q[1] + 1e-6) && overlaps(r0, r1, q[2] - 1e-6, q[3] + 1e-6)) return true;
  }
  if ((double)b.colMax[i] < -1e8) return false;
  const double mn = b.colMin[i], mx = b.colMax[i];
  if (overlaps(r0, r1, mn - 1e-6, mx + 1e-6)) return true;
  if (overlaps(r0, r1, mn + p.twC - 1e-6, mx + p.twC + 1e-6)) return true;
  if (p.xeC >= x0 && p.xeC <= x1 && overlaps(r0, r1, mn, mx + p.twC)) return true;
  if (p.base) {
    if (p.xBaseC >= x0 && p.xBaseC <= x1 && overlaps(r0, r1, mn, p.bodyC)) return true;
    if (x0 < p.xBaseC && p.bodyC >= r0 && p.bodyC <= r1) return true;
  }
  return false;
}

// per column: radius range of the inner wall over [i, i + 1] (cutcell.js:46-51)
__global__ void k_cc_columns(CutParams p, CutBufs b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.Nx) return;
  double mn = 1e9, mx = -1e9;
  for (int a = 0; a <= SUB; ++a) {
    const double r = b.Rtab[i * SUB + a];
    if (r > -1e8) { mn = fmin(mn, r); mx = fmax(mx, r); }
  }
  b.colMin[i] = (float)mn;
  b.colMax[i] = (float)mx;
}

// volume and meridional area (cutcell.js:70-82): 8 x 8 sub-samples where the wall may pass
__global__ void k_cc_cells(CutParams p, CutBufs b) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)p.Nx * p.Nr) return;
  const int i = (int)(k % p.Nx), j = (int)(k / p.Nx);
  float am32 = 0.f, v32 = 0.f;
  unsigned char c = 0;
  if (maybe_cut(p, b, i, j)) {
    double am = 0, v = 0;
    for (int bb = 0; bb < NC; ++bb) {
      const double rc = j + (bb + 0.5) / NC;
      for (int a = 0; a < NC; ++a)
        if (fluid(p, b, i + (a + 0.5) / NC, rc)) { am += 1; v += rc; }
    }
    am32 = (float)(am / (NC * NC));
    v32 = (float)(v / (NC * NC));
    c = (am32 > 0 && am32 < 1) ? 1 : 0;
  } else if (fluid(p, b, i + 0.5, j + 0.5)) {
    am32 = 1.f;
    v32 = (float)(j + 0.5);
  }
  b.Am[k] = am32;
  b.V[k] = v32;
  b.isCut[k] = c;
}

__device__ __forceinline__ int cut_at(const CutParams& p, const CutBufs& b, int i, int j) {
  return (i >= 0 && i < p.Nx && j >= 0 && j < p.Nr) ? b.isCut[(long long)j * p.Nx + i] : 0;
}

// open face areas (cutcell.js:84-104), over the whole (Nx + 1) x (Nr + 1) lattice of each
// array: work items [0, L) are x faces (row Nr holds none and stays 0), [L, 2L) r faces
// (column Nx holds none)
__global__ void k_cc_faces(CutParams p, CutBufs b) {
  const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int W1 = p.Nx + 1;
  const long long L = (long long)W1 * (p.Nr + 1);
  if (f >= 2 * L) return;
  const bool xf = f < L;
  const long long g = xf ? f : f - L;
  const int i = (int)(g % W1), j = (int)(g / W1);
  double a = 0;
  if (xf && j < p.Nr) {
    if (p.cut && (cut_at(p, b, i - 1, j) || cut_at(p, b, i, j) || maybe_cut(p, b, min(i, p.Nx - 1), j) ||
                  (i > 0 && maybe_cut(p, b, i - 1, j)))) {
      for (int bb = 0; bb < NF; ++bb) {
        const double rc = j + (bb + 0.5) / NF;
        if (fluid(p, b, i, rc)) a += rc;
      }
      a /= NF;
    } else {
      a = fluid(p, b, i, j + 0.5) ? j + 0.5 : 0;
    }
  } else if (!xf && i < p.Nx) {
    if (p.cut && (cut_at(p, b, i, j - 1) || cut_at(p, b, i, j) || (j < p.Nr && maybe_cut(p, b, i, j)) ||
                  (j > 0 && maybe_cut(p, b, i, j - 1)))) {
      double fl = 0;
      for (int bb = 0; bb < NF; ++bb)
        if (fluid(p, b, i + (bb + 0.5) / NF, j)) fl++;
      a = j * fl / NF;
    } else {
      a = fluid(p, b, i + 0.5, j) ? (double)j : 0;
    }
  }
  (xf ? b.Ax : b.Ar)[g] = (float)a;
}

// volume fraction and the solid test kappa < 0.02 (cutcell.js:106-113)
__global__ void k_cc_classify(CutParams p, CutBufs b) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)p.Nx * p.Nr) return;
  const int j = (int)(k / p.Nx);
  const float kp = (float)((double)b.V[k] / (j + 0.5));
  b.kap[k] = kp;
  b.solid0[k] = (double)kp < 0.02 ? 1 : 0;
}

// the faces of solid cells closed (cutcell.js:114-120), per face from the two cells it bounds
__global__ void k_cc_close_faces(CutParams p, CutBufs b, const unsigned char* __restrict__ solid) {
  const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int W1 = p.Nx + 1;
  const long long L = (long long)W1 * (p.Nr + 1);
  if (f >= 2 * L) return;
  const bool xf = f < L;
  const long long g = xf ? f : f - L;
  const int i = (int)(g % W1), j = (int)(g / W1);
  if (xf && j < p.Nr) {
    const bool s = (i > 0 && solid[(long long)j * p.Nx + i - 1]) || (i < p.Nx && solid[(long long)j * p.Nx + i]);
    if (s) b.Ax[g] = 0.f;
  } else if (!xf && i < p.Nx) {
    const bool s = (j > 0 && solid[(long long)(j - 1) * p.Nx + i]) || (j < p.Nr && solid[(long long)j * p.Nx + i]);
    if (s) b.Ar[g] = 0.f;
  }
}

// sliver merge (cutcell.js:121-142): a cell with kappa < 0.5 joins the 4-neighbour with
// kappa >= 0.5 through the largest normalised open face (> 0.02), else becomes solid. Only
// cells with kappa >= 0.5 are candidates; they are never slaves nor turned solid, so this
// reads nothing another thread writes.
__global__ void k_cc_merge(CutParams p, CutBufs b) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)p.Nx * p.Nr) return;
  const int i = (int)(k % p.Nx), j = (int)(k / p.Nx), W1 = p.Nx + 1;
  const double KMERGE = 0.5;
  int h = -1;
  unsigned char s = b.solid0[k];
  if (!s && (double)b.kap[k] < KMERGE) {
    const int ci[4] = {i - 1, i + 1, i, i};
    const int cj[4] = {j, j, j - 1, j + 1};
    const double ca[4] = {(double)b.Ax[(long long)j * W1 + i] / (j + 0.5), (double)b.Ax[(long long)j * W1 + i + 1] / (j + 0.5),
                          j > 0 ? (double)b.Ar[(long long)j * W1 + i] / j : 0.0,
                          (double)b.Ar[(long long)(j + 1) * W1 + i] / (j + 1)};
    int best = -1;
    double bestA = 0;
    for (int c = 0; c < 4; ++c) {
      const int ii = ci[c], jj = cj[c];
      if (ii < 0 || ii >= p.Nx || jj < 0 || jj >= p.Nr) continue;
      const long long kk = (long long)jj * p.Nx + ii;
      if (b.solid0[kk] || (double)b.kap[kk] < KMERGE) continue;
      if (ca[c] > bestA + 1e-9) { bestA = ca[c]; best = (int)kk; }
    }
    if (best >= 0 && bestA > 0.02) h = best;
    else s = 1;
  }
  b.host[k] = h;
  b.solid[k] = s;
}

// wall vectors by closure (cutcell.js:154-163): evaluated in double, rounded once
__global__ void k_cc_walls(CutParams p, CutBufs b) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)p.Nx * p.Nr) return;
  const int i = (int)(k % p.Nx), j = (int)(k / p.Nx), W1 = p.Nx + 1;
  float wx = 0.f, wr = 0.f;
  if (!b.solid[k]) {
    wx = (float)((double)b.Ax[(long long)j * W1 + i] - (double)b.Ax[(long long)j * W1 + i + 1]);
    wr = (float)((double)b.Am[k] - ((double)b.Ar[(long long)(j + 1) * W1 + i] - (double)b.Ar[(long long)j * W1 + i]));
    if (fabs((double)wx) < 1e-6) wx = 0.f;
    if (fabs((double)wr) < 1e-6) wr = 0.f;
  }
  b.Wx[k] = wx;
  b.Wr[k] = wr;
}

// group sums, slave bits, geoG / geoM (cutcell.js:164-184): each cell gathers its host's
// members in ascending cell order (south, west, host, east, north) and sums in float32
__global__ void k_cc_groups(CutParams p, CutBufs b) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int Nx = p.Nx, Nr = p.Nr;
  if (k >= (long long)Nx * Nr) return;
  const int i = (int)(k % Nx), j = (int)(k / Nx);
  float* G = b.geoG + k * 4;
  float* M = b.geoM + k * 4;
  if (b.solid[k]) {
    G[0] = 1.f; G[1] = 0.f; G[2] = 0.f; G[3] = 0.f;
    M[0] = 0.f; M[1] = 0.f; M[2] = 0.f; M[3] = 0.f;
    b.mask[k] = 1;
    return;
  }
  b.mask[k] = 0;
  const long long h = b.host[k] >= 0 ? (long long)b.host[k] : k;
  const int hi = (int)(h % Nx), hj = (int)(h / Nx);
  float gV = 0.f, gA = 0.f, gX = 0.f, gR = 0.f;
  const long long cand[5] = {h - Nx, h - 1, h, h + 1, h + Nx};
  const bool valid[5] = {hj > 0, hi > 0, true, hi < Nx - 1, hj < Nr - 1};
  for (int c = 0; c < 5; ++c) {
    if (!valid[c]) continue;
    const long long m = cand[c];
    if (b.solid[m] || !(m == h || b.host[m] == (int)h)) continue;
    gV += b.V[m]; gA += b.Am[m]; gX += b.Wx[m]; gR += b.Wr[m];
  }
  // this cell's own slave bits (a host's; 0 for a slave or a lone cell)
  unsigned int sm = 0;
  if (b.host[k] < 0) {
    const long long nb[4] = {k - 1, k + 1, k - Nx, k + Nx};
    const bool ok[4] = {i > 0, i < Nx - 1, j > 0, j < Nr - 1};
    for (int c = 0; c < 4; ++c)
      if (ok[c] && !b.solid[nb[c]] && b.host[nb[c]] == (int)k) sm |= 1u << c;   // 1 W, 2 E, 4 S, 8 N of the host
  }
  G[0] = gV; G[1] = gA; G[2] = gX; G[3] = gR;
  M[0] = (float)(hi - i); M[1] = (float)(hj - j); M[2] = (float)sm; M[3] = b.kap[k];
  if ((double)b.kap[k] < 0.98) atomicAdd(&b.counts[0], 1u);
  if (b.host[k] >= 0) atomicAdd(&b.counts[1], 1u);
}

// geoA (cutcell.js:144-153): V, Am and the face areas, negated next to a partial cell
__device__ __forceinline__ bool partial_cell(const CutParams& p, const CutBufs& b, int i, int j) {
  if (!(i >= 0 && i < p.Nx && j >= 0 && j < p.Nr)) return false;
  const long long k = (long long)j * p.Nx + i;
  return !b.solid[k] && ((double)b.kap[k] < 0.98 || b.host[k] >= 0);
}
__global__ void k_cc_geoA(CutParams p, CutBufs b) {
  const long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int W1 = p.Nx + 1;
  if (f >= (long long)W1 * (p.Nr + 1)) return;
  const int i = (int)(f % W1), j = (int)(f / W1);
  float* q = b.geoA + f * 4;
  float v = 0.f, am = 0.f, ax = 0.f, ar = 0.f;
  if (i < p.Nx && j < p.Nr) {
    const long long k = (long long)j * p.Nx + i;
    v = b.V[k];
    am = b.Am[k];
  }
  if (j < p.Nr) {
    const float a = b.Ax[f];
    ax = (partial_cell(p, b, i - 1, j) || partial_cell(p, b, i, j)) ? -a : a;
  }
  if (i < p.Nx) {
    const float a = b.Ar[f];
    ar = (partial_cell(p, b, i, j - 1) || partial_cell(p, b, i, j)) ? -a : a;
  }
  q[0] = v; q[1] = am; q[2] = ax; q[3] = ar;
}

// ------------------------------------------------------------------------------------
// host side

// device bytes of a build's inputs and intermediates, laid out back to back at 256-byte
// boundaries; `carve` hands out the pieces of one region in this order
struct Layout {
  size_t off = 0;
  size_t take(size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) / 256 * 256;
    return o;
  }
};
struct Sizes {
  size_t cells, lattice, nTab, npts, nsol;
};
size_t layout_bytes(const Sizes& z, bool withOutputs) {
  Layout L;
  L.take(z.nTab * 4); L.take(z.npts * 8); L.take(z.npts * 8); L.take(std::max<size_t>(z.nsol, 1) * 32);
  const size_t cols = z.nTab / SUB + 1;
  L.take(cols * 4); L.take(cols * 4);
  for (int q = 0; q < 5; ++q) L.take(z.cells * 4);       // V, Am, kap, Wx, Wr
  for (int q = 0; q < 3; ++q) L.take(z.cells);           // isCut, solid0, solid
  L.take(z.cells * 4);                                   // host
  L.take(z.lattice * 4); L.take(z.lattice * 4);          // Ax, Ar
  L.take(16);                                            // counts
  if (withOutputs) { L.take(z.cells); L.take(z.lattice * 16); L.take(z.cells * 16); L.take(z.cells * 16); }
  return L.off;
}
// pointers into `base` (inputs: Rtab, xs, rs, sol first; outputs last when withOutputs)
void carve(char* base, const Sizes& z, bool withOutputs, CutBufs& b, float** Rtab, double** xs, double** rs, double** sol) {
  Layout L;
  *Rtab = reinterpret_cast<float*>(base + L.take(z.nTab * 4));
  *xs = reinterpret_cast<double*>(base + L.take(z.npts * 8));
  *rs = reinterpret_cast<double*>(base + L.take(z.npts * 8));
  *sol = reinterpret_cast<double*>(base + L.take(std::max<size_t>(z.nsol, 1) * 32));
  b.Rtab = *Rtab; b.xs = *xs; b.rs = *rs; b.sol = *sol;
  const size_t cols = z.nTab / SUB + 1;
  b.colMin = reinterpret_cast<float*>(base + L.take(cols * 4));
  b.colMax = reinterpret_cast<float*>(base + L.take(cols * 4));
  float** f5[5] = {&b.V, &b.Am, &b.kap, &b.Wx, &b.Wr};
  for (auto* f : f5) *f = reinterpret_cast<float*>(base + L.take(z.cells * 4));
  b.isCut = reinterpret_cast<unsigned char*>(base + L.take(z.cells));
  b.solid0 = reinterpret_cast<unsigned char*>(base + L.take(z.cells));
  b.solid = reinterpret_cast<unsigned char*>(base + L.take(z.cells));
  b.host = reinterpret_cast<int*>(base + L.take(z.cells * 4));
  b.Ax = reinterpret_cast<float*>(base + L.take(z.lattice * 4));
  b.Ar = reinterpret_cast<float*>(base + L.take(z.lattice * 4));
  b.counts = reinterpret_cast<unsigned int*>(base + L.take(16));
  if (withOutputs) {
    b.mask = reinterpret_cast<unsigned char*>(base + L.take(z.cells));
    b.geoA = reinterpret_cast<float*>(base + L.take(z.lattice * 16));
    b.geoG = reinterpret_cast<float*>(base + L.take(z.cells * 16));
    b.geoM = reinterpret_cast<float*>(base + L.take(z.cells * 16));
  }
}

__global__ void k_cc_zero_counts(unsigned int* c) {
  if (threadIdx.x < 4) c[threadIdx.x] = 0u;
}

// the kernels of one build on stream s (inputs already on the device)
hipError_t enqueue_build(const CutParams& p, const CutBufs& b, hipStream_t s) {
  const long long cells = (long long)p.Nx * p.Nr;
  const long long lattice = (long long)(p.Nx + 1) * (p.Nr + 1);
  const long long faces = 2 * lattice;
  auto grid = [](long long n) { return dim3((unsigned)((n + 255) / 256)); };
  hipLaunchKernelGGL(k_cc_zero_counts, dim3(1), dim3(64), 0, s, b.counts);
  hipLaunchKernelGGL(k_cc_columns, grid(p.Nx), dim3(256), 0, s, p, b);
  hipLaunchKernelGGL(k_cc_cells, grid(cells), dim3(256), 0, s, p, b);
  hipLaunchKernelGGL(k_cc_faces, grid(faces), dim3(256), 0, s, p, b);
  hipLaunchKernelGGL(k_cc_classify, grid(cells), dim3(256), 0, s, p, b);
  hipLaunchKernelGGL(k_cc_close_faces, grid(faces), dim3(256), 0, s, p, b, (const unsigned char*)b.solid0);
  hipLaunchKernelGGL(k_cc_merge, grid(cells), dim3(256), 0, s, p, b);
  hipLaunchKernelGGL(k_cc_close_faces, grid(faces), dim3(256), 0, s, p, b, (const unsigned char*)b.solid);
  hipLaunchKernelGGL(k_cc_walls, grid(cells), dim3(256), 0, s, p, b);
  hipLaunchKernelGGL(k_cc_groups, grid(cells), dim3(256), 0, s, p, b);
  hipLaunchKernelGGL(k_cc_geoA, grid(lattice), dim3(256), 0, s, p, b);
  return hipGetLastError();
}

int check_input(const nz_cutcell_in* in, const char* what, CutParams& p, Sizes& z) {
  if (!in) return nz_fail(-1, std::string(what) + ": null input");
  if (in->Nx < 2 || in->Nr < 2 || (long long)in->Nx * in->Nr > (1LL << 26))
    return nz_fail(-1, std::string(what) + ": grid size out of range");
  if (!in->Rtab || in->npts < 2 || !in->xs || !in->rs || in->nsolids < 0 || (in->nsolids > 0 && !in->solids))
    return nz_fail(-1, std::string(what) + ": missing contour table, polyline or solids");
  if (!(in->dx > 0)) return nz_fail(-1, std::string(what) + ": cell size must be positive");
  p.Nx = in->Nx; p.Nr = in->Nr; p.nTab = in->Nx * SUB + 1; p.npts = in->npts; p.nsol = in->nsolids;
  p.base = in->base ? 1 : 0; p.cut = in->stair ? 0 : 1;
  p.dx = in->dx;
  p.xeC = in->xe / in->dx; p.twC = in->tw / in->dx; p.xBaseC = in->xBase / in->dx; p.bodyC = in->bodyR / in->dx;
  p.Rc = in->Rc; p.Re = in->Re; p.xe = in->xe; p.cavD = in->cavD; p.cavW = in->cavW;
  z.cells = (size_t)in->Nx * in->Nr;
  z.lattice = (size_t)(in->Nx + 1) * (in->Nr + 1);
  z.nTab = (size_t)p.nTab;
  z.npts = (size_t)in->npts;
  z.nsol = (size_t)in->nsolids;
  return 0;
}
// the solids in cells, as cutcell.js:35 divides them
std::vector<double> solids_cells(const nz_cutcell_in* in) {
  std::vector<double> v((size_t)std::max(in->nsolids, 1) * 4, 0.0);
  for (int s = 0; s < in->nsolids * 4; ++s) v[(size_t)s] = in->solids[s] / in->dx;
  return v;
}

}  // namespace nzc

#define HIPCK(call, what)                                                                           \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) { rc = nz_fail(-10, std::string(what) + ": " + hipGetErrorString(e_)); goto done; } \
  } while (0)

extern "C" int nz_cutcell_build(const nz_cutcell_in* in, uint8_t* mask, float* geoA, float* geoG, float* geoM,
                                float* kappa, float* Ax, float* Ar, int32_t* counts) {
  using namespace nzc;
  CutParams p{};
  Sizes z{};
  if (int r = check_input(in, "nz_cutcell_build", p, z)) return r;
  if (!mask || !geoA || !geoG || !geoM) return nz_fail(-1, "nz_cutcell_build: null output");
  int rc = 0;
  char* base = nullptr;
  hipStream_t s = nullptr;
  CutBufs b{};
  float* dR = nullptr;
  double *dxs = nullptr, *drs = nullptr, *dsol = nullptr;
  const std::vector<double> sol = solids_cells(in);
  unsigned int cnt[4] = {0, 0, 0, 0};
  HIPCK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "nz_cutcell_build: stream");
  HIPCK(hipMalloc(&base, layout_bytes(z, true)), "nz_cutcell_build: device buffers");
  carve(base, z, true, b, &dR, &dxs, &drs, &dsol);
  HIPCK(hipMemcpyAsync(dR, in->Rtab, z.nTab * 4, hipMemcpyHostToDevice, s), "nz_cutcell_build: upload");
  HIPCK(hipMemcpyAsync(dxs, in->xs, z.npts * 8, hipMemcpyHostToDevice, s), "nz_cutcell_build: upload");
  HIPCK(hipMemcpyAsync(drs, in->rs, z.npts * 8, hipMemcpyHostToDevice, s), "nz_cutcell_build: upload");
  HIPCK(hipMemcpyAsync(dsol, sol.data(), sol.size() * 8, hipMemcpyHostToDevice, s), "nz_cutcell_build: upload");
  HIPCK(enqueue_build(p, b, s), "nz_cutcell_build: kernels");
  HIPCK(hipMemcpyAsync(mask, b.mask, z.cells, hipMemcpyDeviceToHost, s), "nz_cutcell_build: read-back");
  HIPCK(hipMemcpyAsync(geoA, b.geoA, z.lattice * 16, hipMemcpyDeviceToHost, s), "nz_cutcell_build: read-back");
  HIPCK(hipMemcpyAsync(geoG, b.geoG, z.cells * 16, hipMemcpyDeviceToHost, s), "nz_cutcell_build: read-back");
  HIPCK(hipMemcpyAsync(geoM, b.geoM, z.cells * 16, hipMemcpyDeviceToHost, s), "nz_cutcell_build: read-back");
  if (kappa) HIPCK(hipMemcpyAsync(kappa, b.kap, z.cells * 4, hipMemcpyDeviceToHost, s), "nz_cutcell_build: read-back");
  if (Ax) HIPCK(hipMemcpyAsync(Ax, b.Ax, z.lattice * 4, hipMemcpyDeviceToHost, s), "nz_cutcell_build: read-back");
  if (Ar) HIPCK(hipMemcpyAsync(Ar, b.Ar, z.lattice * 4, hipMemcpyDeviceToHost, s), "nz_cutcell_build: read-back");
  HIPCK(hipMemcpyAsync(cnt, b.counts, 16, hipMemcpyDeviceToHost, s), "nz_cutcell_build: read-back");
  HIPCK(hipStreamSynchronize(s), "nz_cutcell_build: sync");
  if (counts) { counts[0] = (int32_t)cnt[0]; counts[1] = (int32_t)cnt[1]; }
done:
  if (base) (void)hipFree(base);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}

// ------------------------------------------------------------------------------------
// The same build without any allocation, for callers that cannot allocate device memory
// (code under graph capture, serving loops that admit new geometries on the fly). Inputs,
// intermediates and outputs live in one file-scope device array per GPU, held until the
// process ends; grids whose layout does not fit it are refused. Every call on a GPU shares
// that GPU's array: calls that overlap on one GPU race, and each call clears what must start
// at zero (the counters) itself.
namespace nzc {
constexpr size_t kNoAllocScratchBytes = (size_t)36 << 20;   // 36 MiB: 1024 x 384 cells fit
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// the array's device address, looked up once per device (the host-side name of a
// __device__ array is a stand-in, not a device address)
constexpr int kMaxDevices = 64;
static std::mutex g_scratchMu;
static void* g_scratchPtr[kMaxDevices] = {};

int scratch_address(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return nz_fail(-10, "nz_cutcell_build_noalloc: no current HIP device");
  std::lock_guard<std::mutex> lk(g_scratchMu);
  if (!g_scratchPtr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return nz_fail(-10, std::string("nz_cutcell_build_noalloc: scratch address: ") + hipGetErrorString(e));
    g_scratchPtr[dev] = p;
  }
  *out = g_scratchPtr[dev];
  return 0;
}
}  // namespace nzc

extern "C" int nz_cutcell_build_noalloc(const nz_cutcell_in* in, uint8_t* mask, float* geoA, float* geoG, float* geoM,
                                        float* kappa, float* Ax, float* Ar, int32_t* counts) {
  using namespace nzc;
  CutParams p{};
  Sizes z{};
  if (int r = check_input(in, "nz_cutcell_build_noalloc", p, z)) return r;
  if (!mask || !geoA || !geoG || !geoM) return nz_fail(-1, "nz_cutcell_build_noalloc: null output");
  if (in->nsolids > 256) return nz_fail(-1, "nz_cutcell_build_noalloc: more than 256 solids");
  const size_t need = layout_bytes(z, true);
  if (need > kNoAllocScratchBytes)
    return nz_fail(-8, "nz_cutcell_build_noalloc: a " + std::to_string(in->Nx) + " x " + std::to_string(in->Nr) +
                           " grid needs " + std::to_string(need) + " bytes of device scratch, more than its " +
                           std::to_string(kNoAllocScratchBytes) + " (at most about 1024 x 384 cells); use nz_cutcell_build");
  void* base = nullptr;
  if (int r = scratch_address(&base)) return r;
  int rc = 0;
  CutBufs b{};
  float* dR = nullptr;
  double *dxs = nullptr, *drs = nullptr, *dsol = nullptr;
  carve(static_cast<char*>(base), z, true, b, &dR, &dxs, &drs, &dsol);
  // the solids in cells, staged in host memory the caller's call owns (no allocation: at most 256)
  double sol[256 * 4];
  for (int s = 0; s < in->nsolids * 4; ++s) sol[s] = in->solids[s] / in->dx;
  const size_t nsolBytes = (size_t)std::max(in->nsolids, 1) * 32;
  if (in->nsolids == 0) std::memset(sol, 0, 32);
  unsigned int cnt[4] = {0, 0, 0, 0};
  const hipStream_t s = nullptr;   // the null stream: no stream is created
  HIPCK(hipMemcpyAsync(dR, in->Rtab, z.nTab * 4, hipMemcpyHostToDevice, s), "nz_cutcell_build_noalloc: upload");
  HIPCK(hipMemcpyAsync(dxs, in->xs, z.npts * 8, hipMemcpyHostToDevice, s), "nz_cutcell_build_noalloc: upload");
  HIPCK(hipMemcpyAsync(drs, in->rs, z.npts * 8, hipMemcpyHostToDevice, s), "nz_cutcell_build_noalloc: upload");
  HIPCK(hipMemcpyAsync(dsol, sol, nsolBytes, hipMemcpyHostToDevice, s), "nz_cutcell_build_noalloc: upload");
  HIPCK(enqueue_build(p, b, s), "nz_cutcell_build_noalloc: kernels");
  HIPCK(hipMemcpyAsync(mask, b.mask, z.cells, hipMemcpyDeviceToHost, s), "nz_cutcell_build_noalloc: read-back");
  HIPCK(hipMemcpyAsync(geoA, b.geoA, z.lattice * 16, hipMemcpyDeviceToHost, s), "nz_cutcell_build_noalloc: read-back");
  HIPCK(hipMemcpyAsync(geoG, b.geoG, z.cells * 16, hipMemcpyDeviceToHost, s), "nz_cutcell_build_noalloc: read-back");
  HIPCK(hipMemcpyAsync(geoM, b.geoM, z.cells * 16, hipMemcpyDeviceToHost, s), "nz_cutcell_build_noalloc: read-back");
  if (kappa) HIPCK(hipMemcpyAsync(kappa, b.kap, z.cells * 4, hipMemcpyDeviceToHost, s), "nz_cutcell_build_noalloc: read-back");
  if (Ax) HIPCK(hipMemcpyAsync(Ax, b.Ax, z.lattice * 4, hipMemcpyDeviceToHost, s), "nz_cutcell_build_noalloc: read-back");
  if (Ar) HIPCK(hipMemcpyAsync(Ar, b.Ar, z.lattice * 4, hipMemcpyDeviceToHost, s), "nz_cutcell_build_noalloc: read-back");
  HIPCK(hipMemcpyAsync(cnt, b.counts, 16, hipMemcpyDeviceToHost, s), "nz_cutcell_build_noalloc: read-back");
  HIPCK(hipStreamSynchronize(s), "nz_cutcell_build_noalloc: sync");
  if (counts) { counts[0] = (int32_t)cnt[0]; counts[1] = (int32_t)cnt[1]; }
done:
  return rc;
}
