// runtime.hip — the GPUSolver / GPUSpray C ABI over HIP (include/nozzle.h).
//
// Owns every device buffer of one solver instance, converts the reference's
// texture-shaped inputs (geoA/geoG/geoM RGBA, state A/B RGBA) into the SoA
// device layout, enqueues steps on its own HIP stream without host waits
// (GPUSolver.step is asynchronous, solver.js:849-891), and serves synchronous
// and fenced asynchronous read-backs (solver.js:1016-1075).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "nozzle.h"
#include "nz_internal.h"
#include "nz_launch.h"

using nzk::Consts;
using nzk::DevInst;
using nzk::InstState;

// code word of a cell whose host is itself (offset fields hold dx+1 = dy+1 = 1)
static constexpr uint32_t kZeroOffset = (1u << nzk::C_HX_SHIFT) | (1u << nzk::C_HY_SHIFT);

#define HIPCHK(call, what)                                                                          \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) return nz_fail(-10, std::string(what) + ": " + hipGetErrorString(e_));   \
  } while (0)

struct nz_solver {
  int prec = 32;
  size_t rsz = 4;
  int device = 0;
  int Nx = 0, Nr = 0;
  long long Ns = 0;
  int TX = 64, TY = 4, tilesX = 0, tilesY = 0;
  hipStream_t stream = nullptr;
  void* U0 = nullptr;
  void* U1 = nullptr;
  float *ax = nullptr, *ar = nullptr, *lt = nullptr, *jw = nullptr, *src = nullptr, *parc = nullptr, *tab = nullptr;
  float4* gg = nullptr;
  uint32_t* code = nullptr;
  int* exStart = nullptr;    // per tile: faces of merged cells outside the host's tile
  uint2* exFaces = nullptr;
  int exMax = 0;             // most extra faces in one tile
  int* fcStart = nullptr;    // per tile: cells those faces reach beyond the 2-cell halo
  int2* fcCells = nullptr;
  int fcMax = 0;
  void* exFlux = nullptr;    // 16 values per extra face (inviscid 9 + viscous 7), state precision
  InstState* st = nullptr;
  double2* part = nullptr;   // per-tile partials of the stage-2 reductions
  DevInst* dinst = nullptr;
  unsigned* done = nullptr;
  double* dint = nullptr;
  float* wallrow = nullptr;
  float* dstage = nullptr;   // device staging (pack / unpack)
  size_t dstageCap = 0;
  float* hpin = nullptr;     // pinned host staging
  size_t hpinCap = 0;
  std::vector<uint8_t> mask;
  std::vector<float> jWall;
  bool haveGeo = false, haveState = false, haveTab = false;
  int NP = 65536;
  int64_t stepCount = 0;
  double nGroups = 0;
  DevInst h{};               // host copy of the descriptor last sent to the device
  bool instSent = false;
  // async read-back (requestAsync / pollAsync)
  bool asyncPending = false;
  hipEvent_t asyncEv = nullptr;
  int32_t asyncTag = 0;
  size_t asyncFloats = 0;
  float* asyncHost = nullptr;
  size_t asyncCap = 0;
  nz_probes probes{};
  bool haveProbes = false;
};

namespace {

template <class T>
void hfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}
void vfree(void*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void free_buffers(nz_solver* s) {
  vfree(s->U0); vfree(s->U1);
  hfree(s->ax); hfree(s->ar); hfree(s->lt); hfree(s->jw); hfree(s->src); hfree(s->parc); hfree(s->tab);
  hfree(s->gg); hfree(s->code); hfree(s->exStart); hfree(s->exFaces); hfree(s->fcStart); hfree(s->fcCells); vfree(s->exFlux); hfree(s->st); hfree(s->part); hfree(s->dinst); hfree(s->done); hfree(s->dint);
  hfree(s->wallrow); hfree(s->dstage);
  if (s->hpin) (void)hipHostFree(s->hpin);
  s->hpin = nullptr;
  s->hpinCap = s->dstageCap = 0;
  s->instSent = false;
  s->haveGeo = s->haveState = false;
}

int ensure_staging(nz_solver* s, size_t floats) {
  if (floats > s->dstageCap) {
    hfree(s->dstage);
    HIPCHK(hipMalloc(&s->dstage, floats * 4), "hipMalloc staging");
    s->dstageCap = floats;
  }
  if (floats > s->hpinCap) {
    if (s->hpin) (void)hipHostFree(s->hpin);
    s->hpin = nullptr;
    HIPCHK(hipHostMalloc(&s->hpin, floats * 4, hipHostMallocDefault), "hipHostMalloc staging");
    s->hpinCap = floats;
  }
  return 0;
}

void fill_consts(const nz_solver* s, const nz_params& P, Consts& c) {
  std::memset(&c, 0, sizeof(c));
  for (int k = 0; k < 4; ++k) {
    c.amb[k] = P.amb[k]; c.inlet[k] = P.inlet[k]; c.ambCons[k] = P.ambCons[k]; c.ambB[k] = P.ambB[k];
    c.gas[k] = P.gas[k]; c.chemK[k] = P.chemK[k]; c.fuelVec[k] = P.fuelVec[k]; c.oxVec[k] = P.oxVec[k];
  }
  for (int k = 0; k < 3; ++k) { c.res[k] = P.res[k]; c.Rs[k] = P.Rs[k]; }
  for (int k = 0; k < 2; ++k) { c.gamR[k] = P.gamR[k]; c.chamberNu[k] = P.chamberNu[k]; c.cap[k] = P.cap[k]; }
  c.pStart = P.pStart; c.ramp = P.ramp; c.RcCells = P.RcCells; c.inj = P.inj ? 1 : 0;
  c.order2 = P.order2 ? 1 : 0; c.hllBlend = P.hllBlend;
  // uBar as GPUSolver._updatePass computes it (solver.js:935)
  c.uBar[0] = P.inlet[0];
  c.uBar[1] = (float)(1.6 * (double)P.uInlet0);
  c.uBar[2] = std::max(P.ramp, 0.f);
  c.vMax = P.vMax; c.cfl = P.cfl; c.dtMax = P.dtMax; c.wFloor = P.wFloor;
  c.visc = P.visc ? 1 : 0; c.chem = P.chem ? 1 : 0; c.Tw = P.Tw;
  c.sprayMhat = P.sprayMhat; c.sprayTScale = P.sprayTScale; c.eF = P.eF; c.eO = P.eO; c.Ge = P.Ge;
  c.fan = (float)(25 * M_PI / 180);
  // residual per unit grid time, relative to the chamber density P0 / (R T0)
  c.resScale = P.res[0] > 0 ? (float)((double)P.res[2] * P.res[1] / P.res[0]) : 1.f;
  c.spray = (P.inj && s->haveTab) ? 1 : 0;
  c.NP = s->NP;
}

int send_inst(nz_solver* s) {
  HIPCHK(nzk::launch_set_inst(s->dinst, s->h, s->stream), "set instance");
  s->instSent = true;
  return 0;
}

void base_inst(nz_solver* s) {
  DevInst& I = s->h;
  I.Nx = s->Nx; I.Nr = s->Nr;
  I.tilesX = s->tilesX; I.tilesY = s->tilesY; I.tile0 = 0;
  I.pblk0 = 0; I.npblk = (s->NP + 255) / 256;
  const long long N = (long long)s->Nx * s->Nr;
  I.cblk0 = 0; I.ncblk = (int)std::min<long long>((N + 255) / 256, 1024);
  I.Ns = s->Ns;
  I.U0 = s->U0; I.U1 = s->U1;
  I.ax = s->ax; I.ar = s->ar; I.gg = s->gg; I.code = s->code; I.lt = s->lt;
  I.exStart = s->exStart; I.exFaces = s->exFaces;
  I.fcStart = s->fcStart; I.fcCells = s->fcCells; I.exFlux = s->exFlux;
  I.src = s->src; I.parc = s->parc; I.tab = s->tab; I.st = s->st; I.part = s->part;
  I.c.NP = s->NP;
}

int check_solver(nz_solver* s, const char* what) {
  if (!s) return nz_fail(-1, std::string(what) + ": null solver");
  if (s->Nx <= 0) return nz_fail(-4, std::string(what) + ": resize() has not been called");
  return 0;
}

int alloc_spray(nz_solver* s) {
  hfree(s->parc);
  HIPCHK(hipMalloc(&s->parc, (size_t)s->NP * 8 * 4), "hipMalloc parcels");
  HIPCHK(nzk::launch_spray_reset(s->parc, s->NP, s->src, s->Ns, s->stream), "spray reset");
  base_inst(s);
  return send_inst(s);
}

}  // namespace

extern "C" {

int nz_device_available(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, 0) != hipSuccess) return 0;
  return std::strncmp(p.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

int nz_create_ex(nz_solver** out, int precision) {
  if (!out) return nz_fail(-1, "nz_create: null output");
  *out = nullptr;
  if (precision != 32 && precision != 64) return nz_fail(-1, "nz_create: precision must be 32 or 64");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return nz_fail(-5, "nz_create: no HIP device available");
  nz_solver* s = new nz_solver();
  s->prec = precision;
  s->rsz = precision == 64 ? 8 : 4;
  (void)hipGetDevice(&s->device);
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return nz_fail(-10, "nz_create: hipStreamCreate failed");
  }
  if (hipEventCreateWithFlags(&s->asyncEv, hipEventDisableTiming) != hipSuccess) {
    (void)hipStreamDestroy(s->stream);
    delete s;
    return nz_fail(-10, "nz_create: hipEventCreate failed");
  }
  *out = s;
  return 0;
}

int nz_create(nz_solver** out) { return nz_create_ex(out, 32); }

int nz_destroy(nz_solver* s) {
  if (!s) return 0;
  (void)hipStreamSynchronize(s->stream);
  free_buffers(s);
  if (s->asyncHost) (void)hipHostFree(s->asyncHost);
  (void)hipEventDestroy(s->asyncEv);
  (void)hipStreamDestroy(s->stream);
  delete s;
  return 0;
}

int nz_resize(nz_solver* s, int Nx, int Nr) {
  if (!s) return nz_fail(-1, "nz_resize: null solver");
  if (Nx < 2 || Nr < 2 || (long long)Nx * Nr > (1LL << 28)) return nz_fail(-1, "nz_resize: grid size out of range");
  nz_read_async_cancel(s);
  HIPCHK(hipStreamSynchronize(s->stream), "nz_resize: sync");
  free_buffers(s);
  s->Nx = Nx; s->Nr = Nr;
  const long long N = (long long)Nx * Nr;
  s->Ns = (N + 63) / 64 * 64;
  nzk::tile_shape(s->prec == 64, &s->TX, &s->TY);
  s->tilesX = (Nx + s->TX - 1) / s->TX;
  s->tilesY = (Nr + s->TY - 1) / s->TY;
  const size_t R = s->rsz;
  HIPCHK(hipMalloc(&s->U0, (size_t)s->Ns * 8 * R), "hipMalloc state");
  HIPCHK(hipMalloc(&s->U1, (size_t)s->Ns * 8 * R), "hipMalloc state");
  HIPCHK(hipMemset(s->U0, 0, (size_t)s->Ns * 8 * R), "memset");
  HIPCHK(hipMemset(s->U1, 0, (size_t)s->Ns * 8 * R), "memset");
  HIPCHK(hipMalloc(&s->ax, (size_t)Nr * (Nx + 1) * 4), "hipMalloc ax");
  HIPCHK(hipMalloc(&s->ar, (size_t)(Nr + 1) * Nx * 4), "hipMalloc ar");
  HIPCHK(hipMalloc(&s->gg, (size_t)N * 16), "hipMalloc gg");
  HIPCHK(hipMalloc(&s->code, (size_t)N * 4), "hipMalloc code");
  HIPCHK(hipMalloc(&s->exStart, ((size_t)s->tilesX * s->tilesY + 1) * sizeof(int)), "hipMalloc extra starts");
  HIPCHK(hipMemset(s->exStart, 0, ((size_t)s->tilesX * s->tilesY + 1) * sizeof(int)), "memset extra starts");
  HIPCHK(hipMalloc(&s->exFaces, sizeof(uint2)), "hipMalloc extra faces");
  HIPCHK(hipMalloc(&s->fcStart, ((size_t)s->tilesX * s->tilesY + 1) * sizeof(int)), "hipMalloc far starts");
  HIPCHK(hipMemset(s->fcStart, 0, ((size_t)s->tilesX * s->tilesY + 1) * sizeof(int)), "memset far starts");
  HIPCHK(hipMalloc(&s->fcCells, sizeof(int2)), "hipMalloc far cells");
  HIPCHK(hipMalloc(&s->exFlux, 16 * 8), "hipMalloc extra fluxes");
  HIPCHK(hipMalloc(&s->lt, (size_t)N * 4), "hipMalloc lt");
  HIPCHK(hipMalloc(&s->jw, (size_t)Nx * 4 * 4), "hipMalloc jWall");
  HIPCHK(hipMalloc(&s->src, (size_t)s->Ns * 4 * 4), "hipMalloc src");
  HIPCHK(hipMalloc(&s->tab, 256 * 3 * 4 * 4), "hipMalloc tab");
  HIPCHK(hipMalloc(&s->st, sizeof(InstState)), "hipMalloc state info");
  HIPCHK(hipMalloc(&s->part, (size_t)s->tilesX * s->tilesY * sizeof(double2)), "hipMalloc partials");
  HIPCHK(hipMalloc(&s->dinst, sizeof(DevInst)), "hipMalloc instance");
  HIPCHK(hipMalloc(&s->done, 64), "hipMalloc counter");
  HIPCHK(hipMalloc(&s->dint, 16 * 8), "hipMalloc integrals");
  HIPCHK(hipMalloc(&s->wallrow, (size_t)Nx * 4 * 4), "hipMalloc wall row");
  // a mask of all-fluid cells and empty geometry until setGeometry
  HIPCHK(hipMemset(s->ax, 0, (size_t)Nr * (Nx + 1) * 4), "memset");
  HIPCHK(hipMemset(s->ar, 0, (size_t)(Nr + 1) * Nx * 4), "memset");
  HIPCHK(hipMemset(s->gg, 0, (size_t)N * 16), "memset");
  HIPCHK(hipMemsetD32((hipDeviceptr_t)s->code, (int)kZeroOffset, (size_t)N), "memset code");
  HIPCHK(hipMemset(s->lt, 0, (size_t)N * 4), "memset");
  HIPCHK(hipMemset(s->jw, 0, (size_t)Nx * 16), "memset");
  HIPCHK(hipMemset(s->src, 0, (size_t)s->Ns * 16), "memset");
  HIPCHK(hipMemset(s->tab, 0, 256 * 3 * 16), "memset");
  HIPCHK(hipMemset(s->done, 0, 64), "memset");
  HIPCHK(hipMemset(s->wallrow, 0, (size_t)Nx * 16), "memset");
  InstState is{};
  is.seed = 1;  // GPUSpray starts at seed 1 (spray.js:221)
  HIPCHK(hipMemcpy(s->st, &is, sizeof(is), hipMemcpyHostToDevice), "init state info");
  s->mask.assign((size_t)N, 0);
  s->jWall.assign((size_t)Nx * 4, 0.f);
  s->haveTab = false;
  s->stepCount = 0;
  std::memset(&s->h, 0, sizeof(s->h));
  base_inst(s);
  if (int rc = alloc_spray(s)) return rc;
  HIPCHK(hipStreamSynchronize(s->stream), "nz_resize");
  return 0;
}

int nz_get_size(const nz_solver* s, int* Nx, int* Nr) {
  if (!s || !Nx || !Nr) return nz_fail(-1, "nz_get_size: null argument");
  *Nx = s->Nx;
  *Nr = s->Nr;
  return 0;
}

// For the tile shape in use, every merged cell whose host lies in another tile
// contributes its three faces that are not faces of the host's tile; the stage
// kernels evaluate those as extra work items of the host's workgroup.
static int build_extra_faces(nz_solver* s, const std::vector<uint32_t>& code) {
  const int Nx = s->Nx, Nr = s->Nr, TX = s->TX, TY = s->TY;
  const int ntiles = s->tilesX * s->tilesY;
  std::vector<std::vector<uint2>> lists((size_t)ntiles);
  auto inTileFace = [&](bool dirR, int fi, int fj, int i0, int j0) {
    const int fx = fi - i0, fy = fj - j0;
    return dirR ? (fx >= 0 && fx < TX && fy >= 0 && fy <= TY) : (fx >= 0 && fx <= TX && fy >= 0 && fy < TY);
  };
  for (int j = 0; j < Nr; ++j)
    for (int i = 0; i < Nx; ++i) {
      const uint32_t c = code[(size_t)j * Nx + i];
      if ((c & nzk::C_SOLID) || (c & nzk::C_IS_SLAVE)) continue;
      const uint32_t sl = (c >> nzk::C_SLAVE_SHIFT) & 15u;
      if (!sl) continue;
      const int tx = i / TX, ty = j / TY, i0 = tx * TX, j0 = ty * TY;
      std::vector<uint2>& L = lists[(size_t)ty * s->tilesX + tx];
      const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
      for (int b = 0; b < 4; ++b) {
        if (!(sl & (1u << b))) continue;
        const int si = i + di[b], sj = j + dj[b];
        if (si >= i0 && si < i0 + TX && sj >= j0 && sj < j0 + TY) continue;
        const bool fdir[4] = {false, false, true, true};
        const int fi[4] = {si, si + 1, si, si}, fj[4] = {sj, sj, sj, sj + 1};
        for (int f = 0; f < 4; ++f) {
          if (inTileFace(fdir[f], fi[f], fj[f], i0, j0)) continue;
          const uint2 e = make_uint2((uint32_t)fi[f] | (fdir[f] ? 0x80000000u : 0u), (uint32_t)fj[f]);
          bool dup = false;
          for (const uint2& q : L) dup |= (q.x == e.x && q.y == e.y);
          if (!dup) L.push_back(e);
        }
      }
    }
  // cells the extra faces' stencils reach outside the tile's 2-cell halo: inviscid
  // stencil + tangential pressures, and the 5-point gradient stencils of both cells
  // of a viscous face when a cell lies outside the 1-cell ring of the gradient tile
  std::vector<std::vector<int2>> far((size_t)ntiles);
  for (int t = 0; t < ntiles; ++t) {
    const int i0 = (t % s->tilesX) * TX, j0 = (t / s->tilesX) * TY;
    auto inHalo = [&](int i, int j) { return i >= i0 - 2 && i <= i0 + TX + 1 && j >= j0 - 2 && j <= j0 + TY + 1; };
    auto inRing = [&](int i, int j) { return i >= i0 - 1 && i <= i0 + TX && j >= j0 - 1 && j <= j0 + TY; };
    std::vector<int2>& F = far[t];
    auto need = [&](int i, int j) {
      if (inHalo(i, j)) return;
      for (const int2& q : F)
        if (q.x == i && q.y == j) return;
      F.push_back(make_int2(i, j));
    };
    for (const uint2& e : lists[t]) {
      const bool dirR = (e.x >> 31) != 0;
      const int fi = (int)(e.x & 0x7fffffffu), fj = (int)e.y;
      const int dx = dirR ? 0 : 1, dy = dirR ? 1 : 0, tx = dy, ty = dx;
      need(fi - 2 * dx, fj - 2 * dy); need(fi - dx, fj - dy); need(fi, fj); need(fi + dx, fj + dy);
      need(fi - dx + tx, fj - dy + ty); need(fi - dx - tx, fj - dy - ty); need(fi + tx, fj + ty); need(fi - tx, fj - ty);
      const int ci[2] = {dirR ? fi : fi - 1, fi}, cj[2] = {dirR ? fj - 1 : fj, fj};
      for (int q = 0; q < 2; ++q) {
        if (inRing(ci[q], cj[q])) continue;
        need(ci[q], cj[q]); need(ci[q] + 1, cj[q]); need(ci[q] - 1, cj[q]); need(ci[q], cj[q] + 1); need(ci[q], cj[q] - 1);
      }
    }
  }
  std::vector<int> start((size_t)ntiles + 1, 0), fstart((size_t)ntiles + 1, 0);
  std::vector<uint2> faces;
  std::vector<int2> cells;
  s->exMax = s->fcMax = 0;
  for (int t = 0; t < ntiles; ++t) {
    start[t] = (int)faces.size();
    faces.insert(faces.end(), lists[t].begin(), lists[t].end());
    s->exMax = std::max(s->exMax, (int)lists[t].size());
    fstart[t] = (int)cells.size();
    cells.insert(cells.end(), far[t].begin(), far[t].end());
    s->fcMax = std::max(s->fcMax, (int)far[t].size());
  }
  start[ntiles] = (int)faces.size();
  fstart[ntiles] = (int)cells.size();
  hfree(s->exFaces);
  hfree(s->fcCells);
  vfree(s->exFlux);
  HIPCHK(hipMalloc(&s->exFaces, std::max<size_t>(faces.size(), 1) * sizeof(uint2)), "hipMalloc extra faces");
  HIPCHK(hipMalloc(&s->fcCells, std::max<size_t>(cells.size(), 1) * sizeof(int2)), "hipMalloc far cells");
  HIPCHK(hipMalloc(&s->exFlux, std::max<size_t>(faces.size(), 1) * 16 * s->rsz), "hipMalloc extra fluxes");
  if (!faces.empty()) HIPCHK(hipMemcpy(s->exFaces, faces.data(), faces.size() * sizeof(uint2), hipMemcpyHostToDevice), "upload extra faces");
  if (!cells.empty()) HIPCHK(hipMemcpy(s->fcCells, cells.data(), cells.size() * sizeof(int2), hipMemcpyHostToDevice), "upload far cells");
  HIPCHK(hipMemcpy(s->exStart, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice), "upload extra starts");
  HIPCHK(hipMemcpy(s->fcStart, fstart.data(), fstart.size() * sizeof(int), hipMemcpyHostToDevice), "upload far starts");
  base_inst(s);
  return send_inst(s);
}

// convert the reference's texture-shaped geometry into the device layout
static int upload_geometry(nz_solver* s, const uint8_t* mask, const float* geoA, const float* geoG, const float* geoM) {
  const int Nx = s->Nx, Nr = s->Nr;
  const long long N = (long long)Nx * Nr;
  const int W1 = Nx + 1;
  std::vector<float> ax((size_t)Nr * W1), ar((size_t)(Nr + 1) * Nx), lt((size_t)N);
  std::vector<float4> gg((size_t)N);
  std::vector<uint32_t> code((size_t)N, 0u);
  for (int j = 0; j < Nr; ++j)
    for (int i = 0; i <= Nx; ++i) ax[(size_t)j * W1 + i] = geoA[((size_t)j * W1 + i) * 4 + 2];
  for (int j = 0; j <= Nr; ++j)
    for (int i = 0; i < Nx; ++i) ar[(size_t)j * Nx + i] = geoA[((size_t)j * W1 + i) * 4 + 3];
  double groups = 0;
  for (long long k = 0; k < N; ++k) {
    gg[k] = make_float4(geoG[k * 4], geoG[k * 4 + 1], geoG[k * 4 + 2], geoG[k * 4 + 3]);
    lt[k] = geoM[k * 4 + 3];
    // solid cells carry a zero host offset so that decoding it never leaves the cell
    if (mask[k]) { code[k] = nzk::C_SOLID | kZeroOffset; continue; }
    const long dx = std::lround(geoM[k * 4]), dy = std::lround(geoM[k * 4 + 1]);
    const long z = std::lround(geoM[k * 4 + 2]);
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || (dx != 0 && dy != 0))
      return nz_fail(-6, "nz_set_geometry: merged-cell host offset outside the 4-neighbourhood");
    uint32_t c = ((uint32_t)(z & 15) << nzk::C_SLAVE_SHIFT) | ((z & 16) ? nzk::C_BL : 0u) |
                 ((uint32_t)(dx + 1) << nzk::C_HX_SHIFT) | ((uint32_t)(dy + 1) << nzk::C_HY_SHIFT);
    if (dx != 0 || dy != 0) c |= nzk::C_IS_SLAVE;
    else groups += 1;
    code[k] = c;
  }
  // every slave must be listed by its host, which must itself be a fluid host
  for (long long k = 0; k < N; ++k) {
    if (!(code[k] & nzk::C_IS_SLAVE)) continue;
    const int i = (int)(k % Nx), j = (int)(k / Nx);
    const int dx = (int)((code[k] >> nzk::C_HX_SHIFT) & 3u) - 1, dy = (int)((code[k] >> nzk::C_HY_SHIFT) & 3u) - 1;
    const int hi = i + dx, hj = j + dy;
    if (hi < 0 || hi >= Nx || hj < 0 || hj >= Nr) return nz_fail(-6, "nz_set_geometry: merged-cell host outside the grid");
    const uint32_t hc = code[(size_t)hj * Nx + hi];
    if ((hc & nzk::C_SOLID) || (hc & nzk::C_IS_SLAVE)) return nz_fail(-6, "nz_set_geometry: merged-cell host is solid or merged");
    const uint32_t bit = i < hi ? 1u : i > hi ? 2u : j < hj ? 4u : 8u;
    if (!((hc >> nzk::C_SLAVE_SHIFT) & bit)) return nz_fail(-6, "nz_set_geometry: merged cell missing from its host's slave mask");
  }
  // and every slave bit of a host must name a neighbour inside the grid that is its slave
  for (long long k = 0; k < N; ++k) {
    const uint32_t c = code[k];
    if ((c & nzk::C_SOLID) || (c & nzk::C_IS_SLAVE)) continue;
    const uint32_t sl = (c >> nzk::C_SLAVE_SHIFT) & 15u;
    if (!sl) continue;
    const int i = (int)(k % Nx), j = (int)(k / Nx);
    const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
    for (int b = 0; b < 4; ++b) {
      if (!(sl & (1u << b))) continue;
      const int si = i + di[b], sj = j + dj[b];
      if (si < 0 || si >= Nx || sj < 0 || sj >= Nr) return nz_fail(-6, "nz_set_geometry: slave bit points outside the grid");
      const uint32_t sc = code[(size_t)sj * Nx + si];
      const int hx = (int)((sc >> nzk::C_HX_SHIFT) & 3u) - 1, hy = (int)((sc >> nzk::C_HY_SHIFT) & 3u) - 1;
      if ((sc & nzk::C_SOLID) || !(sc & nzk::C_IS_SLAVE) || si + hx != i || sj + hy != j)
        return nz_fail(-6, "nz_set_geometry: slave bit names a cell that is not merged into this host");
    }
  }
  HIPCHK(hipMemcpy(s->ax, ax.data(), ax.size() * 4, hipMemcpyHostToDevice), "upload ax");
  HIPCHK(hipMemcpy(s->ar, ar.data(), ar.size() * 4, hipMemcpyHostToDevice), "upload ar");
  HIPCHK(hipMemcpy(s->gg, gg.data(), gg.size() * 16, hipMemcpyHostToDevice), "upload gg");
  HIPCHK(hipMemcpy(s->code, code.data(), code.size() * 4, hipMemcpyHostToDevice), "upload code");
  HIPCHK(hipMemcpy(s->lt, lt.data(), lt.size() * 4, hipMemcpyHostToDevice), "upload lt");
  s->nGroups = groups;
  HIPCHK(hipMemcpy(&s->st->nGroups, &s->nGroups, sizeof(double), hipMemcpyHostToDevice), "upload groups");
  std::memcpy(s->mask.data(), mask, (size_t)N);
  return build_extra_faces(s, code);
}

int nz_set_geometry(nz_solver* s, const uint8_t* mask, const float* geoA, const float* geoG, const float* geoM,
                    const float* jWall) {
  if (int rc = check_solver(s, "nz_set_geometry")) return rc;
  if (!mask || !geoA || !geoG || !geoM) return nz_fail(-1, "nz_set_geometry: null array");
  HIPCHK(hipStreamSynchronize(s->stream), "nz_set_geometry: sync");
  if (int rc = upload_geometry(s, mask, geoA, geoG, geoM)) return rc;
  if (jWall) {
    std::memcpy(s->jWall.data(), jWall, (size_t)s->Nx * 16);
    HIPCHK(hipMemcpy(s->jw, jWall, (size_t)s->Nx * 16, hipMemcpyHostToDevice), "upload jWall");
  }
  s->haveGeo = true;
  return 0;
}

int nz_set_mask(nz_solver* s, const uint8_t* mask) {
  if (int rc = check_solver(s, "nz_set_mask")) return rc;
  if (!mask) return nz_fail(-1, "nz_set_mask: null mask");
  HIPCHK(hipStreamSynchronize(s->stream), "nz_set_mask: sync");
  const long long N = (long long)s->Nx * s->Nr;
  std::vector<uint32_t> code((size_t)N);
  HIPCHK(hipMemcpy(code.data(), s->code, (size_t)N * 4, hipMemcpyDeviceToHost), "read code");
  for (long long k = 0; k < N; ++k) code[k] = mask[k] ? (code[k] | nzk::C_SOLID) : (code[k] & ~nzk::C_SOLID);
  HIPCHK(hipMemcpy(s->code, code.data(), (size_t)N * 4, hipMemcpyHostToDevice), "upload code");
  std::memcpy(s->mask.data(), mask, (size_t)N);
  return 0;
}

int nz_set_state(nz_solver* s, const float* A, const float* B) {
  if (int rc = check_solver(s, "nz_set_state")) return rc;
  if (!A || !B) return nz_fail(-1, "nz_set_state: null array");
  const long long N = (long long)s->Nx * s->Nr;
  nz_read_async_cancel(s);
  HIPCHK(hipStreamSynchronize(s->stream), "nz_set_state: sync");
  if (int rc = ensure_staging(s, (size_t)N * 8)) return rc;
  HIPCHK(hipMemcpy(s->dstage, A, (size_t)N * 16, hipMemcpyHostToDevice), "upload A");
  HIPCHK(hipMemcpy(s->dstage + N * 4, B, (size_t)N * 16, hipMemcpyHostToDevice), "upload B");
  HIPCHK(nzk::launch_unpack(s->prec == 64, s->U0, s->Ns, N, s->dstage, s->dstage + N * 4, s->stream), "unpack");
  HIPCHK(nzk::launch_unpack(s->prec == 64, s->U1, s->Ns, N, s->dstage, s->dstage + N * 4, s->stream), "unpack");
  // zero step info and time; the spray is reset, its seed carries on (spray.js:242-254)
  InstState is{};
  HIPCHK(hipMemcpy(&is, s->st, sizeof(is), hipMemcpyDeviceToHost), "read state info");
  const uint32_t seed = is.seed;
  std::memset(&is, 0, sizeof(is));
  is.seed = seed;
  is.nGroups = s->nGroups;
  HIPCHK(hipMemcpy(s->st, &is, sizeof(is), hipMemcpyHostToDevice), "reset state info");
  HIPCHK(nzk::launch_spray_reset(s->parc, s->NP, s->src, s->Ns, s->stream), "spray reset");
  if (!s->instSent) { if (int rc = send_inst(s)) return rc; }
  HIPCHK(nzk::launch_cfl_init(s->prec == 64, s->dinst, 1, s->h.ncblk, s->done, s->stream), "cfl init");
  HIPCHK(hipStreamSynchronize(s->stream), "nz_set_state");
  s->stepCount = 0;
  s->haveState = true;
  return 0;
}

static int apply_params(nz_solver* s, const nz_params* P) {
  Consts c;
  fill_consts(s, *P, c);
  if (!s->instSent || std::memcmp(&c, &s->h.c, sizeof(c)) != 0) {
    s->h.c = c;
    return send_inst(s);
  }
  return 0;
}

int nz_step(nz_solver* s, int n, const nz_params* P) {
  if (int rc = check_solver(s, "nz_step")) return rc;
  if (!P) return nz_fail(-1, "nz_step: null params");
  if (n < 0) return nz_fail(-1, "nz_step: negative step count");
  if (!s->haveGeo) return nz_fail(-4, "nz_step: setGeometry() has not been called");
  if (int rc = apply_params(s, P)) return rc;
  const bool fp64 = s->prec == 64;
  const bool phys = P->visc || P->chem;
  const bool spray = s->h.c.spray != 0;
  const int ntiles = s->tilesX * s->tilesY;
  for (int k = 0; k < n; ++k) {
    HIPCHK(nzk::launch_stage(1, false, fp64, s->dinst, 1, ntiles, s->done, s->stream), "stage 1");
    if (spray) HIPCHK(nzk::launch_spray(fp64, s->dinst, 1, s->h.npblk, s->stream), "spray");
    HIPCHK(nzk::launch_stage(2, phys, fp64, s->dinst, 1, ntiles, s->done, s->stream), "stage 2");
    HIPCHK(nzk::launch_finalize(s->dinst, 1, s->stream), "finalize");
  }
  s->stepCount += n;
  return 0;
}

int nz_step_timed(nz_solver* s, int n, const nz_params* P, double* ms) {
  if (int rc = check_solver(s, "nz_step_timed")) return rc;
  if (!P || !ms) return nz_fail(-1, "nz_step_timed: null argument");
  if (n <= 0) return nz_fail(-1, "nz_step_timed: step count must be positive");
  if (!s->haveGeo) return nz_fail(-4, "nz_step_timed: setGeometry() has not been called");
  if (int rc = apply_params(s, P)) return rc;
  const bool fp64 = s->prec == 64;
  const bool phys = P->visc || P->chem;
  const bool spray = s->h.c.spray != 0;
  const int ntiles = s->tilesX * s->tilesY;
  std::vector<hipEvent_t> ev((size_t)n * 4);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e), "event create");
  for (int k = 0; k < n; ++k) {
    HIPCHK(hipEventRecord(ev[k * 4], s->stream), "event");
    HIPCHK(nzk::launch_stage(1, false, fp64, s->dinst, 1, ntiles, s->done, s->stream), "stage 1");
    HIPCHK(hipEventRecord(ev[k * 4 + 1], s->stream), "event");
    if (spray) HIPCHK(nzk::launch_spray(fp64, s->dinst, 1, s->h.npblk, s->stream), "spray");
    HIPCHK(hipEventRecord(ev[k * 4 + 2], s->stream), "event");
    HIPCHK(nzk::launch_stage(2, phys, fp64, s->dinst, 1, ntiles, s->done, s->stream), "stage 2");
    HIPCHK(hipEventRecord(ev[k * 4 + 3], s->stream), "event");
    HIPCHK(nzk::launch_finalize(s->dinst, 1, s->stream), "finalize");
  }
  HIPCHK(hipStreamSynchronize(s->stream), "nz_step_timed sync");
  double t[3] = {0, 0, 0};
  for (int k = 0; k < n; ++k)
    for (int q = 0; q < 3; ++q) {
      float e = 0.f;
      HIPCHK(hipEventElapsedTime(&e, ev[k * 4 + q], ev[k * 4 + q + 1]), "event elapsed");
      t[q] += e;
    }
  float all = 0.f;
  HIPCHK(hipEventElapsedTime(&all, ev[0], ev[(size_t)n * 4 - 1]), "event elapsed");
  for (auto& e : ev) (void)hipEventDestroy(e);
  ms[0] = t[0]; ms[1] = t[1]; ms[2] = t[2]; ms[3] = all;
  s->stepCount += n;
  return 0;
}

int nz_rescale_ambient(nz_solver* s, float fp, float fT, float du, const nz_params* P) {
  if (int rc = check_solver(s, "nz_rescale_ambient")) return rc;
  if (!P) return nz_fail(-1, "nz_rescale_ambient: null params");
  if (int rc = apply_params(s, P)) return rc;
  DevInst I = s->h;
  HIPCHK(nzk::launch_rescale(s->prec == 64, I, fp, fT, du, s->stream), "rescale");
  HIPCHK(nzk::launch_cfl_init(s->prec == 64, s->dinst, 1, s->h.ncblk, s->done, s->stream), "cfl init");
  return 0;
}

int nz_draw_wall_row(nz_solver* s) {
  if (int rc = check_solver(s, "nz_draw_wall_row")) return rc;
  HIPCHK(nzk::launch_wall_row(s->prec == 64, s->h, s->jw, s->wallrow, s->stream), "wall row");
  return 0;
}

// enqueue one rectangle read into device staging at `off` (floats); returns its size or < 0
static long long enqueue_rect(nz_solver* s, int field, int x, int y, int w, int h, float* dst) {
  const int Nx = s->Nx, Nr = s->Nr;
  int FW = Nx, FH = Nr;
  if (field == NZ_FIELD_STEP) { FW = 1; FH = 1; }
  else if (field == NZ_FIELD_WALLROW) { FW = Nx; FH = 1; }
  else if (field != NZ_FIELD_A && field != NZ_FIELD_B && field != NZ_FIELD_T2) return nz_fail(-1, "read: unknown field");
  if (w <= 0 || h <= 0 || x < 0 || y < 0 || x + w > FW || y + h > FH) return nz_fail(-1, "read: rectangle outside the field");
  const bool fp64 = s->prec == 64;
  if (field == NZ_FIELD_A || field == NZ_FIELD_B) {
    if (nzk::launch_pack(fp64, s->U0, s->Ns, Nx, field == NZ_FIELD_A ? 0 : 4, x, y, w, h, dst, s->stream) != hipSuccess)
      return nz_fail(-10, "read: pack failed");
  } else if (field == NZ_FIELD_STEP) {
    if (nzk::launch_step_info(s->st, dst, s->stream) != hipSuccess) return nz_fail(-10, "read: step info failed");
  } else if (field == NZ_FIELD_WALLROW) {
    for (int r = 0; r < h; ++r)
      if (hipMemcpyAsync(dst + (size_t)r * w * 4, s->wallrow + (size_t)x * 4, (size_t)w * 16, hipMemcpyDeviceToDevice, s->stream) != hipSuccess)
        return nz_fail(-10, "read: wall row copy failed");
  } else {
    // T2 of the last physics pass, recomputed into the tail of the staging buffer
    float* t2 = s->dstage + s->dstageCap - (size_t)Nx * Nr * 4;
    if (nzk::launch_t2(fp64, s->h, t2, s->stream) != hipSuccess) return nz_fail(-10, "read: T2 failed");
    for (int r = 0; r < h; ++r)
      if (hipMemcpyAsync(dst + (size_t)r * w * 4, t2 + ((size_t)(y + r) * Nx + x) * 4, (size_t)w * 16, hipMemcpyDeviceToDevice, s->stream) != hipSuccess)
        return nz_fail(-10, "read: T2 copy failed");
  }
  return (long long)w * h * 4;
}

int nz_read_rect(nz_solver* s, int field, int x, int y, int w, int h, float* out) {
  if (int rc = check_solver(s, "nz_read_rect")) return rc;
  if (!out) return nz_fail(-1, "nz_read_rect: null output");
  const size_t n = (size_t)w * h * 4;
  const size_t extra = field == NZ_FIELD_T2 ? (size_t)s->Nx * s->Nr * 4 : 0;
  if (int rc = ensure_staging(s, n + extra)) return rc;
  const long long r = enqueue_rect(s, field, x, y, w, h, s->dstage);
  if (r < 0) return (int)r;
  HIPCHK(hipMemcpyAsync(s->hpin, s->dstage, n * 4, hipMemcpyDeviceToHost, s->stream), "read copy");
  HIPCHK(hipStreamSynchronize(s->stream), "read sync");
  std::memcpy(out, s->hpin, n * 4);
  return 0;
}

int nz_read_step_info(nz_solver* s, float* out) { return nz_read_rect(s, NZ_FIELD_STEP, 0, 0, 1, 1, out); }

int nz_read_state_f64(nz_solver* s, double* A, double* B) {
  if (int rc = check_solver(s, "nz_read_state_f64")) return rc;
  if (!A || !B) return nz_fail(-1, "nz_read_state_f64: null output");
  const long long N = (long long)s->Nx * s->Nr;
  if (int rc = ensure_staging(s, (size_t)N * 16)) return rc;
  double* d = reinterpret_cast<double*>(s->dstage);
  HIPCHK(nzk::launch_pack_f64(s->prec == 64, s->U0, s->Ns, N, d, d + N * 4, s->stream), "pack f64");
  HIPCHK(hipMemcpyAsync(A, d, (size_t)N * 32, hipMemcpyDeviceToHost, s->stream), "copy A");
  HIPCHK(hipMemcpyAsync(B, d + N * 4, (size_t)N * 32, hipMemcpyDeviceToHost, s->stream), "copy B");
  HIPCHK(hipStreamSynchronize(s->stream), "nz_read_state_f64");
  return 0;
}

int nz_step_count(const nz_solver* s, int64_t* out) {
  if (!s || !out) return nz_fail(-1, "nz_step_count: null argument");
  *out = s->stepCount;
  return 0;
}

int nz_sync(nz_solver* s) {
  if (!s) return nz_fail(-1, "nz_sync: null solver");
  HIPCHK(hipStreamSynchronize(s->stream), "nz_sync");
  return 0;
}

int nz_read_async_begin(nz_solver* s, int nreq, const nz_read_req* reqs, int32_t tag) {
  if (int rc = check_solver(s, "nz_read_async_begin")) return rc;
  if (s->asyncPending) return 0;
  if (nreq <= 0 || !reqs) return nz_fail(-1, "nz_read_async_begin: no requests");
  size_t total = 0;
  bool t2 = false;
  for (int k = 0; k < nreq; ++k) { total += (size_t)reqs[k].w * reqs[k].h * 4; t2 |= reqs[k].field == NZ_FIELD_T2; }
  const size_t extra = t2 ? (size_t)s->Nx * s->Nr * 4 : 0;
  if (int rc = ensure_staging(s, total + extra)) return rc;
  if (total > s->asyncCap) {
    if (s->asyncHost) (void)hipHostFree(s->asyncHost);
    s->asyncHost = nullptr;
    HIPCHK(hipHostMalloc(&s->asyncHost, total * 4, hipHostMallocDefault), "hipHostMalloc async");
    s->asyncCap = total;
  }
  size_t off = 0;
  for (int k = 0; k < nreq; ++k) {
    const long long r = enqueue_rect(s, reqs[k].field, reqs[k].x, reqs[k].y, reqs[k].w, reqs[k].h, s->dstage + off);
    if (r < 0) return (int)r;
    off += (size_t)r;
  }
  HIPCHK(hipMemcpyAsync(s->asyncHost, s->dstage, total * 4, hipMemcpyDeviceToHost, s->stream), "async copy");
  HIPCHK(hipEventRecord(s->asyncEv, s->stream), "async event");
  s->asyncPending = true;
  s->asyncTag = tag;
  s->asyncFloats = total;
  return 1;
}

int nz_read_async_poll(nz_solver* s, float* out, int32_t* tag) {
  if (!s) return nz_fail(-1, "nz_read_async_poll: null solver");
  if (!s->asyncPending) return 0;
  const hipError_t e = hipEventQuery(s->asyncEv);
  if (e == hipErrorNotReady) return 0;
  if (e != hipSuccess) return nz_fail(-10, std::string("nz_read_async_poll: ") + hipGetErrorString(e));
  if (out) std::memcpy(out, s->asyncHost, s->asyncFloats * 4);
  if (tag) *tag = s->asyncTag;
  s->asyncPending = false;
  return 1;
}

int nz_read_async_cancel(nz_solver* s) {
  if (!s) return 0;
  if (s->asyncPending) (void)hipEventSynchronize(s->asyncEv);
  s->asyncPending = false;
  return 0;
}

int nz_read_async_pending(const nz_solver* s) { return s && s->asyncPending ? 1 : 0; }

int nz_spray_set_injector(nz_solver* s, const float* tab) {
  if (int rc = check_solver(s, "nz_spray_set_injector")) return rc;
  if (!tab) return 0;  // setInjector(null) leaves the table as it was (spray.js:237)
  HIPCHK(hipStreamSynchronize(s->stream), "set injector sync");
  HIPCHK(hipMemcpy(s->tab, tab, 256 * 3 * 16, hipMemcpyHostToDevice), "upload injector table");
  s->haveTab = true;
  s->instSent = false;  // spray flag of the constants may change
  return 0;
}

int nz_spray_reset(nz_solver* s) {
  if (int rc = check_solver(s, "nz_spray_reset")) return rc;
  HIPCHK(nzk::launch_spray_reset(s->parc, s->NP, s->src, s->Ns, s->stream), "spray reset");
  InstState is{};
  HIPCHK(hipStreamSynchronize(s->stream), "spray reset sync");
  HIPCHK(hipMemcpy(&is, s->st, sizeof(is), hipMemcpyDeviceToHost), "read state info");
  is.head = 0;
  HIPCHK(hipMemcpy(s->st, &is, sizeof(is), hipMemcpyHostToDevice), "write state info");
  return 0;
}

int nz_spray_set_count(nz_solver* s, int n) {
  if (int rc = check_solver(s, "nz_spray_set_count")) return rc;
  if (n < 4 || n > (1 << 22) || n % 4 != 0) return nz_fail(-1, "nz_spray_set_count: count must be a multiple of 4 in [4, 2^22]");
  HIPCHK(hipStreamSynchronize(s->stream), "spray count sync");
  s->NP = n;
  if (int rc = alloc_spray(s)) return rc;
  return nz_spray_reset(s);
}

int nz_spray_read(nz_solver* s, float* PA, float* PB) {
  if (int rc = check_solver(s, "nz_spray_read")) return rc;
  const int NP = s->NP;
  std::vector<float> p((size_t)NP * 8);
  HIPCHK(hipStreamSynchronize(s->stream), "spray read sync");
  HIPCHK(hipMemcpy(p.data(), s->parc, p.size() * 4, hipMemcpyDeviceToHost), "spray read");
  for (int k = 0; k < NP; ++k)
    for (int q = 0; q < 4; ++q) {
      if (PA) PA[(size_t)k * 4 + q] = p[(size_t)q * NP + k];
      if (PB) PB[(size_t)k * 4 + q] = p[(size_t)(4 + q) * NP + k];
    }
  return 0;
}

int nz_spray_stats(nz_solver* s, nz_spray_stats_t* out) {
  if (!out) return nz_fail(-1, "nz_spray_stats: null output");
  const int NP = s ? s->NP : 0;
  std::vector<float> PB((size_t)NP * 4);
  if (int rc = nz_spray_read(s, nullptr, PB.data())) return rc;
  double f = 0, o = 0, dS = 0;
  int64_t n = 0;
  for (int k = 0; k < NP; ++k) {
    const double m = PB[(size_t)k * 4 + 1];
    if (m > 0) {
      n++;
      dS += PB[(size_t)k * 4];
      if (PB[(size_t)k * 4 + 2] < 0.5) f += m; else o += m;
    }
  }
  out->fuel = f; out->ox = o; out->parcels = n; out->meanD = n ? dS / n : 0;
  return 0;
}

int nz_set_probes(nz_solver* s, const nz_probes* pr) {
  if (int rc = check_solver(s, "nz_set_probes")) return rc;
  if (!pr) return nz_fail(-1, "nz_set_probes: null probes");
  if (pr->iExit < 0 || pr->iExit + 1 > s->Nx || pr->jExit < 0 || pr->jExit > s->Nr || pr->iThroat < 0 ||
      pr->iThroat >= s->Nx || pr->iChamber < 0 || pr->iChamber >= s->Nx)
    return nz_fail(-1, "nz_set_probes: probe outside the grid");
  s->probes = *pr;
  s->haveProbes = true;
  return 0;
}

int nz_integrals_read(nz_solver* s, const nz_params* P, nz_integrals* out) {
  if (int rc = check_solver(s, "nz_integrals_read")) return rc;
  if (!out || !P) return nz_fail(-1, "nz_integrals_read: null argument");
  if (!s->haveProbes) return nz_fail(-4, "nz_integrals_read: probes not set");
  if (int rc = apply_params(s, P)) return rc;
  const nz_probes& pr = s->probes;
  HIPCHK(nzk::launch_integrals(s->prec == 64, s->h, pr.iExit, pr.jExit, pr.iThroat, pr.iChamber, s->dint, s->stream), "integrals");
  double v[16];
  HIPCHK(hipMemcpyAsync(v, s->dint, 11 * 8, hipMemcpyDeviceToHost, s->stream), "integrals copy");
  HIPCHK(hipStreamSynchronize(s->stream), "integrals sync");
  const double dx2 = pr.dx * pr.dx;
  const double Fsi = v[0] * dx2 * nz::P_REF;
  const double mdotSi = v[1] * dx2 * nz::RHO_REF * nz::U_REF;
  out->F = Fsi;
  out->mdot = mdotSi;
  out->Me = v[1] > 1e-12 ? v[2] / v[1] : 0;
  out->pe = (v[4] > 0 ? v[3] / v[4] : 0) * nz::P_REF;
  out->Isp = mdotSi > 1e-9 ? Fsi / (mdotSi * nz::G0) : 0;
  out->mdotThroat = v[5] * dx2 * nz::RHO_REF * nz::U_REF;
  out->mass = v[6];
  out->energy = v[7];
  out->Pc = v[8] * nz::P_REF;
  out->Tc = v[9] * nz::T_REF;
  out->residual = v[10];
  return 0;
}

}  // extern "C"
