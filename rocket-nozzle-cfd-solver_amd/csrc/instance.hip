// instance.hip — one solve's device buffers: allocation, the reference's
// texture-shaped geometry converted into the SoA device layout (setGeometry,
// solver.js:801-814), state upload (setState, 816-828), spray buffers
// (GPUSpray.resize / reset / setInjector, spray.js:224-254), and the per-tile
// lists of merged-cell faces the stage kernels evaluate as extra work items.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <iterator>

#include "nz_inst.h"
#include "nz_launch.h"

namespace nz {

using nzk::Consts;
using nzk::DevInst;
using nzk::InstState;

int DevArena::create(size_t bytes) {
  destroy();
  bytes = (bytes + 255) / 256 * 256;
  HIPCHK(hipMalloc(&base_, bytes), "device slab");
  size_ = bytes;
  used_ = 0;
  free_[0] = bytes;
  return 0;
}

void DevArena::destroy() {
  if (base_) (void)hipFree(base_);
  base_ = nullptr;
  size_ = used_ = 0;
  free_.clear();
  live_.clear();
}

void* DevArena::alloc(size_t bytes) {
  bytes = std::max<size_t>((bytes + 255) / 256 * 256, 256);
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < bytes) continue;
    const size_t off = it->first, rest = it->second - bytes;
    free_.erase(it);
    if (rest) free_[off + bytes] = rest;
    live_[off] = bytes;
    used_ += bytes;
    return base_ + off;
  }
  failed_ = true;
  return nullptr;
}

void DevArena::release(void* p) {
  if (!p) return;
  const size_t off = (size_t)(static_cast<char*>(p) - base_);
  auto it = live_.find(off);
  if (it == live_.end()) return;
  size_t start = off, bytes = it->second;
  live_.erase(it);
  used_ -= bytes;
  auto next = free_.lower_bound(start);
  if (next != free_.end() && next->first == start + bytes) {
    bytes += next->second;
    next = free_.erase(next);
  }
  if (next != free_.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == start) {
      start = prev->first;
      bytes += prev->second;
      free_.erase(prev);
    }
  }
  free_[start] = bytes;
}

hipError_t inst_dalloc(Inst& I, void** p, size_t bytes) {
  if (!I.arena) return hipMalloc(p, bytes);
  *p = I.arena->alloc(bytes);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}

void inst_dfree(Inst& I, void* p) {
  if (!p) return;
  if (I.arena) I.arena->release(p);
  else (void)hipFree(p);
}

namespace {
template <class T>
hipError_t dalloc(Inst& I, T** p, size_t bytes) {
  return inst_dalloc(I, reinterpret_cast<void**>(p), bytes);
}
template <class T>
void dfree(Inst& I, T*& p) {
  inst_dfree(I, (void*)p);
  p = nullptr;
}
}  // namespace

// the most tiles any tile shape of this precision cuts the grid into: the capacity of the
// per-tile partials, which carry the wave-speed max from one step to the next and so stay
// in place when the tile shape changes
size_t part_capacity(int Nx, int Nr, int prec) {
  size_t most = 1;
  for (int k = 0; k < nzk::tile_shape_count(prec == 64); ++k) {
    int TX = 0, TY = 0;
    nzk::tile_shape(prec == 64, k, &TX, &TY);
    most = std::max(most, (size_t)((Nx + TX - 1) / TX) * ((Nr + TY - 1) / TY));
  }
  return most;
}

size_t inst_bytes(int Nx, int Nr, int prec, int NP) {
  const size_t N = (size_t)Nx * Nr, Ns = (N + 63) / 64 * 64, R = prec == 64 ? 8 : 4;
  const size_t tiles = part_capacity(Nx, Nr, prec);
  return Ns * 8 * R * 2                                   // U0, U1
         + (size_t)Nr * (Nx + 1) * 4 + (size_t)(Nr + 1) * Nx * 4  // ax, ar
         + N * (16 + 4 + 4)                               // gg, code, lt
         + Ns * 32                                        // spray sources (4 x int64)
         + (size_t)NP * 32                                // parcels
         + tiles * (16 + 8) + (size_t)Nx * 32 + 4096;     // partials, starts, wall row, small
}

int inst_alloc(Inst& I, int Nx, int Nr, int prec, InstState* st, int shape, DevArena* arena) {
  inst_free(I);
  I.arena = arena;
  if (!nzk::tile_shape(prec == 64, shape, &I.TX, &I.TY)) return nz_fail(-1, "unknown tile shape");
  I.shape = shape;
  I.prec = prec;
  I.rsz = prec == 64 ? 8 : 4;
  I.Nx = Nx;
  I.Nr = Nr;
  const long long N = (long long)Nx * Nr;
  I.Ns = (N + 63) / 64 * 64;
  I.tilesX = (Nx + I.TX - 1) / I.TX;
  I.tilesY = (Nr + I.TY - 1) / I.TY;
  const size_t tiles = (size_t)I.tilesX * I.tilesY, R = I.rsz;
  HIPCHK(dalloc(I, &I.U0, (size_t)I.Ns * 8 * R), "allocate state");
  HIPCHK(dalloc(I, &I.U1, (size_t)I.Ns * 8 * R), "allocate state");
  HIPCHK(hipMemset(I.U0, 0, (size_t)I.Ns * 8 * R), "memset");
  HIPCHK(hipMemset(I.U1, 0, (size_t)I.Ns * 8 * R), "memset");
  HIPCHK(dalloc(I, &I.ax, (size_t)Nr * (Nx + 1) * 4), "allocate ax");
  HIPCHK(dalloc(I, &I.ar, (size_t)(Nr + 1) * Nx * 4), "allocate ar");
  HIPCHK(dalloc(I, &I.gg, (size_t)N * 16), "allocate gg");
  HIPCHK(dalloc(I, &I.code, (size_t)N * 4), "allocate code");
  HIPCHK(dalloc(I, &I.lt, (size_t)N * 4), "allocate lt");
  HIPCHK(dalloc(I, &I.jw, (size_t)Nx * 16), "allocate jWall");
  HIPCHK(dalloc(I, &I.src, (size_t)I.Ns * 32), "allocate src");
  HIPCHK(dalloc(I, &I.tab, 256 * 3 * 16), "allocate tab");
  I.partCap = part_capacity(Nx, Nr, prec);
  HIPCHK(dalloc(I, &I.part, I.partCap * sizeof(double2)), "allocate partials");
  HIPCHK(hipMemset(I.part, 0, I.partCap * sizeof(double2)), "memset");
  HIPCHK(dalloc(I, &I.exStart, (tiles + 1) * sizeof(int)), "allocate extra starts");
  HIPCHK(dalloc(I, &I.fcStart, (tiles + 1) * sizeof(int)), "allocate far starts");
  HIPCHK(dalloc(I, &I.exFaces, sizeof(uint2)), "allocate extra faces");
  HIPCHK(dalloc(I, &I.fcCells, sizeof(int2)), "allocate far cells");
  HIPCHK(dalloc(I, &I.exFlux, 16 * 8), "allocate extra fluxes");
  HIPCHK(dalloc(I, &I.wallrow, (size_t)Nx * 16), "allocate wall row");
  HIPCHK(dalloc(I, &I.dint, (16 + 2 * (size_t)nzk::INTEG_BLOCKS_MAX) * 8), "allocate integrals");
  if (st) {
    I.st = st;
    I.ownSt = false;
  } else {
    HIPCHK(dalloc(I, &I.st, sizeof(InstState)), "allocate state info");
    I.ownSt = true;
  }
  // empty geometry (all faces closed, every cell its own host) until setGeometry
  HIPCHK(hipMemset(I.ax, 0, (size_t)Nr * (Nx + 1) * 4), "memset");
  HIPCHK(hipMemset(I.ar, 0, (size_t)(Nr + 1) * Nx * 4), "memset");
  HIPCHK(hipMemset(I.gg, 0, (size_t)N * 16), "memset");
  HIPCHK(hipMemsetD32((hipDeviceptr_t)I.code, (int)kZeroOffset, (size_t)N), "memset code");
  HIPCHK(hipMemset(I.lt, 0, (size_t)N * 4), "memset");
  HIPCHK(hipMemset(I.jw, 0, (size_t)Nx * 16), "memset");
  HIPCHK(hipMemset(I.src, 0, (size_t)I.Ns * 32), "memset");
  HIPCHK(hipMemset(I.tab, 0, 256 * 3 * 16), "memset");
  HIPCHK(hipMemset(I.exStart, 0, (tiles + 1) * sizeof(int)), "memset");
  HIPCHK(hipMemset(I.fcStart, 0, (tiles + 1) * sizeof(int)), "memset");
  HIPCHK(hipMemset(I.wallrow, 0, (size_t)Nx * 16), "memset");
  InstState is{};
  is.seed = 1;  // GPUSpray starts at seed 1 (spray.js:221)
  HIPCHK(hipMemcpy(I.st, &is, sizeof(is), hipMemcpyHostToDevice), "init state info");
  I.mask.assign((size_t)N, 0);
  I.codeH.assign((size_t)N, kZeroOffset);
  I.jWall.assign((size_t)Nx * 4, 0.f);
  I.haveGeo = I.haveTab = false;
  I.bytes = inst_bytes(Nx, Nr, prec, I.NP);
  return 0;
}

void inst_free(Inst& I) {
  dfree(I, I.U0); dfree(I, I.U1);
  dfree(I, I.ax); dfree(I, I.ar); dfree(I, I.lt); dfree(I, I.jw); dfree(I, I.src); dfree(I, I.parc); dfree(I, I.tab);
  dfree(I, I.gg); dfree(I, I.code); dfree(I, I.exStart); dfree(I, I.exFaces); dfree(I, I.fcStart); dfree(I, I.fcCells);
  dfree(I, I.exFlux); dfree(I, I.part); dfree(I, I.wallrow); dfree(I, I.dint);
  if (I.ownSt) dfree(I, I.st);
  I.st = nullptr;
  I.haveGeo = false;
}

// For the tile shape in use, every merged cell whose host lies in another tile
// contributes its three faces that are not faces of the host's tile; the stage
// kernels evaluate those as extra work items of the host's workgroup. They are
// listed host by host, three per such slave (its W, E, S, N faces without the one
// it shares with the host), and the host's code word records the slot of its
// first one, so the kernel finds a slave's faces without searching. A host whose
// slaves do not fit the tile's list gets no slot and evaluates their faces itself.
// The cells those faces' stencils reach beyond the tile's 2-cell halo are
// preloaded into LDS (up to MAX_FAR_CELLS, inside a FAR_HALO window).
static int build_extra_faces(Inst& I, std::vector<uint32_t>& code) {
  const int Nx = I.Nx, Nr = I.Nr, TX = I.TX, TY = I.TY;
  const int ntiles = I.tilesX * I.tilesY;
  std::vector<std::vector<uint2>> lists((size_t)ntiles);
  const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
  for (int j = 0; j < Nr; ++j)
    for (int i = 0; i < Nx; ++i) {
      uint32_t& c = code[(size_t)j * Nx + i];
      c &= ~(255u << nzk::C_EXS_SHIFT);
      if ((c & nzk::C_SOLID) || (c & nzk::C_IS_SLAVE)) continue;
      const uint32_t sl = (c >> nzk::C_SLAVE_SHIFT) & 15u;
      if (!sl) continue;
      const int tx = i / TX, ty = j / TY, i0 = tx * TX, j0 = ty * TY;
      std::vector<uint2>& L = lists[(size_t)ty * I.tilesX + tx];
      int outside[4], nout = 0;
      for (int b = 0; b < 4; ++b) {
        if (!(sl & (1u << b))) continue;
        const int si = i + di[b], sj = j + dj[b];
        if (si >= i0 && si < i0 + TX && sj >= j0 && sj < j0 + TY) continue;
        outside[nout++] = b;
      }
      if (!nout || L.size() + 3 * (size_t)nout > (size_t)nzk::MAX_EXTRA_FACES) continue;
      c |= (uint32_t)(L.size() / 3 + 1) << nzk::C_EXS_SHIFT;
      for (int q = 0; q < nout; ++q) {
        const int b = outside[q], si = i + di[b], sj = j + dj[b];
        const int shared = b ^ 1;  // W, E, S, N = 0..3; the face toward the host
        for (int f = 0; f < 4; ++f) {
          if (f == shared) continue;
          const bool dirR = f >= 2;
          const int fi = si + (f == 1 ? 1 : 0), fj = sj + (f == 3 ? 1 : 0);
          L.push_back(make_uint2((uint32_t)fi | (dirR ? 0x80000000u : 0u), (uint32_t)fj));
        }
      }
    }
  // cells the extra faces' stencils reach outside the tile's 2-cell halo: inviscid
  // stencil + tangential pressures, and the 5-point gradient stencils of both cells
  // of a viscous face when a cell lies outside the 1-cell ring of the gradient tile
  std::vector<std::vector<int2>> far((size_t)ntiles);
  for (int t = 0; t < ntiles; ++t) {
    const int i0 = (t % I.tilesX) * TX, j0 = (t / I.tilesX) * TY;
    const int H = nzk::FAR_HALO;
    auto inHalo = [&](int i, int j) { return i >= i0 - 2 && i <= i0 + TX + 1 && j >= j0 - 2 && j <= j0 + TY + 1; };
    auto inRing = [&](int i, int j) { return i >= i0 - 1 && i <= i0 + TX && j >= j0 - 1 && j <= j0 + TY; };
    auto inWin = [&](int i, int j) { return i >= i0 - H && i < i0 + TX + H && j >= j0 - H && j < j0 + TY + H; };
    std::vector<int2>& F = far[t];
    auto need = [&](int i, int j) {
      if (inHalo(i, j) || !inWin(i, j) || F.size() >= (size_t)nzk::MAX_FAR_CELLS) return;
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
  I.exMax = I.fcMax = 0;
  for (int t = 0; t < ntiles; ++t) {
    start[t] = (int)faces.size();
    faces.insert(faces.end(), lists[t].begin(), lists[t].end());
    I.exMax = std::max(I.exMax, (int)lists[t].size());
    fstart[t] = (int)cells.size();
    cells.insert(cells.end(), far[t].begin(), far[t].end());
    I.fcMax = std::max(I.fcMax, (int)far[t].size());
  }
  start[ntiles] = (int)faces.size();
  fstart[ntiles] = (int)cells.size();
  dfree(I, I.exFaces);
  dfree(I, I.fcCells);
  dfree(I, I.exFlux);
  HIPCHK(dalloc(I, &I.exFaces, std::max<size_t>(faces.size(), 1) * sizeof(uint2)), "allocate extra faces");
  HIPCHK(dalloc(I, &I.fcCells, std::max<size_t>(cells.size(), 1) * sizeof(int2)), "allocate far cells");
  HIPCHK(dalloc(I, &I.exFlux, std::max<size_t>(faces.size(), 1) * 16 * I.rsz), "allocate extra fluxes");
  if (!faces.empty()) HIPCHK(hipMemcpy(I.exFaces, faces.data(), faces.size() * sizeof(uint2), hipMemcpyHostToDevice), "upload extra faces");
  if (!cells.empty()) HIPCHK(hipMemcpy(I.fcCells, cells.data(), cells.size() * sizeof(int2), hipMemcpyHostToDevice), "upload far cells");
  HIPCHK(hipMemcpy(I.exStart, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice), "upload extra starts");
  HIPCHK(hipMemcpy(I.fcStart, fstart.data(), fstart.size() * sizeof(int), hipMemcpyHostToDevice), "upload far starts");
  return 0;
}

int inst_set_geometry(Inst& I, const uint8_t* mask, const float* geoA, const float* geoG, const float* geoM,
                      const float* jWall) {
  const int Nx = I.Nx, Nr = I.Nr;
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
      return nz_fail(-6, "setGeometry: merged-cell host offset outside the 4-neighbourhood");
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
    if (hi < 0 || hi >= Nx || hj < 0 || hj >= Nr) return nz_fail(-6, "setGeometry: merged-cell host outside the grid");
    const uint32_t hc = code[(size_t)hj * Nx + hi];
    if ((hc & nzk::C_SOLID) || (hc & nzk::C_IS_SLAVE)) return nz_fail(-6, "setGeometry: merged-cell host is solid or merged");
    const uint32_t bit = i < hi ? 1u : i > hi ? 2u : j < hj ? 4u : 8u;
    if (!((hc >> nzk::C_SLAVE_SHIFT) & bit)) return nz_fail(-6, "setGeometry: merged cell missing from its host's slave mask");
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
      if (si < 0 || si >= Nx || sj < 0 || sj >= Nr) return nz_fail(-6, "setGeometry: slave bit points outside the grid");
      const uint32_t sc = code[(size_t)sj * Nx + si];
      const int hx = (int)((sc >> nzk::C_HX_SHIFT) & 3u) - 1, hy = (int)((sc >> nzk::C_HY_SHIFT) & 3u) - 1;
      if ((sc & nzk::C_SOLID) || !(sc & nzk::C_IS_SLAVE) || si + hx != i || sj + hy != j)
        return nz_fail(-6, "setGeometry: slave bit names a cell that is not merged into this host");
    }
  }
  if (int rc = build_extra_faces(I, code)) return rc;
  HIPCHK(hipMemcpy(I.ax, ax.data(), ax.size() * 4, hipMemcpyHostToDevice), "upload ax");
  HIPCHK(hipMemcpy(I.ar, ar.data(), ar.size() * 4, hipMemcpyHostToDevice), "upload ar");
  HIPCHK(hipMemcpy(I.gg, gg.data(), gg.size() * 16, hipMemcpyHostToDevice), "upload gg");
  HIPCHK(hipMemcpy(I.code, code.data(), code.size() * 4, hipMemcpyHostToDevice), "upload code");
  HIPCHK(hipMemcpy(I.lt, lt.data(), lt.size() * 4, hipMemcpyHostToDevice), "upload lt");
  I.codeH = std::move(code);
  I.nGroups = groups;
  HIPCHK(hipMemcpy(&I.st->nGroups, &I.nGroups, sizeof(double), hipMemcpyHostToDevice), "upload groups");
  std::memcpy(I.mask.data(), mask, (size_t)N);
  if (jWall) {
    std::memcpy(I.jWall.data(), jWall, (size_t)Nx * 16);
    HIPCHK(hipMemcpy(I.jw, jWall, (size_t)Nx * 16, hipMemcpyHostToDevice), "upload jWall");
  }
  I.haveGeo = true;
  return 0;
}

int inst_set_mask(Inst& I, const uint8_t* mask) {
  const long long N = (long long)I.Nx * I.Nr;
  std::vector<uint32_t>& code = I.codeH;
  for (long long k = 0; k < N; ++k) code[k] = mask[k] ? (code[k] | nzk::C_SOLID) : (code[k] & ~nzk::C_SOLID);
  HIPCHK(hipMemcpy(I.code, code.data(), (size_t)N * 4, hipMemcpyHostToDevice), "upload code");
  std::memcpy(I.mask.data(), mask, (size_t)N);
  return 0;
}

int inst_set_shape(Inst& I, int shape) {
  int TX = 0, TY = 0;
  if (!nzk::tile_shape(I.prec == 64, shape, &TX, &TY)) return nz_fail(-1, "unknown tile shape");
  if (shape == I.shape) return 0;
  // the last step's partials, of the old tiling, folded into the form every tiling reads
  HIPCHK(nzk::launch_part_canon(I.part, I.tilesX * I.tilesY, (int)I.partCap, nullptr), "partials");
  HIPCHK(hipDeviceSynchronize(), "partials");
  I.shape = shape;
  I.TX = TX;
  I.TY = TY;
  I.tilesX = (I.Nx + TX - 1) / TX;
  I.tilesY = (I.Nr + TY - 1) / TY;
  const size_t tiles = (size_t)I.tilesX * I.tilesY;
  dfree(I, I.exStart);
  dfree(I, I.fcStart);
  HIPCHK(dalloc(I, &I.exStart, (tiles + 1) * sizeof(int)), "allocate extra starts");
  HIPCHK(dalloc(I, &I.fcStart, (tiles + 1) * sizeof(int)), "allocate far starts");
  if (int rc = build_extra_faces(I, I.codeH)) return rc;
  HIPCHK(hipMemcpy(I.code, I.codeH.data(), I.codeH.size() * 4, hipMemcpyHostToDevice), "upload code");
  return 0;
}

int inst_spray_alloc(Inst& I, int NP, hipStream_t s) {
  dfree(I, I.parc);
  I.NP = NP;
  HIPCHK(dalloc(I, &I.parc, (size_t)NP * 8 * 4), "allocate parcels");
  I.bytes = inst_bytes(I.Nx, I.Nr, I.prec, NP);
  return inst_spray_reset(I, s);
}

int inst_spray_reset(Inst& I, hipStream_t s) {
  HIPCHK(nzk::launch_spray_reset(I.parc, I.NP, I.src, I.Ns, s), "spray reset");
  HIPCHK(hipStreamSynchronize(s), "spray reset sync");
  InstState is{};
  HIPCHK(hipMemcpy(&is, I.st, sizeof(is), hipMemcpyDeviceToHost), "read state info");
  is.head = 0;
  HIPCHK(hipMemcpy(I.st, &is, sizeof(is), hipMemcpyHostToDevice), "write state info");
  return 0;
}

int inst_set_injector(Inst& I, const float* tab) {
  HIPCHK(hipMemcpy(I.tab, tab, 256 * 3 * 16, hipMemcpyHostToDevice), "upload injector table");
  I.haveTab = true;
  I.tabSpeed = 0.f;
  for (int t = 0; t < 256; ++t) I.tabSpeed = std::max(I.tabSpeed, std::fabs(tab[t * 4 + 3]));
  return 0;
}

DevInst inst_desc(const Inst& I) {
  DevInst d;
  std::memset(&d, 0, sizeof(d));
  d.Nx = I.Nx; d.Nr = I.Nr;
  d.tilesX = I.tilesX; d.tilesY = I.tilesY; d.tile0 = 0;
  d.pblk0 = 0; d.npblk = (I.NP + 255) / 256;
  const long long N = (long long)I.Nx * I.Nr;
  d.cblk0 = 0; d.ncblk = (int)std::min<long long>((N + 255) / 256, 1024);
  d.Ns = I.Ns;
  d.U0 = I.U0; d.U1 = I.U1;
  d.ax = I.ax; d.ar = I.ar; d.gg = I.gg; d.code = I.code; d.lt = I.lt;
  d.exStart = I.exStart; d.exFaces = I.exFaces; d.fcStart = I.fcStart; d.fcCells = I.fcCells; d.exFlux = I.exFlux;
  d.src = I.src; d.parc = I.parc; d.tab = I.tab; d.st = I.st; d.part = I.part;
  d.c = I.c;
  d.c.NP = I.NP;
  return d;
}

int inst_cfl_init(Inst& I, hipStream_t s) {
  HIPCHK(nzk::launch_cfl_init(I.prec == 64, inst_desc(I), (int)I.partCap, s), "cfl init");
  return 0;
}

int inst_set_state(Inst& I, const float* A, const float* B, hipStream_t s) {
  const long long N = (long long)I.Nx * I.Nr;
  float* d = nullptr;
  HIPCHK(dalloc(I, &d, (size_t)N * 32), "allocate upload");
  int rc = 0;
  do {
    if (hipMemcpy(d, A, (size_t)N * 16, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + N * 4, B, (size_t)N * 16, hipMemcpyHostToDevice) != hipSuccess) {
      rc = nz_fail(-10, "setState: upload failed");
      break;
    }
    if (nzk::launch_unpack(I.prec == 64, I.U0, I.Ns, N, d, d + N * 4, s) != hipSuccess ||
        nzk::launch_unpack(I.prec == 64, I.U1, I.Ns, N, d, d + N * 4, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      rc = nz_fail(-10, "setState: unpack failed");
      break;
    }
  } while (false);
  dfree(I, d);  // the stream is done with it (synchronised above, or a copy failed first)
  if (rc) return rc;
  // zero step info and time; the spray is reset, its seed carries on (spray.js:242-254)
  InstState is{};
  HIPCHK(hipMemcpy(&is, I.st, sizeof(is), hipMemcpyDeviceToHost), "read state info");
  const uint32_t seed = is.seed;
  std::memset(&is, 0, sizeof(is));
  is.seed = seed;
  is.nGroups = I.nGroups;
  HIPCHK(hipMemcpy(I.st, &is, sizeof(is), hipMemcpyHostToDevice), "reset state info");
  if (I.parc) HIPCHK(nzk::launch_spray_reset(I.parc, I.NP, I.src, I.Ns, s), "spray reset");
  if (int r = inst_cfl_init(I, s)) return r;
  HIPCHK(hipStreamSynchronize(s), "setState");
  return 0;
}

void fill_consts(const Inst& I, const nz_params& P, Consts& c, double tEndGrid, double resTol, int minSteps,
                 int maxSteps) {
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
  c.spray = (P.inj && I.haveTab) ? 1 : 0;
  c.NP = I.NP;
  // parcel speeds stay within the injection speeds and the gas speeds (drag relaxes towards
  // the gas, whose speed the update caps at vMax; bounces and the fan do not speed up)
  c.depVel = std::max({P.vMax, I.tabSpeed, std::fabs(P.amb[1]) + std::fabs(P.amb[2]), 1e-6f});
  int lg = 0;
  while ((1LL << lg) < (long long)std::max(I.NP, 1)) ++lg;
  c.depExp = 60 - lg;
  c.tEnd = tEndGrid > 0 ? (float)tEndGrid : 0.f;
  c.resTol = resTol > 0 ? (float)resTol : 0.f;
  c.minSteps = minSteps;
  c.maxSteps = maxSteps;
}

nzk::IntegCtx integ_ctx(const DevInst& d, const nz_probes& pr) {
  nzk::IntegCtx q{};
  q.iExit = pr.iExit; q.jExit = pr.jExit; q.iThroat = pr.iThroat; q.iChamber = pr.iChamber;
  q.inj = d.c.inj;
  if (pr.exact) {
    q.pa = pr.pa;
    for (int k = 0; k < 3; ++k) q.Rs[k] = pr.Rs[k];
    q.RaN = pr.RaN;
  } else {
    q.pa = d.c.amb[3];
    for (int k = 0; k < 3; ++k) q.Rs[k] = d.c.Rs[k];
    q.RaN = d.c.gas[0];
  }
  return q;
}

int inst_integrals(const Inst& I, const DevInst& d, const nz_probes& pr, hipStream_t s, nz_integrals* out) {
  const nzk::IntegCtx q = integ_ctx(d, pr);
  HIPCHK(nzk::launch_integrals(I.prec == 64, d, q, I.dint + 16, I.dint, s), "integrals");
  double v[16];
  HIPCHK(hipMemcpyAsync(v, I.dint, 13 * 8, hipMemcpyDeviceToHost, s), "integrals copy");
  HIPCHK(hipStreamSynchronize(s), "integrals sync");
  fill_integrals(v, pr, out);
  return 0;
}

void fill_integrals(const double* v, const nz_probes& pr, nz_integrals* out) {
  // the device sums leave dx^2 and the reference units out (simsetup.js:333-336, 248-249)
  const double dx2 = pr.dx * pr.dx;
  const double Fsi = v[0] * dx2 * P_REF;
  const double mdotSi = v[1] * dx2 * RHO_REF * U_REF;
  const double f = dx2 * RHO_REF * U_REF;
  const double target = (pr.exact && pr.sprayMdot > 0) ? pr.sprayMdot : 1.0;
  out->F = Fsi;
  out->mdot = mdotSi;
  out->Me = v[1] > 1e-12 ? v[2] / v[1] : 0;
  out->pe = (v[4] > 0 ? v[3] / v[4] : 0) * P_REF;
  out->Isp = mdotSi > 1e-9 ? Fsi / (mdotSi * G0) : 0;
  out->mdotThroat = v[5] * f;
  out->mass = v[6];
  out->energy = v[7];
  out->Pc = v[8] * P_REF;
  out->Tc = v[9] * T_REF;
  out->residual = v[10];
  out->unburntOx = v[11] * f / std::max(target, 1e-9);
  out->fuelVapour = v[12] * f / std::max(target, 1e-9);
}

}  // namespace nz
