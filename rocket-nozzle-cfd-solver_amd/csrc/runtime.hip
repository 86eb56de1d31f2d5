// runtime.hip — the GPUSolver / GPUSpray C ABI over HIP (include/nozzle.h).
//
// One solver instance (nz::Inst, instance.hip) driven on its own HIP stream:
// steps are enqueued without host waits (GPUSolver.step is asynchronous,
// solver.js:849-891), and synchronous and fenced asynchronous read-backs are
// served from device staging (solver.js:1016-1075).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "nz_inst.h"
#include "nz_launch.h"

using nzk::Consts;
using nzk::DevInst;
using nzk::InstState;

struct nz_solver {
  int prec = 32;
  int device = 0;
  hipStream_t stream = nullptr;
  nz::Inst I;
  DevInst* dinst = nullptr;  // device copy of the descriptor the kernels read
  DevInst h{};               // host copy of the descriptor last sent to the device
  bool instSent = false;
  float* dstage = nullptr;   // device staging (pack / unpack)
  size_t dstageCap = 0;
  float* hpin = nullptr;     // pinned host staging
  size_t hpinCap = 0;
  bool haveState = false;
  int64_t stepCount = 0;
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

void free_buffers(nz_solver* s) {
  nz::inst_free(s->I);
  if (s->dinst) (void)hipFree(s->dinst);
  if (s->dstage) (void)hipFree(s->dstage);
  if (s->hpin) (void)hipHostFree(s->hpin);
  s->dinst = nullptr;
  s->dstage = nullptr;
  s->hpin = nullptr;
  s->hpinCap = s->dstageCap = 0;
  s->instSent = false;
  s->haveState = false;
}

int ensure_staging(nz_solver* s, size_t floats) {
  if (floats > s->dstageCap) {
    if (s->dstage) (void)hipFree(s->dstage);
    s->dstage = nullptr;
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

// descriptor from the instance's current buffers, keeping the constants last sent
int send_inst(nz_solver* s) {
  const Consts c = s->h.c;
  s->h = nz::inst_desc(s->I);
  s->h.c = c;
  s->h.c.NP = s->I.NP;
  HIPCHK(nzk::launch_set_inst(s->dinst, s->h, s->stream), "set instance");
  s->instSent = true;
  return 0;
}

int check_solver(nz_solver* s, const char* what) {
  if (!s) return nz_fail(-1, std::string(what) + ": null solver");
  if (s->I.Nx <= 0) return nz_fail(-4, std::string(what) + ": resize() has not been called");
  return 0;
}

int apply_params(nz_solver* s, const nz_params* P) {
  Consts c;
  nz::fill_consts(s->I, *P, c, 0.0, 0.0, 0, 0);
  if (!s->instSent || std::memcmp(&c, &s->h.c, sizeof(c)) != 0) {
    s->h.c = c;
    return send_inst(s);
  }
  return 0;
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
  // the parcel count and the tile shape carry over a resize
  const int NP = s->I.NP, shape = s->I.shape;
  free_buffers(s);
  s->I.NP = NP;
  if (int rc = nz::inst_alloc(s->I, Nx, Nr, s->prec, nullptr, shape)) return rc;
  HIPCHK(hipMalloc(&s->dinst, sizeof(DevInst)), "hipMalloc instance");
  s->stepCount = 0;
  std::memset(&s->h, 0, sizeof(s->h));
  if (int rc = nz::inst_spray_alloc(s->I, NP, s->stream)) return rc;
  if (int rc = send_inst(s)) return rc;
  HIPCHK(hipStreamSynchronize(s->stream), "nz_resize");
  return 0;
}

int nz_get_size(const nz_solver* s, int* Nx, int* Nr) {
  if (!s || !Nx || !Nr) return nz_fail(-1, "nz_get_size: null argument");
  *Nx = s->I.Nx;
  *Nr = s->I.Nr;
  return 0;
}

int nz_set_geometry(nz_solver* s, const uint8_t* mask, const float* geoA, const float* geoG, const float* geoM,
                    const float* jWall) {
  if (int rc = check_solver(s, "nz_set_geometry")) return rc;
  if (!mask || !geoA || !geoG || !geoM) return nz_fail(-1, "nz_set_geometry: null array");
  HIPCHK(hipStreamSynchronize(s->stream), "nz_set_geometry: sync");
  if (int rc = nz::inst_set_geometry(s->I, mask, geoA, geoG, geoM, jWall)) return rc;
  if (int rc = send_inst(s)) return rc;
  // a running flow carries on (app.js:254-270): the next time step comes from the current
  // state under the new mask, as the reference's reduction takes it every step
  if (s->haveState) return nz::inst_cfl_init(s->I, s->stream);
  return 0;
}

int nz_set_mask(nz_solver* s, const uint8_t* mask) {
  if (int rc = check_solver(s, "nz_set_mask")) return rc;
  if (!mask) return nz_fail(-1, "nz_set_mask: null mask");
  HIPCHK(hipStreamSynchronize(s->stream), "nz_set_mask: sync");
  return nz::inst_set_mask(s->I, mask);
}

int nz_set_state(nz_solver* s, const float* A, const float* B) {
  if (int rc = check_solver(s, "nz_set_state")) return rc;
  if (!A || !B) return nz_fail(-1, "nz_set_state: null array");
  nz_read_async_cancel(s);
  HIPCHK(hipStreamSynchronize(s->stream), "nz_set_state: sync");
  if (int rc = nz::inst_set_state(s->I, A, B, s->stream)) return rc;
  s->stepCount = 0;
  s->haveState = true;
  return 0;
}

// ev (NULL or 6 events): the start / end of stage 1, the spray and stage 2 kernels
static int enqueue_step(nz_solver* s, bool phys, bool spray, hipEvent_t* ev) {
  const bool fp64 = s->prec == 64;
  const int ntiles = s->I.tilesX * s->I.tilesY;
  HIPCHK(nzk::launch_stage(1, false, fp64, s->I.shape, s->dinst, 1, ntiles, s->I.tilesX, nullptr, s->stream,
                           ev ? ev[0] : nullptr, ev ? ev[1] : nullptr), "stage 1");
  if (spray)
    HIPCHK(nzk::launch_spray(fp64, s->dinst, 1, s->h.npblk, s->stream, ev ? ev[2] : nullptr, ev ? ev[3] : nullptr),
           "spray");
  HIPCHK(nzk::launch_stage(2, phys, fp64, s->I.shape, s->dinst, 1, ntiles, s->I.tilesX, nullptr, s->stream,
                           ev ? ev[4] : nullptr, ev ? ev[5] : nullptr), "stage 2");
  // more partials than the next stage 1 folds itself: fold them now
  if ((unsigned)ntiles > nzk::FUSED_CLOSE_MAX) HIPCHK(nzk::launch_close(s->dinst, 1, s->stream), "close");
  return 0;
}

int nz_step(nz_solver* s, int n, const nz_params* P) {
  if (int rc = check_solver(s, "nz_step")) return rc;
  if (!P) return nz_fail(-1, "nz_step: null params");
  if (n < 0) return nz_fail(-1, "nz_step: negative step count");
  if (!s->I.haveGeo) return nz_fail(-4, "nz_step: setGeometry() has not been called");
  if (int rc = apply_params(s, P)) return rc;
  const bool phys = P->visc || P->chem;
  const bool spray = s->h.c.spray != 0;
  for (int k = 0; k < n; ++k)
    if (int rc = enqueue_step(s, phys, spray, nullptr)) return rc;
  // fold the last step's partials, so that the step state the host reads is current
  if (n > 0) HIPCHK(nzk::launch_close(s->dinst, 1, s->stream), "close");
  s->stepCount += n;
  return 0;
}

int nz_step_timed(nz_solver* s, int n, const nz_params* P, double* ms) {
  if (int rc = check_solver(s, "nz_step_timed")) return rc;
  if (!P || !ms) return nz_fail(-1, "nz_step_timed: null argument");
  if (n <= 0) return nz_fail(-1, "nz_step_timed: step count must be positive");
  if (!s->I.haveGeo) return nz_fail(-4, "nz_step_timed: setGeometry() has not been called");
  if (int rc = apply_params(s, P)) return rc;
  const bool phys = P->visc || P->chem;
  const bool spray = s->h.c.spray != 0;
  // per step: the kernels' own start / end (hipExtLaunchKernel events), and the span
  std::vector<hipEvent_t> ev((size_t)n * 6 + 2);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e), "event create");
  hipEvent_t* span = &ev[(size_t)n * 6];
  int rc = 0;
  if (hipEventRecord(span[0], s->stream) != hipSuccess) rc = nz_fail(-10, "event");
  for (int k = 0; k < n && !rc; ++k) rc = enqueue_step(s, phys, spray, &ev[(size_t)k * 6]);
  if (!rc && hipEventRecord(span[1], s->stream) != hipSuccess) rc = nz_fail(-10, "event");
  if (!rc && nzk::launch_close(s->dinst, 1, s->stream) != hipSuccess) rc = nz_fail(-10, "close");
  HIPCHK(hipStreamSynchronize(s->stream), "nz_step_timed sync");
  double t[3] = {0, 0, 0};
  float all = 0.f;
  if (!rc) {
    for (int k = 0; k < n; ++k)
      for (int q = 0; q < 3; ++q) {
        if (q == 1 && !spray) continue;
        float e = 0.f;
        if (hipEventElapsedTime(&e, ev[(size_t)k * 6 + 2 * q], ev[(size_t)k * 6 + 2 * q + 1]) != hipSuccess)
          rc = nz_fail(-10, "event elapsed");
        t[q] += e;
      }
    if (hipEventElapsedTime(&all, span[0], span[1]) != hipSuccess) rc = nz_fail(-10, "event elapsed");
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc) return rc;
  ms[0] = t[0]; ms[1] = t[1]; ms[2] = t[2]; ms[3] = all;
  s->stepCount += n;
  return 0;
}

int nz_tile_shapes(int precision, int* tx, int* ty, int* waves, int cap) {
  if (precision != 32 && precision != 64) return nz_fail(-1, "nz_tile_shapes: precision must be 32 or 64");
  const int n = nzk::tile_shape_count(precision == 64);
  for (int k = 0; k < n && k < cap; ++k) {
    int a = 0, b = 0, w = 1;
    nzk::tile_shape(precision == 64, k, &a, &b, &w);
    if (tx) tx[k] = a;
    if (ty) ty[k] = b;
    if (waves) waves[k] = w;
  }
  return n;
}

int nz_set_tile_shape(nz_solver* s, int shape) {
  if (int rc = check_solver(s, "nz_set_tile_shape")) return rc;
  HIPCHK(hipStreamSynchronize(s->stream), "nz_set_tile_shape: sync");
  if (int rc = nz::inst_set_shape(s->I, shape)) return rc;
  return send_inst(s);
}

int nz_get_tile_shape(const nz_solver* s, int* shape) {
  if (!s || !shape) return nz_fail(-1, "nz_get_tile_shape: null argument");
  *shape = s->I.shape;
  return 0;
}

namespace {
// device copy of everything a step changes, for trial runs that must leave no trace
struct Snapshot {
  void* U0 = nullptr;
  void* U1 = nullptr;
  float* parc = nullptr;
  void* src = nullptr;
  void* part = nullptr;  // the last step's tile partials, folded by the next step
  InstState st{};
  ~Snapshot() {
    if (U0) (void)hipFree(U0);
    if (U1) (void)hipFree(U1);
    if (parc) (void)hipFree(parc);
    if (src) (void)hipFree(src);
    if (part) (void)hipFree(part);
  }
};
int snap_take(nz_solver* s, Snapshot& z) {
  const nz::Inst& I = s->I;
  // the partials in the form every tiling reads, so that a restore under another shape holds
  HIPCHK(nzk::launch_part_canon(I.part, I.tilesX * I.tilesY, (int)I.partCap, s->stream), "snapshot partials");
  const size_t us = (size_t)I.Ns * 8 * I.rsz, ps = (size_t)I.NP * 8 * 4, ss = (size_t)I.Ns * 32;
  HIPCHK(hipMalloc(&z.U0, us), "snapshot");
  HIPCHK(hipMalloc(&z.U1, us), "snapshot");
  HIPCHK(hipMalloc(&z.parc, ps), "snapshot");
  HIPCHK(hipMalloc(&z.src, ss), "snapshot");
  HIPCHK(hipMemcpyAsync(z.U0, I.U0, us, hipMemcpyDeviceToDevice, s->stream), "snapshot");
  HIPCHK(hipMemcpyAsync(z.U1, I.U1, us, hipMemcpyDeviceToDevice, s->stream), "snapshot");
  HIPCHK(hipMemcpyAsync(z.parc, I.parc, ps, hipMemcpyDeviceToDevice, s->stream), "snapshot");
  HIPCHK(hipMemcpyAsync(z.src, I.src, ss, hipMemcpyDeviceToDevice, s->stream), "snapshot");
  HIPCHK(hipMalloc(&z.part, I.partCap * sizeof(double2)), "snapshot");
  HIPCHK(hipMemcpyAsync(z.part, I.part, I.partCap * sizeof(double2), hipMemcpyDeviceToDevice, s->stream), "snapshot");
  HIPCHK(hipMemcpyAsync(&z.st, I.st, sizeof(InstState), hipMemcpyDeviceToHost, s->stream), "snapshot");
  HIPCHK(hipStreamSynchronize(s->stream), "snapshot");
  return 0;
}
int snap_restore(nz_solver* s, const Snapshot& z) {
  const nz::Inst& I = s->I;
  const size_t us = (size_t)I.Ns * 8 * I.rsz, ps = (size_t)I.NP * 8 * 4, ss = (size_t)I.Ns * 32;
  HIPCHK(hipMemcpyAsync(I.U0, z.U0, us, hipMemcpyDeviceToDevice, s->stream), "restore");
  HIPCHK(hipMemcpyAsync(I.U1, z.U1, us, hipMemcpyDeviceToDevice, s->stream), "restore");
  HIPCHK(hipMemcpyAsync(I.parc, z.parc, ps, hipMemcpyDeviceToDevice, s->stream), "restore");
  HIPCHK(hipMemcpyAsync(I.src, z.src, ss, hipMemcpyDeviceToDevice, s->stream), "restore");
  HIPCHK(hipMemcpyAsync(I.part, z.part, I.partCap * sizeof(double2), hipMemcpyDeviceToDevice, s->stream), "restore");
  HIPCHK(hipMemcpyAsync(I.st, &z.st, sizeof(InstState), hipMemcpyHostToDevice, s->stream), "restore");
  HIPCHK(hipStreamSynchronize(s->stream), "restore");
  return 0;
}
}  // namespace

int nz_autotune(nz_solver* s, const nz_params* P, int steps, int* best, double* us_per_step) {
  if (int rc = check_solver(s, "nz_autotune")) return rc;
  if (!P) return nz_fail(-1, "nz_autotune: null params");
  if (steps < 1) return nz_fail(-1, "nz_autotune: step count must be positive");
  if (!s->I.haveGeo || !s->haveState) return nz_fail(-4, "nz_autotune: setGeometry() and setState() first");
  nz_read_async_cancel(s);
  HIPCHK(hipStreamSynchronize(s->stream), "nz_autotune: sync");
  Snapshot z;
  if (int rc = snap_take(s, z)) return rc;
  const int n = nzk::tile_shape_count(s->prec == 64), start = s->I.shape;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0), "event");
  if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return nz_fail(-10, "event"); }
  int pick = start, rc = 0;
  double bestUs = 1e300;
  for (int k = 0; k < n && !rc; ++k) {
    if ((rc = nz_set_tile_shape(s, k))) break;
    if ((rc = apply_params(s, P))) break;
    const bool phys = P->visc || P->chem, spray = s->h.c.spray != 0;
    // one untimed step loads the shape's code object
    if ((rc = enqueue_step(s, phys, spray, nullptr)) || (rc = snap_restore(s, z))) break;
    if (hipEventRecord(e0, s->stream) != hipSuccess) { rc = nz_fail(-10, "event"); break; }
    for (int q = 0; q < steps && !rc; ++q) rc = enqueue_step(s, phys, spray, nullptr);
    if (rc) break;
    float ms = 0.f;
    if (hipEventRecord(e1, s->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = nz_fail(-10, "autotune timing"); break; }
    const double us = ms * 1e3 / steps;
    if (us_per_step) us_per_step[k] = us;
    if (us < bestUs) { bestUs = us; pick = k; }
    rc = snap_restore(s, z);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {
    (void)nz_set_tile_shape(s, start);
    (void)snap_restore(s, z);
    return rc;
  }
  if (int r = nz_set_tile_shape(s, pick)) return r;
  if (best) *best = pick;
  return 0;
}

int nz_rescale_ambient(nz_solver* s, float fp, float fT, float du, const nz_params* P) {
  if (int rc = check_solver(s, "nz_rescale_ambient")) return rc;
  if (!P) return nz_fail(-1, "nz_rescale_ambient: null params");
  if (int rc = apply_params(s, P)) return rc;
  HIPCHK(nzk::launch_rescale(s->prec == 64, s->h, fp, fT, du, s->stream), "rescale");
  HIPCHK(nzk::launch_cfl_init(s->prec == 64, s->h, (int)s->I.partCap, s->stream), "cfl init");
  return 0;
}

int nz_draw_wall_row(nz_solver* s) {
  if (int rc = check_solver(s, "nz_draw_wall_row")) return rc;
  HIPCHK(nzk::launch_wall_row(s->prec == 64, s->h, s->I.jw, s->I.wallrow, s->stream), "wall row");
  return 0;
}

// enqueue one rectangle read into device staging at `off` (floats); returns its size or < 0
static long long enqueue_rect(nz_solver* s, int field, int x, int y, int w, int h, float* dst) {
  const int Nx = s->I.Nx, Nr = s->I.Nr;
  int FW = Nx, FH = Nr;
  if (field == NZ_FIELD_STEP) { FW = 1; FH = 1; }
  else if (field == NZ_FIELD_WALLROW) { FW = Nx; FH = 1; }
  else if (field != NZ_FIELD_A && field != NZ_FIELD_B && field != NZ_FIELD_T2) return nz_fail(-1, "read: unknown field");
  if (w <= 0 || h <= 0 || x < 0 || y < 0 || x + w > FW || y + h > FH) return nz_fail(-1, "read: rectangle outside the field");
  const bool fp64 = s->prec == 64;
  if (field == NZ_FIELD_A || field == NZ_FIELD_B) {
    if (nzk::launch_pack(fp64, s->I.U0, s->I.Ns, Nx, field == NZ_FIELD_A ? 0 : 4, x, y, w, h, dst, s->stream) != hipSuccess)
      return nz_fail(-10, "read: pack failed");
  } else if (field == NZ_FIELD_STEP) {
    if (nzk::launch_step_info(s->I.st, dst, s->stream) != hipSuccess) return nz_fail(-10, "read: step info failed");
  } else if (field == NZ_FIELD_WALLROW) {
    for (int r = 0; r < h; ++r)
      if (hipMemcpyAsync(dst + (size_t)r * w * 4, s->I.wallrow + (size_t)x * 4, (size_t)w * 16, hipMemcpyDeviceToDevice, s->stream) != hipSuccess)
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
  const size_t extra = field == NZ_FIELD_T2 ? (size_t)s->I.Nx * s->I.Nr * 4 : 0;
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
  const long long N = (long long)s->I.Nx * s->I.Nr;
  if (int rc = ensure_staging(s, (size_t)N * 16)) return rc;
  double* d = reinterpret_cast<double*>(s->dstage);
  HIPCHK(nzk::launch_pack_f64(s->prec == 64, s->I.U0, s->I.Ns, N, d, d + N * 4, s->stream), "pack f64");
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
  const size_t extra = t2 ? (size_t)s->I.Nx * s->I.Nr * 4 : 0;
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
  if (int rc = nz::inst_set_injector(s->I, tab)) return rc;
  s->instSent = false;  // spray flag of the constants may change
  return 0;
}

int nz_spray_reset(nz_solver* s) {
  if (int rc = check_solver(s, "nz_spray_reset")) return rc;
  return nz::inst_spray_reset(s->I, s->stream);
}

int nz_spray_set_count(nz_solver* s, int n) {
  if (int rc = check_solver(s, "nz_spray_set_count")) return rc;
  if (n < 4 || n > (1 << 22) || n % 4 != 0) return nz_fail(-1, "nz_spray_set_count: count must be a multiple of 4 in [4, 2^22]");
  HIPCHK(hipStreamSynchronize(s->stream), "spray count sync");
  if (int rc = nz::inst_spray_alloc(s->I, n, s->stream)) return rc;
  return send_inst(s);
}

int nz_spray_read(nz_solver* s, float* PA, float* PB) {
  if (int rc = check_solver(s, "nz_spray_read")) return rc;
  const int NP = s->I.NP;
  std::vector<float> p((size_t)NP * 8);
  HIPCHK(hipStreamSynchronize(s->stream), "spray read sync");
  HIPCHK(hipMemcpy(p.data(), s->I.parc, p.size() * 4, hipMemcpyDeviceToHost), "spray read");
  for (int k = 0; k < NP; ++k)
    for (int q = 0; q < 4; ++q) {
      if (PA) PA[(size_t)k * 4 + q] = p[(size_t)q * NP + k];
      if (PB) PB[(size_t)k * 4 + q] = p[(size_t)(4 + q) * NP + k];
    }
  return 0;
}

int nz_spray_stats(nz_solver* s, nz_spray_stats_t* out) {
  if (!out) return nz_fail(-1, "nz_spray_stats: null output");
  const int NP = s ? s->I.NP : 0;
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
  const int Nx = s->I.Nx, Nr = s->I.Nr;
  if (pr->iExit < 0 || pr->iExit + 1 > Nx || pr->jExit < 0 || pr->jExit > Nr || pr->iThroat < 0 ||
      pr->iThroat >= Nx || pr->iChamber < 0 || pr->iChamber >= Nx)
    return nz_fail(-1, "nz_set_probes: probe outside the grid");
  s->probes = *pr;
  s->haveProbes = true;
  return 0;
}

int nz_read_derived(nz_solver* s, const nz_derived_params* d, float* out) {
  if (int rc = check_solver(s, "nz_read_derived")) return rc;
  if (!d || !out) return nz_fail(-1, "nz_read_derived: null argument");
  if (d->field < 0 || d->field > 13) return nz_fail(-1, "nz_read_derived: field must be in [0, 13]");
  const size_t N = (size_t)s->I.Nx * s->I.Nr;
  if (int rc = ensure_staging(s, N * 8)) return rc;
  nzk::DerivedC c;
  c.field = d->field;
  c.r0 = d->range[0]; c.r1 = d->range[1];
  c.RaN = d->RaN;
  for (int q = 0; q < 3; ++q) c.Rs[q] = d->Rs[q];
  for (int q = 0; q < 4; ++q) c.chemK[q] = d->chemK[q];
  c.Tc = d->Tc; c.Ta = d->Ta; c.schl = d->schlGain; c.dx = d->dx;
  c.vel = (float)nz::U_REF;
  const bool fp64 = s->prec == 64;
  float* t2 = s->dstage + N * 4;
  if (d->field == 9 || d->field == 13) HIPCHK(nzk::launch_t2(fp64, s->h, t2, s->stream), "derived: T2");
  HIPCHK(nzk::launch_derived(fp64, s->h, c, t2, s->dstage, s->stream), "derived");
  HIPCHK(hipMemcpyAsync(s->hpin, s->dstage, N * 16, hipMemcpyDeviceToHost, s->stream), "derived copy");
  HIPCHK(hipStreamSynchronize(s->stream), "derived sync");
  std::memcpy(out, s->hpin, N * 16);
  return 0;
}

int nz_integrals_read(nz_solver* s, const nz_params* P, nz_integrals* out) {
  if (int rc = check_solver(s, "nz_integrals_read")) return rc;
  if (!out || !P) return nz_fail(-1, "nz_integrals_read: null argument");
  if (!s->haveProbes) return nz_fail(-4, "nz_integrals_read: probes not set");
  if (int rc = apply_params(s, P)) return rc;
  return nz::inst_integrals(s->I, s->h, s->probes, s->stream, out);
}

}  // extern "C"
