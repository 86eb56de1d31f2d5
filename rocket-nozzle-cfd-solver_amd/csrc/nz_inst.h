// nz_inst.h — one solve's device-resident state (host side), shared by the
// single-solver C ABI (runtime.hip) and the batched runner (batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "nozzle.h"
#include "nz_device.h"
#include "nz_internal.h"

#define HIPCHK(call, what)                                                                          \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) return nz_fail(-10, std::string(what) + ": " + hipGetErrorString(e_));   \
  } while (0)

namespace nz {

// One device allocation carved into an instance's buffers: first fit on 256-byte
// boundaries, freed blocks coalesced with their neighbours. Allocation and release are
// host bookkeeping only (no runtime call, so no implicit device synchronisation); the
// caller releases a block only once no queued work can still touch it.
class DevArena {
 public:
  int create(size_t bytes);
  void destroy();
  void* alloc(size_t bytes);   // nullptr when no free block is large enough
  void release(void* p);
  size_t used() const { return used_; }
  size_t size() const { return size_; }
  // whether an allocation failed since clear_failed (a caller's "no room now")
  bool failed() const { return failed_; }
  void clear_failed() { failed_ = false; }

 private:
  char* base_ = nullptr;
  size_t size_ = 0, used_ = 0;
  bool failed_ = false;
  std::map<size_t, size_t> free_;  // offset -> bytes
  std::map<size_t, size_t> live_;  // offset -> bytes
};

// code word of a cell whose host is itself (offset fields hold dx+1 = dy+1 = 1)
constexpr uint32_t kZeroOffset = (1u << nzk::C_HX_SHIFT) | (1u << nzk::C_HY_SHIFT);

struct Inst {
  int prec = 32;
  size_t rsz = 4;
  int Nx = 0, Nr = 0;
  long long Ns = 0;
  int shape = 0;                 // tile shape of the stage kernels (stage_shapes.h)
  int TX = 64, TY = 4, tilesX = 0, tilesY = 0;
  void* U0 = nullptr;
  void* U1 = nullptr;
  float *ax = nullptr, *ar = nullptr, *lt = nullptr, *jw = nullptr, *parc = nullptr, *tab = nullptr;
  unsigned long long* src = nullptr;  // spray sources, fixed point
  float tabSpeed = 0.f;               // largest injection speed of the element table
  float4* gg = nullptr;
  uint32_t* code = nullptr;
  int* exStart = nullptr;
  uint2* exFaces = nullptr;
  int* fcStart = nullptr;
  int2* fcCells = nullptr;
  void* exFlux = nullptr;
  nzk::InstState* st = nullptr;  // owned unless placed in a batch's status array
  bool ownSt = true;
  double2* part = nullptr;       // per-tile partials of the last stage 2 (capacity partCap, every shape)
  size_t partCap = 0;
  float* wallrow = nullptr;
  double* dint = nullptr;
  std::vector<uint8_t> mask;
  std::vector<uint32_t> codeH;   // host copy of the code words (tile-shape changes rebuild from it)
  std::vector<float> jWall;
  bool haveGeo = false, haveTab = false;
  int NP = 65536;
  double nGroups = 0;
  int exMax = 0, fcMax = 0;
  nzk::Consts c{};
  size_t bytes = 0;  // device bytes held
  DevArena* arena = nullptr;  // buffers carved from it (batched solves); else hipMalloc / hipFree
};

// device bytes an instance of this size needs (admission by HBM budget)
size_t inst_bytes(int Nx, int Nr, int prec, int NP);
// arena: NULL for buffers of their own, else every buffer comes from it
int inst_alloc(Inst& I, int Nx, int Nr, int prec, nzk::InstState* st, int shape = 0, DevArena* arena = nullptr);
// device memory of an instance: from its arena, else the runtime's allocator
hipError_t inst_dalloc(Inst& I, void** p, size_t bytes);
void inst_dfree(Inst& I, void* p);
// switch the tile shape: the per-tile lists are rebuilt for it
int inst_set_shape(Inst& I, int shape);
void inst_free(Inst& I);
int inst_set_geometry(Inst& I, const uint8_t* mask, const float* geoA, const float* geoG, const float* geoM,
                      const float* jWall);
int inst_set_mask(Inst& I, const uint8_t* mask);
int inst_spray_alloc(Inst& I, int NP, hipStream_t s);
int inst_spray_reset(Inst& I, hipStream_t s);
int inst_set_injector(Inst& I, const float* tab);
// upload A/B and reset step info, time and spray; then the wave-speed max of the state
int inst_set_state(Inst& I, const float* A, const float* B, hipStream_t s);
// recompute the wave-speed snapshot of the current state (after setState / rescale)
int inst_cfl_init(Inst& I, hipStream_t s);
// parameters -> device constants; tEndGrid, resTol, maxSteps <= 0 disable that stop test
void fill_consts(const Inst& I, const nz_params& P, nzk::Consts& c, double tEndGrid, double resTol, int minSteps,
                 int maxSteps);
// descriptor with tile/parcel/cell block offsets 0 (single-instance launches)
nzk::DevInst inst_desc(const Inst& I);
// the integrals kernels' probe positions and gas data (the probes' own when exact)
nzk::IntegCtx integ_ctx(const nzk::DevInst& d, const nz_probes& pr);
// exitPerformance / chamberPerformance sums on the device (FP64), scaled to SI;
// `d` carries the constants in use (ambient pressure); synchronises `s`
int inst_integrals(const Inst& I, const nzk::DevInst& d, const nz_probes& pr, hipStream_t s, nz_integrals* out);
// the 13 device sums of the integrals kernels -> SI quantities
void fill_integrals(const double* v, const nz_probes& pr, nz_integrals* out);

}  // namespace nz
