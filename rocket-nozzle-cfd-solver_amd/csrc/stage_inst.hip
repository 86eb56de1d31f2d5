// stage_inst.hip — one tile shape's stage kernels and their launcher. Compiled
// once per entry of stage_shapes.h with -DNZ_STAGE_NAME=<name> -DNZ_STAGE_R=<type>
// -DNZ_STAGE_TX=<tx> -DNZ_STAGE_TY=<ty> -DNZ_STAGE_W=<waves>, so the shapes build in parallel.
#include <hip/hip_ext.h>

#include "nz_stage.h"

#if !defined(NZ_STAGE_NAME) || !defined(NZ_STAGE_R) || !defined(NZ_STAGE_TX) || !defined(NZ_STAGE_TY) || \
    !defined(NZ_STAGE_W)
#error "stage_inst.hip needs NZ_STAGE_NAME, NZ_STAGE_R, NZ_STAGE_TX, NZ_STAGE_TY, NZ_STAGE_W"
#endif

#define NZ_CAT2(a, b) a##b
#define NZ_CAT(a, b) NZ_CAT2(a, b)

namespace nzk {

namespace {
// with events, the kernel's own start / end are recorded (the dispatch packet carries them)
template <class K>
hipError_t go(K k, dim3 g, dim3 b, hipStream_t s, hipEvent_t e0, hipEvent_t e1, const DevInst* d, int ninst, int chunk,
              const unsigned short* ti) {
  if (e0 && e1) hipExtLaunchKernelGGL(k, g, b, 0, s, e0, e1, 0, d, ninst, chunk, ti);
  else hipLaunchKernelGGL(k, g, b, 0, s, d, ninst, chunk, ti);
  return hipGetLastError();
}
}  // namespace

hipError_t NZ_CAT(stage_launch_, NZ_STAGE_NAME)(int stage, bool phys, const DevInst* d, int ninst, int ntiles, int chunk,
                                                 const unsigned short* ti, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  using R = NZ_STAGE_R;
  constexpr int TX = NZ_STAGE_TX, TY = NZ_STAGE_TY, W = NZ_STAGE_W;
  const dim3 g(ntiles), b(TX * TY);
  if (stage == 1) return go(k_stage<R, TX, TY, 1, false, W>, g, b, s, e0, e1, d, ninst, chunk, ti);
  if (phys) return go(k_stage<R, TX, TY, 2, true, W>, g, b, s, e0, e1, d, ninst, chunk, ti);
  return go(k_stage<R, TX, TY, 2, false, W>, g, b, s, e0, e1, d, ninst, chunk, ti);
}

}  // namespace nzk
