// nz_launch.h — host-callable launchers of the device kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "nz_device.h"

namespace nzk {

// tile shapes of the fused stage kernels (TX cells wide in x, TY in r; one thread
// per cell), listed per precision in stage_shapes.h; shape 0 is the default
int tile_shape_count(bool fp64);
bool tile_shape(bool fp64, int shape, int* tx, int* ty, int* waves = nullptr);
hipError_t launch_set_inst(DevInst* dst, const DevInst& src, hipStream_t s);
// tilesX0: tile columns of the first instance (sets the XCD chunk); tileInst: instance of
// each of the ntiles launch tiles (device memory; may be null when ninst is 1). With events
// e0/e1, the kernel's own start and end are recorded into them.
hipError_t launch_stage(int stage, bool phys, bool fp64, int shape, const DevInst* d, int ninst, int ntiles,
                        int tilesX0, const unsigned short* tileInst, hipStream_t s, hipEvent_t e0 = nullptr,
                        hipEvent_t e1 = nullptr);
hipError_t launch_close(const DevInst* d, int ninst, hipStream_t s);
hipError_t launch_spray(bool fp64, const DevInst* d, int ninst, int nblk, hipStream_t s, hipEvent_t e0 = nullptr,
                        hipEvent_t e1 = nullptr);
// wave-speed max of the current state; leaves the partials folded into entry 0, zeros up to partCap
hipError_t launch_cfl_init(bool fp64, const DevInst& I, int partCap, hipStream_t s);
// the n partials of a tiling folded into entry 0, zeros up to cap (before a tile-shape change)
hipError_t launch_part_canon(double2* part, int n, int cap, hipStream_t s);
hipError_t launch_rescale(bool fp64, const DevInst& I, float fp, float fT, float du, hipStream_t s);
hipError_t launch_wall_row(bool fp64, const DevInst& I, const float* jw, float* out, hipStream_t s);
hipError_t launch_t2(bool fp64, const DevInst& I, float* out, hipStream_t s);
hipError_t launch_pack(bool fp64, const void* U, long long Ns, int Nx, int p0, int x, int y, int w, int h, float* out, hipStream_t s);
hipError_t launch_unpack(bool fp64, void* U, long long Ns, long long N, const float* A, const float* B, hipStream_t s);
hipError_t launch_step_info(const InstState* st, float* out, hipStream_t s);
hipError_t launch_pack_f64(bool fp64, const void* U, long long Ns, long long N, double* A, double* B, hipStream_t s);
hipError_t launch_spray_reset(float* parc, int NP, unsigned long long* src, long long Ns, hipStream_t s);
hipError_t launch_derived(bool fp64, const DevInst& I, const DerivedC& d, const float* T2, float* out, hipStream_t s);
// scratch: 2 * INTEG_BLOCKS_MAX doubles; out: 13 doubles (see inst_integrals)
hipError_t launch_integrals(bool fp64, const DevInst& I, const IntegCtx& q, double* scratch, double* out, hipStream_t s);

}  // namespace nzk
