// stage_shapes.h — the tile shapes the stage kernels are built for, per precision.
// X(name, R, TX, TY, W): stage_inst.hip is compiled once per entry (the build reads
// this list) and defines nzk::stage_launch_<name>. W is the minimum number of waves
// per SIMD the kernels without the viscous split step are compiled for (register
// budget 512 / W; 1 leaves the compiler free). Shape k of a precision is its k-th
// entry here; entry 0 is the default.
#pragma once

#define NZ_STAGE_SHAPES_F32(X)  \
  X(f32_64x4, float, 64, 4, 1)    \
  X(f32_32x8, float, 32, 8, 1)    \
  X(f32_16x16, float, 16, 16, 1)  \
  X(f32_32x4, float, 32, 4, 1)    \
  X(f32_128x2, float, 128, 2, 1)  \
  X(f32_32x8w4, float, 32, 8, 4)  \
  X(f32_16x16w4, float, 16, 16, 4)

#define NZ_STAGE_SHAPES_F64(X)  \
  X(f64_32x4, double, 32, 4, 1)   \
  X(f64_16x8, double, 16, 8, 1)   \
  X(f64_16x4, double, 16, 4, 1)
