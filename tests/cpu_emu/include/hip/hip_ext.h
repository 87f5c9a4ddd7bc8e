// hip_ext.h for the CPU emulation: a launch with start / stop events records
// them around the (synchronous) emulated kernel.
#pragma once
#include <hip/hip_runtime.h>

#define hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, ev0, ev1, flags, ...) \
  do {                                                                                \
    if (ev0) (void)hipEventRecord((ev0), (stream));                                   \
    hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);              \
    if (ev1) (void)hipEventRecord((ev1), (stream));                                   \
  } while (0)
