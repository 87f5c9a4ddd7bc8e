// fl_internal.h -- shared declarations of the library's translation units.
#pragma once
#include "../../include/fastlem_hip.h"

struct fl_rng;

#ifdef __cplusplus
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

// fl_order.hip: minimum spanning forest with the outlets merged, sorted edges.
int fl_device_msf(hipStream_t s, uint32_t n, uint32_t nnz, const uint32_t* rp, const uint32_t* col, const double* len,
                  const uint8_t* is_outlet, uint32_t r0, std::vector<uint32_t>& a_sorted,
                  std::vector<uint32_t>& b_sorted, bool* applies, std::string* err);
#endif
