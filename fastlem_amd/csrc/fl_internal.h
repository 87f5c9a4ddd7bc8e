// fl_internal.h -- shared declarations of the library's translation units.
#pragma once
#include "../../include/fastlem_hip.h"

struct fl_rng;

#ifdef __cplusplus
#include <hip/hip_runtime.h>

// Set the calling thread's fl_last_error() message.
void fl_internal_set_error(const char* msg);

#include <string>
#include <vector>

// fl_order.hip: minimum spanning forest with the outlets merged, sorted edges.
int fl_device_msf(hipStream_t s, uint32_t n, uint32_t nnz, const uint32_t* rp, const uint32_t* col, const double* len,
                  const std::vector<uint32_t>& init_comp, std::vector<uint32_t>& a_sorted,
                  std::vector<uint32_t>& b_sorted, bool* applies, std::string* err);

// fl_setup.cpp: per-terrain forms on a union graph whose terrain occupies
// sites [s0, s0 + nl) (its rows reference only those sites).
int fl_outlet_pop_order_range(uint32_t s0, uint32_t nl, const uint32_t* row_ptr, const uint32_t* col,
                              const uint32_t* outlets, size_t k, uint32_t* order_out);
int fl_lake_visit_order_range(uint32_t s0, uint32_t nl, const uint32_t* row_ptr, const uint32_t* col,
                              const double* len, const uint32_t* outlets, size_t k, uint32_t rank_base,
                              uint32_t* rank_out);
// Non-outlet visit ranks from the spanning forest (outlets in any order).
// With site_off (groups + 1 entries: group g owns sites [site_off[g],
// site_off[g+1]) and no edge joins two groups) the groups are processed
// independently on the threads; without it, as one forest.
int fl_krt_ranks(size_t n, size_t groups, const uint32_t* outlet_off, const uint32_t* outlets,
                 const uint32_t* rank_off, const uint32_t* site_off, const uint32_t* msf_a, const uint32_t* msf_b,
                 size_t m, uint32_t* rank_out, unsigned threads);
// The one-forest form in two steps: the working arrays set up from the
// sites and outlets alone (while the forest is still being built), then the
// ranks from the forest.  fl_krt_prepare returns null and sets *rc on invalid
// outlets; fl_krt_ranks_prepared consumes the arrays (release them after).
struct fl_krt_work;
fl_krt_work* fl_krt_prepare(size_t n, size_t groups, const uint32_t* outlet_off, const uint32_t* outlets,
                            unsigned threads, int* rc);
int fl_krt_ranks_prepared(fl_krt_work* work, size_t groups, const uint32_t* outlet_off, const uint32_t* outlets,
                          const uint32_t* rank_off, const uint32_t* msf_a, const uint32_t* msf_b, size_t m,
                          uint32_t* rank_out, unsigned threads);
void fl_krt_release(fl_krt_work* work);
// Host threads for the one-time setup work (at most 16).
unsigned fl_host_threads();
extern "C" int fl_ranks_from_msf_groups(size_t n, size_t groups, const uint32_t* outlet_off,
                                        const uint32_t* outlet_order, const uint32_t* rank_off,
                                        const uint32_t* msf_a, const uint32_t* msf_b, size_t m,
                                        uint32_t* rank_out);
#endif
