// fl_tuning_table.h -- launch tunings by workload class, chosen by the
// autotuner (fastlem_amd/autotune.py, which calls fl_autotune on each class
// and writes this file).  Rows are tried in order; the first whose
// max_sites and terrain kind match a solve gives its tuning.  A tuning
// shapes launches only: every row gives the same results.
//
// Written by fastlem_amd/autotune.py on 2026-12-03 (one MI355X).
#pragma once

#include <stdint.h>

struct fl_tuning_row {
  uint64_t max_sites;  // union sites up to this many
  uint32_t batch;      // 0: one terrain, 1: several
  uint32_t site_block, long_chain, tail_units, pipelining, compact_min_idle;
};

static const fl_tuning_row fl_tuning_table[] = {
    {32768ull, 0, 128, 6, 32, 1, 65536u},  // terrain 16384 sites: 2.41 -> 2.35 ms
    {131072ull, 0, 256, 4, 32, 1, 65536u},  // terrain 65536 sites: 5.46 -> 5.46 ms
    {524288ull, 0, 256, 6, 32, 1, 65536u},  // terrain 262144 sites: 18.23 -> 18.23 ms
    {UINT64_MAX, 0, 256, 6, 16, 1, 65536u},  // terrain 1048576 sites: 48.06 -> 48.06 ms
    {1317794ull, 1, 128, 6, 16, 1, 65536u},  // batch of 32 terrains, 1000..65536 sites (540264 in all): 33.04 -> 32.47 ms
    {UINT64_MAX, 1, 256, 6, 16, 1, 65536u},  // batch of 64 terrains, 1000..262144 sites (3214323 in all): 216.62 -> 216.62 ms
};
