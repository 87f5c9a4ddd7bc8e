// fl_tuning_table.h -- launch tunings by workload class, chosen by the
// autotuner (fastlem_amd/autotune.py, which calls fl_autotune on each class
// and writes this file).  Rows are tried in order; the first whose
// max_sites and terrain kind match a solve gives its tuning.  A tuning
// shapes launches only: every row gives the same results.
//
// Initial contents: the hand-swept values (per-site workgroups of 256,
// wavefront chains from 6 sites, tail launches for rounds of up to 16 fold
// units, pipelined iterations, batches compacted once 65536 sites idle).
#pragma once

#include <stdint.h>

struct fl_tuning_row {
  uint64_t max_sites;  // union sites up to this many
  uint32_t batch;      // 0: one terrain, 1: several
  uint32_t site_block, long_chain, tail_units, pipelining, compact_min_idle;
};

static const fl_tuning_row fl_tuning_table[] = {
    {UINT64_MAX, 0, 256, 6, 16, 1, 65536},
    {UINT64_MAX, 1, 256, 6, 16, 1, 65536},
};
