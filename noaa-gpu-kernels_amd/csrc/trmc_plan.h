// Host-side schedule planning for the time-skewed MC sweep.
//
// A river network arrives in the reference's reach CSR + level layout
// (wavefront_mc.cu:577-632).  The plan re-cuts every reach into work units of
// at most `unit_segments` consecutive segments.  Splitting a reach is exact:
// segment i > 0 of a reach takes qup = previous-step qdc of segment i-1 and
// quc = current-step qdc of segment i-1 (wavefront_mc.cu:388-389), which is
// precisely what a single-upstream reach boundary supplies.  Internal unit
// boundaries pass the value directly (no 0.0 + x summation), reach heads sum
// their upstream reaches in CSR order starting from 0.0 (:523-529).
//
// Units get ALAP levels on the unit DAG (sinks at L-1, every other unit one
// level above its earliest downstream unit), so in a tree every edge spans
// exactly one level.  Unit u at step t is scheduled at super-level
// k = level(u) + t; all items of one super-level are independent, giving
// L + T - 1 super-levels instead of L * T level sweeps.  Unit outflows live in
// a ring of depth max_span + 2 indexed by (t + 1) mod ring (slot of t = -1
// holds the initial value).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace trmc {

struct Plan {
    int n_reaches = 0, n_segments = 0;
    int unit_segments = 1;
    int n_units = 0, n_levels = 0;   // unit levels (ALAP)
    int max_span = 1, ring = 3;
    int max_upstream = 0;
    std::vector<int> level_ptr;      // [n_levels+1] device unit ranges per level
    std::vector<int> unit_level;     // [U]
    std::vector<int> unit_seg;       // [U+1] device segment offsets
    std::vector<int> up_ptr;         // [U+1]
    std::vector<int> up_idx;         // [nnz] device unit ids, reference CSR order
    std::vector<int> unit_flags;     // bit0: internal boundary (single upstream, no sum)
    std::vector<int> unit_reach;     // [U] reach id if unit is its reach's last chunk, else -1
    std::vector<int> unit_init;      // [U] source of the t=-1 outflow: >=0 original
                                     // segment (qdp0), <0: reach_last0[-(v+1)]
    std::vector<int> seg_of_dev;     // [N_cov] original segment of each device segment
    std::vector<int> dev_of_seg;     // [N] device segment or -1 (segment in no reach)
    std::vector<int> reach_last_unit;// [R] device unit holding each reach's last segment
    int n_dev_segments = 0;          // covered segments
    // packed per-unit record for the dataflow kernels, 8 ints per unit:
    // level, first segment, end segment, flags | n_up << 8, up[0..3]
    // (flags bit 0: single-upstream internal boundary; bit 1: no slack, the
    // unit lies on a longest dependency chain)
    // (upstream entries beyond n_up are 0; empty unless max_upstream <= 4)
    std::vector<int> unit_rec;
};

// Validates the reference layout and builds the plan.  Returns 0 or a
// negative TRMC_ERR_* code with a message in err.
int build_plan(const int *reach_seg_start, const int *reach_seg_len,
               const int *reach_up_ptr, const int *reach_up_idx,
               const int *level_ptr, const int *level_reach,
               int n_reaches, int n_levels, int n_segments, int unit_segments,
               Plan &out, std::string &err);

// Items (unit x member) of the widest super-level for T steps.
int64_t max_window_items(const Plan &p, int n_steps, int n_members);

}  // namespace trmc
