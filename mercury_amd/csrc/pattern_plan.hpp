// Host-side task plan of one of the P1..P8 kernels: every directed edge
// (the reference's COO task list, graph.cc:336-365; P4 only those with
// v1 < v0) and, for P1, every vertex, heavy ones split into shares (a share
// takes every nsub-th index of the task's outer loops), ordered by the
// pattern's own bound on its pair tests and packed into queue groups
// (heavy tasks alone, light ones together).
#pragma once

#include <cstdint>
#include <vector>

#include "host_graph.hpp"

namespace mgpm {

// One share of a task: edge (v0, v1), or vertex v0 (v1 unused).
struct PatternTask {
  int32_t v0, v1;
  int32_t sub, nsub;  // share sub of nsub (0 <= sub < nsub)
};

// Splitting: a task whose estimated pair tests exceed this many is cut
// into up to kPatternMaxSplit shares.
constexpr uint64_t kPatternSplitWork = uint64_t(1) << 20;
constexpr int32_t kPatternMaxSplit = 1024;

struct PatternPlan {
  std::vector<PatternTask> edges;
  std::vector<int32_t> edge_groups;  // offsets, size ngroups + 1
  std::vector<uint64_t> edge_group_cost;
  std::vector<PatternTask> verts;
  std::vector<int32_t> vert_groups;
  std::vector<uint64_t> vert_group_cost;
  int32_t max_degree = 0;
};

// Plan of pattern `pattern` (1..8).  split_work = 0 takes
// kPatternSplitWork; a huge value disables splits.
void build_pattern_plan(const CsrGraph &g, int pattern, PatternPlan &plan,
                        uint64_t split_work = 0);

}  // namespace mgpm
