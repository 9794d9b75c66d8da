// Host-side task plan of the P1..P8 kernels: every directed edge (the
// reference's COO task list, graph.cc:336-365) and, for P1, every vertex,
// ordered by an estimate of their set-algebra cost and packed into queue
// groups (heavy tasks alone, light ones together).
#pragma once

#include <cstdint>
#include <vector>

#include "host_graph.hpp"

namespace mgpm {

struct PatternTaskEdge {
  int32_t v0, v1;
};

struct PatternPlan {
  std::vector<PatternTaskEdge> edges;
  std::vector<int32_t> edge_groups;  // offsets, size ngroups + 1
  std::vector<uint64_t> edge_group_cost;
  std::vector<int32_t> verts;
  std::vector<int32_t> vert_groups;
  std::vector<uint64_t> vert_group_cost;
  int32_t max_degree = 0;
};

void build_pattern_plan(const CsrGraph &g, PatternPlan &plan);

}  // namespace mgpm
