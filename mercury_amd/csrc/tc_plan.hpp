// Host-side triangle-count plan: which side of every DAG arc is hashed into
// an LDS table and which side is streamed against it, cut into tasks.
//
// The triangle count is  sum over undirected edges {a,b} of |N+(a) ∩ N+(b)|
// where N+ is the out-list of the degree-oriented DAG (graph.cc:196-235):
// each triangle x<y<z (in DAG order) is seen exactly once, on edge {x,y}.
// The reference (tc_challenge.cuh) hashes N+(v) of the endpoint with the
// larger DAG out-degree only for "large" edges (graph.cc:253-263,
// REVERSE_DAG) and binary-searches heads for "small" ones; here every edge
// uses that rule: the owner T has the larger d+ (ties: larger id), its
// N+(T) becomes the table, and N+ of the other endpoint (the partner) is
// streamed.  Probes = sum over edges of min(d+(a), d+(b)).
#pragma once

#include <cstdint>
#include <vector>

#include "host_graph.hpp"

namespace mgpm {

// One partner: the DAG out-list of the streamed endpoint.
struct TcPartner {
  uint32_t beg;  // offset into the DAG column array
  uint32_t len;  // d+ of the streamed endpoint (>= 1)
};

// One task: a table (the owner's DAG out-list) and a contiguous range of
// its partners.  Heavy owners are split into several tasks.
struct TcTask {
  uint32_t tbeg;  // owner's list offset in the DAG column array
  int32_t tlen;   // d+(owner)
  uint32_t pbeg;  // first partner index
  int32_t pcnt;   // number of partners
};

// Task classes by table size (d+ of the owner).
enum TcClass : int {
  kTcSmall = 0,   // d+ <= small_max: wave-private LDS table (<= 256 slots)
  kTcMedium = 1,  // d+ <= wave_max: wave-private LDS table (<= 2048 slots)
  kTcBlock = 2,   // d+ <= block_max: one LDS table shared by a workgroup
  kTcGlobal = 3,  // larger: binary search in the global sorted list
  kTcNumClasses = 4
};

// Cost model used to cut and group tasks: each streamed item costs 1, each
// partner kTcPartnerCost (its record load and share of the segment scan),
// each task's table build kTcTableCost plus its size.
constexpr int64_t kTcPartnerCost = 16;
constexpr int64_t kTcTableCost = 64;
// Queue entries of a class aim at this many (about 16 waves per CU of a
// 256-CU device), and are never cut below kTcMinGroupWork of cost.
constexpr int64_t kTcTargetGroups = 4096;
constexpr int64_t kTcMinGroupWork = 1024;

struct TcPlanParams {
  int32_t small_max = 128;
  int32_t wave_max = 1024;
  int32_t block_max = 16384;
  int64_t chunk_work = 65536;  // max cost of one task (heavy owners split)
  int64_t group_work = 32768;  // consecutive light tasks grouped up to this
};

struct TcPlan {
  int32_t n = 0;
  int64_t m = 0;       // undirected edges
  int64_t m_plus = 0;  // DAG arcs
  uint64_t s_plus = 0; // sum over arcs u->v of d+(v)
  uint64_t probes = 0; // sum over partners of len
  int32_t max_out = 0;
  std::vector<int32_t> dag_col;         // DAG column array (m_plus)
  std::vector<TcPartner> partners;      // grouped by owner, task-contiguous
  std::vector<TcTask> tasks[kTcNumClasses];  // each sorted by cost, desc
  // Tasks [group_off[c][g], group_off[c][g+1]) form queue entry g of class
  // c: heavy tasks alone, light ones packed up to group_work.
  std::vector<int32_t> group_off[kTcNumClasses];
  std::vector<int64_t> group_cost[kTcNumClasses];  // cost of each group
  int32_t class_max_tlen[kTcNumClasses] = {0, 0, 0, 0};
  uint64_t bytes_alg() const {
    return 8ull * (uint64_t(n) + 1) + 20ull * uint64_t(m_plus) + 4ull * s_plus;
  }
};

// Builds the plan.  Returns an MGPM status (MGPM_ERR_CAPACITY when the DAG
// has 2^32 or more arcs, which the 32-bit offsets cannot address).
int build_tc_plan(const CsrGraph &g, const TcPlanParams &p, TcPlan &plan);

}  // namespace mgpm
