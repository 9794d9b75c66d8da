// `sm <graph prefix> <P1..P8> [devices per node] [--device N]
// [--set-algebra]` -- drop-in for the reference's subgraph-matching CLI
// (subgraph_matching/main_mpi.cpp:5-38 + sm_multigpu.cu:97-239) on one
// GPU, without MPI.  The third argument means what it means there:
// devices per node, 0 for all, and the one process (rank 0) takes device
// 0 % count (sm_multigpu.cu:100-104); `--device N` picks another device.
// An unknown pattern prints "Not support." and exits with -1 as
// sm_multigpu.cu:194-198 does.  The count is printed under the reference's
// label "total_num_triangles".  Patterns with a closed form are counted
// from a subgraph census unless --set-algebra asks for the reference's
// per-edge set algebra (the same counts).
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "mgpm.h"

static int die(int rc, const char *what) {
  std::fprintf(stderr, "%s: %s (%d): %s\n", what, mgpm_strerror(rc), rc,
               mgpm_last_error());
  return 1;
}

int main(int argc, char **argv) {
  bool set_algebra = false;
  int device = -1;
  // Options may follow the positional arguments in any order.
  int npos = 0;
  char *pos[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int i = 0; i < argc; ++i) {
    if (std::strcmp(argv[i], "--set-algebra") == 0) {
      set_algebra = true;
    } else if (std::strcmp(argv[i], "--device") == 0 && i + 1 < argc) {
      device = std::atoi(argv[++i]);
    } else if (npos < 4) {
      pos[npos++] = argv[i];
    }
  }
  if (npos < 3) {
    std::printf("Usage: %s <graph> <pattern> [devices per node] "
                "[--device N] [--set-algebra]\n", argv[0]);
    std::printf("Example: %s /graph_inputs/mico/graph P1 4\n", argv[0]);
    return 1;
  }
  std::printf("Subgraph matching \n");
  int pattern = -1;
  if (std::strlen(pos[2]) == 2 && pos[2][0] == 'P' && pos[2][1] >= '1' &&
      pos[2][1] <= '8')
    pattern = pos[2][1] - '0';
  mgpm_graph *g = nullptr;
  int rc = mgpm_graph_load(pos[1], &g);
  if (rc) return die(rc, "load");
  if (pattern < 0) {
    std::printf("Not support.\n");
    std::exit(-1);
  }
  if (device < 0) {
    int per_node = npos >= 4 ? std::atoi(pos[3]) : 0;
    if (per_node <= 0) {
      rc = mgpm_device_count(&per_node);
      if (rc) return die(rc, "device count");
      if (per_node <= 0) return die(MGPM_ERR_NODEVICE, "device count");
    }
    device = 0 % per_node;  // rank 0 of one process
  }
  mgpm_dgraph *dg = nullptr;
  rc = mgpm_dgraph_upload(g, device, nullptr, &dg);
  if (rc) return die(rc, "upload");
  uint64_t total = 0;
  mgpm_stats st;
  mgpm_tune tune;
  std::memset(&tune, 0, sizeof(tune));
  tune.pattern_census = -1;
  rc = mgpm_pattern_count(dg, pattern, set_algebra ? &tune : nullptr, nullptr,
                          &total, &st);
  if (rc == MGPM_ERR_PATTERN) {
    std::printf("Not support.\n");
    std::exit(-1);
  }
  if (rc) return die(rc, "count");
  std::printf("Rank:0 runtime  = %f sec\n", st.kernel_ms / 1e3);
  std::printf("Total runtime  = %f sec\n", st.kernel_ms / 1e3);
  std::printf("total_num_triangles = %" PRIu64 "\n", total);
  mgpm_dgraph_free(dg);
  mgpm_graph_free(g);
  return 0;
}
