// `sm <graph prefix> <P1..P8> [device] [--set-algebra]` — drop-in for the
// reference's subgraph-matching CLI (subgraph_matching/main_mpi.cpp:5-38 +
// sm_multigpu.cu:97-239) on one GPU, without MPI.  An unknown pattern
// prints "Not support." and exits with -1 as sm_multigpu.cu:194-198 does.
// The count is printed under the reference's label "total_num_triangles".
// P1, P4, P7 and P8 are counted in closed form from a clique/cycle census
// unless --set-algebra asks for the reference's per-edge set algebra (the
// same counts).
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
  if (argc > 3 && std::strcmp(argv[argc - 1], "--set-algebra") == 0) {
    set_algebra = true;
    --argc;
  }
  if (argc < 3) {
    std::printf("Usage: %s <graph> <pattern> [device] [--set-algebra]\n",
                argv[0]);
    std::printf("Example: %s /graph_inputs/mico/graph P1\n", argv[0]);
    return 1;
  }
  std::printf("Subgraph matching \n");
  int pattern = -1;
  if (std::strlen(argv[2]) == 2 && argv[2][0] == 'P' && argv[2][1] >= '1' &&
      argv[2][1] <= '8')
    pattern = argv[2][1] - '0';
  const int device = argc >= 4 ? std::atoi(argv[3]) : 0;
  mgpm_graph *g = nullptr;
  int rc = mgpm_graph_load(argv[1], &g);
  if (rc) return die(rc, "load");
  if (pattern < 0) {
    std::printf("Not support.\n");
    std::exit(-1);
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
