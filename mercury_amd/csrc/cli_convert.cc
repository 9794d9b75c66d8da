// `convert <edge file> <graph prefix> [base]` -- writes the reference's CSR
// files (<prefix>.meta.txt / .vertex.bin / .edge.bin) from a MatrixMarket
// (.mtx) file or an edge list (GraphChallenge .tsv, one "src dst [...]" per
// line), symmetrized with self-loops and duplicates dropped.  Takes the
// place of datasets/converter.sh ("converter mtx <in> <out>/graph 1 0 0 0")
// and datasets/tsv2mtx.py.  `base` is the id of the first vertex (default
// 1, as in MatrixMarket and GraphChallenge files; 0 for 0-based lists).
#include <cinttypes>
#include <cstdio>
#include <cstdlib>

#include "mgpm.h"

int main(int argc, char **argv) {
  if (argc < 3) {
    std::printf("Usage: %s <input.mtx|input.tsv> <output prefix> [base]\n",
                argv[0]);
    std::printf("Example: %s amazon0302.tsv /graph_inputs/amazon/graph\n",
                argv[0]);
    return 1;
  }
  const int base = argc >= 4 ? std::atoi(argv[3]) : -1;
  mgpm_graph *g = nullptr;
  int rc = mgpm_graph_read_edges(argv[1], 0, base, &g);
  if (rc) {
    std::fprintf(stderr, "read %s: %s (%d): %s\n", argv[1], mgpm_strerror(rc),
                 rc, mgpm_last_error());
    return 1;
  }
  rc = mgpm_graph_write(g, argv[2]);
  if (rc) {
    std::fprintf(stderr, "write %s: %s (%d): %s\n", argv[2],
                 mgpm_strerror(rc), rc, mgpm_last_error());
    mgpm_graph_free(g);
    return 1;
  }
  mgpm_graph_info info;
  mgpm_graph_info_get(g, &info);
  std::printf("vertices %d, undirected edges %" PRId64 ", max degree %d\n",
              info.n, info.nnz / 2, info.max_degree);
  mgpm_graph_free(g);
  return 0;
}
