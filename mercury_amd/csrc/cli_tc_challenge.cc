// `tc_challenge <graph prefix>` — drop-in for the reference CLI (main.cc:5-20
// + tc_challenge.cu:11-118).  Prints the reference's lines ("Triangle
// Counting", "Runtime  = <s> sec", "total_num_triangles = <n>") and one
// JSON line with the measurements.  Runtime is the device time of the
// counting kernels (load, DAG and upload excluded, as tc_challenge.cu:98-110).
// The DAG and the triangle plan are built on the device
// (MGPM_UPLOAD_DEVICE_PLAN); `--host-plan` builds them on the host instead.
#include <chrono>
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
  int32_t flags = MGPM_UPLOAD_DEVICE_PLAN;
  if (argc >= 2 && std::strcmp(argv[argc - 1], "--host-plan") == 0) {
    flags = 0;
    --argc;
  }
  if (argc < 2) {
    std::printf("Usage: %s <graph> [device] [--host-plan]\n", argv[0]);
    std::printf("Example: %s /graph_inputs/mico/graph\n", argv[0]);
    return 1;
  }
  const int device = argc >= 3 ? std::atoi(argv[2]) : 0;
  std::printf("Triangle Counting\n");
  const auto t0 = std::chrono::steady_clock::now();
  mgpm_graph *g = nullptr;
  int rc = mgpm_graph_load(argv[1], &g);
  if (rc) return die(rc, "load");
  mgpm_graph_info gi;
  mgpm_graph_info_get(g, &gi);
  std::printf("|V| %d |E| %" PRId64 "\n", gi.n, gi.nnz);
  const auto t1 = std::chrono::steady_clock::now();
  mgpm_dgraph *dg = nullptr;
  rc = mgpm_dgraph_upload_ex(g, device, nullptr, flags, &dg);
  if (rc) return die(rc, "upload");
  const auto t2 = std::chrono::steady_clock::now();
  uint64_t total = 0;
  mgpm_stats st;
  rc = mgpm_tc_count(dg, nullptr, nullptr, &total, &st);
  if (rc) return die(rc, "count");
  const auto t3 = std::chrono::steady_clock::now();
  auto sec = [](auto a, auto b) {
    return std::chrono::duration<double>(b - a).count();
  };
  std::printf("Runtime  = %f sec\n", st.kernel_ms / 1e3);
  std::printf("total_num_triangles = %" PRIu64 "\n", total);
  std::printf(
      "{\"count\": %" PRIu64 ", \"kernel_ms\": %.4f, \"load_s\": %.3f, "
      "\"prep_upload_s\": %.3f, \"e2e_s\": %.3f, \"m\": %" PRIu64
      ", \"m_plus\": %" PRIu64 ", \"s_plus\": %" PRIu64
      ", \"edges_per_s\": %.4e, \"bytes_alg\": %" PRIu64
      ", \"alg_GBps\": %.1f, \"roofline_frac\": %.4f}\n",
      total, st.kernel_ms, sec(t0, t1), sec(t1, t2), sec(t0, t3), st.m,
      st.m_plus, st.s_plus, st.m / (st.kernel_ms / 1e3), st.bytes_alg,
      st.bytes_alg / (st.kernel_ms / 1e3) / 1e9,
      st.bytes_alg / (st.kernel_ms / 1e3) / 8e12);
  mgpm_dgraph_free(dg);
  mgpm_graph_free(g);
  return 0;
}
