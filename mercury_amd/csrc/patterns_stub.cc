// Placeholder until the P1..P8 kernels land: triangle counting only.
#include "capi_common.hpp"

extern "C" int mgpm_pattern_count(mgpm_dgraph *dg, int pattern,
                                  const mgpm_tune *tune, void *stream,
                                  uint64_t *count, mgpm_stats *stats) {
  if (pattern == MGPM_TC) return mgpm_tc_count(dg, tune, stream, count, stats);
  return mgpm::fail(MGPM_ERR_PATTERN, "Not support.");
}
