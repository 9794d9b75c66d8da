// Triangle-count plan built on the device: launchers of tc_dplan_kernels.hip
// and the argument block they share (tc_dplan.cc runs them).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

#include "tc_plan.hpp"

namespace mgpm {

struct TcDplanClasses {
  int32_t small_max = 0, wave_max = 0, block_max = 0;
};

// Scalars the kernels produce; read by the host once, at the end.
struct TcDplanSummary {
  int64_t class_cnt[kTcNumClasses];
  int64_t class_off[kTcNumClasses];
  int64_t class_limit[kTcNumClasses];
  int64_t ngroups[kTcNumClasses];
  unsigned long long class_work[kTcNumClasses];
  int32_t class_max_tlen[kTcNumClasses];
  unsigned long long s_plus;
  int32_t max_out;
  int32_t pad;
};

struct TcDplanArgs {
  int cus = 256;
  // Undirected CSR and its arc sources; DAG row offsets and out-degrees.
  const int64_t *rowptr = nullptr;
  const int32_t *col = nullptr;
  const int32_t *src = nullptr;
  int64_t nnz = 0;
  const int64_t *dag_rowptr = nullptr;
  const int32_t *nd = nullptr;
  int64_t chunk_work = 65536, group_work = 32768;
  TcDplanClasses cls;
  // Per arc (nnz + 1): owned flags / later task-start flags, and their scan
  // (partner index / later task index).
  int64_t *val = nullptr, *pos = nullptr, *tix = nullptr;
  void *scan_tmp = nullptr;
  size_t scan_bytes = 0;
  // Per partner.
  TcPartner *partners = nullptr;
  int32_t *owner = nullptr;
  int64_t *cost = nullptr, *cpre = nullptr;  // np + 1
  // Per task.
  int32_t *start = nullptr;                  // nt + 1
  TcTask *tasks_all = nullptr;
  uint64_t *keys = nullptr, *keys_sorted = nullptr;
  int32_t *vals = nullptr, *order = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_bytes = 0;
  int64_t *ca = nullptr, *cb = nullptr, *cas = nullptr, *cbs = nullptr;
  int64_t *work = nullptr, *wpre = nullptr, *gflag = nullptr, *gix = nullptr;
  // Outputs: tasks by class (class c at class_off[c]), group offsets (class
  // c at class_off[c] + c, nt + 4 entries in all), summary.
  TcTask *tasks_out = nullptr;
  int32_t *groups_out = nullptr;
  TcDplanSummary *summary = nullptr;
};

// Temporary bytes of an exclusive int64 scan / a descending pair sort.
int dplan_scan_bytes(int64_t count, size_t *bytes);
int dplan_sort_bytes(int64_t count, size_t *bytes);

// Owned flags, their scan (pos; the partner count lands in pos[nnz]) and
// the partner records.  Returns 0, or -1 if a launch failed.
int dplan_partners(const TcDplanArgs &a, hipStream_t s);
// Cost prefix and task-start flags over `np` partners; the task count
// lands in tix[np].
int dplan_task_flags(const TcDplanArgs &a, int64_t np, hipStream_t s);
// Task records, sort, class partition and groups for `nt` tasks.
int dplan_tasks(const TcDplanArgs &a, int64_t np, int64_t nt, hipStream_t s);
// S+ and the largest out-degree into the summary.
int dplan_dag_sums(const int32_t *dcol, int64_t mplus, const int32_t *nd,
                   int32_t n, int cus, TcDplanSummary *sum, hipStream_t s);

}  // namespace mgpm
