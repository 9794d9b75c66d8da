// Device-side degree-oriented DAG and small/large division (dag_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace mgpm {

struct DagDevice {
  const int64_t *rowptr = nullptr;  // undirected CSR, n + 1
  const int32_t *col = nullptr;     // nnz
  int32_t n = 0;
  int64_t nnz = 0;                  // < 2^32 (packed division counts)
  int32_t deg_thd = 128;
  bool division = true;             // also build the small/large arrays
  int cus = 256;
  // Work arrays.
  int32_t *deg = nullptr;           // n
  int32_t *nd = nullptr;            // n: DAG out-degrees
  int32_t *src = nullptr;           // nnz
  int64_t *val = nullptr;           // nnz + 1
  int64_t *pos = nullptr;           // nnz + 1
  void *scan_tmp = nullptr;
  size_t scan_bytes = 0;
  // Outputs: DAG, small, large.
  int64_t *out_rowptr[3] = {nullptr, nullptr, nullptr};  // n + 1 each
  int32_t *out_col[3] = {nullptr, nullptr, nullptr};     // nnz / 2 each
};

// Temporary bytes of one exclusive scan over nnz + 1 int64 values.
int dag_scan_bytes(int64_t nnz, size_t *bytes);

// Enqueues both passes on `s`.  The column arrays must hold nnz / 2 entries
// each (the DAG has exactly nnz / 2 arcs; small + large at most that).
// Returns 0, or -1 if a launch failed.
int dag_device_build(const DagDevice &d, hipStream_t s);

}  // namespace mgpm
