// Test-infrastructure driver compiled together with the reference's own
// host sources (graph.cc, VertexSet.cc) into oracle/_ref/libmercury_ref.so.
// It only calls the reference's public Graph API and copies the arrays out,
// so tests can compare the product's preprocessing with the reference's.
// Not part of the product; only tests/ load it.
#include <cstring>
#include <string>

#include "graph.h"

extern "C" {

// Graph(prefix) + orientation_with_division(deg_thd) (graph.cc:6-92,
// 196-321).  sizes[0..2] = DAG, small, large arc counts; arrays may be NULL.
int ref_orientation(const char *prefix, int deg_thd, int64_t sizes[3],
                    int64_t *dag_rowptr, int32_t *dag_colidx,
                    int64_t *small_rowptr, int32_t *small_colidx,
                    int64_t *large_rowptr, int32_t *large_colidx) {
  Graph g(prefix);
  g.orientation_with_division(deg_thd);
  const int64_t n = g.num_vertices();
  sizes[0] = g.num_edges();
  sizes[1] = g.num_small_edges();
  sizes[2] = g.num_large_edges();
  const size_t rp = size_t(n + 1) * sizeof(int64_t);
  if (dag_rowptr) std::memcpy(dag_rowptr, g.out_rowptr(), rp);
  if (dag_colidx) std::memcpy(dag_colidx, g.out_colidx(), sizes[0] * 4);
  if (small_rowptr) std::memcpy(small_rowptr, g.small_out_rowptr(), rp);
  if (small_colidx)
    std::memcpy(small_colidx, g.small_out_colidx(), sizes[1] * 4);
  if (large_rowptr) std::memcpy(large_rowptr, g.large_out_rowptr(), rp);
  if (large_colidx)
    std::memcpy(large_colidx, g.large_out_colidx(), sizes[2] * 4);
  return (int)g.get_max_degree();
}

// Graph(prefix) + init_edgelist() (graph.cc:336-365): the COO task order
// of sm_multigpu.cu:140.  Returns the number of directed edges.
int64_t ref_edgelist(const char *prefix, int32_t *src, int32_t *dst) {
  Graph g(prefix);
  const int64_t nnz = g.init_edgelist();
  if (src) std::memcpy(src, g.get_src_ptr(), nnz * 4);
  if (dst) std::memcpy(dst, g.get_dst_ptr(), nnz * 4);
  return nnz;
}

}  // extern "C"
