// C ABI: the degree-oriented DAG and its small/large division built on the
// device from a device graph's CSR (graph.cc:196-321 on the GPU).
#include <algorithm>
#include <cstring>

#include "capi_common.hpp"
#include "dag_device.hpp"

using namespace mgpm;

extern "C" int mgpm_dgraph_orientation(
    mgpm_dgraph *dg, int32_t deg_thd, void *stream, int64_t sizes[3],
    int64_t *dag_rowptr, int32_t *dag_colidx, int64_t *small_rowptr,
    int32_t *small_colidx, int64_t *large_rowptr, int32_t *large_colidx,
    mgpm_stats *stats) {
  if (!dg || !sizes) return fail(MGPM_ERR_INVALID, "null argument");
  if (deg_thd < 0) return fail(MGPM_ERR_INVALID, "negative degree threshold");
  try {
    const auto t0 = std::chrono::steady_clock::now();
    int rc = set_device(dg->device);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const CsrGraph &G = *dg->host;
    DagResident &r = dg->dag;
    if (G.nnz >= (int64_t(1) << 32))
      return fail(MGPM_ERR_CAPACITY, "device orientation needs < 2^32 arcs");
    const size_t n = static_cast<size_t>(G.n);
    const size_t nnz = static_cast<size_t>(G.nnz);
    const size_t m = std::max<size_t>(nnz / 2, 1);
    const size_t n1 = n + 1;
    auto need = [](DevBuf &b, size_t bytes) -> hipError_t {
      return b.bytes >= bytes ? hipSuccess : b.alloc(bytes);
    };
    for (int k = 0; k < 3; ++k) {
      MGPM_HIP(need(r.rowptr[k], (n + 1) * sizeof(int64_t)));
      MGPM_HIP(need(r.col[k], m * sizeof(int32_t)));
    }
    MGPM_HIP(need(r.deg, std::max<size_t>(n, 1) * sizeof(int32_t)));
    MGPM_HIP(need(r.nd, std::max<size_t>(n, 1) * sizeof(int32_t)));
    MGPM_HIP(need(r.src, std::max<size_t>(nnz, 1) * sizeof(int32_t)));
    MGPM_HIP(need(r.val, (nnz + 1) * sizeof(int64_t)));
    MGPM_HIP(need(r.pos, (nnz + 1) * sizeof(int64_t)));
    size_t tb = 0;
    if (dag_scan_bytes(G.nnz, &tb)) return fail(MGPM_ERR_HIP, "scan size query");
    MGPM_HIP(need(r.scan, std::max<size_t>(tb, 1)));
    DagDevice d;
    d.rowptr = dg->csr.rowptr.as<int64_t>();
    d.col = dg->csr.colidx.as<int32_t>();
    d.n = G.n;
    d.nnz = G.nnz;
    d.deg_thd = deg_thd;
    d.division = true;
    d.cus = dg->cus;
    d.deg = r.deg.as<int32_t>();
    d.nd = r.nd.as<int32_t>();
    d.src = r.src.as<int32_t>();
    d.val = r.val.as<int64_t>();
    d.pos = r.pos.as<int64_t>();
    d.scan_tmp = r.scan.p;
    d.scan_bytes = tb;
    for (int k = 0; k < 3; ++k) {
      d.out_rowptr[k] = r.rowptr[k].as<int64_t>();
      d.out_col[k] = r.col[k].as<int32_t>();
    }
    MGPM_HIP(hipEventRecord(dg->ev0, s));
    if (dag_device_build(d, s))
      return fail(MGPM_ERR_HIP, "orientation kernel launch failed");
    MGPM_HIP(hipEventRecord(dg->ev1, s));
    for (int k = 0; k < 3; ++k)
      MGPM_HIP(hipMemcpyAsync(&r.sizes[k], d.out_rowptr[k] + G.n,
                              sizeof(int64_t), hipMemcpyDeviceToHost, s));
    MGPM_HIP(hipStreamSynchronize(s));
    float kms = 0;
    MGPM_HIP(hipEventElapsedTime(&kms, dg->ev0, dg->ev1));
    r.deg_thd = deg_thd;
    r.ready = true;
    for (int k = 0; k < 3; ++k) sizes[k] = r.sizes[k];
    int64_t *rp_out[3] = {dag_rowptr, small_rowptr, large_rowptr};
    int32_t *col_out[3] = {dag_colidx, small_colidx, large_colidx};
    for (int k = 0; k < 3; ++k) {
      if (rp_out[k])
        MGPM_HIP(hipMemcpy(rp_out[k], d.out_rowptr[k], n1 * sizeof(int64_t),
                           hipMemcpyDeviceToHost));
      if (col_out[k] && r.sizes[k] > 0)
        MGPM_HIP(hipMemcpy(col_out[k], d.out_col[k],
                           static_cast<size_t>(r.sizes[k]) * sizeof(int32_t),
                           hipMemcpyDeviceToHost));
    }
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->kernel_ms = kms;
      stats->n = static_cast<uint64_t>(G.n);
      stats->m = static_cast<uint64_t>(G.nnz / 2);
      stats->m_plus = static_cast<uint64_t>(r.sizes[0]);
      stats->path = 200;
      stats->e2e_ms = ms_since(t0);
    }
    return MGPM_OK;
  } catch (const std::bad_alloc &) {
    return fail(MGPM_ERR_NOMEM, "out of host memory");
  } catch (const std::exception &ex) {
    return fail(MGPM_ERR_INTERNAL, ex.what());
  }
}
