// Triangle-count plan built on the device from the device graph's CSR:
// DAG (dag_kernels.hip, DAG pass only), then partners, tasks, LPT order,
// classes and queue groups (tc_dplan_kernels.hip).  Replaces the host
// builder (tc_plan.cc) and its host DAG for callers who want the graph
// ready for counting without host preprocessing; the counts are the same,
// the task cut and grouping follow the same rules in a parallel form.
#include <algorithm>
#include <cstring>

#include "capi_common.hpp"
#include "dag_device.hpp"
#include "tc_dplan.hpp"

namespace mgpm {

int tc_make_resident_device(mgpm_dgraph *dg, const TcPlanParams &params,
                            hipStream_t s) {
  const CsrGraph &G = *dg->host;
  const int64_t nnz = G.nnz;
  const int32_t n = G.n;
  const int64_t m = nnz / 2;  // the DAG keeps each undirected edge once
  if (m >= (int64_t(1) << 32) || nnz >= (int64_t(1) << 32))
    return fail(MGPM_ERR_CAPACITY, "triangle plan: DAG has 2^32 or more arcs");
  TcResident &r = dg->tc;
  const size_t n1 = static_cast<size_t>(n) + 1;
  // The new plan is built into these and swapped into `r` only when it is
  // complete, so a failure leaves the resident plan untouched.
  DevBuf new_col, new_partners, new_tasks, new_groups;
  const size_t a1 = static_cast<size_t>(nnz) + 1;

  // DAG into the resident column array.
  DevBuf deg, nd, src, val, pos, dag_rowptr, scan;
  MGPM_HIP(new_col.alloc(std::max<int64_t>(m, 1) * sizeof(int32_t)));
  MGPM_HIP(deg.alloc(std::max<int32_t>(n, 1) * sizeof(int32_t)));
  MGPM_HIP(nd.alloc(std::max<int32_t>(n, 1) * sizeof(int32_t)));
  MGPM_HIP(src.alloc(std::max<int64_t>(nnz, 1) * sizeof(int32_t)));
  MGPM_HIP(val.alloc(a1 * sizeof(int64_t)));
  MGPM_HIP(pos.alloc(a1 * sizeof(int64_t)));
  MGPM_HIP(dag_rowptr.alloc(n1 * sizeof(int64_t)));
  size_t scan_bytes = 0, b = 0;
  if (dag_scan_bytes(nnz, &b)) return fail(MGPM_ERR_HIP, "scan size query");
  scan_bytes = std::max(scan_bytes, b);
  if (dplan_scan_bytes(static_cast<int64_t>(a1), &b))
    return fail(MGPM_ERR_HIP, "scan size query");
  scan_bytes = std::max(scan_bytes, b);
  MGPM_HIP(scan.alloc(std::max<size_t>(scan_bytes, 1)));
  DagDevice d;
  d.rowptr = dg->csr.rowptr.as<int64_t>();
  d.col = dg->csr.colidx.as<int32_t>();
  d.n = n;
  d.nnz = nnz;
  d.division = false;
  d.cus = dg->cus;
  d.deg = deg.as<int32_t>();
  d.nd = nd.as<int32_t>();
  d.src = src.as<int32_t>();
  d.val = val.as<int64_t>();
  d.pos = pos.as<int64_t>();
  d.scan_tmp = scan.p;
  d.scan_bytes = scan_bytes;
  d.out_rowptr[0] = dag_rowptr.as<int64_t>();
  d.out_col[0] = new_col.as<int32_t>();
  if (dag_device_build(d, s)) return fail(MGPM_ERR_HIP, "DAG kernel launch failed");

  // Partners: the count is needed to size the rest.
  TcDplanArgs a;
  a.cus = dg->cus;
  a.rowptr = d.rowptr;
  a.col = d.col;
  a.src = d.src;
  a.nnz = nnz;
  a.dag_rowptr = d.out_rowptr[0];
  a.nd = d.nd;
  a.chunk_work = std::max<int64_t>(params.chunk_work, 1);
  a.group_work = std::max<int64_t>(params.group_work, 1);
  a.cls.small_max = params.small_max;
  a.cls.wave_max = params.wave_max;
  a.cls.block_max = params.block_max;
  a.val = d.val;
  a.pos = d.pos;
  a.tix = d.pos;  // pos is dead once the task flags are computed
  a.scan_tmp = scan.p;
  a.scan_bytes = scan_bytes;
  const int64_t cap = std::max<int64_t>(m, 1);  // partners <= m
  DevBuf owner, cost, cpre;
  MGPM_HIP(new_partners.alloc(cap * sizeof(TcPartner)));
  MGPM_HIP(owner.alloc(cap * sizeof(int32_t)));
  MGPM_HIP(cost.alloc((cap + 1) * sizeof(int64_t)));
  MGPM_HIP(cpre.alloc((cap + 1) * sizeof(int64_t)));
  a.partners = new_partners.as<TcPartner>();
  a.owner = owner.as<int32_t>();
  a.cost = cost.as<int64_t>();
  a.cpre = cpre.as<int64_t>();
  if (dplan_partners(a, s)) return fail(MGPM_ERR_HIP, "plan kernel launch failed");
  int64_t np = 0;
  MGPM_HIP(hipMemcpyAsync(&np, d.pos + nnz, sizeof(np), hipMemcpyDeviceToHost, s));
  MGPM_HIP(hipStreamSynchronize(s));
  if (np < 0 || np > m) return fail(MGPM_ERR_INTERNAL, "triangle plan: bad partner count");

  int64_t nt = 0;
  if (np > 0) {
    if (dplan_task_flags(a, np, s))
      return fail(MGPM_ERR_HIP, "plan kernel launch failed");
    MGPM_HIP(hipMemcpyAsync(&nt, a.tix + np, sizeof(nt), hipMemcpyDeviceToHost, s));
    MGPM_HIP(hipStreamSynchronize(s));
    if (nt < 1 || nt > np) return fail(MGPM_ERR_INTERNAL, "triangle plan: bad task count");
  }

  // Tasks, LPT order, classes, groups.
  const size_t t1 = static_cast<size_t>(nt) + 1;
  DevBuf start, tasks_all, keys, keys_sorted, vals, order, sort_tmp, ca, cb,
      cas, cbs, work, wpre, gflag, gix, summary;
  MGPM_HIP(summary.alloc(sizeof(TcDplanSummary)));
  MGPM_HIP(hipMemsetAsync(summary.p, 0, sizeof(TcDplanSummary), s));
  a.summary = summary.as<TcDplanSummary>();
  MGPM_HIP(new_tasks.alloc(std::max<size_t>(nt, 1) * sizeof(TcTask)));
  MGPM_HIP(new_groups.alloc((static_cast<size_t>(nt) + kTcNumClasses) * sizeof(int32_t)));
  a.tasks_out = new_tasks.as<TcTask>();
  a.groups_out = new_groups.as<int32_t>();
  if (nt > 0) {
    MGPM_HIP(start.alloc(t1 * sizeof(int32_t)));
    MGPM_HIP(tasks_all.alloc(nt * sizeof(TcTask)));
    MGPM_HIP(keys.alloc(nt * sizeof(uint64_t)));
    MGPM_HIP(keys_sorted.alloc(nt * sizeof(uint64_t)));
    MGPM_HIP(vals.alloc(nt * sizeof(int32_t)));
    MGPM_HIP(order.alloc(nt * sizeof(int32_t)));
    size_t sb = 0;
    if (dplan_sort_bytes(nt, &sb)) return fail(MGPM_ERR_HIP, "sort size query");
    MGPM_HIP(sort_tmp.alloc(std::max<size_t>(sb, 1)));
    DevBuf *per_task[] = {&ca, &cb, &cas, &cbs, &work, &wpre, &gflag, &gix};
    for (DevBuf *x : per_task) MGPM_HIP(x->alloc(t1 * sizeof(int64_t)));
    a.start = start.as<int32_t>();
    a.tasks_all = tasks_all.as<TcTask>();
    a.keys = keys.as<uint64_t>();
    a.keys_sorted = keys_sorted.as<uint64_t>();
    a.vals = vals.as<int32_t>();
    a.order = order.as<int32_t>();
    a.sort_tmp = sort_tmp.p;
    a.sort_bytes = sb;
    a.ca = ca.as<int64_t>();
    a.cb = cb.as<int64_t>();
    a.cas = cas.as<int64_t>();
    a.cbs = cbs.as<int64_t>();
    a.work = work.as<int64_t>();
    a.wpre = wpre.as<int64_t>();
    a.gflag = gflag.as<int64_t>();
    a.gix = gix.as<int64_t>();
    if (dplan_tasks(a, np, nt, s)) return fail(MGPM_ERR_HIP, "plan kernel launch failed");
  }
  if (dplan_dag_sums(new_col.as<int32_t>(), m, d.nd, n, dg->cus, a.summary, s))
    return fail(MGPM_ERR_HIP, "plan kernel launch failed");
  TcDplanSummary sum;
  int64_t cost_total = 0;
  MGPM_HIP(hipMemcpyAsync(&sum, summary.p, sizeof(sum), hipMemcpyDeviceToHost, s));
  MGPM_HIP(hipMemcpyAsync(&cost_total, a.cpre + np, sizeof(cost_total),
                          hipMemcpyDeviceToHost, s));
  MGPM_HIP(hipStreamSynchronize(s));
  if (np == 0) cost_total = 0;

  TcDevicePlan p;
  const TcTask *tb = new_tasks.as<TcTask>();
  const int32_t *gb = new_groups.as<int32_t>();
  int64_t total = 0;
  for (int c = 0; c < kTcNumClasses; ++c) {
    const int64_t cnt = nt > 0 ? sum.class_cnt[c] : 0;
    const int64_t off = nt > 0 ? sum.class_off[c] : 0;
    const int64_t ng = nt > 0 ? sum.ngroups[c] : 0;
    if (cnt < 0 || off < 0 || off + cnt > nt || ng < 0 || ng > cnt)
      return fail(MGPM_ERR_INTERNAL, "triangle plan: bad class summary");
    total += cnt;
    p.tasks[c] = tb + off;
    p.ntasks[c] = static_cast<int>(cnt);
    p.groups[c] = gb + off + c;
    p.ngroups[c] = static_cast<int>(ng);
    p.class_max_tlen[c] = nt > 0 ? sum.class_max_tlen[c] : 0;
  }
  if (total != nt) return fail(MGPM_ERR_INTERNAL, "triangle plan: classes lost tasks");
  p.col = new_col.as<int32_t>();
  p.ncol = static_cast<uint32_t>(m);
  p.partners = new_partners.as<TcPartner>();
  p.queues = dg->scratch.as<unsigned>();
  p.result = reinterpret_cast<unsigned long long *>(dg->scratch.as<char>() + 64);

  // Complete: install it (the previous plan's buffers die with the locals).
  r.col.swap(new_col);
  r.partners.swap(new_partners);
  r.tasks[0].swap(new_tasks);
  r.groups[0].swap(new_groups);
  for (int c = 1; c < kTcNumClasses; ++c) {
    r.tasks[c].reset();
    r.groups[c].reset();
  }
  r.dev = p;
  r.params = params;
  r.generation = next_generation();
  r.drop_layout();
  r.npart = static_cast<uint64_t>(np);
  r.s_plus = sum.s_plus;
  r.probes = static_cast<uint64_t>(cost_total - kTcPartnerCost * np);
  r.m_plus = static_cast<uint64_t>(m);
  r.max_out = sum.max_out;
  if (!r.tuned) r.items = default_tc_items(r.probes, static_cast<uint64_t>(np));
  r.on_device = true;
  return MGPM_OK;
}

}  // namespace mgpm
