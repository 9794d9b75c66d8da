// Allocation-free counting: mgpm_pattern_count_noalloc.  No counterpart in
// the reference.  The same counting kernels as mgpm_tc_count and the
// per-edge path of mgpm_pattern_count, with every scratch buffer they use
// -- queue heads, result counters and the per-wave frontier slab of the
// P1..P8 kernels -- taken from one file-scope device array of fixed size
// instead of memory the library allocates.  For callers that cannot call
// hipMalloc: code under HIP graph capture, tight serving loops, harnesses
// that forbid allocation.
//
// The array is per GPU (each device holds its own copy at its own
// address, from the first use of this code object on it until the process
// ends).  Its device address is looked up once per device with
// hipGetSymbolAddress, with that device current, and only that address is
// ever handed to kernels: the array's name in host code is a host-side
// stand-in, not a device pointer.  Every call on a GPU shares that GPU's
// array and leaves its contents behind; the queue heads and counters are
// cleared by a fill kernel at the start of each call, and the slab is
// written before it is read.  Calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "capi_common.hpp"

namespace mgpm {

namespace {

constexpr size_t kNoAllocBytes = MGPM_NOALLOC_SCRATCH_BYTES;
// Head of the array: triangle queues at 0 and result at 64, pattern
// queues at 128 and counters at 192; the pattern slab from 256 on.
constexpr size_t kHeadBytes = 256;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_base[kMaxDevices] = {};

// Device address of `scratch` on the current device `device`, looked up on
// the first call for that device.
int scratch_base(int device, void **out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(MGPM_ERR_NODEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_base[device]) {
    void *p = nullptr;
    const hipError_t e =
        hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(MGPM_ERR_HIP, "hipGetSymbolAddress(scratch): null");
    g_base[device] = p;
  }
  *out = g_base[device];
  return MGPM_OK;
}

}  // namespace

}  // namespace mgpm

using namespace mgpm;

extern "C" int mgpm_pattern_count_noalloc(mgpm_dgraph *dg, int pattern,
                                          void *stream, uint64_t *count) {
  if (!dg || !count) return fail(MGPM_ERR_INVALID, "null argument");
  if (pattern < MGPM_TC || pattern > MGPM_P8)
    return fail(MGPM_ERR_PATTERN, "Not support.");
  try {
    int rc = set_device(dg->device);
    if (rc) return rc;
    void *base = nullptr;
    if ((rc = scratch_base(dg->device, &base))) return rc;
    char *b = static_cast<char *>(base);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (pattern == MGPM_TC) {
      TcDevicePlan p = dg->tc.dev;
      p.queues = reinterpret_cast<unsigned *>(b);
      p.result = reinterpret_cast<unsigned long long *>(b + 64);
      TcLaunch l;
      l.cus = dg->cus;
      l.items = dg->tc.items;
      l.grid_cap = dg->tc.grid_cap;
      l.cursor_min = dg->tc.cursor_min;
      l.reset = true;  // a fill kernel clears the queues and the result
      if (tc_device_launch(p, l, s))
        return fail(MGPM_ERR_HIP, "triangle kernel launch failed");
      MGPM_HIP(hipMemcpyAsync(dg->pinned, p.result, 8, hipMemcpyDeviceToHost,
                              s));
      MGPM_HIP(hipStreamSynchronize(s));
      *count = dg->pinned[0];
      return MGPM_OK;
    }
    const PatternPlanResident &pr = dg->pat.plan[pattern];
    // Slab per workgroup: 8 lists of `cap` ints per wave (cap >= the
    // largest degree), as the single-query path lays it out.
    const int64_t cap =
        pr.ready ? dg->pat.cap
                 : std::max<int64_t>(dg->host->max_degree(), 64);
    const uint64_t per_block =
        uint64_t(pattern_waves_per_block()) * 8 * uint64_t(cap) * 4;
    const uint64_t fit = (kNoAllocBytes - kHeadBytes) / per_block;
    if (fit < 1)
      return fail(MGPM_ERR_CAPACITY,
                  "allocation-free count: a workgroup's frontier slab needs " +
                      std::to_string(per_block) + " bytes (largest degree " +
                      std::to_string(cap) + "), the fixed scratch holds " +
                      std::to_string(kNoAllocBytes - kHeadBytes));
    if (!pr.ready)
      return fail(MGPM_ERR_INVALID,
                  "allocation-free count: the pattern's plan is not resident "
                  "(count it once with mgpm_pattern_count first)");
    int grid = static_cast<int>(
        std::min<uint64_t>(uint64_t(dg->cus) * 4, fit));
    if (dg->pat.grid_cap > 0) grid = std::min(grid, dg->pat.grid_cap);
    PatternDevicePlan p = pattern_device_plan(dg, pattern);
    p.queues = reinterpret_cast<unsigned *>(b + 128);
    p.counters = reinterpret_cast<unsigned long long *>(b + 192);
    p.slab = reinterpret_cast<int32_t *>(b + kHeadBytes);
    p.cap = cap;
    // reset = true: a fill kernel clears queues and counters first.
    if (pattern_device_launch(p, pattern, grid, s, true))
      return fail(MGPM_ERR_HIP, "pattern kernel launch failed");
    MGPM_HIP(hipMemcpyAsync(dg->pinned, p.counters, 16,
                            hipMemcpyDeviceToHost, s));
    MGPM_HIP(hipStreamSynchronize(s));
    const uint64_t c1 = dg->pinned[0], c2 = dg->pinned[1];
    // The counter combination of sm_multigpu.cu:209-232 on one rank.
    *count = (pattern == MGPM_P4 || pattern == MGPM_P6) ? c1 : c1 - c2;
    return MGPM_OK;
  } catch (const std::bad_alloc &) {
    return fail(MGPM_ERR_NOMEM, "out of host memory");
  } catch (const std::exception &ex) {
    return fail(MGPM_ERR_INTERNAL, ex.what());
  }
}
