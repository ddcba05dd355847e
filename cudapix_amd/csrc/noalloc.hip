// noalloc.hip -- pi(x) for a batch without any allocation: the batch's plan
// and output arrays live in a static device array of the library instead of
// the context's batch arena, for callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
#include "engine.hpp"
#include <mutex>

// The region: each GPU the library runs on holds its own copy of this array
// from module load until the process ends.
#define CPX_NOALLOC_BYTES (128ull << 20)
static __device__ __align__(256) unsigned char scratch[CPX_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu[kMaxDevices];       // one call at a time per GPU (they share its array)
void *g_ptr[kMaxDevices];           // the array's device address on each GPU (null: not looked up yet)
}  // namespace

// The array's address on the current device, looked up once per device.
// (In host code `scratch` names a host-side stand-in, never the device
// memory: every host use goes through this address.)
static int scratch_ptr(cpx_ctx *c, void **out)
{
  if (c->device < 0 || c->device >= kMaxDevices) return ctx_fail(c, CPX_EINVAL, "device ordinal out of range");
  if (!g_ptr[c->device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(c, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress");
    g_ptr[c->device] = p;
  }
  *out = g_ptr[c->device];
  return CPX_OK;
}

extern "C" int cpx_pi_batch_noalloc(cpx_ctx *c, const cpx_u128 *xs, size_t n, cpx_u128 *out)
{
  if (!c || (n && (!xs || !out))) return CPX_EINVAL;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
  if (!n) return CPX_OK;
  std::vector<HostQuery> hq(n);
  for (size_t i = 0; i < n; i++) {
    const int rc = make_host_query(c, ((u128)xs[i].hi << 64) | xs[i].lo, 0, 0, &hq[i]);
    if (rc) return rc;
  }
  std::lock_guard<std::mutex> g(g_mu[c->device]);
  void *region = nullptr;
  int rc = scratch_ptr(c, &region);
  if (rc) return rc;
  std::vector<ResultDev> res;
  rc = run_queries_at(c, hq, region, CPX_NOALLOC_BYTES, res);
  if (rc == CPX_ENOMEM) return ctx_fail(c, CPX_ENOMEM, "batch state exceeds the 128 MiB no-allocation region");
  if (rc) return rc;
  for (size_t i = 0; i < n; i++) out[i] = {res[i].pi[0], res[i].pi[1]};
  return CPX_OK;
}
