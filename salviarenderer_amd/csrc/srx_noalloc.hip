// srx_draw_noalloc: srx_draw for callers that cannot allocate device memory (code under HIP graph
// capture, serving loops and harnesses that forbid hipMalloc).  Every buffer the draw needs beyond
// its inputs and outputs -- vertex-shader outputs, triangle setup, per-triangle record counts and
// their scan space, binned records, sort histograms, tile ranges and order, the visibility work
// table and per-sample winners -- is carved from `scratch`, one fixed region per GPU that lives
// until the process ends.  The renderer's own small blocks (device counters, statistics, pinned
// staging) were allocated when it was created.  No counterpart in the reference.
#include "srx_runtime_internal.h"

namespace {

// The region; in host code this name is only a stand-in for the symbol: device addresses come
// from hipGetSymbolAddress, once per GPU, and only those are passed to kernels or copies.
__device__ __align__(256) uint8_t scratch[SRX_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
uint8_t* g_ptr[MAX_DEVICES] = {};

// The region's device address on `dev`, which must be the current device.
int32_t scratch_on(int dev, uint8_t** out) {
  if (dev < 0 || dev >= MAX_DEVICES) return srx_rt::fail(SRX_E_UNSUPPORTED, "device ordinal beyond 63");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return srx_rt::hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return srx_rt::fail(SRX_E_HIP, "hipGetSymbolAddress(scratch) returned no address");
    g_ptr[dev] = static_cast<uint8_t*>(p);
  }
  *out = g_ptr[dev];
  return SRX_OK;
}

}  // namespace

using namespace srx_rt;

extern "C" int32_t srx_draw_noalloc(srx_renderer rr, size_t start, size_t primcnt) {
  auto* r = as<renderer_obj>(rr, K_RENDERER);
  if (!r) return SRX_INVALID_PARAMETER;
  (void)start;  // the index fetcher ignores the start of non-indexed draws (index_fetcher.cpp:103-107)
  if (primcnt > MAX_DRAW_PRIMITIVES) return fail(SRX_E_UNSUPPORTED, "more than 1,431,650,304 primitives in one draw");
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != r->device) return fail(SRX_INVALID_PARAMETER, "call srx_draw_noalloc with the renderer's device current");
  uint8_t* base = nullptr;
  RCHK(scratch_on(dev, &base));
  const fixed_scratch fx{base, SRX_NOALLOC_SCRATCH_BYTES};
  return execute_draw(r, false, 0, static_cast<uint32_t>(primcnt), 0, &fx);
}
