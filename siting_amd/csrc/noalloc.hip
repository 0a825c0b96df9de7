// noalloc.hip -- SITE on caller-owned device arrays, allocating nothing.
//
// siting_site_it_noalloc runs the same kernels as siting_site_it (bucketing,
// first pass, persistent greedy loop), but its working memory -- candidate
// areas, the cumulative bitset, the SITE state and the bucketing/loop
// scratch -- is carved from `scratch`, a fixed device array of this file,
// instead of hipMalloc'd buffers, and it creates no stream.  It is meant for
// callers that cannot allocate: code under HIP graph capture, tight serving
// loops, harnesses that forbid device allocation.
//
// The array lives on every device the library runs on until the process
// ends.  Host code never passes the array's name to a kernel (on the host it
// is only a stand-in for the symbol): the device address is looked up once
// per device with hipGetSymbolAddress, with that device current, and only
// that address is used.  Every call on one device shares its array, so calls
// that overlap on one device race; the kernels clear what must start at zero
// on every call.
#include <cstdio>
#include <cstring>
#include <mutex>

#include "job.hpp"

namespace siting {

namespace {

// The largest working set a call may use: 64 MiB.  A call needs about 31-70
// bytes per candidate, nrows^2 / 8 bytes of cumulative bitset and
// 256 (2 roi + 1)(nwpr + 1) bytes of per-round pick bits: 60 MiB for 1.4 M
// candidates on an 8192^2 terrain at roi 25, 19 MiB for the 262144
// candidates of an 8192^2 terrain at roi 200, 5.4 MiB for 65536 on 4096^2
// at roi 100.
constexpr size_t kScratchBytes = size_t(64) << 20;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[kMaxDevices];  // device address of scratch, per device

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

// Working-set layout of one call (byte offsets from the array's start).
struct NoAllocLayout {
  size_t area, cum, rowmax, order, state, site, total;
};

NoAllocLayout noalloc_layout(int nrows, int roi, int nsheds, int intervis, bool own_order) {
  NoAllocLayout L{};
  size_t off = 0;
  auto put = [&](size_t &field, size_t bytes) {
    field = off;
    off += align256(bytes ? bytes : 1);
  };
  const size_t cw = ((size_t)nrows + 63) / 64;
  put(L.area, 4 * (size_t)nsheds);
  put(L.cum, 8 * cum_words(nrows, (int)cw));
  put(L.rowmax, 2 * (size_t)site_key_rows(nrows, roi) * 8);
  put(L.order, own_order ? 4 * (size_t)nsheds : 0);
  put(L.state, sizeof(SiteState));
  put(L.site, site_scratch_bytes(nrows, roi, nsheds, intervis));
  L.total = off;
  return L;
}

// The device address of this device's array (looked up on first use).
siting_status scratch_ptr(int device, uint8_t **out) {
  if (device < 0 || device >= kMaxDevices)
    return hip_fail(hipErrorInvalidDevice, "siting_site_it_noalloc (device index)",
                    __FILE__, __LINE__);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                      "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    g_scratch_ptr[device] = (uint8_t *)p;
  }
  *out = g_scratch_ptr[device];
  return SITING_OK;
}

}  // namespace

}  // namespace siting

using namespace siting;

extern "C" {

siting_status siting_site_it_noalloc(int nrows, int roi, int intervis, int nsheds,
                                     const int32_t *obs, const uint64_t *sheds,
                                     uint8_t *selected, int32_t *order,
                                     siting_stats *stats, void *stream) {
  if (!obs || !sheds || !selected || nrows <= 0 || nrows > 32767 || roi < 1 ||
      roi > kMaxRoi || nsheds < 1 || (intervis != 0 && intervis != 1))
    return SITING_EINVAL;
  set_limit_error(nullptr);
  const NoAllocLayout L = noalloc_layout(nrows, roi, nsheds, intervis, order == nullptr);
  if (L.total > kScratchBytes) {
    char msg[200];
    std::snprintf(msg, sizeof msg,
                  "siting_site_it_noalloc: this call needs %zu bytes of working "
                  "memory, more than its fixed %zu bytes",
                  L.total, kScratchBytes);
    set_limit_error(msg);
    return SITING_ENOMEM;
  }
  int device = 0;
  SITING_HIP(hipGetDevice(&device));
  uint8_t *base = nullptr;
  siting_status st = scratch_ptr(device, &base);
  if (st != SITING_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  Geometry g{};
  g.p.nrows = nrows;
  g.p.roi = roi;
  g.p.intervis = intervis;
  g.d.nwanted = nsheds;
  g.d.nr = 2 * roi + 1;
  g.d.nwpr = (g.d.nr + 63) / 64;
  g.d.cumnwpr = (nrows + 63) / 64;
  g.d.shed_words = (int64_t)g.d.nr * g.d.nwpr;
  SiteState *state = (SiteState *)(base + L.state);
  SiteBuffers b{(int32_t *)(base + L.area), selected, (uint64_t *)(base + L.cum),
                (uint64_t *)(base + L.rowmax),
                order ? order : (int32_t *)(base + L.order), state, base + L.site};
  st = launch_site(g, obs, sheds, b, device, s);
  if (st != SITING_OK || !stats) return st;
  SiteState h;
  SITING_HIP(hipMemcpyAsync(&h, state, sizeof h, hipMemcpyDeviceToHost, s));
  SITING_HIP(hipStreamSynchronize(s));
  std::memset(stats, 0, sizeof *stats);
  stats->nused = h.nused;
  stats->niter = h.niter;
  stats->cumarea = (int64_t)h.cumarea;
  stats->sum_u = (int64_t)h.sum_u;
  stats->rounds = h.rounds;
  stats->coverage = 100.0 * (double)h.cumarea / ((double)nrows * (double)nrows);
  if (h.niter < 0) return site_loop_failed(h.niter);
  return SITING_OK;
}

}  // extern "C"
