// site.hip -- SITE stage: greedy max-coverage selection over the viewshed
// bitsets, with the sequential reference's exact semantics (siting.cpp:
// 362-626, sequential argmax 553-562).
//
// Reference loop, per iteration:
//   U = unused candidates (all of them on the first iteration, else those
//       within 2*roi of the last pick)                       cpp:540-544
//   area[c] for c in U = new bits of shed c vs the cumulative bitset, or 0
//       under intervis when c's own cell is not covered yet   cpp:430-499
//   others keep their area (stale under intervis, by design)
//   pick max area, lowest index; stop at area 0               cpp:555-592
//   OR the pick into the cumulative bitset; stop above 95%    cpp:595-604
//
// Device design (one job):
//   site_bucket_*   counting sort of the candidates into square cells of
//                   edge cs; a grid row of cells is a "lane" of candidates
//   site_first_pass area = popcount(shed) for every candidate (cum is empty,
//                   no mask), streaming the whole shed array once
//   site_rowmax_init  per cell row, the best (area, -index) key
//   site_loop       persistent kernel, one grid barrier per
//                   iteration.  Each workgroup owns the cell rows r with
//                   r % G == g.  Iteration: every workgroup reduces the
//                   per-row keys to the winner w (identically), stages w's
//                   shed shifted into terrain words in LDS ("delta"), ORs its
//                   share of delta rows into cum, and recomputes the areas of
//                   the U candidates in its own rows against cum | delta
//                   (so reading cum while other workgroups are OR-ing the
//                   same delta in gives the same bits), then its rows' keys
//                   into the other half of a double-buffered key array.
// Keys: (area << 32) | ~index, 0 for a used candidate, so the maximum key is
// the largest area with the lowest index; a maximum with area 0 ends the
// loop.
#include "common.hpp"

namespace siting {

namespace {

constexpr int kLoopThreads = 1024;
constexpr int kMaxList = 1024;
constexpr int kDeltaLdsWords = 8192;  // 64 KiB

struct SiteArgs {
  const int32_t *obs;
  const uint64_t *sheds;
  int32_t *area;
  uint8_t *used;
  uint64_t *cum;
  uint64_t *rowmax;   // [2][ncr]
  int32_t *order;
  SiteState *st;
  const int32_t *sid;   // candidate ids sorted by cell
  const uint32_t *spos; // (x << 16 | y) in the same order
  const int32_t *cstart; // [ncells + 1]
  int n, roi, nr, nwpr, cw, nsheds;
  int cs, ncr;  // cell edge, cells per side
  int intervis;
  double npts;
};

__device__ __forceinline__ int cell_of(int x, int y, int cs, int ncr) {
  return (x / cs) * ncr + (y / cs);
}

__device__ __forceinline__ uint64_t cand_key(int area, int c) {
  return ((uint64_t)(uint32_t)area << 32) | (uint64_t)(~(uint32_t)c);
}

// ---- bucketing -----------------------------------------------------------

__global__ void fill_zero_i32(int32_t *p, int count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

__global__ void site_bucket_count(const int32_t *obs, int nsheds, int n, int cs,
                                  int ncr, int32_t *count) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nsheds) return;
  int o = obs[c];
  atomicAdd(&count[cell_of(o / n, o % n, cs, ncr)], 1);
}

// single workgroup exclusive scan of count[0..ncells) into start[0..ncells]
__global__ __launch_bounds__(1024) void site_bucket_scan(const int32_t *count,
                                                         int ncells,
                                                         int32_t *start,
                                                         int32_t *cursor) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (ncells + 1023) / 1024;
  const int lo = t * per, hi = min(lo + per, ncells);
  int s = 0;
  for (int i = lo; i < hi; i++) s += count[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? part[t - 1] : 0;
  for (int i = lo; i < hi; i++) {
    start[i] = run;
    cursor[i] = run;
    run += count[i];
  }
  if (t == 1023) start[ncells] = part[1023];
}

__global__ void site_bucket_scatter(const int32_t *obs, int nsheds, int n,
                                    int cs, int ncr, int32_t *cursor,
                                    int32_t *sid, uint32_t *spos) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nsheds) return;
  int o = obs[c];
  int x = o / n, y = o % n;
  int at = atomicAdd(&cursor[cell_of(x, y, cs, ncr)], 1);
  sid[at] = c;
  spos[at] = ((uint32_t)x << 16) | (uint32_t)y;
}

// ---- first iteration -----------------------------------------------------

// area[c] = popcount(shed c): with an empty cumulative bitset the shifted
// shed's new bits are exactly its bits (all inside the terrain), and no mask
// applies on the first iteration (cpp:452, nusedsheds == 0).
__global__ __launch_bounds__(256) void site_first_pass(SiteArgs a) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
  // clear the cumulative bitset and state (grid-stride)
  const size_t cum_words = (size_t)a.n * a.cw;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < cum_words;
       i += (size_t)gridDim.x * 256)
    a.cum[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.st->cumarea = 0;
    a.st->sum_u = (unsigned long long)a.nsheds;
    a.st->nused = 0;
    a.st->niter = 1;
    a.st->barrier = 0;
  }
  if (c >= a.nsheds) return;
  const int words = a.nr * a.nwpr;
  const uint64_t *sh = a.sheds + (size_t)c * words;
  int s = 0;
  for (int i = lane; i < words; i += 64) s += __popcll(sh[i]);
  s = wave_sum_i32(s);
  if (lane == 0) {
    a.area[c] = s;
    a.used[c] = 0;
  }
}

__device__ uint64_t block_max_u64(uint64_t v, uint64_t *red) {
  v = wave_max_u64(v);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  const int nw = blockDim.x >> 6;
  uint64_t r = 0;
  for (int i = 0; i < nw; i++) r = red[i] > r ? red[i] : r;
  return r;
}

__global__ __launch_bounds__(256) void site_rowmax_init(SiteArgs a) {
  __shared__ uint64_t red[16];
  const int r = blockIdx.x;
  const int lo = a.cstart[r * a.ncr], hi = a.cstart[(r + 1) * a.ncr];
  uint64_t m = 0;
  for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    const int c = a.sid[i];
    const uint64_t k = cand_key(a.area[c], c);
    m = k > m ? k : m;
  }
  m = block_max_u64(m, red);
  if (threadIdx.x == 0) a.rowmax[r] = m;
}

// ---- persistent loop -----------------------------------------------------

__device__ __forceinline__ void grid_sync(unsigned int *ctr, unsigned int &target,
                                          unsigned int G) {
  __syncthreads();
  if (threadIdx.x == 0) {
    target += G;
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    while ((int)(__hip_atomic_load(ctr, __ATOMIC_ACQUIRE,
                                   __HIP_MEMORY_SCOPE_AGENT) - target) < 0)
      __builtin_amdgcn_s_sleep(1);
  }
  __syncthreads();
}

// Shifted word dk (0..nwpr) of row dr of the winner's shed, i.e. the bits of
// terrain word w0w+dk that the winner covers (siting.cpp:408-417).
__device__ __forceinline__ uint64_t winner_word(const uint64_t *sw, int nwpr,
                                                int fbw, int dr, int dk) {
  const uint64_t *row = sw + (size_t)dr * nwpr;
  uint64_t cur = dk < nwpr ? row[dk] : 0;
  uint64_t v = cur >> fbw;
  if (fbw != 0 && dk > 0) v |= row[dk - 1] << (64 - fbw);
  return v;
}

template <bool kDeltaInLds>
__global__ __launch_bounds__(kLoopThreads) void site_loop(SiteArgs a) {
  extern __shared__ uint64_t delta[];  // nr * (nwpr + 1) when kDeltaInLds
  __shared__ uint64_t red[16];
  __shared__ int list_i[kMaxList];     // index into sid/spos
  __shared__ int list_acc[kMaxList];
  __shared__ int list_n;

  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
  const int n = a.n, roi = a.roi, nr = a.nr, nwpr = a.nwpr, cw = a.cw;
  const int dw = nwpr + 1;
  const int r2 = (2 * roi) * (2 * roi);

  int nused = 0, niter = 1, parity = 0;
  long long cumarea = 0, my_u = 0;
  unsigned int bar_target = 0;
  if (tid == 0) list_n = 0;

  for (;;) {
    // ---- argmax over the per-row keys (identical in every workgroup)
    uint64_t m = 0;
    const uint64_t *rm = a.rowmax + (size_t)parity * a.ncr;
    for (int i = tid; i < a.ncr; i += kLoopThreads) m = rm[i] > m ? rm[i] : m;
    m = block_max_u64(m, red);
    const int tw = (int)(m >> 32);
    if (tw == 0) break;  // cpp:589-592
    const int w = (int)(~(uint32_t)m);
    nused++;
    cumarea += tw;
    if (g == 0 && tid == 0) {
      a.used[w] = 1;
      a.order[nused - 1] = w;
    }
    const int wo = a.obs[w];
    const int wx = wo / n, wy = wo - (wo / n) * n;
    int w0w, fbw;
    col_split(wy - roi, w0w, fbw);
    const uint64_t *sw = a.sheds + (size_t)w * nr * nwpr;

    if (kDeltaInLds) {
      for (int i = tid; i < nr * dw; i += kLoopThreads) {
        const int dr = i / dw, dk = i - (i / dw) * dw;
        delta[i] = winner_word(sw, nwpr, fbw, dr, dk);
      }
      __syncthreads();
    }

    // ---- union: this workgroup's share of the winner's rows (cpp:393-419)
    {
      const int my_rows = (nr - g + G - 1) / G;  // rows g, g+G, ...
      for (int i = tid; i < my_rows * dw; i += kLoopThreads) {
        const int t = i / dw, dk = i - (i / dw) * dw;
        const int dr = g + t * G;
        const int x = wx - roi + dr, j = w0w + dk;
        if (x >= 0 && x < n && j >= 0 && j < cw) {
          const uint64_t v = kDeltaInLds ? delta[dr * dw + dk]
                                         : winner_word(sw, nwpr, fbw, dr, dk);
          a.cum[(size_t)x * cw + j] |= v;
        }
      }
    }
    if (100.0 * (double)cumarea / a.npts > 95.0) break;  // cpp:601-604
    niter++;

    // ---- update list rows: cells within 2*roi of the winner
    const int crlo = max(0, (wx - 2 * roi) / a.cs);
    const int crhi = min(a.ncr - 1, (wx + 2 * roi) / a.cs);
    const int cclo = max(0, (wy - 2 * roi) / a.cs);
    const int cchi = min(a.ncr - 1, (wy + 2 * roi) / a.cs);
    uint64_t *rm_next = a.rowmax + (size_t)(parity ^ 1) * a.ncr;

    // untouched rows carry their key over to the other buffer
    for (int r = g + tid * G; r < a.ncr; r += kLoopThreads * G)
      if (r < crlo || r > crhi) rm_next[r] = rm[r];

    for (int r = crlo + ((g - crlo % G) + G) % G; r <= crhi; r += G) {
      const int ulo = a.cstart[r * a.ncr + cclo];
      const int uhi = a.cstart[r * a.ncr + cchi + 1];
      for (int base = ulo; base < uhi; base += kMaxList) {
        const int cnt_max = min(kMaxList, uhi - base);
        // build the list of update candidates of this batch (cpp:540-544)
        if (tid < cnt_max) {
          const int i = base + tid;
          const int c = a.sid[i];
          const uint32_t ps = a.spos[i];
          const int cx = (int)(ps >> 16), cy = (int)(ps & 0xffffu);
          const int ex = cx - wx, ey = cy - wy;
          if (c != w && !a.used[c] && ex * ex + ey * ey <= r2) {
            int acc = 0;
            if (a.intervis) {  // cpp:452-457: own cell covered by cum'?
              uint64_t word = a.cum[(size_t)cx * cw + (cy >> 6)];
              const int dr = cx - wx + roi, dk = (cy >> 6) - w0w;
              if (dr >= 0 && dr < nr && dk >= 0 && dk <= nwpr)
                word |= kDeltaInLds ? delta[dr * dw + dk]
                                    : winner_word(sw, nwpr, fbw, dr, dk);
              if (!((word >> (63 - (cy & 63))) & 1u)) acc = -1;
            }
            const int at = atomicAdd(&list_n, 1);
            list_i[at] = i;
            list_acc[at] = acc;
          }
        }
        __syncthreads();
        const int cnt = list_n;
        my_u += cnt;
        // one task per (candidate, shed row)
        for (int t = tid; t < cnt * nr; t += kLoopThreads) {
          const int e = t / nr, row = t - (t / nr) * nr;
          if (list_acc[e] < 0) continue;
          const int i = list_i[e];
          const uint32_t ps = a.spos[i];
          const int cx = (int)(ps >> 16), cy = (int)(ps & 0xffffu);
          const int x = cx - roi + row;
          if (x < 0 || x >= n) continue;
          int w0c, fbc;
          col_split(cy - roi, w0c, fbc);
          const int jlo = max(w0c, 0), jhi = min((cy + roi) / 64, cw - 1);
          const uint64_t *sc = a.sheds + (size_t)a.sid[i] * nr * nwpr +
                               (size_t)row * nwpr;
          const uint64_t *crow = a.cum + (size_t)x * cw;
          const int dr = x - wx + roi;
          const bool wrow = dr >= 0 && dr < nr;
          int kk = jlo - w0c;
          uint64_t prev = kk > 0 ? sc[kk - 1] : 0;
          int s = 0;
          for (int j = jlo; j <= jhi; j++, kk++) {
            const uint64_t cur = kk < nwpr ? sc[kk] : 0;
            uint64_t moved = cur >> fbc;
            if (fbc != 0) moved |= prev << (64 - fbc);
            prev = cur;
            uint64_t cv = crow[j];
            const int dk = j - w0w;
            if (wrow && dk >= 0 && dk <= nwpr)
              cv |= kDeltaInLds ? delta[dr * dw + dk]
                                : winner_word(sw, nwpr, fbw, dr, dk);
            s += __popcll(moved & ~cv);
          }
          if (s) atomicAdd(&list_acc[e], s);
        }
        __syncthreads();
        if (tid < cnt) {
          const int acc = list_acc[tid];
          a.area[a.sid[list_i[tid]]] = acc < 0 ? 0 : acc;
        }
        __syncthreads();
        if (tid == 0) list_n = 0;
        __syncthreads();
      }
      // new key of row r
      const int lo = a.cstart[r * a.ncr], hi = a.cstart[(r + 1) * a.ncr];
      uint64_t km = 0;
      for (int i = lo + tid; i < hi; i += kLoopThreads) {
        const int c = a.sid[i];
        if (c == w || a.used[c]) continue;
        const uint64_t k = cand_key(a.area[c], c);
        km = k > km ? k : km;
      }
      km = block_max_u64(km, red);
      if (tid == 0) rm_next[r] = km;
    }
    parity ^= 1;
    grid_sync(&a.st->barrier, bar_target, (unsigned)G);
  }

  if (tid == 0 && my_u) atomicAdd(&a.st->sum_u, (unsigned long long)my_u);
  if (g == 0 && tid == 0) {
    a.st->nused = nused;
    a.st->niter = niter;
    a.st->cumarea = (unsigned long long)cumarea;
  }
}

}  // namespace

// Cell edge for the candidate buckets: about 30 cell rows across the 4*roi
// update window, so the rows an iteration touches spread over ~30
// workgroups.
static int site_cell_edge(int roi) {
  int cs = (4 * roi + 29) / 30;
  return cs < 4 ? 4 : cs;
}

int site_key_rows(int nrows, int roi) {
  const int cs = site_cell_edge(roi);
  return (nrows + cs - 1) / cs;
}

size_t site_scratch_bytes(int nrows, int roi, int nsheds) {
  const int cs = site_cell_edge(roi);
  const size_t ncr = (size_t)((nrows + cs - 1) / cs);
  const size_t ncells = ncr * ncr;
  size_t b = 0;
  b += ((ncells + 1) * 4 + 255) / 256 * 256 * 3;  // count, start, cursor
  b += ((size_t)nsheds * 4 + 255) / 256 * 256 * 2;  // sid, spos
  return b;
}

siting_status launch_site(const Geometry &g, const int32_t *obs,
                          const uint64_t *sheds, const SiteBuffers &b,
                          int device, hipStream_t s) {
  const int n = g.p.nrows, roi = g.p.roi, nsheds = g.d.nwanted;
  SiteArgs a{};
  a.obs = obs;
  a.sheds = sheds;
  a.area = b.area;
  a.used = b.used;
  a.cum = b.cum;
  a.rowmax = b.rowmax;
  a.order = b.order;
  a.st = b.state;
  a.n = n;
  a.roi = roi;
  a.nr = g.d.nr;
  a.nwpr = g.d.nwpr;
  a.cw = g.d.cumnwpr;
  a.nsheds = nsheds;
  a.cs = site_cell_edge(roi);
  a.ncr = (n + a.cs - 1) / a.cs;
  a.intervis = g.p.intervis;
  a.npts = (double)n * (double)n;
  const int ncells = a.ncr * a.ncr;

  // scratch carve-out (laid out by site_scratch_bytes)
  uint8_t *p = b.scratch;
  auto take = [&](size_t bytes) {
    uint8_t *q = p;
    p += (bytes + 255) / 256 * 256;
    return q;
  };
  int32_t *count = (int32_t *)take((size_t)(ncells + 1) * 4);
  int32_t *start = (int32_t *)take((size_t)(ncells + 1) * 4);
  int32_t *cursor = (int32_t *)take((size_t)(ncells + 1) * 4);
  int32_t *sid = (int32_t *)take((size_t)nsheds * 4);
  uint32_t *spos = (uint32_t *)take((size_t)nsheds * 4);
  a.sid = sid;
  a.spos = spos;
  a.cstart = start;

  hipLaunchKernelGGL(fill_zero_i32, dim3((ncells + 255) / 256), dim3(256), 0, s,
                     count, ncells);
  const int tb = 256, nbk = (nsheds + tb - 1) / tb;
  hipLaunchKernelGGL(site_bucket_count, dim3(nbk), dim3(tb), 0, s, obs, nsheds,
                     n, a.cs, a.ncr, count);
  hipLaunchKernelGGL(site_bucket_scan, dim3(1), dim3(1024), 0, s, count, ncells,
                     start, cursor);
  hipLaunchKernelGGL(site_bucket_scatter, dim3(nbk), dim3(tb), 0, s, obs,
                     nsheds, n, a.cs, a.ncr, cursor, sid, spos);
  hipLaunchKernelGGL(site_first_pass, dim3((nsheds + 3) / 4), dim3(256), 0, s, a);
  hipLaunchKernelGGL(site_rowmax_init, dim3(a.ncr), dim3(256), 0, s, a);
  SITING_HIP(hipGetLastError());

  // Persistent loop: as many workgroups as cell rows an update window can
  // touch, and at most half the CUs.  The grid barrier needs every
  // workgroup resident at once; a grid this far below the device's capacity
  // (one 1024-thread group per CU) is, since any other kernel that holds CUs
  // when it starts does not wait on it and retires.  A plain launch is used:
  // a cooperative launch reserves a device-wide resource that a second
  // process on the same GPU cannot get.
  const int touched = (4 * roi) / a.cs + 2;
  int cus = 0;
  SITING_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                   device));
  const bool lds_delta = (size_t)g.d.nr * (g.d.nwpr + 1) <= (size_t)kDeltaLdsWords;
  const size_t shmem = lds_delta ? sizeof(uint64_t) * g.d.nr * (g.d.nwpr + 1) : 0;
  int grid = touched;
  grid = grid < a.ncr ? grid : a.ncr;
  const int cap = cus / 2 > 0 ? cus / 2 : 1;
  grid = grid < cap ? grid : cap;
  grid = grid < 1 ? 1 : grid;
  if (lds_delta)
    hipLaunchKernelGGL(site_loop<true>, dim3(grid), dim3(kLoopThreads), shmem, s, a);
  else
    hipLaunchKernelGGL(site_loop<false>, dim3(grid), dim3(kLoopThreads), 0, s, a);
  SITING_HIP(hipGetLastError());
  return SITING_OK;
}

}  // namespace siting
