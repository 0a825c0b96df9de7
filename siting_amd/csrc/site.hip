// site.hip -- SITE stage: greedy max-coverage selection over the viewshed
// bitsets, with the sequential reference's exact semantics (siting.cpp:
// 362-626, sequential argmax 553-562).
//
// Reference loop, per iteration:
//   U = unused candidates (all of them on the first iteration, else those
//       within 2*roi of the last pick)                       cpp:540-544
//   area[c] for c in U = new bits of shed c vs the cumulative bitset, or 0
//       under intervis when c's own cell is not covered yet   cpp:430-499
//   every other candidate keeps its last area (so under intervis a far
//       candidate keeps its unmasked first-iteration area)
//   pick max area, lowest index; stop at area 0               cpp:555-592
//   OR the pick into the cumulative bitset; stop above 95%    cpp:595-604
//
// Device design (one job):
//   Candidates are bucketed by position.  Square cells of edge cs tile the
//   terrain; a cell row is split into Q column classes (cell column mod Q);
//   a (cell row, class) pair is a "unit".  Units are dealt to the G
//   workgroups of the persistent loop round-robin, and a counting sort puts
//   every workgroup's candidates contiguously, unit by unit, cell by cell.
//   site_first_pass  area = popcount(shed) for every candidate (cum is empty
//                    and no mask applies), streaming the whole shed array once
//   site_loop        persistent kernel, one grid barrier per iteration.  A
//                    workgroup keeps its candidates' ids, positions, areas and
//                    used flags in LDS (or, when they do not fit, in a global
//                    slice of the same layout) and the best key of each of its
//                    units.  Iteration: every workgroup reduces the unit keys
//                    to the winner w (identically); the owner marks w used;
//                    every workgroup stages w's shed, shifted into terrain
//                    words, in LDS ("delta"), ORs its share of delta rows into
//                    cum, recomputes the areas of the update candidates of its
//                    own units against cum | delta (reading cum while others
//                    OR the same delta in gives the same bits), and publishes
//                    the new unit keys into the other half of a double-buffered
//                    key array.
// Keys: (area << 32) | ~index, 0 for a used candidate, so the maximum key is
// the largest area with the lowest index; a maximum with area 0 ends the
// loop.  Each key travels with its candidate's position.
#include <cstdlib>

#include "common.hpp"

namespace siting {

namespace {

constexpr int kLoopThreads = 1024;
constexpr int kLoopWaves = kLoopThreads / kWave;
constexpr int kMaxList = 1024;
constexpr size_t kLdsBudget = 140 * 1024;  // dynamic LDS of the loop kernel
constexpr int kMaxGroups = 128;

struct SiteArgs {
  const int32_t *obs;
  const uint64_t *sheds;
  int32_t *area;     // [nsheds] first-pass area by candidate id
  uint8_t *used;     // [nsheds] selected flags (the output mask)
  uint64_t *cum;     // [n * cw]
  uint64_t *keys;    // [2][nunits]
  uint32_t *kpos;    // [2][nunits]
  int32_t *order;    // [nsheds]
  SiteState *st;
  int32_t *bcount;   // [nbuckets] counts, then scatter cursors
  int32_t *bstart;   // [nbuckets + 1]
  int32_t *chunk;    // [nchunks] scan partial sums
  int32_t *sid;      // [nsheds] candidate ids, bucket order
  uint32_t *spos;    // [nsheds] (x << 16 | y), bucket order
  int32_t *sT;       // [nsheds] area state of workgroups in global mode
  uint8_t *sused;    // [nsheds] used state of workgroups in global mode
  int n, roi, nr, nwpr, cw, nsheds, intervis;
  int cs, ncr, Q, nunits, G, upw, kpu, nbuckets, cap;
  // dynamic LDS layout (byte offsets)
  int o_delta, o_ukey, o_upos, o_dirty, o_off, o_c, o_pos, o_t, o_used;
  int off_lds;  // unit cell offsets kept in LDS (else read from bstart)
  double npts;
};

__device__ __forceinline__ uint64_t cand_key(int area, int c) {
  return ((uint64_t)(uint32_t)area << 32) | (uint64_t)(~(uint32_t)c);
}

__device__ __forceinline__ int bucket_of(const SiteArgs &a, int x, int y) {
  const int r = x / a.cs, cc = y / a.cs;
  const int q = cc % a.Q, k = cc / a.Q;
  const int u = r * a.Q + q;
  return ((u % a.G) * a.upw + u / a.G) * a.kpu + k;
}

// ---- bucketing (counting sort) -------------------------------------------

__global__ void site_zero_counts(int32_t *p, int count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

__global__ void site_bucket_count(SiteArgs a) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.nsheds) return;
  int o = a.obs[c];
  atomicAdd(&a.bcount[bucket_of(a, o / a.n, o % a.n)], 1);
}

// inclusive scan of one value per thread over a 1024-thread block
__device__ int block_inclusive_scan(int v, int *wsum) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int o = __shfl_up(v, off, kWave);
    if (lane >= off) v += o;
  }
  if (lane == 63) wsum[wid] = v;
  __syncthreads();
  if (wid == 0) {
    int s = lane < kLoopWaves ? wsum[lane] : 0;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      int o = __shfl_up(s, off, kWave);
      if (lane >= off) s += o;
    }
    if (lane < kLoopWaves) wsum[lane] = s;
  }
  __syncthreads();
  if (wid > 0) v += wsum[wid - 1];
  return v;
}

__global__ __launch_bounds__(1024) void site_scan_chunks(SiteArgs a) {
  __shared__ int red[kLoopWaves];
  const int i = blockIdx.x * 1024 + threadIdx.x;
  int v = i < a.nbuckets ? a.bcount[i] : 0;
  v = wave_sum_i32(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < kLoopWaves; w++) s += red[w];
    a.chunk[blockIdx.x] = s;
  }
}

// bstart = exclusive scan of bcount; bcount becomes the scatter cursor
__global__ __launch_bounds__(1024) void site_scan_apply(SiteArgs a) {
  __shared__ int wsum[kLoopWaves];
  __shared__ int prefix;
  int p = 0;
  for (int b = threadIdx.x; b < (int)blockIdx.x; b += 1024) p += a.chunk[b];
  p = wave_sum_i32(p);
  if (threadIdx.x == 0) prefix = 0;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) atomicAdd(&prefix, p);
  __syncthreads();
  const int i = blockIdx.x * 1024 + threadIdx.x;
  const int v = i < a.nbuckets ? a.bcount[i] : 0;
  const int inc = block_inclusive_scan(v, wsum);
  const int ex = prefix + inc - v;
  if (i < a.nbuckets) {
    a.bstart[i] = ex;
    a.bcount[i] = ex;
  }
  if (i == a.nbuckets - 1) a.bstart[a.nbuckets] = ex + v;
}

__global__ void site_bucket_scatter(SiteArgs a) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.nsheds) return;
  int o = a.obs[c];
  int x = o / a.n, y = o % a.n;
  int at = atomicAdd(&a.bcount[bucket_of(a, x, y)], 1);
  a.sid[at] = c;
  a.spos[at] = ((uint32_t)x << 16) | (uint32_t)y;
}

// ---- first iteration -----------------------------------------------------

// area[c] = popcount(shed c): with an empty cumulative bitset the shifted
// shed's new bits are exactly its bits (all inside the terrain), and no mask
// applies on the first iteration (cpp:452, nusedsheds == 0).
__global__ __launch_bounds__(256) void site_first_pass(SiteArgs a) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6);
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

// ---- persistent loop -----------------------------------------------------

// Grid barrier: release on arrival, relaxed spin, one acquire fence after.
// (An acquire load inside the spin would invalidate the cache on every poll,
// and the polling groups would keep evicting the working set of the groups
// still computing.)
__device__ __forceinline__ void grid_sync(unsigned int *ctr, unsigned int &target,
                                          unsigned int G) {
  __syncthreads();
  if (threadIdx.x == 0) {
    target += G;
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    while ((int)(__hip_atomic_load(ctr, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT) - target) < 0)
      __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
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

struct Winner {
  int w, wx, wy, w0w, fbw, nr, nwpr, dw, roi;
  const uint64_t *sw;
  const uint64_t *delta;  // LDS rows of dw words, or null
  __device__ __forceinline__ uint64_t word(int dr, int dk) const {
    return delta ? delta[dr * dw + dk] : winner_word(sw, nwpr, fbw, dr, dk);
  }
};

// New bits of candidate (cx, cy)'s shed against cum | winner, summed over
// the wavefront (siting.cpp:456-491).  Work items are (shed row, terrain
// word) pairs inside the terrain; lane l takes items l, l+64, ..., so a load
// instruction reads consecutive words of the shed (rows are contiguous).
// kBatch items per lane are loaded before any is used.
constexpr int kBatch = 8;

__device__ __forceinline__ int wave_new_bits(const uint64_t *sh, int cx, int cy,
                                             const SiteArgs &a, const Winner &W,
                                             int lane) {
  const int roi = a.roi, nwpr = a.nwpr, cw = a.cw, n = a.n;
  int w0c, fbc;
  col_split(cy - roi, w0c, fbc);
  const int klo = max(w0c, 0) - w0c;                      // first word slot
  const int khi = min((cy + roi) >> 6, cw - 1) - w0c;     // last word slot
  const int nk = khi - klo + 1;                           // 1 .. nwpr + 1
  const int rlo = max(0, cx - roi) - (cx - roi);
  const int rhi = min(n - 1, cx + roi) - (cx - roi);
  const int items = (rhi - rlo + 1) * nk;
  const float inv_nk = 1.0f / (float)nk;
  // (it + 0.5) * fl(1/nk) is within 0.5/nk of (it + 0.5)/nk while it < 2^21
  const bool fdiv = items < (1 << 21);
  const int xo = cx - roi, dro = xo - W.wx + roi;
  int s = 0;
  for (int base = 0; base < items; base += kWave * kBatch) {
    uint64_t cur[kBatch], prv[kBatch], cv[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      const int it = base + u * kWave + lane;
      const bool ok = it < items;
      const int q = fdiv ? (int)(((float)it + 0.5f) * inv_nk) : it / nk;
      const int row = rlo + q, k = klo + (it - q * nk);
      const uint64_t *srow = sh + (size_t)row * nwpr;
      cur[u] = (ok && k < nwpr) ? srow[k] : 0;
      prv[u] = (ok && k > 0) ? srow[k - 1] : 0;
      uint64_t c = ok ? a.cum[(size_t)(xo + row) * cw + (w0c + k)] : ~0ull;
      const int dr = dro + row, dk = w0c + k - W.w0w;
      if (ok && dr >= 0 && dr < W.nr && dk >= 0 && dk <= W.nwpr) c |= W.word(dr, dk);
      cv[u] = c;
    }
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      uint64_t moved = cur[u] >> fbc;
      if (fbc != 0) moved |= prv[u] << (64 - fbc);
      s += __popcll(moved & ~cv[u]);
    }
  }
  return wave_sum_i32(s);
}

template <int MW>
__global__ __launch_bounds__(kLoopThreads) void site_loop(SiteArgs a) {
  extern __shared__ __align__(16) uint8_t smem[];
  __shared__ uint64_t redk[kLoopWaves];
  __shared__ uint32_t redp[kLoopWaves];
  __shared__ int list_i[kMaxList];
  __shared__ int list_acc[kMaxList];
  __shared__ int list_n;

  constexpr bool kDeltaLds = MW > 0;
  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
  const int lane = tid & 63, wid = tid >> 6;
  const int n = a.n, roi = a.roi, nr = a.nr, nwpr = a.nwpr, cw = a.cw;
  const int Q = a.Q, upw = a.upw, kpu = a.kpu;
  const int r2 = (2 * roi) * (2 * roi);
  const size_t S = (size_t)nr * nwpr;

  uint64_t *delta = (uint64_t *)(smem + a.o_delta);
  uint64_t *ukey = (uint64_t *)(smem + a.o_ukey);
  uint32_t *upos = (uint32_t *)(smem + a.o_upos);
  uint8_t *dirty = smem + a.o_dirty;
  int32_t *off_l = (int32_t *)(smem + a.o_off);

  // ---- this workgroup's candidates
  const int B0 = a.bstart[(size_t)g * upw * kpu];
  const int B1 = a.bstart[(size_t)(g + 1) * upw * kpu];
  const int cnt = B1 - B0;
  const bool in_lds = cnt <= a.cap;
  const int32_t *C = in_lds ? (const int32_t *)(smem + a.o_c) : a.sid + B0;
  const uint32_t *P = in_lds ? (const uint32_t *)(smem + a.o_pos) : a.spos + B0;
  int32_t *T = in_lds ? (int32_t *)(smem + a.o_t) : a.sT + B0;
  uint8_t *U = in_lds ? (uint8_t *)(smem + a.o_used) : a.sused + B0;
  for (int i = tid; i < cnt; i += kLoopThreads) {
    const int c = a.sid[B0 + i];
    if (in_lds) {
      ((int32_t *)(smem + a.o_c))[i] = c;
      ((uint32_t *)(smem + a.o_pos))[i] = a.spos[B0 + i];
    }
    T[i] = a.area[c];
    U[i] = 0;
  }
  // start of cell k of owned unit ul, relative to B0 (k == kpu: unit end)
  const int32_t *bst = a.bstart + (size_t)g * upw * kpu;
  auto OFF = [&](int ul, int k) -> int {
    return a.off_lds ? off_l[ul * (kpu + 1) + k] : bst[(size_t)ul * kpu + k] - B0;
  };
  if (a.off_lds)
    for (int e = tid; e < upw * (kpu + 1); e += kLoopThreads) {
      const int ul = e / (kpu + 1), k = e - ul * (kpu + 1);
      off_l[e] = bst[(size_t)ul * kpu + k] - B0;
    }
  if (tid == 0) list_n = 0;
  __syncthreads();

  // best key of owned unit ul -> ukey/upos (one wavefront per unit)
  auto unit_key = [&](int ul, int w) {
    const int lo = OFF(ul, 0), hi = OFF(ul, kpu);
    uint64_t best = 0;
    uint32_t bpos = 0;
    for (int i = lo + lane; i < hi; i += kWave) {
      const int c = C[i];
      if (U[i] || c == w) continue;
      const uint64_t k = cand_key(T[i], c);
      if (k > best) {
        best = k;
        bpos = P[i];
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint64_t ok = __shfl_xor(best, o, kWave);
      const uint32_t op = __shfl_xor(bpos, o, kWave);
      if (ok > best) {
        best = ok;
        bpos = op;
      }
    }
    if (lane == 0) {
      ukey[ul] = best;
      upos[ul] = bpos;
    }
  };

  for (int ul = wid; ul < upw; ul += kLoopWaves) unit_key(ul, -1);
  __syncthreads();
  for (int ul = tid; ul < upw; ul += kLoopThreads) {
    const int u = ul * G + g;
    if (u < a.nunits) {
      a.keys[u] = ukey[ul];
      a.kpos[u] = upos[ul];
    }
  }
  unsigned int bar_target = 0;
  grid_sync(&a.st->barrier, bar_target, (unsigned)G);

  int nused = 0, niter = 1, parity = 0;
  long long cumarea = 0, my_u = 0;
  const int dw = nwpr + 1;

  for (;;) {
    // ---- argmax over the unit keys (identical in every workgroup)
    const uint64_t *kk = a.keys + (size_t)parity * a.nunits;
    const uint32_t *kp = a.kpos + (size_t)parity * a.nunits;
    uint64_t m = 0;
    uint32_t mp = 0;
    for (int i = tid; i < a.nunits; i += kLoopThreads) {
      const uint64_t v = kk[i];
      if (v > m) {
        m = v;
        mp = kp[i];
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint64_t ok = __shfl_xor(m, o, kWave);
      const uint32_t op = __shfl_xor(mp, o, kWave);
      if (ok > m) {
        m = ok;
        mp = op;
      }
    }
    if (lane == 0) {
      redk[wid] = m;
      redp[wid] = mp;
    }
    __syncthreads();
    m = 0;
    mp = 0;
    for (int i = 0; i < kLoopWaves; i++)
      if (redk[i] > m) {
        m = redk[i];
        mp = redp[i];
      }
    const int tw = (int)(m >> 32);
    if (tw == 0) break;  // cpp:589-592
    if (nused >= a.nsheds) {  // cannot happen: every pick retires a candidate
      niter = -1;
      break;
    }
    const int w = (int)(~(uint32_t)m);
    nused++;
    cumarea += tw;
    if (g == 0 && tid == 0) {
      a.used[w] = 1;
      a.order[nused - 1] = w;
    }
    Winner W;
    W.w = w;
    W.wx = (int)(mp >> 16);
    W.wy = (int)(mp & 0xffffu);
    col_split(W.wy - roi, W.w0w, W.fbw);
    W.nr = nr;
    W.nwpr = nwpr;
    W.dw = dw;
    W.roi = roi;
    W.sw = a.sheds + (size_t)w * S;
    W.delta = kDeltaLds ? delta : nullptr;

    // the owner marks w used in its state
    const int w_r = W.wx / a.cs, w_cc = W.wy / a.cs;
    const int w_u = w_r * Q + w_cc % Q;
    if (w_u % G == g) {
      const int ul = w_u / G, k = w_cc / Q;
      const int lo = OFF(ul, k), hi = OFF(ul, k + 1);
      for (int i = lo + tid; i < hi; i += kLoopThreads)
        if (C[i] == w) U[i] = 1;
    }
    if (kDeltaLds)
      for (int i = tid; i < nr * dw; i += kLoopThreads) {
        const int dr = i / dw, dk = i - (i / dw) * dw;
        delta[i] = winner_word(W.sw, nwpr, W.fbw, dr, dk);
      }
    __syncthreads();

    // ---- union: this workgroup's share of the winner's rows (cpp:393-419)
    {
      const int my_rows = (nr - g + G - 1) / G;
      for (int i = tid; i < my_rows * dw; i += kLoopThreads) {
        const int t = i / dw, dk = i - (i / dw) * dw;
        const int dr = g + t * G;
        const int x = W.wx - roi + dr, j = W.w0w + dk;
        if (x >= 0 && x < n && j >= 0 && j < cw) a.cum[(size_t)x * cw + j] |= W.word(dr, dk);
      }
    }
    if (100.0 * (double)cumarea / a.npts > 95.0) break;  // cpp:601-604
    niter++;

    // ---- update candidates of the owned units near the winner
    const int crlo = max(0, (W.wx - 2 * roi) / a.cs);
    const int crhi = min(a.ncr - 1, (W.wx + 2 * roi) / a.cs);
    const int cclo = max(0, (W.wy - 2 * roi) / a.cs);
    const int cchi = min(a.ncr - 1, (W.wy + 2 * roi) / a.cs);
    for (int ul = 0; ul < upw; ul++) {
      const int u = ul * G + g;
      const int r = u / Q;
      const bool touched = u < a.nunits && r >= crlo && r <= crhi;
      if (tid == 0) dirty[ul] = touched;
      if (!touched) continue;
      const int q = u - r * Q;
      const int klo = cclo <= q ? 0 : (cclo - q + Q - 1) / Q;
      const int khi = cchi < q ? -1 : (cchi - q) / Q;
      if (klo > khi) continue;
      const int ulo = OFF(ul, klo), uhi = OFF(ul, khi + 1);
      for (int base = ulo; base < uhi; base += kMaxList) {
        const int span = min(kMaxList, uhi - base);
        if (tid < span) {  // update list, cpp:540-544
          const int i = base + tid;
          const uint32_t ps = P[i];
          const int cx = (int)(ps >> 16), cy = (int)(ps & 0xffffu);
          const int ex = cx - W.wx, ey = cy - W.wy;
          if (!U[i] && C[i] != w && ex * ex + ey * ey <= r2) {
            int acc = 0;
            if (a.intervis) {  // cpp:452-457: own cell covered by cum'?
              uint64_t word = a.cum[(size_t)cx * cw + (cy >> 6)];
              const int dr = cx - W.wx + roi, dk = (cy >> 6) - W.w0w;
              if (dr >= 0 && dr < nr && dk >= 0 && dk <= nwpr) word |= W.word(dr, dk);
              if (!((word >> (63 - (cy & 63))) & 1u)) acc = -1;
            }
            const int at = atomicAdd(&list_n, 1);
            list_i[at] = i;
            list_acc[at] = acc;
          }
        }
        __syncthreads();
        const int lcnt = list_n;
        my_u += lcnt;
        // one wavefront per update candidate, lanes over consecutive words
        for (int e = wid; e < lcnt; e += kLoopWaves) {
          if (list_acc[e] < 0) continue;
          const int i = list_i[e];
          const uint32_t ps = P[i];
          const int s = wave_new_bits(a.sheds + (size_t)C[i] * S, (int)(ps >> 16),
                                      (int)(ps & 0xffffu), a, W, lane);
          if (lane == 0) list_acc[e] = s;
        }
        __syncthreads();
        if (tid < lcnt) {
          const int acc = list_acc[tid];
          T[list_i[tid]] = acc < 0 ? 0 : acc;
        }
        __syncthreads();
        if (tid == 0) list_n = 0;
        __syncthreads();
      }
    }
    // the winner's unit changed too (w is used now)
    if (w_u % G == g && tid == 0) dirty[w_u / G] = 1;
    __syncthreads();
    for (int ul = wid; ul < upw; ul += kLoopWaves)
      if (dirty[ul]) unit_key(ul, w);
    __syncthreads();
    uint64_t *kn = a.keys + (size_t)(parity ^ 1) * a.nunits;
    uint32_t *pn = a.kpos + (size_t)(parity ^ 1) * a.nunits;
    for (int ul = tid; ul < upw; ul += kLoopThreads) {
      const int u = ul * G + g;
      if (u < a.nunits) {
        kn[u] = ukey[ul];
        pn[u] = upos[ul];
      }
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

// ---- host-side plan --------------------------------------------------------

struct Plan {
  int cs, ncr, Q, nunits, G, upw, kpu, nbuckets, nchunks, cap, mw, off_lds;
  size_t lds;
  int o_delta, o_ukey, o_upos, o_dirty, o_off, o_c, o_pos, o_t, o_used;
};

int env_q() {
  const char *e = std::getenv("SITING_SITE_Q");
  int q = e ? std::atoi(e) : 2;
  return q < 1 ? 1 : (q > 8 ? 8 : q);
}

// Cell edge: about 30 cell rows across the 4*roi update window, and at most
// 512 cells per side.
int cell_edge(int n, int roi) {
  int cs = (4 * roi + 29) / 30;
  const int floor_n = (n + 511) / 512;
  cs = cs < floor_n ? floor_n : cs;
  return cs < 4 ? 4 : cs;
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

Plan make_plan(int n, int roi, int nsheds, int groups_cap) {
  Plan p{};
  p.cs = cell_edge(n, roi);
  p.ncr = (n + p.cs - 1) / p.cs;
  p.Q = env_q();
  if (p.Q > p.ncr) p.Q = p.ncr;
  p.nunits = p.ncr * p.Q;
  const int touched = ((4 * roi) / p.cs + 2) * p.Q;
  int G = touched < p.nunits ? touched : p.nunits;
  G = G < groups_cap ? G : groups_cap;
  G = G < kMaxGroups ? G : kMaxGroups;
  p.G = G < 1 ? 1 : G;
  p.upw = (p.nunits + p.G - 1) / p.G;
  p.kpu = (p.ncr + p.Q - 1) / p.Q;
  p.nbuckets = p.G * p.upw * p.kpu;
  p.nchunks = (p.nbuckets + 1023) / 1024;
  const int nwpr = (2 * roi + 1 + 63) / 64;
  p.mw = nwpr + 1 <= 3 ? 3 : (nwpr + 1 <= 5 ? 5 : (nwpr + 1 <= 8 ? 8 : 0));
  size_t o = 0;
  p.o_delta = (int)o;
  if (p.mw) o += sizeof(uint64_t) * (size_t)(2 * roi + 1) * (nwpr + 1);
  p.o_ukey = (int)o;
  o += sizeof(uint64_t) * p.upw;
  p.o_upos = (int)o;
  o += sizeof(uint32_t) * p.upw;
  p.o_dirty = (int)o;
  o = align_up(o + (size_t)p.upw, 16);
  p.o_off = (int)o;
  const size_t off_bytes = sizeof(int32_t) * (size_t)p.upw * (p.kpu + 1);
  p.off_lds = o + off_bytes <= kLdsBudget / 2;
  if (p.off_lds) o += off_bytes;
  o = align_up(o, 16);
  const size_t fixed = o;
  const size_t room = kLdsBudget > fixed ? kLdsBudget - fixed : 0;
  int cap = (int)(room / 13);
  cap = cap > 16 ? (cap - 16) / 16 * 16 : 0;
  p.cap = cap;
  p.o_c = (int)o;
  o += 4 * (size_t)cap;
  p.o_pos = (int)o;
  o += 4 * (size_t)cap;
  p.o_t = (int)o;
  o += 4 * (size_t)cap;
  p.o_used = (int)o;
  o += (size_t)cap;
  p.lds = align_up(o, 16);
  return p;
}

struct Scratch {
  int32_t *bcount, *bstart, *chunk, *sid, *sT;
  uint32_t *spos;
  uint8_t *sused;
  uint64_t *keys;
  uint32_t *kpos;
};

size_t scratch_layout(const Plan &p, int nsheds, uint8_t *base, Scratch *out) {
  size_t o = 0;
  auto take = [&](size_t bytes) {
    uint8_t *q = base ? base + o : nullptr;
    o += align_up(bytes ? bytes : 1, 256);
    return q;
  };
  Scratch s{};
  // bucket arrays sized for any group count up to kMaxGroups:
  // G * ceil(nunits / G) <= nunits + G - 1
  const size_t nb_max = ((size_t)p.nunits + kMaxGroups) * p.kpu;
  s.bcount = (int32_t *)take(4 * nb_max);
  s.bstart = (int32_t *)take(4 * (nb_max + 1));
  s.chunk = (int32_t *)take(4 * ((nb_max + 1023) / 1024));
  s.sid = (int32_t *)take(4 * (size_t)nsheds);
  s.spos = (uint32_t *)take(4 * (size_t)nsheds);
  s.sT = (int32_t *)take(4 * (size_t)nsheds);
  s.sused = (uint8_t *)take((size_t)nsheds);
  s.keys = (uint64_t *)take(8 * 2 * (size_t)p.nunits);
  s.kpos = (uint32_t *)take(4 * 2 * (size_t)p.nunits);
  if (out) *out = s;
  return o;
}

}  // namespace

// The plan's group count depends on the device; the scratch size is taken
// for the largest group count, which bounds every smaller one.
size_t site_scratch_bytes(int nrows, int roi, int nsheds) {
  return scratch_layout(make_plan(nrows, roi, nsheds, kMaxGroups), nsheds, nullptr,
                        nullptr);
}

int site_key_rows(int nrows, int roi) { return 1; }

siting_status launch_site(const Geometry &g, const int32_t *obs,
                          const uint64_t *sheds, const SiteBuffers &b,
                          int device, hipStream_t s) {
  const int n = g.p.nrows, roi = g.p.roi, nsheds = g.d.nwanted;
  int cus = 0;
  SITING_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                   device));
  // The grid barrier needs every workgroup of the loop resident at once; a
  // grid of at most half the CUs (one 1024-thread group per CU) is, since any
  // other kernel holding CUs when it starts does not wait on it and retires.
  // A plain launch is used: a cooperative launch reserves a device-wide
  // resource that a second process on the same GPU cannot get.
  const Plan p = make_plan(n, roi, nsheds, cus / 2 > 0 ? cus / 2 : 1);
  Scratch sc;
  scratch_layout(p, nsheds, b.scratch, &sc);

  SiteArgs a{};
  a.obs = obs;
  a.sheds = sheds;
  a.area = b.area;
  a.used = b.used;
  a.cum = b.cum;
  a.keys = sc.keys;
  a.kpos = sc.kpos;
  a.order = b.order;
  a.st = b.state;
  a.bcount = sc.bcount;
  a.bstart = sc.bstart;
  a.chunk = sc.chunk;
  a.sid = sc.sid;
  a.spos = sc.spos;
  a.sT = sc.sT;
  a.sused = sc.sused;
  a.n = n;
  a.roi = roi;
  a.nr = g.d.nr;
  a.nwpr = g.d.nwpr;
  a.cw = g.d.cumnwpr;
  a.nsheds = nsheds;
  a.intervis = g.p.intervis;
  a.cs = p.cs;
  a.ncr = p.ncr;
  a.Q = p.Q;
  a.nunits = p.nunits;
  a.G = p.G;
  a.upw = p.upw;
  a.kpu = p.kpu;
  a.nbuckets = p.nbuckets;
  a.cap = p.cap;
  a.o_delta = p.o_delta;
  a.o_ukey = p.o_ukey;
  a.o_upos = p.o_upos;
  a.o_dirty = p.o_dirty;
  a.o_off = p.o_off;
  a.off_lds = p.off_lds;
  a.o_c = p.o_c;
  a.o_pos = p.o_pos;
  a.o_t = p.o_t;
  a.o_used = p.o_used;
  a.npts = (double)n * (double)n;

  const int tb = 256, nbk = (nsheds + tb - 1) / tb;
  hipLaunchKernelGGL(site_zero_counts, dim3((p.nbuckets + 255) / 256), dim3(256), 0, s,
                     sc.bcount, p.nbuckets);
  hipLaunchKernelGGL(site_bucket_count, dim3(nbk), dim3(tb), 0, s, a);
  hipLaunchKernelGGL(site_scan_chunks, dim3(p.nchunks), dim3(1024), 0, s, a);
  hipLaunchKernelGGL(site_scan_apply, dim3(p.nchunks), dim3(1024), 0, s, a);
  hipLaunchKernelGGL(site_bucket_scatter, dim3(nbk), dim3(tb), 0, s, a);
  hipLaunchKernelGGL(site_first_pass, dim3((nsheds + 3) / 4), dim3(256), 0, s, a);
  SITING_HIP(hipGetLastError());
  switch (p.mw) {
    case 3:
      hipLaunchKernelGGL(site_loop<3>, dim3(p.G), dim3(kLoopThreads), p.lds, s, a);
      break;
    case 5:
      hipLaunchKernelGGL(site_loop<5>, dim3(p.G), dim3(kLoopThreads), p.lds, s, a);
      break;
    case 8:
      hipLaunchKernelGGL(site_loop<8>, dim3(p.G), dim3(kLoopThreads), p.lds, s, a);
      break;
    default:
      hipLaunchKernelGGL(site_loop<0>, dim3(p.G), dim3(kLoopThreads), p.lds, s, a);
      break;
  }
  SITING_HIP(hipGetLastError());
  return SITING_OK;
}

}  // namespace siting
