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
//   site_loop        persistent kernel over G workgroups.  A workgroup keeps
//                    its candidates' ids, positions, areas, used flags and a
//                    seen-this-round bitmap in LDS (or, when they do not fit,
//                    in a global slice of the same layout) and the two best
//                    keys of each of its units.  A round:
//                    1. every workgroup stages all published unit keys and
//                       walks, identically, down the global key order (at
//                       most 32 picks).  A key more than 2*roi from every
//                       earlier pick of the round is exact and is taken.
//                       A nearer key is re-keyed to its exact area after
//                       those picks (computed by the whole workgroup; under
//                       intervis, 0 while its own cell is uncovered) and
//                       goes back into the order, as the sequential loop
//                       recomputed it after each pick.  Intervis 1: a masked
//                       candidate may gain area when a pick uncovers its
//                       cell, so a key is taken only if it beats every such
//                       candidate's bound (per-cell records of the masked
//                       bounds, the largest computed exactly on demand).  A
//                       walk that read cum is followed by a grid barrier
//                       before step 3 writes it.
//                    2. counters advance pick by pick; the 95% rule may end
//                       the round after any pick.
//                    3. the picks' sheds, shifted to terrain words, are ORed
//                       into cum with atomicOr; the bits each OR newly set are
//                       kept (nb), then one grid barrier.
//                    4. every workgroup walks the candidates of its units that
//                       lie within 2*roi of a pick (cell ranges), each once,
//                       and lowers its area by the pick bits it loses
//                       (intervis 0; one below a threshold keeps its old area
//                       as a stale bound, re-evaluated once the round's head
//                       falls to it), or recomputes it against cum | delta
//                       (intervis 1; one pick needs no barrier in step 3).
//                    5. dirty units re-rank their top two, publish them into
//                       the other half of a double-buffered key array, grid
//                       barrier.
//   site_loop_queue  the same body for many jobs: slots of workgroups pop
//                    job indices from a queue and run one job each.
// Keys: (area << 32) | ~index, 0 for a used candidate, so the maximum key is
// the largest area with the lowest index; a maximum with area 0 ends the
// loop.  Each key travels with its candidate's position.
//
// Barriers are plain atomics on a counter, so every workgroup of a loop must
// be resident at once: the host caps G at half the CUs (one workgroup per
// CU), and the work-queue launch sizes its slots the same way.
#include <algorithm>
#include <cstdlib>

#include "common.hpp"

namespace siting {

// Phase profile of the greedy loop (the profiling build only, compiled with
// SITING_PHASE_PROFILE; the product build has none of it).  Thread 0 of every
// loop workgroup adds the device-clock time (100 MHz) of each phase of each
// round; the walking wavefront adds the time spent in its exact area
// computations, its steps, and why each round's walk ended.  Counters
// (g_site_prof, summed over workgroups and rounds):
//   0 pick walk (after the key gather), 1 bookkeeping and the walk barrier,
//   2 union, 3 update lists, 4 re-evaluation, 5 top-2 and publish,
//   6 end-of-round barrier, 7 key gather, 8 loop workgroups,
//   9 exact computations, 10 their time, 11 walk steps, 12 picks,
//   13-19 walk ends: computation budget, rekeyed set full, bound entry,
//   stale key, no area left, pick cap, round without a pick;
//   20 the merge and sort of the walk's entries (the walk, 0, follows it).
// The profiling build also traces the first kTraceRounds rounds of a loop
// of up to kTraceGroups workgroups: thread 0 of each workgroup stores the
// device clock at the round's start, after the key gather, after the walk,
// after the publish, after the bookkeeping, the union, the update lists and
// the re-evaluation (g_site_trace[mark][round][group], marks 0-7 in that
// order), which shows how far the workgroups drift apart within a round.
#ifdef SITING_PHASE_PROFILE
constexpr int kProfWords = 32;
constexpr int kTraceMarks = 8, kTraceRounds = 256, kTraceGroups = 128;
__device__ unsigned long long g_site_prof[kProfWords];
__device__ unsigned long long g_site_trace[kTraceMarks][kTraceRounds][kTraceGroups];
#define SITE_PROF(stmt) stmt
#else
#define SITE_PROF(stmt)
#endif

namespace {

constexpr int kScanThreads = 1024;  // the bucketing scan kernels
// The loop's workgroups are 1024 threads (four wavefronts per SIMD, so at
// most 128 VGPRs) with up to 132 KiB of dynamic LDS: one per CU.  (512-thread
// groups that leave room for the other stages were measured 6-35 % slower.)
constexpr int kLoopThreads = 1024;
constexpr int kLoopWaves = kLoopThreads / kWave;
constexpr int kMaxList = kLoopThreads;  // list windows: one item per thread
constexpr size_t kLdsBudget = 132 * 1024;  // dynamic LDS of the loop kernel
constexpr int kMaxGroups = 128;
// Candidates one loop workgroup keeps in LDS comfortably (the budget holds
// about 9000 at 14 bytes each).
constexpr int kFitPerGroup = 8000;
constexpr int kMaxBatch = kSiteMaxBatch;
// Wavefronts that gather the published unit keys for the pick walk, and the
// position bit (bit 15: y < 32768) that flags a key whose successor in its
// unit or lane list is unknown.
constexpr int kPickWaves = 4;
// Dirty cells a workgroup lists per round (intervis multi-pick); beyond
// this it republishes every cell it owns.
constexpr int kMaxDirtyCells = 1024;
// Candidates the pick walk may re-key in a round (intervis 0).
constexpr int kMaxRekey = 8;
// Chunks of 64 cells the intervis walk's window scan reads per round trip
#ifndef SITING_WIN_CHUNKS
#define SITING_WIN_CHUNKS 4
#endif
constexpr int kWinChunks = SITING_WIN_CHUNKS;
static_assert(kMaxBatch <= kWave, "lane j of the walk holds pick j");
// Position of a bound entry in a lane's key list: a key that stands for
// keys the lane does not hold (coordinates are below 32767, so no real
// position).
constexpr uint32_t kGhost = 0xffffffffu;


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

// inclusive scan of one value per thread over a kThreads-thread block
template <int kThreads>
__device__ int block_inclusive_scan(int v, int *wsum) {
  constexpr int kWaves = kThreads / kWave;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int o = __shfl_up(v, off, kWave);
    if (lane >= off) v += o;
  }
  if (lane == 63) wsum[wid] = v;
  __syncthreads();
  if (wid == 0) {
    int s = lane < kWaves ? wsum[lane] : 0;
#pragma unroll
    for (int off = 1; off < kWaves; off <<= 1) {
      int o = __shfl_up(s, off, kWave);
      if (lane >= off) s += o;
    }
    if (lane < kWaves) wsum[lane] = s;
  }
  __syncthreads();
  if (wid > 0) v += wsum[wid - 1];
  return v;
}

__global__ __launch_bounds__(kScanThreads) void site_scan_chunks(SiteArgs a) {
  __shared__ int red[kScanThreads / kWave];
  const int i = blockIdx.x * kScanThreads + threadIdx.x;
  int v = i < a.nbuckets ? a.bcount[i] : 0;
  v = wave_sum_i32(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < kScanThreads / kWave; w++) s += red[w];
    a.chunk[blockIdx.x] = s;
  }
}

// bstart = exclusive scan of bcount; bcount becomes the scatter cursor
__global__ __launch_bounds__(kScanThreads) void site_scan_apply(SiteArgs a) {
  __shared__ int wsum[kScanThreads / kWave];
  __shared__ int prefix;
  int p = 0;
  for (int b = threadIdx.x; b < (int)blockIdx.x; b += kScanThreads) p += a.chunk[b];
  p = wave_sum_i32(p);
  if (threadIdx.x == 0) prefix = 0;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) atomicAdd(&prefix, p);
  __syncthreads();
  const int i = blockIdx.x * kScanThreads + threadIdx.x;
  const int v = i < a.nbuckets ? a.bcount[i] : 0;
  const int inc = block_inclusive_scan<kScanThreads>(v, wsum);
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
  const size_t ncum = cum_words(a.n, a.cw);
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < ncum;
       i += (size_t)gridDim.x * 256)
    a.cum[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.st->cumarea = 0;
    a.st->sum_u = (unsigned long long)a.nsheds;
    a.st->nused = 0;
    a.st->niter = 1;
    a.st->rounds = 0;
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

// Grid barrier without fences.
// Data the workgroups of a loop exchange (the published unit keys, the bits
// each pick newly covered, the union cum) is only ever read and written with
// device-coherent accesses: relaxed agent-scope atomic loads and stores
// (sc1 on gfx950; cum takes device atomics), which need no cache
// write-back or invalidation.  The grid barrier therefore only has to wait
// for each wavefront's stores to complete before arriving, and uses no
// agent-scope fence (on gfx950 those write back and invalidate the whole of
// the XCD's L2, for every workgroup of every loop running).  A relaxed
// barrier is sound only because no exchanged word is ever read with a plain
// load: a plain load may hit a stale line in this CU's L1 or another XCD's
// L2, and workgroups that read different keys pick differently.  (An early
// version read the published keys with plain loads under a relaxed barrier
// and hung: the workgroups' picks diverged.)  The loop also compares a hash
// of every workgroup's picks each round, so a divergence that kept the
// barrier counts equal would end the loop with an error instead of a wrong
// selection.
template <class T>
__device__ __forceinline__ T ld_dev(const T *p) {
  return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_dev(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Polls of the barrier counter before a workgroup gives the loop up (about
// a second): a barrier that never completes ends the loop with niter = -1
// instead of hanging the device.
constexpr unsigned int kSpinLimit = 1u << 20;

// Grid barrier over G resident workgroups, in two halves so that work that
// needs nothing from the other workgroups can run in between: arrive once
// this workgroup's stores are done, wait for every arrival (false if the
// wait timed out; a limit of 0 gives up at once).  target is thread 0's.
__device__ __forceinline__ void grid_arrive(unsigned int *ctr, unsigned int &target,
                                            unsigned int G) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wavefront's stores are done
  __syncthreads();
  if (threadIdx.x == 0) {
    target += G;
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__device__ __forceinline__ bool grid_wait(unsigned int *ctr, unsigned int target,
                                          unsigned int limit) {
  __shared__ int timed_out;
  if (threadIdx.x == 0) {
    unsigned int spins = 0;
    while (limit != 0 && (int)(ld_dev(ctr) - target) < 0 && ++spins < limit)
      __builtin_amdgcn_s_sleep(1);
    timed_out = spins >= limit;
  }
  __syncthreads();
  return !timed_out;
}

__device__ __forceinline__ bool grid_sync(unsigned int *ctr, unsigned int &target,
                                          unsigned int G, unsigned int limit = kSpinLimit) {
  grid_arrive(ctr, target, G);
  return grid_wait(ctr, target, limit);
}

// Shifted word dk (0..nwpr) of row dr of the winner's shed, i.e. the bits of
// terrain word w0w+dk that the winner covers (siting.cpp:408-417).
// Both loads are unconditional (from clamped slots), so neither is waited
// for inside a branch.
__device__ __forceinline__ uint64_t winner_word(const uint64_t *sw, int nwpr,
                                                int fbw, int dr, int dk) {
  const uint64_t *row = sw + (size_t)dr * nwpr;
  const uint64_t cur = row[dk < nwpr ? dk : nwpr - 1];
  const uint64_t prv = row[dk > 0 ? dk - 1 : 0];
  uint64_t v = dk < nwpr ? cur >> fbw : 0;
  if (fbw != 0 && dk > 0) v |= prv << (64 - fbw);
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

// New bits of candidate (cx, cy)'s shed against cum | winner
// (siting.cpp:456-491): this lane's share.  Work items are (shed row,
// terrain word) pairs inside the terrain; lane gl of a group of gs lanes
// takes items gl, gl+gs, ..., so a load instruction reads consecutive words
// of the shed (rows are contiguous).  kBatch items per lane are loaded
// before any is used.
constexpr int kBatch = 4;
// items per lane in flight when re-evaluating (8 spills and measured slower)
constexpr int kEvalBatch = 4;
// rows per lane in flight in the row-dealt re-evaluation
#ifndef SITING_EVAL_ROWS
#define SITING_EVAL_ROWS 4
#endif
constexpr int kEvalRows = SITING_EVAL_ROWS;

template <int kB = kBatch>
__device__ __forceinline__ int lane_new_bits(const uint64_t *sh, int cx, int cy,
                                             const SiteArgs &a, const Winner &W,
                                             int gl, int gs, bool use_w) {
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
  for (int base = 0; base < items; base += gs * kB) {
    uint64_t cur[kB], prv[kB], cv[kB];
#pragma unroll
    for (int u = 0; u < kB; u++) {
      const int it = base + u * gs + gl;
      const bool ok = it < items;
      const int q = fdiv ? (int)(((float)it + 0.5f) * inv_nk) : it / nk;
      const int row = rlo + q, k = klo + (it - q * nk);
      const uint64_t *srow = sh + (size_t)row * nwpr;
      cur[u] = (ok && k < nwpr) ? srow[k] : 0;
      prv[u] = (ok && k > 0) ? srow[k - 1] : 0;
      // unconditional: a load under a branch is waited for inside it, which
      // would serialise the batch (a lane past the end has no shed bits, so
      // the word it reads does not matter)
      uint64_t c = ld_dev(ok ? &a.cum[cum_index(xo + row, w0c + k, cw)] : a.cum);
      if (use_w) {
        const int dr = dro + row, dk = w0c + k - W.w0w;
        const bool in = ok && dr >= 0 && dr < W.nr && dk >= 0 && dk <= W.nwpr;
        const uint64_t wv = W.word(in ? dr : 0, in ? dk : 0);
        c |= in ? wv : 0;
      }
      cv[u] = c;
    }
#pragma unroll
    for (int u = 0; u < kB; u++) {
      uint64_t moved = cur[u] >> fbc;
      if (fbc != 0) moved |= prv[u] << (64 - fbc);
      s += __popcll(moved & ~cv[u]);
    }
  }
  return s;
}

// The same count (against cum alone) with rows dealt to lanes: lane l of
// each wavefront of a group of gs lanes takes word slot l mod nk of row
// l / nk + (64 / nk) * (its wavefront in the group), stepping by the
// group's rows, so its addresses only advance by a stride; the word before
// its slot comes from the lane before it.  For boxes whose first word slot
// is inside the terrain and whose rows have at most 64 slots
// (eval_rows_fit).
__device__ __forceinline__ bool eval_rows_fit(int cy, int roi, int nwpr) {
  return cy - roi >= 0 && nwpr + 1 <= kWave;
}

template <int kB>
__device__ __forceinline__ int lane_new_bits_rows(const uint64_t *sh, int cx, int cy,
                                                  const SiteArgs &a, int gl, int gs) {
  const int roi = a.roi, nwpr = a.nwpr, cw = a.cw, n = a.n;
  int w0c, fbc;
  col_split(cy - roi, w0c, fbc);
  const int nk = min((cy + roi) >> 6, cw - 1) - w0c + 1;  // word slots 0 .. nk-1
  const int lane = gl & (kWave - 1), wi = gl >> 6;
  const int rpw = kWave / nk;                             // rows per wavefront and step
  const int mk = lane % nk, mr = lane / nk;
  const bool act = mr < rpw;
  const int rstep = (gs >> 6) * rpw;
  const int rlo = max(0, cx - roi) - (cx - roi);
  const int rhi = min(n - 1, cx + roi) - (cx - roi);
  const int r0 = rlo + wi * rpw + mr;
  const uint64_t *sp = sh + (size_t)r0 * nwpr + mk;
  const size_t ss = (size_t)rstep * nwpr;
  const int xo = cx - roi, jc = w0c + mk;
  int s = 0;
  for (int r = r0; r - r0 + rlo <= rhi; r += kB * rstep) {
    uint64_t cur[kB], cv[kB];
#pragma unroll
    for (int u = 0; u < kB; u++) {
      const bool ok = act && r + u * rstep <= rhi;
      cur[u] = (ok && mk < nwpr) ? sp[u * ss] : 0;
      // unconditional (lane_new_bits)
      cv[u] = ld_dev(ok ? &a.cum[cum_index(xo + r + u * rstep, jc, cw)] : a.cum);
    }
#pragma unroll
    for (int u = 0; u < kB; u++) {
      uint64_t moved = cur[u];
      if (fbc != 0) {
        // every lane takes part in the shuffle; a row's first slot has none
        const uint64_t up = (uint64_t)__shfl_up((unsigned long long)cur[u], 1, kWave);
        const uint64_t prv = mk > 0 ? up : 0;
        moved = (cur[u] >> fbc) | (prv << (64 - fbc));
      }
      s += __popcll(moved & ~cv[u]);
    }
    sp += kB * ss;
  }
  return s;
}

__device__ __forceinline__ int wave_new_bits(const uint64_t *sh, int cx, int cy,
                                             const SiteArgs &a, const Winner &W,
                                             int lane, bool use_w) {
  return wave_sum_i32(lane_new_bits(sh, cx, cy, a, W, lane, kWave, use_w));
}

// Bits of candidate (cx, cy)'s shed that this round's picks newly covered,
// i.e. the drop of its area, summed over the wavefront.  Only picks within
// 2*roi share bits with it.  Items are (row, terrain word) pairs of the
// overlap of the two boxes; a word the pick did not change is skipped
// without loading the candidate's shed.
__device__ __forceinline__ int wave_lost_bits(const uint64_t *sh, int cx, int cy,
                                              const SiteArgs &a, int k,
                                              const uint32_t *bw_pos, int lane) {
  const int roi = a.roi, nr = a.nr, nwpr = a.nwpr, cw = a.cw, n = a.n;
  const int dw = nwpr + 1;
  const int r2 = 4 * roi * roi;
  int w0c, fbc;
  col_split(cy - roi, w0c, fbc);
  const int wlc = (cy + roi) >> 6;
  int s = 0;
  for (int p = 0; p < k; p++) {
    const int wx = (int)(bw_pos[p] >> 16), wy = (int)(bw_pos[p] & 0xffffu);
    const int ex = cx - wx, ey = cy - wy;
    if (ex * ex + ey * ey > r2) continue;
    int w0p, fbp;
    col_split(wy - roi, w0p, fbp);
    const int wlp = (wy + roi) >> 6;
    const int xlo = max(max(cx, wx) - roi, 0), xhi = min(min(cx, wx) + roi, n - 1);
    const int jlo = max(max(w0c, w0p), 0), jhi = min(min(wlc, wlp), cw - 1);
    if (xlo > xhi || jlo > jhi) continue;
    const int nk = jhi - jlo + 1;
    const int items = (xhi - xlo + 1) * nk;  // <= nr * (nwpr + 1)
    const float inv_nk = 1.0f / (float)nk;
    const bool fdiv = items < (1 << 21);
    const uint64_t *nbp = a.nb + (size_t)p * nr * dw;
    for (int base = 0; base < items; base += kWave * kBatch) {
      uint64_t nbv[kBatch], cur[kBatch], prv[kBatch];
      int soff[kBatch], kk[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; u++) {
        const int it = base + u * kWave + lane;
        const bool ok = it < items;
        const int q = fdiv ? (int)(((float)it + 0.5f) * inv_nk) : it / nk;
        const int x = xlo + q, j = jlo + (it - q * nk);
        // unconditional (lane_new_bits); a lane past the end keeps nothing
        const uint64_t v = ld_dev(ok ? &nbp[(size_t)(x - wx + roi) * dw + (j - w0p)] : nbp);
        nbv[u] = ok ? v : 0;
        soff[u] = ok ? (x - cx + roi) * nwpr : 0;
        kk[u] = j - w0c;
      }
#pragma unroll
      for (int u = 0; u < kBatch; u++) {
        const bool live = nbv[u] != 0;
        cur[u] = (live && kk[u] < nwpr) ? sh[soff[u] + kk[u]] : 0;
        prv[u] = (live && kk[u] > 0) ? sh[soff[u] + kk[u] - 1] : 0;
      }
#pragma unroll
      for (int u = 0; u < kBatch; u++) {
        uint64_t moved = cur[u] >> fbc;
        if (fbc != 0) moved |= prv[u] << (64 - fbc);
        s += __popcll(moved & nbv[u]);
      }
    }
  }
  return wave_sum_i32(s);
}

// New bits of candidate (cx, cy)'s shed against cum | the sheds of np
// picks (positions ppos[], shed indices pid[], in LDS), i.e. its area once
// they are taken (intervis 0) or, for one pick, its area bound after that
// pick (intervis 1): lane gl of gs lanes' share.  Items are (shed row,
// terrain word) pairs as in lane_new_bits, kIt per lane in flight; the
// picks' words are read kPG picks at a time, and every load of a group is
// issued before any is used (one round trip per group).
template <int kIt, int kPG>
__device__ __forceinline__ int lane_new_bits_picks(const uint64_t *csh, int cx, int cy, int np,
                                                   const uint32_t *ppos, const uint32_t *pid,
                                                   const SiteArgs &a, int gl, int gs) {
  const int roi = a.roi, nwpr = a.nwpr, cw = a.cw, n = a.n, nr = 2 * roi + 1;
  const size_t S = (size_t)nr * nwpr;
  int w0c, fbc;
  col_split(cy - roi, w0c, fbc);
  const int klo = max(w0c, 0) - w0c;
  const int khi = min((cy + roi) >> 6, cw - 1) - w0c;
  const int nk = khi - klo + 1;
  const int rlo = max(0, cx - roi) - (cx - roi);
  const int rhi = min(n - 1, cx + roi) - (cx - roi);
  const int items = (rhi - rlo + 1) * nk;
  const float inv_nk = 1.0f / (float)nk;
  const bool fdiv = items < (1 << 21);
  int s = 0;
  for (int base = 0; base < items; base += gs * kIt) {
    uint64_t cur[kIt], prv[kIt], mv[kIt];
    int xr[kIt], jw[kIt];
#pragma unroll
    for (int u = 0; u < kIt; u++) {
      const int it = base + u * gs + gl;
      const bool ok = it < items;
      const int q = fdiv ? (int)(((float)it + 0.5f) * inv_nk) : it / nk;
      const int row = rlo + q, k = klo + (it - q * nk);
      const uint64_t *srow = csh + (size_t)row * nwpr;
      cur[u] = (ok && k < nwpr) ? srow[k] : 0;
      prv[u] = (ok && k > 0) ? srow[k - 1] : 0;
      xr[u] = cx - roi + row;
      jw[u] = w0c + k;
      // unconditional loads, so that none waits on another (lane_new_bits)
      mv[u] = ld_dev(ok ? &a.cum[cum_index(xr[u], jw[u], cw)] : a.cum);
    }
    for (int j0 = 0; j0 < np; j0 += kPG) {
      uint64_t pc[kIt][kPG], pv[kIt][kPG];
      int w0p[kPG], fbp[kPG], pxs[kPG];
#pragma unroll
      for (int t = 0; t < kPG; t++) {
        const int j = j0 + t < np ? j0 + t : np - 1;  // a missing pick repeats one
        const uint32_t pq = ppos[j];
        pxs[t] = (int)(pq >> 16);
        col_split((int)(pq & 0xffffu) - roi, w0p[t], fbp[t]);
        const uint64_t *psh = a.sheds + (size_t)pid[j] * S;
#pragma unroll
        for (int u = 0; u < kIt; u++) {
          const int dr = xr[u] - pxs[t] + roi, dk = jw[u] - w0p[t];
          const bool in = dr >= 0 && dr < nr && dk >= 0 && dk <= nwpr;
          const uint64_t *prow = psh + (size_t)(in ? dr : 0) * nwpr;
          const int kc = in ? dk : 0;
          pc[u][t] = prow[kc < nwpr ? kc : nwpr - 1];
          pv[u][t] = prow[kc > 0 ? kc - 1 : 0];
        }
      }
#pragma unroll
      for (int t = 0; t < kPG; t++) {
#pragma unroll
        for (int u = 0; u < kIt; u++) {
          const int dr = xr[u] - pxs[t] + roi, dk = jw[u] - w0p[t];
          const bool in = j0 + t < np && dr >= 0 && dr < nr && dk >= 0 && dk <= nwpr;
          uint64_t pw = dk < nwpr ? pc[u][t] >> fbp[t] : 0;
          if (fbp[t] != 0 && dk > 0) pw |= pv[u][t] << (64 - fbp[t]);
          mv[u] |= in ? pw : 0;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kIt; u++) {
      uint64_t c = cur[u] >> fbc;
      if (fbc != 0) c |= prv[u] << (64 - fbc);
      s += __popcll(c & ~mv[u]);
    }
  }
  return s;
}

// lane_new_bits_picks with the picks' words read in groups of as many picks
// as there are, up to four (most of the walk's computations have one or two
// picks near the candidate; a group of four for one pick loads and shifts
// three repeats: the 4096^2 job's computations 2.7 -> 1.6 us).
__device__ __forceinline__ int lane_new_bits_picks_np(const uint64_t *csh, int cx, int cy, int np,
                                                      const uint32_t *ppos, const uint32_t *pid,
                                                      const SiteArgs &a, int gl, int gs) {
  if (np <= 1) return lane_new_bits_picks<1, 1>(csh, cx, cy, np, ppos, pid, a, gl, gs);
  if (np == 2) return lane_new_bits_picks<1, 2>(csh, cx, cy, np, ppos, pid, a, gl, gs);
  if (np == 3) return lane_new_bits_picks<1, 3>(csh, cx, cy, np, ppos, pid, a, gl, gs);
  return lane_new_bits_picks<1, 4>(csh, cx, cy, np, ppos, pid, a, gl, gs);
}

// Maximum of a 64-bit value over the wavefront, in every lane, with DPP row
// operations (no LDS traffic): quad swaps, half-row and row mirrors, then the
// row broadcasts that fold rows 0-3 into lane 63.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ uint64_t dpp_max_step(uint64_t v) {
  const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  const uint32_t olo = (uint32_t)__builtin_amdgcn_update_dpp((int)lo, (int)lo, kCtrl,
                                                             kRowMask, 0xf, false);
  const uint32_t ohi = (uint32_t)__builtin_amdgcn_update_dpp((int)hi, (int)hi, kCtrl,
                                                             kRowMask, 0xf, false);
  const uint64_t o = ((uint64_t)ohi << 32) | olo;
  return o > v ? o : v;
}

__device__ __forceinline__ uint64_t wave_max_u64_dpp(uint64_t v) {
  v = dpp_max_step<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v = dpp_max_step<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v = dpp_max_step<0x141, 0xf>(v);  // row_half_mirror
  v = dpp_max_step<0x140, 0xf>(v);  // row_mirror
  v = dpp_max_step<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
  v = dpp_max_step<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
  return ((uint64_t)hi << 32) | lo;
}

// The three largest (key, position) pairs pushed, in descending order; keys
// are distinct and 0 marks an empty slot.
struct Top3 {
  uint64_t k[3];
  uint32_t p[3];
  uint64_t over;  // the largest key pushed and not kept
};

__device__ __forceinline__ void top3_push(Top3 &t, uint64_t k, uint32_t p) {
  const uint64_t out = k > t.k[2] ? t.k[2] : k;  // the one that falls off
  t.over = out > t.over ? out : t.over;
  if (k > t.k[2]) {
    if (k > t.k[1]) {
      t.k[2] = t.k[1];
      t.p[2] = t.p[1];
      if (k > t.k[0]) {
        t.k[1] = t.k[0];
        t.p[1] = t.p[0];
        t.k[0] = k;
        t.p[0] = p;
      } else {
        t.k[1] = k;
        t.p[1] = p;
      }
    } else {
      t.k[2] = k;
      t.p[2] = p;
    }
  }
}

// Best and second best (key, position) and the third best key over the
// unused candidates of owned unit ul; one wavefront.
struct Top2 {
  uint64_t k1, k2, k3;
  uint32_t p1, p2;
};

__device__ __forceinline__ void top2_push(Top2 &t, uint64_t k, uint32_t p) {
  if (k > t.k1) {
    t.k3 = t.k2;
    t.k2 = t.k1;
    t.p2 = t.p1;
    t.k1 = k;
    t.p1 = p;
  } else if (k > t.k2) {
    t.k3 = t.k2;
    t.k2 = k;
    t.p2 = p;
  } else if (k > t.k3) {
    t.k3 = k;
  }
}

__device__ __forceinline__ Top2 wave_top2(Top2 t) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    Top2 u;
    u.k1 = __shfl_xor(t.k1, o, kWave);
    u.k2 = __shfl_xor(t.k2, o, kWave);
    u.k3 = __shfl_xor(t.k3, o, kWave);
    u.p1 = __shfl_xor(t.p1, o, kWave);
    u.p2 = __shfl_xor(t.p2, o, kWave);
    top2_push(t, u.k1, u.p1);
    top2_push(t, u.k2, u.p2);
    top2_push(t, u.k3, 0);
  }
  return t;
}

// The greedy loop of one job, run by workgroups g = 0..G-1 (all resident).
// kIvMode: 0 or 1, the job's intervis flag fixed at compile time, so that an
// intervis-0 loop carries none of the intervis code and its registers; 2,
// read from a.intervis (a work queue whose jobs differ in it).
template <int kIvMode>
__device__ __forceinline__ void site_loop_body(const SiteArgs &a, int g, int G) {
  const bool kIv = kIvMode == 2 ? a.intervis != 0 : kIvMode == 1;
  extern __shared__ __align__(16) uint8_t smem[];
#ifdef SITING_PHASE_PROFILE
  __shared__ unsigned long long pf_acc[kProfWords], pf_last;
#define SITE_PROF_MARK(k)                                  \
  if (threadIdx.x == 0) {                                  \
    const unsigned long long t_ = wall_clock64();          \
    pf_acc[k] += t_ - pf_last;                             \
    pf_last = t_;                                          \
  }
#define SITE_PROF_ADD(k, v) \
  if ((threadIdx.x & 63) == 0) pf_acc[k] += (unsigned long long)(v);
#define SITE_TRACE(m)                                              \
  if (threadIdx.x == 0 && pf_r < kTraceRounds && g < kTraceGroups) \
    g_site_trace[m][pf_r][g] = wall_clock64();
#else
#define SITE_TRACE(m)
#define SITE_PROF_MARK(k)
#define SITE_PROF_ADD(k, v)
#endif
  __shared__ int list_i[kMaxList];
  __shared__ int list_acc[kMaxList];
  __shared__ int list_n;
  __shared__ int rng_lo[kMaxList];
  __shared__ int rng_end[kMaxList];
  __shared__ int rng_n;
  __shared__ int wsum[kLoopWaves];
  __shared__ uint64_t bw_key[kMaxBatch];
  __shared__ uint32_t bw_pos[kMaxBatch];
  __shared__ int bw_n;
  __shared__ int bw_head;
  __shared__ uint64_t r_key[kMaxRekey];  // the walk's rekeyed candidates
  __shared__ uint32_t r_pos[kMaxRekey];
  __shared__ int r_mark[kMaxRekey];
  __shared__ int bw_npair[kWave];  // earlier picks of the round within 2*roi of pick i
  __shared__ int bw_readcum;       // the walk read cum
  // The walk's area computations are done by the rest of the workgroup:
  // the walking wavefront posts a request (candidate, picks) and waits, the
  // others, waiting in a service loop, each take a share and add it to
  // mb_acc with their arrival (arrivals << 40 | sum).  mb_seq numbers the
  // requests (-1: the walk is over).
  __shared__ int mb_seq;
  __shared__ unsigned long long mb_acc;
  __shared__ uint32_t mb_cid, mb_cpos;
  __shared__ int mb_np;
  __shared__ uint32_t mb_ppos[kMaxBatch], mb_pid[kMaxBatch];
  __shared__ uint32_t ev_pos[kMaxBatch], ev_id[kMaxBatch];  // intervis walk: listed
  __shared__ int ev_bd[kMaxBatch], ev_pick[kMaxBatch], ev_exact[kMaxBatch];
  __shared__ uint64_t xk[3 * (kPickWaves - 1)][kWave];  // pick lists of wavefronts 1..
  __shared__ uint32_t xp[3 * (kPickWaves - 1)][kWave];
  __shared__ uint64_t xo[kPickWaves - 1][kWave];  // their largest keys not kept

  const int tid = threadIdx.x;
  // (the wavefront index by readfirstlane: wave-uniform, so that branches
  // on it are scalar branches)
  const int lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int n = a.n, roi = a.roi, nr = a.nr, nwpr = a.nwpr, cw = a.cw;
  const int Q = a.Q, upw = a.upw, kpu = a.kpu, nunits = a.nunits;
  const int r2 = (2 * roi) * (2 * roi);
  const size_t S = (size_t)nr * nwpr;
  const int dw = nwpr + 1;
  // intervis rounds of several picks (else one pick per round)
  const bool ivm = kIv && a.iv_multi;
  const int kmax = kIv && !ivm ? 1 : kMaxBatch;
  const bool hybrid = !kIv && a.tau_shift > 0;

  uint64_t *delta = (uint64_t *)(smem + a.o_delta);
  uint64_t *ukey = (uint64_t *)(smem + a.o_ukey);   // [upw][3]
  uint32_t *upos = (uint32_t *)(smem + a.o_upos);   // [upw][2]
  uint8_t *dirty = smem + a.o_dirty;                 // [upw]
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
  // seen-this-round bits over the owned candidates (global mode: a slice of
  // a global bitmap, one word per 32 candidates from this group's start)
  unsigned int *seen = in_lds ? (unsigned int *)(smem + a.o_seen)
                              : a.sseen + (size_t)g * ((a.nsheds + 31) / 32 + 1);
  // intervis multi-pick: Bd[i] bounds candidate i's area from above however
  // the visibility mask turns (its new bits against cum, unmasked, when it
  // was last computed: cum only grows), and the cells whose bounds changed
  // this round and the last are listed for republishing
  int32_t *Bd = !ivm ? nullptr : in_lds ? (int32_t *)(smem + a.o_b) : a.sB + B0;
  int *dcur = ivm ? (int *)(smem + a.o_dcur) : nullptr;
  int *dprev = ivm ? (int *)(smem + a.o_dprev) : nullptr;
  unsigned int *dbits = ivm ? (unsigned int *)(smem + a.o_dbits) : nullptr;
  __shared__ int dcur_n, dprev_n;
  const int ncl = upw * kpu;  // owned cells, local index ul * kpu + k
  for (int i = tid; i < cnt; i += kLoopThreads) {
    const int c = a.sid[B0 + i];
    if (in_lds) {
      ((int32_t *)(smem + a.o_c))[i] = c;
      ((uint32_t *)(smem + a.o_pos))[i] = a.spos[B0 + i];
    }
    T[i] = a.area[c];
    U[i] = 0;
    if (ivm) Bd[i] = a.area[c];
  }
  if (ivm) {
    for (int i = tid; i < (ncl + 31) / 32; i += kLoopThreads) dbits[i] = 0;
    if (tid == 0) {
      dcur_n = 0;
      dprev_n = 0;
    }
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
  for (int ul = tid; ul < upw; ul += kLoopThreads) dirty[ul] = 0;
  for (int i = tid; i < (cnt + 31) / 32; i += kLoopThreads) seen[i] = 0;
  __syncthreads();

  // U[i]: bit 0 used, bit 1 stale (T[i] is then an upper bound of the area,
  // not the area).  A published position carries the stale bit in bit 31
  // (positions are below 2^31).
  auto unit_top2 = [&](int ul) {
    const int lo = OFF(ul, 0), hi = OFF(ul, kpu);
    Top2 t{0, 0, 0, 0, 0};
    // two candidates per lane in flight, every load unconditional (a load
    // under a branch is waited for inside it); a lane past the end re-reads
    // the unit's first candidate and pushes nothing (key 0)
    for (int i0 = lo + lane; i0 < hi; i0 += 2 * kWave) {
      const int i1 = i0 + kWave;
      const int j1 = i1 < hi ? i1 : lo;
      const uint8_t u0 = U[i0], u1 = U[j1];
      const int t0 = T[i0], t1 = T[j1];
      const int c0 = C[i0], c1 = C[j1];
      const uint32_t p0 = P[i0], p1 = P[j1];
      top2_push(t, (u0 & 1) ? 0 : cand_key(t0, c0), p0 | ((u0 & 2) ? 0x80000000u : 0u));
      top2_push(t, (i1 >= hi || (u1 & 1)) ? 0 : cand_key(t1, c1),
                p1 | ((u1 & 2) ? 0x80000000u : 0u));
    }
    t = wave_top2(t);
    if (lane == 0) {
      ukey[3 * ul] = t.k1;
      ukey[3 * ul + 1] = t.k2;
      ukey[3 * ul + 2] = t.k3;
      upos[2 * ul] = t.p1;
      upos[2 * ul + 1] = t.p2;
    }
  };
  // intervis multi-pick: the bound record of owned cell lc in buffer buf,
  // over the cell's unused candidates whose last area was masked (U bit 2;
  // only those can gain area, when a pick uncovers their cell -- the
  // others' areas only shrink): the two largest bounds Bd with their
  // candidates' positions and ids, and the third largest bound --
  // (b1 << 32 | pos1, b2 << 32 | id1, b3 << 32 | pos2, id2)
  auto cell_record = [&](int lc, int buf) {
    const int ul = lc / kpu, k = lc - ul * kpu;
    const int u = ul * G + g;
    if (u >= nunits) return;
    const int r = u / Q, cc = (u - r * Q) + k * Q;
    if (cc >= a.ncr) return;
    const int lo = OFF(ul, k), hi = OFF(ul, k + 1);
    uint32_t b1 = 0, b2 = 0, b3 = 0, p1 = 0, p2 = 0, id1 = 0, id2 = 0;
    for (int i = lo; i < hi; i++) {
      if ((U[i] & 5) != 4) continue;
      const uint32_t b = (uint32_t)Bd[i];
      if (b > b1) {
        b3 = b2;
        b2 = b1;
        p2 = p1;
        id2 = id1;
        b1 = b;
        p1 = P[i];
        id1 = (uint32_t)C[i];
      } else if (b > b2) {
        b3 = b2;
        b2 = b;
        p2 = P[i];
        id2 = (uint32_t)C[i];
      } else if (b > b3) {
        b3 = b;
      }
    }
    uint64_t *rec = a.cbnd + ((size_t)buf * a.ncr * a.ncr + (size_t)r * a.ncr + cc) * 4;
    st_dev(&rec[0], ((uint64_t)b1 << 32) | p1);
    st_dev(&rec[1], ((uint64_t)b2 << 32) | id1);
    st_dev(&rec[2], ((uint64_t)b3 << 32) | p2);
    st_dev(&rec[3], (uint64_t)id2);
  };
  // lists owned candidate (cx, cy)'s cell as changed this round
  auto mark_cell = [&](int cx, int cy) {
    const int r = cx / a.cs, cc = cy / a.cs;
    const int u = r * Q + cc % Q;
    const int lc = (u / G) * kpu + cc / Q;
    const unsigned int bit = 1u << (lc & 31);
    if (!(atomicOr(&dbits[lc >> 5], bit) & bit)) {
      const int at = atomicAdd(&dcur_n, 1);
      if (at < kMaxDirtyCells) dcur[at] = lc;
    }
  };
  // republishes, into buffer buf, the cells changed this round and the last
  // (buffer buf last received them two rounds ago), then starts a new list
  auto publish_cells = [&](int buf) {
    __syncthreads();
    const int nc = dcur_n, np = dprev_n;
    if (nc > kMaxDirtyCells || np > kMaxDirtyCells) {
      for (int lc = tid; lc < ncl; lc += kLoopThreads) cell_record(lc, buf);
    } else {
      for (int e = tid; e < nc + np; e += kLoopThreads)
        cell_record(e < nc ? dcur[e] : dprev[e - nc], buf);
    }
    __syncthreads();
    if (nc > kMaxDirtyCells) {
      for (int i = tid; i < (ncl + 31) / 32; i += kLoopThreads) dbits[i] = 0;
    } else {
      for (int e = tid; e < nc; e += kLoopThreads) {
        dbits[dcur[e] >> 5] = 0;
        dprev[e] = dcur[e];
      }
    }
    __syncthreads();
    if (tid == 0) {
      dprev_n = nc;
      dcur_n = 0;
    }
  };
  auto publish = [&](int parity) {
    uint64_t *kn = a.keys + (size_t)parity * nunits * 3;
    uint32_t *pn = a.kpos + (size_t)parity * nunits * 2;
    for (int e = tid; e < 3 * upw; e += kLoopThreads) {
      const int ul = e / 3, x = e - ul * 3;
      const int u = ul * G + g;
      if (u < nunits) {
        st_dev(&kn[3 * u + x], ukey[e]);
        if (x < 2) st_dev(&pn[2 * u + x], upos[2 * ul + x]);
      }
    }
  };

  for (int ul = wid; ul < upw; ul += kLoopWaves) unit_top2(ul);
  __syncthreads();
  publish(0);
  if (ivm)
    for (int lc = tid; lc < ncl; lc += kLoopThreads) {
      cell_record(lc, 0);
      cell_record(lc, 1);
    }
  unsigned int bar_target = 0;
  const bool started = grid_sync(&a.st->barrier, bar_target, (unsigned)G, a.spin_limit);

  int nused = 0, niter = started ? 1 : -1, parity = 0;
  long long cumarea = 0, my_u = 0;
  // Divergence check: every workgroup computes the round's picks from the
  // same published keys, so all must agree.  The last wavefront publishes a
  // hash of this workgroup's picks each round and, during the next round's
  // key gather (which it takes no part in), compares every workgroup's hash
  // with its own.  All workgroups read the same G values, so either all
  // continue or all leave the loop at the same point (niter = -2), without
  // a barrier that some would never reach.
  // (kept in LDS: registers are the loop's scarcest resource)
  __shared__ int ph_bad, ph_round;
  __shared__ uint64_t ph_prev;  // this group's hash of the previous round
  if (tid == 0) ph_round = 0;
  __syncthreads();

  // the walk's helpers (see mb_seq): rounds that may compute areas
  // (a shed box of at most kWave * kBatch items is computed by the walking
  // wavefront alone: the handoff would cost more than it saves)
  const bool walk_help =
      (kIv ? ivm && (a.walk_evals > 0 || a.near_evals > 0) : a.near_evals > 0) &&
                         (2 * roi + 1) * (nwpr + 1) > kWave * kBatch;
  SITE_PROF(if (tid == 0) {
    for (int q = 0; q < kProfWords; q++) pf_acc[q] = 0;
    pf_last = wall_clock64();
  })
  SITE_PROF(int pf_r = 0;)
  while (started) {
    SITE_PROF_MARK(6)
    SITE_PROF(if (tid == 0) pf_r = ph_round;)
    SITE_TRACE(0)
    if (tid == 0) mb_seq = 0;
    if (wid == kLoopWaves - 1) {
      bool bad = false;
      if (ph_round > 0) {
        const uint64_t *ph = a.phash + (size_t)(parity ^ 1) * kMaxGroups;
        const uint64_t mine = ph_prev;
        for (int w = lane; w < G; w += kWave) bad |= ld_dev(&ph[w]) != mine;
      }
      const bool any_bad = __any(bad);
      if (lane == 0) ph_bad = any_bad;
    }
    // ---- pick the round's winners (identical in every workgroup).  Winners
    // follow the descending order of all published unit keys, as the
    // sequential loop would pick them: a key far (> 2*roi) from every
    // earlier pick of the round is exact; a nearer one is re-keyed to its
    // exact area after those picks (a positive key's area only shrinks, so
    // it goes back into the order); under intervis a key is taken only if
    // the cell bounds of the masked candidates admit it.  Each unit
    // publishes its two largest keys with positions and its third largest
    // key, which stands for all the unit's keys below its second (a bound
    // entry: it ends the round if it comes first).
    //
    // Lane l of wavefronts 0..3 keeps the three largest entries of units
    // l, l + 64, ... (strided by 256) in registers, and the largest it
    // dropped; wavefronts 1..3 hand theirs to wavefront 0 through LDS, whose
    // lanes then hold the three largest of units l + 64 j.  Once a lane's
    // three are taken, its largest dropped key follows as a bound entry.
    {
      const uint64_t *kk = a.keys + (size_t)parity * nunits * 3;
      const uint32_t *kp = a.kpos + (size_t)parity * nunits * 2;
      Top3 h{};
      if (wid < kPickWaves) {
        for (int u0 = wid * kWave + lane; u0 < nunits; u0 += 2 * kPickWaves * kWave) {
          const int u1 = u0 + kPickWaves * kWave;
          const bool has1 = u1 < nunits;
          // every load unconditional (a load under a branch is waited for
          // inside it); a missing second unit reads the first's and drops it
          const int u1s = has1 ? u1 : u0;
          const uint64_t a0 = ld_dev(&kk[3 * u0]), a1 = ld_dev(&kk[3 * u0 + 1]);
          const uint64_t a2 = ld_dev(&kk[3 * u0 + 2]);
          const uint32_t q0 = ld_dev(&kp[2 * u0]), q1 = ld_dev(&kp[2 * u0 + 1]);
          const uint64_t b0v = ld_dev(&kk[3 * u1s]), b1v = ld_dev(&kk[3 * u1s + 1]);
          const uint64_t b2v = ld_dev(&kk[3 * u1s + 2]);
          const uint32_t r0v = ld_dev(&kp[2 * u1s]), r1v = ld_dev(&kp[2 * u1s + 1]);
          const uint64_t b0 = has1 ? b0v : 0, b1 = has1 ? b1v : 0, b2 = has1 ? b2v : 0;
          const uint32_t r0 = has1 ? r0v : 0, r1 = has1 ? r1v : 0;
          // a unit's third key stands for every key it holds below its
          // second: a bound entry
          top3_push(h, a0, q0);
          top3_push(h, a1, q1);
          top3_push(h, a2, kGhost);
          top3_push(h, b0, r0);
          top3_push(h, b1, r1);
          top3_push(h, b2, kGhost);
        }
        if (wid > 0) {
          const int s = (wid - 1) * 3;
          xk[s][lane] = h.k[0];
          xk[s + 1][lane] = h.k[1];
          xk[s + 2][lane] = h.k[2];
          xp[s][lane] = h.p[0];
          xp[s + 1][lane] = h.p[1];
          xp[s + 2][lane] = h.p[2];
          xo[wid - 1][lane] = h.over;
        }
      }
      __syncthreads();
      SITE_PROF_MARK(7)
      SITE_TRACE(1)
      // The walk's order, sorted once: wavefront 0's lanes hold 256 entries
      // (each lane's three keys and its largest dropped key, a bound entry),
      // and the order in which the walk would take the lanes' largest keys
      // one by one is their descending order.  Every thread counts, for one
      // entry, the larger entries among a quarter of them (keys are distinct,
      // 0 an empty slot); the 64 largest go to sk/sp by rank, and the walk
      // steps through them with readlanes.  (The rows of xk/xp are free once
      // wavefront 0 has merged them.)
      static_assert(3 * (kPickWaves - 1) >= 9 && kLoopThreads == 16 * kWave,
                    "the sort's rows of xk/xp and its four quarters of 256 entries");
      uint64_t *ek = &xk[0][0];         // entry s * kWave + l: lane l's s-th
      uint32_t *ep = &xp[0][0];
      int *erank = (int *)&xp[4][0];    // larger entries than entry e
      uint64_t *sk = &xk[4][0];         // the 64 largest, descending (0: none)
      uint32_t *sp = &xp[8][0];
      if (wid == 0) {
#pragma unroll
        for (int s = 0; s < 3 * (kPickWaves - 1); s++) top3_push(h, xk[s][lane], xp[s][lane]);
#pragma unroll
        for (int s = 0; s < kPickWaves - 1; s++) h.over = xo[s][lane] > h.over ? xo[s][lane] : h.over;
        // (each lane reads its own slots above before writing them here)
        ek[lane] = h.k[0];
        ek[kWave + lane] = h.k[1];
        ek[2 * kWave + lane] = h.k[2];
        ek[3 * kWave + lane] = h.over;
        ep[lane] = h.p[0];
        ep[kWave + lane] = h.p[1];
        ep[2 * kWave + lane] = h.p[2];
        ep[3 * kWave + lane] = kGhost;
#pragma unroll
        for (int s = 0; s < 4; s++) erank[s * kWave + lane] = 0;
        sk[lane] = 0;
      }
      __syncthreads();
      {
        const int e = tid & (4 * kWave - 1), q = tid / (4 * kWave);
        const uint64_t mine = ek[e];
        int c = 0;
#pragma unroll 16
        for (int j = 0; j < kWave; j++) c += ek[q * kWave + j] > mine ? 1 : 0;
        if (mine != 0 && c != 0) atomicAdd(&erank[e], c);
      }
      __syncthreads();
      if (wid == 0) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
          const int e = s * kWave + lane, r = erank[e];
          const uint64_t kv = ek[e];
          if (kv != 0 && r < kWave) {
            sk[r] = kv;
            sp[r] = ep[e];
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const uint64_t skr = sk[lane];
        const uint32_t spr = sp[lane];
        int si = 0;  // the next entry of the order
        SITE_PROF_MARK(20)
        // Lane j < k holds pick j.
        uint64_t pk = 0;
        uint32_t pp = 0;
        int k = 0;
        int head = 0;  // area (or bound) of the largest key
        // intervis: a masked candidate gains area only if a pick's
        // viewshed covers its cell, so only masked candidates within roi of
        // a pick count.  From the last round's records of the cells within
        // roi of each pick taken: every cell's third largest masked bound
        // (its second largest, beyond roi 40) gives `bound`; each of its two
        // largest (its largest, beyond roi 40), when the pick sees its cell,
        // is listed (ev_*) and computed exactly against cum and the pick's
        // bits only when it could beat the key at hand (its area then is at
        // most that: other picks only cover more), and when the pick does
        // not see its cell it stays masked.  Picks are never masked (a
        // masked area is 0).
        // Area of candidate cid (at cpos) against cum | the picks of the
        // lanes in `picks`, computed with the helpers.
        int mb_n = 0;
        auto walk_area_core = [&](uint32_t cid, uint32_t cpos, uint64_t picks) -> int {
          if (!walk_help) {  // a small shed: this wavefront alone
            const int np = __popcll(picks);
            const int slot = (int)__builtin_amdgcn_mbcnt_hi(
                (uint32_t)(picks >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)picks, 0));
            if ((picks >> lane) & 1) {
              mb_ppos[slot] = pp;
              mb_pid[slot] = ~(uint32_t)pk;
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
            __builtin_amdgcn_wave_barrier();
            return wave_sum_i32(lane_new_bits_picks<kBatch, 1>(a.sheds + (size_t)cid * S, (int)(cpos >> 16),
                                                    (int)(cpos & 0xffffu), np, mb_ppos, mb_pid, a,
                                                    lane, kWave));
          }
          const int slot = (int)__builtin_amdgcn_mbcnt_hi(
              (uint32_t)(picks >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)picks, 0));
          if ((picks >> lane) & 1) {
            mb_ppos[slot] = pp;
            mb_pid[slot] = ~(uint32_t)pk;
          }
          if (lane == 0) {
            mb_cid = cid;
            mb_cpos = cpos;
            mb_np = __popcll(picks);
            mb_acc = 0;
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
          mb_n++;
          if (lane == 0) __hip_atomic_store(&mb_seq, mb_n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          unsigned long long acc = 0;
          if (lane == 0)
            while (((acc = __hip_atomic_load(&mb_acc, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) >> 40) <
                   (unsigned long long)(kLoopWaves - 1))
              __builtin_amdgcn_s_sleep(1);
          // (lane 0's sum by readlane: a wave-uniform value, so that the
          // walk's control flow that depends on it stays scalar)
          return __builtin_amdgcn_readlane((int)(uint32_t)(acc & ((1ull << 40) - 1)), 0);
        };
        auto walk_area = [&](uint32_t cid, uint32_t cpos, uint64_t picks) -> int {
          SITE_PROF(const unsigned long long t0 = wall_clock64();)
          const int r = walk_area_core(cid, cpos, picks);
          SITE_PROF_ADD(10, wall_clock64() - t0)
          SITE_PROF_ADD(9, 1)
          return r;
        };
        uint32_t bound = 0;
        int nwin = 0, ev_n = 0;
        auto window_scan = [&](int j) {
          const int px = __builtin_amdgcn_readlane((int)(pp >> 16), j);
          const int py = __builtin_amdgcn_readlane((int)(pp & 0xffffu), j);
          const uint32_t pid = ~(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)pk, j);
          const uint64_t *psh = a.sheds + (size_t)pid * S;
          const int r0 = max(0, px - roi) / a.cs, r1 = min(n - 1, px + roi) / a.cs;
          const int c0 = max(0, py - roi) / a.cs, c1 = min(n - 1, py + roi) / a.cs;
          const int ncw = c1 - c0 + 1, cnt_w = (r1 - r0 + 1) * ncw;
          const uint64_t *cb = a.cbnd + (size_t)parity * a.ncr * a.ncr * 4;
          // a cell's second largest is listed too for roi up to 40 (a 7594^2
          // roi-25 request: 26575 -> 17249 rounds, SITE 647 -> 573 ms), else
          // its bound counts (the 2048^2 roi-80 job: 113 vs 118 ms listed)
          const bool two = roi <= 40;
          // does the pick see the cell at position pos (within its disk)?
          // every load unconditional, from the pick's centre when outside
          auto sees = [&](uint32_t pos, bool live) -> bool {
            const int dx = (int)(pos >> 16) - px, dy = (int)(pos & 0xffffu) - py;
            const bool in = live && dx * dx + dy * dy <= roi * roi;
            const int row = in ? dx + roi : roi, col = in ? dy + roi : roi;
            const uint64_t w = psh[(size_t)row * nwpr + (col >> 6)];
            return in && ((w >> (63 - (col & 63))) & 1u);
          };
          uint32_t m = 0;
          // list a seen candidate (beyond the list, its bound counts)
          auto list = [&](bool seen, uint32_t pos, uint32_t id, uint32_t bd) {
            const uint64_t sv = __ballot(seen);
            const int at = ev_n + (int)__builtin_amdgcn_mbcnt_hi(
                                      (uint32_t)(sv >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)sv, 0));
            if (seen) {
              if (at < kMaxBatch) {
                ev_pos[at] = pos;
                ev_id[at] = id;
                ev_bd[at] = (int)bd;
                ev_pick[at] = j;
                ev_exact[at] = -1;
              } else {
                m = bd > m ? bd : m;
              }
            }
            ev_n += __popcll(sv);
          };
          // kWinChunks chunks of 64 cells a pass: their records, then the
          // pick's shed words at their candidates' cells, each in one round
          // trip (listed in the same order as one chunk at a time)
          for (int e0 = 0; e0 < cnt_w; e0 += kWinChunks * kWave) {
            uint64_t w[kWinChunks][4];
            bool ok[kWinChunks];
#pragma unroll
            for (int h = 0; h < kWinChunks; h++) {
              const int e = e0 + h * kWave + lane;
              ok[h] = e < cnt_w;
              const int r = r0 + e / ncw, cc = c0 + e % ncw;
              const uint64_t *rec = cb + (ok[h] ? (size_t)r * a.ncr + cc : 0) * 4;
              w[h][0] = ld_dev(&rec[0]);
              w[h][1] = ld_dev(&rec[1]);
              w[h][2] = ld_dev(&rec[2]);
              w[h][3] = ld_dev(&rec[3]);
            }
            uint32_t b1[kWinChunks], b2[kWinChunks];
            bool s1[kWinChunks], s2[kWinChunks];
#pragma unroll
            for (int h = 0; h < kWinChunks; h++) {
              b1[h] = ok[h] ? (uint32_t)(w[h][0] >> 32) : 0;
              b2[h] = ok[h] ? (uint32_t)(w[h][1] >> 32) : 0;
              const uint32_t b3 = ok[h] ? (uint32_t)(w[h][2] >> 32) : 0;
              m = b3 > m ? b3 : m;  // the rest of the cell's masked candidates
              // the two largest: only if the pick sees its cell can it gain
              s1[h] = sees((uint32_t)w[h][0], b1[h] != 0);
              s2[h] = sees((uint32_t)w[h][2], two && b2[h] != 0);
              if (!two) m = b2[h] > m ? b2[h] : m;
            }
#pragma unroll
            for (int h = 0; h < kWinChunks; h++) {
              list(s1[h], (uint32_t)w[h][0], (uint32_t)w[h][1], b1[h]);
              list(s2[h], (uint32_t)w[h][2], (uint32_t)w[h][3], b2[h]);
            }
          }
          if (ev_n > kMaxBatch) ev_n = kMaxBatch;
          const uint32_t wm = (uint32_t)wave_max_u64_dpp((uint64_t)m);
          bound = wm > bound ? wm : bound;
        };
        // true if a listed candidate could reach area A (computed on demand,
        // at most a.walk_evals per round: one wavefront computes them while
        // the others wait)
        int ev_evals = 0;
        auto listed_block = [&](int A) -> bool {
          for (int e = 0; e < ev_n; e++) {
            if (ev_bd[e] < A) continue;
            if (ev_exact[e] < 0) {
              if (ev_evals >= a.walk_evals) return true;
              ev_evals++;
              const int s_exact = walk_area(ev_id[e], ev_pos[e], 1ull << ev_pick[e]);
              if (lane == 0) ev_exact[e] = s_exact;
              __builtin_amdgcn_wave_barrier();
            }
            if (ev_exact[e] >= A) return true;
          }
          return false;
        };
        // A key within 2*roi of an earlier pick of the round is re-keyed --
        // its exact area once the earlier picks within 2*roi are taken, as
        // the sequential loop recomputed it after each of them -- and joins
        // the rekeyed set (r_*, exact through the first r_mark picks)
        // instead of ending the round; the largest of the lanes' keys and the
        // rekeyed set goes next (a positive key's area only shrinks, so a
        // lane's key bounds the area of its candidate; masked candidates,
        // which may gain, are covered by the cell bounds).  At most
        // a.near_evals such computations per round.
        int rn = 0, nevals = 0, npair = 0;
        // A re-keyed candidate's area once the earlier picks within 2*roi of
        // it are taken.  Intervis: 0 unless its own cell is covered by cum or
        // one of those picks' sheds (siting.cpp:452-457; only picks within
        // roi can cover it); its area can then only have shrunk, since a
        // candidate whose key is positive was not masked.
        auto rekey_area = [&](uint32_t cid, uint32_t cpos, uint64_t nearp) -> int {
          if (kIv) {
            const int ccx = (int)(cpos >> 16), ccy = (int)(cpos & 0xffffu);
            bool cov = false;
            if ((nearp >> lane) & 1) {
              const int dr = ccx - (int)(pp >> 16) + roi, dc = ccy - (int)(pp & 0xffffu) + roi;
              if (dr >= 0 && dr < nr && dc >= 0 && dc < nr) {
                const uint64_t *psh = a.sheds + (size_t)(~(uint32_t)pk) * S;
                cov = (psh[(size_t)dr * nwpr + (dc >> 6)] >> (63 - (dc & 63))) & 1u;
              }
            }
            const uint64_t cwd = ld_dev(&a.cum[cum_index(ccx, ccy >> 6, cw)]);
            const bool vis = ((cwd >> (63 - (ccy & 63))) & 1u) || __ballot(cov) != 0;
            if (!vis) return 0;
          }
          return walk_area(cid, cpos, nearp);
        };
        // the largest key of the rekeyed set (0 while it is empty), kept
        // across steps and refreshed only when the set changes
        uint64_t rmax = 0;
        auto rmax_of_set = [&]() -> uint64_t {
          return wave_max_u64_dpp(lane < rn ? r_key[lane] : 0);
        };
        while (k < kmax) {
          SITE_PROF_ADD(11, 1)
          // the largest key not yet taken; past the 64th, the 64th stands
          // for the rest as a bound entry
          const int sj = si < kWave ? si : kWave - 1;
          const uint64_t bv =
              ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(skr >> 32), sj) << 32) |
              (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)skr, sj);
          if (rn > 0 && rmax >= bv) {
            // (equal only to the bound a rekeyed lane left, which is its own)
            const uint64_t rv = rmax;
            const uint64_t rkv = lane < rn ? r_key[lane] : 0;
            {
              const int e = (int)__builtin_ctzll(__ballot(rkv == rv));
              const uint32_t rpos = r_pos[e];
              const int mark = r_mark[e];
              const int rx = (int)(rpos >> 16), ry = (int)(rpos & 0xffffu);
              bool fresh = false, near_any = false;
              if (lane < k) {
                const int ex = rx - (int)(pp >> 16), ey = ry - (int)(pp & 0xffffu);
                near_any = ex * ex + ey * ey <= r2;
                fresh = near_any && lane >= mark;
              }
              const uint64_t nearp = __ballot(near_any);
              if (__ballot(fresh)) {  // picks since its last computation reach it
                if (nevals >= a.near_evals) {
                  SITE_PROF_ADD(13, 1)
                  break;
                }
                nevals++;
                const int area = __builtin_amdgcn_readfirstlane(rekey_area(~(uint32_t)rv, rpos, nearp));
                if (lane == 0) {
                  r_key[e] = ((uint64_t)(uint32_t)area << 32) | (uint32_t)rv;
                  r_mark[e] = k;
                }
                __builtin_amdgcn_wave_barrier();
                rmax = rmax_of_set();
                continue;
              }
              if ((rv >> 32) == 0) {  // no candidate adds area
                SITE_PROF_ADD(17, 1)
                break;
              }
              if (ivm) {  // it must beat every masked candidate's bound, as below
                for (; nwin < k; nwin++) window_scan(nwin);
                if ((uint32_t)(rv >> 32) <= bound || listed_block((int)(rv >> 32))) break;
              }
              // take it; it was in the update list of each earlier pick
              // within 2*roi (cpp:540-544), which leaves every pick out
              if (lane == k) {
                pk = rv;
                pp = rpos;
                npair = __popcll(nearp);
              }
              if (lane == 0) {
                r_key[e] = r_key[rn - 1];
                r_pos[e] = r_pos[rn - 1];
                r_mark[e] = r_mark[rn - 1];
              }
              __builtin_amdgcn_wave_barrier();
              rn--;
              rmax = rmax_of_set();
              k++;
              continue;
            }
          }
          if ((bv >> 32) == 0) {  // no candidate adds area
            SITE_PROF_ADD(17, 1)
            break;
          }
          if (k == 0) head = (int)(bv >> 32);
          const uint32_t pos =
              si < kWave ? (uint32_t)__builtin_amdgcn_readlane((int)spr, sj) : kGhost;
          // a bound entry (keys the lane does not hold may come next), or a
          // stale key (intervis 0: its area is only a bound)
          if (pos >> 31) {
            SITE_PROF_ADD(pos == kGhost ? 15 : 16, 1)
            break;
          }
          const int px = (int)(pos >> 16), py = (int)(pos & 0xffffu);
          bool near = false;
          if (lane < k) {
            const int ex = px - (int)(pp >> 16);
            const int ey = py - (int)(pp & 0xffffu);
            near = ex * ex + ey * ey <= r2;
          }
          const uint64_t nearp = __ballot(near);
          if (nearp) {
            if (rn == kMaxRekey || nevals >= a.near_evals) {
              SITE_PROF_ADD(rn == kMaxRekey ? 14 : 13, 1)
              break;
            }
            nevals++;
            const int area = __builtin_amdgcn_readfirstlane(rekey_area(~(uint32_t)bv, pos, nearp));
            const uint64_t rk = ((uint64_t)(uint32_t)area << 32) | (uint32_t)bv;
            if (lane == 0) {
              r_key[rn] = rk;
              r_pos[rn] = pos;
              r_mark[rn] = k;
            }
            __builtin_amdgcn_wave_barrier();
            rn++;
            rmax = rk > rmax ? rk : rmax;
            si++;
            continue;
          }
          if (ivm && k > 0) {
            // a candidate near an earlier pick may gain area when the pick
            // uncovers its cell (siting.cpp:452-457), up to its bound: this
            // key must beat every such bound
            for (; nwin < k; nwin++) window_scan(nwin);
            if ((uint32_t)(bv >> 32) <= bound || listed_block((int)(bv >> 32))) break;
          }
          if (lane == k) {
            pk = bv;
            pp = pos;
          }
          si++;
          k++;
        }
        SITE_PROF_ADD(12, k)
        SITE_PROF_ADD(18, k >= kmax)
        SITE_PROF_ADD(19, k == 0)
        if (lane < k) {
          bw_key[lane] = pk;
          bw_pos[lane] = pp;
          bw_npair[lane] = npair;
        }
        if (lane == 0) {
          bw_n = k;
          bw_head = head;
          bw_readcum = nevals + ev_evals > 0;
          if (walk_help) __hip_atomic_store(&mb_seq, -1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      } else if (walk_help) {
        // the walk's helpers: the 15 other wavefronts share each area
        // computation it posts (the walking wavefront only waits; with a
        // share of its own the loop took 1-2 % longer)
        int seen = 0;
        for (;;) {
          int q;
          while ((q = __hip_atomic_load(&mb_seq, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)) == seen)
            __builtin_amdgcn_s_sleep(1);
          if (q < 0) break;
          seen = q;
          const uint32_t cpos = mb_cpos;
          const int part = wave_sum_i32(lane_new_bits_picks_np(
              a.sheds + (size_t)mb_cid * S, (int)(cpos >> 16), (int)(cpos & 0xffffu), mb_np,
              mb_ppos, mb_pid, a, (wid - 1) * kWave + lane, kLoopThreads - kWave));
          if (lane == 0)
            __hip_atomic_fetch_add(&mb_acc, (1ull << 40) + (unsigned long long)part, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
    }
    __syncthreads();
    if (ph_bad) {
      niter = -2;
      break;
    }
    SITE_PROF_MARK(0)
    SITE_TRACE(2)
    int k = bw_n;
    // no pick while the largest key is stale: the round only re-evaluates
    const bool eval_only = k == 0 && bw_head > 0;
    if (k == 0 && !eval_only) break;  // cpp:589-592
    // A walk that read cum (exact areas) must end in every workgroup before
    // any of them ORs this round's picks into it.  Every workgroup's walk is
    // the same up to its first read of cum, so either all arrive here or
    // none does.
    const bool walk_sync = bw_readcum;
    if (walk_sync) grid_arrive(&a.st->barrier, bar_target, (unsigned)G);
    if (wid == kLoopWaves - 1) {
      // hash of (picks, count, head); stored before this round's barriers
      uint64_t v = 0;
      if (lane < k) {
        v = bw_key[lane] ^ ((uint64_t)bw_pos[lane] * 0x9E3779B97F4A7C15ull) ^
            ((uint64_t)(lane + 1) << 58);
        v ^= v >> 31;
        v *= 0xBF58476D1CE4E5B9ull;
        v ^= v >> 29;
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v ^= __shfl_xor(v, o, kWave);
      v ^= ((uint64_t)(uint32_t)k << 32) | (uint32_t)bw_head;
      const int round = ph_round + 1;
      // test knob: the last workgroup's hash of round hash_fault is altered
      if (a.hash_fault == round && g == G - 1) v ^= 1;
      if (lane == 0) {
        ph_prev = v;
        ph_round = round;
        st_dev(&a.phash[(size_t)parity * kMaxGroups + g], v);
      }
    }
    // update candidates below tau become stale; stale ones reaching tau are
    // re-evaluated (areas only shrink, so a bound never falls below the area)
    const int tau = hybrid ? bw_head - (bw_head >> a.tau_shift) : 0;
    if (nused + k > a.nsheds) {  // cannot happen: every pick retires a candidate
      niter = -1;
      break;
    }

    // ---- bookkeeping in pick order; the 95% rule may cut the round short
    bool stop = false;
    bool union_pending = false;  // arrived at the union barrier, not waited
    int kc = k;  // picks whose update lists the sequential loop evaluates
    {
      // every wavefront: lane i holds pick i's area, an inclusive prefix sum
      // gives the cumulative area after each pick, and the first pick that
      // reaches the stop area ends the round (cpp:601-604)
      long long c = lane < k ? (long long)(bw_key[lane] >> 32) : 0;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const long long up = __shfl_up(c, o, kWave);
        if (lane >= o) c += up;
      }
      const uint64_t hits = __ballot(lane < k && cumarea + c >= a.stop_area);
      const int nused0 = nused;
      if (hits) {
        const int i = (int)__builtin_ctzll(hits);
        stop = true;
        k = i + 1;
        kc = i;
      }
      nused += k;
      niter += kc;
      cumarea += __shfl(c, k - 1, kWave);
      if (g == 0 && tid < k) {
        const int w = (int)(~(uint32_t)bw_key[tid]);
        a.used[w] = 1;
        a.order[nused0 + tid] = w;
        my_u += bw_npair[tid];  // in the update lists of earlier picks
      }
    }
    // owners mark the picks used: wavefront i takes picks i, i + 16, ...
    for (int pi = wid; pi < k; pi += kLoopWaves) {
      const int w = (int)(~(uint32_t)bw_key[pi]);
      const int wx = (int)(bw_pos[pi] >> 16), wy = (int)(bw_pos[pi] & 0xffffu);
      const int w_r = wx / a.cs, w_cc = wy / a.cs;
      const int w_u = w_r * Q + w_cc % Q;
      if (w_u % G == g) {
        const int ul = w_u / G, kc2 = w_cc / Q;
        const int lo = OFF(ul, kc2), hi = OFF(ul, kc2 + 1);
        for (int j = lo + lane; j < hi; j += kWave)
          if (C[j] == w) U[j] = 1;
        if (lane == 0) {
          dirty[ul] = 1;
          if (ivm) mark_cell(wx, wy);
        }
      }
    }

    if (walk_sync && !grid_wait(&a.st->barrier, bar_target, a.spin_limit)) {
      niter = -1;
      break;
    }

    SITE_PROF_MARK(1)
    SITE_TRACE(4)
    // ---- union of the picks into cum (cpp:393-419)
    Winner W;
    {
      const int w = (int)(~(uint32_t)bw_key[0]);
      W.w = w;
      W.wx = (int)(bw_pos[0] >> 16);
      W.wy = (int)(bw_pos[0] & 0xffffu);
      col_split(W.wy - roi, W.w0w, W.fbw);
      W.nr = nr;
      W.nwpr = nwpr;
      W.dw = dw;
      W.roi = roi;
      W.sw = a.sheds + (size_t)w * S;
      W.delta = nullptr;
    }
    if (kIv && k == 1) {
      // one pick: areas are recomputed against cum | delta while the union
      // is written (a reader sees either word, the OR is the same)
      if (a.delta_lds) {
        for (int i = tid; i < nr * dw; i += kLoopThreads) {
          const int dr = i / dw, dk = i - (i / dw) * dw;
          delta[i] = winner_word(W.sw, nwpr, W.fbw, dr, dk);
        }
        W.delta = delta;
      }
      __syncthreads();
      const int total = nr * dw;
      const int chunk = (total + G - 1) / G;
      const int t0 = g * chunk, t1 = min(total, t0 + chunk);
      for (int t = t0 + tid; t < t1; t += kLoopThreads) {
        const int dr = t / dw, dk = t - (t / dw) * dw;
        const int x = W.wx - roi + dr, j = W.w0w + dk;
        if (x < 0 || x >= n || j < 0 || j >= cw) continue;
        const uint64_t v = W.word(dr, dk);
        if (v) atomicOr((unsigned long long *)&a.cum[cum_index(x, j, cw)], (unsigned long long)v);
      }
    } else if (k > 0) {
      // every pick: OR its shifted shed into cum and keep the bits that were
      // new; picks within 2*roi of each other share bits, and the atomic
      // gives each new bit to exactly one of them, so the kept bits of the
      // picks are disjoint and together the union's new bits
      __syncthreads();
      const int total = k * nr * dw;
      const int chunk = (total + G - 1) / G;
      const int t0 = g * chunk, t1 = min(total, t0 + chunk);
      for (int t = t0 + tid; t < t1; t += kLoopThreads) {
        const int i = t / (nr * dw), rem = t - i * (nr * dw);
        const int dr = rem / dw, dk = rem - (rem / dw) * dw;
        const int wx = (int)(bw_pos[i] >> 16), wy = (int)(bw_pos[i] & 0xffffu);
        const int x = wx - roi + dr;
        int w0w, fbw;
        col_split(wy - roi, w0w, fbw);
        const int j = w0w + dk;
        uint64_t nbv = 0;
        if (x >= 0 && x < n && j >= 0 && j < cw) {
          const uint64_t *sw = a.sheds + (size_t)(int)(~(uint32_t)bw_key[i]) * S;
          const uint64_t v = winner_word(sw, nwpr, fbw, dr, dk);
          if (v && a.full_at > 1) {
            const uint64_t old = atomicOr((unsigned long long *)&a.cum[cum_index(x, j, cw)],
                                          (unsigned long long)v);
            nbv = v & ~old;
          } else if (v) {
            atomicOr((unsigned long long *)&a.cum[cum_index(x, j, cw)], (unsigned long long)v);
          }
        }
        // the new bits are read only by decrements (full_at > 1)
        if (a.full_at > 1) st_dev(&a.nb[t], nbv);
      }
      // the update lists are built before the wait; nb and cum are read
      // after it
      grid_arrive(&a.st->barrier, bar_target, (unsigned)G);
      union_pending = true;
    }
    bool failed = false;
    auto union_wait = [&]() {
      if (union_pending) {
        union_pending = false;
        failed = !grid_wait(&a.st->barrier, bar_target, a.spin_limit);
      }
    };
    // intervis rounds of several picks recompute their update candidates
    // against the whole union: its barrier completes first
    const bool ivk = kIv && k > 1;
    if (ivk) union_wait();

    // ---- update lists of picks 0..kc-1 among this workgroup's candidates.
    // Ranges: for each pick, the cells of owned units inside the pick's
    // 2*roi window (unit/cell offsets); their concatenation is walked in
    // windows of kMaxList items; a candidate in several ranges is listed
    // once (seen bitmap, cleared at the end of the round).
    SITE_PROF_MARK(2)
    SITE_TRACE(5)
    if (tid == 0) rng_n = 0;
    __syncthreads();
    for (int t = tid; t < kc * upw; t += kLoopThreads) {
      const int p = t / upw, ul = t - p * upw;
      const int u = ul * G + g;
      if (u >= nunits) continue;
      const int r = u / Q, q = u - r * Q;
      const int wx = (int)(bw_pos[p] >> 16), wy = (int)(bw_pos[p] & 0xffffu);
      if (r < max(0, (wx - 2 * roi) / a.cs) || r > (wx + 2 * roi) / a.cs) continue;
      const int cclo = max(0, (wy - 2 * roi) / a.cs);
      const int cchi = min(a.ncr - 1, (wy + 2 * roi) / a.cs);
      const int klo = cclo <= q ? 0 : (cclo - q + Q - 1) / Q;
      const int khi = cchi < q ? -1 : (cchi - q) / Q;
      if (klo > khi) continue;
      const int lo = OFF(ul, klo), hi = OFF(ul, khi + 1);
      if (lo >= hi) continue;
      const int at = atomicAdd(&rng_n, 1);
      if (at < kMaxList) {
        rng_lo[at] = lo;
        rng_end[at] = hi - lo;
      }
    }
    __syncthreads();
    int nrng = rng_n;
    if (nrng > kMaxList) {  // too many ranges: scan every owned candidate
      __syncthreads();
      if (tid == 0) {
        rng_lo[0] = 0;
        rng_end[0] = cnt;
        rng_n = 1;
      }
      __syncthreads();
      nrng = 1;
    }
    {
      const int v = tid < nrng ? rng_end[tid] : 0;
      const int inc = block_inclusive_scan<kLoopThreads>(v, wsum);
      __syncthreads();
      if (tid < nrng) rng_end[tid] = inc;
    }
    __syncthreads();
    const int total = nrng > 0 ? rng_end[nrng - 1] : 0;
    bool reeval_rest = true;  // stale candidates left for the re-evaluation below
    for (int w0 = 0; w0 < total; w0 += kMaxList) {
      const int item = w0 + tid;
      if (item < total) {  // cpp:540-544, once per list
        int lo = 0, hi = nrng - 1;  // first range whose end exceeds item
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (rng_end[mid] > item) hi = mid; else lo = mid + 1;
        }
        const int i = rng_lo[lo] + item - (lo > 0 ? rng_end[lo - 1] : 0);
        const unsigned int bit = 1u << (i & 31);
        // the candidate's flags, position and area in one round trip (the
        // empty asm keeps the loads together, before any use)
        const uint8_t ui = U[i];
        const uint32_t ps = P[i];
        const int ti = T[i];
        asm volatile("" ::"v"((int)ui), "v"(ps), "v"(ti));
        if (!(ui & 1) && !(atomicOr(&seen[i >> 5], bit) & bit)) {
          const int cx = (int)(ps >> 16), cy = (int)(ps & 0xffffu);
          int m = 0;
          for (int j = 0; j < kc; j++) {
            const int ex = cx - (int)(bw_pos[j] >> 16);
            const int ey = cy - (int)(bw_pos[j] & 0xffffu);
            m += ex * ex + ey * ey <= r2;
          }
          if (m && hybrid && ((ui & 2) || ti < tau)) {
            my_u += m;  // keep the old area as a bound
            U[i] |= 2;
            dirty[((cx / a.cs) * Q + (cy / a.cs) % Q) / G] = 1;
          } else if (m) {
            int acc = m;
            if (!stop && kIv) {  // cpp:452-457
              uint64_t word = ld_dev(&a.cum[cum_index(cx, cy >> 6, cw)]);
              if (!ivk) {  // one pick: cum | its bits
                const int dr = cx - W.wx + roi, dk = (cy >> 6) - W.w0w;
                if (dr >= 0 && dr < nr && dk >= 0 && dk <= nwpr) word |= W.word(dr, dk);
              }
              if (!((word >> (63 - (cy & 63))) & 1u)) acc = -m;
            }
            const int at = atomicAdd(&list_n, 1);
            list_i[at] = i;
            list_acc[at] = acc;
            dirty[((cx / a.cs) * Q + (cy / a.cs) % Q) / G] = 1;
            if (ivm) mark_cell(cx, cy);
          }
        }
      }
      __syncthreads();
      const int lcnt = list_n;
      for (int e = tid; e < lcnt; e += kLoopThreads) {
        const int v = list_acc[e];
        my_u += v < 0 ? -v : v;
      }
      __syncthreads();  // counts read before the areas overwrite them
      // Intervis 0, the first window: the stale candidates whose bound
      // reaches tau join the window's list and all are counted in one pass
      // (a group of wavefronts an item when there are few), instead of the
      // update counts and then the re-evaluation.  Counts are exact against
      // cum after the union, so a later window that lists a re-evaluated
      // candidate again counts it again, to the same value.
      const bool merged = !kIv && hybrid && !stop && w0 == 0 && a.full_at <= 1;
      if (merged) {
        for (int i = tid; i < cnt; i += kLoopThreads)
          if ((U[i] & 3) == 2 && T[i] >= tau) {
            const int at = atomicAdd(&list_n, 1);
            if (at < kMaxList) list_i[at] = i;
          }
        __syncthreads();
        const int want = list_n;
        const int nall = want < kMaxList ? want : kMaxList;
        reeval_rest = want > kMaxList;
        if (nall > 0) union_wait();
        if (nall > 0 && !failed) {
          for (int e = tid; e < nall; e += kLoopThreads) list_acc[e] = 0;
          __syncthreads();
          int gw = 1;
          while (2 * gw * nall <= kLoopWaves) gw *= 2;
          const int ngroups = kLoopWaves / gw;
          const int grp = wid / gw, gl = (wid - grp * gw) * kWave + lane;
          for (int e0 = grp; e0 < nall; e0 += kWave * ngroups) {
            const int em = e0 + lane * ngroups;
            uint32_t mps = 0, mid = 0;
            if (em < nall) {
              const int i = list_i[em];
              mps = P[i];
              mid = (uint32_t)C[i];
            }
            const int ncand = min(kWave, (nall - e0 + ngroups - 1) / ngroups);
            for (int j = 0; j < ncand; j++) {
              const uint32_t ps = (uint32_t)__builtin_amdgcn_readlane((int)mps, j);
              const uint32_t cid = (uint32_t)__builtin_amdgcn_readlane((int)mid, j);
              const int cx = (int)(ps >> 16), cy = (int)(ps & 0xffffu);
              const uint64_t *sh = a.sheds + (size_t)cid * S;
              const int s = wave_sum_i32(
                  eval_rows_fit(cy, roi, nwpr)
                      ? lane_new_bits_rows<kEvalRows>(sh, cx, cy, a, gl, gw * kWave)
                      : lane_new_bits<kEvalBatch>(sh, cx, cy, a, W, gl, gw * kWave, false));
              if (lane == 0) atomicAdd(&list_acc[e0 + j * ngroups], s);
            }
          }
          __syncthreads();
          for (int e = tid; e < nall; e += kLoopThreads) {
            const int i = list_i[e];
            T[i] = list_acc[e];
            if (e >= lcnt) {  // re-evaluated: exact again
              U[i] = 0;
              const uint32_t ps = P[i];
              dirty[((int)(ps >> 16) / a.cs * Q + ((int)(ps & 0xffffu) / a.cs) % Q) / G] = 1;
            }
          }
        }
      }
      // (an intervis update reads cum | the winner's bits, the same before,
      // during and after the union: it needs no wait)
      if (!merged && !stop && lcnt > 0 && !kIv) union_wait();
      if (!merged && !stop && lcnt > 0 && !failed) {
        // one wavefront per update candidate
        for (int e = wid; e < lcnt; e += kLoopWaves) {
          if (list_acc[e] < 0) continue;
          const int i = list_i[e];
          const uint32_t ps = P[i];
          const int cx = (int)(ps >> 16), cy = (int)(ps & 0xffffu);
          const uint64_t *sh = a.sheds + (size_t)C[i] * S;
          int s;
          // The exact area: a count of the candidate's box against cum
          // once the union is complete (intervis rounds of one pick: against
          // cum | the pick's bits, while the union is written).  Its area
          // minus the bits this round's picks newly covered is the same
          // number, but reads each near pick's new bits and then the shed
          // words they changed, two dependent round trips a pick: slower
          // even for candidates near one pick (the uneven batch's SITE, one
          // job at a time, 480 ms with decrements, 461 with them only near
          // one pick, 443 with counts; the 4096^2 job 4.53 -> 4.25 ms), and
          // under intervis too (82.2 -> 84.9-91.6 ms), where a candidate
          // unmasked when last computed also holds its exact new bits.
          // SITING_SITE_FULL_AT=m decrements intervis-0 candidates near
          // fewer than m picks.
          const int mp = list_acc[e];
          if (!kIv && mp < a.full_at)
            s = T[i] - wave_lost_bits(sh, cx, cy, a, k, bw_pos, lane);
          else if (kIv && !ivk)
            s = wave_new_bits(sh, cx, cy, a, W, lane, true);
          else
            s = wave_sum_i32(eval_rows_fit(cy, roi, nwpr)
                                 ? lane_new_bits_rows<kEvalRows>(sh, cx, cy, a, lane, kWave)
                                 : lane_new_bits<kEvalBatch>(sh, cx, cy, a, W, lane, kWave, false));
          if (lane == 0) list_acc[e] = s;
        }
        __syncthreads();
        if (tid < lcnt) {
          const int acc = list_acc[tid];
          T[list_i[tid]] = acc < 0 ? 0 : acc;
          if (ivm) {
            // masked: the bound stays and the candidate is flagged masked
            const int i = list_i[tid];
            if (acc >= 0) {
              Bd[i] = acc;
              U[i] &= (uint8_t)~4u;
            } else {
              U[i] |= 4;
            }
          }
        }
      }
      __syncthreads();
      if (tid == 0) list_n = 0;
      __syncthreads();
    }
    SITE_PROF_MARK(3)
    SITE_TRACE(6)
    for (int i = tid; i < (cnt + 31) / 32; i += kLoopThreads) seen[i] = 0;
    if (stop) break;
    // before the re-evaluation reads cum (intervis rounds re-evaluate
    // nothing; the end-of-round barrier still counts every union)
    if (!kIv) union_wait();
    if (failed) {
      niter = -1;
      break;
    }
    __syncthreads();
    if (hybrid && reeval_rest) {
      // re-evaluate the stale candidates whose bound reaches tau against cum
      // (complete: the union ended with a grid barrier); a group of waves
      // shares a candidate when there are few
      Winner W0{};
      for (;;) {
        if (tid == 0) list_n = 0;
        __syncthreads();
        for (int i = tid; i < cnt; i += kLoopThreads)
          if ((U[i] & 3) == 2 && T[i] >= tau) {
            const int at = atomicAdd(&list_n, 1);
            if (at < kMaxList) list_i[at] = i;
          }
        __syncthreads();
        const int want = list_n;
        const int lcnt = want < kMaxList ? want : kMaxList;
        if (lcnt == 0) break;
        for (int e = tid; e < lcnt; e += kLoopThreads) list_acc[e] = 0;
        __syncthreads();
        // waves per candidate: kLoopWaves / lcnt rounded down to a power of two
        int gw = 1;
        while (2 * gw * lcnt <= kLoopWaves) gw *= 2;
        const int ngroups = kLoopWaves / gw;
        const int grp = wid / gw, gl = (wid - grp * gw) * kWave + lane;
        // the wavefront's candidates' positions and ids, 64 at a time, in one
        // round trip; then one candidate after another
        for (int e0 = grp; e0 < lcnt; e0 += kWave * ngroups) {
          const int em = e0 + lane * ngroups;
          uint32_t mps = 0, mid = 0;
          if (em < lcnt) {
            const int i = list_i[em];
            mps = P[i];
            mid = (uint32_t)C[i];
          }
          const int ncand = min(kWave, (lcnt - e0 + ngroups - 1) / ngroups);
          for (int j = 0; j < ncand; j++) {
            const uint32_t ps = (uint32_t)__builtin_amdgcn_readlane((int)mps, j);
            const uint32_t cid = (uint32_t)__builtin_amdgcn_readlane((int)mid, j);
            const int cx = (int)(ps >> 16), cy = (int)(ps & 0xffffu);
            const uint64_t *sh = a.sheds + (size_t)cid * S;
            const int s =
                wave_sum_i32(eval_rows_fit(cy, roi, nwpr)
                                 ? lane_new_bits_rows<kEvalRows>(sh, cx, cy, a, gl, gw * kWave)
                                 : lane_new_bits<kEvalBatch>(sh, cx, cy, a, W0, gl, gw * kWave, false));
            if (lane == 0) atomicAdd(&list_acc[e0 + j * ngroups], s);
          }
        }
        __syncthreads();
        for (int e = tid; e < lcnt; e += kLoopThreads) {
          const int i = list_i[e];
          T[i] = list_acc[e];
          U[i] = 0;
          const uint32_t ps = P[i];
          dirty[((int)(ps >> 16) / a.cs * Q + ((int)(ps & 0xffffu) / a.cs) % Q) / G] = 1;
        }
        __syncthreads();
        if (want <= kMaxList) break;
      }
      // the update walk of the next round starts from an empty list
      if (tid == 0) list_n = 0;
    }
    SITE_PROF_MARK(4)
    SITE_TRACE(7)
    for (int ul = wid; ul < upw; ul += kLoopWaves)
      if (dirty[ul]) unit_top2(ul);
    __syncthreads();
    for (int ul = tid; ul < upw; ul += kLoopThreads) dirty[ul] = 0;
    parity ^= 1;
    publish(parity);
    if (ivm) publish_cells(parity);
    SITE_PROF_MARK(5)
    SITE_TRACE(3)
    if (!grid_sync(&a.st->barrier, bar_target, (unsigned)G, a.spin_limit)) {
      niter = -1;
      break;
    }
  }
  SITE_PROF(if (tid == 0) {
    pf_acc[8] = 1;
    for (int q = 0; q < kProfWords; q++)
      if (pf_acc[q]) atomicAdd(&g_site_prof[q], pf_acc[q]);
  })

  // sum_u: each thread counted its own list entries
  {
    unsigned long long v = (unsigned long long)my_u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    if (lane == 0 && v) atomicAdd(&a.st->sum_u, v);
  }
  if (g == 0 && tid == 0) {
    a.st->nused = nused;
    a.st->niter = niter;
    a.st->rounds = ph_round;
    a.st->cumarea = (unsigned long long)cumarea;
  }
}

template <int kIvMode>
__global__ __launch_bounds__(kLoopThreads) void site_loop(SiteArgs a) {
  site_loop_body<kIvMode>(a, blockIdx.x, gridDim.x);
}

// Work queue of jobs.  The grid is cut into slots of `slot` workgroups; the
// first group of a slot takes the next job index with an atomic and the slot
// runs that job's loop (a job planned for G < slot groups leaves the rest
// idle), then takes another, until the queue is empty.  The job index is
// double-buffered by the slot's turn so a late reader never sees the next
// turn's value.
// Slots of workgroups pop jobs from their class's queue (plan layout in
// site_queue_plan_ints), then from their class's fallback queue, and run
// each job's loop with the slot's first G workgroups (a fallback job is
// planned for fewer groups than the slot has); jobt[j] receives the loop's
// start and end (device clock).
template <int kIvMode>
__global__ __launch_bounds__(kLoopThreads) void site_loop_queue(
    const SiteArgs *__restrict__ jobs, const int *plan, unsigned int *ctl,
    unsigned long long *jobt) {
  __shared__ int job;
  const int nslots = plan[0], ncls = plan[1];
  const int *start = plan + 2, *size = start + nslots, *cls = size + nslots;
  const int *off = cls + nslots, *order = off + ncls + 1, *fb = order + off[ncls];
  int sl = 0;  // slots are listed by first workgroup
  while (sl + 1 < nslots && (int)blockIdx.x >= start[sl + 1]) sl++;
  const int g = blockIdx.x - start[sl], slot = size[sl], c = cls[sl];
  unsigned int *queue = ctl + c;
  unsigned int *slot_job = ctl + ncls + 2 * sl;
  unsigned int *slot_bar = ctl + ncls + 2 * nslots + sl;
  const unsigned int qlen = (unsigned int)(off[c + 1] - off[c]);
  const int f = fb[c];
  constexpr unsigned int kDone = 0xffffffffu;
  unsigned int target = 0;
  for (int turn = 0;; turn++) {
    unsigned int *cell = &slot_job[turn & 1];
    if (g == 0 && threadIdx.x == 0) {
      // the job's position in order[]: the next of the class's queue, else
      // of its fallback class's queue, else none
      const unsigned int q = atomicAdd(queue, 1u);
      unsigned int pos = q < qlen ? (unsigned int)off[c] + q : kDone;
      if (pos == kDone && f >= 0) {
        const unsigned int q2 = atomicAdd(ctl + f, 1u);
        if (q2 < (unsigned int)(off[f + 1] - off[f])) pos = (unsigned int)off[f] + q2;
      }
      __hip_atomic_store(cell, pos, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (!grid_sync(slot_bar, target, (unsigned)slot)) break;
    if (threadIdx.x == 0)
      job = (int)__hip_atomic_load(cell, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const unsigned int pos = (unsigned int)job;
    if (pos == kDone) break;
    // a wave-uniform job index lets the job's arguments be scalar loads
    const int j = __builtin_amdgcn_readfirstlane(order[pos]);
    const SiteArgs &a = jobs[j];
    const unsigned long long t0 = wall_clock64();
    if (g < a.G) site_loop_body<kIvMode>(a, g, a.G);
    __syncthreads();
    if (g == 0 && threadIdx.x == 0) {
      jobt[2 * j] = t0;
      jobt[2 * j + 1] = wall_clock64();
    }
  }
}

// ---- host-side plan --------------------------------------------------------

struct Plan {
  int cs, ncr, Q, nunits, G, upw, kpu, nbuckets, nchunks, cap, delta_lds, off_lds;
  size_t lds, nbwords;
  int o_delta, o_ukey, o_upos, o_dirty, o_off, o_c, o_pos, o_t, o_used, o_seen;
  int iv, o_b, o_dcur, o_dprev, o_dbits;  // intervis multi-pick (o_*: -1 if off)
};

// Column classes per cell row.  More classes spread a pick's update window
// over more workgroups but lengthen the selection scan; three pay only when
// a pick's expected update list pi*(2roi)^2 * nsheds/n^2 reaches 3000
// (measured).  A job's tuning (1 to 8; the tuner tries 2, 3 and 4)
// replaces the rule; SITING_SITE_Q overrides both.
int choose_q(int n, int roi, int nsheds, int tuned = 0) {
  const char *e = std::getenv("SITING_SITE_Q");
  if (e && *e) {
    const int q = std::atoi(e);
    return q < 1 ? 1 : (q > 8 ? 8 : q);
  }
  if (tuned >= 1 && tuned <= 8) return tuned;
  const double list = 3.14159265358979 * 4.0 * (double)roi * roi * (double)nsheds /
                      ((double)n * (double)n);
  return list >= 3000.0 ? 3 : 2;
}

// Update candidates whose area is below tau = head - head / 2^s (head: the
// round's largest key's area) are not decremented but kept as stale bounds
// (s = 3: over the uneven 64-job batch's jobs one by one, SITE summed 502,
// 480, 493 and 494 ms with s = 2, 3, 4 and 5, and the 1000^2 roi-50 job
// took 1.65, 1.69, 2.24 and 2.24 ms; a job's tuning may choose 1..20;
// SITING_SITE_TAU_SHIFT overrides, 0 turning it off).
int tau_shift(int tuned = 0) {
  const char *e = std::getenv("SITING_SITE_TAU_SHIFT");
  const int t = e && *e ? std::atoi(e) : (tuned >= 1 && tuned <= 20 ? tuned : 3);
  return t < 0 ? 0 : (t > 20 ? 20 : t);
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

// The stop rule 100.0 * cumarea / (n * n) > 95 (siting.cpp:601-604), in
// double arithmetic, is monotone in cumarea: the least area that meets it,
// or n * n + 1 when none does.
long long stop_area(int n) {
  const double npts = (double)n * (double)n;
  long long lo = 0, hi = (long long)n * n + 1;  // the answer is in [lo, hi]
  while (lo < hi) {
    const long long mid = lo + (hi - lo) / 2;
    if (100.0 * (double)mid / npts > 95.0)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

Plan make_plan(int n, int roi, int nsheds, int groups_cap, int fixed_groups = 0,
               int q_tuned = 0, int intervis = 0) {
  Plan p{};
  p.iv = intervis;
  p.cs = cell_edge(n, roi);
  p.ncr = (n + p.cs - 1) / p.cs;
  p.Q = choose_q(n, roi, nsheds, q_tuned);
  if (p.Q > p.ncr) p.Q = p.ncr;
  p.nunits = p.ncr * p.Q;
  // units in a pick's 4*roi window, times workgroups per such unit, and at
  // least enough workgroups that each one's candidates fit its LDS lists
  const int touched = ((4 * roi) / p.cs + 2) * p.Q;
  const int fit = (nsheds + kFitPerGroup - 1) / kFitPerGroup;
  int G = fixed_groups > 0 ? fixed_groups : (touched > fit ? touched : fit);
  G = G < p.nunits ? G : p.nunits;
  G = G < groups_cap ? G : groups_cap;
  G = G < kMaxGroups ? G : kMaxGroups;
  p.G = G < 1 ? 1 : G;
  p.upw = (p.nunits + p.G - 1) / p.G;
  p.kpu = (p.ncr + p.Q - 1) / p.Q;
  p.nbuckets = p.G * p.upw * p.kpu;
  p.nchunks = (p.nbuckets + kScanThreads - 1) / kScanThreads;
  const int nwpr = (2 * roi + 1 + 63) / 64;
  const size_t delta_bytes = sizeof(uint64_t) * (size_t)(2 * roi + 1) * (nwpr + 1);
  p.nbwords = (size_t)kMaxBatch * (2 * roi + 1) * (nwpr + 1);
  p.delta_lds = delta_bytes <= 40 * 1024;
  size_t o = 0;
  p.o_delta = (int)o;
  if (p.delta_lds) o += delta_bytes;
  p.o_ukey = (int)o;
  o += sizeof(uint64_t) * 3 * (size_t)p.upw;
  p.o_upos = (int)o;
  o += sizeof(uint32_t) * 2 * (size_t)p.upw;
  p.o_dirty = (int)o;
  o = align_up(o + (size_t)p.upw, 16);
  p.o_off = (int)o;
  const size_t off_bytes = sizeof(int32_t) * (size_t)p.upw * (p.kpu + 1);
  p.off_lds = o + off_bytes <= kLdsBudget / 2;
  if (p.off_lds) o += off_bytes;
  o = align_up(o, 16);
  p.o_b = p.o_dcur = p.o_dprev = p.o_dbits = -1;
  if (intervis) {
    // dirty cell lists of this round and the last, and a bit per owned cell
    p.o_dcur = (int)o;
    o += 4 * (size_t)kMaxDirtyCells;
    p.o_dprev = (int)o;
    o += 4 * (size_t)kMaxDirtyCells;
    p.o_dbits = (int)o;
    o = align_up(o + 4 * (((size_t)p.upw * p.kpu + 31) / 32), 16);
  }
  const size_t fixed = o;
  const size_t room = kLdsBudget > fixed ? kLdsBudget - fixed : 0;
  // id, pos, area, used, seen bit (rounded up), and intervis area bounds
  int cap = (int)(room / (intervis ? 18 : 14));
  cap = cap > 16 ? (cap - 16) / 16 * 16 : 0;
  p.cap = cap;
  p.o_c = (int)o;
  o += 4 * (size_t)cap;
  p.o_pos = (int)o;
  o += 4 * (size_t)cap;
  p.o_t = (int)o;
  o += 4 * (size_t)cap;
  if (intervis) {
    p.o_b = (int)o;
    o += 4 * (size_t)cap;
  }
  p.o_used = (int)o;
  o += (size_t)cap;
  o = align_up(o, 16);
  p.o_seen = (int)o;
  o += 4 * (((size_t)cap + 31) / 32);
  p.lds = align_up(o, 16);
  return p;
}

struct Scratch {
  int32_t *bcount, *bstart, *chunk, *sid, *sT;
  uint32_t *spos;
  uint8_t *sused;
  uint64_t *keys;
  uint32_t *kpos;
  uint64_t *nb;
  unsigned int *sseen;
  uint64_t *phash;
  int32_t *sB;
  uint64_t *cbnd;
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
  s.chunk = (int32_t *)take(4 * ((nb_max + kScanThreads - 1) / kScanThreads));
  s.sid = (int32_t *)take(4 * (size_t)nsheds);
  s.spos = (uint32_t *)take(4 * (size_t)nsheds);
  s.sT = (int32_t *)take(4 * (size_t)nsheds);
  s.sused = (uint8_t *)take((size_t)nsheds);
  s.nb = (uint64_t *)take(8 * p.nbwords);
  // global-mode seen bitmaps: one slice of ceil(nsheds/32)+1 words per group
  s.sseen = (unsigned int *)take(4 * ((size_t)nsheds / 32 + 2) * kMaxGroups);
  s.keys = (uint64_t *)take(8 * 6 * (size_t)p.nunits);
  s.kpos = (uint32_t *)take(4 * 4 * (size_t)p.nunits);
  s.phash = (uint64_t *)take(8 * 2 * (size_t)kMaxGroups);
  if (p.iv) {
    s.sB = (int32_t *)take(4 * (size_t)nsheds);
    s.cbnd = (uint64_t *)take(8 * 2 * 4 * (size_t)p.ncr * p.ncr);
  }
  if (out) *out = s;
  return o;
}

}  // namespace

// The plan's group count depends on the device and its column classes on
// the job's tuning; the scratch size is taken for the largest group count,
// which bounds every smaller one, and the largest over the built-in rule and
// both tunable class counts.
size_t site_scratch_bytes(int nrows, int roi, int nsheds, int intervis) {
  size_t m = 0;
  for (int q : {0, 2, 3})
    m = std::max(m, scratch_layout(make_plan(nrows, roi, nsheds, kMaxGroups, 0, q, intervis),
                                   nsheds, nullptr, nullptr));
  return m;
}

int site_key_rows(int nrows, int roi) { return 1; }

siting_status site_make_args(const Geometry &g, const int32_t *obs,
                             const uint64_t *sheds, const SiteBuffers &b, int slot,
                             SiteArgs *out, size_t *lds) {
  const int n = g.p.nrows, roi = g.p.roi, nsheds = g.d.nwanted;
  const Plan p = make_plan(n, roi, nsheds, slot, slot, g.t.site_q, g.p.intervis);
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
  a.nb = sc.nb;
  a.sseen = sc.sseen;
  a.phash = sc.phash;
  a.sB = sc.sB;
  a.cbnd = sc.cbnd;
  a.o_b = p.o_b;
  a.o_dcur = p.o_dcur;
  a.o_dprev = p.o_dprev;
  a.o_dbits = p.o_dbits;
  {
    // intervis rounds of several picks (SITING_SITE_IV_MULTI=0: one per round)
    const char *e = std::getenv("SITING_SITE_IV_MULTI");
    a.iv_multi = g.p.intervis && !(e && *e && std::atoi(e) == 0);
    // rule: 16 for roi up to 40 (measured on a 7594^2 roi-25 request: SITE
    // 850, 614 and 555 ms with 2, 8 and 16), 8 above 75 (the tuner's roi-100
    // to 200 classes: 5-12 % faster than none), none between (a 7857^2
    // roi-50 request: 255 ms with none, 310-416 with 2-16); a job's tuning
    // replaces the rule, SITING_SITE_WALK_EVALS overrides both
    int ev = roi <= 40 ? 16 : roi > 75 ? 8 : 0;
    if (g.t.site_walk_evals >= 1 && g.t.site_walk_evals <= kMaxBatch + 1)
      ev = g.t.site_walk_evals - 1;
    const char *w = std::getenv("SITING_SITE_WALK_EVALS");
    if (w && *w) ev = std::max(0, std::min(kMaxBatch, std::atoi(w)));
    a.walk_evals = ev;
    // rekeys per round: 16 (intervis 0, measured on roi 25-150 jobs of 894^2
    // to 4096^2: 10-33 % less SITE time than none, 2 to 4 behind; with the
    // computations' pick groups matched to their pick count, 16 against 8:
    // the uneven batch's SITE launch 65.1 -> 64.5 ms, a 2897^2 roi-100 job
    // 35.7 -> 33.7 ms, the 4096^2 job flat); under intervis only where the
    // walk's exact computations pay, by the same choice (the 2048^2 roi-80
    // job: 2867 -> 1560 rounds, SITE 137 -> 113 ms with 8; 91.0 -> 88.7 ms
    // with 16 over 8; a 7857^2 roi-50 request, where they are off: 2 %
    // slower with them).  SITING_SITE_NEAR_EVALS overrides.
    int ne = g.p.intervis ? (a.iv_multi && ev > 0 ? 16 : 0) : 16;
    const char *ne_env = std::getenv("SITING_SITE_NEAR_EVALS");
    if (ne_env && *ne_env) ne = std::max(0, std::min(kMaxBatch, std::atoi(ne_env)));
    a.near_evals = g.p.intervis && !a.iv_multi ? 0 : ne;
    // intervis-0 update candidates within 2*roi of this many of a round's
    // picks are counted in full against cum, those near fewer decremented
    // by the bits each pick took: counts throughout (SITING_SITE_FULL_AT
    // overrides)
    int fa = 1;
    const char *fa_env = std::getenv("SITING_SITE_FULL_AT");
    if (fa_env && *fa_env) fa = std::max(1, std::atoi(fa_env));
    a.full_at = fa;
  }
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
  a.delta_lds = p.delta_lds;
  a.o_c = p.o_c;
  a.o_pos = p.o_pos;
  a.o_t = p.o_t;
  a.o_used = p.o_used;
  a.o_seen = p.o_seen;
  a.tau_shift = tau_shift(g.t.site_tau_shift);
  a.stop_area = stop_area(n);
  // (a test forces the give-up path with a tiny limit, every kernel still
  // exiting normally)
  const char *sl = std::getenv("SITING_SITE_SPIN_LIMIT");
  a.spin_limit = sl && *sl && std::atoi(sl) >= 0 ? (unsigned)std::atoi(sl) : kSpinLimit;
  const char *hf = std::getenv("SITING_SITE_HASH_FAULT");
  a.hash_fault = hf && *hf && std::atoi(hf) > 0 ? std::atoi(hf) : 0;
  *out = a;
  if (lds) *lds = p.lds;
  return SITING_OK;
}

siting_status site_launch_prep(const SiteArgs &a, hipStream_t s) {
  const int tb = 256, nbk = (a.nsheds + tb - 1) / tb;
  const int nchunks = (a.nbuckets + kScanThreads - 1) / kScanThreads;
  hipLaunchKernelGGL(site_zero_counts, dim3((a.nbuckets + 255) / 256), dim3(256), 0, s,
                     a.bcount, a.nbuckets);
  hipLaunchKernelGGL(site_bucket_count, dim3(nbk), dim3(tb), 0, s, a);
  hipLaunchKernelGGL(site_scan_chunks, dim3(nchunks), dim3(kScanThreads), 0, s, a);
  hipLaunchKernelGGL(site_scan_apply, dim3(nchunks), dim3(kScanThreads), 0, s, a);
  hipLaunchKernelGGL(site_bucket_scatter, dim3(nbk), dim3(tb), 0, s, a);
  hipLaunchKernelGGL(site_first_pass, dim3((a.nsheds + 3) / 4), dim3(256), 0, s, a);
  SITING_HIP(hipGetLastError());
  return SITING_OK;
}

int site_loop_groups(const Geometry &g, int device_cus, int groups_cap) {
  int cap = device_cus / 2 > 0 ? device_cus / 2 : 1;
  if (groups_cap > 0 && groups_cap < cap) cap = groups_cap;
  return make_plan(g.p.nrows, g.p.roi, g.d.nwanted, cap, 0, g.t.site_q, g.p.intervis).G;
}

siting_status launch_site(const Geometry &g, const int32_t *obs,
                          const uint64_t *sheds, const SiteBuffers &b,
                          int device, hipStream_t s, int groups_cap) {
  int cus = 0;
  SITING_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                   device));
  // The grid barrier needs every workgroup of the loop resident at once; a
  // grid of at most half the CUs (one group per CU) is, since any other
  // kernel holding CUs when it starts does not wait on it and retires.  A
  // plain launch is used: a cooperative launch reserves a device-wide
  // resource that a second process on the same GPU cannot get.
  int cap = cus / 2 > 0 ? cus / 2 : 1;
  if (groups_cap > 0 && groups_cap < cap) cap = groups_cap;
  if (const char *e = std::getenv("SITING_SITE_GROUPS"))  // tuning override
    if (std::atoi(e) > 0 && std::atoi(e) < cap) cap = std::atoi(e);
  const Plan p = make_plan(g.p.nrows, g.p.roi, g.d.nwanted, cap, 0, g.t.site_q, g.p.intervis);
  SiteArgs a;
  size_t lds = 0;
  siting_status st = site_make_args(g, obs, sheds, b, p.G, &a, &lds);
  if (st != SITING_OK) return st;
  if ((st = site_launch_prep(a, s)) != SITING_OK) return st;
  if (a.intervis)
    hipLaunchKernelGGL(site_loop<1>, dim3(a.G), dim3(kLoopThreads), lds, s, a);
  else
    hipLaunchKernelGGL(site_loop<0>, dim3(a.G), dim3(kLoopThreads), lds, s, a);
  SITING_HIP(hipGetLastError());
  return SITING_OK;
}

__global__ void site_zero_ctl(unsigned int *ctl, int words) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < words) ctl[i] = 0;
}

size_t site_queue_ctl_words(int nslots, int nclasses) {
  return (size_t)nclasses + 3 * (size_t)nslots;
}

size_t site_queue_plan_ints(int nslots, int nclasses, int njobs) {
  return 2 + 3 * (size_t)nslots + (size_t)nclasses + 1 + (size_t)njobs + (size_t)nclasses;
}

siting_status site_launch_queue(const SiteArgs *jobs_dev, const int *plan_dev,
                                int nslots, int nclasses, int total_groups, size_t lds,
                                unsigned int *ctl, unsigned long long *jobt, int iv_mode,
                                hipStream_t s) {
  const int words = (int)site_queue_ctl_words(nslots, nclasses);
  hipLaunchKernelGGL(site_zero_ctl, dim3((words + 255) / 256), dim3(256), 0, s, ctl,
                     words);
#define SITING_QUEUE(M)                                                                   \
  hipLaunchKernelGGL(site_loop_queue<M>, dim3(total_groups), dim3(kLoopThreads), lds, s, \
                     jobs_dev, plan_dev, ctl, jobt)
  if (iv_mode == 0) SITING_QUEUE(0);
  else if (iv_mode == 1) SITING_QUEUE(1);
  else SITING_QUEUE(2);
#undef SITING_QUEUE
  SITING_HIP(hipGetLastError());
  return SITING_OK;
}

}  // namespace siting

#ifdef SITING_PHASE_PROFILE
// Profiling build only: copies the greedy loop's phase counters (n words,
// at most 32, in the order listed at g_site_prof) and optionally clears them.
extern "C" siting_status siting_phase_profile_read(unsigned long long *out, int n, int reset) {
  if (!out || n < 1 || n > siting::kProfWords) return SITING_EINVAL;
  SITING_HIP(hipDeviceSynchronize());
  SITING_HIP(hipMemcpyFromSymbol(out, HIP_SYMBOL(siting::g_site_prof),
                                 sizeof(unsigned long long) * (size_t)n));
  if (reset) {
    unsigned long long z[siting::kProfWords] = {0};
    SITING_HIP(hipMemcpyToSymbol(HIP_SYMBOL(siting::g_site_prof), z, sizeof(z)));
  }
  return SITING_OK;
}

// Profiling build only: copies the round trace (kTraceMarks x kTraceRounds
// x kTraceGroups words, 8 x 256 x 128, in that order; 0 where no
// workgroup stored) and clears it.
extern "C" siting_status siting_phase_trace_read(unsigned long long *out, size_t n) {
  constexpr size_t kWords = (size_t)siting::kTraceMarks * siting::kTraceRounds *
                            siting::kTraceGroups;
  if (!out || n != kWords) return SITING_EINVAL;
  SITING_HIP(hipDeviceSynchronize());
  SITING_HIP(hipMemcpyFromSymbol(out, HIP_SYMBOL(siting::g_site_trace),
                                 sizeof(unsigned long long) * kWords));
  std::vector<unsigned long long> z(kWords, 0);
  SITING_HIP(hipMemcpyToSymbol(HIP_SYMBOL(siting::g_site_trace), z.data(),
                               sizeof(unsigned long long) * kWords));
  return SITING_OK;
}
#endif
