// findmax.hip -- FINDMAX stage: the nwpb best vix points of every block
// become tentative observers (siting.cpp:200-241).
//
// One wavefront per block.  The reference scans the block row-major and
// replaces its running best iff (v1 < v2) or (v1 == v2 and h1 < h2), with
// h = x*(x+y)*0x41041041 in wrapping int32 (siting.cpp:213-229); the winner
// is therefore the maximum of (vix, hash) and, among equal pairs, the first
// cell in scan order.  Each lane folds its strided subset into a
// (vix, hash ^ 0x80000000, scan index) best and the wavefront reduces it
// with that order.  The winner's vix is then zeroed (siting.cpp:231), and
// the next round rescans; reads go through device-coherent atomic loads so
// a round sees the previous round's zero.  (The CUDA counterpart,
// siting.cu:164-224, lets threads read outside their block; not followed.)
#include "common.hpp"

namespace siting {

namespace {

constexpr int kWavesPerGroup = 4;

__device__ __forceinline__ uint32_t load_vix_word(const uint8_t *vix, size_t idx) {
  const uint32_t *w = (const uint32_t *)(vix + (idx & ~(size_t)3));
  return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kWave *kWavesPerGroup) void findmax_kernel(
    uint8_t *__restrict__ vix, int32_t *__restrict__ obs, int n, float bs,
    int nb, int nwpb) {
  const int lane = threadIdx.x & (kWave - 1);
  const int blk = blockIdx.x * kWavesPerGroup + (threadIdx.x >> 6);
  if (blk >= nb * nb) return;  // whole wavefronts exit together
  const int bx = blk / nb, by = blk - (blk / nb) * nb;
  const int x0 = (int)(bs * (float)bx);
  const int x1 = min((int)(bs * (float)(bx + 1)), n);
  const int y0 = (int)(bs * (float)by);
  const int y1 = min((int)(bs * (float)(by + 1)), n);
  const int width = y1 - y0;
  const int np = (x1 - x0) * width;
  if (np <= 0) return;  // derive() admits no empty block; kept as a guard

  for (int round = 0; round < nwpb; round++) {
    // best = (hi: vix<<32 | hash^0x80000000, j: scan index), larger hi wins,
    // equal hi -> smaller j
    uint64_t hi = 0;
    int j = 0x7fffffff;
    for (int q = lane; q < np; q += kWave) {
      const int x = x0 + q / width;
      const int y = y0 + q % width;
      const size_t idx = (size_t)x * n + y;
      const uint32_t word = load_vix_word(vix, idx);
      const uint32_t v = (word >> (8 * (idx & 3))) & 0xffu;
      const uint64_t h = (uint64_t)v << 32 |
                         (uint32_t)((uint32_t)block_hash(x, y) ^ 0x80000000u);
      if (h > hi || (h == hi && q < j)) {
        hi = h;
        j = q;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const uint64_t oh = __shfl_xor(hi, off, kWave);
      const int oj = __shfl_xor(j, off, kWave);
      if (oh > hi || (oh == hi && oj < j)) {
        hi = oh;
        j = oj;
      }
    }
    const int x = x0 + j / width;
    const int y = y0 + j % width;
    const size_t idx = (size_t)x * n + y;
    if (lane == 0) {
      obs[(size_t)blk * nwpb + round] = (int32_t)idx;
      uint32_t *w = (uint32_t *)(vix + (idx & ~(size_t)3));
      atomicAnd(w, ~(0xffu << (8 * (idx & 3))));
    }
    // the zeroing completes (s_waitcnt) before the next round's loads issue
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
  }
}

// Blocks of at most 64*PER cells: every lane keeps the (vix, hash) keys of
// its cells in registers, read once; a round's winner is zeroed in the
// owning lane's copy (and with a byte store in memory), so later rounds
// never re-read vix.
template <int PER>
__global__ __launch_bounds__(kWave *kWavesPerGroup) void findmax_regs_kernel(
    uint8_t *__restrict__ vix, int32_t *__restrict__ obs, int n, float bs,
    int nb, int nwpb) {
  const int lane = threadIdx.x & (kWave - 1);
  const int blk = blockIdx.x * kWavesPerGroup + (threadIdx.x >> 6);
  if (blk >= nb * nb) return;  // whole wavefronts exit together
  const int bx = blk / nb, by = blk - (blk / nb) * nb;
  const int x0 = (int)(bs * (float)bx);
  const int x1 = min((int)(bs * (float)(bx + 1)), n);
  const int y0 = (int)(bs * (float)by);
  const int y1 = min((int)(bs * (float)(by + 1)), n);
  const int width = y1 - y0;
  const int np = (x1 - x0) * width;
  if (np <= 0) return;  // derive() admits no empty block; kept as a guard

  uint64_t key[PER];  // vix << 32 | hash ^ 0x80000000; cell q = lane + 64*t
#pragma unroll
  for (int t = 0; t < PER; t++) {
    const int q = lane + kWave * t;
    key[t] = 0;
    if (q < np) {
      const int x = x0 + q / width, y = y0 + q % width;
      key[t] = (uint64_t)vix[(size_t)x * n + y] << 32 |
               (uint32_t)((uint32_t)block_hash(x, y) ^ 0x80000000u);
    }
  }
  for (int round = 0; round < nwpb; round++) {
    uint64_t hi = 0;
    int j = 0x7fffffff;  // cells past np keep j > np, so any real cell wins
#pragma unroll
    for (int t = 0; t < PER; t++) {
      const int q = lane + kWave * t;
      if (q < np && (key[t] > hi || (key[t] == hi && q < j))) {
        hi = key[t];
        j = q;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const uint64_t oh = __shfl_xor(hi, off, kWave);
      const int oj = __shfl_xor(j, off, kWave);
      if (oh > hi || (oh == hi && oj < j)) {
        hi = oh;
        j = oj;
      }
    }
    if ((j & (kWave - 1)) == lane) {  // owner: this cell's vix becomes 0
#pragma unroll
      for (int t = 0; t < PER; t++)
        if (lane + kWave * t == j) key[t] &= 0xffffffffull;
    }
    if (lane == 0) {
      const int x = x0 + j / width, y = y0 + j % width;
      const size_t idx = (size_t)x * n + y;
      obs[(size_t)blk * nwpb + round] = (int32_t)idx;
      vix[idx] = 0;  // siting.cpp:231
    }
  }
}

}  // namespace

siting_status launch_findmax(const Geometry &g, uint8_t *vix, int32_t *obs,
                             hipStream_t s) {
  const int nb = g.d.nblocks;
  const int nblk = nb * nb;
  const int groups = (nblk + kWavesPerGroup - 1) / kWavesPerGroup;
  // largest block: edge (int)(bs*(b+1)) - (int)(bs*b) <= ceil(bs) + 1
  const int edge = (int)g.d.blocksize + 2;
  const long long maxcells = (long long)edge * edge;
  const dim3 grid(groups), block(kWave * kWavesPerGroup);
  if (maxcells <= 2 * kWave)
    hipLaunchKernelGGL(findmax_regs_kernel<2>, grid, block, 0, s, vix, obs, g.p.nrows,
                       g.d.blocksize, nb, g.d.nwpb);
  else if (maxcells <= 8 * kWave)
    hipLaunchKernelGGL(findmax_regs_kernel<8>, grid, block, 0, s, vix, obs, g.p.nrows,
                       g.d.blocksize, nb, g.d.nwpb);
  else
    hipLaunchKernelGGL(findmax_kernel, grid, block, 0, s, vix, obs, g.p.nrows,
                       g.d.blocksize, nb, g.d.nwpb);
  SITING_HIP(hipGetLastError());
  return SITING_OK;
}

}  // namespace siting
