// vix.hip -- VIX stage: approximate visibility index of every terrain point
// (siting.cpp:54-198; the CUDA counterpart siting.cu:51-158 divides where the
// sequential reference multiplies by a rounded double reciprocal, so it is
// not followed).
//
// Work decomposition: a workgroup owns a 16x16 tile of observer points and
// walks all (point, target) pairs of the tile, so lanes of one wavefront
// test different targets of the same few points.  Target k of point p needs
// the generator state after 2k+1 and 2k+2 steps from seed p; a host-built
// jump-ahead table (A_m, C_m) with r_m = (A_m*p + C_m) mod 2^31 lets every
// lane draw its own target without the serial chain.  Visible targets are
// counted with LDS atomics; the tile's vix bytes are written at the end.
//
// Tiles are assigned XCD-major: workgroup b runs on XCD b%8, and each XCD
// gets one contiguous band of tiles so the DEM window it gathers from stays
// in its own L2.
#include <cstdlib>

#include "common.hpp"

namespace siting {

namespace {

constexpr int kTile = 16;           // tile edge (points)
constexpr int kThreads = kTile * kTile;
// The tile's DEM window with a halo of kHalo cells stays in LDS: samples at
// steps up to kHalo from an observer of the tile (the minor offset is never
// larger than the step) and the observers' own elevations are read from it.
template <int kHalo>
constexpr int win_edge() { return kTile + 2 * kHalo; }

// siting.cpp:157,160, evaluated in float in the reference's order:
// ((float)(2roi) + 0.99999f) * (float)r / 2^31, truncated.
__device__ __forceinline__ int target_offset(float span, uint32_t r) {
  float prod = span * (float)(int)r;
  float q = prod / 2147483648.0f;
  return (int)q;
}

template <int kGroup, int kHalo>
__global__ __launch_bounds__(kThreads) void vix_kernel(
    const uint16_t *__restrict__ elevs, uint8_t *__restrict__ vix,
    const float *__restrict__ recip, const uint32_t *__restrict__ jump,
    int n, int roi, int oht, int tht, int ntests, int tiles_side,
    int tiles_per_xcd) {
  extern __shared__ uint32_t lds[];
  uint32_t *jl = lds;                      // 4*ntests: A1 C1 A2 C2 per target
  int *count = (int *)(lds + 4 * ntests);  // kThreads
  constexpr int kWin = win_edge<kHalo>();
  uint16_t *win = (uint16_t *)(count + kThreads);  // kWin x kWin

  const int b = blockIdx.x;
  const int tile = (b & 7) * tiles_per_xcd + (b >> 3);
  if (tile >= tiles_side * tiles_side) return;
  const int tx0 = (tile / tiles_side) * kTile;
  const int ty0 = (tile % tiles_side) * kTile;

  for (int i = threadIdx.x; i < 4 * ntests; i += kThreads) jl[i] = jump[i];
  count[threadIdx.x] = 0;
  const int wx0 = tx0 - kHalo, wy0 = ty0 - kHalo;
  for (int i = threadIdx.x; i < kWin * kWin; i += kThreads) {
    const int wx = wx0 + i / kWin, wy = wy0 + i % kWin;
    // cells outside the terrain are never read (samples lie between an
    // observer and a target, both inside)
    win[i] = wx >= 0 && wx < n && wy >= 0 && wy < n ? elevs[wx * n + wy] : 0;
  }
  __syncthreads();

  const float span = (float)(2 * roi) + 0.99999f;
  const int rr = roi * roi;
  const int total = kThreads * ntests;
  // pair q = lp * ntests + k, advanced by kThreads per step
  int lp = threadIdx.x / ntests;
  int k = threadIdx.x - lp * ntests;
  const int step_lp = kThreads / ntests;
  const int step_k = kThreads - step_lp * ntests;

  for (int q = threadIdx.x; q < total; q += kThreads) {
    const int px = tx0 + (lp >> 4);
    const int py = ty0 + (lp & 15);
    bool vis = false;
    if (px < n && py < n) {
      const uint32_t p = (uint32_t)(px * n + py);
      const uint32_t r1 = (jl[4 * k] * p + jl[4 * k + 1]) & 0x7fffffffu;
      const uint32_t r2 = (jl[4 * k + 2] * p + jl[4 * k + 3]) & 0x7fffffffu;
      int tx = target_offset(span, r1) + (px - roi);
      int ty = target_offset(span, r2) + (py - roi);
      int ddx = tx - px, ddy = ty - py;
      if (ddx * ddx + ddy * ddy > rr) {  // siting.cpp:162-166
        tx = px + ddx / 3;
        ty = py + ddy / 3;
      }
      if (tx >= 0 && tx < n && ty >= 0 && ty < n) {
        const int dx = tx - px, dy = ty - py;
        const int adx = dx < 0 ? -dx : dx, ady = dy < 0 ? -dy : dy;
        if (adx <= 1 && ady <= 1) {
          vis = true;
        } else {  // siting.cpp:87-133
          const bool alx = adx >= ady;
          const int major = alx ? dx : dy;
          const int minor = alx ? dy : dx;
          const int len = alx ? adx : ady;
          float rec = recip[len];
          if (major < 0) rec = -rec;  // (float)(1.0/-d) == -(float)(1.0/d)
          const float slope = (float)minor * rec;
          const int dir = major > 0 ? 1 : -1;
          // The reference samples id = dir*1, dir*2, dir*4, ... while |id| <
          // len and stops at the first sample above the sight line, so the
          // target is visible iff no sample is above it: the samples can be
          // read in any order.  kGroup at a time, the first group in flight
          // together with the observer and target elevations.
          const int oz_raw = (int)win[(px - wx0) * kWin + (py - wy0)];
          const int tz_raw = (int)elevs[tx * n + ty];
          float zz[kGroup];
          auto load_group = [&](int s0) {
#pragma unroll
            for (int u = 0; u < kGroup; u++) {
              const int sv = s0 << u;
              const int id = dir * sv;
              const int off = round_half_away(slope * (float)id);
              const int sx = alx ? px + id : px + off;
              const int sy = alx ? py + off : py + id;
              zz[u] = sv >= len ? 0.0f
                      : sv <= kHalo ? (float)win[(sx - wx0) * kWin + (sy - wy0)]
                                    : (float)elevs[sx * n + sy];
            }
          };
          load_group(1);
          const int oz = oz_raw + oht;
          const int tz = tz_raw + tht;
          const float zs = (float)(tz - oz) * rec;
          const float foz = (float)oz;
          vis = true;
          for (int s0 = 1;;) {
#pragma unroll
            for (int u = 0; u < kGroup; u++) {
              const int sv = s0 << u;
              const float fid = (float)(dir * sv);
              const float lim = foz + fid * zs;
              if (sv < len && zz[u] > lim) vis = false;
            }
            s0 <<= kGroup;
            if (!vis || s0 >= len) break;
            load_group(s0);
          }
        }
      }
    }
    // visible pairs per point: the wavefront's pairs are consecutive, so a
    // point's pairs are a run of lanes; one atomic per run
    {
      const uint64_t vm = __ballot(vis);
      const int lane = threadIdx.x & (kWave - 1);
      const int lp_first = __builtin_amdgcn_readfirstlane(lp);
      const int lp_last = __shfl(lp, kWave - 1, kWave);
      for (int j = lp_first; j <= lp_last; j++) {
        const uint64_t run = __ballot(lp == j);
        const int c = __popcll(vm & run);
        if (c && lane == (int)__builtin_ctzll(run)) atomicAdd(&count[j], c);
      }
    }
    lp += step_lp;
    k += step_k;
    if (k >= ntests) {
      k -= ntests;
      lp += 1;
    }
  }
  __syncthreads();

  const int lx = tx0 + (threadIdx.x >> 4);
  const int ly = ty0 + (threadIdx.x & 15);
  if (lx < n && ly < n) {
    const float v = (float)count[threadIdx.x] / (float)ntests;  // siting.cpp:186
    const int byte = (int)(v * 255.999f);                        // siting.cpp:189
    vix[lx * n + ly] = (uint8_t)(byte < 255 ? byte : 255);
  }
}

}  // namespace

void fill_recip_table(int roi, float *out) {
  out[0] = 0.0f;
  for (int d = 1; d <= roi; d++) out[d] = (float)(1.0 / (double)d);  // siting.cpp:91
}

void fill_jump_table(int ntests, uint32_t *out) {
  // r_m = (A_m * r_0 + C_m) mod 2^31, A_1 = a, C_1 = c,
  // A_{m+1} = a*A_m, C_{m+1} = a*C_m + c (all mod 2^32, masked at use).
  uint32_t A = 1103515245u, C = 12345u;
  for (int m = 1; m <= 2 * ntests; m++) {
    out[2 * (m - 1)] = A;
    out[2 * (m - 1) + 1] = C;
    C = 1103515245u * C + 12345u;
    A = 1103515245u * A;
  }
}

siting_status launch_vix(const Geometry &g, const uint16_t *elevs, uint8_t *vix,
                         const VixTables &t, hipStream_t s) {
  const int n = g.p.nrows;
  const int side = (n + kTile - 1) / kTile;
  const int ntiles = side * side;
  const int per_xcd = (ntiles + 7) / 8;
  // samples read per dependent step and LDS halo (tuning overrides
  // SITING_VIX_GROUP, SITING_VIX_HALO)
  int group = 1, halo = 16;
  if (const char *e = std::getenv("SITING_VIX_GROUP")) group = std::atoi(e);
  if (const char *e = std::getenv("SITING_VIX_HALO")) halo = std::atoi(e);
  const size_t base = sizeof(uint32_t) * (4 * (size_t)g.p.ntests + kThreads);
#define SITING_VIX(G, H)                                                             \
  hipLaunchKernelGGL((vix_kernel<G, H>), dim3(8 * per_xcd), dim3(kThreads),             \
                     base + sizeof(uint16_t) * win_edge<H>() * win_edge<H>(), s, elevs, \
                     vix, t.recip, t.jump, n, g.p.roi, g.p.oht, g.p.tht, g.p.ntests,   \
                     side, per_xcd)
  if (group == 4) SITING_VIX(4, 16);
  else if (group == 2) SITING_VIX(2, 16);
  else if (halo == 8) SITING_VIX(1, 8);
  else if (halo == 32) SITING_VIX(1, 32);
  else SITING_VIX(1, 16);
#undef SITING_VIX
  SITING_HIP(hipGetLastError());
  return SITING_OK;
}

}  // namespace siting
