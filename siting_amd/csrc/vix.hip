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

// The reciprocal table (roi + 1 floats) is copied to LDS up to this roi.
constexpr int kRecipLdsMax = 2048;

// x / 3 truncated toward zero (siting.cpp:164-165) as trunc(fl(x * fl(1/3)))
// with fl(1/3) = 0x3eaaaaab, slightly above 1/3: exact for every |x| <= 2^16
// (checked exhaustively; here |x| <= 2 roi + 1 < 2^15).
__device__ __forceinline__ int div3(int x) {
  return (int)__fmul_rn((float)x, 0.33333334f);
}

// Word 3 of a raw buffer resource on gfx9-family parts (32-bit data format,
// no swizzle).
constexpr int kBufferFlags = 0x00020000;

typedef const __attribute__((address_space(3))) uint16_t LdsU16;

// a * b + c with 24-bit signed a and b, as one v_mad_i32_i24 (the compiler
// otherwise fuses the add into a 64-bit multiply-add)
__device__ __forceinline__ int mad24(int a, int b, int c) {
  int r;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// siting.cpp:157,160: (int)(fl(fl(span * (float)r) / 2^31)) with span =
// (float)(2roi) + 0.99999f.  Dividing by 2^31 is exact, and scaling by a
// power of two commutes with rounding for normal floats, so this is
// (int)fl(span31 * (float)r) with span31 = span * 2^-31 (exact): one
// multiply instead of two.
__device__ __forceinline__ int target_offset(float span31, uint32_t r) {
  return (int)__fmul_rn(span31, (float)(int)r);
}

template <int kHalo, bool kRecipLds>
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
  float *rl = (float *)(count + kThreads + (kWin * kWin + 1) / 2);  // roi + 1 (kRecipLds)

  const int b = blockIdx.x;
  const int tile = (b & 7) * tiles_per_xcd + (b >> 3);
  if (tile >= tiles_side * tiles_side) return;
  const int tx0 = (tile / tiles_side) * kTile;
  const int ty0 = (tile % tiles_side) * kTile;

  for (int i = threadIdx.x; i < 4 * ntests; i += kThreads) jl[i] = jump[i];
  if (kRecipLds)
    for (int i = threadIdx.x; i <= roi; i += kThreads) rl[i] = recip[i];
  const float *rtab = kRecipLds ? rl : recip;
  count[threadIdx.x] = 0;
  const int wx0 = tx0 - kHalo, wy0 = ty0 - kHalo;
  for (int i = threadIdx.x; i < kWin * kWin; i += kThreads) {
    const int wx = wx0 + i / kWin, wy = wy0 + i % kWin;
    // cells outside the terrain are never read (samples lie between an
    // observer and a target, both inside)
    win[i] = wx >= 0 && wx < n && wy >= 0 && wy < n ? elevs[wx * n + wy] : 0;
  }
  __syncthreads();

  // the window's LDS byte address, folded into each pair's base so that a
  // sample's address is a shift-add and a multiply-add
  const int wlds = (int)(uint32_t)(uintptr_t)(const LdsU16 *)win;
  constexpr int kHaloLog = kHalo == 8 ? 3 : kHalo == 16 ? 4 : 5;
  static_assert(kHalo == 1 << kHaloLog, "the halo is a power of two");
  // the DEM as a raw buffer of 2 n^2 bytes: byte offsets below 2^31
  const __amdgpu_buffer_rsrc_t dres =
      __builtin_amdgcn_make_buffer_rsrc((void *)elevs, (short)0, 2 * n * n, kBufferFlags);
  const float span31 = ((float)(2 * roi) + 0.99999f) * 4.656612873077393e-10f;  // * 2^-31
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
      // (24-bit products throughout: coordinates, n and distances are below
      // 2^15, so every product below fits 32 bits)
      const uint32_t p = __umul24((uint32_t)px, (uint32_t)n) + (uint32_t)py;
      const uint32_t r1 = (jl[4 * k] * p + jl[4 * k + 1]) & 0x7fffffffu;
      const uint32_t r2 = (jl[4 * k + 2] * p + jl[4 * k + 3]) & 0x7fffffffu;
      int dx = target_offset(span31, r1) - roi;
      int dy = target_offset(span31, r2) - roi;
      if (__mul24(dx, dx) + __mul24(dy, dy) > rr) {  // siting.cpp:162-166
        dx = div3(dx);
        dy = div3(dy);
      }
      const int tx = px + dx, ty = py + dy;
      if ((uint32_t)tx < (uint32_t)n && (uint32_t)ty < (uint32_t)n) {
        const int adx = dx < 0 ? -dx : dx, ady = dy < 0 ? -dy : dy;
        if (adx <= 1 && ady <= 1) {
          vis = true;
        } else {  // siting.cpp:87-133
          const bool alx = adx >= ady;
          const int major = alx ? dx : dy;
          const int minor = alx ? dy : dx;
          const int len = alx ? adx : ady;
          float rec = rtab[len];
          if (major < 0) rec = -rec;  // (float)(1.0/-d) == -(float)(1.0/d)
          const float slope = (float)minor * rec;
          const int dir = major > 0 ? 1 : -1;
          const int wcell = __mul24(px - wx0, kWin) + (py - wy0);
          const int oz = (int)win[wcell] + oht;
          const int tz = (int)elevs[__umul24((uint32_t)tx, (uint32_t)n) + (uint32_t)ty] + tht;
          const float zs = (float)(tz - oz) * rec;
          const float foz = (float)oz;
          // The reference samples id = dir*1, dir*2, dir*4, ... while |id| <
          // len and stops at the first sample above the sight line, so the
          // target is visible iff no sample is above it.  With |id| = 2^j,
          // slope*id and id*zs are exact scalings of slope and zs: both
          // double from one sample to the next, exactly.  The minor offset
          // rounds slope*id half away from zero, i.e. sign *
          // trunc(fl(|slope*id| + (0.5 - 2^-25))).  Samples up to kHalo
          // steps away are read from the LDS window, the rest from the DEM.
          const bool neg = (slope < 0.0f) != (dir < 0);  // sign of slope*id
          float xs = fabsf(slope);                        // |slope*id| at |id| = 1
          float ts = dir > 0 ? zs : -zs;                  // id*zs at |id| = 1
          // sample j's cell, in bytes: base + (major step << j) + |off| *
          // (signed minor step), in the window and in the DEM
          const int w2 = 2 * kWin, g2 = 2 * n;
          const int wmaj = alx ? (dir > 0 ? w2 : -w2) : (dir > 0 ? 2 : -2);
          const int wmin = alx ? (neg ? -2 : 2) : (neg ? -w2 : w2);
          const int gmaj = alx ? (dir > 0 ? g2 : -g2) : (dir > 0 ? 2 : -2);
          const int gmin = alx ? (neg ? -2 : 2) : (neg ? -g2 : g2);
          const int wbase = 2 * wcell + wlds;
          const int gbase = 2 * (int)p;
          // samples: |id| = 2^j < len, j < nsm; the first nlds in the window
          const int nsm = 32 - __clz(len - 1);
          const int nlds = nsm < kHaloLog + 1 ? nsm : kHaloLog + 1;
          vis = true;
          int j = 0;
          for (; j < nlds; j++) {
            const int oa = (int)__fadd_rn(xs, kHalfDown);
            const int cell = mad24(oa, wmin, (int)((uint32_t)wmaj << j) + wbase);
            const float z = (float)*(const LdsU16 *)(uintptr_t)(uint32_t)cell;
            if (z > __fadd_rn(foz, ts)) {
              vis = false;
              break;
            }
            xs = __fadd_rn(xs, xs);
            ts = __fadd_rn(ts, ts);
          }
          if (vis) {
            for (; j < nsm; j++) {
              const int oa = (int)__fadd_rn(xs, kHalfDown);
              const uint32_t cell =
                  (uint32_t)mad24(oa, gmin, (int)((uint32_t)gmaj << j) + gbase);
              const float z = (float)__builtin_amdgcn_raw_buffer_load_b16(dres, cell, 0, 0);
              if (z > __fadd_rn(foz, ts)) {
                vis = false;
                break;
              }
              xs = __fadd_rn(xs, xs);
              ts = __fadd_rn(ts, ts);
            }
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
  // LDS halo: 32 cells when samples reach 32 steps and a tile has enough
  // pairs to repay the larger window (measured: roi 100-150 with 64 targets
  // 2-6 % faster, roi 25 or 32 targets slower), else 16; a job's tuning
  // (8, 16 or 32) replaces the rule, SITING_VIX_HALO overrides both
  int halo = g.p.roi >= 64 && g.p.ntests >= 64 ? 32 : 16;
  if (g.t.vix_halo == 8 || g.t.vix_halo == 16 || g.t.vix_halo == 32) halo = g.t.vix_halo;
  if (const char *e = std::getenv("SITING_VIX_HALO")) halo = std::atoi(e);
  const bool rlds = g.p.roi <= kRecipLdsMax;
  const size_t base = sizeof(uint32_t) * (4 * (size_t)g.p.ntests + kThreads);
#define SITING_VIX(H, L)                                                                \
  hipLaunchKernelGGL((vix_kernel<H, L>), dim3(8 * per_xcd), dim3(kThreads),                \
                     base + sizeof(uint32_t) * ((win_edge<H>() * win_edge<H>() + 1) / 2) + \
                         (L ? sizeof(float) * (g.p.roi + 1) : 0),                          \
                     s, elevs, vix, t.recip, t.jump, n, g.p.roi, g.p.oht, g.p.tht,         \
                     g.p.ntests, side, per_xcd)
  if (halo == 8) {
    if (rlds) SITING_VIX(8, true);
    else SITING_VIX(8, false);
  } else if (halo == 32) {
    if (rlds) SITING_VIX(32, true);
    else SITING_VIX(32, false);
  } else {
    if (rlds) SITING_VIX(16, true);
    else SITING_VIX(16, false);
  }
#undef SITING_VIX
  SITING_HIP(hipGetLastError());
  return SITING_OK;
}

}  // namespace siting
