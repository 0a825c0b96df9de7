// viewshed.hip -- VIEWSHED stage: one exact radial-sweep bitset per
// tentative observer (siting.cpp:243-360; siting.cu:250-370 has the same
// arithmetic).
//
// A workgroup owns one observer.  Its lanes take the 8*roi rays to the cells
// of the box perimeter (each corner once, siting.cpp:292-304); adjacent lanes
// take adjacent perimeter cells, so at every step the wavefront reads a
// short arc of neighbouring DEM cells.  Along a ray the running horizon
// slope is kept in float exactly as the reference computes it: the IEEE
// quotient for the ray slope and the point slope (the latter via an exact
// double reciprocal, see below), round-half-away for the minor coordinate,
// no fused multiply-add (siting.cpp:323-344).  Visible
// cells are OR'ed into the shed bitset held in LDS (MSB-first 64-bit words,
// siting.cpp:247-250); the finished bitset goes to HBM with coalesced
// 64-bit stores, zero words included, so no separate clear is needed.  When
// the bitset does not fit the LDS budget (roi above ~350) the rays OR
// straight into HBM with device atomics after the workgroup clears it.
//
// Observers are assigned XCD-major (workgroup b runs on XCD b%8, each XCD
// takes a contiguous run of observers, which FINDMAX emitted block by block,
// i.e. spatially adjacent), so each XCD's L2 serves one band of the DEM.
#include <cstdlib>
#include <type_traits>

#include "common.hpp"

namespace siting {

namespace {

constexpr int kMaxThreads = 1024;
constexpr int kLdsWords = 8192;          // 64 KiB bitset budget
constexpr size_t kSplitBytes = 48 * 1024;  // budget of the two split bitsets

// Where a workgroup keeps its observer's bitset while the rays run:
//   kGlobal  straight in HBM (device atomics), for bitsets above kLdsWords
//   kLds     one LDS bitset in the output layout
//   kSplit   two LDS bitsets of 32-bit words, least significant bit first,
//            with an odd row stride of ps dwords (so that neighbouring rows
//            fall in different banks): rows for the rays whose major axis is
//            y, and a transposed one (a run of words per column) for the rays
//            whose major axis is x.  At step a every x-major ray lies on one
//            of two rows, so in the output layout their lanes would all OR
//            into the same few words; transposed, each lane has a word of its
//            own.  Both are turned into the output's MSB-first 64-bit words
//            (the transposed one 64x64 bits per wavefront) when the output is
//            written.
enum { kGlobal = 0, kLds = 1, kSplit = 2 };

__device__ __host__ inline int split_stride(int nwpr) { return 2 * nwpr + 1; }

// 64x64 bit transpose across a wavefront: lane l holds row l (bit c = column
// c); returns lane l's column l as a row, i.e. result bit c = input bit l of
// lane c.  Six butterfly stages swap the off-diagonal j x j quadrants of
// every 2j x 2j block.
__device__ __forceinline__ uint64_t wave_transpose64(uint64_t x, int lane) {
  const uint64_t m[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                         0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
  for (int s = 0; s < 6; s++) {
    const int j = 32 >> s;
    const uint64_t t = __shfl_xor(x, j, kWave);
    x = (lane & j) ? ((x & ~m[s]) | ((t & ~m[s]) >> j)) : ((x & m[s]) | ((t & m[s]) << j));
  }
  return x;
}

__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off, kWave);
    v = o < v ? o : v;
  }
  return v;
}

// The major axis a wavefront's rays walk, in one direction (kAxisMixed:
// each lane's own).
enum { kAxisY = 0, kAxisX = 1, kAxisMixed = 2 };

// max(x, y) as one v_max_f32 (fmaxf adds a canonicalising max of the
// running value; no operand here is ever NaN)
__device__ __forceinline__ float fmax_plain(float x, float y) {
  float r;
  asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(y));
  return r;
}

// Word 3 of a raw buffer resource on gfx9-family parts (32-bit data format,
// no swizzle).
constexpr int kBufferFlags = 0x00020000;


__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(v, off, kWave);
    v = o > v ? o : v;
  }
  return v;
}

template <int kMode>
__global__ __launch_bounds__(kMaxThreads) void viewshed_kernel(
    const uint16_t *__restrict__ elevs, const uint16_t *__restrict__ elevs_t,
    const int32_t *__restrict__ obs,
    uint64_t *__restrict__ sheds, const double *__restrict__ rd, int n, int roi,
    int oht, int tht, int nwpr, int nsheds, int per_xcd) {
  const int nthreads = blockDim.x;
  extern __shared__ uint64_t lds_dyn[];  // the bitset(s) (kLds, kSplit)
  const int b = blockIdx.x;
  const int k = (b & 7) * per_xcd + (b >> 3);
  if (k >= nsheds) return;
  const int nr = 2 * roi + 1;
  const int words = nr * nwpr;
  const int ps = split_stride(nwpr);
  uint64_t *out = sheds + (size_t)k * words;
  uint64_t *bits = kMode == kGlobal ? out : lds_dyn;
  uint32_t *b32 = (uint32_t *)lds_dyn;          // kSplit: row-major bitset
  uint32_t *b32_t = b32 + (size_t)nr * ps;        // kSplit: transposed bitset

  const int clear = kMode == kSplit ? nr * ps : words;  // 64-bit words
  for (int i = threadIdx.x; i < clear; i += nthreads) bits[i] = 0;
  if (kMode == kGlobal) __threadfence();
  __syncthreads();

  const int o = obs[k];
  const int ox = o / n, oy = o - (o / n) * n;
  const int oalt = (int)elevs[o] + oht;
  const float foalt = (float)oalt;
  const int rr = roi * roi;

  if (threadIdx.x == 0) {  // observer sees itself, siting.cpp:279
    if (kMode == kSplit)
      atomicOr(&b32[roi * ps + (roi >> 5)], 1u << (roi & 31));
    else
      atomicOr((unsigned long long *)&bits[roi * nwpr + (roi >> 6)], 1ull << (63 - (roi & 63)));
  }

  // Rays are dealt in passes of nthreads, the y-major ones first (the
  // interior cells of the rows dy = -roi and dy = roi), then the x-major ones
  // (the columns dx = -roi and dx = roi, corners included), so that every
  // wavefront but one walks a single major axis.  Lanes past the last ray
  // repeat it, OR'ing the same bits.  A wavefront walks its rays' steps
  // together: the step index is uniform across it and 1.0/step comes from
  // the table through scalar loads.
  const int ny = 4 * roi - 2;
  for (int base = 0; base < 8 * roi; base += nthreads) {
    const int ray = min(base + (int)threadIdx.x, 8 * roi - 1);
    int dx, dy;
    if (ray < ny) {
      const bool lo = ray < 2 * roi - 1;
      dx = -roi + 1 + (lo ? ray : ray - (2 * roi - 1));
      dy = lo ? -roi : roi;
    } else {
      const int r = ray - ny;
      const bool lo = r < 2 * roi + 1;
      dx = lo ? -roi : roi;
      dy = -roi + (lo ? r : r - (2 * roi + 1));
    }
    const bool along_y = ray < ny;  // |dx| < |dy| exactly for these
    const int major = along_y ? dy : dx;
    const int minor = along_y ? dx : dy;
    const int len = major < 0 ? -major : major;
    const int dir = major > 0 ? 1 : -1;
    const float slope = (float)minor / (float)major;  // siting.cpp:323
    // The minor offset of step a is fl(i * slope) rounded half away from
    // zero, i = dir * a (siting.cpp:330).  Its magnitude is fl(a * |slope|)
    // rounded the same way, which for every float below 2^15 equals
    // trunc(fl(x + (0.5 - 2^-25))) (checked exhaustively); its sign is that
    // of dir * slope.
    const float aslope = fabsf(slope);
    const int sgn = (dir > 0) == (slope >= 0.0f) ? 1 : -1;
    auto offset_abs = [&](float fa) {
      const int r = (int)__fadd_rn(__fmul_rn(fa, aslope), kHalfDown);
      __builtin_assume(r >= 0 && r < 32768);  // |offset| <= a <= roi < 2^14
      return r;
    };
    // The reference walks a = 1..len and stops at the first cell outside the
    // terrain or the roi disk (siting.cpp:333-334).  Both tests are monotone
    // in a: i moves away from the observer, and so does the minor offset (a
    // float product of a fixed slope, rounded half away from zero, is
    // monotone in i).  The steps taken are therefore 1..last, with last
    // found by bisection; the walk below then needs no exit tests.
    int last = 0;
    {
      int hi = len + 1;  // cell(last) inside (a = 0 is the observer), cell(hi) not
      while (hi - last > 1) {
        const int mid = (last + hi) >> 1;
        const int i = dir * mid, off = sgn * offset_abs((float)mid);
        const int qx = along_y ? off : i, qy = along_y ? i : off;
        const int px = ox + qx, py = oy + qy;
        const bool in = px >= 0 && px < n && py >= 0 && py < n && qx * qx + qy * qy <= rr;
        if (in) last = mid; else hi = mid;
      }
    }
    float horizon = -99999.0f;  // siting.cpp:325
    const int wmin = __builtin_amdgcn_readfirstlane(wave_min_i32(last));
    const int wlast = __builtin_amdgcn_readfirstlane(wave_max_i32(last));

    // kCls: kAxisY or kAxisX when every lane of the wavefront walks that
    // major axis in one direction (the step's major coordinate, the DEM row
    // it reads and its bit position are then wavefront-uniform), else
    // kAxisMixed (each lane selects).
    auto walk = [&](auto cls_c) {
      constexpr int kCls = decltype(cls_c)::value;
      constexpr bool kUni = kCls != kAxisMixed;
      const bool ay = kCls == kAxisY ? true : kCls == kAxisX ? false : along_y;
      const int sdir = kUni ? __builtin_amdgcn_readfirstlane(dir) : dir;
      const int sps4 = sgn * ps * 4;
      // siting.cpp:337-344 for one step.  The point slope is siting.cpp:340,
      // (float)(pel - oalt) / (float)a: with |pel - oalt| < 2^21 and
      // a < 2^14 the quotient is never closer than 2^-39 (relative) to a
      // float rounding midpoint, and never on one, while the double product
      // below is within 2^-52 of it, so rounding it to float gives the
      // correctly rounded fp32 quotient, bit for bit.  (The running maximum
      // compares as the reference's "if (horizon < s) horizon = s": the
      // slopes are never NaN and the sign of a zero horizon changes
      // nothing downstream.)
      auto step = [&](bool take, int a, float fa, double rda, int pel, int offa) {
        const float sl = (float)((double)(pel - oalt) * rda);
        horizon = take ? fmax_plain(horizon, sl) : horizon;
        const float rise = __fmul_rn(horizon, fa);
        const float halt = __fadd_rn(foalt, rise);  // siting.cpp:342
        if (take && (float)(pel + tht) >= halt) {
          // (24-bit products: rows, columns and strides are below 2^15)
          // the major coordinate in the box (wavefront-uniform in kUni)
          const int c = kUni ? __builtin_amdgcn_readfirstlane(roi + sdir * a) : roi + sdir * a;
          if (kMode == kSplit && kUni) {
            // y-major: row off + roi, column c; x-major, transposed: column
            // off + roi, row c
            const int wb = 4 * ((ay ? roi * ps : nr * ps + roi * ps) + (c >> 5));
            atomicOr((uint32_t *)((uint8_t *)b32 + wb + __mul24(offa, sps4)), 1u << (c & 31));
          } else {
            const int off = __mul24(sgn, offa);
            const int row = ay ? off + roi : c, col = ay ? c : off + roi;
            if (kMode == kSplit) {
              const int w = ay ? (int)__umul24(row, ps) + (col >> 5)
                               : nr * ps + (int)__umul24(col, ps) + (row >> 5);
              atomicOr(&b32[w], 1u << ((ay ? col : row) & 31));
            } else {
              atomicOr((unsigned long long *)&bits[(int)__umul24(row, nwpr) + (col >> 6)],
                       1ull << (63 - (col & 63)));
            }
          }
        }
      };
      // Four steps at a time, their DEM loads issued before any is used.
      // The four reciprocals are read first (scalar loads share their wait
      // counter with the LDS atomics, so a read just before its use would
      // also wait for every atomic in flight; the table has three spare
      // entries past roi for the last group).  The y-major rays read the
      // transposed DEM, so that a wavefront's reads share a row either way
      // (in a uniform wavefront, the one row).  DEM offsets are 32-bit with
      // 24-bit products (n < 2^15).
      const uint16_t *dem = ay ? elevs_t : elevs;
      const int omaj = ay ? oy : ox, omin = ay ? ox : oy;
      const int sgn2 = 2 * sgn;
      // the DEM as a buffer of 2 n^2 bytes (raw, no stride)
      const __amdgpu_buffer_rsrc_t dres =
          __builtin_amdgcn_make_buffer_rsrc((void *)dem, (short)0, 2 * n * n, kBufferFlags);
      auto group = [&](int a, auto checked_c) {
        constexpr bool kChecked = decltype(checked_c)::value;
        double rda[4];
#pragma unroll
        for (int u = 0; u < 4; u++) rda[u] = rd[a + u];
        int offa[4], pel[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          offa[u] = offset_abs((float)(a + u));
          int v = 0;
          if (!kChecked || a + u <= last) {
            if (kUni) {
              // a buffer load: the row's byte offset as the scalar offset,
              // the lane's cell in the row as the vector one (both >= 0)
              const int soff = 2 * (int)__umul24((uint32_t)(omaj + sdir * (a + u)), (uint32_t)n);
              v = __builtin_amdgcn_raw_buffer_load_b16(dres, __mul24(offa[u], sgn2) + 2 * omin, soff, 0);
            } else {
              const uint32_t prow = (uint32_t)(omaj + sdir * (a + u));
              const uint32_t pcol = (uint32_t)(omin + __mul24(sgn, offa[u]));
              v = dem[__umul24(prow, (uint32_t)n) + pcol];
            }
          }
          pel[u] = v;
        }
        // (no instruction: makes the four reciprocals ready here, after the
        // DEM loads are issued and before the steps' conditional blocks, so
        // that none of those waits again)
        asm volatile("" ::"s"(rda[0]), "s"(rda[1]), "s"(rda[2]), "s"(rda[3]));
#pragma unroll
        for (int u = 0; u < 4; u++)
          step(!kChecked || a + u <= last, a + u, (float)(a + u), rda[u], pel[u], offa[u]);
      };
      int a = 1;
      // steps every lane takes, then the rest, each lane up to its own last
      for (; a + 3 <= wmin; a += 4) group(a, std::integral_constant<bool, false>{});
      for (; a <= wlast; a += 4) group(a, std::integral_constant<bool, true>{});
    };
    const int cls = (along_y ? 0 : 2) + (dir > 0 ? 1 : 0);
    if (__ballot(cls == __builtin_amdgcn_readfirstlane(cls)) != ~0ull)
      walk(std::integral_constant<int, kAxisMixed>{});
    else if (along_y)
      walk(std::integral_constant<int, kAxisY>{});
    else
      walk(std::integral_constant<int, kAxisX>{});
  }

  if (kMode == kLds) {
    __syncthreads();
    for (int i = threadIdx.x; i < words; i += nthreads) out[i] = bits[i];
  } else if (kMode == kSplit) {
    // block (R, W): rows 64R..64R+63, output word W (columns 64W..64W+63).
    // Lane l loads column 64W+l's two transposed dwords for those rows (bit r
    // for row 64R+r): row l of a matrix whose bit r is row 64R+r.
    // Transposed, lane l holds row 64R+l with bit c for column 64W+c, which
    // bit-reversed is the output word (bit 63-c), as is the row-major pair.
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave, nwaves = nthreads / kWave;
    auto pair = [](const uint32_t *p) { return (uint64_t)p[0] | ((uint64_t)p[1] << 32); };
    for (int blk = wid; blk < nwpr * nwpr; blk += nwaves) {
      const int R = blk / nwpr, W = blk - R * nwpr;
      const int c = 64 * W + lane;
      const uint64_t tw = c < nr ? pair(&b32_t[c * ps + 2 * R]) : 0;
      const uint64_t mine = wave_transpose64(tw, lane);
      const int row = 64 * R + lane;
      if (row < nr) out[row * nwpr + W] = __brevll(pair(&b32[row * ps + 2 * W]) | mine);
    }
  }
}

// elevs_t[y * n + x] = elevs[x * n + y], through 64 x 64 LDS tiles: rows of
// a tile are read and columns written with consecutive lanes on consecutive
// cells.
constexpr int kTile = 64, kTileRows = 4;
__global__ __launch_bounds__(kTile * kTileRows) void dem_transpose(
    const uint16_t *__restrict__ in, uint16_t *__restrict__ out, int n) {
  __shared__ uint16_t tile[kTile][kTile + 2];
  const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  const int x0 = blockIdx.y * kTile, y0 = blockIdx.x * kTile;
  for (int r = ty; r < kTile; r += kTileRows) {
    const int x = x0 + r, y = y0 + tx;
    if (x < n && y < n) tile[r][tx] = in[(size_t)x * n + y];
  }
  __syncthreads();
  for (int r = ty; r < kTile; r += kTileRows) {
    const int y = y0 + r, x = x0 + tx;
    if (x < n && y < n) out[(size_t)y * n + x] = tile[tx][r];
  }
}

}  // namespace

void fill_rdiv_table(int roi, double *out) {
  out[0] = 0.0;
  for (int a = 1; a <= roi; a++) out[a] = 1.0 / (double)a;
}

siting_status launch_viewshed(const Geometry &g, const uint16_t *elevs, uint16_t *elevs_t,
                              const int32_t *obs, uint64_t *sheds,
                              const double *rdiv, hipStream_t s) {
  const int nsheds = g.d.nwanted;
  if (nsheds <= 0) return SITING_OK;
  const int tiles = (g.p.nrows + kTile - 1) / kTile;
  hipLaunchKernelGGL(dem_transpose, dim3(tiles, tiles), dim3(kTile * kTileRows), 0, s, elevs,
                     elevs_t, g.p.nrows);
  const int per_xcd = (nsheds + 7) / 8;
  const int words = g.d.nr * g.d.nwpr;
  // 256 threads (4 wavefronts) per observer, rays dealt round-robin; the
  // count can be overridden for tuning
  int threads = 256;
  if (const char *e = std::getenv("SITING_VIEWSHED_THREADS")) {
    const int t = std::atoi(e);
    if (t >= kWave && t <= kMaxThreads && t % kWave == 0) threads = t;
  }
  const size_t split_bytes = 2 * sizeof(uint32_t) * g.d.nr * (size_t)split_stride(g.d.nwpr);
  const int mode = split_bytes <= kSplitBytes ? kSplit : words <= kLdsWords ? kLds : kGlobal;
  const size_t lds = mode == kSplit ? split_bytes
                     : mode == kLds ? sizeof(uint64_t) * words : 0;
#define SITING_VS(A)                                                               \
  hipLaunchKernelGGL((viewshed_kernel<A>), dim3(8 * per_xcd), dim3(threads), lds, s, \
                     elevs, elevs_t, obs, sheds, rdiv, g.p.nrows, g.p.roi, g.p.oht,    \
                     g.p.tht, g.d.nwpr, nsheds, per_xcd)
  if (mode == kSplit) SITING_VS(kSplit);
  else if (mode == kLds) SITING_VS(kLds);
  else SITING_VS(kGlobal);
#undef SITING_VS
  SITING_HIP(hipGetLastError());
  return SITING_OK;
}

}  // namespace siting
