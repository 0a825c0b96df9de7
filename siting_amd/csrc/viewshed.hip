// viewshed.hip -- VIEWSHED stage: one exact radial-sweep bitset per
// tentative observer (siting.cpp:243-360; siting.cu:250-370 has the same
// arithmetic).
//
// A workgroup owns one observer.  Its lanes take the 8*roi rays to the cells
// of the box perimeter (each corner once, siting.cpp:292-304), dealt by side
// of the box in whole wavefronts, so that a wavefront's rays share major
// axis and direction and adjacent lanes take adjacent perimeter cells: at
// every step the wavefront reads a short arc of one DEM row (y-major rays
// read a transposed copy), and the step's row, reciprocal and bit position
// are wavefront-uniform.  Along a ray the running horizon slope is kept in
// float exactly as the reference computes it: the IEEE quotient for the ray
// slope and the point slope (the latter via an exact double reciprocal, see
// below), round-half-away for the minor coordinate, no fused multiply-add
// (siting.cpp:323-344).  Everything a step needs that depends only on roi
// comes from per-roi tables filled on the host (vs_tables): the minor offsets
// of every ray, the reciprocals, the bit positions, the lanes that share a
// cell, and each ray's last step inside the roi disk.  Visible cells are
// OR'ed into bitsets in LDS: up to roi ~200 two of 32-bit words (one
// transposed for the x-major rays), larger ones in the output layout (the
// MSB-first 64-bit words of siting.cpp:247-250), and beyond the LDS budget
// (roi above ~350) straight into HBM with device atomics.  The finished
// bitset goes to HBM with coalesced 64-bit stores, zero words included, so
// no separate clear is needed.
//
// Observers are assigned XCD-major (workgroup b runs on XCD b%8, each XCD
// takes a contiguous run of observers, which FINDMAX emitted block by block,
// i.e. spatially adjacent), so each XCD's L2 serves one band of the DEM.
#include <cmath>
#include <cstdlib>
#include <type_traits>

#include "common.hpp"

namespace siting {

namespace {

constexpr int kMaxThreads = 1024;
constexpr int kLdsWords = 8192;          // 64 KiB bitset budget

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

__device__ __host__ inline int split_stride(int nwpr) { return vs_split_stride(nwpr); }

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

// The major axis a wavefront's rays walk (all in one direction).
enum { kAxisY = 0, kAxisX = 1 };

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

// a lane's four offset-table entries of one group
typedef uint32_t VsOff4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) uint32_t LdsU32;
// the step records, read-only for the kernel's life: the constant address
// space, so that they are read with scalar loads (the host pass, which only
// parses the kernel, gets a plain pointer)
#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(4))) VsStep4 ConstStep4;
#else
typedef const VsStep4 ConstStep4;
#endif


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
    uint64_t *__restrict__ sheds, const uint8_t *__restrict__ tabs, int n, int roi,
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

  const VsTables tb = vs_tables(roi);
  const int cols = 2 * roi + 1;
  // the offset table as a buffer: group g at byte 16 g cols (the scalar
  // offset), a lane's four entries at 16 (m + roi) (the vector one)
  const __amdgpu_buffer_rsrc_t tres = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(tabs + tb.off), (short)0, (int)(tb.step - tb.off), kBufferFlags);
  // 2 ps in a vector register: the bit address t * 2ps + wofs is then one
  // multiply-add with a single scalar operand
  int ps2v;
  asm("v_mov_b32 %0, %1" : "=v"(ps2v) : "s"(2 * ps));
  // the split bitsets as a byte pointer in the LDS address space
  __attribute__((address_space(3))) uint8_t *lds8 =
      (__attribute__((address_space(3))) uint8_t *)b32;

  const int o = obs[k];
  const int ox = o / n, oy = o - (o / n) * n;
  const int oalt = (int)elevs[o] + oht;
  const float foalt = (float)oalt;
  const int rr = roi * roi;
  // an observer at least roi cells from every edge of the terrain
  const bool interior = ox >= roi && oy >= roi && ox + roi < n && oy + roi < n;

  if (threadIdx.x == 0) {  // observer sees itself, siting.cpp:279
    if (kMode == kSplit)
      atomicOr(&b32[roi * ps + (roi >> 5)], 1u << (roi & 31));
    else
      atomicOr((unsigned long long *)&bits[roi * nwpr + (roi >> 6)], 1ull << (63 - (roi & 63)));
  }

  // The rays are dealt by side of the box, each side in whole wavefronts:
  // the rows dy = -roi and dy = roi without their corners (2roi-1 y-major
  // rays each), then the columns dx = -roi and dx = roi with them (2roi+1
  // x-major rays each).  A wavefront's rays therefore share major axis and
  // direction.  A lane past the end of its side repeats the ray 64 places
  // back (walked by the wavefront before, so that the two never OR into one
  // word in one instruction), or on a side of fewer than 64 rays one of its
  // own; the repeated bits change nothing.  Wavefront wv of the 2(wy+wx)
  // goes to wavefront wv mod (nthreads/64) of the workgroup.  A wavefront
  // walks its rays' steps together, so the step index is uniform across it
  // and the step's constants come from the tables through scalar loads.
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int my = 2 * roi - 1, mx = 2 * roi + 1;
  const int wy = (my + kWave - 1) / kWave, wx = (mx + kWave - 1) / kWave;
  for (int wv = wid; wv < 2 * (wy + wx); wv += nthreads / kWave) {
    const bool along_y = wv < 2 * wy;  // |dx| < |dy| exactly for these
    const int h = along_y ? wv : wv - 2 * wy, wside = along_y ? wy : wx;
    const bool lo = h < wside;
    const int m = along_y ? my : mx;
    const int chunk = lo ? h : h - wside;  // the wavefront's 64 rays of its side
    const int j0 = chunk * kWave + lane;
    const int j = j0 < m ? j0 : m >= kWave ? j0 - kWave : j0 % m;
    int dx, dy;
    if (along_y) {
      dx = -roi + 1 + j;
      dy = lo ? -roi : roi;
    } else {
      dx = lo ? -roi : roi;
      dy = -roi + j;
    }
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
    if (interior) {
      // no ray leaves the terrain within roi steps: only the disk ends it
      last = (int)((const uint32_t *)(tabs + tb.disk))[minor + roi];
    } else {
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
    // the lane's entries in the offset table (bytes)
    const int tcol = 16 * (minor + roi);

    // kCls: the major axis every lane of the wavefront walks, in one
    // direction, so the step's major coordinate, the DEM row it reads and
    // its bit position are wavefront-uniform.
    auto walk = [&](auto cls_c) {
      constexpr bool ay = decltype(cls_c)::value == kAxisY;
      const int sdir = __builtin_amdgcn_readfirstlane(dir);
      ConstStep4 *rec = (ConstStep4 *)(tabs + tb.step) +
                        ((size_t)((ay ? 0 : 2) + (sdir > 0)) * tb.chunks +
                         (kMode == kSplit ? chunk : 0)) * tb.groups;
      // siting.cpp:337-344 for one step.  The point slope is siting.cpp:340,
      // (float)(pel - oalt) / (float)a: with |pel - oalt| < 2^21 and
      // a < 2^14 the quotient is never closer than 2^-39 (relative) to a
      // float rounding midpoint, and never on one, while the double product
      // below is within 2^-52 of it, so rounding it to float gives the
      // correctly rounded fp32 quotient, bit for bit.  (The running maximum
      // compares as the reference's "if (horizon < s) horizon = s": the
      // slopes are never NaN and the sign of a zero horizon changes
      // nothing downstream.)  t: the offset table's entry 2 (off + roi).
      auto step = [&](bool take, int a, float fa, double rda, uint32_t wofs, uint32_t wbit,
                      uint64_t same, int pel, int t) {
        const float sl = (float)((double)(pel - oalt) * rda);
        horizon = take ? fmax_plain(horizon, sl) : horizon;
        const float rise = __fmul_rn(horizon, fa);
        const float halt = __fadd_rn(foalt, rise);  // siting.cpp:342
        const bool vis = take && (float)(pel + tht) >= halt;
        if (kMode == kSplit) {
          // A lane whose left neighbour is visible in the same cell leaves
          // the OR to it (the first visible lane of each run of lanes on one
          // cell ORs; lanes on one cell OR the same bit of the same word, so
          // their ORs would only serialise).
          const uint64_t v = __ballot(vis);
          if (__builtin_amdgcn_inverse_ballot_w64(v & ~((v << 1) & same))) {
            // y-major: row off + roi, column c; x-major, transposed: column
            // off + roi, row c; the row's byte offset (off + roi) 4 ps is t 2ps
            __hip_atomic_fetch_or((LdsU32 *)(lds8 + wofs + __umul24(t, ps2v)), wbit,
                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        } else if (vis) {
          // (24-bit products: rows, columns and strides are below 2^15)
          const int c = __builtin_amdgcn_readfirstlane(roi + sdir * a);
          const int row = ay ? t >> 1 : c, col = ay ? c : t >> 1;
          atomicOr((unsigned long long *)&bits[(int)__umul24(row, nwpr) + (col >> 6)],
                   1ull << (63 - (col & 63)));
        }
      };
      // Four steps at a time, their DEM loads issued before any is used.
      // Their minor offsets come from the offset table (one 16-byte load
      // per group, no arithmetic); the next group's are loaded while this
      // group's steps run.  The group's uniform values are read with scalar
      // loads after the DEM loads are issued (scalar loads share their wait
      // counter with the LDS atomics, so a read just before its use would
      // also wait for every atomic in flight).  The y-major rays read the
      // transposed DEM, so that a wavefront's reads share a row either way
      // (the one row).
      const uint16_t *dem = ay ? elevs_t : elevs;
      const int omaj = ay ? oy : ox, omin = ay ? ox : oy;
      // The DEM as a raw buffer whose byte 0 is column omin - roi of row 0:
      // a cell's row is the scalar offset and the table entry 2 (off + roi)
      // the vector one.  Every access is a cell of the terrain (the walk
      // stops at its edge), so the base may lie before the DEM's first cell
      // without anything being read there.
      const __amdgpu_buffer_rsrc_t dres = __builtin_amdgcn_make_buffer_rsrc(
          (void *)(dem + (omin - roi)), (short)0, 2 * n * n + 4 * roi, kBufferFlags);
      auto tload = [&](int g) {
        return __builtin_amdgcn_raw_buffer_load_b128(tres, tcol, g * (16 * cols), 0);
      };
      // byte offset of the DEM row of step 1, and from one step to the next
      const int sd = 2 * n * sdir;
      int srow = 2 * n * omaj + sd;
      // group g: steps 4g+1..4g+4; tc its table entries, tn receives the
      // next group's
      auto group = [&](int g, auto checked_c, const VsOff4 &tc, VsOff4 &tn, ConstStep4 *rg) {
        constexpr bool kChecked = decltype(checked_c)::value;
        const int a = 4 * g + 1;
        int pel[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          int v = 0;
          if (!kChecked || a + u <= last)
            v = __builtin_amdgcn_raw_buffer_load_b16(dres, tc[u], srow + u * sd, 0);
          pel[u] = v;
        }
        // (unconditional, so that the waits for this group's DEM loads let
        // it stay in flight)
        tn = tload(g + 1);
        const VsStep4 r = *rg;
        // (no instruction: makes the scalar values ready here, before the
        // steps' conditional blocks, so that none of those waits again)
        asm volatile("" ::"s"(r.rd[0]), "s"(r.rd[1]), "s"(r.rd[2]), "s"(r.rd[3]),
                     "s"(r.fa[0]), "s"(r.fa[1]), "s"(r.fa[2]), "s"(r.fa[3]), "s"(r.wofs[0]),
                     "s"(r.wofs[1]), "s"(r.wofs[2]), "s"(r.wofs[3]), "s"(r.wbit[0]),
                     "s"(r.wbit[1]), "s"(r.wbit[2]), "s"(r.wbit[3]), "s"(r.same[0]),
                     "s"(r.same[1]), "s"(r.same[2]), "s"(r.same[3]));
#pragma unroll
        for (int u = 0; u < 4; u++)
          step(!kChecked || a + u <= last, a + u, r.fa[u], r.rd[u], r.wofs[u], r.wbit[u],
               r.same[u], pel[u], (int)tc[u]);
        srow += 4 * sd;
      };
      // steps every lane takes, then the rest, each lane up to its own last;
      // the unchecked groups in pairs that swap the two entry buffers, the
      // others with the next entries loaded into ta and this group's copied
      const std::integral_constant<bool, false> unchecked;
      const std::integral_constant<bool, true> checked;
      // (rg: group g's record; kept opaque to the loop optimiser so that
      // the records are read at small positive offsets from it)
      VsOff4 ta = tload(0), tb4;
      ConstStep4 *rg = rec;
      int g = 0;
      for (; 4 * g + 8 <= wmin; g += 2) {
        group(g, unchecked, ta, tb4, rg);
        group(g + 1, unchecked, tb4, ta, rg + 1);
        rg += 2;
        asm volatile("" : "+s"(rg));
      }
      for (; 4 * g + 4 <= wmin; g++, rg++) {
        const VsOff4 tc = ta;
        group(g, unchecked, tc, ta, rg);
      }
      for (; 4 * g + 1 <= wlast; g++, rg++) {
        const VsOff4 tc = ta;
        group(g, checked, tc, ta, rg);
      }
    };
    if (along_y)
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

// the minor offset of step a on the ray whose perimeter cell has minor
// coordinate m: the kernel's ray slope is fl(minor / major), |major| = roi;
// the offset's magnitude is fl(a |slope|) rounded half away from zero, its
// sign that of m (on every side of the box)
static int vs_offset(int roi, int a, int m) {
  const float aslope = std::fabs((float)m / (float)roi);
  const volatile float prod = (float)a * aslope;  // one rounding each
  const volatile float biased = prod + kHalfDown;
  const int o = (int)biased;
  return m < 0 ? -o : o;
}

void fill_vs_tables(int roi, void *out) {
  const VsTables t = vs_tables(roi);
  uint8_t *b = (uint8_t *)out;
  uint32_t *off = (uint32_t *)(b + t.off);
  VsStep4 *rec = (VsStep4 *)(b + t.step);
  const int cols = 2 * roi + 1, nr = 2 * roi + 1;
  const int ps = split_stride((nr + 63) / 64);
  for (int g = 0; g < t.groups; g++) {
    for (int u = 0; u < 4; u++) {
      const int a = 4 * g + 1 + u;
      const bool in = a <= roi;
      for (int m = -roi; m <= roi; m++)
        off[((size_t)g * cols + (m + roi)) * 4 + u] =
            (uint32_t)(2 * ((in ? vs_offset(roi, a, m) : 0) + roi));
      for (int cls = 0; cls < 4; cls++) {
        const bool ay = cls < 2;
        const int dir = (cls & 1) ? 1 : -1;
        const int c = roi + dir * a;
        // the side's rays as the kernel deals them: chunk ch, lane l takes
        // ray j of the side, minor coordinate m = m0 + j
        const int mside = ay ? 2 * roi - 1 : 2 * roi + 1, m0 = ay ? -roi + 1 : -roi;
        for (int ch = 0; ch < t.chunks; ch++) {
          VsStep4 &r = rec[((size_t)cls * t.chunks + ch) * t.groups + g];
          r.rd[u] = in ? 1.0 / (double)a : 0.0;
          r.fa[u] = (float)a;
          r.wofs[u] = in ? 4u * (uint32_t)((ay ? 0 : nr * ps) + (c >> 5)) : 0u;
          r.wbit[u] = in ? 1u << (c & 31) : 0u;
          uint64_t same = 0;
          if (in && vs_split(roi)) {
            int prev = 0;
            for (int l = 0; l < 64; l++) {
              const int j0 = ch * 64 + l;
              const int j = j0 < mside ? j0 : mside >= 64 ? j0 - 64 : j0 % mside;
              const int o = vs_offset(roi, a, m0 + j);
              if (l > 0 && o == prev) same |= 1ull << l;
              prev = o;
            }
          }
          r.same[u] = same;
        }
      }
    }
  }
  // the last step inside the disk: a^2 + off^2 <= roi^2 holds for a prefix
  // of the steps (both terms grow with a), found by bisection
  uint32_t *disk = (uint32_t *)(b + t.disk);
  for (int m = -roi; m <= roi; m++) {
    int last = 0, hi = roi + 1;  // step last inside (0: the observer), hi not
    while (hi - last > 1) {
      const int a = (last + hi) >> 1, o = vs_offset(roi, a, m);
      if (a * a + o * o <= roi * roi) last = a; else hi = a;
    }
    disk[m + roi] = (uint32_t)last;
  }
}

siting_status launch_viewshed(const Geometry &g, const uint16_t *elevs, uint16_t *elevs_t,
                              const int32_t *obs, uint64_t *sheds,
                              const uint8_t *tables, hipStream_t s) {
  const int nsheds = g.d.nwanted;
  if (nsheds <= 0) return SITING_OK;
  const int tiles = (g.p.nrows + kTile - 1) / kTile;
  hipLaunchKernelGGL(dem_transpose, dim3(tiles, tiles), dim3(kTile * kTileRows), 0, s, elevs,
                     elevs_t, g.p.nrows);
  const int per_xcd = (nsheds + 7) / 8;
  const int words = g.d.nr * g.d.nwpr;
  // 256 threads (4 wavefronts) per observer, rays dealt by box side; a
  // job's tuning replaces the count, SITING_VIEWSHED_THREADS overrides both
  int threads = 256;
  auto valid_threads = [](int t) { return t >= kWave && t <= kMaxThreads && t % kWave == 0; };
  if (valid_threads(g.t.viewshed_threads)) threads = g.t.viewshed_threads;
  if (const char *e = std::getenv("SITING_VIEWSHED_THREADS")) {
    const int t = std::atoi(e);
    if (valid_threads(t)) threads = t;
  }
  const size_t split_bytes = 2 * sizeof(uint32_t) * g.d.nr * (size_t)split_stride(g.d.nwpr);
  // bitsets: the split pair when it fits, else one LDS bitset when it fits,
  // else HBM; a job's tuning may pick another mode that fits
  int mode = vs_split(g.p.roi) ? kSplit : words <= kLdsWords ? kLds : kGlobal;
  if (g.t.viewshed_mode == 1 && vs_split(g.p.roi)) mode = kSplit;
  else if (g.t.viewshed_mode == 2 && words <= kLdsWords) mode = kLds;
  else if (g.t.viewshed_mode == 3) mode = kGlobal;
  const size_t lds = mode == kSplit ? split_bytes
                     : mode == kLds ? sizeof(uint64_t) * words : 0;
#define SITING_VS(A)                                                               \
  hipLaunchKernelGGL((viewshed_kernel<A>), dim3(8 * per_xcd), dim3(threads), lds, s, \
                     elevs, elevs_t, obs, sheds, tables, g.p.nrows, g.p.roi, g.p.oht,  \
                     g.p.tht, g.d.nwpr, nsheds, per_xcd)
  if (mode == kSplit) SITING_VS(kSplit);
  else if (mode == kLds) SITING_VS(kLds);
  else SITING_VS(kGlobal);
#undef SITING_VS
  SITING_HIP(hipGetLastError());
  return SITING_OK;
}

}  // namespace siting
