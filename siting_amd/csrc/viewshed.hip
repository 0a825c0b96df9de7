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

#include "common.hpp"

namespace siting {

namespace {

constexpr int kMaxThreads = 1024;
constexpr int kLdsWords = 8192;  // 64 KiB bitset budget

template <bool kInLds, bool kRdLds>
__global__ __launch_bounds__(kMaxThreads) void viewshed_kernel(
    const uint16_t *__restrict__ elevs, const int32_t *__restrict__ obs,
    uint64_t *__restrict__ sheds, const double *__restrict__ rdiv, int n, int roi,
    int oht, int tht, int nwpr, int nsheds, int per_xcd) {
  const int nthreads = blockDim.x;
  extern __shared__ uint64_t lds_dyn[];  // [roi + 1] reciprocals?, bitset?
  const int b = blockIdx.x;
  const int k = (b & 7) * per_xcd + (b >> 3);
  if (k >= nsheds) return;
  const int nr = 2 * roi + 1;
  const int words = nr * nwpr;
  uint64_t *out = sheds + (size_t)k * words;
  double *rd_l = (double *)lds_dyn;
  const double *rd = kRdLds ? rd_l : rdiv;
  uint64_t *bits = kInLds ? lds_dyn + (kRdLds ? roi + 1 : 0) : out;

  if (kRdLds)
    for (int i = threadIdx.x; i <= roi; i += nthreads) rd_l[i] = rdiv[i];
  for (int i = threadIdx.x; i < words; i += nthreads) bits[i] = 0;
  if (!kInLds) __threadfence();
  __syncthreads();

  const int o = obs[k];
  const int ox = o / n, oy = o - (o / n) * n;
  const int oalt = (int)elevs[o] + oht;
  const float foalt = (float)oalt;
  const int rr = roi * roi;
  const int twor = 2 * roi;

  if (threadIdx.x == 0)  // observer sees itself, siting.cpp:279
    atomicOr((unsigned long long *)&bits[roi * nwpr + (roi >> 6)],
             1ull << (63 - (roi & 63)));

  for (int ray = threadIdx.x; ray < 8 * roi; ray += nthreads) {
    const int side = ray / twor;
    const int j = ray - side * twor;
    int dx, dy;
    if (side == 0) {
      dx = -roi + j;
      dy = -roi;
    } else if (side == 1) {
      dx = -roi + 1 + j;
      dy = roi;
    } else if (side == 2) {
      dx = -roi;
      dy = -roi + 1 + j;
    } else {
      dx = roi;
      dy = -roi + j;
    }
    const int adx = dx < 0 ? -dx : dx, ady = dy < 0 ? -dy : dy;
    const bool along_y = adx < ady;
    const int major = along_y ? dy : dx;
    const int minor = along_y ? dx : dy;
    const int len = along_y ? ady : adx;
    const int dir = major > 0 ? 1 : -1;
    const float slope = (float)minor / (float)major;  // siting.cpp:323
    // cell of step a (siting.cpp:330): the minor offset is the rounded float
    // product, as the reference computes it
    auto cell = [&](int a, int &qx, int &qy) {
      const int i = dir * a;
      const int off = round_half_away((float)i * slope);
      qx = along_y ? off : i;
      qy = along_y ? i : off;
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
        int qx, qy;
        cell(mid, qx, qy);
        const int px = ox + qx, py = oy + qy;
        const bool in = px >= 0 && px < n && py >= 0 && py < n && qx * qx + qy * qy <= rr;
        if (in) last = mid; else hi = mid;
      }
    }
    float horizon = -99999.0f;  // siting.cpp:325
    // siting.cpp:337-344 for one step.  The point slope is siting.cpp:340,
    // (float)(pel - oalt) / (float)a: with |pel - oalt| < 2^21 and a < 2^14
    // the quotient is never closer than 2^-39 (relative) to a float rounding
    // midpoint, and never on one, while the double product below is within
    // 2^-52 of it, so rounding it to float gives the correctly rounded fp32
    // quotient, bit for bit.
    auto step = [&](int a, int pel, int qx, int qy) {
      const float sl = (float)((double)(pel - oalt) * rd[a]);
      if (horizon < sl) horizon = sl;
      const float rise = horizon * (float)a;
      const float halt = foalt + rise;  // siting.cpp:342
      if ((float)(pel + tht) >= halt) {
        const int row = qx + roi, col = qy + roi;
        atomicOr((unsigned long long *)&bits[row * nwpr + (col >> 6)],
                 1ull << (63 - (col & 63)));
      }
    };
    // four steps per pass: their DEM loads are issued before any is used
    int a = 1;
    for (; a + 3 <= last; a += 4) {
      int qx[4], qy[4], pel[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        cell(a + u, qx[u], qy[u]);
        pel[u] = (int)elevs[(ox + qx[u]) * n + (oy + qy[u])];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) step(a + u, pel[u], qx[u], qy[u]);
    }
    for (; a <= last; a++) {
      int qx, qy;
      cell(a, qx, qy);
      step(a, (int)elevs[(ox + qx) * n + (oy + qy)], qx, qy);
    }
  }

  if (kInLds) {
    __syncthreads();
    for (int i = threadIdx.x; i < words; i += nthreads) out[i] = bits[i];
  }
}

}  // namespace

void fill_rdiv_table(int roi, double *out) {
  out[0] = 0.0;
  for (int a = 1; a <= roi; a++) out[a] = 1.0 / (double)a;
}

siting_status launch_viewshed(const Geometry &g, const uint16_t *elevs,
                              const int32_t *obs, uint64_t *sheds,
                              const double *rdiv, hipStream_t s) {
  const int nsheds = g.d.nwanted;
  if (nsheds <= 0) return SITING_OK;
  const int per_xcd = (nsheds + 7) / 8;
  const int words = g.d.nr * g.d.nwpr;
  // 256 threads (4 wavefronts) per observer, rays dealt round-robin; the
  // count can be overridden for tuning
  int threads = 256;
  if (const char *e = std::getenv("SITING_VIEWSHED_THREADS")) {
    const int t = std::atoi(e);
    if (t >= kWave && t <= kMaxThreads && t % kWave == 0) threads = t;
  }
  const bool rd_lds = g.p.roi < 4096;  // 32 KiB table
  const size_t rd_bytes = rd_lds ? sizeof(double) * ((size_t)g.p.roi + 1) : 0;
  const bool in_lds = words <= kLdsWords;
  const size_t lds = rd_bytes + (in_lds ? sizeof(uint64_t) * words : 0);
#define SITING_VS(A, B)                                                             \
  hipLaunchKernelGGL((viewshed_kernel<A, B>), dim3(8 * per_xcd), dim3(threads), lds, \
                     s, elevs, obs, sheds, rdiv, g.p.nrows, g.p.roi, g.p.oht, g.p.tht, \
                     g.d.nwpr, nsheds, per_xcd)
  if (in_lds && rd_lds) SITING_VS(true, true);
  else if (in_lds) SITING_VS(true, false);
  else if (rd_lds) SITING_VS(false, true);
  else SITING_VS(false, false);
#undef SITING_VS
  SITING_HIP(hipGetLastError());
  return SITING_OK;
}

}  // namespace siting
