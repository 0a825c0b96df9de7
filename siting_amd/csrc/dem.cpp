// dem.cpp -- terrain input: the raw uint16 DEM format the reference reads
// (siting.cpp:726-730: nrows*nrows native-endian uint16, row-major,
// elevs[x*nrows+y]) and a seeded synthetic terrain generator.
//
// The generator is diamond-square midpoint displacement (Fournier, Fussell &
// Carpenter 1982) on a (2^k+1)^2 grid with roughness H (amplitude scales by
// 2^-H per halving), cropped to nrows^2, rescaled linearly to [hmin, hmax],
// plus Gaussian hills, rounded and clamped to uint16.  Everything is plain
// IEEE double arithmetic with no libm transcendental call (exp is computed
// here), so the same seed gives the same bytes on any x86-64 host.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "siting.h"

namespace {

struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [0, 1) with 53 random bits
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double symmetric() { return 2.0 * uniform() - 1.0; }
};

// exp(x) for x <= 0 by range reduction x = k*ln2 + r, |r| <= ln2/2, and a
// degree-16 Taylor polynomial (error far below double rounding for the
// uint16 output).  Deterministic: only +,-,*,/ on doubles.
double det_exp(double x) {
  if (x < -700.0) return 0.0;
  const double ln2 = 0.69314718055994530942;
  double kf = std::floor(x / ln2 + 0.5);
  double r = x - kf * ln2;
  double term = 1.0, sum = 1.0;
  for (int i = 1; i <= 16; i++) {
    term = term * r / (double)i;
    sum = sum + term;
  }
  int k = (int)kf;
  return std::ldexp(sum, k);
}

}  // namespace

extern "C" siting_status siting_dem_generate(int nrows, double rough_h,
                                             int hmin, int hmax, int nhills,
                                             double hill_peak,
                                             double hill_sigma, uint64_t seed,
                                             uint16_t *out) {
  if (nrows <= 0 || nrows > 65536 || !out || hmax < hmin || nhills < 0)
    return SITING_EINVAL;
  int k = 0;
  while ((1 << k) + 1 < nrows) k++;
  const int64_t g = ((int64_t)1 << k) + 1;
  std::vector<double> z((size_t)(g * g), 0.0);
  SplitMix rng{seed ^ 0x5173696E67ull};
  auto at = [&](int64_t x, int64_t y) -> double & { return z[(size_t)(x * g + y)]; };

  at(0, 0) = rng.symmetric();
  at(0, g - 1) = rng.symmetric();
  at(g - 1, 0) = rng.symmetric();
  at(g - 1, g - 1) = rng.symmetric();
  double amp = 1.0;
  const double decay = det_exp(-rough_h * 0.69314718055994530942);  // 2^-H
  for (int64_t step = g - 1; step > 1; step /= 2) {
    int64_t half = step / 2;
    // diamond: square centres
    for (int64_t x = half; x < g; x += step)
      for (int64_t y = half; y < g; y += step) {
        double a = at(x - half, y - half) + at(x - half, y + half) +
                   at(x + half, y - half) + at(x + half, y + half);
        at(x, y) = 0.25 * a + amp * rng.symmetric();
      }
    // square: edge midpoints, averaging the neighbours that exist
    for (int64_t x = 0; x < g; x += half)
      for (int64_t y = ((x / half) % 2 == 0) ? half : 0; y < g; y += step) {
        double a = 0.0;
        int c = 0;
        if (x >= half) { a += at(x - half, y); c++; }
        if (x + half < g) { a += at(x + half, y); c++; }
        if (y >= half) { a += at(x, y - half); c++; }
        if (y + half < g) { a += at(x, y + half); c++; }
        at(x, y) = a / (double)c + amp * rng.symmetric();
      }
    amp *= decay;
  }

  // crop and rescale linearly to [hmin, hmax]
  double lo = at(0, 0), hi = at(0, 0);
  for (int x = 0; x < nrows; x++)
    for (int y = 0; y < nrows; y++) {
      double v = at(x, y);
      lo = v < lo ? v : lo;
      hi = v > hi ? v : hi;
    }
  double scale = hi > lo ? (double)(hmax - hmin) / (hi - lo) : 0.0;
  std::vector<double> f((size_t)nrows * nrows);
  for (int x = 0; x < nrows; x++)
    for (int y = 0; y < nrows; y++)
      f[(size_t)x * nrows + y] = (double)hmin + (at(x, y) - lo) * scale;
  z.clear();
  z.shrink_to_fit();

  // Gaussian hills, each touching the cells within 5 sigma of its centre
  for (int h = 0; h < nhills; h++) {
    double cx = rng.uniform() * (double)nrows;
    double cy = rng.uniform() * (double)nrows;
    double inv = 1.0 / (2.0 * hill_sigma * hill_sigma);
    int reach = (int)(5.0 * hill_sigma) + 1;
    int x0 = (int)cx - reach, x1 = (int)cx + reach;
    int y0 = (int)cy - reach, y1 = (int)cy + reach;
    x0 = x0 < 0 ? 0 : x0;
    y0 = y0 < 0 ? 0 : y0;
    x1 = x1 > nrows - 1 ? nrows - 1 : x1;
    y1 = y1 > nrows - 1 ? nrows - 1 : y1;
    for (int x = x0; x <= x1; x++)
      for (int y = y0; y <= y1; y++) {
        double ddx = (double)x - cx, ddy = (double)y - cy;
        f[(size_t)x * nrows + y] += hill_peak * det_exp(-(ddx * ddx + ddy * ddy) * inv);
      }
  }

  for (size_t i = 0; i < f.size(); i++) {
    double v = std::floor(f[i] + 0.5);
    v = v < 0.0 ? 0.0 : (v > 65535.0 ? 65535.0 : v);
    out[i] = (uint16_t)v;
  }
  return SITING_OK;
}

// siting.cpp:726-730: read exactly nrows^2 uint16 values; a short file fails.
extern "C" siting_status siting_dem_read(const char *path, int nrows,
                                         uint16_t *out) {
  if (!path || nrows <= 0 || !out) return SITING_EINVAL;
  FILE *fp = std::fopen(path, "rb");
  if (!fp) return SITING_EIO;
  size_t want = (size_t)nrows * (size_t)nrows;
  size_t got = std::fread(out, sizeof(uint16_t), want, fp);
  std::fclose(fp);
  return got == want ? SITING_OK : SITING_EIO;
}

extern "C" siting_status siting_dem_write(const char *path, int nrows,
                                          const uint16_t *dem) {
  if (!path || nrows <= 0 || !dem) return SITING_EINVAL;
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return SITING_EIO;
  size_t want = (size_t)nrows * (size_t)nrows;
  size_t put = std::fwrite(dem, sizeof(uint16_t), want, fp);
  int rc = std::fclose(fp);
  return (put == want && rc == 0) ? SITING_OK : SITING_EIO;
}
