/*
 * siting_oracle.c -- CPU restatement of the reference's sequential siting
 * pipeline (siting.cpp, -DSEQ semantics).  TEST INFRASTRUCTURE ONLY: this is
 * the checker for the HIP path and the CPU timing baseline; the product never
 * calls it.
 *
 * Pinning (what checks each function besides its reference citation):
 *   so_lcg            closed-form known answers; 2-step jump composition
 *   so_derive         closed-form float32 values; reference build (_ref)
 *   so_calc_vix       flat terrain -> vix = table[#in-terrain targets];
 *                     reference build (_ref) on generated and sample DEMs
 *   so_find_max       brute-force key order on tiny DEMs; hash known answers;
 *                     reference build (_ref)
 *   so_calc_sheds     flat terrain -> shed = disk(roi) cap terrain exactly;
 *                     reference build (_ref)
 *   so_site           naive exact greedy over explicit boolean grids
 *                     (intervis=0); cumarea == popcount(cumshed); reference
 *                     build (_ref) for intervis 0 and 1
 *
 * Built with -O2 -ffp-contract=off (no fused multiply-add, as the reference's
 * x86-64 baseline g++ build).
 */
#include "siting_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double wall_now(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

static inline int iabs_(int v) { return v < 0 ? -v : v; }
static inline int imin_(int a, int b) { return a < b ? a : b; }

/* siting.cpp:29-33: r' = (1103515245 * (r mod 2^31) + 12345) mod 2^31 */
int32_t so_lcg(int32_t seed) {
  uint32_t s = (uint32_t)seed & 0x7fffffffu;
  return (int32_t)((s * 1103515245u + 12345u) & 0x7fffffffu);
}

/* siting.cpp:689-695.  Returns 0; 1 when the last block is too small
 * (siting.cpp:694-695); 2 when nrows/bw rounds to zero blocks. */
int so_derive(int nrows, int bw, int nwanted0, so_derived *out) {
  float fn = (float)nrows;
  int nb_req = (int)(fn / (float)bw + 0.5f);
  if (nb_req <= 0) return 2; /* the reference divides by zero here */
  float bs = fn / (float)nb_req;
  int nb = (int)(fn / bs + 0.5f);
  int nbsq = nb * nb;
  int per = (int)((float)nwanted0 / (float)nbsq + 0.99999f);
  out->blocksize = bs;
  out->nblocks = nb;
  out->nwpb = per;
  out->nwanted = per * nbsq;
  out->lastsize = nrows - (int)(bs * (float)(nb - 1));
  return (out->lastsize * out->lastsize < per) ? 1 : 0;
}

/* siting.cpp:680-686 */
int so_check_args(int nrows, int roi, int ht, int ntests, int bw, int nwanted0,
                  int intervis) {
  if (nrows <= 0 || nrows > 20000) return 1;
  if (roi < 1 || roi > 10000) return 2;
  if (ht < 0 || ht > 1000000) return 3;
  if (ntests < 1 || ntests > 1000) return 4;
  if (bw < 10 || bw > 2000) return 5;
  if (nwanted0 < 100 || nwanted0 > 2000000) return 6;
  if (intervis != 0 && intervis != 1) return 7;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* VIX                                                                      */
/* ------------------------------------------------------------------------ */

/* siting.cpp:54-135.  Samples at major-axis offsets 1,2,4,... strictly inside
 * the segment; blocked when the terrain sample is above the straight line of
 * sight.  The reciprocal of the major extent is a double division rounded to
 * float (siting.cpp:91); slope and zslope are float products with it. */
static int los_test(int n, const uint16_t *e, int ox, int oy, int oz, int tx,
                    int ty, int tz, int64_t *nsamp) {
  int dx = tx - ox, dy = ty - oy;
  if (iabs_(dx) <= 1 && iabs_(dy) <= 1) return 1;
  int along_x = iabs_(dx) >= iabs_(dy);
  int major = along_x ? dx : dy;
  int minor = along_x ? dy : dx;
  float rec = (float)(1.0 / (double)major);
  float slope = (float)minor * rec;
  float zslope = (float)(tz - oz) * rec;
  int dir = major > 0 ? 1 : -1;
  int len = iabs_(major);
  for (int k = 1; k < len; k <<= 1) {
    int id = dir * k;
    float fid = (float)id;
    float pm = slope * fid;
    int off = (int)roundf(pm);
    int px = along_x ? ox + id : ox + off;
    int py = along_x ? oy + off : oy + id;
    float z = (float)e[px * n + py];
    float step = fid * zslope;
    float lim = (float)oz + step;
    if (nsamp) (*nsamp)++;
    if (z > lim) return 0;
  }
  return 1;
}

int so_test_one_target(int nrows, const uint16_t *elevs, int ox, int oy,
                       int oz, int tx, int ty, int tz) {
  return los_test(nrows, elevs, ox, oy, oz, tx, ty, tz, NULL);
}

/* siting.cpp:157,160: offset = (int)(((float)(2roi) + 0.99999f) * (float)r
 * / 2^31), evaluated in float in that order. */
static inline int target_offset(int roi, int32_t r) {
  float span = (float)(2 * roi) + 0.99999f;
  float prod = span * (float)r;
  float q = prod / 2147483648.0f;
  return (int)q;
}

/* siting.cpp:137-190 */
static uint8_t vix_one(int n, const uint16_t *e, int roi, int oht, int tht,
                       int ntests, int p, int64_t *nsamp) {
  int ox = p / n, oy = p % n;
  int oz = (int)e[p] + oht;
  int32_t r = p;
  int nvis = 0;
  int rr = roi * roi;
  for (int t = 0; t < ntests; t++) {
    r = so_lcg(r);
    int tx = target_offset(roi, r) + (ox - roi);
    r = so_lcg(r);
    int ty = target_offset(roi, r) + (oy - roi);
    int ddx = tx - ox, ddy = ty - oy;
    if (ddx * ddx + ddy * ddy > rr) { /* siting.cpp:162-166, C truncation */
      tx = ox + ddx / 3;
      ty = oy + ddy / 3;
    }
    if (tx < 0 || tx >= n || ty < 0 || ty >= n) continue; /* counted, not seen */
    int tz = (int)e[tx * n + ty] + tht;
    nvis += los_test(n, e, ox, oy, oz, tx, ty, tz, nsamp);
  }
  float v = (float)nvis / (float)ntests; /* siting.cpp:186,189 */
  int b = (int)(v * 255.999f);
  return (uint8_t)imin_(255, b);
}

void so_calc_vix(int nrows, const uint16_t *elevs, int roi, int oht, int tht,
                 int ntests, uint8_t *vix, int nthreads, int64_t *nsamples) {
  int64_t total = 0;
  int npts = nrows * nrows;
  if (nthreads > 1) {
#pragma omp parallel for schedule(guided) num_threads(nthreads) reduction(+ : total)
    for (int p = 0; p < npts; p++) {
      int64_t c = 0;
      vix[p] = vix_one(nrows, elevs, roi, oht, tht, ntests, p,
                       nsamples ? &c : NULL);
      total += c;
    }
  } else {
    for (int p = 0; p < npts; p++)
      vix[p] = vix_one(nrows, elevs, roi, oht, tht, ntests, p,
                       nsamples ? &total : NULL);
  }
  if (nsamples) *nsamples = total;
}

/* ------------------------------------------------------------------------ */
/* FINDMAX                                                                  */
/* ------------------------------------------------------------------------ */

/* siting.cpp:217,222: h = x*(x+y)*0x41041041 with 32-bit two's-complement
 * wrap (computed unsigned, so no signed overflow here). */
static inline int32_t block_hash(int x, int y) {
  return (int32_t)((uint32_t)x * (uint32_t)(x + y) * 0x41041041u);
}

/* siting.cpp:204-233: nwpb rounds; each round keeps the first cell of the
 * row-major scan that no later cell beats on (vix, hash), records it and
 * zeroes its vix. */
static void findmax_block(int n, uint8_t *vix, float bs, int bx, int by,
                          int nwpb, int32_t *out) {
  int x0 = (int)(bs * (float)bx);
  int x1 = imin_((int)(bs * (float)(bx + 1)), n);
  int y0 = (int)(bs * (float)by);
  int y1 = imin_((int)(bs * (float)(by + 1)), n);
  for (int k = 0; k < nwpb; k++) {
    int bestx = x0, besty = y0;
    uint8_t bv = vix[x0 * n + y0];
    int32_t bh = block_hash(x0, y0);
    for (int x = x0; x < x1; x++)
      for (int y = y0; y < y1; y++) {
        uint8_t v = vix[x * n + y];
        int32_t h = block_hash(x, y);
        if (bv < v || (bv == v && bh < h)) {
          bv = v;
          bh = h;
          bestx = x;
          besty = y;
        }
      }
    out[k] = bestx * n + besty;
    vix[bestx * n + besty] = 0;
  }
}

void so_find_max(int nrows, uint8_t *vix, float blocksize, int nblocks,
                 int nwpb, int32_t *obs, int nthreads) {
  int nb2 = nblocks * nblocks;
  if (nthreads > 1) {
#pragma omp parallel for schedule(guided) num_threads(nthreads)
    for (int b = 0; b < nb2; b++)
      findmax_block(nrows, vix, blocksize, b / nblocks, b % nblocks, nwpb,
                    obs + (size_t)b * nwpb);
  } else {
    for (int b = 0; b < nb2; b++)
      findmax_block(nrows, vix, blocksize, b / nblocks, b % nblocks, nwpb,
                    obs + (size_t)b * nwpb);
  }
}

/* ------------------------------------------------------------------------ */
/* VIEWSHED                                                                 */
/* ------------------------------------------------------------------------ */

/* MSB-first bit (row, col) of an nr x nwpr word shed, siting.cpp:247-250 */
static inline void shed_set(uint64_t *shed, int nwpr, int row, int col) {
  shed[row * nwpr + (col >> 6)] |= 1ull << (63 - (col & 63));
}

/* One ray from the observer to the perimeter cell at (ox+dx, oy+dy),
 * siting.cpp:313-345.  Steps every cell of the major axis up to and including
 * the target; stops at the terrain edge or outside the circle of radius roi. */
static int64_t cast_ray(int n, const uint16_t *e, int roi, int tht, int ox,
                        int oy, int oalt, int dx, int dy, uint64_t *shed,
                        int nwpr) {
  if (dx == 0 && dy == 0) return 0;
  int along_y = iabs_(dx) < iabs_(dy);
  int major = along_y ? dy : dx;
  int minor = along_y ? dx : dy;
  float slope = (float)minor / (float)major;
  int dir = major > 0 ? 1 : -1;
  float horizon = -99999.0f;
  int rr = roi * roi;
  int64_t steps = 0;
  for (int i = dir; i != major + dir; i += dir) {
    float fi = (float)i;
    float pm = fi * slope;
    int off = (int)roundf(pm);
    int px = along_y ? ox + off : ox + i;
    int py = along_y ? oy + i : oy + off;
    if (px < 0 || px >= n || py < 0 || py >= n) break;
    int qx = px - ox, qy = py - oy;
    if (qx * qx + qy * qy > rr) break;
    int pel = (int)e[px * n + py];
    int ai = iabs_(i);
    float fai = (float)ai;
    float s = (float)(pel - oalt) / fai;
    if (horizon < s) horizon = s;
    float rise = horizon * fai;
    float halt = (float)oalt + rise;
    steps++;
    if ((float)(pel + tht) >= halt) shed_set(shed, nwpr, qx + roi, qy + roi);
  }
  return steps;
}

/* siting.cpp:252-347: rays to the 8*roi cells of the box perimeter
 * (each corner once), plus the observer's own bit. */
static int64_t shed_one(int n, const uint16_t *e, int roi, int oht, int tht,
                        int32_t o, uint64_t *shed, int nwpr) {
  int ox = o / n, oy = o % n;
  int oalt = (int)e[o] + oht;
  int64_t steps = 0;
  shed_set(shed, nwpr, roi, roi);
  for (int j = 0; j < 2 * roi; j++) {
    /* bottom edge (y = -roi) x from -roi..roi-1; top edge (y = +roi) x from
     * -roi+1..roi; left edge (x = -roi) y from -roi+1..roi; right edge
     * (x = +roi) y from -roi..roi-1 (siting.cpp:292-304) */
    steps += cast_ray(n, e, roi, tht, ox, oy, oalt, -roi + j, -roi, shed, nwpr);
    steps += cast_ray(n, e, roi, tht, ox, oy, oalt, -roi + 1 + j, roi, shed, nwpr);
    steps += cast_ray(n, e, roi, tht, ox, oy, oalt, -roi, -roi + 1 + j, shed, nwpr);
    steps += cast_ray(n, e, roi, tht, ox, oy, oalt, roi, -roi + j, shed, nwpr);
  }
  return steps;
}

void so_calc_sheds(int nrows, const uint16_t *elevs, int roi, int oht,
                   int tht, int nsheds, const int32_t *obs, uint64_t *sheds,
                   int nthreads, int64_t *nsteps) {
  int nr = 2 * roi + 1, nwpr = (nr + 63) / 64;
  size_t sw = (size_t)nr * nwpr;
  int64_t total = 0;
  if (nthreads > 1) {
#pragma omp parallel for schedule(guided) num_threads(nthreads) reduction(+ : total)
    for (int k = 0; k < nsheds; k++) {
      uint64_t *sh = sheds + (size_t)k * sw;
      memset(sh, 0, sw * sizeof(uint64_t));
      total += shed_one(nrows, elevs, roi, oht, tht, obs[k], sh, nwpr);
    }
  } else {
    memset(sheds, 0, (size_t)nsheds * sw * sizeof(uint64_t));
    for (int k = 0; k < nsheds; k++)
      total += shed_one(nrows, elevs, roi, oht, tht, obs[k],
                        sheds + (size_t)k * sw, nwpr);
  }
  if (nsteps) *nsteps = total;
}

/* ------------------------------------------------------------------------ */
/* SITE                                                                     */
/* ------------------------------------------------------------------------ */

/* Split the column offset oy-roi into a cumulative-bitset word and a bit
 * shift with floor semantics (siting.cpp:400-405, 460-465). */
static inline void col_split(int c, int *word, int *bit) {
  int w = c / 64, b = c % 64;
  if (b < 0) {
    w -= 1;
    b += 64;
  }
  *word = w;
  *bit = b;
}

/* siting.cpp:393-419 */
void so_calc_union(int nrows, int roi, int ox, int oy, const uint64_t *shed,
                   uint64_t *cum) {
  int nr = 2 * roi + 1, nwpr = (nr + 63) / 64, cw = (nrows + 63) / 64;
  int w0, fb;
  col_split(oy - roi, &w0, &fb);
  for (int row = 0; row < nr; row++) {
    int x = ox - roi + row;
    if (x < 0 || x >= nrows) continue;
    uint64_t *dst = cum + (size_t)x * cw;
    const uint64_t *src = shed + (size_t)row * nwpr;
    for (int w = 0; w < nwpr; w++) {
      int c = w0 + w;
      if (c >= 0 && c < cw) dst[c] |= src[w] >> fb;
      if (fb != 0 && c + 1 >= 0 && c + 1 < cw) dst[c + 1] |= src[w] << (64 - fb);
    }
  }
}

/* siting.cpp:421-428 */
static inline int cum_test(const uint64_t *cum, int cw, int x, int y) {
  return (int)((cum[(size_t)x * cw + (y >> 6)] >> (63 - (y & 63))) & 1u);
}

/* siting.cpp:456-491 for one candidate: popcount of the shed bits, shifted
 * into terrain columns, that the cumulative bitset does not hold yet. */
int so_union_area(int nrows, int roi, int ox, int oy, const uint64_t *shed,
                  const uint64_t *cum) {
  int nr = 2 * roi + 1, nwpr = (nr + 63) / 64, cw = (nrows + 63) / 64;
  int w0, fb;
  col_split(oy - roi, &w0, &fb);
  int wlast = (oy + roi) / 64;
  int total = 0;
  for (int row = 0; row < nr; row++) {
    int x = ox - roi + row;
    if (x < 0 || x >= nrows) continue;
    const uint64_t *src = shed + (size_t)row * nwpr;
    const uint64_t *crow = cum + (size_t)x * cw;
    for (int c = (w0 > 0 ? w0 : 0); c <= wlast && c < cw; c++) {
      int w = c - w0;
      uint64_t cur = w < nwpr ? src[w] : 0;
      uint64_t prev = (w > 0) ? src[w - 1] : 0;
      uint64_t moved = cur >> fb;
      if (fb != 0) moved |= prev << (64 - fb);
      total += __builtin_popcountll(moved & ~crow[c]);
    }
  }
  return total;
}

/* siting.cpp:501-626 with the sequential argmax (555-562).  The update list
 * (540-544) is every unused candidate on the first iteration and, after
 * that, the unused candidates within 2*roi of the last pick; only they get a
 * new area, and under intervis an invisible one gets 0 (452-457).  Every
 * other candidate keeps its last area. */
void so_site(int nrows, int roi, int intervis, int nsheds, const int32_t *obs,
             const uint64_t *sheds, uint8_t *selected, int32_t *order,
             so_site_stats *st, int nthreads) {
  int nr = 2 * roi + 1, nwpr = (nr + 63) / 64, cw = (nrows + 63) / 64;
  size_t sw = (size_t)nr * nwpr;
  uint8_t *used = (uint8_t *)calloc((size_t)nsheds, 1);
  int32_t *area = (int32_t *)calloc((size_t)nsheds, sizeof(int32_t));
  uint64_t *cum = (uint64_t *)calloc((size_t)nrows * cw, sizeof(uint64_t));
  int nused = 0, niter = 0, lastx = 0, lasty = 0;
  int64_t cumarea = 0, sum_u = 0;
  int r2 = (2 * roi) * (2 * roi);
  double npts = (double)nrows * (double)nrows;
  int nth = nthreads > 1 ? nthreads : 1;

  for (;;) {
    niter++;
    int64_t ucount = 0;
#pragma omp parallel for schedule(guided) num_threads(nth) reduction(+ : ucount) if (nth > 1)
    for (int b = 0; b < nsheds; b++) {
      if (used[b]) continue;
      int bx = obs[b] / nrows, by = obs[b] % nrows;
      if (nused != 0) {
        int ex = bx - lastx, ey = by - lasty;
        if (ex * ex + ey * ey > r2) continue;
      }
      ucount++;
      if (intervis && nused > 0 && !cum_test(cum, cw, bx, by))
        area[b] = 0;
      else
        area[b] = so_union_area(nrows, roi, bx, by, sheds + (size_t)b * sw, cum);
    }
    sum_u += ucount;

    /* max area, lowest index (strictly greater than 0) */
    int best = -1, bestarea = 0;
    if (nth > 1) {
      int tbest[1024], tarea[1024];
      int nt = nth > 1024 ? 1024 : nth;
#pragma omp parallel num_threads(nt)
      {
        int id = 0, cnt = 1;
#ifdef _OPENMP
        id = omp_get_thread_num();
        cnt = omp_get_num_threads();
#endif
        int chunk = (nsheds + cnt - 1) / cnt;
        int lo = id * chunk, hi = lo + chunk < nsheds ? lo + chunk : nsheds;
        int mb = -1, ma = 0;
        for (int b = lo; b < hi; b++)
          if (!used[b] && area[b] > ma) {
            ma = area[b];
            mb = b;
          }
        tbest[id] = mb;
        tarea[id] = ma;
        if (id == 0)
          for (int t = cnt; t < nt; t++) tbest[t] = -1, tarea[t] = 0;
      }
      for (int t = 0; t < nt; t++)
        if (tarea[t] > bestarea) {
          bestarea = tarea[t];
          best = tbest[t];
        }
    } else {
      for (int b = 0; b < nsheds; b++)
        if (!used[b] && area[b] > bestarea) {
          bestarea = area[b];
          best = b;
        }
    }
    if (bestarea == 0) break;

    int wx = obs[best] / nrows, wy = obs[best] % nrows;
    so_calc_union(nrows, roi, wx, wy, sheds + (size_t)best * sw, cum);
    if (order) order[nused] = best;
    nused++;
    used[best] = 1;
    lastx = wx;
    lasty = wy;
    cumarea += bestarea;
    if (100.0 * (double)cumarea / npts > 95.0) break;
  }

  for (int b = 0; b < nsheds; b++) selected[b] = used[b];
  if (st) {
    st->nused = nused;
    st->niter = niter;
    st->cumarea = cumarea;
    st->sum_u = sum_u;
    st->coverage = 100.0 * (double)cumarea / npts;
  }
  free(used);
  free(area);
  free(cum);
}

/* The four stages in order, as main() runs them (siting.cpp:737-755). */
int so_run(int nrows, const uint16_t *elevs, int roi, int ht, int ntests,
           int bw, int nwanted0, int intervis, uint8_t *vix, int32_t *obs,
           uint64_t *sheds, uint8_t *selected, so_site_stats *stats,
           double *stage_s, int nthreads) {
  so_derived d;
  if (so_derive(nrows, bw, nwanted0, &d)) return 1;
  double t0 = wall_now();
  so_calc_vix(nrows, elevs, roi, ht, ht, ntests, vix, nthreads, NULL);
  double t1 = wall_now();
  so_find_max(nrows, vix, d.blocksize, d.nblocks, d.nwpb, obs, nthreads);
  double t2 = wall_now();
  so_calc_sheds(nrows, elevs, roi, ht, ht, d.nwanted, obs, sheds, nthreads, NULL);
  double t3 = wall_now();
  so_site(nrows, roi, intervis, d.nwanted, obs, sheds, selected, NULL, stats,
          nthreads);
  double t4 = wall_now();
  if (stage_s) {
    stage_s[0] = t1 - t0;
    stage_s[1] = t2 - t1;
    stage_s[2] = t3 - t2;
    stage_s[3] = t4 - t3;
  }
  return 0;
}
