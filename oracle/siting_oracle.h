/*
 * siting_oracle.h -- CPU restatement of the reference's sequential siting
 * pipeline (siting.cpp built with -DSEQ).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  The product (siting_amd/) never links, loads or calls it.
 *
 * Every function cites the siting.cpp lines it restates.  The arithmetic is
 * kept in the reference's order and precision and the library is compiled
 * with -ffp-contract=off, so no multiply-add is fused (siting.cpp is built
 * for x86-64 baseline by g++ -O2, which emits no FMA).
 *
 * The *_threads arguments select the OpenMP variant (nthreads > 1) used as
 * the CPU timing baseline; every variant returns the same bits as the
 * sequential one (the SITE argmax reduction keeps the lowest-index tie rule
 * of the sequential path, siting.cpp:555-562).
 */
#ifndef SITING_ORACLE_H
#define SITING_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  float blocksize;  /* perturbed block edge, siting.cpp:689 */
  int nblocks;      /* blocks per side, siting.cpp:690 */
  int nwpb;         /* tentative observers per block, siting.cpp:691 */
  int nwanted;      /* nwpb * nblocks^2, siting.cpp:692 */
  int lastsize;     /* edge of the last block, siting.cpp:693 */
} so_derived;

typedef struct {
  int nused;          /* selected observers (nusedsheds) */
  int niter;          /* greedy loop iterations, including the final one */
  int64_t cumarea;    /* covered points */
  int64_t sum_u;      /* sum over iterations of |update list| */
  double coverage;    /* 100 * cumarea / nrows^2 */
} so_site_stats;

/* glibc TYPE_0 step, siting.cpp:29-33 */
int32_t so_lcg(int32_t seed);

/* Parameter derivation, siting.cpp:689-695.  Returns 0, or 1 when the last
 * block is too small (the reference dies there).  No range checks. */
int so_derive(int nrows, int bw, int nwanted0, so_derived *out);

/* Range checks of main(), siting.cpp:680-686.  Returns 0 or the 1-based
 * index of the first failing check (nrows, roi, tht, ntests, bw, nwanted,
 * intervis). */
int so_check_args(int nrows, int roi, int ht, int ntests, int bw, int nwanted0,
                  int intervis);

/* VIX: siting.cpp:54-198.  Optional counter: LOS samples read (may be NULL). */
void so_calc_vix(int nrows, const uint16_t *elevs, int roi, int oht, int tht,
                 int ntests, uint8_t *vix, int nthreads, int64_t *nsamples);

/* Line-of-sight test of one target, siting.cpp:54-135 (exposed for tests). */
int so_test_one_target(int nrows, const uint16_t *elevs, int ox, int oy,
                       int oz, int tx, int ty, int tz);

/* FINDMAX: siting.cpp:204-241.  vix is mutated (picked cells set to 0). */
void so_find_max(int nrows, uint8_t *vix, float blocksize, int nblocks,
                 int nwpb, int32_t *obs, int nthreads);

/* VIEWSHED: siting.cpp:247-360.  sheds holds nsheds*nr*nwpr words and is
 * cleared first.  Optional counter: ray steps taken (may be NULL). */
void so_calc_sheds(int nrows, const uint16_t *elevs, int roi, int oht,
                   int tht, int nsheds, const int32_t *obs, uint64_t *sheds,
                   int nthreads, int64_t *nsteps);

/* SITE: siting.cpp:366-626, sequential argmax rule.  order (may be NULL,
 * else nsheds entries) receives the selection order. */
void so_site(int nrows, int roi, int intervis, int nsheds, const int32_t *obs,
             const uint64_t *sheds, uint8_t *selected, int32_t *order,
             so_site_stats *stats, int nthreads);

/* Extra area of one candidate against a cumulative bitset, siting.cpp:430-499
 * with intervis=0 (exposed for tests). */
int so_union_area(int nrows, int roi, int ox, int oy, const uint64_t *shed,
                  const uint64_t *cumshed);

/* OR one shed into the cumulative bitset, siting.cpp:393-419. */
void so_calc_union(int nrows, int roi, int ox, int oy, const uint64_t *shed,
                   uint64_t *cumshed);

/* Whole pipeline on one DEM (the seqsite run without file I/O).
 * vix (nrows^2), obs/selected (nwanted) and sheds (nwanted*nr*nwpr) are
 * caller-allocated; pass the so_derive() result's nwanted.  stage_s
 * (may be NULL) receives 4 wall times in seconds. */
int so_run(int nrows, const uint16_t *elevs, int roi, int ht, int ntests,
           int bw, int nwanted0, int intervis, uint8_t *vix, int32_t *obs,
           uint64_t *sheds, uint8_t *selected, so_site_stats *stats,
           double *stage_s, int nthreads);

#ifdef __cplusplus
}
#endif

#endif
