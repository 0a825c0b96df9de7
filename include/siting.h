/*
 * siting.h -- C ABI of the MI355X-native siting library (libsiting_amd.so).
 *
 * Four stages, as in the reference (siting.cpp / siting.cu):
 *   VIX       approximate visibility index per terrain point
 *   FINDMAX   top vix points per block become tentative observers
 *   VIEWSHED  one packed MSB-first bitset per tentative observer
 *   SITE      greedy max-coverage selection over the bitsets
 *
 * Plain pointers and sizes only.  Results are bit-exact against the
 * reference's sequential build (siting.cpp -DSEQ): vix bytes, obs indices,
 * every shed word, the selected mask and the coverage.
 *
 * Errors are returned as siting_status codes instead of the reference's
 * die()/exit(1) (siting.cpp:22-25); siting_status_str() gives the message.
 *
 * Data layouts (identical to the reference's host arrays):
 *   elevs     uint16[nrows*nrows], elevs[x*nrows+y], x = row     (cpp:709)
 *   vix       uint8 [nrows*nrows]                                (cpp:710)
 *   obs       int32 [nwanted], obs[(bx*nblocks+by)*nwpb+k]=x*nrows+y (cpp:711)
 *   sheds     uint64[nwanted*nr*nwpr], nr = 2*roi+1, nwpr = ceil(nr/64);
 *             bit (row,col) of shed k is word k*nr*nwpr + row*nwpr + col/64,
 *             bit 63 - col%64; row r is terrain row x_obs - roi + r  (cpp:247-250)
 *   selected  uint8 [nwanted], 1 for a selected observer          (cpp:713)
 */
#ifndef SITING_H
#define SITING_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SITING_OK = 0,
  SITING_EINVAL = 1,      /* bad pointer / size / parameter outside what the
                             library can run */
  SITING_ERANGE = 2,      /* parameter outside the reference CLI's ranges
                             (siting.cpp:680-686); siting_check_cli only */
  SITING_ESMALLBLOCK = 3, /* "The last block is too small for
                             nwantedperblock." (siting.cpp:694-695) */
  SITING_ENOMEM = 4,      /* device or host allocation failed */
  SITING_EDEVICE = 5,     /* HIP runtime error or no usable GPU */
  SITING_EIO = 6,         /* "Input failed" (siting.cpp:726-729) / write error */
  SITING_EBUSY = 7,       /* server: no admission possible right now */
  SITING_ENOTFOUND = 8    /* server: unknown job id */
} siting_status;

const char *siting_status_str(siting_status s);

/* Run parameters: the reference's argv (siting.cpp:668-674).  The CLI sets
 * oht = tht = height (cpp:670); the library keeps them separate. */
typedef struct {
  int nrows;
  int roi;
  int oht;
  int tht;
  int ntests;   /* "targets" */
  int bw;       /* requested block width (blocksize0) */
  int nwanted0; /* requested tentative observers */
  int intervis; /* 0 or 1 */
} siting_params;

/* Derived sizes (siting.cpp:689-702). */
typedef struct {
  float blocksize; /* (float)n / (int)((float)n/bw + 0.5f)        cpp:689 */
  int nblocks;     /* (int)(n/blocksize + 0.5f)                   cpp:690 */
  int nwpb;        /* (int)((float)nwanted0/nblocks^2 + 0.99999f) cpp:691 */
  int nwanted;     /* nwpb * nblocks^2                            cpp:692 */
  int lastsize;    /* n - (int)(blocksize*(nblocks-1))            cpp:693 */
  int nr;          /* 2*roi + 1                                   cpp:701 */
  int nwpr;        /* (nr + 63) / 64                              cpp:702 */
  int cumnwpr;     /* (n + 63) / 64                               cpp:505 */
  int64_t shed_words; /* nr * nwpr                                          */
} siting_derived;

/* Per-job counters and device stage times. */
typedef struct {
  int nused;        /* selected observers (" nusedsheds:N", cpp:617) */
  int niter;        /* SITE loop iterations, the final one included */
  int64_t cumarea;  /* covered terrain points */
  int64_t sum_u;    /* sum over iterations of |update list| (cpp:540-544) */
  double coverage;  /* 100.0*cumarea/nrows^2 (" coverage:C", cpp:617) */
  float ms_vix, ms_findmax, ms_viewshed, ms_site; /* HIP event times */
  int rounds;       /* rounds of the device's greedy loop: each takes one or
                       more picks (at most one under intervis) */
} siting_stats;

/* Kernel shape choices of one job.  Any choice gives the same bits; only
 * the speed differs.  0 in a field means the built-in rule. */
typedef struct {
  int vix_halo;          /* VIX: cells of DEM halo around a 16x16 tile kept
                            in LDS: 8, 16 or 32 */
  int viewshed_threads;  /* VIEWSHED: threads per observer, a multiple of 64
                            in 64..1024 */
  int viewshed_mode;     /* VIEWSHED bitsets: 1 two LDS bitsets (row-major and
                            transposed), 2 one LDS bitset, 3 in HBM; a mode
                            whose bitsets do not fit falls back to the rule */
  int site_q;            /* SITE: column classes per cell row, 1..8 */
  int site_tau_shift;    /* SITE: update candidates below head - head/2^s
                            become stale bounds, s in 1..20 */
  int site_walk_evals;   /* SITE, intervis: 1 + the exact area computations
                            a round's pick walk may make to admit another
                            pick (1: none), 1..33 */
} siting_tuning;

/* ---- parameters ------------------------------------------------------- */

/* siting.cpp:689-702 without the CLI's range checks (so bw < 10 runs, as a
 * library call of the reference's stage functions does).  Returns
 * SITING_ESMALLBLOCK where the reference dies with "last block too small",
 * SITING_EINVAL for nonsensical values (nrows, roi, ntests < 1, bw < 1,
 * roi > 10000 (the CLI's bound), intervis not 0/1, nrows/bw rounding to
 * zero blocks, sizes whose arrays overflow 32-bit indexing). */
siting_status siting_derive(const siting_params *p, siting_derived *d);

/* The reference CLI's range checks (siting.cpp:680-686), in its order.
 * Returns SITING_OK or SITING_ERANGE with *which = 1..7 (nrows, roi, tht,
 * ntests, bw, nwanted0, intervis) and *msg = the reference's die() text. */
siting_status siting_check_cli(const siting_params *p, int *which,
                               const char **msg);

/* ---- stage functions on host arrays (drop-in for the reference's) -----
 * Each uploads its inputs, runs the HIP kernels on the library's default
 * context (device 0) and downloads its outputs, like cusite's host
 * wrappers (siting.cu:133,226,336,596) but with status returns. */

/* calc_vix, siting.cpp:192 / siting.cu:133 */
siting_status siting_calc_vix(int nrows, const uint16_t *elevs, int roi,
                              int oht, int tht, int ntests, uint8_t *vix);

/* find_max, siting.cpp:235 / siting.cu:226.  vix is mutated as in the
 * sequential reference (picked cells set to 0, cpp:231). */
siting_status siting_find_max(int nrows, uint8_t *vix, float blocksize,
                              int nwanted, int nblocks, int nwpb,
                              int32_t *obs);

/* calc_sheds, siting.cpp:349 / siting.cu:336 */
siting_status siting_calc_sheds(int nrows, const uint16_t *elevs, int roi,
                                int oht, int tht, int nsheds,
                                const int32_t *obs, uint64_t *sheds);

/* site_it, siting.cpp:501 / siting.cu:596, without the stdout print (the
 * counters come back in *stats, which may be NULL). */
siting_status siting_site_it(int nrows, int roi, int intervis, int nsheds,
                             const int32_t *obs, const uint64_t *sheds,
                             uint8_t *selected, siting_stats *stats);

/* site_it on device arrays, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory or create
 * streams -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation.  It runs the same kernels as siting_site_it with
 * the same results, taking its working memory (candidate areas, the
 * cumulative bitset, the loop's state and scratch) from a fixed 64 MiB
 * device array the library holds on every device it runs on until the
 * process ends; a call that needs more returns SITING_ENOMEM, and
 * siting_status_str then names the size it needed.  A call uses about
 * 31-70 bytes per candidate, nrows^2/8 bytes and 256 (2 roi + 1)
 * (ceil((2 roi + 1)/64) + 1) bytes (1.4 M candidates on 8192^2 at roi 25
 * fit; so do 262144 on 8192^2 at roi 200).
 *   obs, sheds  device arrays of the current device (layouts above); obs
 *               entries must lie in [0, nrows^2), as FINDMAX writes them
 *               (they are not checked)
 *   selected    device uint8[nsheds], the output mask
 *   order       device int32[nsheds] receiving the selection order, or NULL
 *   stats       NULL: the call only enqueues work on `stream` and returns;
 *               else it waits for the stream and fills the counters
 *   stream      a hipStream_t of the current device, or NULL (null stream)
 * All calls on one device share the one array: calls that overlap on a
 * device (other streams, other threads) race and must be serialised by the
 * caller.  With stats NULL the call can be captured into a HIP graph. */
siting_status siting_site_it_noalloc(int nrows, int roi, int intervis, int nsheds,
                                     const int32_t *obs, const uint64_t *sheds,
                                     uint8_t *selected, int32_t *order,
                                     siting_stats *stats, void *stream);

/* ---- device-resident jobs --------------------------------------------
 * A job owns its device buffers (DEM, vix, obs, sheds, SITE state), so the
 * four stages run back to back in HBM with no host round trip.  All work of
 * a job goes to one HIP stream owned by the job. */

typedef struct siting_job siting_job;

/* Device memory a job needs (bytes), for admission control. */
siting_status siting_job_bytes(const siting_params *p, size_t *bytes);

siting_status siting_job_create(int device, const siting_params *p,
                                siting_job **out);
void siting_job_destroy(siting_job *job);

/* Copy a host DEM (nrows^2 uint16) into the job's device buffer. */
siting_status siting_job_upload_dem(siting_job *job, const uint16_t *elevs);

/* Enqueue the stages (asynchronous on the job's stream). */
siting_status siting_job_vix(siting_job *job);
siting_status siting_job_findmax(siting_job *job);
siting_status siting_job_viewshed(siting_job *job);
siting_status siting_job_site(siting_job *job);
/* All four, in order. */
siting_status siting_job_run(siting_job *job);

/* Wait for the job's stream. */
siting_status siting_job_sync(siting_job *job);

/* Buffers that can be read back (after siting_job_sync). */
typedef enum {
  SITING_BUF_ELEVS = 0,
  SITING_BUF_VIX = 1,
  SITING_BUF_OBS = 2,
  SITING_BUF_SHEDS = 3,
  SITING_BUF_SELECTED = 4,
  SITING_BUF_ORDER = 5 /* int32[nused]: selection order */
} siting_buffer;

siting_status siting_job_download(siting_job *job, siting_buffer which,
                                  void *host, size_t bytes);
/* Device address of a job buffer (for callers that keep data in HBM). */
siting_status siting_job_device_ptr(siting_job *job, siting_buffer which,
                                    void **dptr, size_t *bytes);
/* Kernel shape choices for the job's later stage launches (NULL: the
 * built-in rules).  A job starts with the tuning cache's entry for its
 * parameters, when there is one (siting_tuning_lookup). */
siting_status siting_job_set_tuning(siting_job *job, const siting_tuning *t);
siting_status siting_job_get_tuning(siting_job *job, siting_tuning *t);
/* Counters + per-stage device times of the last run (after sync). */
siting_status siting_job_stats(siting_job *job, siting_stats *st);
siting_status siting_job_derived(siting_job *job, siting_derived *d);

/* ---- batched jobs -------------------------------------------------------
 * Many independent jobs (any mix of sizes and parameters) run as one unit:
 * the per-job VIX, FINDMAX, VIEWSHED and SITE-preparation kernels go to a
 * small pool of streams so they overlap, and every job's greedy SITE loop
 * runs in a single persistent work-queue launch whose slots of workgroups
 * take jobs, largest first, as they free up.  The slot plan is tuned by
 * measurement over the first synced runs (equal slots, then the three best
 * mixed plans of a model of each job's measured loop time, then the job
 * that ran longest moved onto a larger slot of its own, up to seven runs),
 * and the fastest plan measured is kept.  The prep sequence is captured into a HIP graph and
 * replayed.  Results are bit-identical to running each job
 * alone, whatever slot runs it. */

typedef struct siting_batch siting_batch;

typedef struct {
  int streams;   /* stream pool size (0: 4) */
  int slot;      /* workgroups per SITE slot (0: tuned default) */
  int nslots;    /* SITE slots (0: CUs / slot) */
  int use_graph; /* 1: capture the run into a HIP graph on first use */
  int fixed_slots; /* 1: keep equal slots; 0: tune the slot plan over the
                      first synced runs from each job's measured SITE loop
                      time and keep the fastest plan measured */
} siting_batch_opts;

siting_status siting_batch_create(int device, int njobs, const siting_params *params,
                                  const siting_batch_opts *opts, siting_batch **out);
void siting_batch_destroy(siting_batch *b);
siting_status siting_batch_upload_dem(siting_batch *b, int job, const uint16_t *elevs);
/* Enqueue all stages of all jobs (asynchronous). */
siting_status siting_batch_run(siting_batch *b);
siting_status siting_batch_sync(siting_batch *b);
/* Device time of the last run, from events around it (after sync). */
siting_status siting_batch_elapsed_ms(siting_batch *b, float *ms);
/* Device time of the last run's SITE work-queue launch alone (after sync). */
siting_status siting_batch_site_ms(siting_batch *b, float *ms);
/* Start and end of every job's SITE loop in the last run (after sync), in
 * ms from the first loop's start, on the device clock: start_end[2*i],
 * start_end[2*i+1] for job i (njobs pairs). */
siting_status siting_batch_site_times(siting_batch *b, double *start_end, int njobs);
/* Borrowed handle of job i, for downloads and stats (valid until destroy). */
siting_status siting_batch_job(siting_batch *b, int job, siting_job **out);

/* ---- serving -----------------------------------------------------------
 * A request loop on one GPU: submitted jobs are admitted as device memory
 * and SITE workgroups free up, run alongside the jobs already on the device
 * (each on its own stream, all data resident in HBM) and retire into host
 * result buffers.  First-come first-served with backfill of smaller requests
 * while the head of the queue waits (up to backfill_ms).  A job's results
 * are bit-identical to running it alone. */

typedef struct siting_server siting_server;

typedef struct {
  size_t hbm_bytes;   /* device memory for jobs (0: 80% of free memory) */
  int max_active;     /* jobs on the device at once (0: 8) */
  int site_groups;    /* workgroups of each job's SITE loop (0: 32) */
  int backfill_ms;    /* how long smaller requests may pass a blocked head (0: 200) */
} siting_server_opts;

typedef enum {
  SITING_JOB_QUEUED = 0,
  SITING_JOB_RUNNING = 1,
  SITING_JOB_DONE = 2,
  SITING_JOB_FAILED = 3
} siting_job_state;

typedef struct {
  siting_job_state state;
  siting_status status;   /* why it failed, when it did */
  siting_stats stats;     /* SITE counters (stage times are not recorded) */
  double t_submit, t_admit, t_done; /* seconds on the server's clock */
} siting_job_info;

typedef struct {
  int queued, running, groups_in_use;
  size_t hbm_bytes;
} siting_server_load_info;

siting_status siting_server_create(int device, const siting_server_opts *opts,
                                   siting_server **out);
/* Drops queued requests (FAILED, SITING_EBUSY), waits for running ones,
 * then for every thread already blocked in siting_server_wait to return
 * (they see their request DONE or FAILED).  No other call on s may start
 * once destroy has begun. */
void siting_server_destroy(siting_server *s);
/* Copies the DEM.  Rejects malformed parameters (SITING_EINVAL /
 * SITING_ESMALLBLOCK) and jobs larger than the whole budget
 * (SITING_ENOMEM) at once; the server keeps running. */
siting_status siting_server_submit(siting_server *s, const siting_params *p,
                                   const uint16_t *elevs, uint64_t *id);
siting_status siting_server_info(siting_server *s, uint64_t id, siting_job_info *info);
/* Blocks until done or failed (timeout_s < 0: no limit; SITING_EBUSY on
 * timeout). */
siting_status siting_server_wait(siting_server *s, uint64_t id, double timeout_s,
                                 siting_job_info *info);
/* selected: uint8[nwanted]; obs: int32[nwanted] (either may be NULL). */
siting_status siting_server_result(siting_server *s, uint64_t id, uint8_t *selected,
                                   size_t nsel, int32_t *obs, size_t nobs);
/* Removes a queued request (SITING_EBUSY if it already started). */
siting_status siting_server_cancel(siting_server *s, uint64_t id);
/* Frees a finished request's results. */
siting_status siting_server_release(siting_server *s, uint64_t id);
siting_status siting_server_load(siting_server *s, siting_server_load_info *out);

/* ---- terrain input (raw uint16 format, siting.cpp:726-730) ------------ */

/* Exactly nrows*nrows native-endian uint16 values; SITING_EIO if short. */
siting_status siting_dem_read(const char *path, int nrows, uint16_t *out);
siting_status siting_dem_write(const char *path, int nrows,
                               const uint16_t *dem);
/* Seeded diamond-square terrain with roughness rough_h, rescaled linearly to
 * [hmin, hmax], plus nhills Gaussian hills (peak, sigma in cells). */
siting_status siting_dem_generate(int nrows, double rough_h, int hmin,
                                  int hmax, int nhills, double hill_peak,
                                  double hill_sigma, uint64_t seed,
                                  uint16_t *out);

/* ---- tuning cache --------------------------------------------------------
 * Kernel shape choices measured by the project's tuner (python -m
 * siting_amd.tune) for classes of jobs, kept in a text file: one line per
 * (device, roi class, targets class, edge class), written by the tuner.
 * The file is the one SITING_TUNE_CACHE names ("off" disables the cache),
 * else siting_amd/tune/tuning.txt beside the library.  Classes: roi 25
 * (roi <= 37), 50 (<= 75), 100 (<= 125), 150 (<= 175), 200 (<= 300);
 * targets 32 (<= 48) or 64; edge 1024 (nrows <= 1536), 2048 (<= 3072) or
 * 4096. */

/* "<arch>:<CUs>" of a device, e.g. "gfx950:256", the cache's device key. */
siting_status siting_device_key(int device, char *buf, size_t size);
/* The cached choices for parameters p on the device with this key:
 * SITING_OK, or SITING_ENOTFOUND (no cache, or no entry; *out = zeros, the
 * built-in rules). */
siting_status siting_tuning_lookup(const char *device_key, const siting_params *p,
                                   siting_tuning *out);
/* The cache file in use ("" when disabled). */
const char *siting_tuning_cache_path(void);

/* ---- library info ----------------------------------------------------- */

/* Number of visible GPUs (0 when none); never fails. */
int siting_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* SITING_H */
