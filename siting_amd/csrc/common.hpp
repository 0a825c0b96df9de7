// common.hpp -- shared definitions of the siting HIP library (gfx950).
//
// Every kernel is compiled with -ffp-contract=off: the reference's
// sequential build (g++ -O2, x86-64 baseline) never fuses a multiply and an
// add, so no kernel may either (siting.cpp:96,106,119,129,342).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "siting.h"

namespace siting {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
#define SITING_HIP(call)                                   \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return ::siting::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

siting_status hip_fail(hipError_t e, const char *what, const char *file, int line);
// The SITE loop gave up (niter < 0): -1 a grid barrier timed out, -2 its
// workgroups disagreed on a round's picks.  Sets the message, returns
// SITING_EDEVICE.
siting_status site_loop_failed(int niter);
// Message of the next SITING_ENOMEM on this thread (nullptr or "" clears it).
void set_limit_error(const char *msg);

// ---------------------------------------------------------------------------
// integer helpers shared by host and device
// ---------------------------------------------------------------------------

// glibc TYPE_0 generator step on 31 bits, siting.cpp:29-33.
__host__ __device__ inline uint32_t lcg_step(uint32_t r) {
  return (1103515245u * (r & 0x7fffffffu) + 12345u) & 0x7fffffffu;
}

// FINDMAX tie hash h = x*(x+y)*0x41041041 with int32 wrap (siting.cpp:217).
__host__ __device__ inline int32_t block_hash(int x, int y) {
  return (int32_t)((uint32_t)x * (uint32_t)(x + y) * 0x41041041u);
}

// Floor split of a (possibly negative) column into a 64-bit word index and a
// bit shift (siting.cpp:400-405).
__host__ __device__ inline void col_split(int c, int &word, int &bit) {
  int w = c / 64, b = c % 64;
  if (b < 0) {
    w -= 1;
    b += 64;
  }
  word = w;
  bit = b;
}

// 0.5 - 2^-25, the largest float below one half.  libm round() of a
// non-negative float x below 2^15 (siting.cpp:95,330 round half away from
// zero) is trunc(fl(x + kHalfDown)): checked for every such float.
constexpr float kHalfDown = 0.49999997f;

// The largest roi the library accepts: the reference CLI's bound
// (siting.cpp:680-686).  It keeps VIEWSHED's per-roi tables (about
// 0.8 GB at the bound) within the 32-bit offsets of buffer loads.
constexpr int kMaxRoi = 10000;

// ---------------------------------------------------------------------------
// wave / block reductions
// ---------------------------------------------------------------------------
// Sum over the wavefront, in every lane (all lanes active), with DPP row
// operations: quad swaps, half-row and row mirrors (every lane then holds
// its row's sum), then the row broadcasts that fold rows 0-3 into lane 63.
// A lane of a row a broadcast leaves out adds 0.
template <int kCtrl, int kRowMask>
__device__ __forceinline__ int dpp_add_step(int v) {
  return v + __builtin_amdgcn_update_dpp(0, v, kCtrl, kRowMask, 0xf, false);
}

__device__ __forceinline__ int wave_sum_i32(int v) {
  v = dpp_add_step<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v = dpp_add_step<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v = dpp_add_step<0x141, 0xf>(v);  // row_half_mirror
  v = dpp_add_step<0x140, 0xf>(v);  // row_mirror
  v = dpp_add_step<0x142, 0xa>(v);  // row_bcast:15 into rows 1, 3
  v = dpp_add_step<0x143, 0xc>(v);  // row_bcast:31 into rows 2, 3
  return __builtin_amdgcn_readlane(v, 63);
}

// ---------------------------------------------------------------------------
// device-side SITE state (one per job)
// ---------------------------------------------------------------------------
struct SiteState {
  unsigned long long cumarea;
  unsigned long long sum_u;
  int nused;
  int niter;
  unsigned int barrier;  // grid barrier arrival counter
  int rounds;            // rounds of the persistent loop (picks are batched)
};

// ---------------------------------------------------------------------------
// job geometry (host side), derived once per job
// ---------------------------------------------------------------------------
struct Geometry {
  siting_params p;
  siting_derived d;
  siting_tuning t;  // kernel shape choices (zeros: the built-in rules)
};

// Class of a job in the tuning cache, and the lookup (api.hip).
siting_status tuning_lookup(const char *device_key, const siting_params &p,
                            siting_tuning *out);

siting_status derive(const siting_params *p, siting_derived *d);

// ---------------------------------------------------------------------------
// kernel launchers (one per stage), all asynchronous on `stream`
// ---------------------------------------------------------------------------
struct VixTables {
  const float *recip;      // (float)(1.0/d) for d = 0..roi (entry 0 unused)
  const uint32_t *jump;    // LCG jump-ahead: (A_m, C_m) for m = 1..2*ntests
};

siting_status launch_vix(const Geometry &g, const uint16_t *elevs, uint8_t *vix,
                         const VixTables &t, hipStream_t s);
siting_status launch_findmax(const Geometry &g, uint8_t *vix, int32_t *obs,
                             hipStream_t s);
// VIEWSHED's per-roi tables, one block of vs_tables(roi).bytes bytes.  The
// walk takes a ray's steps in groups of four, group g holding steps
// a = 4g+1..4g+4, g = 0..groups-1 (it reads group g+1's offsets while it
// runs group g, so one group past the last that holds a step <= roi).
//   off   per group, per signed minor coordinate m = -roi..roi of a ray's
//         perimeter cell (major coordinate +-roi), four uint32, one per
//         step: 2 (o + roi), o = the ray's minor offset at step a,
//         sign(m) * round-half-away(fl(a * fl(|m| / roi)))
//         (siting.cpp:323-330); 2 roi past roi
//   disk  per signed minor coordinate m, uint32: the last step of the ray
//         inside the roi disk (the steps inside form a prefix)
//   step  per ray class (major axis y or x, times direction -1 or +1),
//         chunk of 64 rays of a side (in the split-bitset mode; one chunk
//         otherwise) and group, a VsStep4 of the group's wavefront-uniform
//         values
struct VsStep4 {
  double rd[4];      // 1.0 / a (0 past roi)
  float fa[4];       // (float)a
  uint32_t wofs[4];  // split bitsets: byte offset of the word of major
                     // coordinate c = roi + dir * a within a row, plus the
                     // transposed bitset's base for x-major rays (0 past roi)
  uint32_t wbit[4];  // its bit, 1 << (c & 31) (0 past roi)
  uint64_t same[4];  // split bitsets: bit l set when lane l's cell at step a
                     // is lane l-1's (0 past roi)
};
// The split-bitset mode: two LDS bitsets of 32-bit words, 2 nwpr + 1 dwords
// a row, when both fit in 48 KiB.
__host__ __device__ inline int vs_split_stride(int nwpr) { return 2 * nwpr + 1; }
__host__ __device__ inline bool vs_split(int roi) {
  const int nr = 2 * roi + 1;
  return 2 * 4 * (size_t)nr * vs_split_stride((nr + 63) / 64) <= 48 * 1024;
}
struct VsTables {
  size_t off, step, disk, bytes;
  int groups, chunks;
};
__host__ __device__ inline VsTables vs_tables(int roi) {
  VsTables t;
  t.groups = (roi - 1) / 4 + 2;
  t.chunks = vs_split(roi) ? (2 * roi + 1 + 63) / 64 : 1;
  t.off = 0;
  t.step = (size_t)t.groups * (2 * (size_t)roi + 1) * 16;
  t.disk = t.step + 4 * (size_t)t.chunks * t.groups * sizeof(VsStep4);
  t.bytes = t.disk + (2 * (size_t)roi + 1) * 4;
  return t;
}
void fill_vs_tables(int roi, void *out);  // vs_tables(roi).bytes bytes
// elevs_t: n * n scratch that receives the DEM transposed (the rays whose
// major axis is y read it, so that a wavefront's reads share rows).
siting_status launch_viewshed(const Geometry &g, const uint16_t *elevs, uint16_t *elevs_t,
                              const int32_t *obs, uint64_t *sheds,
                              const uint8_t *tables, hipStream_t s);

// The cumulative bitset is kept in tiles of kCumTile terrain rows: the
// words of one terrain word column in kCumTile consecutive rows are
// consecutive, so a shed box's words span fewer lines than in a row-major
// bitset.  Measured on SITE (tile height 1, 4, 8, 16): 4 took roi 25-150
// jobs at 8 workgroups 8-21 % and the 4096^2 job 6 % below row-major, 16
// only half as far, and intervis jobs were flat at 4 and 8 but 9 % slower
// at 16.  Word j of terrain row x (x, j >= 0):
#ifndef SITING_CUM_TILE
#define SITING_CUM_TILE 4
#endif
constexpr unsigned kCumTile = SITING_CUM_TILE;
__host__ __device__ inline size_t cum_index(int x, int j, int cw) {
  const unsigned ux = (unsigned)x;
  return ((size_t)(ux / kCumTile) * (size_t)cw + (size_t)j) * kCumTile + (size_t)(ux % kCumTile);
}
// Words the tiled bitset of an n-row terrain takes (rows rounded up to a tile).
__host__ __device__ inline size_t cum_words(int n, int cw) {
  return (size_t)(((unsigned)n + kCumTile - 1) / kCumTile) * kCumTile * (size_t)cw;
}

struct SiteBuffers {
  int32_t *area;      // [nwanted] testshedarea
  uint8_t *used;      // [nwanted] usedq == selected
  uint64_t *cum;      // [cum_words(nrows, cumnwpr)] cumulative bitset, tiled
  uint64_t *rowmax;   // [2 * nblocks] double-buffered per-block-row argmax keys
  int32_t *order;     // [nwanted] selection order
  SiteState *state;
  uint8_t *scratch;   // site_scratch_bytes(): candidate buckets
};

// Everything one job's SITE kernels need (device pointers, geometry, plan).
constexpr int kSiteMaxBatch = 32;  // picks per round (intervis 0)
struct SiteArgs {
  const int32_t *obs;
  const uint64_t *sheds;
  int32_t *area;     // [nsheds] first-pass area by candidate id
  uint8_t *used;     // [nsheds] selected flags (the output mask)
  uint64_t *cum;     // [n * cw]
  uint64_t *keys;    // [2][nunits][2] best two keys per unit
  uint32_t *kpos;    // [2][nunits][2] their positions
  int32_t *order;    // [nsheds]
  SiteState *st;
  int32_t *bcount;   // [nbuckets] counts, then scatter cursors
  int32_t *bstart;   // [nbuckets + 1]
  int32_t *chunk;    // [nchunks] scan partial sums
  int32_t *sid;      // [nsheds] candidate ids, bucket order
  uint32_t *spos;    // [nsheds] (x << 16 | y), bucket order
  int32_t *sT;       // [nsheds] area state of workgroups in global mode
  uint8_t *sused;    // [nsheds] used state of workgroups in global mode
  uint64_t *nb;      // [kSiteMaxBatch][nr][nwpr + 1] bits each pick newly covered
  unsigned int *sseen;  // [G][nsheds/32 + 1] seen bitmaps of global-mode groups
  uint64_t *phash;   // [2][kMaxGroups] per-round hash of each workgroup's picks
  int32_t *sB;       // intervis: [nsheds] area bounds of global-mode workgroups
  uint64_t *cbnd;    // intervis: [2][ncr][ncr][4] per-cell bounds of masked candidates
  int n, roi, nr, nwpr, cw, nsheds, intervis;
  int cs, ncr, Q, nunits, G, upw, kpu, nbuckets, cap;
  // dynamic LDS layout (byte offsets)
  int o_delta, o_ukey, o_upos, o_dirty, o_off, o_c, o_pos, o_t, o_used, o_seen;
  int off_lds;    // unit cell offsets kept in LDS (else read from bstart)
  int delta_lds;  // the winner's shifted shed staged in LDS (else recomputed)
  int tau_shift;  // intervis 0: update candidates whose area is below
                  // head - head / 2^tau_shift (head: the round's largest
                  // key's area) become stale bounds instead of being
                  // decremented (0: never)
  long long stop_area;  // least cumulative area c with 100.0 * c / (n * n) > 95
  unsigned int spin_limit;  // barrier polls before the loop gives up
  int hash_fault;  // test knob: round whose pick hash the last workgroup
                   // alters (0: none), to exercise the divergence check
  int iv_multi;    // intervis: rounds take several picks under cell bounds
  int walk_evals;  // intervis: exact area computations per round's walk
  int near_evals;  // intervis 0: exact area computations per round's walk
  int full_at;     // intervis 0: update candidates near this many picks of
                   // a round are counted in full instead of decremented
  int o_b, o_dcur, o_dprev, o_dbits;  // intervis LDS: bounds, dirty cells
};

// Device scratch SITE needs beyond SiteBuffers' arrays, and the size of the
// double-buffered per-row key array (entries).
size_t site_scratch_bytes(int nrows, int roi, int nsheds, int intervis);
int site_key_rows(int nrows, int roi);

// groups_cap: most workgroups the loop may use (0: half the CUs)
siting_status launch_site(const Geometry &g, const int32_t *obs,
                          const uint64_t *sheds, const SiteBuffers &b,
                          int device, hipStream_t s, int groups_cap = 0);
// Workgroups launch_site's loop will use for this job under groups_cap.
int site_loop_groups(const Geometry &g, int device_cus, int groups_cap);

// Batched SITE: site_make_args plans one job for a work-queue slot of `slot`
// workgroups (no launch); site_launch_prep enqueues its preparation kernels;
// site_launch_queue runs every job's loop in one persistent launch of
// total_groups workgroups split into slots.
//   plan (device ints, site_queue_plan_ints of them): nslots, nclasses,
//     start[nslots] (first workgroup, increasing), size[nslots],
//     cls[nslots], off[nclasses + 1], order[njobs], fb[nclasses]: class c's
//     queue is order[off[c] .. off[c+1]), indices into the SiteArgs array,
//     and every job of class c is planned for at most the size of class c's
//     slots; fb[c] is a class whose jobs c's slots take once c's queue is
//     empty (its slots no larger than c's), or -1.
//   ctl: site_queue_ctl_words(nslots, nclasses) words of queue state.
//   jobt: 2 words per job, its loop's start and end on the device clock
//     (wall_clock64, 100 MHz).
//   iv_mode: 0 or 1 when every job's intervis flag is that (the loop is
//     compiled for it), 2 when they differ.
siting_status site_make_args(const Geometry &g, const int32_t *obs,
                             const uint64_t *sheds, const SiteBuffers &b, int slot,
                             SiteArgs *out, size_t *lds);
siting_status site_launch_prep(const SiteArgs &a, hipStream_t s);
size_t site_queue_ctl_words(int nslots, int nclasses);
size_t site_queue_plan_ints(int nslots, int nclasses, int njobs);
siting_status site_launch_queue(const SiteArgs *jobs_dev, const int *plan_dev,
                                int nslots, int nclasses, int total_groups, size_t lds,
                                unsigned int *ctl, unsigned long long *jobt, int iv_mode,
                                hipStream_t s);

// host tables for VIX
void fill_recip_table(int roi, float *out);               // roi+1 entries
void fill_jump_table(int ntests, uint32_t *out);          // 4*ntests entries

}  // namespace siting
