// api.hip -- the C ABI (include/siting.h): parameter derivation, device
// resident jobs, and the reference-shaped stage functions on host arrays.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <dlfcn.h>

#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "job.hpp"
#include "slot_plan.hpp"

namespace siting {

static thread_local char g_last_error[256];
// Set only when a call on this thread rejected a size beyond a fixed working
// memory (siting_site_it_noalloc); cleared by its next call and by any HIP
// failure, so that SITING_ENOMEM's message names the limit that was hit.
static thread_local char g_limit_error[256];

void set_limit_error(const char *msg) {
  std::snprintf(g_limit_error, sizeof g_limit_error, "%s", msg ? msg : "");
}

siting_status hip_fail(hipError_t e, const char *what, const char *file, int line) {
  g_limit_error[0] = 0;
  std::snprintf(g_last_error, sizeof g_last_error, "%s failed: %s (%s:%d)", what,
                hipGetErrorString(e), file, line);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return SITING_ENOMEM;
  return SITING_EDEVICE;
}

// The SITE loop gave up (niter < 0): -1 when a grid barrier did not
// complete, -2 when its workgroups disagreed on a round's picks.
siting_status site_loop_failed(int niter) {
  std::snprintf(g_last_error, sizeof g_last_error, "%s",
                niter == -2 ? "SITE loop: workgroups disagreed on a round's picks"
                            : "SITE loop: a grid barrier timed out (workgroups not "
                              "co-resident)");
  return SITING_EDEVICE;
}

// siting.cpp:689-702, in float32 exactly as the reference evaluates it.
siting_status derive(const siting_params *p, siting_derived *d) {
  if (!p || !d) return SITING_EINVAL;
  const int n = p->nrows;
  if (n <= 0 || n > 32767 || p->roi < 1 || p->roi > kMaxRoi || p->ntests < 1 ||
      p->ntests > 4000 || p->bw < 1 || p->nwanted0 < 1 || p->oht < 0 ||
      p->tht < 0 || p->oht > 1000000 || p->tht > 1000000 ||
      (p->intervis != 0 && p->intervis != 1))
    return SITING_EINVAL;
  const float fn = (float)n;
  const int nb_req = (int)(fn / (float)p->bw + 0.5f);
  if (nb_req <= 0) return SITING_EINVAL;
  const float bs = fn / (float)nb_req;
  const int nb = (int)(fn / bs + 0.5f);
  if (nb <= 0) return SITING_EINVAL;
  const long long nb2 = (long long)nb * nb;
  const int nwpb = (int)((float)p->nwanted0 / (float)nb2 + 0.99999f);
  const long long nwanted = (long long)nwpb * nb2;
  const int lastsize = n - (int)(bs * (float)(nb - 1));
  d->blocksize = bs;
  d->nblocks = nb;
  d->nwpb = nwpb;
  d->nwanted = (int)nwanted;
  d->lastsize = lastsize;
  d->nr = 2 * p->roi + 1;
  d->nwpr = (d->nr + 63) / 64;
  d->cumnwpr = (n + 63) / 64;
  d->shed_words = (int64_t)d->nr * d->nwpr;
  if ((long long)lastsize * lastsize < nwpb) return SITING_ESMALLBLOCK;
  if (nwanted > 0x7fffffffLL || nwpb < 1 || lastsize < 1) return SITING_EINVAL;
  // every block must hold at least one cell (blocksize >= 1 guarantees it)
  if (!(bs >= 1.0f)) return SITING_EINVAL;
  return SITING_OK;
}

// ---- tuning cache ---------------------------------------------------------

namespace {

struct TuneEntry {
  std::string key;
  int roi, ntests, edge;
  siting_tuning t;
};

std::mutex g_tune_mu;
bool g_tune_loaded = false;
std::string g_tune_path;
std::vector<TuneEntry> g_tune;

// SITING_TUNE_CACHE, or siting_amd/tune/tuning.txt beside the library
std::string tune_cache_path() {
  const char *e = std::getenv("SITING_TUNE_CACHE");
  if (e && *e) return std::string(e) == "off" ? std::string() : std::string(e);
  Dl_info info;
  if (dladdr((void *)&tune_cache_path, &info) && info.dli_fname) {
    std::string lib = info.dli_fname;
    const size_t slash = lib.rfind('/');
    const std::string dir = slash == std::string::npos ? "." : lib.substr(0, slash);
    return dir + "/../tune/tuning.txt";
  }
  return std::string();
}

// (Re)reads the cache when its path changed; caller holds g_tune_mu.
void tune_load() {
  const std::string path = tune_cache_path();
  if (g_tune_loaded && path == g_tune_path) return;
  g_tune_loaded = true;
  g_tune_path = path;
  g_tune.clear();
  if (path.empty()) return;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line.resize(hash);
    std::istringstream in(line);
    TuneEntry e{};
    if (in >> e.key >> e.roi >> e.ntests >> e.edge >> e.t.vix_halo >> e.t.viewshed_threads >>
        e.t.viewshed_mode >> e.t.site_q >> e.t.site_tau_shift) {
      if (!(in >> e.t.site_walk_evals)) e.t.site_walk_evals = 0;  // older lines
      g_tune.push_back(e);
    }
  }
}

int roi_class(int roi) {
  return roi <= 37 ? 25 : roi <= 75 ? 50 : roi <= 125 ? 100 : roi <= 175 ? 150
                                                                         : roi <= 300 ? 200 : 0;
}
int ntests_class(int ntests) { return ntests <= 48 ? 32 : 64; }
int edge_class(int n) { return n <= 1536 ? 1024 : n <= 3072 ? 2048 : 4096; }

std::mutex g_key_mu;
std::map<int, std::string> g_device_keys;

}  // namespace

siting_status tuning_lookup(const char *device_key, const siting_params &p,
                            siting_tuning *out) {
  *out = siting_tuning{};
  if (!device_key) return SITING_EINVAL;
  const int rc = roi_class(p.roi), tc = ntests_class(p.ntests), ec = edge_class(p.nrows);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_load();
  for (const TuneEntry &e : g_tune)
    if (e.key == device_key && e.roi == rc && e.ntests == tc && e.edge == ec) {
      *out = e.t;
      return SITING_OK;
    }
  return SITING_ENOTFOUND;
}

// "<arch>:<CUs>" of a device (cached per device).
siting_status device_key(int device, std::string *out) {
  std::lock_guard<std::mutex> lk(g_key_mu);
  auto it = g_device_keys.find(device);
  if (it != g_device_keys.end()) {
    *out = it->second;
    return SITING_OK;
  }
  hipDeviceProp_t prop;
  SITING_HIP(hipGetDeviceProperties(&prop, device));
  std::string arch = prop.gcnArchName;
  const size_t colon = arch.find(':');
  if (colon != std::string::npos) arch.resize(colon);
  *out = arch + ":" + std::to_string(prop.multiProcessorCount);
  g_device_keys[device] = *out;
  return SITING_OK;
}

}  // namespace siting

using namespace siting;

// ---------------------------------------------------------------------------
// job
// ---------------------------------------------------------------------------


namespace {

struct Layout {
  size_t elevs, elevs_t, vix, obs, sheds, area, used, cum, rowmax, order, state, recip,
      jump, vstab, scratch, total;
};

size_t round_up(size_t v) { return (v + 255) / 256 * 256; }

Layout layout_for(const siting_params &p, const siting_derived &d) {
  Layout L{};
  const size_t n2 = (size_t)p.nrows * p.nrows;
  const size_t nw = (size_t)d.nwanted;
  size_t off = 0;
  auto put = [&](size_t &field, size_t bytes) {
    field = off;
    off += round_up(bytes ? bytes : 1);
  };
  put(L.elevs, n2 * 2);
  put(L.elevs_t, n2 * 2);
  put(L.vix, n2);
  put(L.obs, nw * 4);
  put(L.sheds, nw * (size_t)d.shed_words * 8);
  put(L.area, nw * 4);
  put(L.used, nw);
  put(L.cum, cum_words(p.nrows, d.cumnwpr) * 8);
  put(L.rowmax, 2 * (size_t)site_key_rows(p.nrows, p.roi) * 8);
  put(L.order, nw * 4);
  put(L.state, sizeof(SiteState));
  put(L.recip, ((size_t)p.roi + 1) * 4);
  put(L.jump, 4 * (size_t)p.ntests * 4);
  put(L.vstab, vs_tables(p.roi).bytes);
  put(L.scratch, site_scratch_bytes(p.nrows, p.roi, d.nwanted, p.intervis));
  L.total = off;
  return L;
}

siting_status record(hipEvent_t e, hipStream_t s) {
  SITING_HIP(hipEventRecord(e, s));
  return SITING_OK;
}

}  // namespace

extern "C" {

const char *siting_status_str(siting_status s) {
  switch (s) {
    case SITING_OK: return "ok";
    case SITING_EINVAL: return "invalid argument";
    case SITING_ERANGE: return "parameter out of range";
    case SITING_ESMALLBLOCK: return "The last block is too small for nwantedperblock.";
    case SITING_ENOMEM:
      return g_limit_error[0] ? g_limit_error : "Memory exhausted. Program terminates.";
    case SITING_EDEVICE: return g_last_error[0] ? g_last_error : "HIP device error";
    case SITING_EIO: return "Input failed";
    case SITING_EBUSY: return "busy";
    case SITING_ENOTFOUND: return "not found";
  }
  return "unknown status";
}

siting_status siting_derive(const siting_params *p, siting_derived *d) {
  return derive(p, d);
}

siting_status siting_check_cli(const siting_params *p, int *which, const char **msg) {
  static const char *msgs[] = {
      "", "Unreasonable value for nrows", "Unreasonable value for roi.",
      "Unreasonable value for tht.", "Unreasonable value for ntests.",
      "Unreasonable value for blocksize0.", "Unreasonable value for nwanted0.",
      "Unreasonable value for intervis."};
  int w = 0;  // siting.cpp:680-686, in order
  if (!p) return SITING_EINVAL;
  if (p->nrows <= 0 || p->nrows > 20000) w = 1;
  else if (p->roi < 1 || p->roi > 10000) w = 2;
  else if (p->tht < 0 || p->tht > 1000000) w = 3;
  else if (p->ntests < 1 || p->ntests > 1000) w = 4;
  else if (p->bw < 10 || p->bw > 2000) w = 5;
  else if (p->nwanted0 < 100 || p->nwanted0 > 2000000) w = 6;
  else if (p->intervis != 0 && p->intervis != 1) w = 7;
  if (which) *which = w;
  if (msg) *msg = msgs[w];
  return w ? SITING_ERANGE : SITING_OK;
}

int siting_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

}  // extern "C"

namespace siting {

size_t job_bytes(const siting_params &p, const siting_derived &d) {
  return layout_for(p, d).total;
}

siting_status job_create_in(int device, const siting_params *p, void *mem,
                            hipStream_t stream, siting_job **out) {
  if (!out) return SITING_EINVAL;
  *out = nullptr;
  siting_derived d;
  siting_status st = derive(p, &d);
  if (st != SITING_OK) return st;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    std::snprintf(g_last_error, sizeof g_last_error, "no HIP device %d", device);
    return SITING_EDEVICE;
  }
  SITING_HIP(hipSetDevice(device));
  siting_job *j = new siting_job();
  j->device = device;
  j->g.p = *p;
  j->g.d = d;
  {  // the tuning cache's choices for this job, if any
    std::string key;
    if (device_key(device, &key) == SITING_OK) tuning_lookup(key.c_str(), *p, &j->g.t);
  }
  Layout L = layout_for(*p, d);
  j->bytes = L.total;
  auto fail = [&](siting_status s) {
    siting_job_destroy(j);
    return s;
  };
  hipError_t e = hipSuccess;
  if (mem) {
    j->base = mem;
    j->owns_mem = false;
  } else {
    e = hipMalloc(&j->base, L.total);
    if (e != hipSuccess) {
      j->base = nullptr;
      return fail(hip_fail(e, "hipMalloc(job)", __FILE__, __LINE__));
    }
  }
  if (stream) {
    j->stream = stream;
    j->owns_stream = false;
  } else {
    e = hipStreamCreateWithFlags(&j->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(hip_fail(e, "hipStreamCreate", __FILE__, __LINE__));
  }
  for (int i = 0; i < 4; i++) {
    if ((e = hipEventCreate(&j->ev_start[i])) != hipSuccess ||
        (e = hipEventCreate(&j->ev_end[i])) != hipSuccess)
      return fail(hip_fail(e, "hipEventCreate", __FILE__, __LINE__));
  }
  uint8_t *b = (uint8_t *)j->base;
  j->elevs = (uint16_t *)(b + L.elevs);
  j->elevs_t = (uint16_t *)(b + L.elevs_t);
  j->vix = b + L.vix;
  j->obs = (int32_t *)(b + L.obs);
  j->sheds = (uint64_t *)(b + L.sheds);
  j->site.area = (int32_t *)(b + L.area);
  j->site.used = b + L.used;
  j->site.cum = (uint64_t *)(b + L.cum);
  j->site.rowmax = (uint64_t *)(b + L.rowmax);
  j->site.order = (int32_t *)(b + L.order);
  j->site.state = (SiteState *)(b + L.state);
  j->site.scratch = b + L.scratch;
  j->recip = (float *)(b + L.recip);
  j->jump = (uint32_t *)(b + L.jump);
  j->vstab = b + L.vstab;
  std::vector<float> rec((size_t)p->roi + 1);
  std::vector<uint32_t> jmp(4 * (size_t)p->ntests);
  std::vector<uint8_t> vst(vs_tables(p->roi).bytes);
  fill_recip_table(p->roi, rec.data());
  fill_jump_table(p->ntests, jmp.data());
  fill_vs_tables(p->roi, vst.data());
  e = hipMemcpy(j->recip, rec.data(), rec.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(j->jump, jmp.data(), jmp.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(j->vstab, vst.data(), vst.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(hip_fail(e, "hipMemcpy(tables)", __FILE__, __LINE__));
  *out = j;
  return SITING_OK;
}

}  // namespace siting

extern "C" {

siting_status siting_job_bytes(const siting_params *p, size_t *bytes) {
  siting_derived d;
  siting_status st = derive(p, &d);
  if (st != SITING_OK) return st;
  if (bytes) *bytes = layout_for(*p, d).total;
  return SITING_OK;
}

siting_status siting_job_create(int device, const siting_params *p, siting_job **out) {
  return job_create_in(device, p, nullptr, nullptr, out);
}

void siting_job_destroy(siting_job *j) {
  if (!j) return;
  (void)hipSetDevice(j->device);
  if (j->stream) (void)hipStreamSynchronize(j->stream);
  for (int i = 0; i < 4; i++) {
    if (j->ev_start[i]) (void)hipEventDestroy(j->ev_start[i]);
    if (j->ev_end[i]) (void)hipEventDestroy(j->ev_end[i]);
  }
  if (j->stream && j->owns_stream) (void)hipStreamDestroy(j->stream);
  if (j->base && j->owns_mem) (void)hipFree(j->base);
  delete j;
}

siting_status siting_job_upload_dem(siting_job *j, const uint16_t *elevs) {
  if (!j || !elevs) return SITING_EINVAL;
  SITING_HIP(hipSetDevice(j->device));
  const size_t bytes = (size_t)j->g.p.nrows * j->g.p.nrows * 2;
  SITING_HIP(hipMemcpyAsync(j->elevs, elevs, bytes, hipMemcpyHostToDevice, j->stream));
  SITING_HIP(hipStreamSynchronize(j->stream));
  return SITING_OK;
}

#define JOB_STAGE(idx, body)                                         \
  do {                                                               \
    if (!j) return SITING_EINVAL;                                    \
    SITING_HIP(hipSetDevice(j->device));                             \
    siting_status st_ = record(j->ev_start[idx], j->stream);         \
    if (st_ != SITING_OK) return st_;                                \
    st_ = (body);                                                    \
    if (st_ != SITING_OK) return st_;                                \
    st_ = record(j->ev_end[idx], j->stream);                         \
    if (st_ != SITING_OK) return st_;                                \
    j->ran[idx] = true;                                              \
    if (idx == 3) j->site_state = true;                              \
    return SITING_OK;                                                \
  } while (0)

siting_status siting_job_vix(siting_job *j) {
  JOB_STAGE(0, launch_vix(j->g, j->elevs, j->vix, VixTables{j->recip, j->jump}, j->stream));
}
siting_status siting_job_findmax(siting_job *j) {
  JOB_STAGE(1, launch_findmax(j->g, j->vix, j->obs, j->stream));
}
siting_status siting_job_viewshed(siting_job *j) {
  JOB_STAGE(2, launch_viewshed(j->g, j->elevs, j->elevs_t, j->obs, j->sheds, j->vstab,
                               j->stream));
}
siting_status siting_job_site(siting_job *j) {
  JOB_STAGE(3, launch_site(j->g, j->obs, j->sheds, j->site, j->device, j->stream,
                           j->site_groups));
}

siting_status siting_job_run(siting_job *j) {
  siting_status s;
  if ((s = siting_job_vix(j)) != SITING_OK) return s;
  if ((s = siting_job_findmax(j)) != SITING_OK) return s;
  if ((s = siting_job_viewshed(j)) != SITING_OK) return s;
  return siting_job_site(j);
}

siting_status siting_job_sync(siting_job *j) {
  if (!j) return SITING_EINVAL;
  SITING_HIP(hipSetDevice(j->device));
  SITING_HIP(hipStreamSynchronize(j->stream));
  return SITING_OK;
}

siting_status siting_job_device_ptr(siting_job *j, siting_buffer which, void **dptr,
                                    size_t *bytes) {
  if (!j) return SITING_EINVAL;
  const size_t n2 = (size_t)j->g.p.nrows * j->g.p.nrows;
  const size_t nw = (size_t)j->g.d.nwanted;
  void *p = nullptr;
  size_t b = 0;
  switch (which) {
    case SITING_BUF_ELEVS: p = j->elevs; b = n2 * 2; break;
    case SITING_BUF_VIX: p = j->vix; b = n2; break;
    case SITING_BUF_OBS: p = j->obs; b = nw * 4; break;
    case SITING_BUF_SHEDS: p = j->sheds; b = nw * (size_t)j->g.d.shed_words * 8; break;
    case SITING_BUF_SELECTED: p = j->site.used; b = nw; break;
    case SITING_BUF_ORDER: p = j->site.order; b = nw * 4; break;
    default: return SITING_EINVAL;
  }
  if (dptr) *dptr = p;
  if (bytes) *bytes = b;
  return SITING_OK;
}

siting_status siting_job_download(siting_job *j, siting_buffer which, void *host,
                                  size_t bytes) {
  void *p = nullptr;
  size_t b = 0;
  siting_status st = siting_job_device_ptr(j, which, &p, &b);
  if (st != SITING_OK) return st;
  if (!host || bytes > b) return SITING_EINVAL;
  SITING_HIP(hipSetDevice(j->device));
  SITING_HIP(hipMemcpyAsync(host, p, bytes, hipMemcpyDeviceToHost, j->stream));
  SITING_HIP(hipStreamSynchronize(j->stream));
  return SITING_OK;
}

siting_status siting_job_stats(siting_job *j, siting_stats *st) {
  if (!j || !st) return SITING_EINVAL;
  SITING_HIP(hipSetDevice(j->device));
  SITING_HIP(hipStreamSynchronize(j->stream));
  std::memset(st, 0, sizeof *st);
  float *ms[4] = {&st->ms_vix, &st->ms_findmax, &st->ms_viewshed, &st->ms_site};
  for (int i = 0; i < 4; i++)
    if (j->ran[i]) SITING_HIP(hipEventElapsedTime(ms[i], j->ev_start[i], j->ev_end[i]));
  if (j->site_state) {
    SiteState s;
    SITING_HIP(hipMemcpy(&s, j->site.state, sizeof s, hipMemcpyDeviceToHost));
    st->nused = s.nused;
    st->niter = s.niter;
    st->cumarea = (int64_t)s.cumarea;
    st->sum_u = (int64_t)s.sum_u;
    st->rounds = s.rounds;
    const double n2 = (double)j->g.p.nrows * (double)j->g.p.nrows;
    st->coverage = 100.0 * (double)s.cumarea / n2;
    if (s.niter < 0) return site_loop_failed(s.niter);
  }
  return SITING_OK;
}

siting_status siting_job_derived(siting_job *j, siting_derived *d) {
  if (!j || !d) return SITING_EINVAL;
  *d = j->g.d;
  return SITING_OK;
}

siting_status siting_job_set_tuning(siting_job *j, const siting_tuning *t) {
  if (!j) return SITING_EINVAL;
  j->g.t = t ? *t : siting_tuning{};
  return SITING_OK;
}

siting_status siting_job_get_tuning(siting_job *j, siting_tuning *t) {
  if (!j || !t) return SITING_EINVAL;
  *t = j->g.t;
  return SITING_OK;
}

siting_status siting_device_key(int device, char *buf, size_t size) {
  if (!buf || size == 0) return SITING_EINVAL;
  std::string key;
  siting_status st = device_key(device, &key);
  if (st != SITING_OK) return st;
  if (key.size() + 1 > size) return SITING_EINVAL;
  std::memcpy(buf, key.c_str(), key.size() + 1);
  return SITING_OK;
}

siting_status siting_tuning_lookup(const char *device_key, const siting_params *p,
                                   siting_tuning *out) {
  if (!device_key || !p || !out) return SITING_EINVAL;
  return tuning_lookup(device_key, *p, out);
}

const char *siting_tuning_cache_path(void) {
  static thread_local std::string path;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_load();
  path = g_tune_path;
  return path.c_str();
}

// ---------------------------------------------------------------------------
// stage functions on host arrays (the reference's signatures)
// ---------------------------------------------------------------------------

namespace {

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

struct StageStream {
  hipStream_t s = nullptr;
  ~StageStream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

siting_status stage_device(hipStream_t &s, StageStream &holder) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    std::snprintf(g_last_error, sizeof g_last_error, "no HIP device");
    return SITING_EDEVICE;
  }
  SITING_HIP(hipSetDevice(0));
  SITING_HIP(hipStreamCreateWithFlags(&holder.s, hipStreamNonBlocking));
  s = holder.s;
  return SITING_OK;
}

siting_status dalloc(DevBuf &b, size_t bytes) {
  SITING_HIP(hipMalloc(&b.p, bytes ? bytes : 1));
  return SITING_OK;
}

}  // namespace

#define CHECK(x)                             \
  do {                                       \
    siting_status s_ = (x);                  \
    if (s_ != SITING_OK) return s_;          \
  } while (0)

siting_status siting_calc_vix(int nrows, const uint16_t *elevs, int roi, int oht,
                              int tht, int ntests, uint8_t *vix) {
  if (!elevs || !vix || nrows <= 0 || nrows > 32767 || roi < 1 || roi > kMaxRoi ||
      ntests < 1 || ntests > 4000)
    return SITING_EINVAL;
  hipStream_t s;
  StageStream hs;
  CHECK(stage_device(s, hs));
  Geometry g{};
  g.p.nrows = nrows;
  g.p.roi = roi;
  g.p.oht = oht;
  g.p.tht = tht;
  g.p.ntests = ntests;
  const size_t n2 = (size_t)nrows * nrows;
  DevBuf de, dv, dr, dj;
  CHECK(dalloc(de, n2 * 2));
  CHECK(dalloc(dv, n2));
  CHECK(dalloc(dr, ((size_t)roi + 1) * 4));
  CHECK(dalloc(dj, 16 * (size_t)ntests));
  std::vector<float> rec((size_t)roi + 1);
  std::vector<uint32_t> jmp(4 * (size_t)ntests);
  fill_recip_table(roi, rec.data());
  fill_jump_table(ntests, jmp.data());
  SITING_HIP(hipMemcpyAsync(de.p, elevs, n2 * 2, hipMemcpyHostToDevice, s));
  SITING_HIP(hipMemcpyAsync(dr.p, rec.data(), rec.size() * 4, hipMemcpyHostToDevice, s));
  SITING_HIP(hipMemcpyAsync(dj.p, jmp.data(), jmp.size() * 4, hipMemcpyHostToDevice, s));
  CHECK(launch_vix(g, (const uint16_t *)de.p, (uint8_t *)dv.p,
                   VixTables{(const float *)dr.p, (const uint32_t *)dj.p}, s));
  SITING_HIP(hipMemcpyAsync(vix, dv.p, n2, hipMemcpyDeviceToHost, s));
  SITING_HIP(hipStreamSynchronize(s));
  return SITING_OK;
}

siting_status siting_find_max(int nrows, uint8_t *vix, float blocksize, int nwanted,
                              int nblocks, int nwpb, int32_t *obs) {
  (void)nwanted;  // unused by the reference too (siting.cpp:235)
  if (!vix || !obs || nrows <= 0 || nrows > 32767 || nblocks < 1 || nwpb < 0 ||
      !(blocksize >= 1.0f))
    return SITING_EINVAL;
  hipStream_t s;
  StageStream hs;
  CHECK(stage_device(s, hs));
  Geometry g{};
  g.p.nrows = nrows;
  g.d.blocksize = blocksize;
  g.d.nblocks = nblocks;
  g.d.nwpb = nwpb;
  const size_t n2 = (size_t)nrows * nrows;
  const size_t nobs = (size_t)nblocks * nblocks * nwpb;
  DevBuf dv, dobs;
  CHECK(dalloc(dv, n2));
  CHECK(dalloc(dobs, nobs * 4));
  SITING_HIP(hipMemcpyAsync(dv.p, vix, n2, hipMemcpyHostToDevice, s));
  CHECK(launch_findmax(g, (uint8_t *)dv.p, (int32_t *)dobs.p, s));
  SITING_HIP(hipMemcpyAsync(obs, dobs.p, nobs * 4, hipMemcpyDeviceToHost, s));
  SITING_HIP(hipMemcpyAsync(vix, dv.p, n2, hipMemcpyDeviceToHost, s));
  SITING_HIP(hipStreamSynchronize(s));
  return SITING_OK;
}

siting_status siting_calc_sheds(int nrows, const uint16_t *elevs, int roi, int oht,
                                int tht, int nsheds, const int32_t *obs,
                                uint64_t *sheds) {
  if (!elevs || !obs || !sheds || nrows <= 0 || nrows > 32767 || roi < 1 ||
      roi > kMaxRoi || nsheds < 0)
    return SITING_EINVAL;
  const size_t n2 = (size_t)nrows * nrows;
  for (int k = 0; k < nsheds; k++)
    if (obs[k] < 0 || (size_t)obs[k] >= n2) return SITING_EINVAL;
  if (nsheds == 0) return SITING_OK;
  hipStream_t s;
  StageStream hs;
  CHECK(stage_device(s, hs));
  Geometry g{};
  g.p.nrows = nrows;
  g.p.roi = roi;
  g.p.oht = oht;
  g.p.tht = tht;
  g.d.nwanted = nsheds;
  g.d.nr = 2 * roi + 1;
  g.d.nwpr = (g.d.nr + 63) / 64;
  const size_t sbytes = (size_t)nsheds * g.d.nr * g.d.nwpr * 8;
  DevBuf de, det, dobs, dsh, drd;
  CHECK(dalloc(de, n2 * 2));
  CHECK(dalloc(det, n2 * 2));
  CHECK(dalloc(dobs, (size_t)nsheds * 4));
  CHECK(dalloc(dsh, sbytes));
  std::vector<uint8_t> vst(vs_tables(roi).bytes);
  CHECK(dalloc(drd, vst.size()));
  fill_vs_tables(roi, vst.data());
  SITING_HIP(hipMemcpyAsync(de.p, elevs, n2 * 2, hipMemcpyHostToDevice, s));
  SITING_HIP(hipMemcpyAsync(dobs.p, obs, (size_t)nsheds * 4, hipMemcpyHostToDevice, s));
  SITING_HIP(hipMemcpyAsync(drd.p, vst.data(), vst.size(), hipMemcpyHostToDevice, s));
  CHECK(launch_viewshed(g, (const uint16_t *)de.p, (uint16_t *)det.p, (const int32_t *)dobs.p,
                        (uint64_t *)dsh.p, (const uint8_t *)drd.p, s));
  SITING_HIP(hipMemcpyAsync(sheds, dsh.p, sbytes, hipMemcpyDeviceToHost, s));
  SITING_HIP(hipStreamSynchronize(s));
  return SITING_OK;
}

siting_status siting_site_it(int nrows, int roi, int intervis, int nsheds,
                             const int32_t *obs, const uint64_t *sheds,
                             uint8_t *selected, siting_stats *stats) {
  if (!obs || !sheds || !selected || nrows <= 0 || nrows > 32767 || roi < 1 ||
      roi > kMaxRoi || nsheds < 1 || (intervis != 0 && intervis != 1))
    return SITING_EINVAL;
  const size_t n2 = (size_t)nrows * nrows;
  for (int k = 0; k < nsheds; k++)
    if (obs[k] < 0 || (size_t)obs[k] >= n2) return SITING_EINVAL;
  hipStream_t s;
  StageStream hs;
  CHECK(stage_device(s, hs));
  Geometry g{};
  g.p.nrows = nrows;
  g.p.roi = roi;
  g.p.intervis = intervis;
  g.d.nwanted = nsheds;
  g.d.nr = 2 * roi + 1;
  g.d.nwpr = (g.d.nr + 63) / 64;
  g.d.cumnwpr = (nrows + 63) / 64;
  g.d.shed_words = (int64_t)g.d.nr * g.d.nwpr;
  const size_t nw = (size_t)nsheds;
  const size_t sbytes = nw * (size_t)g.d.shed_words * 8;
  DevBuf dobs, dsh, darea, dused, dcum, drm, dord, dst, dscr;
  CHECK(dalloc(dobs, nw * 4));
  CHECK(dalloc(dsh, sbytes));
  CHECK(dalloc(darea, nw * 4));
  CHECK(dalloc(dused, nw));
  CHECK(dalloc(dcum, cum_words(nrows, g.d.cumnwpr) * 8));
  CHECK(dalloc(drm, 2 * (size_t)site_key_rows(nrows, roi) * 8));
  CHECK(dalloc(dord, nw * 4));
  CHECK(dalloc(dst, sizeof(SiteState)));
  CHECK(dalloc(dscr, site_scratch_bytes(nrows, roi, nsheds, intervis)));
  SITING_HIP(hipMemcpyAsync(dobs.p, obs, nw * 4, hipMemcpyHostToDevice, s));
  SITING_HIP(hipMemcpyAsync(dsh.p, sheds, sbytes, hipMemcpyHostToDevice, s));
  SiteBuffers b{(int32_t *)darea.p, (uint8_t *)dused.p, (uint64_t *)dcum.p,
                (uint64_t *)drm.p, (int32_t *)dord.p, (SiteState *)dst.p,
                (uint8_t *)dscr.p};
  CHECK(launch_site(g, (const int32_t *)dobs.p, (const uint64_t *)dsh.p, b, 0, s));
  SITING_HIP(hipMemcpyAsync(selected, dused.p, nw, hipMemcpyDeviceToHost, s));
  SiteState st;
  SITING_HIP(hipMemcpyAsync(&st, dst.p, sizeof st, hipMemcpyDeviceToHost, s));
  SITING_HIP(hipStreamSynchronize(s));
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->nused = st.nused;
    stats->niter = st.niter;
    stats->cumarea = (int64_t)st.cumarea;
    stats->sum_u = (int64_t)st.sum_u;
    stats->rounds = st.rounds;
    stats->coverage = 100.0 * (double)st.cumarea / (double)n2;
  }
  if (st.niter < 0) return site_loop_failed(st.niter);
  return SITING_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
namespace siting {

// A plan the batch has run, with each job's loop time and the SITE launch
// time it measured.
struct TriedPlan {
  SlotClasses classes;
  std::vector<double> dur;
  double span_ms;
};

}  // namespace siting

struct siting_batch {
  int device = 0, cus = 0;
  std::vector<siting_job *> jobs;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
  hipEvent_t ts = nullptr;             // before the SITE work-queue launch
  std::vector<int> order;              // prep order: descending SITE work
  std::vector<int> cap;                // per job: workgroups of its slot class
  SiteArgs *args_dev = nullptr;        // [njobs], by job index
  int *plan_dev = nullptr;             // slot and queue plan
  unsigned int *ctl = nullptr;         // work-queue state
  unsigned long long *jobt = nullptr;  // [2 * njobs] loop start/end per job
  size_t lds = 0;
  int slot = 16, nslots = 1, nclasses = 1, total_groups = 16, max_slots = 1;
  int budget = 16;  // workgroups of the launch (every plan stays within it)
  int iv_mode = 2;  // the jobs' common intervis flag (0, 1), or 2 when they differ
  bool use_graph = false, replan = true;
  int runs = 0, runs_seen = 0;
  // slot tuning: plans tried so far and the current one
  int tune_left = 0;
  siting::QueuePlan cur;
  siting::SlotClasses cur_classes;
  std::vector<siting::TriedPlan> tried;
  std::vector<siting::SlotClasses> pending;  // model plans still to measure
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

// Uploads a plan and re-plans every job for its class's slot size.
siting_status batch_set_plan(siting_batch *b, const QueuePlan &p,
                             const std::vector<int> &cap) {
  const int njobs = (int)b->jobs.size();
  std::vector<SiteArgs> args((size_t)njobs);
  size_t lds_max = 0;
  for (int i = 0; i < njobs; i++) {
    siting_job *j = b->jobs[(size_t)i];
    size_t lds = 0;
    siting_status st = site_make_args(j->g, j->obs, j->sheds, j->site, cap[(size_t)i],
                                      &args[(size_t)i], &lds);
    if (st != SITING_OK) return st;
    lds_max = lds > lds_max ? lds : lds_max;
  }
  std::vector<int> packed;
  packed.push_back((int)p.start.size());
  packed.push_back((int)p.off.size() - 1);
  for (const auto *v : {&p.start, &p.size, &p.cls, &p.off, &p.order, &p.fb})
    packed.insert(packed.end(), v->begin(), v->end());
  SITING_HIP(hipMemcpy(b->args_dev, args.data(), sizeof(SiteArgs) * (size_t)njobs,
                       hipMemcpyHostToDevice));
  SITING_HIP(hipMemcpy(b->plan_dev, packed.data(), sizeof(int) * packed.size(),
                       hipMemcpyHostToDevice));
  b->cap = cap;
  b->cur = p;
  b->lds = lds_max;
  b->nslots = (int)p.start.size();
  b->nclasses = (int)p.off.size() - 1;
  b->total_groups = p.total();
  return SITING_OK;
}

// Every job's VIX, FINDMAX, VIEWSHED and SITE preparation, spread over the
// stream pool and joined back into streams[0] (the part a HIP graph holds).
siting_status batch_enqueue_prep(siting_batch *b) {
  hipStream_t s0 = b->streams[0];
  SITING_HIP(hipEventRecord(b->fork, s0));
  for (size_t k = 1; k < b->streams.size(); k++)
    SITING_HIP(hipStreamWaitEvent(b->streams[k], b->fork, 0));
  for (size_t q = 0; q < b->order.size(); q++) {
    const int i = b->order[q];
    siting_job *j = b->jobs[(size_t)i];
    hipStream_t s = b->streams[q % b->streams.size()];
    siting_status st;
    if ((st = launch_vix(j->g, j->elevs, j->vix, VixTables{j->recip, j->jump}, s)) != SITING_OK)
      return st;
    if ((st = launch_findmax(j->g, j->vix, j->obs, s)) != SITING_OK) return st;
    if ((st = launch_viewshed(j->g, j->elevs, j->elevs_t, j->obs, j->sheds, j->vstab, s)) !=
        SITING_OK)
      return st;
    SiteArgs a;
    if ((st = site_make_args(j->g, j->obs, j->sheds, j->site, b->cap[(size_t)i], &a,
                             nullptr)) != SITING_OK)
      return st;
    if ((st = site_launch_prep(a, s)) != SITING_OK) return st;
  }
  for (size_t k = 1; k < b->streams.size(); k++) {
    SITING_HIP(hipEventRecord(b->joins[k], b->streams[k]));
    SITING_HIP(hipStreamWaitEvent(s0, b->joins[k], 0));
  }
  return SITING_OK;
}

// Every job's greedy loop in one work-queue launch on streams[0], between
// the events that time it.
siting_status batch_enqueue_site(siting_batch *b) {
  hipStream_t s0 = b->streams[0];
  SITING_HIP(hipEventRecord(b->ts, s0));
  siting_status st = site_launch_queue(b->args_dev, b->plan_dev, b->nslots, b->nclasses,
                                       b->total_groups, b->lds, b->ctl, b->jobt, b->iv_mode, s0);
  if (st != SITING_OK) return st;
  SITING_HIP(hipEventRecord(b->t1, s0));
  return SITING_OK;
}

// Captures the prep sequence into a HIP graph (executes nothing).
siting_status batch_capture(siting_batch *b) {
  hipStream_t s0 = b->streams[0];
  SITING_HIP(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
  siting_status st = batch_enqueue_prep(b);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s0, &g);
  if (st != SITING_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
  b->graph = g;
  SITING_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return SITING_OK;
}

void batch_drop_graph(siting_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Slot tuning.  The work-queue launch's slots (how many workgroups each
// job's loop gets, and which jobs take turns on a slot) are tuned on the
// batch's first runs by measurement, since a loop's time on G workgroups
// depends on its job (rounds against per-round work) and on what else runs
// beside it:
//  - run 1 uses equal slots of the base size;
//  - runs 2-4 the three best mixed-slot plans of the square-root model
//    (slot_plan.hpp), measured since the model ranks them only roughly;
//  - each later run, up to kTuneRuns, moves the job whose loop ran longest
//    in the fastest plan measured so far onto a slot of its own twice as
//    large (at least 32 groups), paid for with base slots; a move that does
//    not shorten the launch is undone, and the next tries the job that ran
//    longest but one, and so on.  (Moving the job that finished last was
//    worse: that is usually a small job that waited in a queue, and its slot
//    took base slots from the queue's other small jobs, 77 -> 437 ms);
//  - then the fastest plan measured is kept for good.
// (Alone, a 3365^2 roi-150 job of the uneven batch takes 84 ms on 8 groups,
// 26 on 32 and 18 on 64; a 2593^2 roi-150 one with 2761 rounds 230, 90 and
// 77: loops differ widely in how far more groups help.)
constexpr int kTuneRuns = 7;
constexpr int kModelPlans = 3;

// After a synced run: record the plan's measured launch time and job end
// times, and choose the next plan.
siting_status batch_tune(siting_batch *b) {
  const int njobs = (int)b->jobs.size();
  std::vector<unsigned long long> t((size_t)2 * njobs);
  SITING_HIP(hipMemcpy(t.data(), b->jobt, sizeof(unsigned long long) * t.size(),
                       hipMemcpyDeviceToHost));
  std::vector<double> ms((size_t)njobs), g((size_t)njobs);
  for (int i = 0; i < njobs; i++) {
    const unsigned long long a = t[(size_t)2 * i], e = t[(size_t)2 * i + 1];
    ms[(size_t)i] = e > a ? (double)(e - a) * 1e-5 : 0.0;  // 100 MHz clock
    g[(size_t)i] = site_loop_groups(b->jobs[(size_t)i]->g, b->cus, b->cap[(size_t)i]);
  }
  float span = 0;
  SITING_HIP(hipEventElapsedTime(&span, b->ts, b->t1));
  b->tried.push_back({b->cur_classes, ms, (double)span});
  b->tune_left--;
  size_t best = 0;
  for (size_t k = 1; k < b->tried.size(); k++)
    if (b->tried[k].span_ms < b->tried[best].span_ms) best = k;
  SlotClasses next;
  bool have = false;
  if (b->tune_left > 0) {
    if (b->tried.size() == 1) {
      std::vector<SlotClasses> plans;
      sqrt_plans(ms, g,
                 [&](int i, int size) {
                   return (double)site_loop_groups(b->jobs[(size_t)i]->g, b->cus, size);
                 },
                 b->budget, b->slot, kModelPlans, &plans);
      b->pending.assign(plans.begin(), plans.end());
    }
    while (!have && !b->pending.empty()) {
      next = b->pending.front();
      b->pending.erase(b->pending.begin());
      QueuePlan pp;
      std::vector<int> pc;
      to_plan(next, njobs, b->slot, &pp, &pc);
      have = (int)pp.start.size() <= b->max_slots;
    }
    // local moves from the best plan: its jobs by loop time, longest first,
    // skipping the moves already tried from it
    const auto &bp = b->tried[best];
    std::vector<int> late((size_t)njobs);
    for (int i = 0; i < njobs; i++) late[(size_t)i] = i;
    std::stable_sort(late.begin(), late.end(),
                     [&](int x, int y) { return bp.dur[(size_t)x] > bp.dur[(size_t)y]; });
    for (size_t r = 0; !have && r < late.size() && r < 8; r++) {
      SlotClasses c = bp.classes;
      if (!move_to_own_slot(b->cus / 2, &c, late[r])) continue;
      bool seen = false;
      for (const auto &tp : b->tried)
        if (tp.classes.size == c.size && tp.classes.count == c.count &&
            tp.classes.queue == c.queue)
          seen = true;
      if (seen) continue;
      next = c;
      have = true;
    }
  }
  if (!have) {
    b->tune_left = 0;
    next = b->tried[best].classes;
  }
  QueuePlan p;
  std::vector<int> cap;
  to_plan(next, njobs, b->slot, &p, &cap);
  if ((int)p.start.size() > b->max_slots) {
    b->tune_left = 0;
    next = b->tried[best].classes;
    to_plan(next, njobs, b->slot, &p, &cap);
  }
  if (same_plan(p, b->cur)) return SITING_OK;
  batch_drop_graph(b);
  b->cur_classes = next;
  return batch_set_plan(b, p, cap);
}

}  // namespace

extern "C" {

void siting_batch_destroy(siting_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  for (hipStream_t s : b->streams)
    if (s) (void)hipStreamSynchronize(s);
  batch_drop_graph(b);
  for (hipEvent_t e : b->joins)
    if (e) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->t0) (void)hipEventDestroy(b->t0);
  if (b->t1) (void)hipEventDestroy(b->t1);
  if (b->ts) (void)hipEventDestroy(b->ts);
  for (hipStream_t s : b->streams)
    if (s) (void)hipStreamDestroy(s);
  if (b->args_dev) (void)hipFree(b->args_dev);
  if (b->plan_dev) (void)hipFree(b->plan_dev);
  if (b->ctl) (void)hipFree(b->ctl);
  if (b->jobt) (void)hipFree(b->jobt);
  for (siting_job *j : b->jobs) siting_job_destroy(j);
  delete b;
}

siting_status siting_batch_create(int device, int njobs, const siting_params *params,
                                  const siting_batch_opts *opts, siting_batch **out) {
  if (!out || !params || njobs < 1) return SITING_EINVAL;
  *out = nullptr;
  siting_batch *b = new siting_batch();
  b->device = device;
  auto fail = [&](siting_status s) {
    siting_batch_destroy(b);
    return s;
  };
  for (int i = 0; i < njobs; i++) {
    siting_job *j = nullptr;
    siting_status st = siting_job_create(device, &params[i], &j);
    if (st != SITING_OK) return fail(st);
    j->site_state = true;  // stats read the SITE state
    // the tuning cache's SITE choices are measured for a job's loop alone on
    // up to half the CUs; in work-queue slots of 4-128 groups beside 60
    // other loops the built-in column rule measured faster (uneven 64-job
    // batch: SITE launch 96 ms with the rules, 116 with the cached choices),
    // and a stale threshold of head - head/16 (the uneven batch's SITE
    // launch with shifts 2..6: 81.1, 65.5, 61.7, 65.1, 77.8 ms)
    j->g.t.site_q = 0;
    j->g.t.site_tau_shift = 4;
    b->jobs.push_back(j);
  }
  SITING_HIP(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device));
  const int cus = b->cus;
  b->iv_mode = params[0].intervis ? 1 : 0;
  for (int i = 1; i < njobs; i++)
    if ((params[i].intervis ? 1 : 0) != b->iv_mode) b->iv_mode = 2;
  const int nstreams = opts && opts->streams > 0 ? opts->streams : 4;
  // 12 workgroups per base slot measured best on the uneven 64-job batch
  // with the current loop (SITE launch ms with tuned slots: 6: 72.4, 8: 60.2,
  // 10: 61.0, 12: 57.6, 14: 60.7, 16: 60.6; with the loops of an earlier
  // design, whose rounds cost several times more, 8 had been best)
  b->slot = opts && opts->slot > 0 ? opts->slot : 12;
  if (b->slot > cus) b->slot = cus;
  b->nslots = opts && opts->nslots > 0 ? opts->nslots : cus / b->slot;
  if (b->nslots * b->slot > cus) b->nslots = cus / b->slot;
  if (b->nslots < 1) b->nslots = 1;
  if (b->nslots > njobs) b->nslots = njobs;
  b->use_graph = opts && opts->use_graph;
  b->replan = !(opts && opts->fixed_slots);
  b->tune_left = kTuneRuns;
  hipError_t e;
  for (int k = 0; k < nstreams; k++) {
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
    if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess)
      return fail(hip_fail(e, "hipStreamCreate", __FILE__, __LINE__));
    b->streams.push_back(s);
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
      return fail(hip_fail(e, "hipEventCreate", __FILE__, __LINE__));
    b->joins.push_back(ev);
  }
  if ((e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventCreate(&b->t0)) != hipSuccess || (e = hipEventCreate(&b->t1)) != hipSuccess ||
      (e = hipEventCreate(&b->ts)) != hipSuccess)
    return fail(hip_fail(e, "hipEventCreate", __FILE__, __LINE__));
  // prep order and first queue order: largest SITE work first (candidates x
  // shed words)
  b->order.resize((size_t)njobs);
  for (int i = 0; i < njobs; i++) b->order[(size_t)i] = i;
  std::vector<double> work((size_t)njobs);
  for (int i = 0; i < njobs; i++)
    work[(size_t)i] = (double)b->jobs[(size_t)i]->g.d.nwanted *
                      (double)b->jobs[(size_t)i]->g.d.shed_words;
  std::stable_sort(b->order.begin(), b->order.end(),
                   [&](int x, int y) { return work[(size_t)x] > work[(size_t)y]; });
  // buffers sized for any plan the re-planning may choose: slots of at
  // least 4 workgroups, each with its own queue
  const int max_slots = std::max(b->nslots, b->nslots * b->slot / 4);
  b->max_slots = max_slots;
  if ((e = hipMalloc(&b->args_dev, sizeof(SiteArgs) * (size_t)njobs)) != hipSuccess ||
      (e = hipMalloc(&b->plan_dev,
                     sizeof(int) * site_queue_plan_ints(max_slots, max_slots, njobs))) !=
          hipSuccess ||
      (e = hipMalloc(&b->ctl, 4 * site_queue_ctl_words(max_slots, max_slots))) != hipSuccess ||
      (e = hipMalloc(&b->jobt, sizeof(unsigned long long) * 2 * (size_t)njobs)) != hipSuccess)
    return fail(hip_fail(e, "hipMalloc(batch)", __FILE__, __LINE__));
  b->budget = b->nslots * b->slot;
  // first plan: equal slots of the base size, one queue in prep order
  b->cur_classes = SlotClasses{{b->slot}, {b->nslots}, {b->order}};
  QueuePlan p0;
  std::vector<int> cap0;
  to_plan(b->cur_classes, njobs, b->slot, &p0, &cap0);
  siting_status st = batch_set_plan(b, p0, cap0);
  if (st != SITING_OK) return fail(st);
  *out = b;
  return SITING_OK;
}

siting_status siting_batch_upload_dem(siting_batch *b, int job, const uint16_t *elevs) {
  if (!b || job < 0 || job >= (int)b->jobs.size()) return SITING_EINVAL;
  return siting_job_upload_dem(b->jobs[(size_t)job], elevs);
}

siting_status siting_batch_run(siting_batch *b) {
  if (!b) return SITING_EINVAL;
  SITING_HIP(hipSetDevice(b->device));
  b->runs++;
  hipStream_t s0 = b->streams[0];
  if (b->use_graph && !b->exec) {
    siting_status st = batch_capture(b);
    if (st != SITING_OK) return st;
  }
  SITING_HIP(hipEventRecord(b->t0, s0));
  if (b->use_graph) {
    SITING_HIP(hipGraphLaunch(b->exec, s0));
  } else {
    siting_status st = batch_enqueue_prep(b);
    if (st != SITING_OK) return st;
  }
  return batch_enqueue_site(b);
}

siting_status siting_batch_sync(siting_batch *b) {
  if (!b) return SITING_EINVAL;
  SITING_HIP(hipSetDevice(b->device));
  for (hipStream_t s : b->streams) SITING_HIP(hipStreamSynchronize(s));
  if (b->runs > b->runs_seen) {
    b->runs_seen = b->runs;
    if (b->replan && b->tune_left > 0) {
      siting_status st = batch_tune(b);
      if (st != SITING_OK) return st;
      if (b->use_graph && !b->exec) return batch_capture(b);
    }
  }
  return SITING_OK;
}

siting_status siting_batch_elapsed_ms(siting_batch *b, float *ms) {
  if (!b || !ms) return SITING_EINVAL;
  SITING_HIP(hipEventSynchronize(b->t1));
  SITING_HIP(hipEventElapsedTime(ms, b->t0, b->t1));
  return SITING_OK;
}

siting_status siting_batch_site_ms(siting_batch *b, float *ms) {
  if (!b || !ms) return SITING_EINVAL;
  SITING_HIP(hipEventSynchronize(b->t1));
  SITING_HIP(hipEventElapsedTime(ms, b->ts, b->t1));
  return SITING_OK;
}

siting_status siting_batch_site_times(siting_batch *b, double *start_end, int njobs) {
  if (!b || !start_end || njobs != (int)b->jobs.size()) return SITING_EINVAL;
  SITING_HIP(hipSetDevice(b->device));
  SITING_HIP(hipEventSynchronize(b->t1));
  std::vector<unsigned long long> t((size_t)2 * njobs);
  SITING_HIP(hipMemcpy(t.data(), b->jobt, sizeof(unsigned long long) * t.size(),
                       hipMemcpyDeviceToHost));
  unsigned long long t0 = ~0ull;
  for (int i = 0; i < njobs; i++) t0 = t[(size_t)2 * i] < t0 ? t[(size_t)2 * i] : t0;
  for (size_t k = 0; k < t.size(); k++)
    start_end[k] = t[k] >= t0 ? (double)(t[k] - t0) * 1e-5 : 0.0;  // 100 MHz clock
  return SITING_OK;
}

siting_status siting_batch_job(siting_batch *b, int job, siting_job **out) {
  if (!b || !out || job < 0 || job >= (int)b->jobs.size()) return SITING_EINVAL;
  *out = b->jobs[(size_t)job];
  return SITING_OK;
}

}  // extern "C"
