// server.hip -- request loop that serves siting jobs on one GPU.
//
// Requests (parameters + a host DEM) are queued by siting_server_submit and
// admitted by a scheduler thread as device memory and SITE workgroups free
// up, so jobs that arrive while others run join them on the device
// (continuous batching): each admitted job gets its buffers from one device
// arena (first-fit, no hipMalloc per request), a stream from a pool and a
// reservation of the workgroups its persistent SITE loop needs; its four
// stages are enqueued back to back and stay in HBM; a completion event
// retires it, its selected mask and observers are copied out and its memory
// returns to the arena.
//
// Admission is first-come first-served with backfill: a later, smaller
// request may start while the head of the queue waits for memory, until the
// head has waited `backfill_ms`; then the queue drains in order so the head
// cannot starve.  The SITE reservation keeps the workgroups of all running
// persistent loops within the CU count, so every loop is fully resident.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "arena.hpp"
#include "job.hpp"

using namespace siting;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t id = 0;
  siting_params p{};
  std::vector<uint16_t> dem;
  size_t bytes = 0;
  int groups = 0;
  siting_job_state state = SITING_JOB_QUEUED;
  siting_status status = SITING_OK;
  siting_job *job = nullptr;
  size_t off = 0;
  int stream = -1;
  hipEvent_t done = nullptr;
  std::vector<uint8_t> selected;
  std::vector<int32_t> obs;
  siting_stats stats{};
  double t_submit = 0, t_admit = 0, t_done = 0;
};

}  // namespace

struct siting_server {
  int device = 0;
  int cus = 0;
  siting_server_opts opts{};
  void *block = nullptr;
  Arena arena;
  std::vector<hipStream_t> streams;
  std::vector<bool> stream_busy;
  std::mutex mu;
  std::condition_variable cv_done, cv_work;
  std::map<uint64_t, Request *> reqs;
  std::deque<Request *> pending;
  std::vector<Request *> active;
  uint64_t next_id = 1;
  int groups_in_use = 0;
  int waiters = 0;  // threads blocked in siting_server_wait
  bool stop = false;
  std::thread thread;
  Clock::time_point t0;
  double now() const {
    return std::chrono::duration<double>(Clock::now() - t0).count();
  }
};

namespace {

// Retire a finished request: copy its results out, return its resources.
void finish(siting_server *S, Request *r, hipError_t q) {
  siting_job *j = r->job;
  if (q == hipSuccess) {
    r->selected.resize((size_t)j->g.d.nwanted);
    r->obs.resize((size_t)j->g.d.nwanted);
    siting_status st = siting_job_download(j, SITING_BUF_SELECTED, r->selected.data(),
                                           r->selected.size());
    if (st == SITING_OK)
      st = siting_job_download(j, SITING_BUF_OBS, r->obs.data(), r->obs.size() * 4);
    if (st == SITING_OK) st = siting_job_stats(j, &r->stats);
    r->status = st;
    r->state = st == SITING_OK ? SITING_JOB_DONE : SITING_JOB_FAILED;
  } else {
    r->status = hip_fail(q, "job completion", __FILE__, __LINE__);
    r->state = SITING_JOB_FAILED;
  }
  siting_job_destroy(j);
  r->job = nullptr;
  S->arena.release(r->off, r->bytes);
  S->groups_in_use -= r->groups;
  S->stream_busy[(size_t)r->stream] = false;
  (void)hipEventDestroy(r->done);
  r->done = nullptr;
  r->dem.clear();
  r->dem.shrink_to_fit();
  r->t_done = S->now();
}

// Try to start r; false if memory, workgroups or streams are short.
bool admit(siting_server *S, Request *r) {
  if (S->groups_in_use + r->groups > S->cus) return false;
  int sidx = -1;
  for (size_t k = 0; k < S->streams.size(); k++)
    if (!S->stream_busy[k]) {
      sidx = (int)k;
      break;
    }
  if (sidx < 0) return false;
  size_t off = 0;
  if (!S->arena.alloc(r->bytes, &off)) return false;
  siting_job *j = nullptr;
  siting_status st = job_create_in(S->device, &r->p, (uint8_t *)S->block + off,
                                   S->streams[(size_t)sidx], &j);
  hipError_t e = hipSuccess;
  if (st == SITING_OK) {
    j->site_groups = S->opts.site_groups;
    e = hipEventCreateWithFlags(&r->done, hipEventDisableTiming);
    if (e == hipSuccess)
      e = hipMemcpyAsync(j->elevs, r->dem.data(), r->dem.size() * 2,
                         hipMemcpyHostToDevice, j->stream);
    if (e == hipSuccess) st = siting_job_run(j);
    if (st == SITING_OK && e == hipSuccess) e = hipEventRecord(r->done, j->stream);
  }
  if (st != SITING_OK || e != hipSuccess) {
    if (j) siting_job_destroy(j);
    if (r->done) (void)hipEventDestroy(r->done);
    r->done = nullptr;
    S->arena.release(off, r->bytes);
    r->status = st != SITING_OK ? st : hip_fail(e, "admit", __FILE__, __LINE__);
    r->state = SITING_JOB_FAILED;
    r->t_done = S->now();
    S->cv_done.notify_all();
    return true;  // consumed (failed)
  }
  r->job = j;
  r->off = off;
  r->stream = sidx;
  S->stream_busy[(size_t)sidx] = true;
  S->groups_in_use += r->groups;
  r->state = SITING_JOB_RUNNING;
  r->t_admit = S->now();
  S->active.push_back(r);
  return true;
}

void scheduler(siting_server *S) {
  (void)hipSetDevice(S->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    // retire finished jobs
    bool any_done = false;
    for (size_t i = 0; i < S->active.size();) {
      Request *r = S->active[i];
      const hipError_t q = hipEventQuery(r->done);
      if (q == hipErrorNotReady) {
        i++;
        continue;
      }
      finish(S, r, q);
      S->active.erase(S->active.begin() + (long)i);
      any_done = true;
    }
    if (any_done) S->cv_done.notify_all();
    // admit: in order, backfilling past a blocked head until it has waited
    // backfill_ms
    if (!S->stop) {
      bool head_blocked = false;
      double head_since = 0;
      for (auto it = S->pending.begin(); it != S->pending.end();) {
        if ((int)S->active.size() >= S->opts.max_active) break;
        Request *r = *it;
        if (head_blocked && (S->now() - head_since) * 1000.0 > S->opts.backfill_ms) break;
        if (admit(S, r)) {
          it = S->pending.erase(it);
          continue;
        }
        if (!head_blocked) {
          head_blocked = true;
          head_since = r->t_submit;
        }
        ++it;
      }
    }
    if (S->stop && S->active.empty()) break;
    if (S->active.empty() && S->pending.empty()) {
      S->cv_work.wait(lk);
    } else {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(50));
      lk.lock();
    }
  }
}

}  // namespace

extern "C" {

siting_status siting_server_create(int device, const siting_server_opts *opts,
                                   siting_server **out) {
  if (!out) return SITING_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return SITING_EDEVICE;
  SITING_HIP(hipSetDevice(device));
  siting_server *S = new siting_server();
  S->device = device;
  if (opts) S->opts = *opts;
  if (S->opts.max_active <= 0) S->opts.max_active = 8;
  if (S->opts.site_groups <= 0) S->opts.site_groups = 32;
  if (S->opts.backfill_ms <= 0) S->opts.backfill_ms = 200;
  hipError_t e = hipDeviceGetAttribute(&S->cus, hipDeviceAttributeMultiprocessorCount, device);
  size_t freeb = 0, totalb = 0;
  if (e == hipSuccess) e = hipMemGetInfo(&freeb, &totalb);
  if (e != hipSuccess) {
    delete S;
    return hip_fail(e, "server device query", __FILE__, __LINE__);
  }
  size_t budget = S->opts.hbm_bytes ? S->opts.hbm_bytes : freeb / 10 * 8;
  if (budget > freeb) budget = freeb / 10 * 9;
  S->opts.hbm_bytes = budget;
  e = hipMalloc(&S->block, budget);
  if (e != hipSuccess) {
    delete S;
    return hip_fail(e, "hipMalloc(server arena)", __FILE__, __LINE__);
  }
  S->arena.reset(budget);
  for (int k = 0; k < S->opts.max_active; k++) {
    hipStream_t s = nullptr;
    if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) {
      siting_server_destroy(S);
      return hip_fail(e, "hipStreamCreate", __FILE__, __LINE__);
    }
    S->streams.push_back(s);
    S->stream_busy.push_back(false);
  }
  S->t0 = Clock::now();
  S->thread = std::thread(scheduler, S);
  *out = S;
  return SITING_OK;
}

void siting_server_destroy(siting_server *S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    for (Request *r : S->pending) {
      r->state = SITING_JOB_FAILED;
      r->status = SITING_EBUSY;
      r->t_done = S->now();
    }
    S->pending.clear();
    S->cv_done.notify_all();  // waiters on dropped requests return now
  }
  S->cv_work.notify_all();
  // the scheduler retires the running jobs (waking their waiters) and exits
  if (S->thread.joinable()) S->thread.join();
  {
    // every request is now DONE or FAILED: let the blocked waiters leave
    // before their requests and the server go away
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.notify_all();
    S->cv_done.wait(lk, [&] { return S->waiters == 0; });
  }
  (void)hipSetDevice(S->device);
  for (hipStream_t s : S->streams) (void)hipStreamDestroy(s);
  if (S->block) (void)hipFree(S->block);
  for (auto &kv : S->reqs) delete kv.second;
  delete S;
}

siting_status siting_server_submit(siting_server *S, const siting_params *p,
                                   const uint16_t *elevs, uint64_t *id) {
  if (!S || !p || !elevs || !id) return SITING_EINVAL;
  siting_derived d;
  siting_status st = derive(p, &d);
  if (st != SITING_OK) return st;
  Request *r = new Request();
  r->p = *p;
  r->bytes = job_bytes(*p, d);
  if (Arena::granule(r->bytes) > S->opts.hbm_bytes) {
    delete r;
    return SITING_ENOMEM;  // can never be admitted
  }
  Geometry g{};
  g.p = *p;
  g.d = d;
  r->groups = site_loop_groups(g, S->cus, S->opts.site_groups);
  const size_t n2 = (size_t)p->nrows * p->nrows;
  r->dem.assign(elevs, elevs + n2);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) {
      delete r;
      return SITING_EBUSY;
    }
    r->id = S->next_id++;
    r->t_submit = S->now();
    S->reqs[r->id] = r;
    S->pending.push_back(r);
    *id = r->id;
  }
  S->cv_work.notify_all();
  return SITING_OK;
}

static void fill_info(const Request *r, siting_job_info *info) {
  info->state = r->state;
  info->status = r->status;
  info->stats = r->stats;
  info->t_submit = r->t_submit;
  info->t_admit = r->t_admit;
  info->t_done = r->t_done;
}

siting_status siting_server_info(siting_server *S, uint64_t id, siting_job_info *info) {
  if (!S || !info) return SITING_EINVAL;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return SITING_ENOTFOUND;
  fill_info(it->second, info);
  return SITING_OK;
}

siting_status siting_server_wait(siting_server *S, uint64_t id, double timeout_s,
                                 siting_job_info *info) {
  if (!S) return SITING_EINVAL;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return SITING_ENOTFOUND;
  Request *r = it->second;
  auto pred = [&] { return r->state == SITING_JOB_DONE || r->state == SITING_JOB_FAILED; };
  S->waiters++;
  bool done = true;
  if (timeout_s < 0)
    S->cv_done.wait(lk, pred);
  else
    done = S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
  if (done && info) fill_info(r, info);
  // a destroy in progress waits for the last waiter before freeing anything
  if (--S->waiters == 0 && S->stop) S->cv_done.notify_all();
  return done ? SITING_OK : SITING_EBUSY;
}

siting_status siting_server_result(siting_server *S, uint64_t id, uint8_t *selected,
                                   size_t nsel, int32_t *obs, size_t nobs) {
  if (!S) return SITING_EINVAL;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return SITING_ENOTFOUND;
  Request *r = it->second;
  if (r->state != SITING_JOB_DONE) return r->state == SITING_JOB_FAILED ? r->status
                                                                         : SITING_EBUSY;
  if ((selected && nsel < r->selected.size()) || (obs && nobs < r->obs.size()))
    return SITING_EINVAL;
  if (selected) std::memcpy(selected, r->selected.data(), r->selected.size());
  if (obs) std::memcpy(obs, r->obs.data(), r->obs.size() * 4);
  return SITING_OK;
}

siting_status siting_server_cancel(siting_server *S, uint64_t id) {
  if (!S) return SITING_EINVAL;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return SITING_ENOTFOUND;
  Request *r = it->second;
  for (auto p = S->pending.begin(); p != S->pending.end(); ++p)
    if (*p == r) {
      S->pending.erase(p);
      r->state = SITING_JOB_FAILED;
      r->status = SITING_EBUSY;
      r->t_done = S->now();
      S->cv_done.notify_all();
      return SITING_OK;
    }
  return SITING_EBUSY;  // already running or finished
}

siting_status siting_server_release(siting_server *S, uint64_t id) {
  if (!S) return SITING_EINVAL;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->reqs.find(id);
  if (it == S->reqs.end()) return SITING_ENOTFOUND;
  Request *r = it->second;
  if (r->state == SITING_JOB_QUEUED || r->state == SITING_JOB_RUNNING) return SITING_EBUSY;
  delete r;
  S->reqs.erase(it);
  return SITING_OK;
}

siting_status siting_server_load(siting_server *S, siting_server_load_info *out) {
  if (!S || !out) return SITING_EINVAL;
  std::lock_guard<std::mutex> lk(S->mu);
  out->queued = (int)S->pending.size();
  out->running = (int)S->active.size();
  out->groups_in_use = S->groups_in_use;
  out->hbm_bytes = S->opts.hbm_bytes;
  return SITING_OK;
}

}  // extern "C"
