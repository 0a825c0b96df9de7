// slot_plan.hpp -- host-side plans of the batch's SITE work-queue launch:
// how its workgroups are cut into slots and which jobs each slot runs, one
// after another.  Plain C++ with no HIP dependency (built under the
// sanitizers by tests/native); the batch (api.hip) measures plans on the
// device and keeps the fastest.
#pragma once

#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

namespace siting {

// Device form of a plan (site_loop_queue reads it; layout in common.hpp,
// site_queue_plan_ints): slot first workgroups and sizes, each slot's queue
// class, the classes' job queues concatenated, and each class's fallback:
// the class of the smallest slots, whose jobs its slots take once their own
// queue is empty (-1: none), so that a plan that sized its slots wrongly
// does not leave big slots idle while small jobs wait in line.
struct QueuePlan {
  std::vector<int> start, size, cls, off, order, fb;
  int total() const {
    int t = 0;
    for (int v : size) t += v;
    return t;
  }
};

// A plan as classes of equal slots, each class with one shared queue: the
// class's slots pop its jobs in order.
struct SlotClasses {
  std::vector<int> size, count;
  std::vector<std::vector<int>> queue;
};

inline bool same_plan(const QueuePlan &x, const QueuePlan &y) {
  return x.start == y.start && x.size == y.size && x.cls == y.cls && x.off == y.off &&
         x.order == y.order && x.fb == y.fb;
}

// The device form of c; cap[i] = the slot size of job i's class (jobs in no
// queue keep `base`).
inline void to_plan(const SlotClasses &c, int njobs, int base, QueuePlan *p,
                    std::vector<int> *cap) {
  *p = QueuePlan{};
  cap->assign((size_t)njobs, base);
  p->off = {0};
  int start = 0;
  for (size_t k = 0; k < c.size.size(); k++) {
    for (int n = 0; n < c.count[k]; n++) {
      p->start.push_back(start);
      p->size.push_back(c.size[k]);
      p->cls.push_back((int)k);
      start += c.size[k];
    }
    for (int i : c.queue[k]) {
      p->order.push_back(i);
      (*cap)[(size_t)i] = c.size[k];
    }
    p->off.push_back((int)p->order.size());
  }
  size_t small = 0;
  for (size_t k = 1; k < c.size.size(); k++)
    if (c.size[k] < c.size[small]) small = k;
  for (size_t k = 0; k < c.size.size(); k++)
    p->fb.push_back(k != small && c.size[k] >= c.size[small] ? (int)small : -1);
}

// The mixed-slot plan from one run's loop times: job i took ms[i] on g0[i]
// workgroups and is modelled to take ms[i] sqrt(g0[i] / groups(i, size)) on
// a slot of `size` (groups(i, size): the workgroups its loop takes there).
// The mixes of an optional slot of 64-128 groups for the longest job, up to
// four big slots (32 to 128 groups) and slots of `base` groups, within
// `budget` workgroups, are ranked by modelled span; the `want` best (fewer
// if fewer beat equal base slots by 10 %) go to out, best first; one queue
// per slot size, longest jobs first.  The model errs (a loop's serial pick
// walk does not shrink with more groups), so the caller measures them.
inline bool sqrt_plans(const std::vector<double> &ms, const std::vector<double> &g0,
                       const std::function<double(int, int)> &groups, int budget, int base,
                       int want, std::vector<SlotClasses> *out) {
  const int njobs = (int)ms.size();
  out->clear();
  if (njobs == 0 || (int)g0.size() != njobs || base < 1 || want < 1) return false;
  auto time_on = [&](int i, int size) {
    const double g = groups(i, size);
    return g > 0 ? ms[(size_t)i] * std::sqrt(g0[(size_t)i] / g) : 1e300;
  };
  std::vector<int> by_time((size_t)njobs);
  for (int i = 0; i < njobs; i++) by_time[(size_t)i] = i;
  std::stable_sort(by_time.begin(), by_time.end(),
                   [&](int x, int y) { return ms[(size_t)x] > ms[(size_t)y]; });
  auto sizes_of = [&](int s1, int nb, int sb) {
    std::vector<int> sz;
    if (s1 > 0) sz.push_back(s1);
    for (int k = 0; k < nb; k++) sz.push_back(sb);
    const int used = (s1 > 0 ? s1 : 0) + nb * sb;
    for (int k = 0; k < (budget - used) / base; k++) sz.push_back(base);
    return sz;
  };
  auto schedule = [&](const std::vector<int> &size, std::vector<int> *slot_of) {
    std::vector<double> fin(size.size(), 0.0);
    if (slot_of) slot_of->assign((size_t)njobs, base);
    for (int i : by_time) {
      size_t best = 0;
      double bf = 1e300;
      for (size_t k = 0; k < size.size(); k++) {
        const double f = fin[k] + time_on(i, size[k]);
        if (f < bf) {
          bf = f;
          best = k;
        }
      }
      fin[best] = bf;
      if (slot_of) (*slot_of)[(size_t)i] = size[best];
    }
    double m = 0;
    for (double f : fin) m = f > m ? f : m;
    return m;
  };
  if (budget < base) return false;
  const double m0 = schedule(sizes_of(0, 0, base), nullptr);
  std::vector<std::pair<double, std::vector<int>>> ranked;
  for (int s1 : {0, 64, 96, 128})
    for (int sb : {32, 48, 64, 96, 128})
      for (int nb = 0; nb <= 4; nb++) {
        if (sb <= base || (s1 > 0 ? s1 : 0) + nb * sb > budget - base) continue;
        if (s1 == 0 && nb == 0) continue;
        std::vector<int> sz = sizes_of(s1, nb, sb);
        const double m = schedule(sz, nullptr);
        if (m <= 0.9 * m0) ranked.push_back({m, std::move(sz)});
      }
  std::stable_sort(ranked.begin(), ranked.end(),
                   [](const auto &x, const auto &y) { return x.first < y.first; });
  for (const auto &r : ranked) {
    if ((int)out->size() == want) break;
    std::vector<int> slot_of;
    schedule(r.second, &slot_of);
    SlotClasses c;
    for (int sz : r.second) {
      size_t k = 0;
      while (k < c.size.size() && c.size[k] != sz) k++;
      if (k == c.size.size()) {
        c.size.push_back(sz);
        c.count.push_back(0);
        c.queue.push_back({});
      }
      c.count[k]++;
    }
    for (int i : by_time) {
      size_t k = 0;
      while (c.size[k] != slot_of[(size_t)i]) k++;
      c.queue[k].push_back(i);
    }
    bool dup = false;  // another size mix can schedule to the same plan
    for (const auto &o : *out)
      dup |= o.size == c.size && o.count == c.count && o.queue == c.queue;
    if (!dup) out->push_back(std::move(c));
  }
  return !out->empty();
}

// The best plan of sqrt_plans.
inline bool sqrt_plan(const std::vector<double> &ms, const std::vector<double> &g0,
                      const std::function<double(int, int)> &groups, int budget, int base,
                      SlotClasses *out) {
  std::vector<SlotClasses> v;
  if (!sqrt_plans(ms, g0, groups, budget, base, 1, &v)) return false;
  *out = v[0];
  return true;
}

// Moves job j onto a slot of its own twice its current size (at least 32
// groups, at most max_size), paid for with slots of the smallest class (at
// most half of them, while that class has jobs), and with its own class's
// slots when it was that class's only job (any surplus going back to the
// smallest class); false (c unchanged) if it cannot.
inline bool move_to_own_slot(int max_size, SlotClasses *c, int j) {
  size_t from = 0;
  for (; from < c->queue.size(); from++) {
    const auto &q = c->queue[from];
    if (std::find(q.begin(), q.end(), j) != q.end()) break;
  }
  if (from == c->queue.size()) return false;
  const int cur = c->size[from];
  const int want = std::min(std::max(2 * cur, 32), max_size);
  if (want <= cur) return false;
  size_t base = 0;
  for (size_t k = 1; k < c->size.size(); k++)
    if (c->size[k] < c->size[base]) base = k;
  // when j is its class's only job, the class's slots come back
  const bool only = c->queue[from].size() == 1;
  if (only && base == from) return false;
  const int own = only ? c->count[from] * cur : 0;
  const int need = want - own, bs = c->size[base];
  if (need % bs) return false;
  // the base class keeps at least one slot while it has jobs, and gives up
  // at most half of its slots (more serialised its small jobs: a trial of a
  // 64-job batch took 429 ms against 69)
  if (c->count[base] - need / bs < (c->queue[base].empty() ? 0 : 1)) return false;
  if (!c->queue[base].empty() && need / bs > c->count[base] / 2) return false;
  c->count[base] -= need / bs;
  if (only) {
    c->size[from] = want;
    c->count[from] = 1;
    return true;
  }
  auto &q = c->queue[from];
  q.erase(std::find(q.begin(), q.end(), j));
  c->size.push_back(want);
  c->count.push_back(1);
  c->queue.push_back({j});
  return true;
}

}  // namespace siting
