// host_checks.cpp -- the library's host-side code and the CPU oracle under
// AddressSanitizer and UndefinedBehaviorSanitizer (see Makefile):
//   the server's device-memory arena (arena.hpp) against a byte-map model,
//   the batch's slot plans (slot_plan.hpp) on random job sets,
//   the DEM generator, reader and writer (dem.cpp),
//   the oracle (siting_oracle.c) on small terrains, sequential and OpenMP,
//   with edge observers, roi wider than the terrain and duplicate observers.
// Prints one line per part and exits non-zero on the first failed check
// (the sanitizers abort on their own findings).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <set>
#include <string>
#include <vector>

#include "arena.hpp"
#include "siting.h"
#include "slot_plan.hpp"
extern "C" {
#include "siting_oracle.h"
}

#define CHECK(c)                                                         \
  do {                                                                   \
    if (!(c)) {                                                          \
      std::fprintf(stderr, "check failed: %s (%s:%d)\n", #c, __FILE__, __LINE__); \
      std::exit(1);                                                      \
    }                                                                    \
  } while (0)

static void check_arena() {
  std::mt19937_64 rng(7);
  const size_t G = 65536, total = 512 * G;
  siting::Arena a;
  a.reset(total);
  std::vector<uint8_t> used(total / G, 0);  // granule map
  struct Block {
    size_t off, bytes;
  };
  std::vector<Block> live;
  for (int step = 0; step < 20000; step++) {
    if (live.empty() || rng() % 3 != 0) {
      const size_t bytes = 1 + rng() % (24 * G);
      size_t off = 0;
      const size_t gr = siting::Arena::granule(bytes) / G;
      // the model: is there a free run of gr granules?
      size_t run = 0;
      bool fits = false;
      for (size_t k = 0; k < used.size() && !fits; k++) {
        run = used[k] ? 0 : run + 1;
        fits = run >= gr;
      }
      const bool got = a.alloc(bytes, &off);
      CHECK(got == fits);  // first fit finds a run whenever one exists
      if (!got) continue;
      CHECK(off % G == 0 && off + gr * G <= total);
      for (size_t k = off / G; k < off / G + gr; k++) {
        CHECK(!used[k]);
        used[k] = 1;
      }
      live.push_back({off, bytes});
    } else {
      const size_t i = rng() % live.size();
      const Block b = live[i];
      live[i] = live.back();
      live.pop_back();
      a.release(b.off, b.bytes);
      for (size_t k = b.off / G; k < b.off / G + siting::Arena::granule(b.bytes) / G; k++)
        used[k] = 0;
    }
    size_t free_g = 0, ranges = 0;
    for (size_t k = 0; k < used.size(); k++) {
      free_g += !used[k];
      ranges += !used[k] && (k == 0 || used[k - 1]);
    }
    CHECK(a.free_bytes() == free_g * G);
    CHECK(a.free_ranges() == ranges);  // freed neighbours always coalesce
  }
  for (const Block &b : live) a.release(b.off, b.bytes);
  CHECK(a.free_bytes() == total && a.free_ranges() == 1);
  std::printf("arena ok\n");
}

static void check_plan_shape(const siting::SlotClasses &c, int njobs, int budget, int base) {
  siting::QueuePlan p;
  std::vector<int> cap;
  siting::to_plan(c, njobs, base, &p, &cap);
  CHECK(p.total() <= budget);
  std::vector<int> seen((size_t)njobs, 0);
  for (int i : p.order) {
    CHECK(i >= 0 && i < njobs);
    seen[(size_t)i]++;
  }
  for (int v : seen) CHECK(v == 1);
  CHECK(p.off.size() == c.size.size() + 1 && p.off.back() == (int)p.order.size());
  int start = 0;
  for (size_t s = 0; s < p.start.size(); s++) {
    CHECK(p.start[s] == start);
    CHECK(p.cls[s] >= 0 && p.cls[s] < (int)c.size.size());
    CHECK(p.size[s] == c.size[(size_t)p.cls[s]]);
    start += p.size[s];
  }
  for (size_t k = 0; k < c.size.size(); k++) {
    CHECK(c.count[k] >= 0);
    if (!c.queue[k].empty()) CHECK(c.count[k] >= 1);
    for (int i : c.queue[k]) CHECK(cap[(size_t)i] == c.size[k]);
  }
  // fallback queues: another class with slots no larger, and one class
  // (the smallest slots) with none
  CHECK(p.fb.size() == c.size.size());
  int none = 0;
  for (size_t k = 0; k < p.fb.size(); k++) {
    const int f = p.fb[k];
    if (f < 0) {
      none++;
      continue;
    }
    CHECK(f < (int)c.size.size() && f != (int)k && c.size[(size_t)f] <= c.size[k]);
    CHECK(p.fb[(size_t)f] < 0);
  }
  CHECK(none >= 1);
}

static void check_slot_plans() {
  std::mt19937_64 rng(11);
  int made = 0, moved = 0;
  for (int trial = 0; trial < 300; trial++) {
    const int njobs = 1 + (int)(rng() % 80);
    const int cus = 256, base = 8, budget = (cus / base) * base;
    std::vector<double> ms((size_t)njobs), g0((size_t)njobs, base);
    std::vector<int> gmax((size_t)njobs);
    for (int i = 0; i < njobs; i++) {
      ms[(size_t)i] = 0.5 + (double)(rng() % 100000) / 1000.0;
      gmax[(size_t)i] = 4 + (int)(rng() % 125);
    }
    auto groups = [&](int i, int size) { return (double)std::min(size, gmax[(size_t)i]); };
    // the model's best few plans: each a valid plan, all distinct, the first
    // the one sqrt_plan gives
    std::vector<siting::SlotClasses> top;
    const bool any = siting::sqrt_plans(ms, g0, groups, budget, base, 3, &top);
    CHECK(any == !top.empty() && top.size() <= 3);
    for (size_t x = 0; x < top.size(); x++) {
      check_plan_shape(top[x], njobs, budget, base);
      for (size_t y = 0; y < x; y++)
        CHECK(!(top[x].size == top[y].size && top[x].count == top[y].count &&
                top[x].queue == top[y].queue));
    }
    siting::SlotClasses c;
    if (!siting::sqrt_plan(ms, g0, groups, budget, base, &c)) {
      CHECK(!any);
      c = siting::SlotClasses{{base}, {std::min(budget / base, njobs)}, {{}}};
      for (int i = 0; i < njobs; i++) c.queue[0].push_back(i);
    } else {
      CHECK(any && c.size == top[0].size && c.count == top[0].count && c.queue == top[0].queue);
      made++;
    }
    check_plan_shape(c, njobs, budget, base);
    for (int m = 0; m < 6; m++) {
      siting::SlotClasses d = c;
      const int j = (int)(rng() % njobs);
      if (siting::move_to_own_slot(cus / 2, &d, j)) {
        moved++;
        check_plan_shape(d, njobs, budget, base);
        c = d;
      } else {
        // a refused move leaves the plan as it was
        CHECK(d.size == c.size && d.count == c.count && d.queue == c.queue);
      }
    }
  }
  CHECK(made > 0 && moved > 0);
  std::printf("slot plans ok (%d mixed plans, %d moves)\n", made, moved);
}

static void check_dem() {
  const char *path = "host_checks_dem.bin";
  for (int n : {1, 2, 3, 17, 64, 65, 257, 300}) {
    std::vector<uint16_t> a((size_t)n * n), b((size_t)n * n), c((size_t)n * n);
    CHECK(siting_dem_generate(n, 0.7, 0, 20000, 8, 5000.0, 80.0, 42, a.data()) == SITING_OK);
    CHECK(siting_dem_generate(n, 0.7, 0, 20000, 8, 5000.0, 80.0, 42, c.data()) == SITING_OK);
    CHECK(a == c);  // seeded: the same bytes every time
    CHECK(siting_dem_write(path, n, a.data()) == SITING_OK);
    CHECK(siting_dem_read(path, n, b.data()) == SITING_OK);
    CHECK(a == b);
    if (n > 1) {
      std::vector<uint16_t> big((size_t)(n + 1) * (n + 1));
      CHECK(siting_dem_read(path, n + 1, big.data()) == SITING_EIO);  // short file
    }
  }
  std::remove(path);
  std::vector<uint16_t> x(4);
  CHECK(siting_dem_generate(0, 0.7, 0, 100, 0, 0, 1, 1, x.data()) != SITING_OK);
  CHECK(siting_dem_read("/nonexistent/dem.bin", 2, x.data()) == SITING_EIO);
  std::printf("dem ok\n");
}

static void check_oracle() {
  struct Case {
    int n, roi, ht, nt, bw, nw0, iv;
  };
  const Case cases[] = {{48, 6, 100, 16, 8, 64, 0},  {60, 9, 50, 40, 10, 100, 1},
                        {33, 3, 100, 8, 3, 1000, 1}, {64, 80, 100, 16, 10, 100, 1},
                        {97, 5, 0, 1, 7, 200, 0},    {120, 30, 100, 70, 12, 300, 0}};
  for (const Case &c : cases) {
    so_derived d;
    CHECK(so_derive(c.n, c.bw, c.nw0, &d) == 0);
    const int nr = 2 * c.roi + 1, nwpr = (nr + 63) / 64;
    std::vector<uint16_t> dem((size_t)c.n * c.n);
    CHECK(siting_dem_generate(c.n, 0.7, 0, 20000, 8, 5000.0, 80.0, 1234 + c.n, dem.data()) ==
          SITING_OK);
    std::vector<uint8_t> vix[2], sel[2];
    std::vector<int32_t> obs[2];
    std::vector<uint64_t> sheds[2];
    so_site_stats st[2];
    for (int v = 0; v < 2; v++) {
      vix[v].resize((size_t)c.n * c.n);
      obs[v].resize((size_t)d.nwanted);
      sheds[v].resize((size_t)d.nwanted * nr * nwpr);
      sel[v].resize((size_t)d.nwanted);
      CHECK(so_run(c.n, dem.data(), c.roi, c.ht, c.nt, c.bw, c.nw0, c.iv, vix[v].data(),
                   obs[v].data(), sheds[v].data(), sel[v].data(), &st[v], nullptr,
                   v == 0 ? 1 : 4) == 0);
    }
    CHECK(vix[0] == vix[1] && obs[0] == obs[1] && sheds[0] == sheds[1] && sel[0] == sel[1]);
    CHECK(st[0].nused == st[1].nused && st[0].cumarea == st[1].cumarea &&
          st[0].sum_u == st[1].sum_u && st[0].niter == st[1].niter);
    // SITE on every cell as an observer, duplicates included
    std::vector<int32_t> all;
    for (int k = 0; k < c.n * c.n; k += 3) all.push_back(k);
    all.push_back(0);
    all.push_back(c.n * c.n - 1);
    std::vector<uint64_t> sh(all.size() * (size_t)nr * nwpr);
    int64_t steps = 0;
    so_calc_sheds(c.n, dem.data(), c.roi, c.ht, c.ht, (int)all.size(), all.data(), sh.data(),
                  2, &steps);
    CHECK(steps > 0);
    std::vector<uint8_t> s2(all.size());
    std::vector<int32_t> order(all.size());
    so_site_stats st2;
    so_site(c.n, c.roi, c.iv, (int)all.size(), all.data(), sh.data(), s2.data(), order.data(),
            &st2, 1);
    CHECK(st2.nused >= 1 && st2.cumarea <= (int64_t)c.n * c.n);
  }
  std::printf("oracle ok\n");
}

int main() {
  check_arena();
  check_slot_plans();
  check_dem();
  check_oracle();
  std::printf("host checks ok\n");
  return 0;
}
