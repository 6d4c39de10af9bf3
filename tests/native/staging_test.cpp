// CPU test of the server's staging bookkeeping (tgmm_amd/csrc/tgmm_staging.h):
// page-locked blocks handed to an upload stay out of the pool until the
// upload's event completes, are recycled by capacity afterwards, and are all
// returned on shutdown. Plain host memory and fake events stand in for
// hipHostMalloc and hipEvent_t. Prints "ok" and exits 0 on success.
#include <cstdio>
#include <cstdlib>
#include <set>

#include "../../tgmm_amd/csrc/tgmm_staging.h"

using tgmm::HostBlock;

static int g_allocs = 0, g_frees = 0;
static bool alloc_fn(void** p, size_t n) {
  *p = std::malloc(n);
  g_allocs++;
  return *p != nullptr;
}
static void free_fn(void* p) {
  std::free(p);
  g_frees++;
}

struct FakeEvent {
  int id;
};
static std::set<int> g_done;  // completed events
static int query(FakeEvent e) { return g_done.count(e.id) ? 1 : 0; }

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::printf("FAILED %s at line %d\n", #c, __LINE__);         \
      return 1;                                                    \
    }                                                              \
  } while (0)

int main() {
  {
    tgmm::HostPool<bool (*)(void**, size_t), void (*)(void*)> pool(&alloc_fn, &free_fn);
    tgmm::InFlight<FakeEvent> fl;
    // two tickets stage their inputs; capacities round up to powers of two
    HostBlock a, b, c;
    CHECK(pool.get(100000, &a) && a.cap == 131072);
    CHECK(pool.get(70000, &b) && b.cap == 131072);
    CHECK(pool.live() == 2 && g_allocs == 2);
    fl.hold(1, a, FakeEvent{10});
    fl.hold(2, b, FakeEvent{20});
    // nothing completed: nothing comes back, a new request allocates anew
    CHECK(fl.reclaim(pool, query) == 0 && fl.size() == 2 && pool.pooled() == 0);
    CHECK(pool.get(1000, &c) && c.cap == 65536 && g_allocs == 3);
    CHECK(c.p != a.p && c.p != b.p);
    CHECK(fl.held_by(1) == 1 && fl.held_by(2) == 1 && fl.held_by(3) == 0);
    // ticket 2's upload completes: its block is recycled, ticket 1's is not
    g_done.insert(20);
    CHECK(fl.reclaim(pool, query) == 0 && fl.size() == 1 && fl.held_by(2) == 0);
    CHECK(pool.pooled() == 1 && pool.live() == 2);
    HostBlock d;
    CHECK(pool.get(90000, &d) && d.p == b.p && g_allocs == 3);  // reused, not a's
    CHECK(d.p != a.p);
    // a failed query is reported and keeps the block
    tgmm::InFlight<FakeEvent> bad;
    bad.hold(5, d, FakeEvent{99});
    CHECK(bad.reclaim(pool, [](FakeEvent) { return -4; }) == -4 && bad.size() == 1);
    bad.drain(pool, [](FakeEvent) {});
    // shutdown waits for every pending event and returns every block
    int waited = 0;
    fl.drain(pool, [&](FakeEvent e) {
      waited++;
      g_done.insert(e.id);
    });
    CHECK(waited == 1 && fl.size() == 0);
    pool.put(c);
    CHECK(pool.live() == 0 && pool.pooled() == 3);
    pool.put(c);  // an empty block is ignored
    CHECK(pool.pooled() == 3);
  }
  CHECK(g_frees == g_allocs);  // the pool frees what it holds
  std::printf("ok\n");
  return 0;
}
