// tgmm_staging.h -- page-locked staging buffers for asynchronous host-to-
// device copies, and the bookkeeping that keeps each buffer alive until the
// copies reading it have completed.
//
// Every host source of an asynchronous copy issued by the server is one of
// these buffers (or memory the caller lends until the frame is fetched), so
// no copy reads pageable memory that the runtime would have to lock, and no
// buffer is freed or reused while a copy may still read it: a buffer handed
// to hold() goes back to the pool only once reclaim() finds the event that
// follows its copies complete.
//
// Header-only and independent of HIP: the allocator and the event type are
// parameters, so the bookkeeping is tested on the CPU with plain memory and
// fake events (tests/native/staging_test.cpp).
#pragma once
#include <cstddef>
#include <map>
#include <utility>
#include <vector>

namespace tgmm {

struct HostBlock {
  void* p = nullptr;
  size_t cap = 0;
};

// Buffers of power-of-two capacity (at least 64 KiB), recycled by capacity.
// Alloc: bool(void** p, size_t bytes); Free: void(void* p).
template <class Alloc, class Free>
class HostPool {
 public:
  HostPool(Alloc a, Free f) : alloc_(a), free_fn_(f) {}
  ~HostPool() { clear(); }
  HostPool(const HostPool&) = delete;
  HostPool& operator=(const HostPool&) = delete;
  bool get(size_t bytes, HostBlock* out) {
    size_t c = size_t(1) << 16;
    while (c < bytes) c <<= 1;
    auto it = free_.find(c);
    if (it != free_.end()) {
      *out = it->second;
      free_.erase(it);
      live_++;
      return true;
    }
    HostBlock b;
    b.cap = c;
    if (!alloc_(&b.p, c)) return false;
    *out = b;
    live_++;
    return true;
  }
  void put(HostBlock& b) {
    if (b.p) {
      free_.emplace(b.cap, b);
      live_--;
    }
    b = HostBlock{};
  }
  void clear() {
    for (auto& kv : free_) free_fn_(kv.second.p);
    free_.clear();
  }
  size_t live() const { return live_; }      // handed out, not yet returned
  size_t pooled() const { return free_.size(); }

 private:
  Alloc alloc_;
  Free free_fn_;
  std::multimap<size_t, HostBlock> free_;
  size_t live_ = 0;
};

// Buffers whose copies are in flight, each with the event recorded after its
// copies. Query: int(Event) returning 1 complete, 0 pending, < 0 error.
template <class Event>
class InFlight {
 public:
  void hold(long long owner, HostBlock b, Event e) { held_.push_back({owner, b, e}); }
  // Returns buffers whose event has completed to `pool`; the first error
  // code of a query, else 0.
  template <class Pool, class Query>
  int reclaim(Pool& pool, Query query) {
    int err = 0;
    for (size_t i = 0; i < held_.size();) {
      const int q = query(held_[i].e);
      if (q < 0 && !err) err = q;
      if (q == 1) {
        pool.put(held_[i].b);
        held_[i] = held_.back();
        held_.pop_back();
      } else {
        ++i;
      }
    }
    return err;
  }
  // Buffers still held for an owner (a ticket may be fetched only without).
  size_t held_by(long long owner) const {
    size_t n = 0;
    for (const auto& h : held_) n += h.owner == owner ? 1 : 0;
    return n;
  }
  size_t size() const { return held_.size(); }
  // Waits for every event (Wait: void(Event)) and returns all buffers.
  template <class Pool, class Wait>
  void drain(Pool& pool, Wait wait) {
    for (auto& h : held_) {
      wait(h.e);
      pool.put(h.b);
    }
    held_.clear();
  }

 private:
  struct H {
    long long owner;
    HostBlock b;
    Event e;
  };
  std::vector<H> held_;
};

}  // namespace tgmm
