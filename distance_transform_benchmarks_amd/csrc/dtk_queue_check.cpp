// CPU model check of the tile work queue (dtk_queue.hpp): the queue
// functions the persistent kernel runs, run here by many CPU threads on a
// synthetic wake graph.  The wave-level parts of the kernel (the drain of
// the later rings, one item per lane, and the publish of up to nine tiles)
// are restated lane by lane.  Test infrastructure: it runs on the host only
// and is not part of libdtk.
//
// Checks, for the FIFO, KEY and SEED orders (deferral windows 0 and 3) and
// several sizes, thread counts and seeds:
//   - every tile id and bucket read back from the queue is in range (the
//     queue flags ERR_QUEUE otherwise), and no wait hits its spin limit;
//   - no tile is visited by two workers at once;
//   - no wake is lost: every tile's last visit starts after its last wake;
//   - the workers terminate: DONE is raised exactly when pending reaches 0;
// each case runs with and without workers keeping a tile that was woken
// during its visit (end_visit's keep).
//
// Usage: dtk_queue_check [iterations]   (exit status 0 = all checks passed)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "dtk_launch.hpp"
#include "dtk_queue.hpp"

using namespace dtk;

namespace
{
  // A CPU wait counts as stalled after kStallSeconds of wall time (a whole
  // case runs in milliseconds; a thread of an oversubscribed CPU can still be
  // descheduled for a while, so the limit is time, not a poll count).
  constexpr double kStallSeconds = 3.0;
  struct Deadline
  {
    std::chrono::steady_clock::time_point start = std::chrono::steady_clock::now();
    bool expired(uint32_t n) const
    {
      return (n & 1023u) == 1023u &&
             std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count() > kStallSeconds;
    }
  };
  struct HostNap
  {
    Deadline d;
    bool     operator()(uint32_t n) const
    {
      if (n > 64)
        std::this_thread::yield();
      return !d.expired(n);
    }
  };

  struct Model
  {
    int                              tx = 1, ty = 1, ntiles = 1, workers = 1;
    int                              order = ORDER_FIFO;
    int                              window = 0;
    std::vector<uint32_t>            ctrl;
    std::vector<uint32_t>            state, tkey;
    std::vector<uint8_t>             tprio;
    std::vector<unsigned long long>  slots, bslots;
    Queue                            q{};
    // bookkeeping of the check (not part of the queue)
    std::vector<std::atomic<int>>    running;      // visitors inside a tile
    std::vector<std::atomic<uint32_t>> wakes;      // wakes issued to a tile
    std::vector<std::atomic<uint32_t>> covered;    // wakes seen at the start of its latest visit
    std::vector<std::atomic<int>>    budget;       // visits that may still wake neighbours
    std::atomic<uint32_t>            visits{0}, faults{0}, kept{0};
    bool                             keeps = false; // workers may keep a tile woken during its visit
  };

  uint32_t next_pow2(uint32_t v)
  {
    uint32_t p = 1;
    while (p < v)
      p <<= 1;
    return p;
  }

  void setup(Model& m, int tx, int ty, int workers, int order, int window, bool keeps, uint32_t seed)
  {
    m.keeps = keeps;
    m.tx = tx;
    m.ty = ty;
    m.ntiles  = tx * ty;
    m.workers = workers;
    m.order   = order;
    m.window  = window;
    const uint32_t cap = next_pow2((uint32_t)(m.ntiles + workers + 1));
    m.ctrl.assign(CTRL_WORDS + 2, 0u);
    m.state.assign((size_t)m.ntiles * kStateStride, 0u);
    m.tkey.assign(m.ntiles, kInfBits);
    m.tprio.assign(m.ntiles, 0);
    std::mt19937 rng(seed);
    for (auto& p : m.tprio)
      p = (uint8_t)(rng() % kBuckets);
    m.q        = Queue{};
    m.q.ctrl   = m.ctrl.data();
    m.q.state  = m.state.data();
    m.q.ntiles = m.ntiles;
    m.q.order  = order;
    m.q.cap_mask = cap - 1;
    m.q.cap_log2 = 0;
    while ((1u << m.q.cap_log2) < cap)
      ++m.q.cap_log2;
    m.slots.assign((size_t)cap * kSlotStride, 0ull);
    for (uint32_t k = 0; k < cap; ++k)
      m.slots[(size_t)k * kSlotStride] = (unsigned long long)(2u * k) << 32;
    m.q.slots  = m.slots.data();
    m.q.window = order == ORDER_FIFO ? (int)kBuckets : window;
    if (order != ORDER_FIFO)
    {
      m.bslots.assign((size_t)kBuckets * cap, 0ull);
      for (size_t k = 0; k < m.bslots.size(); ++k)
        m.bslots[k] = (unsigned long long)(2u * ((uint32_t)k & (cap - 1))) << 32;
      m.q.bslots = m.bslots.data();
      m.q.tkey   = m.tkey.data();
      m.q.tprio  = m.tprio.data();
    }
    m.running  = std::vector<std::atomic<int>>(m.ntiles);
    m.wakes    = std::vector<std::atomic<uint32_t>>(m.ntiles);
    m.covered  = std::vector<std::atomic<uint32_t>>(m.ntiles);
    m.budget   = std::vector<std::atomic<int>>(m.ntiles);
    for (int t = 0; t < m.ntiles; ++t)
      m.budget[t] = 1 + (int)(rng() % 6);
    m.visits = 0;
    m.faults = 0;
    m.kept   = 0;
  }

  // Initial dirty tiles (k_seed): every third tile, with a random key.
  void seed_queue(Model& m, uint32_t s)
  {
    std::mt19937 rng(s);
    const Queue& q = m.q;
    for (int t = 0; t < m.ntiles; t += 3)
    {
      q.state[(size_t)t * kStateStride] = TILE_DIRTY;
      const float key = (float)(rng() % 100000) * 0.01f;
      uint32_t    kb;
      std::memcpy(&kb, &key, 4);
      const uint32_t bucket = m.order == ORDER_KEY ? key_bucket(kb) : m.order == ORDER_SEED ? q.tprio[t] : 0u;
      if (m.order == ORDER_KEY)
        q.tkey[t] = kb;
      if (!deferred(q, bucket, 0u))
      {
        const uint32_t ticket = (uint32_t)at_add(tp_word(q), (unsigned long long)kOnePendingOneTicket);
        q.slots[(size_t)(ticket & q.cap_mask) * kSlotStride] = ((unsigned long long)(2u * ticket + 1u) << 32) | (uint32_t)t;
      }
      else
      {
        at_add(tp_word(q), 1ull << 32);
        later_push(q, bucket, (uint32_t)t, HostNap{});
      }
    }
  }

  // The kernel's drain_later, lane by lane (the caller holds CTRL_DRAIN).
  // Returns the number of tiles moved.
  uint32_t drain_host(const Queue& q, uint32_t want)
  {
    uint32_t moved = 0;
    uint32_t h[kBuckets], t[kBuckets];
    for (uint32_t b = 0; b < kBuckets; ++b)
      h[b] = at_load(&q.ctrl[CTRL_LHEAD + b]);
    for (uint32_t b = 0; b < kBuckets; ++b)
      t[b] = at_load(&q.ctrl[CTRL_LTAIL + b]);
    int b0 = -1;
    for (uint32_t k = 0; k < kBuckets && b0 < 0; ++k)
      if ((int)(t[k] - h[k]) > 0)
        b0 = (int)k;
    if (b0 < 0)
      return 0;
    const int first = std::min(b0 + q.window, (int)kBuckets - 1);
    for (int b = b0; b < (int)kBuckets; ++b)
    {
      if ((int)(t[b] - h[b]) <= 0)
        continue;
      if (b > first && moved >= want)
        break;
      at_store(&q.ctrl[CTRL_FRONT], (uint32_t)b);
      for (uint32_t base = h[b]; (int)(t[b] - base) > 0; base += 64u)
      {
        const uint32_t n = std::min(64u, t[b] - base);
        int            tiles[64];
        uint32_t       ok = 0;
        for (uint32_t i = 0; i < n; ++i)
        {
          tiles[i] = -1;
          const Deadline dl;
          for (uint32_t spin = 0;; ++spin)
          {
            tiles[i] = later_poll(q, (uint32_t)b, base + i);
            if (tiles[i] != -2)
              break;
            if (dl.expired(spin))
            {
              at_or(&q.ctrl[CTRL_ERROR], (uint32_t)ERR_WATCHDOG);
              return moved; // stalled: report it instead of waiting out every item
            }
            std::this_thread::yield();
          }
          ok += tiles[i] >= 0 ? 1u : 0u;
        }
        uint32_t tk = (uint32_t)at_add(tp_word(q), (unsigned long long)ok);
        for (uint32_t i = 0; i < n; ++i)
          if (tiles[i] >= 0)
          {
            ring_fill(q, &q.slots[(size_t)(tk & q.cap_mask) * kSlotStride], tk, (uint32_t)tiles[i], HostNap{});
            ++tk;
          }
        moved += ok;
      }
      at_store(&q.ctrl[CTRL_LHEAD + b], t[b]);
    }
    return moved;
  }

  // The kernel's fill_main for one item: fill the main-ring slot of
  // `ticket`; while it is still held one lap back, take this worker's own
  // claimed tile as soon as it is filled (*pre).
  void fill_main_host(const Queue& q, uint32_t ticket, uint32_t tile, uint32_t claim, int* pre)
  {
    unsigned long long*      slot = &q.slots[(size_t)(ticket & q.cap_mask) * kSlotStride];
    const unsigned long long full = ((unsigned long long)(2u * ticket + 1u) << 32) | tile;
    const Deadline           dl;
    for (uint32_t spin = 0;; ++spin)
    {
      unsigned long long expect = (unsigned long long)(2u * ticket) << 32;
      if (at_cas(slot, expect, full))
        return;
      if (*pre < 0)
      {
        const int t = fifo_poll(q, claim);
        if (t >= 0)
          *pre = t;
      }
      if (dl.expired(spin))
      {
        at_or(&q.ctrl[CTRL_ERROR], (uint32_t)ERR_WATCHDOG);
        return;
      }
      if (spin > 64)
        std::this_thread::yield();
    }
  }

  void worker(Model& m, uint32_t seed)
  {
    const Queue& q = m.q;
    std::mt19937 rng(seed);
    if (at_load(tp_word(q)) >> 32 == 0)
      return;
    uint32_t ticket = fifo_claim(q);
    int      pre    = -1;
    for (;;)
    {
      // ---- pop (the kernel's pop, with the same poll cadences)
      int            t = -1;
      const Deadline dl;
      for (uint32_t spin = 0;; ++spin)
      {
        t   = pre >= 0 ? pre : fifo_poll(q, ticket);
        pre = -1;
        if (t != -2)
          break;
        if ((spin & 7u) == 7u && is_done(q))
        {
          t = -1;
          break;
        }
        if (q.order != ORDER_FIFO && (spin & 3u) == 3u)
        {
          const uint32_t tail = (uint32_t)at_load(tp_word(q));
          if ((int)(tail - ticket) <= 0 && (tail == ticket || spin > 32u) && drain_try_lock(q))
          {
            const uint32_t want  = at_load(&q.ctrl[CTRL_HEAD]) - (uint32_t)at_load(tp_word(q));
            const uint32_t moved = drain_host(q, (int)want > 0 ? want : 1u);
            drain_unlock(q);
            if (moved)
              continue;
          }
        }
        if (dl.expired(spin))
        {
          at_or(&q.ctrl[CTRL_ERROR], (uint32_t)ERR_WATCHDOG);
          t = -1;
          break;
        }
        std::this_thread::yield();
      }
      if (t < 0)
        break;
      uint32_t own_key = begin_visit(q, t);
      ticket           = fifo_claim(q);
      const uint32_t front   = at_load(&q.ctrl[CTRL_FRONT]);
      int            kept    = 0;
    revisit:
      // ---- visit
      if (m.running[t].fetch_add(1) != 0)
        m.faults.fetch_add(1);
      m.covered[t].store(m.wakes[t].load());
      m.visits.fetch_add(1);
      int      targets[9];
      uint32_t keys[9];
      int      nt      = 0;
      bool     requeue = false;
      if (m.budget[t].fetch_sub(1) > 0)
      {
        const int        x = t % m.tx, y = t / m.tx;
        static const int dx[8] = {-1, 1, 0, 0, -1, 1, -1, 1}, dy[8] = {0, 0, -1, 1, -1, -1, 1, 1};
        for (int k = 0; k < 8; ++k)
        {
          const int nx = x + dx[k], ny = y + dy[k];
          if ((rng() & 1u) || nx < 0 || ny < 0 || nx >= m.tx || ny >= m.ty)
            continue;
          const float key = (float)(rng() % 100000) * 0.01f;
          targets[nt]     = ny * m.tx + nx;
          std::memcpy(&keys[nt], &key, 4);
          ++nt;
        }
        requeue = (rng() % 8u) == 0u;
      }
      if ((rng() % 4u) == 0u)
        std::this_thread::yield();
      if ((rng() % 64u) == 0u) // a slow visit: the other workers run laps of the ring meanwhile
        std::this_thread::sleep_for(std::chrono::microseconds(200));
      m.running[t].fetch_sub(1);
      // ---- publish (lanes 0..nt-1 wake, lane nt ends the visit)
      bool     insert[9], later[9];
      uint32_t bucket[9];
      for (int i = 0; i < nt; ++i)
      {
        m.wakes[targets[i]].fetch_add(1);
        insert[i] = wake_tile(q, targets[i], keys[i], &bucket[i]);
      }
      targets[nt] = t;
      // The kernel keeps a tile woken during its visit (up to kMaxKeeps
      // times in a row); the check keeps at random so both ends are taken.
      bool           keep    = false;
      uint32_t       new_key = kInfBits;
      const bool     keep_ok = m.keeps && kept < (int)kMaxKeeps && (rng() & 1u);
      insert[nt]             = end_visit(q, t, requeue, own_key, &bucket[nt], keep_ok, &keep, &new_key);
      uint32_t n_all = 0, n_main = 0;
      for (int i = 0; i <= nt; ++i)
      {
        later[i] = insert[i] && deferred(q, bucket[i], front);
        n_all += insert[i] ? 1u : 0u;
        n_main += (insert[i] && !later[i]) ? 1u : 0u;
      }
      unsigned long long old = 0;
      if (!keep)
      {
        old = publish_count(q, n_main, n_all);
        if (n_all == 0u && (uint32_t)(old >> 32) == 1u)
          raise_done(q);
      }
      else if (n_all)
        old = publish_count_keep(q, n_main, n_all);
      uint32_t tk = (uint32_t)old;
      for (int i = 0; i <= nt; ++i)
      {
        if (insert[i] && !later[i])
        {
          fill_main_host(q, tk, (uint32_t)targets[i], ticket, &pre);
          ++tk;
        }
        if (later[i])
          later_push(q, bucket[i], (uint32_t)targets[i], HostNap{});
      }
      if (keep)
      {
        // The same tile again, with the absorbed wakes' key.
        own_key = new_key;
        ++kept;
        m.kept.fetch_add(1);
        goto revisit;
      }
    }
  }

  int run_case(int tx, int ty, int threads, int order, int window, bool keeps, uint32_t seed)
  {
    Model m;
    setup(m, tx, ty, threads, order, window, keeps, seed);
    seed_queue(m, seed * 7u + 1u);
    std::vector<std::thread> th;
    for (int i = 0; i < threads; ++i)
      th.emplace_back(worker, std::ref(m), seed * 131u + (uint32_t)i);
    for (auto& t : th)
      t.join();
    int bad = 0;
    const uint32_t err = m.ctrl[CTRL_ERROR];
    const uint32_t pend = (uint32_t)(at_load(tp_word(m.q)) >> 32);
    if (err)
      ++bad, std::printf("  error bits %u\n", err);
    if (m.faults)
      ++bad, std::printf("  %u concurrent visits\n", m.faults.load());
    if (pend != 0)
      ++bad, std::printf("  pending %u at exit\n", pend);
    if (!is_done(m.q) && m.visits > 0)
      ++bad, std::printf("  DONE not raised\n");
    for (int t = 0; t < m.ntiles; ++t)
      if (m.covered[t].load() != m.wakes[t].load())
      {
        ++bad;
        std::printf("  tile %d: %u wakes, last visit saw %u\n", t, m.wakes[t].load(), m.covered[t].load());
        break;
      }
    for (int t = 0; t < m.ntiles; ++t)
      if (m.state[(size_t)t * kStateStride] != 0u)
      {
        ++bad;
        std::printf("  tile %d left in state %u\n", t, m.state[(size_t)t * kStateStride]);
        break;
      }
    std::printf("%s order=%d window=%d keep=%d tiles=%dx%d threads=%d seed=%u visits=%u kept=%u\n", bad ? "FAIL" : "ok  ",
                order, window, (int)keeps, tx, ty, threads, seed, m.visits.load(), m.kept.load());
    return bad;
  }
} // namespace

// Launch shapes that the host chooses without a device (dtk_launch.hpp):
// the grid of the adjoint's tile kernels.
int check_launch_shapes()
{
  int bad = 0;
  auto expect = [&](int n, int W, int H, bool ok, unsigned wx, unsigned wy) {
    unsigned   gx = 0, gy = 0;
    const bool got = backward_tiles(n, W, H, 64, &gx, &gy);
    if (got != ok || (ok && (gx != wx || gy != wy)))
    {
      std::printf("FAIL backward tiles n=%d %dx%d: %d %ux%u, expected %d %ux%u\n", n, W, H, (int)got, gx, gy, (int)ok,
                  wx, wy);
      ++bad;
    }
  };
  expect(32, 1024, 1024, true, 16, 16);      // diff32
  expect(1, 1000, 700, true, 16, 11);        // partial tiles round up
  expect(1, 1, 65535 * 64, true, 1, 65535);  // the tallest image: 65535 tile rows
  expect(1, 1, 65535 * 64 + 1, false, 0, 0); // one row more needs a 65536th
  expect(65535, 64, 64, true, 1, 1);         // the most images per launch
  expect(65536, 64, 64, false, 0, 0);
  expect(0, 64, 64, false, 0, 0);
  {
    // the tile side sets the grid: 32-pixel tiles need twice the blocks per side
    unsigned gx = 0, gy = 0;
    if (!backward_tiles(2, 1000, 700, 32, &gx, &gy) || gx != 32 || gy != 22)
    {
      std::printf("FAIL backward tiles with 32-pixel tiles: %ux%u\n", gx, gy);
      ++bad;
    }
  }
  std::printf(bad ? "launch shape checks FAILED\n" : "launch shape checks passed\n");
  return bad;
}

int main(int argc, char** argv)
{
  if (check_launch_shapes())
    return 1;
  const int iters = argc > 1 ? std::atoi(argv[1]) : 2;
  const int shapes[][2] = {{1, 1}, {2, 1}, {3, 3}, {8, 8}, {40, 25}};
  int       bad = 0;
  for (int it = 0; it < iters; ++it)
    for (int order : {ORDER_FIFO, ORDER_KEY, ORDER_SEED})
      for (const auto& s : shapes)
        for (int threads : {1, 4, 24})
          for (int window : {0, 3})
            for (bool keeps : {false, true})
          {
            if (order == ORDER_FIFO && window != 0)
              continue;
            bad += run_case(s[0], s[1], threads, order, window, keeps,
                            1000u * (uint32_t)it + (uint32_t)(s[0] * 31 + threads));
            if (bad)
            {
              std::printf("queue model check FAILED\n");
              return 1;
            }
          }
  std::printf("queue model check passed\n");
  return 0;
}
