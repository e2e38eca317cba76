// Work queue of dirty tiles: the operations of one lane (or one CPU thread).
//
// Replaces the reference's two-level task queue (dt/src/task_queue.cuh:20-150:
// 64 LIFO priority buckets with a non-empty mask, a bitset that
// de-duplicates enqueues, grid-wide rounds) with a lock-free design that
// needs no grid barrier.  Every function here is __host__ __device__ and
// uses the HIP atomic builtins, which also run on CPU threads: the
// persistent kernel and the CPU model check (csrc/dtk_queue_check.cpp) run
// this same code.
//
// Tile states (one word per tile, changed only by read-modify-writes, so
// every waker's stores are in the release sequence that the next visit
// acquires from):
//   0                  idle
//   TILE_DIRTY         queued (in a ring, not yet popped)
//   TILE_RUNNING       being solved by one worker
//   RUNNING|DIRTY      being solved, woken again meanwhile (re-queued at the end,
//                      or visited again at once by the same worker: end_visit's keep)
// A waker does fetch_or(TILE_DIRTY); the one that sees 0 inserts the tile,
// so a tile is in the queue at most once and never solved by two workers.
//
// Containers.  Workers pop only from the main ring: one ring of tickets
// whose tail and pending count share a 64-bit word (one atomic reserves
// tickets and counts items), claimed with fetch_add and waited on one slot
// per worker, so popping never contends.  The bucket orders give each
// inserted tile a bucket:
//   ORDER_KEY   the smallest improved halo value the wake carries (two
//               buckets per octave of the distance), folded into a per-tile
//               minimum;
//   ORDER_SEED  a static per-tile priority: the city-block tile distance to
//               the nearest seed tile, quantised by its cdf into 64 levels
//               (the reference's dt/src/block_priority.cu:74-155).
// A tile whose bucket lies beyond the front (CTRL_FRONT + window) is
// deferred into that bucket's later ring instead of the main ring.  When a
// waiting worker finds the main ring empty, it drains the lowest non-empty
// later buckets into the main ring and moves the front there.  So tiles
// near the front of final values are solved first, and far-off wakes wait
// until the front reaches them (most of them are improved again before
// then).  ORDER_FIFO never defers.
// A ring slot holds (seq << 32 | tile); seq = 2*ticket while the slot is
// free for that ticket and 2*ticket + 1 once filled.  Tile ids and buckets
// read back from a ring are bounds-checked: a protocol fault flags
// ERR_QUEUE instead of addressing memory out of range.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include "dtk_internal.hpp"

namespace dtk
{
#define DTK_HD __host__ __device__ __forceinline__

  // Relaxed agent-scope atomics (ordering comes from the explicit fences of
  // the callers and from data dependencies between atomics).
  template <typename T>
  DTK_HD T at_load(T* p)
  {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  template <typename T>
  DTK_HD void at_store(T* p, T v)
  {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  template <typename T>
  DTK_HD T at_add(T* p, T v)
  {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  template <typename T>
  DTK_HD T at_or(T* p, T v)
  {
    return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  template <typename T>
  DTK_HD T at_and(T* p, T v)
  {
    return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  template <typename T>
  DTK_HD T at_xchg(T* p, T v)
  {
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  template <typename T>
  DTK_HD T at_min(T* p, T v)
  {
    return __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  template <typename T>
  DTK_HD bool at_cas(T* p, T& expect, T v)
  {
    return __hip_atomic_compare_exchange_strong(p, &expect, v, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
  }

  constexpr uint32_t kInfBits   = 0x7f800000u; // bits of +inf: the key of "no pending wake"
  constexpr uint32_t kNoBucket  = 0xffffffffu;
  constexpr uint64_t kOnePendingOneTicket = (1ull << 32) | 1ull;
  constexpr uint64_t kMinusOnePending     = ~0ull << 32; // adds -1 to the high half

  // Bucket of a key given as the bits of a float in [+0, +inf]: two buckets
  // per octave, bucket 0 below 2^-3, the last bucket from 2^28 (and +inf).
  // The bits of such floats order like the values, so this is monotone.
  DTK_HD uint32_t key_bucket(uint32_t kbits)
  {
    const int b = (int)(kbits >> 22) - ((127 - 3) << 1);
    return (uint32_t)(b < 0 ? 0 : (b > (int)kBuckets - 1 ? (int)kBuckets - 1 : b));
  }

  // ---- waits --------------------------------------------------------------
  // Spin policy of a wait: `nap(n)` is called between polls with the poll
  // count and returns false when the wait should give up (the caller then
  // flags ERR_WATCHDOG).  The device naps on the shader clock and gives up
  // after kSlotWaitLimit polls.
  struct DeviceNap
  {
    __device__ bool operator()(uint32_t n) const
    {
      __builtin_amdgcn_s_sleep(1);
      return n < kSlotWaitLimit;
    }
  };

  // ---- FIFO ring ---------------------------------------------------------
  DTK_HD unsigned long long* tp_word(const Queue& q) { return reinterpret_cast<unsigned long long*>(q.ctrl + CTRL_TP); }

  DTK_HD uint32_t fifo_claim(const Queue& q) { return at_add(&q.ctrl[CTRL_HEAD], 1u); }

  // Reads the tile of `ticket` once its slot is filled and frees the slot for
  // ticket + capacity.  Returns -2 while the slot is not filled yet, -1 after
  // a protocol fault (flagged in CTRL_ERROR).
  DTK_HD int fifo_poll(const Queue& q, uint32_t ticket)
  {
    unsigned long long*      slot = &q.slots[(size_t)(ticket & q.cap_mask) * kSlotStride];
    const unsigned long long s    = at_load(slot);
    if ((uint32_t)(s >> 32) != 2u * ticket + 1u)
      return -2;
    const uint32_t t = (uint32_t)s;
    at_store(slot, (unsigned long long)(2u * (ticket + q.cap_mask + 1u)) << 32);
    if (t >= (uint32_t)q.ntiles)
    {
      at_or(&q.ctrl[CTRL_ERROR], (uint32_t)ERR_QUEUE);
      return -1;
    }
    return (int)t;
  }

  // Fills the slot of a ticket reserved by the caller.  The slot is free for
  // it once the popper of ticket - capacity has read its tile.
  template <class Nap>
  DTK_HD bool ring_fill(const Queue& q, unsigned long long* slot, uint32_t ticket, uint32_t tile, Nap nap)
  {
    const unsigned long long full = ((unsigned long long)(2u * ticket + 1u) << 32) | tile;
    for (uint32_t spin = 0;; ++spin)
    {
      unsigned long long expect = (unsigned long long)(2u * ticket) << 32;
      if (at_cas(slot, expect, full))
        return true;
      if (!nap(spin))
      {
        at_or(&q.ctrl[CTRL_ERROR], (uint32_t)ERR_WATCHDOG);
        return false;
      }
    }
  }

  // ---- later rings (bucket orders) ----------------------------------------
  DTK_HD unsigned long long* later_slot(const Queue& q, uint32_t b, uint32_t ticket)
  {
    return &q.bslots[((size_t)b << q.cap_log2) + (ticket & q.cap_mask)];
  }

  // Given the word `s` just read from the slot of ticket `ticket` of later
  // ring b: when it is filled (its pusher fills it right after reserving the
  // ticket) frees the slot for ticket + capacity and returns the tile.
  // -2: not filled yet; -1: protocol fault.  (Split from the load so that a
  // drainer can issue the loads of several slots together.)
  DTK_HD int later_take(const Queue& q, uint32_t b, uint32_t ticket, unsigned long long s)
  {
    if (b >= kBuckets)
    {
      at_or(&q.ctrl[CTRL_ERROR], (uint32_t)ERR_QUEUE);
      return -1;
    }
    if ((uint32_t)(s >> 32) != 2u * ticket + 1u)
      return -2;
    const uint32_t t = (uint32_t)s;
    at_store(later_slot(q, b, ticket), (unsigned long long)(2u * (ticket + q.cap_mask + 1u)) << 32);
    if (t >= (uint32_t)q.ntiles)
    {
      at_or(&q.ctrl[CTRL_ERROR], (uint32_t)ERR_QUEUE);
      return -1;
    }
    return (int)t;
  }
  DTK_HD int later_poll(const Queue& q, uint32_t b, uint32_t ticket)
  {
    if (b >= kBuckets)
      return later_take(q, b, ticket, 0ull);
    return later_take(q, b, ticket, at_load(later_slot(q, b, ticket)));
  }

  // Defers a tile (already counted as pending) into later ring b.
  template <class Nap>
  DTK_HD bool later_push(const Queue& q, uint32_t b, uint32_t tile, Nap nap)
  {
    if (b >= kBuckets)
      b = kBuckets - 1;
    const uint32_t ticket = at_add(&q.ctrl[CTRL_LTAIL + b], 1u);
    return ring_fill(q, later_slot(q, b, ticket), ticket, tile, nap);
  }

  // The drain lock (CTRL_DRAIN): one worker at a time moves later rings into
  // the main ring.  Taken with acquire and given back with release, so a
  // drainer sees the later-ring heads its predecessor stored (without that
  // order it could read an old head and wait for tickets already drained).
  DTK_HD bool drain_try_lock(const Queue& q)
  {
    if (at_load(&q.ctrl[CTRL_DRAIN]) != 0u)
      return false;
    uint32_t e = 0u;
    return __hip_atomic_compare_exchange_strong(&q.ctrl[CTRL_DRAIN], &e, 1u, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
  }
  DTK_HD void drain_unlock(const Queue& q)
  {
    __hip_atomic_store(&q.ctrl[CTRL_DRAIN], 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }

  // True when an inserted tile of bucket b goes to a later ring (bucket
  // orders only) given the front read by the caller.
  DTK_HD bool deferred(const Queue& q, uint32_t b, uint32_t front)
  {
    return q.order != ORDER_FIFO && b > front + (uint32_t)q.window;
  }

  // ---- tile state machine ----------------------------------------------
  // A popped tile starts its visit: take the key of its queued entry (the
  // wakes' minimum; +inf when none), then queued -> RUNNING.  The exchange
  // continues the release sequence of every earlier state RMW (the wakers'
  // and the previous visit's); the caller's acquire fence then makes their
  // stores visible to the visit.
  DTK_HD uint32_t begin_visit(const Queue& q, int t)
  {
    uint32_t key = kInfBits;
    if (q.order == ORDER_KEY)
      key = at_xchg(&q.tkey[t], kInfBits);
    at_xchg(&q.state[(size_t)t * kStateStride], (uint32_t)TILE_RUNNING);
    return key;
  }

  // A neighbour's visit improved pixels of tile u (key kb: the smallest
  // improved value).  Returns true when the caller must insert u; *bucket
  // receives its bucket (bucket orders).
  DTK_HD bool wake_tile(const Queue& q, int u, uint32_t kb, uint32_t* bucket)
  {
    if (q.order == ORDER_KEY)
    {
      const uint32_t old = at_min(&q.tkey[u], kb);
      *bucket            = key_bucket(old < kb ? old : kb);
    }
    else if (q.order == ORDER_SEED)
      *bucket = q.tprio[u];
    else
      *bucket = 0;
    return at_or(&q.state[(size_t)u * kStateStride], (uint32_t)TILE_DIRTY) == 0u;
  }

  // The owner of tile t ends its visit: RUNNING -> idle, or -> queued when a
  // wake arrived meanwhile (DIRTY stays set) or when the visit stopped at
  // its sweep cap (`requeue`).  own_key: the key the visit was popped with.
  // Returns true when the caller must insert t.
  //
  // Keeping (keep_ok): a wake that arrived during the visit is absorbed by
  // visiting the tile again at once instead of queueing it.  The and above
  // left the state DIRTY alone: no ring holds the tile, and wakers see a
  // non-zero state and insert nothing, so only this owner can change it; the
  // exchange makes it RUNNING again (*keep = true, nothing to insert, the
  // tile's pending item stays).  ORDER_KEY: *own_key receives the key of the
  // absorbed wakes.  The caller needs an acquire fence before the next
  // visit's loads, as after a pop (the and read the wakers' RMWs).
  DTK_HD bool end_visit(const Queue& q, int t, bool requeue, uint32_t own_key, uint32_t* bucket, bool keep_ok = false,
                        bool* keep = nullptr, uint32_t* new_key = nullptr)
  {
    uint32_t*      st  = &q.state[(size_t)t * kStateStride];
    const uint32_t old = at_and(st, ~(uint32_t)TILE_RUNNING);
    bool           ins = false;
    *bucket            = 0;
    if (keep)
      *keep = false;
    if ((old & TILE_DIRTY) && keep_ok)
    {
      at_xchg(st, (uint32_t)TILE_RUNNING);
      *keep = true;
      if (new_key)
        *new_key = q.order == ORDER_KEY ? at_xchg(&q.tkey[t], kInfBits) : kInfBits;
      return false;
    }
    if (old & TILE_DIRTY)
      ins = true;
    else if (requeue)
      ins = at_or(st, (uint32_t)TILE_DIRTY) == 0u;
    if (!ins)
      return false;
    if (q.order == ORDER_KEY)
    {
      const uint32_t k = at_load(&q.tkey[t]);
      *bucket          = key_bucket(k < own_key ? k : own_key);
    }
    else if (q.order == ORDER_SEED)
      *bucket = q.tprio[t];
    return true;
  }

  // The one atomic of a publish: n_main tickets of the main ring, n_all new
  // pending items (main and later) and this visit's item retired.  Returns
  // the word before; its low half is the first ticket.  The caller raises
  // DONE when pending was 1 and n_all is 0.  The new items are counted
  // before any of them is pushed (the pushes wait for this value), so
  // pending cannot reach 0 while one of them is still queued.
  DTK_HD unsigned long long publish_count(const Queue& q, uint32_t n_main, uint32_t n_all)
  {
    return at_add(tp_word(q), (unsigned long long)n_main + ((unsigned long long)n_all << 32) + kMinusOnePending);
  }
  // The same for a visit whose tile is kept (end_visit's keep): its item stays
  // pending, so pending cannot reach 0 here.
  DTK_HD unsigned long long publish_count_keep(const Queue& q, uint32_t n_main, uint32_t n_all)
  {
    return at_add(tp_word(q), (unsigned long long)n_main + ((unsigned long long)n_all << 32));
  }

  DTK_HD void raise_done(const Queue& q) { at_store(&q.ctrl[CTRL_DONE], 1u); }
  // Idle workers stop once DONE is raised, or once any error is flagged (a
  // fault leaves work that never drains; stopping reports it at once).
  DTK_HD bool is_done(const Queue& q)
  {
    return at_load(&q.ctrl[CTRL_DONE]) != 0u || at_load(&q.ctrl[CTRL_ERROR]) != 0u;
  }
} // namespace dtk
