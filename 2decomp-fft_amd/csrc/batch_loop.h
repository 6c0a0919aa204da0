// batch_loop.h -- the device work queues of a persistent batch kernel.
//
// A queue item covers up to per_item consecutive tiles of one BatchGroup
// (one stage of one grid), so that the queue, barrier and descriptor costs
// of an item are spread over enough work.  Items are numbered phase-major
// (stage 0 of every grid, then stage 1, ...), and phase p starts only when
// every item of phase p-1 has completed.  Each phase's items are dealt to
// kBatchQueues queues in chunks of a few consecutive items, round robin, and
// block b takes items in order from queue b % kBatchQueues with an atomic on
// that queue's counter: the counters of different XCDs do not contend, each
// XCD works on neighbouring tiles, and every queue gets a like mix of work.
//
// The synchronisation is per block and per phase, not per item: a block
// counts the items it completes in its current phase and publishes the count
// with one agent-scope release on its queue's counter of that phase when it
// moves to an item of a later phase or finds its queue empty; before its
// first item of phase p it polls the kBatchQueues counters of phase p-1 with
// relaxed agent-scope loads until each equals its queue's item count of the
// phase, then takes one agent-scope acquire.  (One counter per queue and
// phase, each on its own 256-byte line: the publishing atomics of a phase
// spread over kBatchQueues memory channels instead of queuing at one.)  Each gfx950
// XCD has its own L2, so the release writes back that L2 and the acquire
// invalidates it; doing both once per phase instead of once per tile (and
// never inside the poll loop) keeps the L2 warm for the blocks still working.
//
// Progress: a block holds its current item and the positions of the next
// one or two of its queue, all in queue order, so its phases never
// decrease; it finishes every item it holds of phases below p, publishing
// at each change of phase, before it waits for phase p-1.  Inductively over
// the phases every wait ends as long as each queue has a running block.  A
// limit on the wall clock (s_memrealtime ticks) turns a broken schedule
// into an error flag instead of a hang.
//
// Pipelining: the position of the item after next is dequeued while the
// current item runs (run_tile issues the atomic when asked, so that its
// latency overlaps the tile's first loads), and the next item's queue entry
// is a scalar load issued before the current item starts.
//
// The loop control is read back from LDS and made wave-uniform with
// readfirstlane, so every barrier stays in uniform control flow.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.h"

namespace d2dhip {

__device__ __forceinline__ long long uniform_ll(long long v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(v & 0xffffffffll));
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)v >> 32));
  return (long long)(((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ QueueEntry queue_entry(const QueueEntry *__restrict__ e, long long pos,
                                                  long long len, int nphase) {
  if (pos < len) return e[pos];
  return QueueEntry{0, 0, nphase, 0};
}

// Runs the block's queue to its end.  run_tile(group, tile, head) computes
// one tile of a group with the whole block (wave-uniform arguments,
// barriers allowed); when head is not null, thread 0 also takes the next
// position of that queue counter and returns it.
template <class RunTile>
__device__ __forceinline__ void batch_loop(const BatchQueue &Q, RunTile &&run_tile) {
  __shared__ long long s_pos[2];
  __shared__ int s_fail;
  const int tid = threadIdx.x;
  const int qi = (int)(blockIdx.x % kBatchQueues);
  const QueueEntry *ent = Q.entries + Q.off[qi];
  const long long len = Q.len[qi];
  unsigned long long *head = Q.heads + qi * kHeadStride;
  if (tid == 0) {
    s_pos[0] = (long long)atomicAdd(head, 1ull);
    s_pos[1] = (long long)atomicAdd(head, 1ull);
  }
  __syncthreads();
  long long pos = uniform_ll(s_pos[0]), pos_n = uniform_ll(s_pos[1]);
  QueueEntry e = queue_entry(ent, pos, len, Q.nphase);
  int phase = 0;                // phase of the items counted in `held`
  unsigned long long held = 0;  // items of `phase` completed here, not yet published
  int fail = 0;
  for (;;) {
    const QueueEntry e_n = queue_entry(ent, pos_n, len, Q.nphase);  // used next round
    const int ph = __builtin_amdgcn_readfirstlane(e.phase);
    if (ph != phase) {
      if (tid == 0) {
        s_fail = 0;
        if (held)
          __hip_atomic_fetch_add(&Q.done[((long long)phase * kBatchQueues + qi) * kHeadStride], held,
                                 __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        if (ph < Q.nphase && ph > 0) {
          const long long t0 = (long long)__builtin_amdgcn_s_memrealtime();
          // relaxed agent-scope loads see the other XCDs' released counts
          // without invalidating this XCD's L2 on every poll; the acquire
          // fence after the barrier below orders the phase's data.  The
          // polls bypass L2, so they are spaced to leave the counters'
          // memory channels to the blocks still working.
          for (int q = 0; q < kBatchQueues && !s_fail; ++q) {
            const long long c = (long long)(ph - 1) * kBatchQueues + q;
            const unsigned long long need = Q.phase_items[c];
            while (__hip_atomic_load(&Q.done[c * kHeadStride], __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT) < need) {
              __builtin_amdgcn_s_sleep(64);
              if ((long long)__builtin_amdgcn_s_memrealtime() - t0 > Q.wait_ticks) {
                atomicOr(Q.err, 1u);
                s_fail = 1;
                break;
              }
            }
          }
        }
      }
      __syncthreads();
      if (ph < Q.nphase) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      fail = __builtin_amdgcn_readfirstlane(s_fail);
      phase = ph;
      held = 0;
    }
    if (ph >= Q.nphase) break;
    // the position after next: dequeued by the first tile's call
    unsigned long long pos_nn = (unsigned long long)len;
    bool fetched = pos_n >= len;
    if (!fail) {
      const BatchGroup &g = Q.groups[__builtin_amdgcn_readfirstlane(e.group)];
      const long long nt = uniform_ll(g.ntiles);
      const long long t0 = uniform_ll((long long)e.local * g.per_item);
      const long long t1 = t0 + g.per_item < nt ? t0 + g.per_item : nt;
      for (long long tile = t0; tile < t1; ++tile) {
        if (tile > t0) __syncthreads();  // LDS reuse across tiles
        const unsigned long long r = run_tile(g, tile, fetched ? nullptr : head);
        if (!fetched) pos_nn = r;
        fetched = true;
      }
    }
    if (!fetched && tid == 0) pos_nn = atomicAdd(head, 1ull);
    __syncthreads();  // every wave has finished the tile's LDS use
    if (tid == 0) s_pos[0] = (long long)pos_nn;
    __syncthreads();
    pos = pos_n;
    pos_n = uniform_ll(s_pos[0]);
    e = e_n;
    ++held;
  }
}

}  // namespace d2dhip
