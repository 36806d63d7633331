// The serving kernel: one persistent launch whose waves pull work units
// from a queue that the host keeps appending to, so new requests' units are
// merged into the running kernel instead of waiting for a launch.
//
// Queue protocol.  The host writes a request's profiles and pair slots into
// the serving arena (device memory, addresses not used before in this launch,
// by a copy that completed before publication), its unit descriptors into the
// host-coherent ring, and publishes them by raising the host-coherent `tail`.
//  * Block 0 is the poller: it alone reads the host tail and stop flag over
//    the bus (about once a microsecond) and writes the tail into the
//    kSeenCopies device words the other waves read.
//  * A worker takes a ticket (the next unit index) with a fetch-and-add on
//    `head`, which always succeeds, so every wave on every XCD gets units in
//    turn; it scores unit `ticket` once the published tail it reads exceeds
//    the ticket, then increments `done`.  Units [0, done) are exactly the
//    ones finished when a launch ends; the host sets head = done before the
//    next launch.  Device control words are read and written only with
//    read-modify-write atomics, which are performed at the coherence point:
//    a plain load, even an atomic one, may be served from the L2 of the
//    reader's XCD, which keeps a copy until an acquire invalidates it.  The
//    reads add a zero taken from the kernel arguments, so the compiler cannot
//    turn them back into loads.
//  * The claimed unit's descriptor is read with system-scope loads (coherent
//    with the host's writes, which preceded the tail), the unit is scored
//    (the batch kernels' DP pass: SWG + all-diagonal Kadane, packed int16),
//    each pair's gate runs from its group's merged keys (as in the DP kernels)
//    and the unit's result rows (consecutive, 16 bytes each) are written to
//    host-coherent memory with system-scope (write-through) stores arranged
//    so that each store instruction covers 256 contiguous bytes; once those
//    stores have completed (s_waitcnt) the request's remaining-unit counter
//    is decremented.  No cache-wide writeback or invalidation is needed.
//  * The poller ends the launch when the host sets `stop`, or when nothing
//    was published or finished for idle_ticks and every published unit is
//    done: it stamps this launch's epoch and its last tail T into the seen
//    words; every worker finishes the unit it holds (all below T) and exits.
//    No unit at or above T is started in this launch, so done = T at its end
//    and units published after the stamp are scored by the next launch, which
//    the host starts when it sees this one end.
#include <hip/hip_runtime.h>

#include "dp_core.h"
#include "gate_core.h"
#include "serve_args.h"

namespace st {



// uniform 64-bit value of lane 0
__device__ __forceinline__ uint64_t bcast64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// current value of a device control word (read-modify-write: coherent);
// zero is 0 at run time
__device__ __forceinline__ uint64_t rmw_read(uint64_t *p, uint64_t zero) {
  return __hip_atomic_fetch_add(p, zero, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void rmw_write(uint64_t *p, uint64_t v) {
  (void)__hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// [p, p + bytes) inside the serving arena
__device__ __forceinline__ bool in_arena(const ServeArgs &a, const void *p, uint64_t bytes) {
  const uint64_t o = (uint64_t)((const uint8_t *)p - a.sarena);
  return p && (const uint8_t *)p >= a.sarena && o <= a.sarena_bytes && bytes <= a.sarena_bytes - o;
}

// Block 0: mirror the host tail into the seen words, end the launch.
__device__ void serve_poller(const ServeArgs &a, int lane) {
  uint64_t *const done = a.ctl + kCtlDone * kCtlStride;
  uint64_t *const last_work = a.ctl + kCtlLastWork * kCtlStride;
  uint64_t *const mine = a.ctl + (uint64_t)(kCtlSeen + lane) * kCtlStride;  // lane k writes copy k
  const uint64_t quit_mark = (uint64_t)(a.epoch & 0xFFFF) << 48;
  uint64_t last_t = ~0ull, changed = wall_clock64();
  for (;;) {
    uint64_t t = 0;
    uint32_t stop = 0;
    if (lane == 0) {
      t = __hip_atomic_load(a.tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) & kSeenTailMask;
      stop = __hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    t = bcast64(t);
    stop = (uint32_t)__builtin_amdgcn_readfirstlane((int)stop);
    const uint64_t now = wall_clock64();
    // the host stamps each request's units with this clock when publishing
    if (lane == 0) __hip_atomic_store(a.clock, now, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    if (t != last_t) {
      if (lane < kSeenCopies) rmw_write(mine, t);
      last_t = t;
      changed = now;
    }
    int quit = 0;
    if (lane == 0) {
      const uint64_t lw = rmw_read(last_work, a.zero);
      const uint64_t since = lw > changed ? lw : changed;
      quit = stop != 0 || (now > since && now - since > a.idle_ticks && rmw_read(done, a.zero) >= t);
    }
    if (__builtin_amdgcn_readfirstlane(quit)) {
      if (lane < kSeenCopies) rmw_write(mine, quit_mark | t);
      return;
    }
    __builtin_amdgcn_s_sleep(20);
  }
}

template <int R>
__global__ __launch_bounds__(64, 3) void serve_kernel(ServeArgs a) {
  using P = Pk16;
  using T = P::T;
  using K = P::K;
  constexpr int kPer = P::kPer;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  int16_t *const lds = (int16_t *)lds_raw;
  const int lane = threadIdx.x;
  if (blockIdx.x == 0) {
    serve_poller(a, lane);
    return;
  }
  uint4 *const myb = a.passbuf + (uint64_t)blockIdx.x * (2 * a.passbuf_half + kEdgeQuadsPerBlock);
  uint64_t *const head = a.ctl + kCtlHead * kCtlStride;
  uint64_t *const done = a.ctl + kCtlDone * kCtlStride;
  uint64_t *const last_work = a.ctl + kCtlLastWork * kCtlStride;
  uint64_t *const seen_w = a.ctl + (uint64_t)(kCtlSeen + blockIdx.x % kSeenCopies) * kCtlStride;
  const uint64_t quit_mark = a.epoch & 0xFFFF;
  int backoff = 0;
  uint64_t ticket = ~0ull;  // lane 0: the unit this wave holds

  for (;;) {
    // take a ticket if this wave holds none; run it once it is published
    uint64_t ui = ~0ull;
    int quit = 0;
    if (lane == 0) {
      if (ticket == ~0ull) ticket = __hip_atomic_fetch_add(head, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint64_t sw = rmw_read(seen_w, a.zero);
      if ((sw & kSeenTailMask) > ticket) {
        ui = ticket;
        ticket = ~0ull;
      } else if ((sw >> 48) == quit_mark) {
        quit = 1;
      }
    }
    ui = bcast64(ui);
    if (ui == ~0ull) {
      if (__builtin_amdgcn_readfirstlane(quit)) break;
      // back off: 1, 2, 4, 8 sleeps of 127 x 64 clocks
      for (int k = 0; k < (1 << backoff); k++) __builtin_amdgcn_s_sleep(127);
      backoff = backoff < 3 ? backoff + 1 : 3;
      continue;
    }
    backoff = 0;
    // descriptor: lanes 0..15 load one word each (system scope: coherent
    // with the host's writes, which preceded the tail this wave saw)
    const uint32_t *dw = (const uint32_t *)(a.ring + (ui & a.ring_mask));
    uint32_t w = 0;
    if (lane < (int)(sizeof(ServeUnit) / 4)) w = __hip_atomic_load(dw + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    ServeUnit u;
    uint32_t *uw = (uint32_t *)&u;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(ServeUnit) / 4); k++) uw[k] = (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)w, k));
    // a malformed descriptor is reported, never followed
    const bool ok = u.G >= 1 && u.G <= 64 && (u.G & (u.G - 1)) == 0 && u.G * R <= 256 && u.passes >= 1 &&
                    (uint64_t)u.passes * u.G * R <= u.ld && u.m <= u.ld && u.npairs >= 1 &&
                    u.npairs <= 2 * (64 / u.G) && u.req < a.req_slots &&
                    in_arena(a, u.prof16, (uint64_t)kCodes * u.ld * 2) &&
                    in_arena(a, u.prof32, (uint64_t)kCodes * u.ld * 4) &&
                    in_arena(a, u.pairs, (uint64_t)u.npairs * sizeof(PairDev)) &&
                    (u.passes == 1 || (uint64_t)(64 / u.G) * u.n_pad <= a.passbuf_half);
    if (!ok) {
      if (lane == 0) {
        __hip_atomic_store(a.error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_fetch_add(done, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      continue;
    }
    QueryDev q;
    q.prof16 = u.prof16;
    q.prof32 = u.prof32;
    q.m = u.m;
    q.ld = u.ld;
    q.prof8 = nullptr;

    UnitResult<P> r;
    if (a.prio_ticks[0]) {
      const AgeClock ac{u.t_pub, {a.prio_ticks[0], a.prio_ticks[1], a.prio_ticks[2]}};
      score_unit<P, R, true, false, true>(q, u.pairs, u.npairs, u.G, u.n_pad, u.passes, u.open, u.ext, a.arena,
                                          a.toff, a.tlen, lds, a.lds_stride, myb, a.passbuf_half, lane, r, &ac);
      __builtin_amdgcn_s_setprio(0);
    } else {
      score_unit<P, R, true, false>(q, u.pairs, u.npairs, u.G, u.n_pad, u.passes, u.open, u.ext, a.arena, a.toff,
                                    a.tlen, lds, a.lds_stride, myb, a.passbuf_half, lane, r);
    }

    // gate of each pair slot s: lane s mod 64; its group's merged values come
    // from the group's first lane (every lane of a group holds the merge)
    for (int s0 = 0; s0 < (int)u.npairs; s0 += 64) {
      const int s = s0 + lane;
      const int x = s % kPer;
      const int src = ((s / kPer) * (int)u.G) & 63;
      const T bs = (T)__shfl((int)r.best, src);
      K key[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const K k0 = (K)__shfl((int)r.top[0][k], src);
        const K k1 = (K)__shfl((int)r.top[1][k], src);
        key[k] = x ? k1 : k0;
      }
      const bool act = s < (int)u.npairs;
      const PairDev pd = act ? u.pairs[s] : PairDev{0, 0};
      const bool ok_row = act && pd.out < a.result_rows && pd.tid < a.n_targets;
      uint64_t k64[4];
#pragma unroll
      for (int k = 0; k < 4; k++) k64[k] = P::key64(key[k]);
      // int16 pass; a result at the clamp (32767) is rescored by the host
      int32_t val[4] = {P::half(bs, x), 0, 0, 0};
      // the profile region of LDS is free until the next unit's first pass
      gate_wave(ProfGlobal{u.prof32, u.ld}, ok_row, a.arena + (ok_row ? a.toff[pd.tid] : 0), (int)u.m,
                ok_row ? (int)a.tlen[pd.tid] : 0, k64, u.open, u.ext, lane, (GateLds *)lds, val[1], val[2], val[3]);
      const uint32_t orow = ok_row ? pd.out : ~0u;  // this slot's result row, ~0u: none
      // store instruction w writes words 64w .. 64w + 63 of these 64 rows:
      // word 4r + f (row r, field f) comes from lane r
      uint32_t *const words = (uint32_t *)a.results;
#pragma unroll
      for (int w = 0; w < 4; w++) {
        const int src = w * 16 + (lane >> 2), f = lane & 3;
        const uint32_t o = (uint32_t)__shfl((int)orow, src);
        const int32_t v0 = __shfl(val[0], src), v1 = __shfl(val[1], src), v2 = __shfl(val[2], src),
                      v3 = __shfl(val[3], src);
        const int32_t v = f == 0 ? v0 : f == 1 ? v1 : f == 2 ? v2 : v3;
        if (o != ~0u) __hip_atomic_store(words + (uint64_t)o * 4 + f, (uint32_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
    // the write-through stores above have completed before the counter moves
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (lane == 0) {
      __hip_atomic_fetch_max(last_work, (uint64_t)wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(done, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(a.remaining + u.req, (uint32_t)-1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// Resident blocks per CU of the serving kernel with `lds_bytes` of dynamic LDS.
int serve_occupancy(size_t lds_bytes) {
  int n = 0;
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, serve_kernel<8>, 64, lds_bytes);
  return n;
}

hipError_t launch_serve(const ServeArgs &a, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL(serve_kernel<8>, dim3(grid), dim3(64), (size_t)kCodes * a.lds_stride * sizeof(int16_t), s, a);
  return hipGetLastError();
}

}  // namespace st
