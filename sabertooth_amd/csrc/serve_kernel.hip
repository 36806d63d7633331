// The serving kernel: one persistent launch whose waves pull work units
// from a queue that the host keeps appending to, so new requests' units are
// merged into the running kernel instead of waiting for a launch.
//
// Queue protocol (all queue memory is host-coherent pinned memory; the only
// device-memory word is the claim counter):
//  * the host writes ServeUnit descriptors and pair slots, then publishes
//    them by raising `tail` (a release store);
//  * a wave claims unit h by a compare-and-swap of `head` from h to h + 1 while
//    h < tail (acquire load), so a unit is never claimed before it is
//    published and never claimed twice;
//  * the wave scores the unit (the batch kernels' DP pass: SWG + all-diagonal
//    Kadane, packed int16), runs the gate of each of the unit's pairs from the
//    group's merged keys (the gate_pairs rule), writes the pairs' result rows
//    into the host-coherent result pool and decrements its request's
//    remaining-unit counter (system-scope release);
//  * with no unit to claim a wave sleeps briefly and exits once `stop` is set
//    or it has been idle for `idle_ticks` of the 100 MHz wall clock, so every
//    wave ends; the host relaunches the kernel when work arrives after that.
// Query profiles live in a device arena whose addresses are fresh for the
// lifetime of a launch (written by copies that completed before the units
// were published), so no wave can hold a stale cached line of them.
#include <hip/hip_runtime.h>

#include "dp_core.h"
#include "gate_core.h"

namespace st {

struct ServeArgs {
  const ServeUnit *ring;   // host-coherent descriptors, ring_mask + 1 slots
  uint64_t ring_mask;
  const uint64_t *tail;    // host-coherent: units published
  const uint32_t *stop;    // host-coherent: exit when idle
  uint64_t *head;          // device: units claimed
  const PairDev *pairs;    // host-coherent pair pool (tid, out = result row)
  PairResultDev *results;  // host-coherent result pool
  uint32_t *remaining;     // host-coherent: units left per request slot
  uint32_t *error;         // host-coherent: set when a descriptor is malformed
  uint64_t pair_rows;      // pair / result pool rows
  uint32_t req_slots;      // completion counters
  const uint8_t *arena;
  const uint64_t *toff;
  const uint32_t *tlen;
  uint4 *passbuf;
  uint64_t passbuf_half;
  uint64_t idle_ticks;
  int lds_stride;
};

// uniform 64-bit value of lane 0
__device__ __forceinline__ uint64_t bcast64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

template <int R>
__global__ __launch_bounds__(64) void serve_kernel(ServeArgs a) {
  using P = Pk16;
  using T = P::T;
  using K = P::K;
  constexpr int kPer = P::kPer;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  int16_t *const lds = (int16_t *)lds_raw;
  const int lane = threadIdx.x;
  uint4 *const myb = a.passbuf + (uint64_t)blockIdx.x * (2 * a.passbuf_half + kEdgeQuadsPerBlock);
  uint64_t idle_from = wall_clock64();

  for (;;) {
    // claim the next published unit (lane 0), or learn that there is none
    uint64_t ui = ~0ull;
    if (lane == 0) {
      for (int tries = 0; tries < 64; tries++) {
        const uint64_t h = __hip_atomic_load(a.head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t t = __hip_atomic_load(a.tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
        if (h >= t) break;
        uint64_t expect = h;
        if (__hip_atomic_compare_exchange_strong(a.head, &expect, h + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
          ui = h;
          break;
        }
      }
    }
    ui = bcast64(ui);
    if (ui == ~0ull) {
      int quit = 0;
      if (lane == 0)
        quit = __hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0 ||
               wall_clock64() - idle_from > a.idle_ticks;
      if (__builtin_amdgcn_readfirstlane(quit)) break;
      __builtin_amdgcn_s_sleep(32);
      continue;
    }
    // the descriptor and pair slots were written before `tail` was raised
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
    // descriptor: lanes 0..15 load one word each (vector loads), broadcast
    const uint32_t *dw = (const uint32_t *)(a.ring + (ui & a.ring_mask));
    uint32_t w = 0;
    if (lane < (int)(sizeof(ServeUnit) / 4)) w = __hip_atomic_load(dw + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    ServeUnit u;
    uint32_t *uw = (uint32_t *)&u;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(ServeUnit) / 4); k++) uw[k] = (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)w, k));
    // a malformed descriptor is reported, never followed
    const bool ok = u.prof16 && u.prof32 && u.G >= 1 && u.G <= 64 && (u.G & (u.G - 1)) == 0 && u.G * R <= 256 &&
                    u.passes >= 1 && (uint64_t)u.passes * u.G * R <= u.ld && u.npairs >= 1 &&
                    u.npairs <= 2 * (64 / u.G) && (uint64_t)u.pair0 + u.npairs <= a.pair_rows && u.req < a.req_slots &&
                    (u.passes == 1 || (uint64_t)(64 / u.G) * u.n_pad <= a.passbuf_half);
    if (!ok) {
      if (lane == 0) __hip_atomic_store(a.error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      idle_from = wall_clock64();
      continue;
    }
    QueryDev q;
    q.prof16 = u.prof16;
    q.prof32 = u.prof32;
    q.m = u.m;
    q.ld = u.ld;

    UnitResult<P> r;
    score_unit<P, R, true, false>(q, a.pairs + u.pair0, u.npairs, u.G, u.n_pad, u.passes, u.open, u.ext, a.arena,
                                  a.toff, a.tlen, lds, a.lds_stride, myb, a.passbuf_half, lane, r);

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
      const PairDev pd = s < (int)u.npairs ? a.pairs[u.pair0 + s] : PairDev{0, 0};
      if (s < (int)u.npairs && pd.out < a.pair_rows) {
        const int m = (int)u.m;
        const int n = (int)a.tlen[pd.tid];
        int32_t ts[4], td[4], lo[4], hi[4];
        int nt = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const uint64_t k64 = P::key64(key[k]);
          if (k64 == 0 || nt < k) continue;
          ts[nt] = (int32_t)(k64 >> 32);
          td[nt] = (int32_t)(0xFFFFFFFFu - (uint32_t)k64) - (m - 1);
          // whole diagonal (diag_range, src/ungapped.rs:38-46)
          const int d = td[nt];
          const int a0 = -d > 0 ? -d : 0;
          const int b0 = m < n - d ? m : n - d;
          lo[nt] = a0;
          hi[nt] = b0 > a0 ? b0 : a0;
          nt++;
        }
        const uint8_t *t = a.arena + a.toff[pd.tid];
        const GateRes g = gate_eval<ProfGlobal, 4>(ProfGlobal{u.prof32, u.ld}, t, ts, td, lo, hi, nt, u.open,
                                                   u.ext, INT32_MIN, false);
        PairResultDev o;
        o.swg = P::half(bs, x);
        o.ung = nt ? ts[0] : 0;
        o.best_diag = nt ? td[0] : 0;
        o.gate_score = nt ? g.score : 0;
        o.pass = (nt ? g.score : 0) >= u.gate;
        o.kernel = 0;  // int16 pass; a result at the clamp (32767) is rescored by the host
        o.r0 = o.r1 = 0;
        a.results[pd.out] = o;
      }
    }
    // results visible system-wide before the request's counter moves
    __threadfence_system();
    if (lane == 0) __hip_atomic_fetch_add(a.remaining + u.req, (uint32_t)-1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    idle_from = wall_clock64();
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
