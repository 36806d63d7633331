// Arguments of the persistent serving kernel, shared by the kernel and the
// host runtime that launches it.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "st_device.h"

namespace st {

// Device control words, kCtlStride uint64 apart (4 KiB: separate memory
// channels): tickets taken (head), units finished (done), wall clock of the
// last unit finished, and kSeenCopies copies of the published tail written
// by the poller wave (a worker reads the copy of its block index mod
// kSeenCopies).  A seen word holds the tail in its low 48 bits and, in its
// high 16 bits, the epoch of the last launch told to exit.
enum ServeCtl { kCtlHead = 0, kCtlDone = 1, kCtlLastWork = 2, kCtlSeen = 3 };
constexpr int kSeenCopies = 64;
constexpr int kCtlWords = kCtlSeen + kSeenCopies;
constexpr int kCtlStride = 512;
constexpr uint64_t kSeenTailMask = (1ull << 48) - 1;

// A pair's result row in the serving result pool, in plan-slot order (a
// unit's rows are consecutive); the host derives the gate decision from
// gate_score and the request's gate.  16 bytes: 64 rows fill whole lines.
struct ServeRow {
  int32_t swg, ung, best_diag, gate_score;
};

struct ServeArgs {
  const ServeUnit *ring;   // host-coherent descriptors, ring_mask + 1 slots
  uint64_t ring_mask;
  const uint64_t *tail;    // host-coherent: units published
  const uint32_t *stop;    // host-coherent: exit now (after the current units)
  uint64_t *ctl;           // device control words
  ServeRow *results;       // host-coherent result pool
  uint32_t *remaining;     // host-coherent: units left per request slot
  uint32_t *error;         // host-coherent: set when a descriptor is malformed
  uint64_t result_rows;    // result pool rows
  uint32_t req_slots;      // completion counters
  const uint8_t *sarena;   // serving arena: profiles and pair slots
  uint64_t sarena_bytes;
  const uint8_t *arena;    // target residues
  const uint64_t *toff;
  const uint32_t *tlen;
  uint32_t n_targets;
  uint4 *passbuf;
  uint64_t passbuf_half;
  uint64_t idle_ticks;     // exit after this long without work (100 MHz ticks)
  uint64_t *clock;         // host-coherent: the device wall clock, written by the poller
  uint32_t prio_ticks[3];  // a unit's wave priority rises to k + 1 once its request is
                           // prio_ticks[k] ticks old (0: priorities off)
  uint32_t epoch;          // this launch's number, 1..65535
  uint64_t zero;           // 0: operand of the read-modify-write reads
  int lds_stride;
};

// Resident serving-kernel blocks per CU with `lds_bytes` of dynamic LDS.
int serve_occupancy(size_t lds_bytes);
hipError_t launch_serve(const ServeArgs &a, uint32_t grid, hipStream_t s);

}  // namespace st
