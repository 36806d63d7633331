// Inputs of the reverse score pass of align_canonical (src/align.rs:847-866),
// which scores the reversed prefixes rows[..=qend] and target[..=tend] of
// each hit to find its start.
//
//  * mirror_prof_kernel: each query's profile (its int8 or int16 profile
//    where it has one, else its int32 profile) with its rows in reverse order, so that the reversed prefix rows[..=e] of any end row e is the
//    window starting at row m - 1 - e (no profile per hit).  Rows past m are
//    zero, and at least 1024 of them follow the last row, as every pass of
//    the DP kernels reads whole passes of up to 1024 rows.  One launch
//    writes every query's mirror.
//  * rev_prefix_kernel: each hit's reversed target prefix, in the arena
//    layout of a database (16-byte aligned starts, code 21 padding, 64 bytes
//    of code 21 after the last target).
#include <hip/hip_runtime.h>

#include "dp_launch.h"

namespace st {

// every query's mirror in one launch: one block per (256-row chunk of some
// query's mirror, code); chunk_job[x] names the query of chunk x and
// chunk_first[j] its first chunk: dst[c][r] = src[c][m - 1 - r]
__global__ __launch_bounds__(256) void mirror_prof_kernel(const MirrorJob *__restrict__ jobs,
                                                          const uint32_t *__restrict__ chunk_job,
                                                          const uint32_t *__restrict__ chunk_first,
                                                          uint8_t *__restrict__ dst) {
  const uint32_t x = blockIdx.x, c = blockIdx.y;
  const uint32_t j = chunk_job[x];
  const MirrorJob jb = jobs[j];
  const uint32_t r = (x - chunk_first[j]) * 256 + threadIdx.x;
  if (r >= jb.dst_ld) return;
  const uint64_t o = (uint64_t)c * jb.dst_ld + r, i = (uint64_t)c * jb.src_ld + (jb.m - 1 - r);
  if (jb.esize == 1)
    ((int8_t *)(dst + jb.dst_off))[o] = r < jb.m ? ((const int8_t *)jb.src)[i] : (int8_t)0;
  else if (jb.esize == 2)
    ((int16_t *)(dst + jb.dst_off))[o] = r < jb.m ? ((const int16_t *)jb.src)[i] : (int16_t)0;
  else
    ((int32_t *)(dst + jb.dst_off))[o] = r < jb.m ? ((const int32_t *)jb.src)[i] : 0;
}

// one block per reversed prefix k: bytes [doff[k], doff[k + 1]) (and the
// 64-byte tail after the last one)
__global__ __launch_bounds__(256) void rev_prefix_kernel(const uint8_t *__restrict__ arena,
                                                         const uint64_t *__restrict__ toff,
                                                         const RevPrefix *__restrict__ pre, uint32_t n,
                                                         const uint64_t *__restrict__ doff,
                                                         uint8_t *__restrict__ dst) {
  const uint32_t k = blockIdx.x;
  const RevPrefix p = pre[k];
  const uint64_t o = doff[k];
  const uint64_t end = doff[k + 1] + (k + 1 == n ? 64 : 0);
  const uint8_t *t = arena + toff[p.tid];
  for (uint64_t x = threadIdx.x; o + x < end; x += 256)
    dst[o + x] = x < p.len ? t[p.len - 1 - x] : (uint8_t)kSentinel;
}

hipError_t launch_mirror_prof(const MirrorJob *jobs, const uint32_t *chunk_job, const uint32_t *chunk_first,
                              uint32_t nchunks, uint8_t *dst, hipStream_t s) {
  if (nchunks == 0) return hipSuccess;
  hipLaunchKernelGGL(mirror_prof_kernel, dim3(nchunks, kCodes), dim3(256), 0, s, jobs, chunk_job, chunk_first, dst);
  return hipGetLastError();
}

hipError_t launch_rev_prefix(const uint8_t *arena, const uint64_t *toff, const RevPrefix *pre, uint32_t n,
                             const uint64_t *doff, uint8_t *dst, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(rev_prefix_kernel, dim3(n), dim3(256), 0, s, arena, toff, pre, n, doff, dst);
  return hipGetLastError();
}

}  // namespace st
