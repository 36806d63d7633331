// Query profiles of simulated pairs (simulate_scores, src/evalue.rs:537-565):
// every sample has its own random query, so its profile is built on the
// device from the query's residue codes rather than uploaded.  One block
// per (sample, code) writes that code's row of the sample's int16 and int32
// profiles: entry r = BLOSUM62[q[r]][code] + bias[r] for r < m, 0 for the
// padded rows and for code 21 (the layout st_qprof_create uploads).
#include <hip/hip_runtime.h>

#include "dp_launch.h"

namespace st {

__global__ __launch_bounds__(256) void sim_prof_kernel(const uint8_t *__restrict__ q, const int32_t *__restrict__ bias,
                                                       const int8_t *__restrict__ mat, uint32_t m, uint32_t ld,
                                                       int16_t *__restrict__ p16, int32_t *__restrict__ p32) {
  __shared__ int8_t col[kCodes * kCodes];
  const uint32_t s = blockIdx.x / kCodes, c = blockIdx.x % kCodes;
  for (uint32_t k = threadIdx.x; k < (uint32_t)(kCodes * kCodes); k += 256) col[k] = mat[k];
  __syncthreads();
  const uint8_t *qs = q + (uint64_t)s * m;
  const int32_t *bs = bias ? bias + (uint64_t)s * m : nullptr;
  const uint64_t o = ((uint64_t)s * kCodes + c) * ld;
  for (uint32_t r = threadIdx.x; r < ld; r += 256) {
    int32_t v = 0;
    if (r < m && c < kCodes - 1) {
      const uint32_t a = qs[r] < kCodes ? qs[r] : kCodes - 1;
      v = (int32_t)col[a * kCodes + c] + (bs ? bs[r] : 0);
    }
    p32[o + r] = v;
    p16[o + r] = (int16_t)v;
  }
}

hipError_t launch_sim_prof(const uint8_t *q, const int32_t *bias, const int8_t *mat, uint32_t m, uint32_t ld,
                           uint32_t samples, int16_t *p16, int32_t *p32, hipStream_t s) {
  if (samples == 0) return hipSuccess;
  hipLaunchKernelGGL(sim_prof_kernel, dim3(samples * kCodes), dim3(256), 0, s, q, bias, mat, m, ld, p16, p32);
  return hipGetLastError();
}

}  // namespace st
