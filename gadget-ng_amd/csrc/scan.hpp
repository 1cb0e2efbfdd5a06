// scan.hpp -- exclusive scan of unsigned counts by 256-thread workgroups.
//
// An array of nb counts is scanned in three launches: each workgroup scans
// kScanBlock entries and keeps its total (scan_local_body), one workgroup
// scans the totals (scan_partials_body, which also writes the grand total at
// offset[nb]), and each workgroup adds its base back (scan_add_body).
#pragma once
#include <hip/hip_runtime.h>

namespace gng {

constexpr int kScanItems = 16;
constexpr int kScanBlock = 256 * kScanItems;

__device__ __forceinline__ unsigned block_exclusive_scan(unsigned v, unsigned* lds, unsigned& total) {
  // 256 threads: wave scans then wave totals
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned s = v;
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned t = __shfl_up(s, off);
    if (lane >= off) s += t;
  }
  if (lane == 63) lds[wid] = s;
  __syncthreads();
  unsigned wbase = 0, tot = 0;
  for (int w = 0; w < 4; ++w) {
    const unsigned ws = lds[w];
    if (w < wid) wbase += ws;
    tot += ws;
  }
  __syncthreads();
  total = tot;
  return wbase + s - v;
}

__device__ __forceinline__ void scan_local_body(const unsigned* __restrict__ count, unsigned* __restrict__ offset,
                                                    unsigned* __restrict__ partial, long long nb) {
  __shared__ unsigned lds[4];
  const long long base = (long long)blockIdx.x * kScanBlock + (long long)threadIdx.x * kScanItems;
  unsigned v[kScanItems];
  unsigned sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < nb) ? count[base + k] : 0u;
    sum += v[k];
  }
  unsigned total;
  unsigned run = block_exclusive_scan(sum, lds, total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < nb) offset[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__device__ __forceinline__ void scan_partials_body(unsigned* __restrict__ partial, int nblocks,
                                                       unsigned* __restrict__ offset, long long nb) {
  __shared__ unsigned lds[4];
  unsigned carry = 0;
  for (int b0 = 0; b0 < nblocks; b0 += 256) {
    const int b = b0 + (int)threadIdx.x;
    const unsigned v = b < nblocks ? partial[b] : 0u;
    unsigned total;
    const unsigned ex = block_exclusive_scan(v, lds, total);
    if (b < nblocks) partial[b] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) offset[nb] = carry;
}

// scan_add_body without the partials pass: workgroup b sums the totals of
// the nblocks workgroups before it itself (a few hundred at most), and the
// last one also writes the grand total at offset[nb].  partial keeps the
// workgroup totals of scan_local_body.
__device__ __forceinline__ void scan_add_prefix_body(unsigned* __restrict__ offset,
                                                     const unsigned* __restrict__ partial, int nblocks,
                                                     long long nb) {
  __shared__ unsigned lds[4];
  unsigned s = 0;
  for (int j = threadIdx.x; j < (int)blockIdx.x; j += 256) s += partial[j];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
  __syncthreads();
  const unsigned add = (lds[0] + lds[1]) + (lds[2] + lds[3]);
  const long long base = (long long)blockIdx.x * kScanBlock;
  for (int k = threadIdx.x; k < kScanBlock; k += 256)
    if (base + k < nb) offset[base + k] += add;
  if ((int)blockIdx.x == nblocks - 1 && threadIdx.x == 0) offset[nb] = add + partial[blockIdx.x];
}

__device__ __forceinline__ void scan_add_body(unsigned* __restrict__ offset, const unsigned* __restrict__ partial,
                                                  long long nb) {
  const long long base = (long long)blockIdx.x * kScanBlock;
  const unsigned add = partial[blockIdx.x];
  for (int k = threadIdx.x; k < kScanBlock; k += 256)
    if (base + k < nb) offset[base + k] += add;
}

}  // namespace gng
