// sort.hpp -- stable LSD radix sort of (u32 key, u32 value) pairs, 8 key
// bits per pass, deterministic.
//
// Each pass over digit d = (key >> shift) & 255:
//   1. k_digit_hist   tile b of kTile consecutive pairs counts its digits
//                     into hist[d * tiles + b] (digit-major, so that one
//                     exclusive scan over the whole array gives every
//                     (digit, tile) its first output slot);
//   2. scan.hpp's three-phase exclusive scan of hist;
//   3. k_digit_scatter tile b walks its pairs in index order, 256 at a time;
//                     a pair's slot is its (digit, tile) base plus the pairs
//                     of the same digit before it in the tile: lanes of one
//                     wavefront that share the digit are found with eight
//                     ballots, the counts of earlier wavefronts and rounds
//                     come from LDS.
// Equal keys keep their input order in every pass, so after the passes over
// the key's bits the pairs are ordered by key and, within a key, by input
// position.  No atomics decide an output slot, so the result is the same
// from run to run.
#pragma once
#include <hip/hip_runtime.h>

#include "scan.hpp"

namespace gng {
namespace radix {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;
constexpr int kBins = 256;
static_assert(kThreads == kBins, "one digit per thread in the tile bookkeeping");

__global__ __launch_bounds__(kThreads) void k_digit_hist(const unsigned* __restrict__ keys, long long n, int shift,
                                                         unsigned* __restrict__ hist, int tiles) {
  __shared__ unsigned h[kBins];
  const int t = threadIdx.x;
  h[t] = 0u;
  __syncthreads();
  const long long base = (long long)blockIdx.x * kTile;
  unsigned v[kItems];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const long long i = base + k * kThreads + t;
    v[k] = i < n ? keys[i] : 0u;
  }
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const long long i = base + k * kThreads + t;
    if (i < n) atomicAdd(&h[(v[k] >> shift) & (kBins - 1)], 1u);
  }
  __syncthreads();
  hist[(long long)t * tiles + blockIdx.x] = h[t];
}

// vals_in == nullptr stands for the identity (value = input index);
// keys_out == nullptr skips writing the keys (the last pass, when only the
// permutation is wanted).
__global__ __launch_bounds__(kThreads) void k_digit_scatter(const unsigned* __restrict__ keys_in,
                                                            const unsigned* __restrict__ vals_in, long long n,
                                                            int shift, const unsigned* __restrict__ slot,
                                                            int tiles, unsigned* __restrict__ keys_out,
                                                            unsigned* __restrict__ vals_out) {
  __shared__ unsigned run[kBins];                   // next slot of each digit in this tile
  __shared__ unsigned wcnt[kThreads / 64][kBins];   // this round's count per wavefront and digit
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  run[t] = slot[(long long)t * tiles + blockIdx.x];
#pragma unroll
  for (int w = 0; w < kThreads / 64; ++w) wcnt[w][t] = 0u;
  const long long base = (long long)blockIdx.x * kTile;
  unsigned key[kItems], val[kItems];
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const long long i = base + k * kThreads + t;
    key[k] = i < n ? keys_in[i] : 0u;
    val[k] = i < n ? (vals_in ? vals_in[i] : (unsigned)i) : 0u;
  }
  const unsigned long long below = (1ull << lane) - 1ull;
  __syncthreads();
#pragma unroll 1
  for (int k = 0; k < kItems; ++k) {
    const long long i = base + k * kThreads + t;
    const bool live = i < n;
    const unsigned d = (key[k] >> shift) & (kBins - 1);
    unsigned long long same = __ballot(live);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const unsigned long long ones = __ballot((d >> b) & 1u);
      same &= ((d >> b) & 1u) ? ones : ~ones;
    }
    const unsigned rank = (unsigned)__popcll(same & below);
    if (live && rank == 0) wcnt[wave][d] = (unsigned)__popcll(same);
    __syncthreads();
    unsigned pos = 0;
    if (live) {
      pos = run[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
    }
    __syncthreads();
    {
      unsigned add = 0;
#pragma unroll
      for (int w = 0; w < kThreads / 64; ++w) {
        add += wcnt[w][t];
        wcnt[w][t] = 0u;
      }
      run[t] += add;
    }
    if (live) {
      if (keys_out) keys_out[pos] = key[k];
      vals_out[pos] = val[k];
    }
    __syncthreads();
  }
}

// Work space of sort_pairs for n pairs, in bytes.
inline size_t sort_workspace(long long n) {
  const long long tiles = (n + kTile - 1) / kTile;
  const long long nh = tiles * kBins;
  const long long parts = (nh + kScanBlock - 1) / kScanBlock;
  auto up = [](long long b) { return (size_t)((b + 255) / 256 * 256); };
  return up(4 * (nh + 1)) * 2 + up(4 * parts) + up(4 * n) * 2;
}

// Scan kernels of the sort's digit counts.
__global__ __launch_bounds__(256) void k_sort_scan_local(const unsigned* __restrict__ count,
                                                         unsigned* __restrict__ offset,
                                                         unsigned* __restrict__ partial, long long nb) {
  scan_local_body(count, offset, partial, nb);
}
__global__ __launch_bounds__(256) void k_sort_scan_partials(unsigned* __restrict__ partial, int nblocks,
                                                            unsigned* __restrict__ offset, long long nb) {
  scan_partials_body(partial, nblocks, offset, nb);
}
__global__ __launch_bounds__(256) void k_sort_scan_add(unsigned* __restrict__ offset,
                                                       const unsigned* __restrict__ partial, long long nb) {
  scan_add_body(offset, partial, nb);
}

// Stable sort of keys[0..n) (only bits below end_bit are looked at) with
// the input index as the value: perm[j] = index of the j-th pair in key
// order; keys_sorted (may be null) receives the keys in that order.  n must
// be below 2^31; work is sort_workspace(n) bytes of device memory.
inline hipError_t sort_pairs(const unsigned* keys, long long n, int end_bit, unsigned* keys_sorted, unsigned* perm,
                             void* work, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const long long tiles = (n + kTile - 1) / kTile;
  const long long nh = tiles * kBins;
  const int parts = (int)((nh + kScanBlock - 1) / kScanBlock);
  auto up = [](long long b) { return (size_t)((b + 255) / 256 * 256); };
  char* w = static_cast<char*>(work);
  unsigned* hist = reinterpret_cast<unsigned*>(w);
  unsigned* slot = reinterpret_cast<unsigned*>(w + up(4 * (nh + 1)));
  unsigned* part = reinterpret_cast<unsigned*>(w + 2 * up(4 * (nh + 1)));
  unsigned* tk = reinterpret_cast<unsigned*>(w + 2 * up(4 * (nh + 1)) + up(4 * (long long)parts));
  unsigned* tv = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(tk) + up(4 * n));
  const int passes = end_bit <= 8 ? 1 : (end_bit + 7) / 8;
  const unsigned* kin = keys;
  const unsigned* vin = nullptr;
  for (int p = 0; p < passes; ++p) {
    const bool last = p == passes - 1;
    // the last pass lands in the caller's buffers; earlier ones alternate so
    // that a pass never reads the buffer it writes
    const bool to_final = ((passes - 1 - p) & 1) == 0;
    unsigned* kout = to_final ? keys_sorted : tk;
    unsigned* vout = to_final ? perm : tv;
    if (!last && to_final && !keys_sorted) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_digit_hist, dim3((unsigned)tiles), dim3(kThreads), 0, s, kin, n, 8 * p, hist, (int)tiles);
    hipLaunchKernelGGL(k_sort_scan_local, dim3((unsigned)parts), dim3(256), 0, s, hist, slot, part, nh);
    hipLaunchKernelGGL(k_sort_scan_partials, dim3(1), dim3(256), 0, s, part, parts, slot, nh);
    hipLaunchKernelGGL(k_sort_scan_add, dim3((unsigned)parts), dim3(256), 0, s, slot, part, nh);
    hipLaunchKernelGGL(k_digit_scatter, dim3((unsigned)tiles), dim3(kThreads), 0, s, kin, vin, n, 8 * p, slot,
                       (int)tiles, last ? keys_sorted : kout, vout);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    kin = kout;
    vin = vout;
  }
  return hipSuccess;
}

}  // namespace radix
}  // namespace gng
