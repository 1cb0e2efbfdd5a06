// deposit.hpp -- deterministic tiled CIC deposit for gfx950.
//
// Reference semantics: crates/gadget-ng-pm/src/cic.rs:37-93 (mass per cell,
// node-centred trilinear weights, periodic wrap of all 8 indices).
//
// Positions enter as FP32 or FP64 (any stride); every kernel forms the cell
// coordinate c = (x * nm / L) rem_euclid nm in FP64, as cic.rs:50-52 does,
// so FP64 inputs keep their accuracy (cell_fix in pm_kernels.hpp).
//
// Method:
//   1. k_bucket_count  each particle's base cell (i0,j0,k0) gives a tile of
//      T^3 cells and a class: bit a is set when the base cell is on the tile's
//      last layer along axis a, i.e. when the particle also feeds the next
//      tile along that axis.  Bucket = Hilbert rank of the tile * 8 + class.  Counts use one
//      atomic per run of equal buckets within a wavefront (particles that
//      arrive in spatial order share buckets in runs), which also yields each
//      particle's rank inside its bucket.  The max |m| for the fixed-point
//      scale is reduced in the same pass.
//   2. k_scan_*        exclusive scan of the bucket counts.
//   3. k_scatter       particles to bucket order as 24-byte cell records
//      (PartRec: per axis the cell index and a 43-bit in-cell fraction, plus
//      the FP32 mass), and the sum of the fixed-point masses.
//   4. k_deposit_tiles one workgroup per tile owns its T^3 cells outright.  It
//      reads its own particles and, from the 7 lower neighbour tiles, only the
//      classes that reach into it, and adds each of the 8 CIC targets exactly
//      once over all tiles (target offset a lands in tile t + (a & class)).
//      Weights are formed in FP64 from the fractions and become 64-bit fixed
//      point (w * m * 2^s, N max|m| 2^s < 2^62); they accumulate with LDS
//      integer atomics, so the sum is exact and order-independent.  The tile
//      is then written as FP32 mass per cell minus the mean mass per cell
//      (an integer subtracted before the conversion, so a nearly uniform
//      density keeps its contrast to FP32 precision; the constant only
//      touches the k = 0 mode, which the Green's function drops).  No global
//      atomics touch the mesh and no mesh clearing pass is needed.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pm_kernels.hpp"
#include "scan.hpp"

namespace gng {

// Tiles are numbered along a 3-D Hilbert curve: rank[t] is the curve
// position of the tile with linear index t = (tz*nt + ty)*nt + tx and
// order[] the inverse, so bucket order (and the sorted particle order) walks
// the box tile by face-adjacent tile, and the deposit launches its
// workgroups in the same order.
struct TileMap {
  int T;   // tile side in cells
  int nt;  // tiles per axis (nm / T)
  const unsigned* rank;
  const unsigned* order;
  // src[8r + d]: first bucket (rank * 8) of source tile d of the tile of rank
  // r, the tile u - d (bit 0 x, bit 1 y, bit 2 z, periodic), so a deposit
  // workgroup finds its 72 bucket offsets after one table load
  const unsigned* src;
  // per tile rank, kRangeWords words of the flattened layout of its 14
  // bucket ranges, filled by k_tile_ranges for each solve (null: the deposit
  // works it out from the offsets itself)
  unsigned* ranges;
};

// Fill tm.src from rank[] for the tiles t = t0, t0 + step, ...
__device__ __host__ inline void fill_tile_src(int nt, const unsigned* rank, unsigned* src, int t0, int step) {
  const int nt3 = nt * nt * nt;
  for (int t = t0; t < nt3; t += step) {
    const int ux = t % nt, uy = (t / nt) % nt, uz = t / (nt * nt);
    for (int d = 0; d < 8; ++d) {
      const int sx = (d & 1) ? (ux == 0 ? nt - 1 : ux - 1) : ux;
      const int sy = (d & 2) ? (uy == 0 ? nt - 1 : uy - 1) : uy;
      const int sz = (d & 4) ? (uz == 0 ? nt - 1 : uz - 1) : uz;
      src[(size_t)rank[t] * 8 + d] = rank[(sz * nt + sy) * nt + sx] * 8u;
    }
  }
}

// Zero a buffer with vector stores (a fill kernel rather than a memset, so the
// sequence stays safe to capture in a graph).
__device__ __forceinline__ void zero_body(uint4* __restrict__ p, long long n16) {
  const uint4 z = make_uint4(0u, 0u, 0u, 0u);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (long long)gridDim.x * blockDim.x)
    p[i] = z;
}

__global__ __launch_bounds__(256) void k_zero(uint4* __restrict__ p, long long n16) {
  zero_body(p, n16);
}

// The bucket counts cleared and, by the first workgroup, the solve's scalars
// reset: the two clears before a sort in one launch.
__global__ __launch_bounds__(256) void k_sort_reset(uint4* __restrict__ p, long long n16, SolveScalars* sc) {
  if (blockIdx.x == 0) reset_scalars_body(sc);
  zero_body(p, n16);
}

__global__ __launch_bounds__(256) void k_zero_box(const BoxArgs* __restrict__ bx, long long n16) {
  const BoxArgs& b = bx[blockIdx.z];
  zero_body(reinterpret_cast<uint4*>(b.count), n16);
}

__device__ __forceinline__ unsigned bucket_of_cell(int ix, int iy, int iz, const TileMap tm) {
  const int tx = ix / tm.T, ty = iy / tm.T, tz = iz / tm.T;
  const unsigned cls = (unsigned)(ix - tx * tm.T == tm.T - 1) | ((unsigned)(iy - ty * tm.T == tm.T - 1) << 1) |
                       ((unsigned)(iz - tz * tm.T == tm.T - 1) << 2);
  return tm.rank[((unsigned)tz * tm.nt + ty) * tm.nt + tx] * 8u + cls;
}

// Each thread classifies kBucketPer particles (block stride 256, so every
// wavefront still sees 64 consecutive particles).  Counting is two-level:
// runs of equal buckets within a wavefront take one slot of a block-local
// hash table in LDS and one LDS atomic; then each distinct bucket of the
// block takes one global atomic, which returns the block's base rank in that
// bucket.  A block of clustered particles thus costs a few global atomics
// instead of one per run.  Ranks within a bucket stay unique; their order
// differs from call to call, which no result depends on (the deposit sums
// integers).
constexpr int kBucketPer = 4;
constexpr int kBucketSlots = 2048;  // > 256 * kBucketPer keys: probing always ends
constexpr unsigned kEmptyKey = 0xffffffffu;

template <typename T, typename M>
__device__ __forceinline__ void bucket_count_body(const PosSrc<T> src, const M* __restrict__ m, long long n,
                                                      const CellMap cm, const TileMap tm,
                                                      unsigned* __restrict__ count, unsigned* __restrict__ bucket,
                                                      unsigned* __restrict__ rank, SolveScalars* sc) {
  __shared__ unsigned hkey[kBucketSlots], hcnt[kBucketSlots];
  for (int e = threadIdx.x; e < kBucketSlots; e += 256) {
    hkey[e] = kEmptyKey;
    hcnt[e] = 0u;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const long long i0 = (long long)blockIdx.x * (256 * kBucketPer) + threadIdx.x;
  unsigned b[kBucketPer], local[kBucketPer];
  int slot[kBucketPer];
  float am = 0.0f;
#pragma unroll
  for (int k = 0; k < kBucketPer; ++k) {
    const long long i = i0 + 256LL * k;
    const bool live = i < n;
    b[k] = kEmptyKey;
    if (live) {
      const long long o = i * src.stride;
      const double px = (double)src.x[o], py = (double)src.y[o], pz = (double)src.z[o];
      am = fmaxf(am, fabsf((float)m[i]));
      int ix, iy, iz;
      unsigned long long f;
      cell_fix(px, cm, ix, f);
      cell_fix(py, cm, iy, f);
      cell_fix(pz, cm, iz, f);
      b[k] = bucket_of_cell(ix, iy, iz, tm);
    }
    // runs of equal buckets: a lane heads a run when its bucket differs from the previous lane's
    const unsigned prev = __shfl_up(b[k], 1);
    const bool head = live && (lane == 0 || prev != b[k]);
    const unsigned long long heads = __ballot(head);
    const unsigned long long below = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1);
    const int my_head = 63 - __clzll(heads & below);  // highest head at or below this lane
    const unsigned long long above = heads & ~below;
    const unsigned long long live_mask = __ballot(live);
    const int next_head = above ? __ffsll((long long)above) - 1 : 64 - __clzll(live_mask);
    int hs = 0;
    unsigned hb = 0;
    if (head) {
      // insert or find the bucket in the block's table (linear probing)
      hs = (int)((b[k] * 2654435761u) >> 21);  // 11 bits: kBucketSlots = 2048
      for (;;) {
        const unsigned old = atomicCAS(&hkey[hs], kEmptyKey, b[k]);
        if (old == kEmptyKey || old == b[k]) break;
        hs = (hs + 1) & (kBucketSlots - 1);
      }
      hb = atomicAdd(&hcnt[hs], (unsigned)(next_head - lane));
    }
    const int src = my_head < 0 ? 0 : my_head;
    slot[k] = __shfl(hs, src);
    local[k] = __shfl(hb, src) + (unsigned)(lane - my_head);
  }
  __syncthreads();
  // one global atomic per distinct bucket of the block; hcnt becomes the base
#pragma unroll
  for (int e = threadIdx.x; e < kBucketSlots; e += 256) {
    const unsigned key = hkey[e];
    if (key != kEmptyKey) hcnt[e] = atomicAdd(count + key, hcnt[e]);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kBucketPer; ++k) {
    const long long i = i0 + 256LL * k;
    if (i < n) {
      bucket[i] = b[k];
      rank[i] = hcnt[slot[k]] + local[k];
    }
  }
  // max |m|: wave, then block, then one atomic per block only when it raises
  // the running maximum (a plain read filters the common equal-mass case)
  for (int off = 32; off > 0; off >>= 1) am = fmaxf(am, __shfl_xor(am, off));
  __shared__ float wmax[4];
  if (lane == 0) wmax[threadIdx.x >> 6] = am;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float bm = fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
    const unsigned bits = __float_as_uint(bm);
    if (bm > 0.0f && bits > __atomic_load_n(&sc->mmax_bits, __ATOMIC_RELAXED)) atomicMax(&sc->mmax_bits, bits);
  }
}

template <typename T, typename M>
__global__ __launch_bounds__(256) void k_bucket_count(const PosSrc<T> src, const M* __restrict__ m, long long n,
                                                      const CellMap cm, const TileMap tm,
                                                      unsigned* __restrict__ count, unsigned* __restrict__ bucket,
                                                      unsigned* __restrict__ rank, SolveScalars* sc) {
  bucket_count_body<T, M>(src, m, n, cm, tm, count, bucket, rank, sc);
}

__global__ __launch_bounds__(256) void k_bucket_count_box(const BoxArgs* __restrict__ bx, const PosSrc<float> src, const float* __restrict__ m, const TileMap tm) {
  const BoxArgs& b = bx[blockIdx.z];
  bucket_count_body<float, float>(PosSrc<float>{src.x + b.p0, src.y + b.p0, src.z + b.p0, 1}, m + b.p0, b.n, b.cm, tm, b.count, b.bucket, b.rank,
                                  b.sc);
}

// Exclusive scan of `count` (nb entries) into `offset` (nb + 1 entries),
// in the three phases of scan.hpp.
__global__ __launch_bounds__(256) void k_scan_local(const unsigned* __restrict__ count, unsigned* __restrict__ offset,
                                                    unsigned* __restrict__ partial, long long nb) {
  scan_local_body(count, offset, partial, nb);
}

__global__ __launch_bounds__(256) void k_scan_local_box(const BoxArgs* __restrict__ bx, long long nb) {
  const BoxArgs& b = bx[blockIdx.z];
  scan_local_body(b.count, b.offset, b.partial, nb);
}

__global__ __launch_bounds__(256) void k_scan_partials(unsigned* __restrict__ partial, int nblocks,
                                                       unsigned* __restrict__ offset, long long nb) {
  scan_partials_body(partial, nblocks, offset, nb);
}

__global__ __launch_bounds__(256) void k_scan_partials_box(const BoxArgs* __restrict__ bx, int nblocks, long long nb) {
  const BoxArgs& b = bx[blockIdx.z];
  scan_partials_body(b.partial, nblocks, b.offset, nb);
}

__global__ __launch_bounds__(256) void k_scan_add(unsigned* __restrict__ offset, const unsigned* __restrict__ partial,
                                                  long long nb) {
  scan_add_body(offset, partial, nb);
}

__global__ __launch_bounds__(256) void k_scan_add_prefix(unsigned* __restrict__ offset,
                                                         const unsigned* __restrict__ partial, int nblocks,
                                                         long long nb) {
  scan_add_prefix_body(offset, partial, nblocks, nb);
}

__global__ __launch_bounds__(256) void k_scan_add_box(const BoxArgs* __restrict__ bx, long long nb) {
  const BoxArgs& b = bx[blockIdx.z];
  scan_add_body(b.offset, b.partial, nb);
}

// Particle i to its bucket slot as a cell record, and the sum of the
// particles' fixed-point masses round(m 2^s) (one integer atomic per block,
// so the total does not depend on the order): the deposit subtracts the mean
// per cell before it converts to FP32.  Each thread moves kScatterPer
// particles (block stride 256, so a wavefront still reads 64 consecutive
// particles); all of their loads are issued before the dependent offset
// loads, and those before any store, so a thread waits for two memory
// latencies however many particles it moves.
constexpr int kScatterPer = 4;

template <typename T, typename M>
__device__ __forceinline__ void scatter_body(const PosSrc<T> src, const M* __restrict__ m, long long n,
                                                 const CellMap cm, const unsigned* __restrict__ bucket,
                                                 const unsigned* __restrict__ rank,
                                                 const unsigned* __restrict__ offset, PartRec* __restrict__ sorted,
                                                 unsigned* __restrict__ perm, SolveScalars* sc) {
  const long long i0 = (long long)blockIdx.x * (256 * kScatterPer) + threadIdx.x;
  double px[kScatterPer], py[kScatterPer], pz[kScatterPer];
  float mf[kScatterPer];
  unsigned slot[kScatterPer];
#pragma unroll
  for (int k = 0; k < kScatterPer; ++k) {
    const long long i = i0 + 256LL * k;
    if (i < n) {
      const long long o = i * src.stride;
      px[k] = (double)src.x[o];
      py[k] = (double)src.y[o];
      pz[k] = (double)src.z[o];
      mf[k] = (float)m[i];
      slot[k] = bucket[i];
    }
  }
#pragma unroll
  for (int k = 0; k < kScatterPer; ++k) {
    const long long i = i0 + 256LL * k;
    if (i < n) slot[k] = offset[slot[k]] + rank[i];
  }
  const int shift = fixed_shift(sc, n);
  long long q = 0;
#pragma unroll
  for (int k = 0; k < kScatterPer; ++k) {
    const long long i = i0 + 256LL * k;
    if (i < n) {
      int c[3];
      unsigned long long f[3];
      cell_fix(px[k], cm, c[0], f[0]);
      cell_fix(py[k], cm, c[1], f[1]);
      cell_fix(pz[k], cm, c[2], f[2]);
      sorted[slot[k]] = encode_rec(c, f, mf[k]);
      if (perm) perm[slot[k]] = (unsigned)i;  // only for the sorted-order interpolation
      q += rint_small(ldexp((double)mf[k], shift));
    }
  }
  for (int off = 32; off > 0; off >>= 1) q += __shfl_xor(q, off);
  __shared__ long long wsum[4];
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = q;
  __syncthreads();
  if (threadIdx.x == 0) {
    const long long t = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
    if (t) atomicAdd(&sc->qsum[(blockIdx.x % kQSlots) * kQStride], (unsigned long long)t);
  }
}

template <typename T, typename M>
__global__ __launch_bounds__(256) void k_scatter(const PosSrc<T> src, const M* __restrict__ m, long long n,
                                                 const CellMap cm, const unsigned* __restrict__ bucket,
                                                 const unsigned* __restrict__ rank,
                                                 const unsigned* __restrict__ offset, PartRec* __restrict__ sorted,
                                                 unsigned* __restrict__ perm, SolveScalars* sc) {
  scatter_body<T, M>(src, m, n, cm, bucket, rank, offset, sorted, perm, sc);
}

__global__ __launch_bounds__(256) void k_scatter_box(const BoxArgs* __restrict__ bx, const PosSrc<float> src, const float* __restrict__ m) {
  const BoxArgs& b = bx[blockIdx.z];
  scatter_body<float, float>(PosSrc<float>{src.x + b.p0, src.y + b.p0, src.z + b.p0, 1}, m + b.p0, b.n, b.cm, b.bucket, b.rank, b.offset, b.sorted,
                             nullptr, b.sc);
}

// LDS fixed-point add with wavefront aggregation: the lanes whose cell equals
// the first valid lane's cell sum their contributions in registers and one
// lane adds the total; the others add directly.  In a dense cluster most of a
// wavefront hits one cell, so this turns 64 serialised atomics into one; with
// scattered cells the ballot finds a single match and the reduction is skipped.
// Integer sums keep the result independent of the grouping.
__device__ __forceinline__ void lds_add_aggregated(unsigned long long* acc, int cell, long long q, bool valid) {
  const unsigned long long vmask = __ballot(valid);
  if (vmask == 0ull) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)vmask) - 1;
  const int lcell = __shfl(cell, leader);
  const bool mine = valid && cell == lcell;
  const unsigned long long mm = __ballot(mine);
  if (__popcll(mm) > 1) {
    long long v = mine ? q : 0ll;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == leader) atomicAdd(acc + lcell, (unsigned long long)v);
    valid = valid && !mine;
  }
  if (valid) atomicAdd(acc + cell, (unsigned long long)q);
}

// The 14 bucket ranges a tile reads: (d, first class, end class).  d is the
// neighbour offset (bit 0 x, bit 1 y, bit 2 z: the source tile is u - d) and
// the classes are those containing every bit of d, merged where adjacent.
// d of range r is (kRangeD >> 3r) & 7 (packed, so a per-lane r needs no load)
constexpr unsigned char kRangeDList[14] = {0, 1, 1, 1, 1, 2, 2, 3, 3, 4, 5, 5, 6, 7};
constexpr unsigned long long pack_range_d() {
  unsigned long long v = 0;
  for (int i = 0; i < 14; ++i) v |= (unsigned long long)kRangeDList[i] << (3 * i);
  return v;
}
constexpr unsigned long long kRangeD = pack_range_d();
__constant__ unsigned char kRangeLo[14] = {0, 1, 3, 5, 7, 2, 6, 3, 7, 4, 5, 7, 6, 7};
__constant__ unsigned char kRangeHi[14] = {8, 2, 4, 6, 8, 4, 8, 4, 8, 8, 6, 8, 8, 8};

// Heavy tiles over several workgroups.  Workgroup r of the deposit handles
// tile rank r.  A tile whose 14 ranges hold more than kSliceItems particles
// gets S = ceil(items / kSliceItems) slices (at most kMaxSlices): its own
// workgroup takes slice 0, and S - 1 extra workgroups past the ntiles
// primary ones take the others, each depositing an equal share of the
// ranges into its own LDS tile.  The slices add their int64 tiles into a
// scratch tile with global atomics (integer sums: the result does not depend
// on the order or on how a tile is split) and the last slice to finish
// converts the scratch tile and clears it for the next solve.  Clustered
// inputs otherwise leave a few Plummer-core tiles running hundreds of rounds
// on one workgroup each.
constexpr unsigned kSliceItems = 32768;
constexpr unsigned kMaxSlices = 1024;
constexpr unsigned kNoSlice = 0xffffffffu;

struct DepositSplit {
  const unsigned* slices;        // slices of each tile rank (null: no splitting)
  const unsigned* slot;          // scratch tile of each split tile rank
  const unsigned* extra;         // extra workgroup e -> rank | slice << 20, or kNoSlice
  const SolveScalars* sc;        // split_budget: extra workgroups reserved this solve
  unsigned ntiles, slots;
  unsigned long long* scratch;   // slots x T^3 partial tiles, zero between solves
  unsigned* done;                // slots finished-slice counters, zero between solves
};

// The flattened range layout of every tile rank r, written after the scatter
// so that a deposit workgroup reads its whole prologue with scalar loads:
// ranges[kRangeWords r + k] is, for k < 13, the end of range k when the 14
// ranges are laid end to end; word 13 is their total, word 14 the slot delta
// of range 0 (sorted slot = flattened index + delta) and word 15 + k the
// change of that delta from range k to k + 1.  Thread 0 also sums the
// scatter's fixed-point mass slots into sc->qtot.
constexpr int kRangeWords = 32;
__global__ __launch_bounds__(256) void k_tile_ranges(const unsigned* __restrict__ offset, const TileMap tm,
                                                     SolveScalars* sc) {
  const unsigned ntiles = (unsigned)tm.nt * tm.nt * tm.nt;
  const unsigned r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r == 0) {
    unsigned long long q = 0;
    for (int k = 0; k < kQSlots; ++k) q += sc->qsum[k * kQStride];
    sc->qtot = q;
  }
  if (r >= ntiles) return;
  unsigned w[kRangeWords];
  unsigned pre = 0, prev = 0;
#pragma unroll
  for (int k = 0; k < 14; ++k) {
    const int d = (int)((kRangeD >> (3 * k)) & 7ull);
    const unsigned b = tm.src[(size_t)r * 8 + d];
    const unsigned beg = offset[b + kRangeLo[k]];
    const unsigned len = offset[b + kRangeHi[k]] - beg;
    const unsigned delta = beg - pre;
    pre += len;
    if (k < 13) w[k] = pre;
    if (k == 0) w[14] = delta;
    else w[14 + k] = delta - prev;
    prev = delta;
  }
  w[13] = pre;
#pragma unroll
  for (int k = 28; k < kRangeWords; ++k) w[k] = 0u;
  uint4* dst = reinterpret_cast<uint4*>(tm.ranges + (size_t)r * kRangeWords);
#pragma unroll
  for (int k = 0; k < kRangeWords / 4; ++k) dst[k] = make_uint4(w[4 * k], w[4 * k + 1], w[4 * k + 2], w[4 * k + 3]);
}

// The 8 CIC targets of one cell record of source range r (d = the range's
// neighbour offset) into the LDS tile acc of T^3 int64 accumulators: FP64
// weights (the 43-bit fractions convert exactly, 1 - f is exact, and the
// products round far below the fixed-point unit) times m 2^s (scale), as
// 64-bit fixed point.  Every lane of the wavefront calls it (live marks the
// lanes with a record), so the ballots see whole waves.
template <int T>
__device__ __forceinline__ void deposit_record(unsigned long long* acc, const PartRec& p, int r, bool live,
                                               double scale) {
  const int lane = threadIdx.x & 63;
  const int d = (int)((kRangeD >> (3 * r)) & 7ull);
  const int dx = d & 1, dy = (d >> 1) & 1, dz = (d >> 2) & 1;
  const double fx = rec_frac(p, 0), fy = rec_frac(p, 1), fz = rec_frac(p, 2);
  const double ms = (double)rec_mass(p) * scale;
  const double wx[2] = {1.0 - fx, fx}, wy[2] = {1.0 - fy, fy};
  const double wz[2] = {(1.0 - fz) * ms, fz * ms};
  // local base index in the particle's own tile; its class follows from it
  const int lx0 = rec_cell(p, 0) % T, ly0 = rec_cell(p, 1) % T, lz0 = rec_cell(p, 2) % T;
  const int cx = lx0 == T - 1, cy = ly0 == T - 1, cz = lz0 == T - 1;
  // Crowded wavefront: some lane shares its base cell (and source range)
  // with the lane before it, as in dense clumps, where the sorted order puts
  // a cell's particles next to each other.  Only then do the targets go
  // through the wave-level aggregation; otherwise each lane adds directly.
  // Integer sums are exact, so both paths give the same totals.
  const int key = (((lz0 * T + ly0) * T + lx0) << 3) | d;
  const int prev_key = __shfl_up(key, 1);
  const bool crowded = __ballot(live && lane != 0 && prev_key == key) != 0ull;
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int lz = dz ? 0 : lz0 + a;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int ly = dy ? 0 : ly0 + b;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int lx = dx ? 0 : lx0 + c;
        const bool valid = live && (a & cz) == dz && (b & cy) == dy && (c & cx) == dx;
        const long long qv = valid ? rint_small(wx[c] * (wy[b] * wz[a])) : 0ll;
        if (crowded)
          lds_add_aggregated(acc, valid ? (lz * T + ly) * T + lx : 0, qv, valid);
        else if (valid)
          atomicAdd(acc + (lz * T + ly) * T + lx, (unsigned long long)qv);
      }
    }
  }
}

// Particles the 14 ranges of the tile of rank r hold, from the bucket offsets.
__device__ __forceinline__ unsigned tile_items(const unsigned* __restrict__ offset, const TileMap& tm, unsigned r) {
  unsigned w = 0;
#pragma unroll
  for (int k = 0; k < 14; ++k) {
    const int d = (int)((kRangeD >> (3 * k)) & 7ull);
    const unsigned b = tm.src[(size_t)r * 8 + d];
    w += offset[b + kRangeHi[k]] - offset[b + kRangeLo[k]];
  }
  return w;
}

// Slices per tile rank.  A split tile reserves its S - 1 extra workgroups
// from a budget of `slots` and writes their table entries; the first of them
// also names its scratch tile.  A tile that finds the budget spent marks its
// reservation unused and runs unsplit, so no more than ntiles + slots
// workgroups are ever needed.
__global__ __launch_bounds__(256) void k_deposit_plan(const unsigned* __restrict__ offset, const TileMap tm,
                                                      unsigned* __restrict__ slices, unsigned* __restrict__ slot,
                                                      unsigned* __restrict__ extra, SolveScalars* sc,
                                                      unsigned slots) {
  const unsigned ntiles = (unsigned)tm.nt * tm.nt * tm.nt;
  const unsigned r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ntiles) return;
  const unsigned w = tile_items(offset, tm, r);
  unsigned S = w > kSliceItems ? (w + kSliceItems - 1) / kSliceItems : 1u;
  if (S > kMaxSlices) S = kMaxSlices;
  if (S > 1) {
    const unsigned base = atomicAdd(&sc->split_budget, S - 1u);
    if (base + S - 1u > slots) {
      for (unsigned e = base; e < slots; ++e) extra[e] = kNoSlice;
      S = 1;
    } else {
      slot[r] = base;
      for (unsigned k = 1; k < S; ++k) extra[base + k - 1u] = r | (k << 20);
    }
  }
  slices[r] = S;
}
// One workgroup per tile; T is the tile side (T^3 int64 accumulators in LDS).
// Cell (x, y, z) of the shifted mass mesh is rho[z*rho_plane + y*nm + x];
// keep (when not null) receives the unshifted mass per cell, unpadded.
// The 72 bucket offsets of the 8 source tiles are fetched in parallel, the 14
// ranges are laid end to end, and all their particles are processed in one
// flat loop, so the only serial latency is one offset load.  Eight wavefronts
// per workgroup: the LDS accumulators cap a CU at four tiles, and eight waves
// each keep twice as many particle loads and atomics in flight as four.
constexpr int kDepositThreads = 512;

//
// TABLE: the range layout and the mass total come from k_tile_ranges (scalar
// loads, no LDS prologue); otherwise the workgroup computes them.
template <int T, int NT = kDepositThreads, bool SPLIT = false, bool TABLE = false>
__device__ __forceinline__ void deposit_tiles_body(const PartRec* __restrict__ sorted,
                                                       const unsigned* __restrict__ offset, const CellMap cm,
                                                       const TileMap tm, const SolveScalars* sc, long long ntot,
                                                       float* __restrict__ rho, long long rho_plane,
                                                       float* __restrict__ keep,
                                                       const DepositSplit sp = DepositSplit{}) {
  constexpr int T3 = T * T * T;
  __shared__ unsigned long long acc[T3];
  // Prologue scratch inside the accumulator tile, which is cleared only after
  // it has been read: the bucket offsets of the 8 source tiles (72 words),
  // the range prefix sums (15) and slot deltas (14), and the mass total.
  // With no other LDS the 16^3 tile takes exactly 32 KB, so five workgroups
  // share a CU's 160 KB instead of four.
  static_assert(T3 * 2 >= 106, "the prologue scratch must fit in the tile");
  unsigned* const toff = reinterpret_cast<unsigned*>(acc);  // [8][9]
  unsigned* const rpre = toff + 72;                         // [15]
  unsigned* const rdelta = toff + 88;                       // [14]
  unsigned long long* const qtotp = acc + 52;               // words 104-105
  const int tid = threadIdx.x;
  const int nt = tm.nt;
  // this workgroup's tile rank and slice (k of S): the primary workgroups
  // [0, ntiles) take slice 0 of their rank, extra workgroups the others
  unsigned rank_r = blockIdx.x, k_slice = 0, n_slice = 1, my_slot = 0;
  if constexpr (SPLIT) {
    if (blockIdx.x >= sp.ntiles) {
      const unsigned e = blockIdx.x - sp.ntiles;
      if (e >= sp.slots || e >= sp.sc->split_budget) return;  // uniform: no slice for this workgroup
      const unsigned v = sp.extra[e];
      if (v == kNoSlice) return;
      rank_r = v & 0xfffffu;
      k_slice = v >> 20;
    }
    n_slice = sp.slices[rank_r];
    if (n_slice > 1) my_slot = sp.slot[rank_r];
  }
  // the tile itself is needed only for the output rows; its load and the
  // source-bucket load below depend on the rank alone
  const int tile = (int)tm.order[rank_r];
  const int ux = tile % nt, uy = (tile / nt) % nt, uz = tile / (nt * nt);
  // range ends and slot deltas, uniform over the workgroup, in scalar
  // registers
  unsigned rend[13], dstep[13], dbase, all;
  unsigned long long qtot;
  if constexpr (TABLE) {
    const unsigned* __restrict__ w = tm.ranges + (size_t)rank_r * kRangeWords;
#pragma unroll
    for (int k = 0; k < 13; ++k) {
      rend[k] = w[k];
      dstep[k] = w[15 + k];
    }
    all = w[13];
    dbase = w[14];
    qtot = sc->qtot;
  } else {
  if (tid < 72) {
    const int d = tid / 9, j = tid % 9;
    toff[d * 9 + j] = offset[tm.src[(size_t)rank_r * 8 + d] + j];
  }
  // the fixed-point mass total, reduced by the third wavefront while the
  // offsets load (one slot per lane; NT >= 256)
  static_assert(NT >= 192 && kQSlots == 64, "one qsum slot per lane of the third wavefront");
  if (tid >= 128 && tid < 192) {
    unsigned long long v = sc->qsum[(tid - 128) * kQStride];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (tid == 128) *qtotp = v;
  }
  __syncthreads();
  if (tid < 64) {
    unsigned len = 0, beg = 0;
    if (tid < 14) {
      const int d = (int)((kRangeD >> (3 * tid)) & 7ull);
      beg = toff[d * 9 + kRangeLo[tid]];
      len = toff[d * 9 + kRangeHi[tid]] - beg;
    }
    unsigned incl = len;
    for (int off = 1; off < 16; off <<= 1) {
      const unsigned t = __shfl_up(incl, off);
      if (tid >= off) incl += t;
    }
    if (tid < 14) {
      rpre[tid + 1] = incl;
      // slot of particle q of range tid in the sorted array: q + rdelta
      rdelta[tid] = beg - (incl - len);
    }
    if (tid == 0) rpre[0] = 0;
  }
  __syncthreads();
  // then the scratch is free for the accumulators
#pragma unroll
  for (int k = 0; k < 13; ++k) rend[k] = __builtin_amdgcn_readfirstlane(rpre[k + 1]);
  dbase = __builtin_amdgcn_readfirstlane(rdelta[0]);
#pragma unroll
  for (int k = 0; k < 13; ++k) dstep[k] = __builtin_amdgcn_readfirstlane(rdelta[k + 1] - rdelta[k]);
  all = __builtin_amdgcn_readfirstlane(rpre[14]);
  qtot = *qtotp;
  __syncthreads();
  }
  const int s = fixed_shift(sc, ntot);
  // this slice's share [q0, total) of the tile's flattened ranges
  unsigned q0 = 0, total = all;
  if constexpr (SPLIT) {
    const unsigned share = (all + n_slice - 1u) / n_slice;
    q0 = min(all, k_slice * share);
    total = min(all, q0 + share);
  }
  // every lane of a wavefront runs the same number of iterations (lanes past
  // the end carry no particle), so the aggregation ballots see whole waves
  const unsigned rounds = (total - q0 + NT - 1u) / NT;
  // Range r of flattened item q (the number of range ends at or below it, 13
  // past the end) and its record: the sorted slot is q + the delta of its
  // range, summed as masked steps over the range starts at or below q (range
  // ends never decrease; written as arithmetic so that it stays in
  // registers).  Range 0, the tile's own bucket, holds most of its particles:
  // a wavefront whose 64 items all lie in it (the items of a wavefront are
  // consecutive, lane 0's first) skips the per-lane search.
  auto fetch = [&](unsigned q, int& r) {
    const unsigned qw = __builtin_amdgcn_readfirstlane(q);
    unsigned dl = dbase;
    r = 0;
    if (qw + 63u >= rend[0]) {
#pragma unroll
      for (int k = 0; k < 13; ++k) {
        const bool ge = q >= rend[k];
        r += ge ? 1 : 0;
        dl += dstep[k] & (0u - (unsigned)ge);
      }
    }
    PartRec v;
    if (q < total) {
      v = sorted[q + dl];
    } else {
      v.a[0] = v.a[1] = v.a[2] = 0ull;
    }
    return v;
  };
  // m 2^s of a record: the FP32 mass scaled exactly
  const double scale = ldexp(1.0, s);
  auto deposit_one = [&](const PartRec& p, int r, bool live) { deposit_record<T>(acc, p, r, live, scale); };
  // kPrefetch rounds of records are in flight at once: a tile of the uniform
  // headline holds about 2.5 rounds at 256 threads, so all of its loads
  // share one memory latency.  Slot j holds the record of round it + j; the
  // round loop is unrolled by kPrefetch so every slot index is static.
  constexpr int kPrefetch = 3;
  const unsigned qbase = q0 + (unsigned)tid;
  PartRec buf[kPrefetch];
  int rbuf[kPrefetch];
#pragma unroll
  for (int j = 0; j < kPrefetch; ++j) {
    const unsigned qj = qbase + (unsigned)j * NT;
    buf[j] = fetch(qj, rbuf[j]);
  }
  // the accumulators are cleared while the first records load
  for (int c = tid; c < T3; c += NT) acc[c] = 0ull;
  __syncthreads();
  for (unsigned it = 0; it < rounds; it += kPrefetch) {
#pragma unroll
    for (int j = 0; j < kPrefetch; ++j) {
      if (it + j < rounds) {  // uniform over the workgroup
        const unsigned qc = qbase + (it + j) * NT;
        const PartRec p = buf[j];
        const int r = rbuf[j];
        if (it + j + kPrefetch < rounds) {
          const unsigned qn = qc + (unsigned)kPrefetch * NT;
          buf[j] = fetch(qn, rbuf[j]);
        }
        deposit_one(p, r, qc < total);
      }
    }
  }
  __syncthreads();
  unsigned long long* part = nullptr;
  if (SPLIT && n_slice > 1) {
    // add this slice's partial tile into the scratch tile; the last slice
    // to finish converts the sum and clears the scratch tile
    part = sp.scratch + (size_t)my_slot * T3;
    for (int c = tid; c < T3; c += NT) {
      const unsigned long long v = acc[c];
      if (v) atomicAdd(part + c, v);
    }
    __threadfence();
    __syncthreads();
    // the accumulators have been read: their first word holds the flag
    unsigned* const last = reinterpret_cast<unsigned*>(acc);
    if (tid == 0) *last = atomicAdd(sp.done + my_slot, 1u) == n_slice - 1u;
    __syncthreads();
    if (!*last) return;
    __threadfence();
  }
  const double unit = ldexp(1.0, -s);
  // the integer part of the mean fixed-point mass per cell
  const long long qmean = (long long)(qtot / (unsigned long long)((long long)cm.nm * cm.nm * cm.nm));
  for (int c = tid; c < T3; c += NT) {
    const int lx = c % T, ly = (c / T) % T, lz = c / (T * T);
    const size_t row = (size_t)(uy * T + ly) * cm.nm + (ux * T + lx);
    long long v;
    if (SPLIT && part) {
      v = (long long)__hip_atomic_load(part + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      part[c] = 0ull;
    } else {
      v = (long long)acc[c];
    }
    rho[(size_t)(uz * T + lz) * rho_plane + row] = (float)((double)(v - qmean) * unit);
    if (keep) keep[(size_t)(uz * T + lz) * cm.nm * cm.nm + row] = (float)((double)v * unit);
  }
  if (SPLIT && part && tid == 0) sp.done[my_slot] = 0u;
}

// Particle moments per cell from the sorted cell records of the last sort,
// kCellMomParts workgroups per tile, one for each part of its z layers
// (5 int64 LDS accumulators per cell of the part), 1024 threads each, so that
// a dense tile (a Plummer core holds 10^5 particles) is read by several
// workgroups at once and in few rounds: for the particles whose base cell c is in the tile,
// mom[5c + a] = sum m (f_a - 1/2) (a = x, y, z), mom[5c + 3] =
// sum m |f - 1/2|^2, mom[5c + 4] = sum m, with f the 43-bit in-cell fraction
// of the record.  The sums are 64-bit fixed point (the deposit's scale), so
// they do not depend on the particle order; every cell of the tile is
// written once (zero for an empty cell), so no clearing pass and no global
// atomics are needed.
constexpr int kCellMoments = 5;
constexpr int kCellMomThreads = 1024;
constexpr int kCellMomParts = 2;
template <int T>
__global__ __launch_bounds__(kCellMomThreads) void k_cell_moments(const PartRec* __restrict__ sorted,
                                                                  const unsigned* __restrict__ offset,
                                                                  const CellMap cm, const TileMap tm,
                                                                  const SolveScalars* sc, long long ntot,
                                                                  double* __restrict__ mom) {
  constexpr int HL = T / kCellMomParts, HC = T * T * HL, NT = kCellMomThreads;
  static_assert(T % kCellMomParts == 0, "tiles split into parts of whole layers");
  __shared__ unsigned long long acc[kCellMoments][HC];
  const int tid = threadIdx.x;
  const int nt = tm.nt;
  const int tile = blockIdx.x, half = blockIdx.y;
  const int ux = tile % nt, uy = (tile / nt) % nt, uz = tile / (nt * nt);
  const unsigned b0 = tm.rank[tile] * 8u;
  const unsigned beg = offset[b0], end = offset[b0 + 8];
  const int s = fixed_shift(sc, ntot);
  const double scale = ldexp(1.0, s), unit = ldexp(1.0, -s);
  const unsigned rounds = (end - beg + NT - 1u) / NT;
  {
    for (int c = tid; c < HC; c += NT)
#pragma unroll
      for (int k = 0; k < kCellMoments; ++k) acc[k][c] = 0ull;
    __syncthreads();
    // every lane runs the same rounds, so the aggregation ballots see whole waves
    for (unsigned it = 0; it < rounds; ++it) {
      const unsigned q = beg + it * NT + (unsigned)tid;
      bool live = q < end;
      PartRec p;
      p.a[0] = p.a[1] = p.a[2] = 0ull;
      if (live) p = sorted[q];
      const int lx = rec_cell(p, 0) % T, ly = rec_cell(p, 1) % T, lz = rec_cell(p, 2) % T;
      live = live && lz / HL == half;
      const int c = ((lz - half * HL) * T + ly) * T + lx;
      const double fx = rec_frac(p, 0) - 0.5, fy = rec_frac(p, 1) - 0.5, fz = rec_frac(p, 2) - 0.5;
      const double ms = (double)rec_mass(p) * scale;
      const long long v[kCellMoments] = {rint_small(ms * fx), rint_small(ms * fy), rint_small(ms * fz),
                                          rint_small(ms * (fx * fx + fy * fy + fz * fz)), rint_small(ms)};
#pragma unroll
      for (int k = 0; k < kCellMoments; ++k) lds_add_aggregated(acc[k], live ? c : 0, live ? v[k] : 0ll, live);
    }
    __syncthreads();
    for (int c = tid; c < HC; c += NT) {
      const int lx = c % T, ly = (c / T) % T, lz = c / (T * T) + half * HL;
      const size_t cell = ((size_t)(uz * T + lz) * cm.nm + (uy * T + ly)) * cm.nm + (ux * T + lx);
#pragma unroll
      for (int k = 0; k < kCellMoments; ++k) mom[cell * kCellMoments + k] = (double)(long long)acc[k][c] * unit;
    }
  }
}

// At most 96 VGPRs, so five 256-thread workgroups (the LDS limit of a 16^3
// tile) fit a CU.
template <int T, int NT = kDepositThreads, bool SPLIT = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(5))) void k_deposit_tiles(const PartRec* __restrict__ sorted,
                                                       const unsigned* __restrict__ offset, const CellMap cm,
                                                       const TileMap tm, const SolveScalars* sc, long long ntot,
                                                       float* __restrict__ rho, long long rho_plane,
                                                       float* __restrict__ keep, const DepositSplit sp) {
  deposit_tiles_body<T, NT, SPLIT, true>(sorted, offset, cm, tm, sc, ntot, rho, rho_plane, keep, sp);
}

// Two tiles per workgroup: ranks 2w and 2w + 1, neighbours along the
// Hilbert curve, from the range table.  Their rounds of records form one
// sequence, so the prefetch ring runs across the tile boundary: the second
// tile's records load while the first tile is converted and written.  The
// accumulator tile, the per-record work and the result are those of
// k_deposit_tiles (integer sums: bit for bit the same density).
template <int T, int NT>
__device__ __forceinline__ void deposit_pair_body(const PartRec* __restrict__ sorted, const CellMap cm,
                                                  const TileMap tm, const SolveScalars* sc, long long ntot,
                                                  float* __restrict__ rho, long long rho_plane,
                                                  float* __restrict__ keep) {
  constexpr int T3 = T * T * T;
  __shared__ unsigned long long acc[T3];
  const int tid = threadIdx.x;
  const int nt = tm.nt;
  const unsigned ntiles = (unsigned)(nt * nt * nt);
  const unsigned ra = 2u * blockIdx.x, rb = ra + 1u;
  const bool has_b = rb < ntiles;  // uniform: the last workgroup may hold one tile
  const unsigned* __restrict__ wa = tm.ranges + (size_t)ra * kRangeWords;
  const unsigned* __restrict__ wb = tm.ranges + (size_t)(has_b ? rb : ra) * kRangeWords;
  const unsigned tot_a = wa[13], dbase_a = wa[14];
  const unsigned tot_b = has_b ? wb[13] : 0u, dbase_b = wb[14];
  const unsigned rounds_a = (tot_a + NT - 1u) / NT, rounds_b = (tot_b + NT - 1u) / NT;
  const unsigned rounds = rounds_a + rounds_b;
  const int tile_a = (int)tm.order[ra], tile_b = has_b ? (int)tm.order[rb] : 0;
  const int s = fixed_shift(sc, ntot);
  const double scale = ldexp(1.0, s), unit = ldexp(1.0, -s);
  const long long qmean = (long long)(sc->qtot / (unsigned long long)((long long)cm.nm * cm.nm * cm.nm));
  // record of concatenated round g (tile a below rounds_a, then tile b): its
  // range r and whether the lane has one; a wavefront whose 64 items all lie
  // in range 0 skips the per-lane range search, as in k_deposit_tiles
  // (the range words are read through the uniform table row of the tile, so
  // they come from the scalar cache rather than living in registers)
  auto fetch = [&](unsigned g, int& r, bool& live) {
    const bool in_b = g >= rounds_a;  // uniform
    const unsigned* __restrict__ w = in_b ? wb : wa;
    const unsigned q = (in_b ? g - rounds_a : g) * NT + (unsigned)tid;
    const unsigned qw = __builtin_amdgcn_readfirstlane(q);
    unsigned dl = in_b ? dbase_b : dbase_a;
    r = 0;
    if (qw + 63u >= w[0]) {
#pragma unroll
      for (int k = 0; k < 13; ++k) {
        const bool ge = q >= w[k];
        r += ge ? 1 : 0;
        dl += w[15 + k] & (0u - (unsigned)ge);
      }
    }
    live = q < (in_b ? tot_b : tot_a);
    PartRec v;
    if (live) {
      v = sorted[q + dl];
    } else {
      v.a[0] = v.a[1] = v.a[2] = 0ull;
    }
    return v;
  };
  // the accumulated tile as FP32 mass per cell minus the mean (and, with
  // keep, unshifted), then the accumulators cleared for the next tile
  auto write_tile = [&](int tile) {
    const int ux = tile % nt, uy = (tile / nt) % nt, uz = tile / (nt * nt);
    for (int c = tid; c < T3; c += NT) {
      const int lx = c % T, ly = (c / T) % T, lz = c / (T * T);
      const size_t row = (size_t)(uy * T + ly) * cm.nm + (ux * T + lx);
      const long long v = (long long)acc[c];
      rho[(size_t)(uz * T + lz) * rho_plane + row] = (float)((double)(v - qmean) * unit);
      if (keep) keep[(size_t)(uz * T + lz) * cm.nm * cm.nm + row] = (float)((double)v * unit);
    }
  };
  constexpr int kPrefetch = 3;
  PartRec buf[kPrefetch];
  int rbuf[kPrefetch];
  bool lbuf[kPrefetch];
#pragma unroll
  for (int j = 0; j < kPrefetch; ++j) {
    rbuf[j] = 0;
    lbuf[j] = false;
    if ((unsigned)j < rounds) buf[j] = fetch((unsigned)j, rbuf[j], lbuf[j]);
  }
  for (int c = tid; c < T3; c += NT) acc[c] = 0ull;
  __syncthreads();
  for (unsigned it = 0; it < rounds; it += kPrefetch) {
#pragma unroll
    for (int j = 0; j < kPrefetch; ++j) {
      const unsigned g = it + (unsigned)j;
      if (g < rounds) {  // uniform over the workgroup
        if (g == rounds_a) {
          // tile a is complete: write it and start tile b from zero
          __syncthreads();
          write_tile(tile_a);
          __syncthreads();
          for (int c = tid; c < T3; c += NT) acc[c] = 0ull;
          __syncthreads();
        }
        const PartRec p = buf[j];
        const int r = rbuf[j];
        const bool live = lbuf[j];
        if (g + kPrefetch < rounds) buf[j] = fetch(g + kPrefetch, rbuf[j], lbuf[j]);
        deposit_record<T>(acc, p, r, live, scale);
      }
    }
  }
  __syncthreads();
  if (rounds_b == 0u) {
    // tile b had no records (or there is no tile b): tile a is still in acc
    write_tile(tile_a);
    if (!has_b) return;  // uniform
    __syncthreads();
    for (int c = tid; c < T3; c += NT) acc[c] = 0ull;
    __syncthreads();
  }
  write_tile(tile_b);
}

template <int T, int NT = kDepositThreads>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(5))) void k_deposit_pairs(
    const PartRec* __restrict__ sorted, const CellMap cm, const TileMap tm, const SolveScalars* sc, long long ntot,
    float* __restrict__ rho, long long rho_plane, float* __restrict__ keep) {
  deposit_pair_body<T, NT>(sorted, cm, tm, sc, ntot, rho, rho_plane, keep);
}

template <int T, int NT = kDepositThreads>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(5))) void k_deposit_tiles_box(const BoxArgs* __restrict__ bx, const TileMap tm) {
  const BoxArgs& b = bx[blockIdx.z];
  deposit_tiles_body<T, NT>(b.sorted, b.offset, b.cm, tm, b.sc, b.n, reinterpret_cast<float*>(b.P),
                            2 * mesh_plane(b.cm.nm), nullptr);
}

}  // namespace gng
