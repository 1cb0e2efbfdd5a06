// shortrange.hip -- TreePM short-range force by direct summation over a
// chaining mesh.
//
// Reference: crates/gadget-ng-treepm/src/short_range.rs (pair law :384-398,
// minimum image :53-63, periodic walk :232-330).  The reference walks an
// octree and replaces small nodes by their monopole; here every pair inside
// r_cut is summed exactly, which is the quantity that walk approximates.
//
// Method:
//   1. particles are counted into a chaining mesh of nc^3 cells of side
//      L/nc >= r_cut (nc >= 3) and scanned; a stable radix sort by cell
//      (sort.hpp) puts them in cell order with ascending particle index inside
//      each cell, and they are gathered as float4 (position relative to the
//      corner of their unwrapped cell, mass) plus the original index.  The
//      order of every sum is then fixed, so the result is bit-identical from
//      call to call;
//   2. the work list is (cell, chunk of 256 particles); its length is read
//      back once to size the grid, so every workgroup has one chunk and the
//      hardware dispatcher balances cells of very different occupancy;
//   3. a workgroup streams the 27 neighbour cells through LDS 256 particles at
//      a time and each lane sums its particle's pairs: FP32 inside a staged
//      chunk, FP64 across chunks.  A neighbour cell at offset d (cells, each
//      component -1, 0 or 1) holds its particles at rel + d * L/nc from the
//      lane's cell corner, so the displacement needs no per-pair minimum
//      image: every pair within r_cut <= L/nc <= L/3 is its own minimum image
//      in that form, and no pair outside the cut comes inside it.
//
// The split factor uses erfc from Abramowitz & Stegun 7.1.26 (|error| <=
// 1.5e-7), sharing exp(-x^2) with the exact-split term.
#include <hip/hip_runtime.h>

#include <cmath>
#include <new>
#include <string>

#include "../../include/gng_learned.h"
#include "sort.hpp"
#include "status.hpp"

namespace gng {
namespace sr {

constexpr int kChunk = 256;

__global__ void k_fill_u32(unsigned* __restrict__ p, long long n, unsigned v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p) { *p = 0ull; }

// The unwrapped chaining-mesh cell floor(v nc / L) of one coordinate, in FP64
// (cs = nc / L), and its index wrapped into [0, nc); a non-finite coordinate
// lands in cell 0.  k_count and k_gather both start from cell_floor, so a
// particle's cell and its in-cell offset always agree.
__device__ __forceinline__ double cell_floor(float v, double cs) { return floor((double)v * cs); }

__device__ __forceinline__ int wrap_cell(double fl, int nc) {
  double c = fmod(fl, (double)nc);
  if (c < 0.0) c += (double)nc;
  return c >= 0.0 && c < (double)nc ? (int)c : 0;
}

__global__ void k_count(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                        long long n, int nc, double cs, unsigned* __restrict__ cnt, unsigned* __restrict__ cell_of) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int cx = wrap_cell(cell_floor(x[p], cs), nc), cy = wrap_cell(cell_floor(y[p], cs), nc),
            cz = wrap_cell(cell_floor(z[p], cs), nc);
  const unsigned c = (unsigned)(((long long)cz * nc + cy) * nc + cx);
  cell_of[p] = c;
  atomicAdd(&cnt[c], 1u);
}

// One workgroup of 1024 threads: exclusive scans of the counts (off) and of
// the chunk counts ceil(cnt/kChunk) (choff), each with its total at [ncell].
__global__ __launch_bounds__(1024) void k_scan(const unsigned* __restrict__ cnt, long long ncell,
                                               unsigned* __restrict__ off, unsigned* __restrict__ choff) {
  __shared__ unsigned long long s1[1024], s2[1024];
  const int t = threadIdx.x;
  const long long seg = (ncell + 1023) / 1024;
  const long long a = t * seg, b = a + seg < ncell ? a + seg : ncell;
  unsigned long long c1 = 0, c2 = 0;
  for (long long i = a; i < b; ++i) {
    c1 += cnt[i];
    c2 += (cnt[i] + kChunk - 1) / kChunk;
  }
  s1[t] = c1;
  s2[t] = c2;
  __syncthreads();
  for (int d = 1; d < 1024; d <<= 1) {
    const unsigned long long u1 = t >= d ? s1[t - d] : 0, u2 = t >= d ? s2[t - d] : 0;
    __syncthreads();
    s1[t] += u1;
    s2[t] += u2;
    __syncthreads();
  }
  unsigned long long r1 = s1[t] - c1, r2 = s2[t] - c2;
  for (long long i = a; i < b; ++i) {
    off[i] = (unsigned)r1;
    choff[i] = (unsigned)r2;
    r1 += cnt[i];
    r2 += (cnt[i] + kChunk - 1) / kChunk;
  }
  if (t == 1023) {
    off[ncell] = (unsigned)s1[1023];
    choff[ncell] = (unsigned)s2[1023];
  }
}

// sorted[i] = particle order[i] (the cell-sorted permutation): its position
// relative to the corner of its unwrapped cell, v - floor(v nc / L) L / nc,
// formed in FP64 from the same floor as k_count's (so it is also the offset
// from the wrapped cell's corner, whatever image of the box v lies in), and
// its mass
__global__ void k_gather(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                         const float* __restrict__ m, long long n, const unsigned* __restrict__ order, double cs,
                         double cside, float4* __restrict__ sorted, unsigned* __restrict__ sidx) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned p = order[i];
  const float vx = x[p], vy = y[p], vz = z[p];
  sorted[i] = make_float4((float)((double)vx - cell_floor(vx, cs) * cside),
                          (float)((double)vy - cell_floor(vy, cs) * cside),
                          (float)((double)vz - cell_floor(vz, cs) * cside), m[p]);
  sidx[i] = p;
}

struct Params {
  float csize, rc2, eps2, inv_s2rs, g;  // csize: chaining cell side L/nc
  float log2e_s2rs2;                    // log2(e) / (2 r_s^2)
  int nc;
  long long ncell;
};

template <int KERNEL>
__global__ __launch_bounds__(256) void k_force(const float4* __restrict__ sorted, const unsigned* __restrict__ sidx,
                                               const unsigned* __restrict__ off, const unsigned* __restrict__ choff,
                                               Params pr, float* __restrict__ ax, float* __restrict__ ay,
                                               float* __restrict__ az, unsigned long long* __restrict__ pairs) {
  __shared__ float4 js[kChunk];
  // work item -> (cell, chunk): last cell with choff[c] <= blockIdx.x
  const unsigned w = blockIdx.x;
  long long lo = 0, hi = pr.ncell;  // invariant: choff[lo] <= w < choff[hi]
  while (hi - lo > 1) {
    const long long mid = (lo + hi) >> 1;
    if (choff[mid] <= w) lo = mid;
    else hi = mid;
  }
  const long long cell = lo;
  const unsigned q = w - choff[cell];
  const unsigned i = off[cell] + q * kChunk + threadIdx.x;
  const bool live = i < off[cell + 1];
  const float4 pi = live ? sorted[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const int nc = pr.nc;
  const int cx = (int)(cell % nc), cy = (int)((cell / nc) % nc), cz = (int)(cell / ((long long)nc * nc));
  double sx = 0.0, sy = 0.0, sz = 0.0;
  unsigned long long seen = 0;
  for (int nb = 0; nb < 27; ++nb) {
    const int ox = nb % 3 - 1, oy = (nb / 3) % 3 - 1, oz = nb / 9 - 1;
    // the lane's particle relative to this neighbour cell's corner
    const float px = pi.x - (float)ox * pr.csize, py = pi.y - (float)oy * pr.csize, pz = pi.z - (float)oz * pr.csize;
    int ux = cx + ox, uy = cy + oy, uz = cz + oz;
    ux += ux < 0 ? nc : 0;
    ux -= ux >= nc ? nc : 0;
    uy += uy < 0 ? nc : 0;
    uy -= uy >= nc ? nc : 0;
    uz += uz < 0 ? nc : 0;
    uz -= uz >= nc ? nc : 0;
    const long long nbc = ((long long)uz * nc + uy) * nc + ux;
    const unsigned j0 = off[nbc], j1 = off[nbc + 1];
    seen += j1 - j0;
    for (unsigned jb = j0; jb < j1; jb += kChunk) {
      __syncthreads();
      if (jb + threadIdx.x < j1) {
        float4 v = sorted[jb + threadIdx.x];
        v.w *= pr.g;  // G m, staged once per particle rather than per pair
        js[threadIdx.x] = v;
      }
      __syncthreads();
      const int cnt = (int)(j1 - jb < (unsigned)kChunk ? j1 - jb : (unsigned)kChunk);
      float fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma unroll 4
      for (int k = 0; k < cnt; ++k) {
        const float4 pj = js[k];
        const float dx = pj.x - px, dy = pj.y - py, dz = pj.z - pz;
        const float r2b = dx * dx + dy * dy + dz * dz;
        // r2b > 0 skips coincident pairs and the particle itself (for the
        // lane's own cell the offset is 0 * csize = 0, so its displacement
        // is exactly zero)
        if (r2b <= pr.rc2 && r2b > 0.0f) {
          const float r2 = r2b + pr.eps2;
          const float rinv = rsqrtf(r2);
          const float x = r2 * rinv * pr.inv_s2rs;
          // exp(-x^2) with x^2 = r2 / (2 r_s^2), on the hardware exp2
          const float e = __builtin_amdgcn_exp2f(-r2 * pr.log2e_s2rs2);
          // hardware reciprocal (1 ulp), far below the polynomial's 1.5e-7
          const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * x);
          const float poly =
              t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
          const float S = e * (KERNEL == GNG_SR_EXACT ? poly + 1.1283791670955126f * x : poly);
          const float f = pj.w * S * (rinv * rinv * rinv);
          fx += f * dx;
          fy += f * dy;
          fz += f * dz;
        }
      }
      sx += fx;
      sy += fy;
      sz += fz;
    }
  }
  if (live) {
    const unsigned o = sidx[i];
    ax[o] = (float)sx;
    ay[o] = (float)sy;
    az[o] = (float)sz;
  }
  // evaluated pairs: live lanes x particles seen
  if (threadIdx.x == 0) {
    const unsigned nlive = off[cell + 1] - (off[cell] + q * kChunk);
    const unsigned nl = nlive < (unsigned)kChunk ? nlive : (unsigned)kChunk;
    atomicAdd(pairs, (unsigned long long)nl * seen);
  }
}

}  // namespace sr
}  // namespace gng

using namespace gng::sr;
using gng::fail;

struct gng_sr {
  int device = 0;
  long long cap_n = 0;
  float4* d_sorted = nullptr;
  unsigned* d_sidx = nullptr;
  unsigned* d_cell_of = nullptr;
  unsigned* d_cell_sorted = nullptr;  // radix-sort outputs
  unsigned* d_order = nullptr;
  void* d_temp = nullptr;  // radix-sort work space (sized for cap_n)
  size_t temp_bytes = 0;
  long long cap_cells = 0;
  unsigned* d_cnt = nullptr;
  unsigned* d_off = nullptr;
  unsigned* d_choff = nullptr;
  unsigned long long* d_pairs = nullptr;
  long long last_pairs = -1;
  bool pairs_pending = false;
};

namespace {

void free_sr(gng_sr* s) {
  (void)hipFree(s->d_sorted);
  (void)hipFree(s->d_sidx);
  (void)hipFree(s->d_cell_of);
  (void)hipFree(s->d_cell_sorted);
  (void)hipFree(s->d_order);
  (void)hipFree(s->d_temp);
  (void)hipFree(s->d_cnt);
  (void)hipFree(s->d_off);
  (void)hipFree(s->d_choff);
  (void)hipFree(s->d_pairs);
}

}  // namespace

namespace {

// Chaining-mesh bookkeeping shared by the entry points: cell of every
// particle, per-cell counts and offsets, chunk offsets, and the stable
// cell-sorted permutation d_order.  Returns nc through *nc_out.
int cell_sort(gng_sr* sr, const float* d_x, const float* d_y, const float* d_z, int64_t n, double box_size,
              double r_cut, hipStream_t s, int* nc_out) {
  int nc = (int)std::floor(box_size / r_cut);
  if (nc > 256) nc = 256;
  const long long ncell = (long long)nc * nc * nc;
  if (n > sr->cap_n) {
    (void)hipFree(sr->d_sorted);
    (void)hipFree(sr->d_sidx);
    (void)hipFree(sr->d_cell_of);
    (void)hipFree(sr->d_cell_sorted);
    (void)hipFree(sr->d_order);
    (void)hipFree(sr->d_temp);
    sr->d_sorted = nullptr;
    sr->d_sidx = sr->d_cell_of = sr->d_cell_sorted = sr->d_order = nullptr;
    sr->d_temp = nullptr;
    sr->cap_n = 0;
    GNG_HIP(hipMalloc(&sr->d_sorted, sizeof(float4) * (size_t)n));
    GNG_HIP(hipMalloc(&sr->d_sidx, sizeof(unsigned) * (size_t)n));
    GNG_HIP(hipMalloc(&sr->d_cell_of, sizeof(unsigned) * (size_t)n));
    GNG_HIP(hipMalloc(&sr->d_cell_sorted, sizeof(unsigned) * (size_t)n));
    GNG_HIP(hipMalloc(&sr->d_order, sizeof(unsigned) * (size_t)n));
    GNG_HIP(hipMalloc(&sr->d_temp, gng::radix::sort_workspace(n)));
    sr->cap_n = n;
  }
  if (ncell > sr->cap_cells) {
    (void)hipFree(sr->d_cnt);
    (void)hipFree(sr->d_off);
    (void)hipFree(sr->d_choff);
    sr->d_cnt = sr->d_off = sr->d_choff = nullptr;
    sr->cap_cells = 0;
    GNG_HIP(hipMalloc(&sr->d_cnt, sizeof(unsigned) * (size_t)ncell));
    GNG_HIP(hipMalloc(&sr->d_off, sizeof(unsigned) * (size_t)(ncell + 1)));
    GNG_HIP(hipMalloc(&sr->d_choff, sizeof(unsigned) * (size_t)(ncell + 1)));
    sr->cap_cells = ncell;
  }
  const unsigned cb = (unsigned)((ncell + 255) / 256), pb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_fill_u32, dim3(cb), dim3(256), 0, s, sr->d_cnt, ncell, 0u);
  const double cs = (double)nc / box_size;
  hipLaunchKernelGGL(k_count, dim3(pb), dim3(256), 0, s, d_x, d_y, d_z, (long long)n, nc, cs, sr->d_cnt,
                     sr->d_cell_of);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, sr->d_cnt, ncell, sr->d_off, sr->d_choff);
  GNG_HIP(hipGetLastError());
  // stable LSD radix sort over the bits a cell index needs
  int end_bit = 1;
  while ((1LL << end_bit) < ncell) ++end_bit;
  GNG_HIP(gng::radix::sort_pairs(sr->d_cell_of, n, end_bit, sr->d_cell_sorted, sr->d_order, sr->d_temp, s));
  *nc_out = nc;
  return GNG_OK;
}

int check_geometry(int64_t n, double box_size, double r_cut) {
  if (n < 0) return fail(GNG_ERR_INVALID, "negative particle count");
  if (n >= (1LL << 31)) return fail(GNG_ERR_INVALID, "n of 2^31 or more is not supported");
  if (!(box_size > 0.0) || !(r_cut > 0.0)) return fail(GNG_ERR_INVALID, "box_size and r_cut must be positive");
  if (box_size < 3.0 * r_cut) return fail(GNG_ERR_INVALID, "box_size must be at least 3 r_cut");
  return GNG_OK;
}

}  // namespace

extern "C" {

int gng_sr_create(int32_t device, gng_sr** out) {
  if (!out) return fail(GNG_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(GNG_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= ndev) return fail(GNG_ERR_INVALID, "device index out of range");
  GNG_HIP(hipSetDevice(device));
  gng_sr* s = new (std::nothrow) gng_sr();
  if (!s) return fail(GNG_ERR_OOM, "host allocation failed");
  s->device = device;
  if (hipMalloc(&s->d_pairs, sizeof(unsigned long long)) != hipSuccess) {
    delete s;
    return fail(GNG_ERR_OOM, "short-range counters");
  }
  *out = s;
  return GNG_OK;
}

void gng_sr_destroy(gng_sr* sr) {
  if (!sr) return;
  (void)hipSetDevice(sr->device);
  (void)hipDeviceSynchronize();
  free_sr(sr);
  delete sr;
}


int gng_sr_cell_order(gng_sr* sr, const float* d_x, const float* d_y, const float* d_z, int64_t n,
                      double box_size, double r_cut, uint32_t* d_order, void* stream) {
  if (!sr) return fail(GNG_ERR_INVALID, "null short-range context");
  if (int e = check_geometry(n, box_size, r_cut)) return e;
  if (n == 0) return GNG_OK;
  if (!d_x || !d_y || !d_z || !d_order) return fail(GNG_ERR_INVALID, "null array");
  GNG_HIP(hipSetDevice(sr->device));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  int nc = 0;
  if (int e = cell_sort(sr, d_x, d_y, d_z, n, box_size, r_cut, s, &nc)) return e;
  GNG_HIP(hipMemcpyAsync(d_order, sr->d_order, sizeof(unsigned) * (size_t)n, hipMemcpyDeviceToDevice, s));
  return GNG_OK;
}

int gng_sr_accelerations(gng_sr* sr, const float* d_x, const float* d_y, const float* d_z, const float* d_m,
                         int64_t n, double box_size, double r_split, double r_cut, double eps2, double g,
                         int32_t kernel, float* d_ax, float* d_ay, float* d_az, void* stream) {
  if (!sr) return fail(GNG_ERR_INVALID, "null short-range context");
  if (int e = check_geometry(n, box_size, r_cut)) return e;
  if (!(r_split > 0.0) || !(eps2 >= 0.0)) return fail(GNG_ERR_INVALID, "r_split must be positive and eps2 non-negative");
  if (kernel != GNG_SR_EXACT && kernel != GNG_SR_REFERENCE) return fail(GNG_ERR_INVALID, "unknown split kernel");
  if (n == 0) return GNG_OK;
  if (!d_x || !d_y || !d_z || !d_m || !d_ax || !d_ay || !d_az) return fail(GNG_ERR_INVALID, "null array");
  GNG_HIP(hipSetDevice(sr->device));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  int nc = 0;
  if (int e = cell_sort(sr, d_x, d_y, d_z, n, box_size, r_cut, s, &nc)) return e;
  const long long ncell = (long long)nc * nc * nc;
  const unsigned pb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, s, sr->d_pairs);
  // the same cs as the cell count (cell_sort), and the cell side
  const double cs = (double)nc / box_size, cside = box_size / (double)nc;
  const float csize = (float)cside;
  hipLaunchKernelGGL(k_gather, dim3(pb), dim3(256), 0, s, d_x, d_y, d_z, d_m, (long long)n, sr->d_order, cs, cside,
                     sr->d_sorted, sr->d_sidx);
  GNG_HIP(hipGetLastError());
  unsigned chunks = 0;
  GNG_HIP(hipMemcpyAsync(&chunks, sr->d_choff + ncell, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  GNG_HIP(hipStreamSynchronize(s));
  if (chunks == 0) return fail(GNG_ERR_HIP, "short-range work list is empty");
  Params pr;
  pr.csize = csize;
  pr.rc2 = (float)(r_cut * r_cut);
  pr.eps2 = (float)eps2;
  pr.inv_s2rs = (float)(1.0 / (std::sqrt(2.0) * r_split));
  pr.g = (float)g;
  pr.log2e_s2rs2 = (float)(1.4426950408889634 / (2.0 * r_split * r_split));
  pr.nc = nc;
  pr.ncell = ncell;
  if (kernel == GNG_SR_EXACT)
    hipLaunchKernelGGL((k_force<GNG_SR_EXACT>), dim3(chunks), dim3(256), 0, s, sr->d_sorted, sr->d_sidx, sr->d_off,
                       sr->d_choff, pr, d_ax, d_ay, d_az, sr->d_pairs);
  else
    hipLaunchKernelGGL((k_force<GNG_SR_REFERENCE>), dim3(chunks), dim3(256), 0, s, sr->d_sorted, sr->d_sidx,
                       sr->d_off, sr->d_choff, pr, d_ax, d_ay, d_az, sr->d_pairs);
  GNG_HIP(hipGetLastError());
  sr->pairs_pending = true;
  return GNG_OK;
}

int64_t gng_sr_last_pair_count(const gng_sr* sr) {
  if (!sr) return -1;
  gng_sr* m = const_cast<gng_sr*>(sr);
  if (m->pairs_pending) {
    unsigned long long v = 0;
    if (hipSetDevice(m->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(&v, m->d_pairs, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess)
      return -1;
    m->last_pairs = (long long)v;
    m->pairs_pending = false;
  }
  return m->last_pairs;
}

}  // extern "C"
