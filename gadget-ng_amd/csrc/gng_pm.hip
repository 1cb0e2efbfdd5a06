// gng_pm.hip -- plan, dispatch and C ABI of the MI355X PM solver.
//
// The plan owns the device meshes (allocated once at create, as the
// reference's handle does: crates/gadget-ng-hip/hip/pm_gravity.hip:278-334)
// and grows the particle buffers on demand (pm_gravity.hip:376-392).  Every
// solve is a fixed sequence of launches on one stream; host entry points add
// the H2D/D2H copies and a stream synchronisation.
//
// Mesh memory: three packed half-spectrum buffers P, A, B of nm z planes,
// mesh_plane(nm) complex values apart (4 nm^3 bytes each, plus a small pad
// per plane from nm = 512 on).  The deposit writes the FP32 mass per cell into P (or into a
// separate buffer when the density is kept for readback) and every FFT pass
// works in place or ping-pongs among P, A, B, ending with F_x in A, F_y in P
// and F_z in B.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gng_ic.h"
#include "../../include/gng_pm.h"
#include "deposit.hpp"
#include "ic_kernels.hpp"
#include "pk_kernels.hpp"
#include "pm_kernels.hpp"
#include "status.hpp"

using namespace gng;

namespace {
thread_local std::string g_last_error;
}  // namespace

int gng::fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

namespace {

// stages: 0 bucket count, 1 sort (scan + scatter), 2 deposit, 3 x forward,
// 4 y forward, 5 z + Green, 6 y inverse, 7 x inverse, 8 interpolation
constexpr int kStages = 9;
constexpr int kTimingSlots = 64;

// ---------------------------------------------------------------------------
// per-mesh-size kernel table
// ---------------------------------------------------------------------------

constexpr int pow2_floor(int v) {
  int p = 1;
  while (p * 2 <= v) p *= 2;
  return p;
}
constexpr long long pow2_part(long long v) {
  long long p = 1;
  while (v % (p * 2) == 0) p *= 2;
  return p;
}
template <int N>
constexpr int x_rows() {
  constexpr int H = N / 2;
  constexpr int a = pow2_floor(2048 / H > 0 ? 2048 / H : 1);
  constexpr long long b = pow2_part((long long)N * N);
  return a < b ? a : (int)b;
}
// Deposit tile side: the largest of 16, 12, 8, 4 that divides nm.
constexpr int deposit_tile(int nm) { return nm % 16 == 0 ? 16 : nm % 12 == 0 ? 12 : nm % 8 == 0 ? 8 : 4; }

// Launch choices with more than one compiled variant.  Defaults: the
// persistent z pass at nm = 512 (elsewhere it spills or loses occupancy) and
// 512-thread deposit workgroups; gng_pm_autotune() times the alternatives.
struct Tuning {
  int z_persistent = 0;
  int deposit_wide = 1;
  int y_fwd_persistent = 0;
  int y_inv_persistent = 0;
  int tuned = 0;  // 1 when measured (or read from the tuning cache)
  int interp_sorted = 0;  // full-set gather in the deposit's sorted order
  int chunk_fwd = 0;      // z planes per chunk of the forward x/y pass pair (0: whole mesh)
  int chunk_inv = 0;      // same for the inverse y/x pass pair
  int deposit_split = 1;  // heavy deposit tiles over several workgroups
  int y_inv_narrow = 1;   // inverse y pass on the narrower tile where one is compiled
  int deposit_pair = 0;   // unsplit deposit: two tiles per workgroup
};

// Planes per chunk when a pass pair runs chunk by chunk, so that the second
// pass reads what the first wrote from the 256 MiB last-level cache: the
// largest power of two whose planes of one complex field take at most
// budget_mib (the forward pair moves one field, the inverse pair three).
// 0 where the whole mesh is that small.
int chunk_rule(int nm, long long budget_mib) {
  const long long plane_bytes = 8LL * nm * (nm / 2);
  int c = nm;
  while (c > 1 && (long long)c * plane_bytes > (budget_mib << 20)) c /= 2;
  return c < nm && nm % c == 0 ? c : 0;
}
// measured best on MI355X: 256 planes at 512^3 and 64 at 1024^3 for the
// forward pair, 64 and 16 for the inverse pair
constexpr long long kChunkFwdMiB = 256, kChunkInvMiB = 64;

Tuning default_tuning(int nm) {
  Tuning t;
  t.z_persistent = nm == 512;
  t.deposit_wide = 1;
  t.chunk_fwd = chunk_rule(nm, kChunkFwdMiB);
  // at nm >= 1024 the inverse y pass runs as two launches on full lines, and
  // whole-mesh passes measured faster (10.5 against 11.0 ms at 1024^3)
  t.chunk_inv = nm >= 1024 ? 0 : chunk_rule(nm, kChunkInvMiB);
  return t;
}

// knob 8 (two tiles per deposit workgroup) only acts with the split off, so
// it comes after knob 6 in the coordinate descent
constexpr int kKnobs = 9;
int& knob(Tuning& t, int k) {
  switch (k) {
    case 0: return t.z_persistent;
    case 1: return t.deposit_wide;
    case 2: return t.y_fwd_persistent;
    case 3: return t.y_inv_persistent;
    case 4: return t.chunk_fwd;
    case 5: return t.chunk_inv;
    case 6: return t.deposit_split;
    case 7: return t.y_inv_narrow;
    default: return t.deposit_pair;
  }
}

// Process-wide cache of measured choices, keyed by (device architecture, nm).
// When GNG_PM_TUNE_FILE names a file, it is read once (per line: arch, nm and
// the kKnobs knobs in knob() order) and rewritten after every new measurement,
// so later processes reuse the decisions.
std::mutex g_tune_mu;
std::map<std::pair<std::string, int>, Tuning> g_tune;
bool g_tune_loaded = false;

std::string device_arch(int dev) {
  hipDeviceProp_t pr;
  if (hipGetDeviceProperties(&pr, dev) != hipSuccess) return "unknown";
  std::string a = pr.gcnArchName;
  const size_t c = a.find(':');
  return c == std::string::npos ? a : a.substr(0, c);
}

void load_tune_file_locked() {
  g_tune_loaded = true;
  const char* f = std::getenv("GNG_PM_TUNE_FILE");
  if (!f || !*f) return;
  FILE* fp = std::fopen(f, "r");
  if (!fp) return;
  char arch[64];
  int nm, v[kKnobs];
  while (std::fscanf(fp, "%63s %d %d %d %d %d %d %d %d %d %d", arch, &nm, &v[0], &v[1], &v[2], &v[3], &v[4], &v[5],
                     &v[6], &v[7], &v[8]) == 2 + kKnobs) {
    Tuning t;
    for (int k = 0; k < 4; ++k) knob(t, k) = v[k] != 0;
    t.deposit_split = v[6] != 0;
    t.y_inv_narrow = v[7] != 0;
    t.deposit_pair = v[8] != 0;
    t.chunk_fwd = v[4] == chunk_rule(nm, kChunkFwdMiB) ? v[4] : 0;
    t.chunk_inv = v[5] == chunk_rule(nm, kChunkInvMiB) ? v[5] : 0;
    t.tuned = 1;
    g_tune[{std::string(arch), nm}] = t;
  }
  std::fclose(fp);
}

Tuning cached_tuning(int dev, int nm) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (!g_tune_loaded) load_tune_file_locked();
  auto it = g_tune.find({device_arch(dev), nm});
  return it == g_tune.end() ? default_tuning(nm) : it->second;
}

void store_tuning(int dev, int nm, const Tuning& t) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (!g_tune_loaded) load_tune_file_locked();
  g_tune[{device_arch(dev), nm}] = t;
  const char* f = std::getenv("GNG_PM_TUNE_FILE");
  if (!f || !*f) return;
  FILE* fp = std::fopen(f, "w");
  if (!fp) return;
  for (const auto& kv : g_tune)
    std::fprintf(fp, "%s %d %d %d %d %d %d %d %d %d %d\n", kv.first.first.c_str(), kv.first.second,
                 kv.second.z_persistent, kv.second.deposit_wide, kv.second.y_fwd_persistent,
                 kv.second.y_inv_persistent, kv.second.chunk_fwd, kv.second.chunk_inv, kv.second.deposit_split,
                 kv.second.y_inv_narrow, kv.second.deposit_pair);
  std::fclose(fp);
}

// Uniform pseudo-random particles in [0, box)^3 with unit masses, for timing
// (a 32-bit integer hash of the index per coordinate).
__global__ void k_synth_particles(float* __restrict__ x, float* __restrict__ y, float* __restrict__ z,
                                  float* __restrict__ m, long long n, float box) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned h = (unsigned)i * 0x9E3779B9u + 0x7F4A7C15u;
  float c[3];
  for (int a = 0; a < 3; ++a) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    c[a] = (float)(h >> 8) * (1.0f / 16777216.0f) * box;
  }
  x[i] = c[0] < box ? c[0] : 0.0f;
  y[i] = c[1] < box ? c[1] : 0.0f;
  z[i] = c[2] < box ? c[2] : 0.0f;
  m[i] = 1.0f;
}

// 3-D Hilbert key of integer coordinates with `bits` bits each: Skilling's
// transform of the axes to the "transposed" Hilbert index (J. Skilling,
// Programming the Hilbert curve, AIP Conf. Proc. 707, 381 (2004)), then the
// bits interleaved with x most significant.  For bits = 1 the curve visits
// (0,0,0) (0,0,1) (0,1,1) (0,1,0) (1,1,0) (1,1,1) (1,0,1) (1,0,0).
__host__ __device__ uint64_t hilbert_key3(uint32_t x, uint32_t y, uint32_t z, int bits) {
  uint32_t X[3] = {x, y, z};
  if (bits <= 0) return 0;
  const uint32_t M = 1u << (bits - 1);
  for (uint32_t Q = M; Q > 1; Q >>= 1) {
    const uint32_t P = Q - 1;
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; ++i) X[i] ^= t;
  uint64_t key = 0;
  for (int b = bits - 1; b >= 0; --b)
    for (int i = 0; i < 3; ++i) key = (key << 1) | ((X[i] >> b) & 1u);
  return key;
}


// The plan's tables, computed on the device (plan creation and the
// allocation-free entry point share this kernel, so both use the same
// values): tw[m] = exp(-2 pi i m / nm); dc[m] = 1 / sinc^4(pi f / nm) at the
// signed frequency f of index m (the CIC deconvolution); with rank/order,
// the Hilbert rank of every deposit tile and its inverse (nt^3 tiles, by
// counting the smaller keys: for small tile counts only).  One workgroup.
__global__ __launch_bounds__(256) void k_plan_tables(int nm, int nt, int bits, float2* __restrict__ tw,
                                                     float* __restrict__ dc, unsigned* __restrict__ rank,
                                                     unsigned* __restrict__ order, unsigned* __restrict__ src) {
  for (int m = threadIdx.x; m < nm; m += blockDim.x) {
    double sn, cs;
    sincospi(2.0 * (double)m / (double)nm, &sn, &cs);
    tw[m] = make_float2((float)cs, (float)(-sn));
    const int f = m <= nm / 2 ? m : m - nm;
    const double xx = M_PI * (double)f / (double)nm;
    const double sc = f == 0 ? 1.0 : sin(xx) / xx;
    dc[m] = (float)(1.0 / (sc * sc * sc * sc));
  }
  if (!rank) return;
  const int nt3 = nt * nt * nt;
  for (int t = threadIdx.x; t < nt3; t += blockDim.x) {
    const uint64_t kt = hilbert_key3((uint32_t)(t % nt), (uint32_t)((t / nt) % nt), (uint32_t)(t / (nt * nt)), bits);
    unsigned r = 0;
    for (int u = 0; u < nt3; ++u)
      r += hilbert_key3((uint32_t)(u % nt), (uint32_t)((u / nt) % nt), (uint32_t)(u / (nt * nt)), bits) < kt ? 1u : 0u;
    rank[t] = r;
    order[r] = (unsigned)t;
  }
  __syncthreads();
  fill_tile_src(nt, rank, src, threadIdx.x, blockDim.x);
}

struct MeshOps {
  int nm;
  int x_rows;
  int yz_cols;
  int yz_threads;
  int z_cols;  // the per-item z pass's tiles
  int z_threads;
  int tile;
  // deposit[v][split]: workgroups of 256 (v = 0) or 512 (v = 1) threads;
  // split: dense tiles shared by several workgroups
  void (*deposit[2][2])(dim3, dim3, hipStream_t, const PartRec*, const unsigned*, CellMap, TileMap,
                        const SolveScalars*, long long, float*, long long, float*, DepositSplit);
  // deposit_pairs[v]: two tiles per workgroup (unsplit deposit only)
  void (*deposit_pairs[2])(dim3, dim3, hipStream_t, const PartRec*, CellMap, TileMap, const SolveScalars*, long long,
                           float*, long long, float*);
  // per-cell particle moments of the sorted records, one workgroup per tile
  void (*cell_moments)(dim3, dim3, hipStream_t, const PartRec*, const unsigned*, CellMap, TileMap,
                       const SolveScalars*, long long, double*);
  void (*x_fwd)(dim3, dim3, hipStream_t, const float*, float2*, const float2*, long long);
  void (*y_fwd)(dim3, dim3, hipStream_t, float2*, const float2*);
  void (*z_green)(dim3, dim3, hipStream_t, const float2*, float2*, float2*, const float2*, GreenParams);
  void (*z_green_col0)(dim3, dim3, hipStream_t, const float2*, float2*, float2*, const float2*, GreenParams);
  void (*z_green_persist)(dim3, dim3, hipStream_t, const float2*, float2*, float2*, const float2*, GreenParams);
  void (*y_fwd_persist)(dim3, dim3, hipStream_t, float2*, const float2*);
  void (*y_inv_persist)(dim3, dim3, hipStream_t, float2*, float2*, float2*, const float2*);
  void (*y_inv)(dim3, dim3, hipStream_t, float2*, float2*, float2*, const float2*);
  // the inverse y pass on narrower tiles (same grid argument); equal to y_inv
  // where the mesh has no narrower variant (has_narrow_y_inv = false)
  void (*y_inv_narrow)(dim3, dim3, hipStream_t, float2*, float2*, float2*, const float2*);
  bool has_narrow_y_inv;
  void (*x_inv)(dim3, dim3, hipStream_t, float2*, float2*, float2*, float2*, const float2*, int, float);
  // power spectrum back end
  void (*z_pk)(dim3, dim3, hipStream_t, const float2*, const float2*, const double*, double*, long long,
               unsigned long long*);
  void (*pk_col0)(dim3, dim3, hipStream_t, const float2*, const float2*, const double*, double*, long long, long long,
                  unsigned long long*);
  // initial conditions: generated modes through the inverse z pass, and the
  // inverse y and x passes with mode multipliers (second derivatives)
  void (*ic_z)(dim3, dim3, hipStream_t, IcParams, float2*, float2*, const float2*, int2);
  void (*y_inv_mul)(dim3, dim3, hipStream_t, float2*, float2*, float2*, const float2*, int3);
  void (*x_inv_mul)(dim3, dim3, hipStream_t, float2*, float2*, float2*, const float2*, int3);
  // ensembles: every stage for all boxes of this mesh side in one launch
  // (box b of the BoxArgs table on blockIdx.z); grids are per box, z = boxes
  void (*deposit_box[2])(dim3, dim3, hipStream_t, const BoxArgs*, TileMap);
  void (*x_fwd_box)(dim3, dim3, hipStream_t, const BoxArgs*, const float2*);
  void (*y_fwd_box)(dim3, dim3, hipStream_t, const BoxArgs*, const float2*);
  void (*z_green_box)(dim3, dim3, hipStream_t, const BoxArgs*, const float2*);
  void (*z_green_col0_box)(dim3, dim3, hipStream_t, const BoxArgs*, const float2*);
  void (*y_inv_box)(dim3, dim3, hipStream_t, const BoxArgs*, const float2*);
  void (*x_inv_box)(dim3, dim3, hipStream_t, const BoxArgs*, const float2*);
};

template <int N>
MeshOps make_ops() {
  constexpr int R = x_rows<N>();
  constexpr int C = YZCfg<N>::C;
  constexpr int TY = YZCfg<N>::T;
  constexpr int TD = deposit_tile(N);
  static_assert(N % TD == 0, "deposit tiles must cover the mesh");
  static_assert((long long)(N / TD) * (N / TD) * (N / TD) <= (1LL << 20),
                "the split deposit's extra-workgroup table keeps a tile rank in 20 bits");
  MeshOps o;
  o.nm = N;
  o.x_rows = R;
  o.yz_cols = C;
  o.yz_threads = TY;
  o.z_cols = YZCfg<N>::CZ;
  o.z_threads = YZCfg<N>::TZ;
  o.tile = TD;
#define GNG_DEPOSIT(V, NT, SPLIT)                                                                          \
  o.deposit[V][SPLIT] = [](dim3 g, dim3 b, hipStream_t s, const PartRec* srt, const unsigned* off, CellMap cm,     \
                           TileMap tm, const SolveScalars* sc, long long n, float* rho, long long rho_plane,        \
                           float* keep, DepositSplit sp) {                                                          \
    hipLaunchKernelGGL((k_deposit_tiles<TD, NT, SPLIT>), g, b, 0, s, srt, off, cm, tm, sc, n, rho, rho_plane, keep, \
                       sp);                                                                                         \
  };
  GNG_DEPOSIT(0, 256, false)
  GNG_DEPOSIT(1, 512, false)
  GNG_DEPOSIT(0, 256, true)
  GNG_DEPOSIT(1, 512, true)
#undef GNG_DEPOSIT
  o.deposit_pairs[0] = [](dim3 g, dim3 b, hipStream_t s, const PartRec* srt, CellMap cm, TileMap tm,
                          const SolveScalars* sc, long long n, float* rho, long long rho_plane, float* keep) {
    hipLaunchKernelGGL((k_deposit_pairs<TD, 256>), g, b, 0, s, srt, cm, tm, sc, n, rho, rho_plane, keep);
  };
  o.deposit_pairs[1] = [](dim3 g, dim3 b, hipStream_t s, const PartRec* srt, CellMap cm, TileMap tm,
                          const SolveScalars* sc, long long n, float* rho, long long rho_plane, float* keep) {
    hipLaunchKernelGGL((k_deposit_pairs<TD, 512>), g, b, 0, s, srt, cm, tm, sc, n, rho, rho_plane, keep);
  };
  o.cell_moments = [](dim3 g, dim3 b, hipStream_t s, const PartRec* srt, const unsigned* off, CellMap cm,
                      TileMap tm, const SolveScalars* sc, long long n, double* mom) {
    hipLaunchKernelGGL((k_cell_moments<TD>), g, b, 0, s, srt, off, cm, tm, sc, n, mom);
  };

  o.x_fwd = [](dim3 g, dim3 b, hipStream_t s, const float* rho, float2* P, const float2* tw, long long rho_plane) {
    hipLaunchKernelGGL((k_fft_x_fwd<N, R>), g, b, 0, s, rho, P, tw, rho_plane);
  };
  o.y_fwd = [](dim3 g, dim3 b, hipStream_t s, float2* P, const float2* tw) {
    hipLaunchKernelGGL((k_fft_y_fwd<N, C, TY>), g, b, 0, s, P, tw);
  };
  o.z_green = [](dim3 g, dim3 b, hipStream_t s, const float2* P, float2* A, float2* B, const float2* tw,
                 GreenParams gp) {
    hipLaunchKernelGGL((k_fft_z_green<N, YZCfg<N>::CZ, YZCfg<N>::TZ>), g, b, 0, s, P, A, B, tw, gp);
  };
  o.z_green_col0 = [](dim3 g, dim3 b, hipStream_t s, const float2* P, float2* A, float2* B, const float2* tw,
                      GreenParams gp) { hipLaunchKernelGGL((k_fft_z_green_col0<N>), g, b, 0, s, P, A, B, tw, gp); };
  o.z_green_persist = [](dim3 g, dim3 b, hipStream_t s, const float2* P, float2* A, float2* B, const float2* tw,
                         GreenParams gp) {
    hipLaunchKernelGGL((k_fft_z_green_persist<N, C, TY>), g, b, 0, s, P, A, B, tw, gp);
  };
  o.y_fwd_persist = [](dim3 g, dim3 b, hipStream_t s, float2* P, const float2* tw) {
    hipLaunchKernelGGL((k_fft_y_fwd_persist<N, C, TY>), g, b, 0, s, P, tw);
  };
  o.y_inv_persist = [](dim3 g, dim3 b, hipStream_t s, float2* A, float2* B, float2* P, const float2* tw) {
    hipLaunchKernelGGL((k_fft_y_inv_persist<N, C, TY>), g, b, 0, s, A, B, P, tw);
  };
  o.y_inv = [](dim3 g, dim3 b, hipStream_t s, float2* A, float2* B, float2* P, const float2* tw) {
    if constexpr (YZCfg<N>::YSPLIT) {
      // full-line tiles of the z pass's width; the caller's grid counts C-wide tiles
      constexpr int CS = YZCfg<N>::CZ, TS = YZCfg<N>::TZ;
      const dim3 gs((unsigned)(N / 2 / CS), g.y);
      hipLaunchKernelGGL((k_fft_y_inv<N, CS, TS, false, 1>), gs, dim3(TS), 0, s, A, B, P, tw, make_int3(0, 0, 0));
      hipLaunchKernelGGL((k_fft_y_inv<N, CS, TS, false, 2>), gs, dim3(TS), 0, s, A, B, P, tw, make_int3(0, 0, 0));
    } else {
      hipLaunchKernelGGL((k_fft_y_inv<N, C, TY>), g, b, 0, s, A, B, P, tw, make_int3(0, 0, 0));
    }
  };
  o.has_narrow_y_inv = YZCfg<N>::CI != C;
  if constexpr (YZCfg<N>::CI != C) {
    o.y_inv_narrow = [](dim3 g, dim3 b, hipStream_t s, float2* A, float2* B, float2* P, const float2* tw) {
      // narrower tiles than the caller's grid counts
      constexpr int CI = YZCfg<N>::CI, TI = YZCfg<N>::TI;
      (void)b;
      hipLaunchKernelGGL((k_fft_y_inv<N, CI, TI>), dim3((unsigned)(N / 2 / CI), g.y), dim3(TI), 0, s, A, B, P, tw,
                         make_int3(0, 0, 0));
    };
  } else {
    o.y_inv_narrow = o.y_inv;
  }
  o.x_inv = [](dim3 g, dim3 b, hipStream_t s, float2* A, float2* P, float2* B, float2* pot, const float2* tw,
                int field0, float pot_scale) {
    hipLaunchKernelGGL((k_fft_x_inv<N, R>), g, b, 0, s, A, P, B, pot, tw, field0, pot_scale, make_int3(0, 0, 0));
  };
  o.z_pk = [](dim3 g, dim3 b, hipStream_t s, const float2* P, const float2* tw, const double* w, double* part,
              long long nblk, unsigned long long* modes) {
    hipLaunchKernelGGL((k_fft_z_pk<N, C, TY>), g, b, 0, s, P, tw, w, part, nblk, modes);
  };
  o.pk_col0 = [](dim3 g, dim3 b, hipStream_t s, const float2* P, const float2* tw, const double* w, double* part,
                 long long nblk, long long base, unsigned long long* modes) {
    hipLaunchKernelGGL((k_pk_col0<N>), g, b, 0, s, P, tw, w, part, nblk, base, modes);
  };
  o.ic_z = [](dim3 g, dim3 b, hipStream_t s, IcParams ip, float2* A, float2* B, const float2* tw, int2 zc) {
    hipLaunchKernelGGL((k_ic_z<N, C, TY>), g, b, 0, s, ip, A, B, tw, zc);
  };
  o.y_inv_mul = [](dim3 g, dim3 b, hipStream_t s, float2* A, float2* B, float2* P, const float2* tw, int3 codes) {
    hipLaunchKernelGGL((k_fft_y_inv<N, C, TY, true>), g, b, 0, s, A, B, P, tw, codes);
  };
  o.x_inv_mul = [](dim3 g, dim3 b, hipStream_t s, float2* A, float2* P, float2* B, const float2* tw, int3 codes) {
    hipLaunchKernelGGL((k_fft_x_inv<N, R, true>), g, b, 0, s, A, P, B, nullptr, tw, 0, 1.0f, codes);
  };
  o.deposit_box[0] = [](dim3 g, dim3 b, hipStream_t s, const BoxArgs* bx, TileMap tm) {
    hipLaunchKernelGGL((k_deposit_tiles_box<TD, 256>), g, b, 0, s, bx, tm);
  };
  o.deposit_box[1] = [](dim3 g, dim3 b, hipStream_t s, const BoxArgs* bx, TileMap tm) {
    hipLaunchKernelGGL((k_deposit_tiles_box<TD, 512>), g, b, 0, s, bx, tm);
  };
  o.x_fwd_box = [](dim3 g, dim3 b, hipStream_t s, const BoxArgs* bx, const float2* tw) {
    hipLaunchKernelGGL((k_fft_x_fwd_box<N, R>), g, b, 0, s, bx, tw);
  };
  o.y_fwd_box = [](dim3 g, dim3 b, hipStream_t s, const BoxArgs* bx, const float2* tw) {
    hipLaunchKernelGGL((k_fft_y_fwd_box<N, C, TY>), g, b, 0, s, bx, tw);
  };
  o.z_green_box = [](dim3 g, dim3 b, hipStream_t s, const BoxArgs* bx, const float2* tw) {
    hipLaunchKernelGGL((k_fft_z_green_box<N, YZCfg<N>::CZ, YZCfg<N>::TZ>), g, b, 0, s, bx, tw);
  };
  o.z_green_col0_box = [](dim3 g, dim3 b, hipStream_t s, const BoxArgs* bx, const float2* tw) {
    hipLaunchKernelGGL((k_fft_z_green_col0_box<N>), g, b, 0, s, bx, tw);
  };
  o.y_inv_box = [](dim3 g, dim3 b, hipStream_t s, const BoxArgs* bx, const float2* tw) {
    if constexpr (YZCfg<N>::YSPLIT) {
      constexpr int CS = YZCfg<N>::CZ, TS = YZCfg<N>::TZ;
      const dim3 gs((unsigned)(N / 2 / CS), g.y, g.z);
      hipLaunchKernelGGL((k_fft_y_inv_box<N, CS, TS, 1>), gs, dim3(TS), 0, s, bx, tw);
      hipLaunchKernelGGL((k_fft_y_inv_box<N, CS, TS, 2>), gs, dim3(TS), 0, s, bx, tw);
    } else {
      hipLaunchKernelGGL((k_fft_y_inv_box<N, C, TY>), g, b, 0, s, bx, tw);
    }
  };
  o.x_inv_box = [](dim3 g, dim3 b, hipStream_t s, const BoxArgs* bx, const float2* tw) {
    hipLaunchKernelGGL((k_fft_x_inv_box<N, R>), g, b, 0, s, bx, tw);
  };
  return o;
}

const std::vector<MeshOps>& ops_table() {
  static const std::vector<MeshOps> t = {
      make_ops<4>(),   make_ops<8>(),   make_ops<16>(),  make_ops<32>(),  make_ops<64>(),
      make_ops<128>(), make_ops<256>(), make_ops<512>(), make_ops<1024>(), make_ops<12>(),
      make_ops<24>(),  make_ops<48>(),  make_ops<96>(),  make_ops<192>(), make_ops<384>(),
  };
  return t;
}

const MeshOps* find_ops(int nm) {
  for (const MeshOps& o : ops_table())
    if (o.nm == nm) return &o;
  return nullptr;
}

}  // namespace

struct gng_pm_plan {
  gng_pm_config cfg{};
  const MeshOps* ops = nullptr;
  int device = 0;
  int cus = 0;  // compute units of the device (persistent launches)
  Tuning tune{};
  std::mutex mu;

  long long nm3 = 0;
  float2* d_P = nullptr;  // three nm^3/2 complex meshes
  float2* d_A = nullptr;
  float2* d_B = nullptr;
  float* d_rho = nullptr;  // kept density (only with keep_density)
  bool keep_density = false;
  float* d_pot = nullptr;  // mesh potential (only with keep_potential)
  bool keep_potential = false;
  bool potential_valid = false;
  float2* d_tw = nullptr;  // nm-th roots of unity
  float* d_deconv = nullptr;
  SolveScalars* d_sc = nullptr;

  // deposit buckets: nb = tiles * 8 counts, nb + 1 offsets, scan partials
  TileMap tm{};
  long long nb = 0;
  int scan_blocks = 0;
  unsigned* d_count = nullptr;
  unsigned* d_offset = nullptr;
  unsigned* d_partial = nullptr;
  unsigned* d_tile_rank = nullptr;   // tile -> Hilbert rank
  unsigned* d_tile_order = nullptr;  // Hilbert rank -> tile
  unsigned* d_tile_src = nullptr;    // Hilbert rank -> its 8 source tiles' first buckets
  unsigned* d_tile_ranges = nullptr; // Hilbert rank -> its flattened bucket ranges (per solve)

  // split deposit of heavy tiles: per tile rank its slice count, their
  // exclusive scan, its scratch slot; per workgroup its tile rank; the
  // scratch tiles and finished-slice counters (zero between solves)
  unsigned* d_slices = nullptr;
  unsigned* d_slot = nullptr;
  unsigned* d_extra = nullptr;
  unsigned long long* d_split_scratch = nullptr;
  unsigned* d_split_done = nullptr;
  long long split_slots = 0;

  // per-particle work buffers (grown on demand)
  long long work_cap = 0;
  PartRec* d_sorted = nullptr;
  unsigned* d_perm = nullptr;  // original index of each sorted slot
  unsigned* d_bucket = nullptr;
  unsigned* d_rank = nullptr;

  // particle staging for the host entry points
  long long cap = 0;
  float* d_part = nullptr;  // x, y, z, m, ax, ay, az (7 * cap floats)
  long long cap64 = 0;
  double* d_part64 = nullptr;  // pos (3 n), mass (n), out (3 k) of the f64 entry points: 7 * cap64 doubles
  // Ordering across streams: every entry point records done_ev on the stream
  // it used; the next entry point on another stream waits for it first.
  hipEvent_t done_ev = nullptr;
  hipStream_t done_stream = nullptr;
  bool done_valid = false;
  // Host staging of the f64 host entry points: per worker thread a stream,
  // two pinned slots and their events (allocated on first use).
  int pipe_threads = 0;
  hipStream_t pipe_stream[8] = {};
  hipEvent_t pipe_ev[8][2] = {};
  unsigned char* pipe_slot[8][2] = {};
  long long idx_cap = 0;
  long long* d_idx = nullptr;
  hipStream_t stream = nullptr;

  long long last_n = 0;
  bool density_valid = false;
  long long sorted_n = -1;  // particles of the last counting sort (its records are in d_sorted)

  // power spectrum work space (allocated on first use)
  double* d_pk_w = nullptr;     // 1 / sinc(freq(i) / nm)^2, nm values
  double* d_pk_part = nullptr;  // per-workgroup shell sums
  long long pk_part_cap = 0;
  unsigned long long* d_pk_modes = nullptr;
  double* d_pk_msum = nullptr;  // mass partials
  double* d_pk_out = nullptr;   // shell sums times (V / M_tot)^2
  // initial conditions: the per-|n|^2 amplitude table (allocated on first use)
  double* d_ic_sigma = nullptr;
  long long ic_sigma_cap = 0;
  float* d_ic_mask = nullptr;  // 1 per axis index, 0 at the Nyquist index (second order)
  // Stage timing: every solve records kStages+1 events into the next slot of
  // a ring; nothing synchronises until gng_pm_stage_times() reads them.
  bool timing = false;
  long long timed_solves = 0;
  hipEvent_t ev[kTimingSlots][kStages + 1] = {};
};

namespace {

int set_device(gng_pm_plan* p) {
  GNG_HIP(hipSetDevice(p->device));
  return GNG_OK;
}

// Stream order for one entry point's work on a plan: construction makes s
// wait for the previous entry point's work when that ran on another stream;
// destruction records the point the next entry point waits for.  Construct it
// after set_device, under the plan's lock.
struct StreamOrder {
  gng_pm_plan* p;
  hipStream_t s;
  StreamOrder(gng_pm_plan* p_, hipStream_t s_) : p(p_), s(s_) {
    if (p->done_valid && p->done_stream != s) (void)hipStreamWaitEvent(s, p->done_ev, 0);
  }
  ~StreamOrder() {
    if (hipEventRecord(p->done_ev, s) == hipSuccess) {
      p->done_valid = true;
      p->done_stream = s;
    }
  }
};

void free_work(gng_pm_plan* p) {
  for (void* q : {(void*)p->d_slices, (void*)p->d_slot, (void*)p->d_extra, (void*)p->d_split_scratch,
                  (void*)p->d_split_done})
    (void)hipFree(q);
  p->d_slices = p->d_slot = p->d_extra = p->d_split_done = nullptr;
  p->d_split_scratch = nullptr;
  p->split_slots = 0;
  (void)hipFree(p->d_sorted);
  (void)hipFree(p->d_perm);
  (void)hipFree(p->d_bucket);
  (void)hipFree(p->d_rank);
  p->d_sorted = nullptr;
  p->d_perm = nullptr;
  p->d_bucket = p->d_rank = nullptr;
  p->work_cap = 0;
}

void free_plan(gng_pm_plan* p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  for (void* q : {(void*)p->d_P, (void*)p->d_A, (void*)p->d_B, (void*)p->d_rho, (void*)p->d_pot, (void*)p->d_tw, (void*)p->d_deconv,
                  (void*)p->d_sc, (void*)p->d_count, (void*)p->d_offset, (void*)p->d_partial, (void*)p->d_part,
                  (void*)p->d_idx, (void*)p->d_tile_rank, (void*)p->d_tile_order, (void*)p->d_tile_src, (void*)p->d_tile_ranges,
                  (void*)p->d_pk_w,
                  (void*)p->d_pk_part, (void*)p->d_pk_modes, (void*)p->d_pk_msum, (void*)p->d_pk_out,
                  (void*)p->d_ic_sigma, (void*)p->d_ic_mask, (void*)p->d_part64})
    (void)hipFree(q);
  if (p->done_ev) (void)hipEventDestroy(p->done_ev);
  for (int t = 0; t < 8; ++t) {
    for (int k = 0; k < 2; ++k) {
      if (p->pipe_ev[t][k]) (void)hipEventDestroy(p->pipe_ev[t][k]);
      if (p->pipe_slot[t][k]) (void)hipHostFree(p->pipe_slot[t][k]);
    }
    if (p->pipe_stream[t]) (void)hipStreamDestroy(p->pipe_stream[t]);
  }
  free_work(p);
  for (auto& slot : p->ev)
    for (auto& e : slot)
      if (e) (void)hipEventDestroy(e);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

CellMap cell_map(const gng_pm_plan* p) {
  CellMap cm;
  cm.inv_cell = (double)p->cfg.grid_size / p->cfg.box_size;  // cic.rs:262
  cm.nm = p->cfg.grid_size;
  return cm;
}

// Particle inputs of a solve: FP32 or FP64 positions at any element stride,
// FP32 or FP64 masses (stride 1); accelerations out as FP32 or FP64.
struct PartIn {
  int pos64 = 0;
  const void* x = nullptr;
  const void* y = nullptr;
  const void* z = nullptr;
  long long stride = 1;
  int mass64 = 0;
  const void* m = nullptr;
};
struct AccOut {
  int f64 = 0;
  void* x = nullptr;
  void* y = nullptr;
  void* z = nullptr;
  long long stride = 1;
};

PartIn soa32(const float* x, const float* y, const float* z, const float* m) {
  PartIn in;
  in.x = x;
  in.y = y;
  in.z = z;
  in.m = m;
  return in;
}
AccOut acc32(float* x, float* y, float* z) {
  AccOut o;
  o.x = x;
  o.y = y;
  o.z = z;
  return o;
}
PartIn offset_in(PartIn in, long long o) {
  const long long b = o * in.stride * (in.pos64 ? 8 : 4);
  in.x = (const char*)in.x + b;
  in.y = (const char*)in.y + b;
  in.z = (const char*)in.z + b;
  in.m = (const char*)in.m + o * (in.mass64 ? 8 : 4);
  return in;
}
AccOut offset_out(AccOut out, long long o) {
  const long long b = o * out.stride * (out.f64 ? 8 : 4);
  out.x = (char*)out.x + b;
  out.y = (char*)out.y + b;
  out.z = (char*)out.z + b;
  return out;
}

template <typename T>
PosSrc<T> pos_src(const PartIn& in) {
  return PosSrc<T>{(const T*)in.x, (const T*)in.y, (const T*)in.z, in.stride};
}
template <typename T>
AccDst<T> acc_dst(const AccOut& o) {
  return AccDst<T>{(T*)o.x, (T*)o.y, (T*)o.z, o.stride};
}

// The particle kernels for each combination of input types.
template <typename T, typename M>
void launch_bucket_t(const PartIn& in, long long n, const CellMap& cm, const TileMap& tm, unsigned* count,
                     unsigned* bucket, unsigned* rank, SolveScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL((k_bucket_count<T, M>), dim3((unsigned)((n + 256 * kBucketPer - 1) / (256 * kBucketPer))),
                     dim3(256), 0, s, pos_src<T>(in), (const M*)in.m, n, cm, tm, count, bucket, rank, sc);
}
void launch_bucket(const PartIn& in, long long n, const CellMap& cm, const TileMap& tm, unsigned* count,
                   unsigned* bucket, unsigned* rank, SolveScalars* sc, hipStream_t s) {
  if (in.pos64)
    (in.mass64 ? launch_bucket_t<double, double> : launch_bucket_t<double, float>)(in, n, cm, tm, count, bucket,
                                                                                     rank, sc, s);
  else
    (in.mass64 ? launch_bucket_t<float, double> : launch_bucket_t<float, float>)(in, n, cm, tm, count, bucket, rank,
                                                                                   sc, s);
}
template <typename T, typename M>
void launch_scatter_t(const PartIn& in, long long n, const CellMap& cm, const unsigned* bucket,
                      const unsigned* rank, const unsigned* offset, PartRec* sorted, unsigned* perm,
                      SolveScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL((k_scatter<T, M>), dim3((unsigned)((n + 256 * kScatterPer - 1) / (256 * kScatterPer))), dim3(256),
                     0, s, pos_src<T>(in),
                     (const M*)in.m, n, cm, bucket, rank, offset, sorted, perm, sc);
}
void launch_scatter(const PartIn& in, long long n, const CellMap& cm, const unsigned* bucket, const unsigned* rank,
                    const unsigned* offset, PartRec* sorted, unsigned* perm, SolveScalars* sc, hipStream_t s) {
  if (in.pos64)
    (in.mass64 ? launch_scatter_t<double, double> : launch_scatter_t<double, float>)(in, n, cm, bucket, rank, offset,
                                                                                       sorted, perm, sc, s);
  else
    (in.mass64 ? launch_scatter_t<float, double> : launch_scatter_t<float, float>)(in, n, cm, bucket, rank, offset,
                                                                                     sorted, perm, sc, s);
}
template <typename T, typename O>
void launch_interp_t(const PartIn& in, const long long* idx, long long k, const CellMap& cm, long long fplane,
                     const float* fx, const float* fy, const float* fz, const AccOut& out, hipStream_t s) {
  hipLaunchKernelGGL((k_interp<T, O>), dim3((unsigned)((k + 255) / 256)), dim3(256), 0, s, pos_src<T>(in), idx, k,
                     cm, fplane, fx, fy, fz, acc_dst<O>(out));
}
void launch_interp(const PartIn& in, const long long* idx, long long k, const CellMap& cm, long long fplane,
                   const float* fx, const float* fy, const float* fz, const AccOut& out, hipStream_t s) {
  if (in.pos64)
    (out.f64 ? launch_interp_t<double, double> : launch_interp_t<double, float>)(in, idx, k, cm, fplane, fx, fy, fz,
                                                                                   out, s);
  else
    (out.f64 ? launch_interp_t<float, double> : launch_interp_t<float, float>)(in, idx, k, cm, fplane, fx, fy, fz,
                                                                                 out, s);
}

GreenParams green_params(const gng_pm_plan* p, double g) {
  const double L = p->cfg.box_size;
  const double kf = 2.0 * M_PI / L;
  double a = 0.0;
  if (p->cfg.r_split > 0.0) a += 0.5 * kf * kf * p->cfg.r_split * p->cfg.r_split;
  if (p->cfg.plummer_eps > 0.0) a += kf * kf * p->cfg.plummer_eps * p->cfg.plummer_eps;
  GreenParams gp;
  gp.c64 = 2.0 * g / (L * L);
  gp.a64 = a;
  gp.c = (float)gp.c64;
  gp.a = (float)a;
  gp.deconv = p->cfg.deconvolve ? p->d_deconv : nullptr;
  return gp;
}

inline void mark(gng_pm_plan* p, int i, hipStream_t s) {
  if (p->timing) (void)hipEventRecord(p->ev[p->timed_solves % kTimingSlots][i], s);
}

int ensure_work(gng_pm_plan* p, long long n) {
  if (n <= p->work_cap) return GNG_OK;
  free_work(p);
  const size_t m = (size_t)n;
  GNG_HIP(hipMalloc(&p->d_sorted, sizeof(PartRec) * m));
  GNG_HIP(hipMalloc(&p->d_perm, sizeof(unsigned) * m));
  GNG_HIP(hipMalloc(&p->d_bucket, sizeof(unsigned) * m));
  GNG_HIP(hipMalloc(&p->d_rank, sizeof(unsigned) * m));
  // split deposit: extra workgroups (and scratch tiles) for at most 2 n /
  // kSliceItems slices beyond one per tile
  const long long ntiles = p->nb / 8;
  const long long slots = 2 * n / (long long)kSliceItems + 16;
  const size_t t3 = (size_t)p->tm.T * p->tm.T * p->tm.T;
  GNG_HIP(hipMalloc(&p->d_slices, sizeof(unsigned) * (size_t)ntiles));
  GNG_HIP(hipMalloc(&p->d_slot, sizeof(unsigned) * (size_t)ntiles));
  GNG_HIP(hipMalloc(&p->d_extra, sizeof(unsigned) * (size_t)slots));
  GNG_HIP(hipMalloc(&p->d_split_scratch, sizeof(unsigned long long) * t3 * (size_t)slots));
  GNG_HIP(hipMalloc(&p->d_split_done, sizeof(unsigned) * (size_t)slots));
  // the scratch tiles and counters start at zero; every solve leaves them so
  GNG_HIP(hipMemset(p->d_split_scratch, 0, sizeof(unsigned long long) * t3 * (size_t)slots));
  GNG_HIP(hipMemset(p->d_split_done, 0, sizeof(unsigned) * (size_t)slots));
  p->split_slots = slots;
  p->work_cap = n;
  return GNG_OK;
}

// Deposit and the forward x and y transforms (shared by the force solve and
// the power spectrum); the mass mesh is in P as the half spectrum after y.
// The counting sort of the deposit: bucket counts, their scan, and the
// scatter of the particles into bucket order as cell records (perm, when not
// null, receives the original index of every sorted slot).  stage_marks
// records the bucket stage's timing event (a solve's pipeline only).
void run_sort(gng_pm_plan* p, const PartIn& in, long long n, unsigned* perm, bool stage_marks, hipStream_t s) {
  const CellMap cm = cell_map(p);
  {
    const long long n16 = (p->nb + 3) / 4;  // counts padded to 16-byte words
    hipLaunchKernelGGL(k_sort_reset, dim3((unsigned)((n16 + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<uint4*>(p->d_count), n16, p->d_sc);
  }
  if (n > 0) launch_bucket(in, n, cm, p->tm, p->d_count, p->d_bucket, p->d_rank, p->d_sc, s);
  if (stage_marks) mark(p, 1, s);
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)p->scan_blocks), dim3(256), 0, s, p->d_count, p->d_offset,
                     p->d_partial, p->nb);
  hipLaunchKernelGGL(k_scan_add_prefix, dim3((unsigned)p->scan_blocks), dim3(256), 0, s, p->d_offset, p->d_partial,
                     p->scan_blocks, p->nb);
  if (n > 0) launch_scatter(in, n, cm, p->d_bucket, p->d_rank, p->d_offset, p->d_sorted, perm, p->d_sc, s);
  p->sorted_n = n;
}

int run_front(gng_pm_plan* p, const PartIn& in, long long n, hipStream_t s) {
  int rc = ensure_work(p, n > 0 ? n : 1);
  if (rc) return rc;
  const int nm = p->cfg.grid_size;
  const int H = nm / 2;
  const MeshOps& o = *p->ops;
  const CellMap cm = cell_map(p);
  // the shifted mass mesh goes into P as real rows (float z-plane stride
  // fplane); a kept density is also written unshifted and unpadded
  float* rho = reinterpret_cast<float*>(p->d_P);
  const long long cplane = mesh_plane(nm), fplane = 2 * cplane;
  const long long rho_plane = fplane;

  mark(p, 0, s);
  run_sort(p, in, n, p->tune.interp_sorted ? p->d_perm : nullptr, true, s);
  mark(p, 2, s);
  const unsigned ntiles = (unsigned)p->tm.nt * p->tm.nt * p->tm.nt;
  hipLaunchKernelGGL(k_tile_ranges, dim3((ntiles + 255) / 256), dim3(256), 0, s, p->d_offset, p->tm, p->d_sc);
  DepositSplit sp{};
  unsigned deposit_groups = ntiles;
  if (p->split_slots > 0 && p->tune.deposit_split) {
    // heavy tiles over several workgroups: slices per tile, and the table of
    // the extra workgroups past the ntiles primary ones
    hipLaunchKernelGGL(k_deposit_plan, dim3((ntiles + 255) / 256), dim3(256), 0, s, p->d_offset, p->tm, p->d_slices,
                       p->d_slot, p->d_extra, p->d_sc, (unsigned)p->split_slots);
    sp.slices = p->d_slices;
    sp.slot = p->d_slot;
    sp.extra = p->d_extra;
    sp.sc = p->d_sc;
    sp.ntiles = ntiles;
    sp.slots = (unsigned)p->split_slots;
    sp.scratch = p->d_split_scratch;
    sp.done = p->d_split_done;
    deposit_groups = ntiles + (unsigned)p->split_slots;
  }
  if (!sp.slices && p->tune.deposit_pair) {
    o.deposit_pairs[p->tune.deposit_wide](dim3((ntiles + 1u) / 2u), dim3(p->tune.deposit_wide ? 512 : 256), s,
                                          p->d_sorted, cm, p->tm, p->d_sc, n, rho, rho_plane,
                                          p->keep_density ? p->d_rho : nullptr);
  } else {
    o.deposit[p->tune.deposit_wide][sp.slices != nullptr](dim3(deposit_groups),
                                                          dim3(p->tune.deposit_wide ? 512 : 256), s, p->d_sorted,
                                                          p->d_offset, cm, p->tm, p->d_sc, n, rho, rho_plane,
                                                          p->keep_density ? p->d_rho : nullptr, sp);
  }
  mark(p, 3, s);
  const long long rows = (long long)nm * nm;
  const dim3 yz_grid((unsigned)(H / o.yz_cols), (unsigned)nm);
  const dim3 yz_block((unsigned)o.yz_threads);
  const int cz = p->tune.chunk_fwd;
  if (cz > 0 && !p->tune.y_fwd_persistent) {
    // chunks of cz planes: each y pass reads the chunk its x pass just wrote
    // (both passes touch only their own planes; the pointers are offset)
    const long long pl = mesh_plane(nm);
    for (int z0 = 0; z0 < nm; z0 += cz) {
      o.x_fwd(dim3((unsigned)((long long)cz * nm / o.x_rows)), dim3(256), s, rho + rho_plane * z0, p->d_P + pl * z0,
              p->d_tw, rho_plane / 2);
      o.y_fwd(dim3((unsigned)(H / o.yz_cols), (unsigned)cz), yz_block, s, p->d_P + pl * z0, p->d_tw);
    }
    // the two stages share one timing slot (the x stage's)
    mark(p, 4, s);
    mark(p, 5, s);
    return GNG_OK;
  }
  o.x_fwd(dim3((unsigned)(rows / o.x_rows)), dim3(256), s, rho, p->d_P, p->d_tw, rho_plane / 2);
  mark(p, 4, s);
  if (p->tune.y_fwd_persistent && p->cus > 0)
    o.y_fwd_persist(dim3((unsigned)p->cus), yz_block, s, p->d_P, p->d_tw);
  else
    o.y_fwd(yz_grid, yz_block, s, p->d_P, p->d_tw);
  mark(p, 5, s);
  return GNG_OK;
}

// The whole pipeline on device pointers.
int run_pipeline(gng_pm_plan* p, const PartIn& in, long long n, const long long* idx, long long k, double g,
                 const AccOut& out, hipStream_t s) {
  int rc = run_front(p, in, n, s);
  if (rc) return rc;
  const int nm = p->cfg.grid_size;
  const int H = nm / 2;
  const MeshOps& o = *p->ops;
  const CellMap cm = cell_map(p);
  const long long rows = (long long)nm * nm;
  const long long fplane = 2 * mesh_plane(nm);
  const dim3 yz_grid((unsigned)(H / o.yz_cols), (unsigned)nm);
  const dim3 yz_block((unsigned)o.yz_threads);
  const GreenParams gp = green_params(p, g);
  {
    // persistent z pass (one workgroup per CU) or one workgroup per item
    if (p->tune.z_persistent && p->cus > 0)
      o.z_green_persist(dim3((unsigned)p->cus), yz_block, s, p->d_P, p->d_A, p->d_B, p->d_tw, gp);
    else
      o.z_green(dim3((unsigned)(H / o.z_cols), (unsigned)nm), dim3((unsigned)o.z_threads), s, p->d_P, p->d_A, p->d_B,
                p->d_tw, gp);
  }
  o.z_green_col0(dim3((unsigned)(H + 1)), dim3(256), s, p->d_P, p->d_A, p->d_B, p->d_tw, gp);
  mark(p, 6, s);
  const auto y_inv = p->tune.y_inv_narrow ? o.y_inv_narrow : o.y_inv;
  const int cz = p->tune.chunk_inv;
  if (cz > 0 && !p->tune.y_inv_persistent && !p->keep_potential) {
    // chunks of cz planes: each x pass reads the three fields its y pass just wrote
    const long long pl = mesh_plane(nm);
    for (int z0 = 0; z0 < nm; z0 += cz) {
      y_inv(dim3((unsigned)(H / o.yz_cols), (unsigned)cz), yz_block, s, p->d_A + pl * z0, p->d_B + pl * z0,
            p->d_P + pl * z0, p->d_tw);
      o.x_inv(dim3((unsigned)((long long)cz * nm / o.x_rows), 3), dim3(256), s, p->d_A + pl * z0, p->d_P + pl * z0,
              p->d_B + pl * z0, nullptr, p->d_tw, 0, 1.0f);
    }
    // the two stages share one timing slot (the y stage's)
    mark(p, 7, s);
    mark(p, 8, s);
  } else {
  if (p->tune.y_inv_persistent && p->cus > 0)
    o.y_inv_persist(dim3((unsigned)p->cus), yz_block, s, p->d_A, p->d_B, p->d_P, p->d_tw);
  else
    y_inv(yz_grid, yz_block, s, p->d_A, p->d_B, p->d_P, p->d_tw);
  mark(p, 7, s);
  if (p->keep_potential)
    // the potential reads A before the force pass below overwrites it
    o.x_inv(dim3((unsigned)(rows / o.x_rows), 1), dim3(256), s, p->d_A, p->d_P, p->d_B,
            reinterpret_cast<float2*>(p->d_pot), p->d_tw, 3, (float)(p->cfg.box_size / (2.0 * M_PI)));
  o.x_inv(dim3((unsigned)(rows / o.x_rows), 3), dim3(256), s, p->d_A, p->d_P, p->d_B, nullptr, p->d_tw, 0, 1.0f);
  mark(p, 8, s);
  }
  const float* fx = reinterpret_cast<const float*>(p->d_A);
  const float* fy = reinterpret_cast<const float*>(p->d_P);
  const float* fz = reinterpret_cast<const float*>(p->d_B);
  if (k > 0 && !idx && p->tune.interp_sorted) {
    // every particle: walk the deposit's sorted copy (local mesh reads whatever the caller's order)
    const dim3 gi((unsigned)((k + 255) / 256));
    if (out.f64)
      hipLaunchKernelGGL(k_interp_sorted<double>, gi, dim3(256), 0, s, p->d_sorted, p->d_perm, k, cm, fplane, fx, fy,
                         fz, acc_dst<double>(out));
    else
      hipLaunchKernelGGL(k_interp_sorted<float>, gi, dim3(256), 0, s, p->d_sorted, p->d_perm, k, cm, fplane, fx, fy,
                         fz, acc_dst<float>(out));
  } else if (k > 0) {
    launch_interp(in, idx, k, cm, fplane, fx, fy, fz, out, s);
  }
  mark(p, 9, s);
  GNG_HIP(hipGetLastError());
  if (p->timing) ++p->timed_solves;
  p->last_n = n;
  p->density_valid = p->keep_density;
  p->potential_valid = p->keep_potential;
  return GNG_OK;
}

// Average per-stage time (ms[0..kStages-1]) and whole-solve time
// (ms[kStages]) over the solves still held in the timing ring.
int read_timing(gng_pm_plan* p, float* ms) {
  const long long held = p->timed_solves < kTimingSlots ? p->timed_solves : kTimingSlots;
  for (int i = 0; i <= kStages; ++i) ms[i] = 0.0f;
  if (held == 0) return GNG_OK;
  for (long long j = 0; j < held; ++j) {
    hipEvent_t* ev = p->ev[(p->timed_solves - 1 - j) % kTimingSlots];
    GNG_HIP(hipEventSynchronize(ev[kStages]));
    for (int i = 0; i < kStages; ++i) {
      float t = 0.0f;
      GNG_HIP(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
      ms[i] += t;
    }
    float tot = 0.0f;
    GNG_HIP(hipEventElapsedTime(&tot, ev[0], ev[kStages]));
    ms[kStages] += tot;
  }
  for (int i = 0; i <= kStages; ++i) ms[i] /= (float)held;
  return GNG_OK;
}

int ensure_idx(gng_pm_plan* p, long long k) {
  if (k > p->idx_cap) {
    (void)hipFree(p->d_idx);
    p->d_idx = nullptr;
    p->idx_cap = 0;
    GNG_HIP(hipMalloc(&p->d_idx, sizeof(long long) * (size_t)k));
    p->idx_cap = k;
  }
  return GNG_OK;
}

int ensure_particles(gng_pm_plan* p, long long n, long long k, bool need_idx) {
  const long long need = n > k ? n : k;
  if (need > p->cap) {
    (void)hipFree(p->d_part);
    p->d_part = nullptr;
    p->cap = 0;
    GNG_HIP(hipMalloc(&p->d_part, sizeof(float) * 7 * (size_t)need));
    p->cap = need;
  }
  return need_idx ? ensure_idx(p, k) : GNG_OK;
}

// ---- host staging: pageable caller memory <-> device through pinned slots ----
//
// The f64 host entry points move the caller's pageable arrays with
// pipe_threads host threads.  Each owns a stream and two pinned slots of
// kStageBytes: it fills a slot (copying, or converting f64 masses to f32),
// starts the slot's DMA on its stream, and reuses the slot once that DMA's
// event has completed.  Downloads run the other way (DMA into a slot, then
// the thread converts f32 accelerations to f64 into the caller's array), two
// slots in flight per thread.  The compute stream waits for every upload
// stream, and every download stream waits for the compute stream.
constexpr size_t kStageBytes = 4u << 20;

int ensure_pipe(gng_pm_plan* p) {
  if (p->pipe_threads) return GNG_OK;
  unsigned hw = std::thread::hardware_concurrency();
  const int T = (int)std::max(1u, std::min(8u, hw ? hw : 1u));
  for (int t = 0; t < T; ++t) {
    GNG_HIP(hipStreamCreateWithFlags(&p->pipe_stream[t], hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
      GNG_HIP(hipEventCreateWithFlags(&p->pipe_ev[t][k], hipEventDisableTiming));
      GNG_HIP(hipHostMalloc(reinterpret_cast<void**>(&p->pipe_slot[t][k]), kStageBytes, hipHostMallocDefault));
    }
  }
  p->pipe_threads = T;
  return GNG_OK;
}

// One transfer: kind 0 copies `count` bytes host -> device; kind 1 converts
// `count` doubles host -> floats on the device; kind 2 converts `count`
// floats on the device -> doubles host; kind 3 copies `count` bytes device
// -> host.
struct Xfer {
  void* dev;
  const void* src;
  void* dst;
  size_t count;
  int kind;
};

// Splits the transfers into slot-sized chunks and runs them on the worker
// threads; for downloads, `after` (recorded on the compute stream) gates
// every DMA; for uploads, s waits for every worker stream when they are done.
int run_xfers(gng_pm_plan* p, const std::vector<Xfer>& xs, hipStream_t s, hipEvent_t after) {
  struct Chunk {
    const Xfer* x;
    size_t off, n;  // in elements of the transfer (bytes for kind 0)
  };
  std::vector<Chunk> chunks;
  for (const Xfer& x : xs) {
    const size_t per = (x.kind == 0 || x.kind == 3) ? kStageBytes : kStageBytes / 8;
    for (size_t o = 0; o < x.count; o += per) chunks.push_back({&x, o, std::min(per, x.count - o)});
  }
  if (chunks.empty()) return GNG_OK;
  const int T = std::min<int>(p->pipe_threads, (int)chunks.size());
  std::vector<hipError_t> err((size_t)T, hipSuccess);
  auto work = [&](int t) {
    hipError_t e = hipSetDevice(p->device);
    hipStream_t st = p->pipe_stream[t];
    if (e == hipSuccess && after) e = hipStreamWaitEvent(st, after, 0);
    bool used[2] = {false, false};
    int k = 0;
    std::vector<size_t> mine;
    for (size_t c = (size_t)t; c < chunks.size(); c += (size_t)T) mine.push_back(c);
    const bool down = !mine.empty() && chunks[mine[0]].x->kind >= 2;
    auto issue_down = [&](size_t c, int slot) {
      const Chunk& ch = chunks[c];
      const bool raw = ch.x->kind == 3;
      const void* d = raw ? (const void*)(static_cast<const unsigned char*>(ch.x->dev) + ch.off)
                          : (const void*)(static_cast<const float*>(ch.x->dev) + ch.off);
      hipError_t r = hipMemcpyAsync(p->pipe_slot[t][slot], d, raw ? ch.n : ch.n * sizeof(float),
                                    hipMemcpyDeviceToHost, st);
      if (r == hipSuccess) r = hipEventRecord(p->pipe_ev[t][slot], st);
      return r;
    };
    if (down) {
      // two DMAs in flight: chunk i + 1 goes out while chunk i is converted
      if (e == hipSuccess) e = issue_down(mine[0], 0);
      for (size_t i = 0; i < mine.size() && e == hipSuccess; ++i) {
        const int slot = (int)(i & 1);
        if (i + 1 < mine.size()) e = issue_down(mine[i + 1], slot ^ 1);
        if (e == hipSuccess) e = hipEventSynchronize(p->pipe_ev[t][slot]);
        if (e != hipSuccess) break;
        const Chunk& ch = chunks[mine[i]];
        if (ch.x->kind == 3) {
          std::memcpy(static_cast<unsigned char*>(ch.x->dst) + ch.off, p->pipe_slot[t][slot], ch.n);
        } else {
          const float* src = reinterpret_cast<const float*>(p->pipe_slot[t][slot]);
          double* dst = static_cast<double*>(ch.x->dst) + ch.off;
          for (size_t j = 0; j < ch.n; ++j) dst[j] = (double)src[j];
        }
      }
    } else {
      for (size_t i = 0; i < mine.size() && e == hipSuccess; ++i) {
        const int slot = k++ & 1;
        if (used[slot]) e = hipEventSynchronize(p->pipe_ev[t][slot]);
        if (e != hipSuccess) break;
        const Chunk& ch = chunks[mine[i]];
        unsigned char* buf = p->pipe_slot[t][slot];
        size_t bytes;
        void* dev;
        if (ch.x->kind == 0) {
          std::memcpy(buf, static_cast<const unsigned char*>(ch.x->src) + ch.off, ch.n);
          bytes = ch.n;
          dev = static_cast<unsigned char*>(ch.x->dev) + ch.off;
        } else {
          const double* src = static_cast<const double*>(ch.x->src) + ch.off;
          float* f = reinterpret_cast<float*>(buf);
          for (size_t j = 0; j < ch.n; ++j) f[j] = (float)src[j];
          bytes = ch.n * sizeof(float);
          dev = static_cast<float*>(ch.x->dev) + ch.off;
        }
        e = hipMemcpyAsync(dev, buf, bytes, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipEventRecord(p->pipe_ev[t][slot], st);
        used[slot] = true;
      }
    }
    err[(size_t)t] = e;
  };
  std::vector<std::thread> pool;
  int started = 1;
  try {
    for (int t = 1; t < T; ++t, ++started) pool.emplace_back(work, t);
  } catch (...) {
  }
  // workers that could not be started run here, after thread 0's share
  work(0);
  for (int t = started; t < T; ++t) work(t);
  for (auto& th : pool) th.join();
  for (int t = 0; t < T; ++t)
    if (err[(size_t)t] != hipSuccess)
      return fail(err[(size_t)t] == hipErrorOutOfMemory ? GNG_ERR_OOM : GNG_ERR_HIP,
                  std::string("host staging: ") + hipGetErrorString(err[(size_t)t]));
  if (!after) {
    // uploads: the compute stream waits for every worker stream's last DMA
    (void)hipSetDevice(p->device);
    for (int t = 0; t < T; ++t) {
      GNG_HIP(hipEventRecord(p->pipe_ev[t][0], p->pipe_stream[t]));
      GNG_HIP(hipStreamWaitEvent(s, p->pipe_ev[t][0], 0));
    }
  }
  return GNG_OK;
}

// Device staging of the f64 entry points: positions 3 cap64, masses cap64,
// accelerations 3 cap64 doubles.
int ensure_part64(gng_pm_plan* p, long long n, long long k) {
  const long long need = n > k ? n : k;
  if (need <= p->cap64) return GNG_OK;
  (void)hipFree(p->d_part64);
  p->d_part64 = nullptr;
  p->cap64 = 0;
  GNG_HIP(hipMalloc(&p->d_part64, sizeof(double) * 7 * (size_t)need));
  p->cap64 = need;
  return GNG_OK;
}

// The reference's Vec3 arrays: n x 3 interleaved f64 positions, f64 masses.
PartIn aos64(const double* pos, const double* mass) {
  PartIn in;
  in.pos64 = 1;
  in.x = pos;
  in.y = pos + 1;
  in.z = pos + 2;
  in.stride = 3;
  in.mass64 = 1;
  in.m = mass;
  return in;
}
AccOut aos64_out(double* out) {
  AccOut o;
  o.f64 = 1;
  o.x = out;
  o.y = out + 1;
  o.z = out + 2;
  o.stride = 3;
  return o;
}

int validate(gng_pm_plan* p, long long n, const void* idx, long long k) {
  if (!p) return fail(GNG_ERR_INVALID, "null plan");
  if (n < 0 || k < 0) return fail(GNG_ERR_INVALID, "negative particle count");
  if (!idx && k != n) return fail(GNG_ERR_INVALID, "idx is NULL but k != n");
  if (n >= (1LL << 31)) return fail(GNG_ERR_INVALID, "n of 2^31 or more is not supported");
  return GNG_OK;
}

// Host SoA f32 solve (shared by the legacy and the f32/f64 entry points).
int solve_host_soa(gng_pm_plan* p, const float* x, const float* y, const float* z, const float* m, long long n,
                   const int64_t* idx, long long k, double g, float* ax, float* ay, float* az) {
  int rc = ensure_particles(p, n, k, idx != nullptr);
  if (rc) return rc;
  const size_t cap = (size_t)p->cap;
  float* d = p->d_part;
  float *dx = d, *dy = d + cap, *dz = d + 2 * cap, *dm = d + 3 * cap;
  float *dax = d + 4 * cap, *day = d + 5 * cap, *daz = d + 6 * cap;
  hipStream_t s = p->stream;
  StreamOrder order(p, s);
  if (idx)
    for (long long j = 0; j < k; ++j)
      if (idx[j] < 0 || idx[j] >= n) return fail(GNG_ERR_INVALID, "index out of range");
  if ((rc = ensure_pipe(p))) return rc;
  const size_t nb = sizeof(float) * (size_t)n;
  std::vector<Xfer> up;
  if (n > 0)
    for (auto pr : {std::make_pair((void*)dx, (const void*)x), std::make_pair((void*)dy, (const void*)y),
                    std::make_pair((void*)dz, (const void*)z), std::make_pair((void*)dm, (const void*)m)})
      up.push_back({pr.first, pr.second, nullptr, nb, 0});
  if (idx && k > 0) up.push_back({p->d_idx, idx, nullptr, sizeof(long long) * (size_t)k, 0});
  if ((rc = run_xfers(p, up, s, nullptr))) return rc;
  rc = run_pipeline(p, soa32(dx, dy, dz, dm), n, idx ? p->d_idx : nullptr, k, g, acc32(dax, day, daz), s);
  if (rc) return rc;
  GNG_HIP(hipEventRecord(p->done_ev, s));
  const size_t kb = sizeof(float) * (size_t)k;
  std::vector<Xfer> down;
  if (k > 0)
    for (auto pr : {std::make_pair((void*)dax, (void*)ax), std::make_pair((void*)day, (void*)ay),
                    std::make_pair((void*)daz, (void*)az)})
      down.push_back({pr.first, nullptr, pr.second, kb, 3});
  if ((rc = run_xfers(p, down, s, p->done_ev))) return rc;
  GNG_HIP(hipStreamSynchronize(s));
  return GNG_OK;
}

constexpr int kPkMassBlocks = 1024;

// Power spectrum of n particles on device pointers: shell sums (already times
// (V / M_tot)^2) and mode counts, nm/2 shells, copied to the host; the call
// synchronises s.
int run_pk(gng_pm_plan* p, const PartIn& in, long long n, hipStream_t s, double* sums, unsigned long long* modes) {
  const int nm = p->cfg.grid_size, H = nm / 2;
  const MeshOps& o = *p->ops;
  if (!p->d_pk_w) {
    // 1 / sinc(f / nm)^2, sinc(x) = sin(pi x) / (pi x), at the signed
    // frequency f of each index (power_spectrum.rs:130-145)
    std::vector<double> w((size_t)nm);
    for (int i = 0; i < nm; ++i) {
      const int f = i <= H ? i : i - nm;
      const double xv = (double)f / (double)nm;
      double sc = 1.0;
      if (std::fabs(xv) >= 1e-12) {
        const double px = M_PI * xv;
        sc = std::sin(px) / px;
      }
      w[i] = 1.0 / (sc * sc);
    }
    // all four buffers or none: d_pk_w, set last, marks the set as ready
    unsigned long long* modes_b = nullptr;
    double *msum_b = nullptr, *out_b = nullptr, *w_b = nullptr;
    hipError_t e = hipMalloc(&modes_b, sizeof(unsigned long long) * (size_t)H);
    if (e == hipSuccess) e = hipMalloc(&msum_b, sizeof(double) * (kPkMassBlocks + 1));
    if (e == hipSuccess) e = hipMalloc(&out_b, sizeof(double) * (size_t)H);
    if (e == hipSuccess) e = hipMalloc(&w_b, sizeof(double) * (size_t)nm);
    if (e == hipSuccess) e = hipMemcpy(w_b, w.data(), sizeof(double) * (size_t)nm, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      for (void* q : {(void*)modes_b, (void*)msum_b, (void*)out_b, (void*)w_b}) (void)hipFree(q);
      return fail(e == hipErrorOutOfMemory ? GNG_ERR_OOM : GNG_ERR_HIP,
                  std::string("power-spectrum tables: ") + hipGetErrorString(e));
    }
    p->d_pk_modes = modes_b;
    p->d_pk_msum = msum_b;
    p->d_pk_out = out_b;
    p->d_pk_w = w_b;
  }
  const long long zblocks = (long long)(H / o.yz_cols) * nm, nblk = zblocks + H + 1;
  if (nblk > p->pk_part_cap) {
    (void)hipFree(p->d_pk_part);
    p->d_pk_part = nullptr;
    p->pk_part_cap = 0;
    GNG_HIP(hipMalloc(&p->d_pk_part, sizeof(double) * (size_t)nblk * (size_t)H));
    p->pk_part_cap = nblk;
  }
  int rc = run_front(p, in, n, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((H + 255) / 256)), dim3(256), 0, s, p->d_pk_modes, H);
  if (in.mass64)
    hipLaunchKernelGGL(k_mass_partials<double>, dim3(kPkMassBlocks), dim3(256), 0, s, (const double*)in.m, n,
                       p->d_pk_msum);
  else
    hipLaunchKernelGGL(k_mass_partials<float>, dim3(kPkMassBlocks), dim3(256), 0, s, (const float*)in.m, n,
                       p->d_pk_msum);
  hipLaunchKernelGGL(k_mass_total, dim3(1), dim3(1024), 0, s, p->d_pk_msum, kPkMassBlocks);
  o.z_pk(dim3((unsigned)(H / o.yz_cols), (unsigned)nm), dim3((unsigned)o.yz_threads), s, p->d_P, p->d_tw, p->d_pk_w,
         p->d_pk_part, nblk, p->d_pk_modes);
  o.pk_col0(dim3((unsigned)(H + 1)), dim3(256), s, p->d_P, p->d_tw, p->d_pk_w, p->d_pk_part, nblk, zblocks,
            p->d_pk_modes);
  // in the stage timings the z transform and binning take the z stage's
  // slot; the inverse stages and the interpolation are empty
  for (int i = 6; i < kStages + 1; ++i) mark(p, i, s);
  if (p->timing) ++p->timed_solves;
  const double L = p->cfg.box_size;
  hipLaunchKernelGGL(k_pk_reduce, dim3((unsigned)H), dim3(256), 0, s, p->d_pk_part, nblk,
                     p->d_pk_msum + kPkMassBlocks, L * L * L, p->d_pk_out);
  GNG_HIP(hipGetLastError());
  GNG_HIP(hipMemcpyAsync(sums, p->d_pk_out, sizeof(double) * (size_t)H, hipMemcpyDeviceToHost, s));
  GNG_HIP(hipMemcpyAsync(modes, p->d_pk_modes, sizeof(unsigned long long) * (size_t)H, hipMemcpyDeviceToHost, s));
  GNG_HIP(hipStreamSynchronize(s));
  p->density_valid = p->keep_density;
  p->potential_valid = false;
  return GNG_OK;
}

// Initial conditions: the amplitude table to the device.  sigma[n2] is the
// spectrum at |n| = sqrt(n2) for every |n|^2 of a mode without a Nyquist
// component, n2 = 0 .. 3 (nm/2 - 1)^2 (the reference evaluates its spectrum
// closure at sqrt(n2) per mode, ic_zeldovich.rs:315-316).
int ic_upload_sigma(gng_pm_plan* p, const double* sigma, long long sigma_len, hipStream_t s, IcParams& ip,
                    unsigned long long seed) {
  const int nm = p->cfg.grid_size, H = nm / 2;
  if (nm < 4 || nm % 2) return fail(GNG_ERR_INVALID, "initial conditions need an even grid_size of 4 or more");
  const long long need = 3LL * (H - 1) * (H - 1) + 1;
  if (!sigma || sigma_len < need) return fail(GNG_ERR_INVALID, "sigma table shorter than 3 (grid_size/2 - 1)^2 + 1");
  if (need > p->ic_sigma_cap) {
    (void)hipFree(p->d_ic_sigma);
    p->d_ic_sigma = nullptr;
    p->ic_sigma_cap = 0;
    GNG_HIP(hipMalloc(&p->d_ic_sigma, sizeof(double) * (size_t)need));
    p->ic_sigma_cap = need;
  }
  GNG_HIP(hipMemcpyAsync(p->d_ic_sigma, sigma, sizeof(double) * (size_t)need, hipMemcpyHostToDevice, s));
  // the caller's table may go away once this returns
  GNG_HIP(hipStreamSynchronize(s));
  ip.seed = seed;
  ip.sigma = p->d_ic_sigma;
  return GNG_OK;
}

// Zel'dovich particles lo .. lo + cnt into device pos / vel (n x 3 f64):
// the z pass generates the modes, the force pipeline's inverse y and x passes
// finish the three displacement transforms, then the lattice kernel.
int run_ic(gng_pm_plan* p, const IcParams& ip, double vel_factor, double scale, long long lo, long long cnt,
           double* pos, double* vel, hipStream_t s, int wrap = 1) {
  const int nm = p->cfg.grid_size, H = nm / 2;
  const MeshOps& o = *p->ops;
  const long long rows = (long long)nm * nm;
  const dim3 yz_grid((unsigned)(H / o.yz_cols), (unsigned)nm);
  const dim3 yz_block((unsigned)o.yz_threads);
  o.ic_z(yz_grid, yz_block, s, ip, p->d_A, p->d_B, p->d_tw, make_int2(0, 1));  // a, -i n_z a
  o.y_inv(yz_grid, yz_block, s, p->d_A, p->d_B, p->d_P, p->d_tw);
  o.x_inv(dim3((unsigned)(rows / o.x_rows), 3), dim3(256), s, p->d_A, p->d_P, p->d_B, nullptr, p->d_tw, 0, 1.0f);
  p->density_valid = false;
  p->potential_valid = false;
  if (cnt > 0) {
    const double L = p->cfg.box_size;
    const double d = L / (double)nm;
    const double ifft_norm = 1.0 / (double)((long long)nm * nm * nm);
    hipLaunchKernelGGL(k_ic_particles, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s,
                       reinterpret_cast<const float*>(p->d_B), reinterpret_cast<const float*>(p->d_P),
                       reinterpret_cast<const float*>(p->d_A), 2 * mesh_plane(nm), nm, d, L, ifft_norm * d, scale,
                       vel_factor, lo, cnt, pos, vel, wrap);
  }
  GNG_HIP(hipGetLastError());
  return GNG_OK;
}

// Second-order (2LPT) particles (ic_2lpt.rs:348-553): the first-order part as
// above without the wrap; then the six second derivatives of phi in two
// rounds of three fields through the z, y and x passes with mode multipliers
// (round 0: xx yy zz, round 1: xz yz xy), accumulated into the source S; then
// S through the force pipeline with G = 1/|n|^2 and every mode with a Nyquist
// component masked, which leaves Psi2 = Re IFFT(-i n S-hat / |n|^2) in the
// force meshes; then the second-order particle kernel.
int run_ic_2lpt(gng_pm_plan* p, const IcParams& ip, double vel1, double vel2, double d2, double scale, long long lo,
                long long cnt, double* pos, double* vel, hipStream_t s) {
  const int nm = p->cfg.grid_size, H = nm / 2;
  const MeshOps& o = *p->ops;
  const long long rows = (long long)nm * nm, nm3 = rows * nm;
  const dim3 yz_grid((unsigned)(H / o.yz_cols), (unsigned)nm);
  const dim3 yz_block((unsigned)o.yz_threads);
  const long long fplane = 2 * mesh_plane(nm);
  if (!p->d_ic_mask) {
    std::vector<float> mask((size_t)nm, 1.0f);
    mask[(size_t)H] = 0.0f;
    GNG_HIP(hipMalloc(&p->d_ic_mask, sizeof(float) * (size_t)nm));
    const hipError_t e = hipMemcpy(p->d_ic_mask, mask.data(), sizeof(float) * (size_t)nm, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      // leave no uninitialised mask behind for the next call
      (void)hipFree(p->d_ic_mask);
      p->d_ic_mask = nullptr;
      return fail(GNG_ERR_HIP, std::string("2LPT mask upload: ") + hipGetErrorString(e));
    }
  }
  int rc = run_ic(p, ip, vel1, scale, lo, cnt, pos, vel, s, 0);
  if (rc) return rc;
  float* S = nullptr;  // the source, an unpadded real mesh
  GNG_HIP(hipMalloc(&S, sizeof(float) * (size_t)nm3));
  const float* fa = reinterpret_cast<const float*>(p->d_A);
  const float* fp = reinterpret_cast<const float*>(p->d_P);
  const float* fb = reinterpret_cast<const float*>(p->d_B);
  const double inv_n6 = 1.0 / ((double)nm3 * (double)nm3);  // two unnormalised inverse transforms per factor
  const dim3 xg((unsigned)(rows / o.x_rows), 3);
  // round 0: (a, n_z^2 a) -> y (1, n_y^2, 1) -> x (n_x^2, 1, 1) = xx, yy, zz (mesh axes)
  o.ic_z(yz_grid, yz_block, s, ip, p->d_A, p->d_B, p->d_tw, make_int2(0, 3));
  o.y_inv_mul(yz_grid, yz_block, s, p->d_A, p->d_B, p->d_P, p->d_tw, make_int3(0, 3, 0));
  o.x_inv_mul(xg, dim3(256), s, p->d_A, p->d_P, p->d_B, p->d_tw, make_int3(3, 0, 0));
  hipLaunchKernelGGL(k_ic_source, dim3((unsigned)((nm3 + 255) / 256)), dim3(256), 0, s, fa, fp, fb, S, fplane,
                     (long long)rows, nm, inv_n6, 0);
  // round 1: (n_z a, a) -> y (1, n_y, n_y) -> x (n_x, 1, n_x) = xz, yz, xy
  o.ic_z(yz_grid, yz_block, s, ip, p->d_A, p->d_B, p->d_tw, make_int2(2, 0));
  o.y_inv_mul(yz_grid, yz_block, s, p->d_A, p->d_B, p->d_P, p->d_tw, make_int3(0, 2, 2));
  o.x_inv_mul(xg, dim3(256), s, p->d_A, p->d_P, p->d_B, p->d_tw, make_int3(2, 0, 2));
  hipLaunchKernelGGL(k_ic_source, dim3((unsigned)((nm3 + 255) / 256)), dim3(256), 0, s, fa, fp, fb, S, fplane,
                     (long long)rows, nm, inv_n6, 1);
  // S -> Psi2 through the force pipeline's passes
  GreenParams gp;
  gp.c64 = -1.0;  // G = -c / |n|^2 = 1 / |n|^2
  gp.a64 = 0.0;
  gp.c = -1.0f;
  gp.a = 0.0f;
  gp.deconv = p->d_ic_mask;
  o.x_fwd(dim3((unsigned)(rows / o.x_rows)), dim3(256), s, S, p->d_P, p->d_tw, rows / 2);
  o.y_fwd(yz_grid, yz_block, s, p->d_P, p->d_tw);
  o.z_green(dim3((unsigned)(H / o.z_cols), (unsigned)nm), dim3((unsigned)o.z_threads), s, p->d_P, p->d_A, p->d_B,
            p->d_tw, gp);
  o.z_green_col0(dim3((unsigned)(H + 1)), dim3(256), s, p->d_P, p->d_A, p->d_B, p->d_tw, gp);
  o.y_inv(yz_grid, yz_block, s, p->d_A, p->d_B, p->d_P, p->d_tw);
  o.x_inv(xg, dim3(256), s, p->d_A, p->d_P, p->d_B, nullptr, p->d_tw, 0, 1.0f);
  if (cnt > 0) {
    const double L = p->cfg.box_size;
    const double d = L / (double)nm;
    const double ifft_norm = 1.0 / (double)nm3;
    hipLaunchKernelGGL(k_ic_particles_2lpt, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, fb, fp, fa, fplane,
                       nm, L, ifft_norm * d, scale * scale, d2, vel2, lo, cnt, pos, vel);
  }
  hipError_t e = hipGetLastError();
  // S must outlive the kernels that read it
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(S);
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? GNG_ERR_OOM : GNG_ERR_HIP, std::string("2LPT: ") + hipGetErrorString(e));
  return GNG_OK;
}

int ic_check_range(const gng_pm_plan* p, int64_t lo, int64_t hi) {
  const long long n3 = (long long)p->cfg.grid_size * p->cfg.grid_size * p->cfg.grid_size;
  if (lo < 0 || hi < lo || hi > n3) return fail(GNG_ERR_INVALID, "need 0 <= lo <= hi <= grid_size^3");
  if (hi - lo >= (1LL << 31)) return fail(GNG_ERR_INVALID, "2^31 or more particles per call are not supported");
  return GNG_OK;
}

// Non-empty shells in increasing k: k = (s + 1) 2 pi / L, P = sum / modes
// (power_spectrum.rs:115-126).
int pk_bins(const gng_pm_plan* p, const double* sums, const unsigned long long* modes, int32_t capacity, double* k,
            double* pk, uint64_t* n_modes, int32_t* n_bins) {
  const int H = p->cfg.grid_size / 2;
  const double k_fund = 2.0 * M_PI / p->cfg.box_size;
  int nb = 0;
  for (int b = 0; b < H; ++b) {
    if (modes[b] == 0) continue;
    if (nb >= capacity) return fail(GNG_ERR_INVALID, "capacity below the number of shells (grid_size / 2)");
    if (k) k[nb] = ((double)b + 1.0) * k_fund;
    if (pk) pk[nb] = sums[b] / (double)modes[b];
    if (n_modes) n_modes[nb] = modes[b];
    ++nb;
  }
  *n_bins = nb;
  return GNG_OK;
}

}  // namespace

extern "C" {

uint64_t gng_hilbert_key3(uint32_t x, uint32_t y, uint32_t z, int32_t bits) { return hilbert_key3(x, y, z, bits); }

int gng_pm_abi_version(void) { return GNG_PM_ABI_VERSION; }

const char* gng_last_error_string(void) { return g_last_error.c_str(); }

int gng_pm_mesh_supported(int grid_size) { return find_ops(grid_size) ? 1 : 0; }

int gng_pm_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int gng_pm_create(const gng_pm_config* cfg, gng_pm_plan** out) {
  if (!out) return fail(GNG_ERR_INVALID, "null output pointer");
  *out = nullptr;
  if (!cfg) return fail(GNG_ERR_INVALID, "null config");
  if (!(cfg->box_size > 0.0) || !std::isfinite(cfg->box_size)) return fail(GNG_ERR_INVALID, "box_size must be > 0");
  const MeshOps* ops = find_ops(cfg->grid_size);
  if (!ops)
    return fail(GNG_ERR_UNSUPPORTED, "grid_size " + std::to_string(cfg->grid_size) +
                                         " is not compiled in (supported: 2^a 3^b from 4 to 1024)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(GNG_ERR_NO_DEVICE, "no HIP device");
  gng_pm_plan* p = new (std::nothrow) gng_pm_plan();
  if (!p) return fail(GNG_ERR_OOM, "host allocation failed");
  p->cfg = *cfg;
  p->ops = ops;
  if (cfg->device >= 0) {
    if (cfg->device >= ndev) {
      delete p;
      return fail(GNG_ERR_INVALID, "device ordinal out of range");
    }
    p->device = cfg->device;
  } else {
    (void)hipGetDevice(&p->device);
  }
  auto bail = [&](int rc) {
    free_plan(p);
    return rc;
  };
  int rc = set_device(p);
  if (rc) return bail(rc);
  if (hipDeviceGetAttribute(&p->cus, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess) p->cus = 0;
  p->tune = cached_tuning(p->device, cfg->grid_size);
  const int nm = cfg->grid_size;
  p->nm3 = (long long)nm * nm * nm;
  p->tm.T = ops->tile;
  p->tm.nt = nm / ops->tile;
  p->nb = (long long)p->tm.nt * p->tm.nt * p->tm.nt * 8;
  p->scan_blocks = (int)((p->nb + kScanBlock - 1) / kScanBlock);
  const size_t half = sizeof(float2) * (size_t)nm * (size_t)mesh_plane(nm);
#define GNG_ALLOC(ptr, bytes)                                                                             \
  do {                                                                                                    \
    hipError_t e_ = hipMalloc(&(ptr), (bytes));                                                           \
    if (e_ != hipSuccess)                                                                                 \
      return bail(fail(e_ == hipErrorOutOfMemory ? GNG_ERR_OOM : GNG_ERR_HIP,                             \
                       std::string("hipMalloc ") + #ptr + ": " + hipGetErrorString(e_)));                 \
  } while (0)
  GNG_ALLOC(p->d_P, half);
  GNG_ALLOC(p->d_A, half);
  GNG_ALLOC(p->d_B, half);
  GNG_ALLOC(p->d_tw, sizeof(float2) * (size_t)nm);
  GNG_ALLOC(p->d_deconv, sizeof(float) * (size_t)nm);
  GNG_ALLOC(p->d_sc, sizeof(SolveScalars));
  GNG_ALLOC(p->d_count, sizeof(unsigned) * (size_t)((p->nb + 3) / 4 * 4));
  GNG_ALLOC(p->d_offset, sizeof(unsigned) * (size_t)(p->nb + 1));
  GNG_ALLOC(p->d_partial, sizeof(unsigned) * (size_t)p->scan_blocks);
  GNG_ALLOC(p->d_tile_rank, sizeof(unsigned) * (size_t)(p->nb / 8));
  GNG_ALLOC(p->d_tile_order, sizeof(unsigned) * (size_t)(p->nb / 8));
  GNG_ALLOC(p->d_tile_src, sizeof(unsigned) * (size_t)p->nb);
  GNG_ALLOC(p->d_tile_ranges, sizeof(unsigned) * (size_t)(p->nb / 8) * kRangeWords);
#undef GNG_ALLOC
  {
    // Hilbert ranks of the tiles (keys on the enclosing power-of-two grid)
    const int nt = p->tm.nt;
    int bits = 0;
    while ((1 << bits) < nt) ++bits;
    std::vector<std::pair<uint64_t, unsigned>> keys;
    keys.reserve((size_t)nt * nt * nt);
    for (int tz = 0; tz < nt; ++tz)
      for (int ty = 0; ty < nt; ++ty)
        for (int tx = 0; tx < nt; ++tx)
          keys.emplace_back(hilbert_key3(tx, ty, tz, bits), (unsigned)((tz * nt + ty) * nt + tx));
    std::sort(keys.begin(), keys.end());
    std::vector<unsigned> rank(keys.size()), order(keys.size()), src(keys.size() * 8);
    for (size_t r = 0; r < keys.size(); ++r) {
      order[r] = keys[r].second;
      rank[keys[r].second] = (unsigned)r;
    }
    fill_tile_src(nt, rank.data(), src.data(), 0, 1);
    hipError_t e2 = hipMemcpy(p->d_tile_rank, rank.data(), sizeof(unsigned) * rank.size(), hipMemcpyHostToDevice);
    if (e2 == hipSuccess)
      e2 = hipMemcpy(p->d_tile_order, order.data(), sizeof(unsigned) * order.size(), hipMemcpyHostToDevice);
    if (e2 == hipSuccess)
      e2 = hipMemcpy(p->d_tile_src, src.data(), sizeof(unsigned) * src.size(), hipMemcpyHostToDevice);
    if (e2 != hipSuccess) return bail(fail(GNG_ERR_HIP, std::string("tile order upload: ") + hipGetErrorString(e2)));
    p->tm.rank = p->d_tile_rank;
    p->tm.order = p->d_tile_order;
    p->tm.src = p->d_tile_src;
    p->tm.ranges = p->d_tile_ranges;
  }
  hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_plan_tables, dim3(1), dim3(256), 0, p->stream, nm, 0, 0, p->d_tw, p->d_deconv, nullptr,
                       nullptr, nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->done_ev, hipEventDisableTiming);
  for (int j = 0; j < kTimingSlots && e == hipSuccess; ++j)
    for (int i = 0; i <= kStages && e == hipSuccess; ++i) e = hipEventCreate(&p->ev[j][i]);
  if (e != hipSuccess) return bail(fail(GNG_ERR_HIP, std::string("plan setup: ") + hipGetErrorString(e)));
  *out = p;
  return GNG_OK;
}

void gng_pm_destroy(gng_pm_plan* p) { free_plan(p); }

int gng_pm_accelerations_f32(gng_pm_plan* p, const float* x, const float* y, const float* z, const float* mass,
                             int64_t n, const int64_t* idx, int64_t k, double g, float* ax, float* ay, float* az) {
  int rc = validate(p, n, idx, k);
  if (rc) return rc;
  if (n > 0 && (!x || !y || !z || !mass)) return fail(GNG_ERR_INVALID, "null input array");
  if (k > 0 && (!ax || !ay || !az)) return fail(GNG_ERR_INVALID, "null output array");
  std::lock_guard<std::mutex> lk(p->mu);
  if ((rc = set_device(p))) return rc;
  return solve_host_soa(p, x, y, z, mass, n, idx, k, g, ax, ay, az);
}

int gng_pm_accelerations_f64(gng_pm_plan* p, const double* pos, const double* mass, int64_t n, const int64_t* idx,
                             int64_t k, double eps2, double g, double* out) {
  (void)eps2;  // PmSolver ignores eps2 (crates/gadget-ng-pm/src/solver.rs:70-72)
  int rc = validate(p, n, idx, k);
  if (rc) return rc;
  if (n > 0 && (!pos || !mass)) return fail(GNG_ERR_INVALID, "null input array");
  if (k > 0 && !out) return fail(GNG_ERR_INVALID, "null output array");
  if (idx)
    for (long long j = 0; j < k; ++j)
      if (idx[j] < 0 || idx[j] >= n) return fail(GNG_ERR_INVALID, "index out of range");
  std::lock_guard<std::mutex> lk(p->mu);
  if ((rc = set_device(p))) return rc;
  hipStream_t s = p->stream;
  StreamOrder order(p, s);
  if ((rc = ensure_part64(p, n, k))) return rc;
  if (idx && (rc = ensure_idx(p, k))) return rc;
  if ((rc = ensure_pipe(p))) return rc;
  // device layout: f64 positions (3 cap64), f32 masses (cap64 floats), f32
  // accelerations (3 cap64 floats, interleaved like the output)
  double* dpos = p->d_part64;
  float* dm = reinterpret_cast<float*>(dpos + 3 * p->cap64);
  float* dout = dm + p->cap64;
  std::vector<Xfer> up;
  if (n > 0) {
    up.push_back({dpos, pos, nullptr, sizeof(double) * 3 * (size_t)n, 0});
    up.push_back({dm, mass, nullptr, (size_t)n, 1});  // masses enter the deposit as f32
  }
  if (idx && k > 0) up.push_back({p->d_idx, idx, nullptr, sizeof(long long) * (size_t)k, 0});
  if ((rc = run_xfers(p, up, s, nullptr))) return rc;
  PartIn in = aos64(dpos, nullptr);
  in.mass64 = 0;
  in.m = dm;
  AccOut o;
  o.x = dout;
  o.y = dout + 1;
  o.z = dout + 2;
  o.stride = 3;
  rc = run_pipeline(p, in, n, idx ? p->d_idx : nullptr, k, g, o, s);
  if (rc) return rc;
  GNG_HIP(hipEventRecord(p->done_ev, s));
  std::vector<Xfer> down;
  if (k > 0) down.push_back({dout, nullptr, out, 3 * (size_t)k, 2});
  if ((rc = run_xfers(p, down, s, p->done_ev))) return rc;
  // the download streams are done (their last events were waited for); the
  // next entry point orders itself after s
  GNG_HIP(hipStreamSynchronize(s));
  return GNG_OK;
}

int gng_pm_solve_device_f64(gng_pm_plan* p, const double* d_pos, const double* d_mass, int64_t n,
                            const int64_t* d_idx, int64_t k, double g, double* d_out, void* stream) {
  int rc = validate(p, n, d_idx, k);
  if (rc) return rc;
  if (n > 0 && (!d_pos || !d_mass)) return fail(GNG_ERR_INVALID, "null input array");
  if (k > 0 && !d_out) return fail(GNG_ERR_INVALID, "null output array");
  std::lock_guard<std::mutex> lk(p->mu);
  if ((rc = set_device(p))) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  StreamOrder order(p, s);
  return run_pipeline(p, aos64(d_pos, d_mass), n, reinterpret_cast<const long long*>(d_idx), k, g,
                      aos64_out(d_out), s);
}

int gng_pm_deposit_order(gng_pm_plan* p, const double* d_pos, const double* d_mass, int64_t n, uint32_t* d_order,
                         void* stream) {
  int rc = validate(p, n, nullptr, n);
  if (rc) return rc;
  if (n > 0 && (!d_pos || !d_mass || !d_order)) return fail(GNG_ERR_INVALID, "null array");
  if (n == 0) return GNG_OK;
  std::lock_guard<std::mutex> lk(p->mu);
  if ((rc = set_device(p))) return rc;
  if ((rc = ensure_work(p, n))) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  StreamOrder order(p, s);
  run_sort(p, aos64(d_pos, d_mass), n, d_order, false, s);
  GNG_HIP(hipGetLastError());
  return GNG_OK;
}

int gng_pm_solve_device_soa64(gng_pm_plan* p, const double* d_x, const double* d_y, const double* d_z,
                              const float* d_mass, int64_t n, const int64_t* d_idx, int64_t k, double g, float* d_ax,
                              float* d_ay, float* d_az, void* stream) {
  int rc = validate(p, n, d_idx, k);
  if (rc) return rc;
  if (n > 0 && (!d_x || !d_y || !d_z || !d_mass)) return fail(GNG_ERR_INVALID, "null input array");
  if (k > 0 && (!d_ax || !d_ay || !d_az)) return fail(GNG_ERR_INVALID, "null output array");
  std::lock_guard<std::mutex> lk(p->mu);
  if ((rc = set_device(p))) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  StreamOrder order(p, s);
  PartIn in;
  in.pos64 = 1;
  in.x = d_x;
  in.y = d_y;
  in.z = d_z;
  in.m = d_mass;
  return run_pipeline(p, in, n, reinterpret_cast<const long long*>(d_idx), k, g, acc32(d_ax, d_ay, d_az), s);
}

int gng_pm_power_spectrum(gng_pm_plan* p, const float* d_x, const float* d_y, const float* d_z, const float* d_mass,
                          int64_t n, int32_t capacity, double* k, double* pk, uint64_t* n_modes, int32_t* n_bins,
                          void* stream) {
  if (!p || !n_bins) return fail(GNG_ERR_INVALID, "null argument");
  if (n <= 0) return fail(GNG_ERR_INVALID, "the power spectrum needs at least one particle");
  if (n >= (1LL << 31)) return fail(GNG_ERR_INVALID, "n of 2^31 or more is not supported");
  if (!d_x || !d_y || !d_z || !d_mass) return fail(GNG_ERR_INVALID, "null particle array");
  std::lock_guard<std::mutex> lk(p->mu);
  int rc = set_device(p);
  if (rc) return rc;
  const int H = p->cfg.grid_size / 2;
  std::vector<double> sums((size_t)H);
  std::vector<unsigned long long> modes((size_t)H);
  StreamOrder order(p, reinterpret_cast<hipStream_t>(stream));
  if ((rc = run_pk(p, soa32(d_x, d_y, d_z, d_mass), n, reinterpret_cast<hipStream_t>(stream), sums.data(),
                   modes.data())))
    return rc;
  return pk_bins(p, sums.data(), modes.data(), capacity, k, pk, n_modes, n_bins);
}

int gng_pm_power_spectrum_f64(gng_pm_plan* p, const double* pos, const double* mass, int64_t n, int32_t capacity,
                              double* k, double* pk, uint64_t* n_modes, int32_t* n_bins) {
  if (!p || !n_bins) return fail(GNG_ERR_INVALID, "null argument");
  if (n <= 0) return fail(GNG_ERR_INVALID, "the power spectrum needs at least one particle");
  if (n >= (1LL << 31)) return fail(GNG_ERR_INVALID, "n of 2^31 or more is not supported");
  if (!pos || !mass) return fail(GNG_ERR_INVALID, "null input array");
  std::lock_guard<std::mutex> lk(p->mu);
  int rc = set_device(p);
  if (rc) return rc;
  hipStream_t s = p->stream;
  StreamOrder order(p, s);
  if ((rc = ensure_part64(p, n, 0))) return rc;
  double* dpos = p->d_part64;
  double* dm = dpos + 3 * p->cap64;
  GNG_HIP(hipMemcpyAsync(dpos, pos, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, s));
  GNG_HIP(hipMemcpyAsync(dm, mass, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, s));
  const int H = p->cfg.grid_size / 2;
  std::vector<double> sums((size_t)H);
  std::vector<unsigned long long> modes((size_t)H);
  if ((rc = run_pk(p, aos64(dpos, dm), n, s, sums.data(), modes.data()))) return rc;
  return pk_bins(p, sums.data(), modes.data(), capacity, k, pk, n_modes, n_bins);
}

int gng_ic_delta_kspace(gng_pm_plan* p, uint64_t seed, const double* sigma, int64_t sigma_len, double* out) {
  if (!p || !out) return fail(GNG_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(p->mu);
  int rc = set_device(p);
  if (rc) return rc;
  hipStream_t s = p->stream;
  StreamOrder order(p, s);
  IcParams ip{};
  if ((rc = ic_upload_sigma(p, sigma, sigma_len, s, ip, seed))) return rc;
  const int nm = p->cfg.grid_size;
  const long long total = (long long)nm * nm * (nm / 2 + 1);
  double2* d_out = nullptr;
  GNG_HIP(hipMalloc(&d_out, sizeof(double2) * (size_t)total));
  hipLaunchKernelGGL(k_ic_delta, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, ip, nm, d_out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, sizeof(double2) * (size_t)total, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d_out);
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? GNG_ERR_OOM : GNG_ERR_HIP,
                std::string("initial-condition modes: ") + hipGetErrorString(e));
  return GNG_OK;
}

int gng_ic_zeldovich_device(gng_pm_plan* p, uint64_t seed, const double* sigma, int64_t sigma_len,
                            double vel_factor, double scale, int64_t lo, int64_t hi, double* d_pos, double* d_vel,
                            void* stream) {
  if (!p) return fail(GNG_ERR_INVALID, "null plan");
  int rc = ic_check_range(p, lo, hi);
  if (rc) return rc;
  if (hi > lo && (!d_pos || !d_vel)) return fail(GNG_ERR_INVALID, "null output array");
  std::lock_guard<std::mutex> lk(p->mu);
  if ((rc = set_device(p))) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  StreamOrder order(p, s);
  IcParams ip{};
  if ((rc = ic_upload_sigma(p, sigma, sigma_len, s, ip, seed))) return rc;
  return run_ic(p, ip, vel_factor, scale, lo, hi - lo, d_pos, d_vel, s);
}

int gng_ic_zeldovich(gng_pm_plan* p, uint64_t seed, const double* sigma, int64_t sigma_len, double vel_factor,
                     double scale, int64_t lo, int64_t hi, double* pos, double* vel) {
  if (!p) return fail(GNG_ERR_INVALID, "null plan");
  int rc = ic_check_range(p, lo, hi);
  if (rc) return rc;
  if (hi > lo && (!pos || !vel)) return fail(GNG_ERR_INVALID, "null output array");
  std::lock_guard<std::mutex> lk(p->mu);
  if ((rc = set_device(p))) return rc;
  hipStream_t s = p->stream;
  StreamOrder order(p, s);
  IcParams ip{};
  if ((rc = ic_upload_sigma(p, sigma, sigma_len, s, ip, seed))) return rc;
  const long long cnt = hi - lo;
  double* d = nullptr;
  if (cnt > 0) GNG_HIP(hipMalloc(&d, sizeof(double) * 6 * (size_t)cnt));
  rc = run_ic(p, ip, vel_factor, scale, lo, cnt, d, d + 3 * cnt, s);
  hipError_t e = hipSuccess;
  if (!rc && cnt > 0) {
    e = hipMemcpyAsync(pos, d, sizeof(double) * 3 * (size_t)cnt, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(vel, d + 3 * cnt, sizeof(double) * 3 * (size_t)cnt, hipMemcpyDeviceToHost, s);
  }
  if (!rc && e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  if (rc) return rc;
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? GNG_ERR_OOM : GNG_ERR_HIP,
                std::string("initial conditions: ") + hipGetErrorString(e));
  return GNG_OK;
}

int gng_ic_2lpt_device(gng_pm_plan* p, uint64_t seed, const double* sigma, int64_t sigma_len, double vel1,
                       double vel2, double d2_over_d1sq, double scale, int64_t lo, int64_t hi, double* d_pos,
                       double* d_vel, void* stream) {
  if (!p) return fail(GNG_ERR_INVALID, "null plan");
  int rc = ic_check_range(p, lo, hi);
  if (rc) return rc;
  if (hi > lo && (!d_pos || !d_vel)) return fail(GNG_ERR_INVALID, "null output array");
  std::lock_guard<std::mutex> lk(p->mu);
  if ((rc = set_device(p))) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  StreamOrder order(p, s);
  IcParams ip{};
  if ((rc = ic_upload_sigma(p, sigma, sigma_len, s, ip, seed))) return rc;
  return run_ic_2lpt(p, ip, vel1, vel2, d2_over_d1sq, scale, lo, hi - lo, d_pos, d_vel, s);
}

int gng_ic_2lpt(gng_pm_plan* p, uint64_t seed, const double* sigma, int64_t sigma_len, double vel1, double vel2,
                double d2_over_d1sq, double scale, int64_t lo, int64_t hi, double* pos, double* vel) {
  if (!p) return fail(GNG_ERR_INVALID, "null plan");
  int rc = ic_check_range(p, lo, hi);
  if (rc) return rc;
  if (hi > lo && (!pos || !vel)) return fail(GNG_ERR_INVALID, "null output array");
  std::lock_guard<std::mutex> lk(p->mu);
  if ((rc = set_device(p))) return rc;
  hipStream_t s = p->stream;
  StreamOrder order(p, s);
  IcParams ip{};
  if ((rc = ic_upload_sigma(p, sigma, sigma_len, s, ip, seed))) return rc;
  const long long cnt = hi - lo;
  double* d = nullptr;
  if (cnt > 0) GNG_HIP(hipMalloc(&d, sizeof(double) * 6 * (size_t)cnt));
  rc = run_ic_2lpt(p, ip, vel1, vel2, d2_over_d1sq, scale, lo, cnt, d, d + 3 * cnt, s);
  hipError_t e = hipSuccess;
  if (!rc && cnt > 0) {
    e = hipMemcpyAsync(pos, d, sizeof(double) * 3 * (size_t)cnt, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(vel, d + 3 * cnt, sizeof(double) * 3 * (size_t)cnt, hipMemcpyDeviceToHost, s);
  }
  if (!rc && e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  if (rc) return rc;
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? GNG_ERR_OOM : GNG_ERR_HIP,
                std::string("2LPT initial conditions: ") + hipGetErrorString(e));
  return GNG_OK;
}

int gng_pm_get_config(const gng_pm_plan* p, gng_pm_config* out) {
  if (!p || !out) return fail(GNG_ERR_INVALID, "null argument");
  *out = p->cfg;
  out->device = p->device;
  return GNG_OK;
}

int gng_pm_solve_device(gng_pm_plan* p, const float* d_x, const float* d_y, const float* d_z, const float* d_mass,
                        int64_t n, const int64_t* d_idx, int64_t k, double g, float* d_ax, float* d_ay,
                        float* d_az, void* stream) {
  int rc = validate(p, n, d_idx, k);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(p->mu);
  if ((rc = set_device(p))) return rc;
  StreamOrder order(p, reinterpret_cast<hipStream_t>(stream));
  return run_pipeline(p, soa32(d_x, d_y, d_z, d_mass), n, reinterpret_cast<const long long*>(d_idx), k, g,
                      acc32(d_ax, d_ay, d_az), reinterpret_cast<hipStream_t>(stream));
}

int gng_pm_keep_density(gng_pm_plan* p, int enable) {
  if (!p) return fail(GNG_ERR_INVALID, "null plan");
  std::lock_guard<std::mutex> lk(p->mu);
  int rc = set_device(p);
  if (rc) return rc;
  if (enable && !p->d_rho) GNG_HIP(hipMalloc(&p->d_rho, sizeof(float) * (size_t)p->nm3));
  p->keep_density = enable != 0;
  p->density_valid = false;
  return GNG_OK;
}

int gng_pm_density(gng_pm_plan* p, double* out) {
  if (!p || !out) return fail(GNG_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(p->mu);
  if (!p->density_valid) return fail(GNG_ERR_INVALID, "no kept density: call gng_pm_keep_density(p, 1) first");
  int rc = set_device(p);
  if (rc) return rc;
  std::vector<float> f((size_t)p->nm3);
  GNG_HIP(hipDeviceSynchronize());
  GNG_HIP(hipMemcpy(f.data(), p->d_rho, sizeof(float) * (size_t)p->nm3, hipMemcpyDeviceToHost));
  for (long long i = 0; i < p->nm3; ++i) out[i] = (double)f[i];
  return GNG_OK;
}

}  // extern "C"

namespace {

// Coordinate descent over the launch choices from the defaults, timing whole
// solves of `in` (best of `reps` after a warm-up, HIP events on the plan's
// stream); each knob in turn keeps its faster value.  The caller holds the
// plan's lock and has set the device.
int autotune_impl(gng_pm_plan* p, const PartIn& in, long long n, const AccOut& out, int reps) {
  const int nm = p->cfg.grid_size;
  hipStream_t s = p->stream;
  StreamOrder order(p, s);
  const bool timing = p->timing;
  p->timing = false;
  hipEvent_t e0, e1;
  GNG_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(GNG_ERR_HIP, "event creation failed");
  }
  // time one configuration: best of `reps` solves after a warm-up
  auto measure = [&](float& out_ms) -> int {
    int r2 = run_pipeline(p, in, n, nullptr, n, 1.0, out, s);
    float t_min = 1e30f;
    for (int r = 0; r < reps && r2 == GNG_OK; ++r) {
      if (hipEventRecord(e0, s) != hipSuccess) r2 = fail(GNG_ERR_HIP, "event record failed");
      if (r2 == GNG_OK) r2 = run_pipeline(p, in, n, nullptr, n, 1.0, out, s);
      float ms = 0.0f;
      if (r2 == GNG_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                           hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        r2 = fail(GNG_ERR_HIP, "event timing failed");
      if (r2 == GNG_OK && ms < t_min) t_min = ms;
    }
    out_ms = t_min;
    return r2;
  };
  p->tune = default_tuning(nm);
  Tuning best = p->tune;
  float best_ms = 1e30f;
  // Bring the clocks up first: timed from a cold start, the first (default)
  // configuration measured slow and lost to whichever variant came next.
  // Solves run until 100 ms of them have passed (at most 64).
  int rc = GNG_OK;
  {
    float warm = 0.0f;
    for (int w = 0; w < 64 && warm < 100.0f && rc == GNG_OK; ++w) {
      float ms = 0.0f;
      if (hipEventRecord(e0, s) != hipSuccess) rc = fail(GNG_ERR_HIP, "event record failed");
      if (rc == GNG_OK) rc = run_pipeline(p, in, n, nullptr, n, 1.0, out, s);
      if (rc == GNG_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                           hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = fail(GNG_ERR_HIP, "event timing failed");
      warm += ms;
    }
  }
  if (rc == GNG_OK) rc = measure(best_ms);
  // a variant replaces the best only when it is faster by more than the
  // run-to-run spread of best-of-reps timings on a warm GPU (a few 0.1 %)
  constexpr float kMargin = 0.995f;
  for (int k = 0; k < kKnobs && rc == GNG_OK; ++k) {
    if (k == 7 && !p->ops->has_narrow_y_inv) continue;  // one tile width compiled
    Tuning trial = best;
    if (k == 4 || k == 5) {
      // chunked or whole-mesh pass pairs
      const int rule = chunk_rule(nm, k == 4 ? kChunkFwdMiB : kChunkInvMiB);
      knob(trial, k) = knob(best, k) ? 0 : rule;
      if (knob(trial, k) == knob(best, k)) continue;
    } else {
      knob(trial, k) = !knob(trial, k);
    }
    p->tune = trial;
    float ms = 1e30f;
    rc = measure(ms);
    if (rc == GNG_OK && ms < kMargin * best_ms) {
      best_ms = ms;
      best = trial;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  p->timing = timing;
  p->density_valid = false;
  p->potential_valid = false;
  if (rc != GNG_OK) {
    p->tune = cached_tuning(p->device, nm);
    return rc;
  }
  best.tuned = 1;
  p->tune = best;
  store_tuning(p->device, nm, best);
  return GNG_OK;
}

}  // namespace

extern "C" {

int gng_pm_autotune(gng_pm_plan* p, int reps) {
  if (!p) return fail(GNG_ERR_INVALID, "null plan");
  if (reps < 1 || reps > 100) return fail(GNG_ERR_INVALID, "reps must be in [1, 100]");
  std::lock_guard<std::mutex> lk(p->mu);
  int rc = set_device(p);
  if (rc) return rc;
  const int nm = p->cfg.grid_size;
  long long n = (long long)nm * nm * nm / 8;
  if (n < 4096) n = 4096;
  if ((rc = ensure_particles(p, n, n, false))) return rc;
  const size_t cap = (size_t)p->cap;
  float* d = p->d_part;
  {
    StreamOrder order(p, p->stream);
    hipLaunchKernelGGL(k_synth_particles, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, d, d + cap,
                       d + 2 * cap, d + 3 * cap, n, (float)p->cfg.box_size);
    GNG_HIP(hipGetLastError());
  }
  return autotune_impl(p, soa32(d, d + cap, d + 2 * cap, d + 3 * cap), n, acc32(d + 4 * cap, d + 5 * cap, d + 6 * cap),
                       reps);
}

int gng_pm_autotune_f64(gng_pm_plan* p, const double* d_pos, const double* d_mass, int64_t n, double* d_out,
                        int reps) {
  if (!p) return fail(GNG_ERR_INVALID, "null plan");
  if (reps < 1 || reps > 100) return fail(GNG_ERR_INVALID, "reps must be in [1, 100]");
  int rc = validate(p, n, nullptr, n);
  if (rc) return rc;
  if (n > 0 && (!d_pos || !d_mass || !d_out)) return fail(GNG_ERR_INVALID, "null array");
  std::lock_guard<std::mutex> lk(p->mu);
  if ((rc = set_device(p))) return rc;
  return autotune_impl(p, aos64(d_pos, d_mass), n, aos64_out(d_out), reps);
}

int gng_pm_set_tuning(gng_pm_plan* p, const gng_pm_tuning_t* t) {
  if (!p || !t) return fail(GNG_ERR_INVALID, "null argument");
  const int nm = p->cfg.grid_size;
  for (const int c : {t->plane_chunk_fwd, t->plane_chunk_inv})
    if (c < 0 || (c > 0 && (c >= nm || nm % c != 0 || ((long long)c * nm) % p->ops->x_rows != 0)))
      return fail(GNG_ERR_INVALID, "a plane chunk must be 0 or a proper divisor of grid_size whose planes hold whole "
                                   "x-pass workgroups");
  std::lock_guard<std::mutex> lk(p->mu);
  p->tune.z_persistent = t->z_persistent != 0;
  p->tune.deposit_wide = t->deposit_wide != 0;
  p->tune.y_fwd_persistent = t->y_fwd_persistent != 0;
  p->tune.y_inv_persistent = t->y_inv_persistent != 0;
  p->tune.interp_sorted = t->interp_sorted != 0;
  p->tune.chunk_fwd = t->plane_chunk_fwd;
  p->tune.chunk_inv = t->plane_chunk_inv;
  p->tune.deposit_split = t->deposit_split != 0;
  p->tune.y_inv_narrow = t->y_inv_narrow != 0;
  p->tune.deposit_pair = t->deposit_pair != 0;
  p->tune.tuned = 0;
  return GNG_OK;
}

int gng_pm_get_tuning(const gng_pm_plan* p, gng_pm_tuning_t* t) {
  if (!p || !t) return fail(GNG_ERR_INVALID, "null argument");
  std::memset(t, 0, sizeof(*t));
  t->z_persistent = p->tune.z_persistent;
  t->deposit_wide = p->tune.deposit_wide;
  t->y_fwd_persistent = p->tune.y_fwd_persistent;
  t->y_inv_persistent = p->tune.y_inv_persistent;
  t->tuned = p->tune.tuned;
  t->interp_sorted = p->tune.interp_sorted;
  t->plane_chunk_fwd = p->tune.chunk_fwd;
  t->plane_chunk_inv = p->tune.chunk_inv;
  t->deposit_split = p->tune.deposit_split;
  t->y_inv_narrow = p->tune.y_inv_narrow;
  t->deposit_pair = p->tune.deposit_pair;
  return GNG_OK;
}

int gng_pm_keep_potential(gng_pm_plan* p, int enable) {
  if (!p) return fail(GNG_ERR_INVALID, "null plan");
  std::lock_guard<std::mutex> lk(p->mu);
  int rc = set_device(p);
  if (rc) return rc;
  if (enable && !p->d_pot) GNG_HIP(hipMalloc(&p->d_pot, sizeof(float) * (size_t)p->nm3));
  p->keep_potential = enable != 0;
  p->potential_valid = false;
  return GNG_OK;
}

int gng_pm_potential(gng_pm_plan* p, double* out) {
  if (!p || !out) return fail(GNG_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(p->mu);
  if (!p->potential_valid) return fail(GNG_ERR_INVALID, "no kept potential: call gng_pm_keep_potential(p, 1) first");
  int rc = set_device(p);
  if (rc) return rc;
  std::vector<float> f((size_t)p->nm3);
  GNG_HIP(hipDeviceSynchronize());
  GNG_HIP(hipMemcpy(f.data(), p->d_pot, sizeof(float) * (size_t)p->nm3, hipMemcpyDeviceToHost));
  for (long long i = 0; i < p->nm3; ++i) out[i] = (double)f[i];
  return GNG_OK;
}

int gng_pm_device_meshes(gng_pm_plan* p, const float** d_mass, const float** d_potential) {
  if (!p) return fail(GNG_ERR_INVALID, "null plan");
  std::lock_guard<std::mutex> lk(p->mu);
  if (d_mass) {
    if (!p->density_valid) return fail(GNG_ERR_INVALID, "no kept density: call gng_pm_keep_density(p, 1) first");
    *d_mass = p->d_rho;
  }
  if (d_potential) {
    if (!p->potential_valid)
      return fail(GNG_ERR_INVALID, "no kept potential: call gng_pm_keep_potential(p, 1) first");
    *d_potential = p->d_pot;
  }
  return GNG_OK;
}

int gng_pm_cell_moments(gng_pm_plan* p, double* d_mom, void* stream) {
  if (!p || !d_mom) return fail(GNG_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(p->mu);
  if (p->sorted_n < 0) return fail(GNG_ERR_INVALID, "no particles sorted yet: solve first");
  int rc = set_device(p);
  if (rc) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  StreamOrder order(p, s);
  const unsigned ntiles = (unsigned)p->tm.nt * p->tm.nt * p->tm.nt;
  p->ops->cell_moments(dim3(ntiles, kCellMomParts), dim3(kCellMomThreads), s, p->d_sorted, p->d_offset, cell_map(p), p->tm,
                       p->d_sc, p->sorted_n, d_mom);
  GNG_HIP(hipGetLastError());
  return GNG_OK;
}

int gng_pm_force_meshes(gng_pm_plan* p, float* out) {
  if (!p || !out) return fail(GNG_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(p->mu);
  int rc = set_device(p);
  if (rc) return rc;
  GNG_HIP(hipDeviceSynchronize());
  // the meshes' z planes sit mesh_plane(nm) complex values apart
  const size_t nm = (size_t)p->cfg.grid_size, row = sizeof(float) * nm * nm;
  const size_t pitch = sizeof(float2) * (size_t)mesh_plane((int)nm);
  const float2* src[3] = {p->d_A, p->d_P, p->d_B};
  for (int f = 0; f < 3; ++f)
    GNG_HIP(hipMemcpy2D(out + f * p->nm3, row, src[f], pitch, row, nm, hipMemcpyDeviceToHost));
  return GNG_OK;
}

int gng_pm_set_timing(gng_pm_plan* p, int enable) {
  if (!p) return fail(GNG_ERR_INVALID, "null plan");
  std::lock_guard<std::mutex> lk(p->mu);
  p->timing = enable != 0;
  p->timed_solves = 0;
  return GNG_OK;
}

int gng_pm_stage_times(gng_pm_plan* p, float* ms, int capacity) {
  if (!p || (!ms && capacity > 0)) return fail(GNG_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(p->mu);
  int rc = set_device(p);
  if (rc) return rc;
  float t[kStages + 1];
  if ((rc = read_timing(p, t))) return -rc;
  for (int i = 0; i <= kStages && i < capacity; ++i) ms[i] = t[i];
  return kStages;
}

// ---- ensembles of boxes ---------------------------------------------------

}  // extern "C"

struct gng_pm_batch {
  std::vector<gng_pm_plan*> plans;
  int device = 0;
  int branches = 0;
  std::mutex mu;
  hipStream_t stream = nullptr;               // capture origin
  std::vector<hipStream_t> side;              // fork branches
  std::vector<hipEvent_t> fork_ev, join_ev;   // one per branch
  hipEvent_t root_ev = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // key of the captured graph: pointers, offsets and g
  std::vector<uintptr_t> key;
  // box tables: boxes grouped by mesh side, every stage one launch over up to
  // box_split boxes of one side (0: one pipeline per box)
  int box_split = 1 << 30;
  BoxArgs* d_box = nullptr;  // one BoxArgs per box, grouped by mesh side
  long long box_cap = 0;
};

namespace {

void free_batch(gng_pm_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (auto e : b->fork_ev) (void)hipEventDestroy(e);
  for (auto e : b->join_ev) (void)hipEventDestroy(e);
  if (b->root_ev) (void)hipEventDestroy(b->root_ev);
  for (auto s : b->side) (void)hipStreamDestroy(s);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  for (auto p : b->plans) free_plan(p);
  (void)hipFree(b->d_box);
  delete b;
}

// Boxes of one mesh side, in batch order.
struct BoxGroup {
  int nm;
  std::vector<int> boxes;
};

std::vector<BoxGroup> box_groups(const gng_pm_batch* b) {
  std::vector<BoxGroup> gs;
  for (int i = 0; i < (int)b->plans.size(); ++i) {
    const int nm = b->plans[i]->cfg.grid_size;
    auto it = std::find_if(gs.begin(), gs.end(), [&](const BoxGroup& g) { return g.nm == nm; });
    if (it == gs.end()) {
      gs.push_back({nm, {}});
      it = gs.end() - 1;
    }
    it->boxes.push_back(i);
  }
  return gs;
}

// Every box's table entry (synchronous upload; called before capture).
int upload_box_tables(gng_pm_batch* b, const std::vector<BoxGroup>& gs, const int64_t* offsets, double g) {
  std::vector<BoxArgs> h;
  for (const BoxGroup& gr : gs)
    for (int i : gr.boxes) {
      gng_pm_plan* p = b->plans[i];
      BoxArgs a{};
      a.P = p->d_P;
      a.A = p->d_A;
      a.B = p->d_B;
      a.gp = green_params(p, g);
      a.cm = cell_map(p);
      a.count = p->d_count;
      a.offset = p->d_offset;
      a.partial = p->d_partial;
      a.sc = p->d_sc;
      a.sorted = p->d_sorted;
      a.bucket = p->d_bucket;
      a.rank = p->d_rank;
      a.p0 = offsets[i];
      a.n = offsets[i + 1] - offsets[i];
      h.push_back(a);
    }
  if ((long long)h.size() > b->box_cap) {
    (void)hipFree(b->d_box);
    b->d_box = nullptr;
    b->box_cap = 0;
    GNG_HIP(hipMalloc(&b->d_box, sizeof(BoxArgs) * h.size()));
    b->box_cap = (long long)h.size();
  }
  GNG_HIP(hipMemcpy(b->d_box, h.data(), sizeof(BoxArgs) * h.size(), hipMemcpyHostToDevice));
  return GNG_OK;
}

// The whole pipeline of one group of boxes (one mesh side), every stage one
// launch over the group's boxes.  The launch choices are the first box's
// (boxes of one side share them); persistent and plane-chunked variants are
// not used here.
void launch_box_group(gng_pm_batch* b, const BoxGroup& gr, const BoxArgs* bx, long long maxn, const float* x,
                      const float* y, const float* z, const float* m, float* ax, float* ay, float* az,
                      hipStream_t s) {
  const gng_pm_plan* p0 = b->plans[gr.boxes[0]];
  const MeshOps& o = *p0->ops;
  const unsigned nbx = (unsigned)gr.boxes.size();
  const int nm = gr.nm, H = nm / 2;
  const PosSrc<float> src{x, y, z, 1};
  hipLaunchKernelGGL(k_reset_scalars_box, dim3(1, 1, nbx), dim3(64), 0, s, bx);
  const long long n16 = (p0->nb + 3) / 4;
  hipLaunchKernelGGL(k_zero_box, dim3((unsigned)((n16 + 255) / 256), 1, nbx), dim3(256), 0, s, bx, n16);
  if (maxn > 0)
    hipLaunchKernelGGL(k_bucket_count_box,
                       dim3((unsigned)((maxn + 256 * kBucketPer - 1) / (256 * kBucketPer)), 1, nbx), dim3(256), 0,
                       s, bx, src, m, p0->tm);
  hipLaunchKernelGGL(k_scan_local_box, dim3((unsigned)p0->scan_blocks, 1, nbx), dim3(256), 0, s, bx, p0->nb);
  hipLaunchKernelGGL(k_scan_partials_box, dim3(1, 1, nbx), dim3(256), 0, s, bx, p0->scan_blocks, p0->nb);
  hipLaunchKernelGGL(k_scan_add_box, dim3((unsigned)p0->scan_blocks, 1, nbx), dim3(256), 0, s, bx, p0->nb);
  if (maxn > 0)
    hipLaunchKernelGGL(k_scatter_box, dim3((unsigned)((maxn + 256 * kScatterPer - 1) / (256 * kScatterPer)), 1, nbx),
                       dim3(256), 0, s, bx, src, m);
  const unsigned ntiles = (unsigned)p0->tm.nt * p0->tm.nt * p0->tm.nt;
  o.deposit_box[p0->tune.deposit_wide](dim3(ntiles, 1, nbx), dim3(p0->tune.deposit_wide ? 512 : 256), s, bx,
                                       p0->tm);
  const long long rows = (long long)nm * nm;
  o.x_fwd_box(dim3((unsigned)(rows / o.x_rows), 1, nbx), dim3(256), s, bx, p0->d_tw);
  o.y_fwd_box(dim3((unsigned)(H / o.yz_cols), (unsigned)nm, nbx), dim3((unsigned)o.yz_threads), s, bx, p0->d_tw);
  o.z_green_box(dim3((unsigned)(H / o.z_cols), (unsigned)nm, nbx), dim3((unsigned)o.z_threads), s, bx, p0->d_tw);
  o.z_green_col0_box(dim3((unsigned)(H + 1), 1, nbx), dim3(256), s, bx, p0->d_tw);
  o.y_inv_box(dim3((unsigned)(H / o.yz_cols), (unsigned)nm, nbx), dim3((unsigned)o.yz_threads), s, bx, p0->d_tw);
  o.x_inv_box(dim3((unsigned)(rows / o.x_rows), 3, nbx), dim3(256), s, bx, p0->d_tw);
  if (maxn > 0)
    hipLaunchKernelGGL(k_interp_box, dim3((unsigned)((maxn + 255) / 256), 1, nbx), dim3(256), 0, s, bx, src,
                       AccDst<float>{ax, ay, az, 1});
}

}  // namespace

extern "C" {

int gng_pm_batch_set_box_tables(gng_pm_batch* b, int boxes_per_launch) {
  if (!b) return fail(GNG_ERR_INVALID, "null batch");
  if (boxes_per_launch < 0) return fail(GNG_ERR_INVALID, "boxes_per_launch must be >= 0");
  std::lock_guard<std::mutex> lk(b->mu);
  if (b->box_split != boxes_per_launch) {
    b->box_split = boxes_per_launch;
    (void)hipSetDevice(b->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->key.clear();
  }
  return GNG_OK;
}

int gng_pm_batch_get_box_tables(const gng_pm_batch* b) { return b ? b->box_split : -1; }

int gng_pm_batch_create(int n_boxes, const gng_pm_config* cfgs, gng_pm_batch** out) {
  if (!out) return fail(GNG_ERR_INVALID, "null output pointer");
  *out = nullptr;
  if (n_boxes <= 0 || !cfgs) return fail(GNG_ERR_INVALID, "need at least one box configuration");
  gng_pm_batch* b = new (std::nothrow) gng_pm_batch();
  if (!b) return fail(GNG_ERR_OOM, "host allocation failed");
  for (int i = 0; i < n_boxes; ++i) {
    gng_pm_plan* p = nullptr;
    const int rc = gng_pm_create(&cfgs[i], &p);
    if (rc) {
      const std::string msg = "box " + std::to_string(i) + ": " + g_last_error;
      free_batch(b);
      return fail(rc, msg);
    }
    b->plans.push_back(p);
  }
  b->device = b->plans[0]->device;
  b->branches = n_boxes < 8 ? n_boxes : 8;
  hipError_t e = hipSetDevice(b->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->root_ev, hipEventDisableTiming);
  for (int k = 0; k < b->branches && e == hipSuccess; ++k) {
    hipStream_t s;
    hipEvent_t f, j;
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&f, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&j, hipEventDisableTiming);
    if (e == hipSuccess) {
      b->side.push_back(s);
      b->fork_ev.push_back(f);
      b->join_ev.push_back(j);
    }
  }
  if (e != hipSuccess) {
    free_batch(b);
    return fail(GNG_ERR_HIP, std::string("batch setup: ") + hipGetErrorString(e));
  }
  *out = b;
  return GNG_OK;
}

void gng_pm_batch_destroy(gng_pm_batch* b) { free_batch(b); }

int gng_pm_batch_size(const gng_pm_batch* b) { return b ? (int)b->plans.size() : 0; }

int gng_pm_batch_plan(gng_pm_batch* b, int i, gng_pm_plan** out) {
  if (!b || !out || i < 0 || i >= (int)b->plans.size()) return fail(GNG_ERR_INVALID, "bad batch index");
  *out = b->plans[i];
  return GNG_OK;
}

int gng_pm_batch_autotune(gng_pm_batch* b, int reps) {
  if (!b) return fail(GNG_ERR_INVALID, "null batch");
  std::lock_guard<std::mutex> lk(b->mu);
  // one measurement per distinct mesh side, shared by the boxes of that side
  std::map<int, Tuning> done;
  for (gng_pm_plan* p : b->plans) {
    const int nm = p->cfg.grid_size;
    auto it = done.find(nm);
    if (it != done.end()) {
      std::lock_guard<std::mutex> pl(p->mu);
      p->tune = it->second;
      continue;
    }
    const int rc = gng_pm_autotune(p, reps);
    if (rc != GNG_OK) return rc;
    done[nm] = p->tune;
  }
  // the captured graph holds the old launch choices
  (void)hipSetDevice(b->device);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->key.clear();
  return GNG_OK;
}

int gng_pm_batch_solve_device(gng_pm_batch* b, const float* d_x, const float* d_y, const float* d_z,
                              const float* d_mass, const int64_t* offsets, double g, float* d_ax, float* d_ay,
                              float* d_az, void* stream) {
  if (!b || !offsets) return fail(GNG_ERR_INVALID, "null argument");
  const int nb = (int)b->plans.size();
  for (int i = 0; i < nb; ++i)
    if (offsets[i + 1] < offsets[i] || offsets[i] < 0) return fail(GNG_ERR_INVALID, "offsets must be non-decreasing");
  std::lock_guard<std::mutex> lk(b->mu);
  GNG_HIP(hipSetDevice(b->device));
  std::vector<uintptr_t> key = {(uintptr_t)d_x, (uintptr_t)d_y, (uintptr_t)d_z, (uintptr_t)d_mass,
                                (uintptr_t)d_ax, (uintptr_t)d_ay, (uintptr_t)d_az};
  uint64_t gbits;
  std::memcpy(&gbits, &g, sizeof(g));
  key.push_back((uintptr_t)gbits);
  for (int i = 0; i <= nb; ++i) key.push_back((uintptr_t)offsets[i]);
  if (!b->exec || key != b->key) {
    // (re)capture: every allocation happens before capture begins
    for (int i = 0; i < nb; ++i) {
      int rc = ensure_work(b->plans[i], offsets[i + 1] - offsets[i] > 0 ? offsets[i + 1] - offsets[i] : 1);
      if (rc) return rc;
      b->plans[i]->timing = false;
    }
    const std::vector<BoxGroup> groups = box_groups(b);
    if (b->box_split > 0) {
      // a launch of the previous graph may still read the old tables
      GNG_HIP(hipDeviceSynchronize());
      const int rc = upload_box_tables(b, groups, offsets, g);
      if (rc) return rc;
    }
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    GNG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = GNG_OK;
    hipError_t e = hipEventRecord(b->root_ev, b->stream);
    for (int k = 0; k < b->branches && e == hipSuccess; ++k) e = hipStreamWaitEvent(b->side[k], b->root_ev, 0);
    if (b->box_split > 0) {
      // each mesh side in runs of at most box_split boxes (contiguous in the
      // table); every run is one launch per stage, runs go to the branches
      // round robin
      long long first = 0;
      int run = 0;
      for (size_t gi = 0; gi < groups.size() && e == hipSuccess; ++gi) {
        const int ng = (int)groups[gi].boxes.size();
        for (int r0 = 0; r0 < ng; r0 += b->box_split, ++run) {
          BoxGroup part{groups[gi].nm, {}};
          long long maxn = 0;
          for (int r = r0; r < ng && r < r0 + b->box_split; ++r) {
            const int i = groups[gi].boxes[r];
            part.boxes.push_back(i);
            maxn = std::max<long long>(maxn, offsets[i + 1] - offsets[i]);
          }
          launch_box_group(b, part, b->d_box + first + r0, maxn, d_x, d_y, d_z, d_mass, d_ax, d_ay, d_az,
                           b->side[run % b->branches]);
        }
        first += ng;
      }
    } else {
      for (int i = 0; i < nb && e == hipSuccess && rc == GNG_OK; ++i) {
        const long long o = offsets[i], n = offsets[i + 1] - offsets[i];
        rc = run_pipeline(b->plans[i], soa32(d_x + o, d_y + o, d_z + o, d_mass + o), n, nullptr, n, g,
                          acc32(d_ax + o, d_ay + o, d_az + o), b->side[i % b->branches]);
      }
    }
    for (int k = 0; k < b->branches && e == hipSuccess; ++k) {
      e = hipEventRecord(b->join_ev[k], b->side[k]);
      if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->join_ev[k], 0);
    }
    hipGraph_t graph = nullptr;
    const hipError_t ee = hipStreamEndCapture(b->stream, &graph);
    if (e != hipSuccess || ee != hipSuccess || rc != GNG_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      if (rc) return rc;
      return fail(GNG_ERR_HIP, std::string("ensemble capture: ") + hipGetErrorString(e != hipSuccess ? e : ee));
    }
    b->graph = graph;
    GNG_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->key = key;
  }
  GNG_HIP(hipGraphLaunch(b->exec, reinterpret_cast<hipStream_t>(stream)));
  for (int i = 0; i < nb; ++i) b->plans[i]->last_n = offsets[i + 1] - offsets[i];
  return GNG_OK;
}

// Times whole ensemble solves (best of `reps` after one warm-up, HIP events)
// on the caller's particles for several launch splits: one table launch per
// stage per mesh side, halves and quarters of each side's boxes, and one
// pipeline per box; keeps the fastest.
int gng_pm_batch_tune_launches(gng_pm_batch* b, const float* d_x, const float* d_y, const float* d_z,
                               const float* d_mass, const int64_t* offsets, double g, float* d_ax, float* d_ay,
                               float* d_az, int reps, int* chosen) {
  if (!b || !offsets) return fail(GNG_ERR_INVALID, "null argument");
  if (reps < 1 || reps > 100) return fail(GNG_ERR_INVALID, "reps must be in [1, 100]");
  int biggest = 1;
  for (const BoxGroup& gr : box_groups(b)) biggest = std::max(biggest, (int)gr.boxes.size());
  std::vector<int> cand = {biggest, (biggest + 1) / 2, (biggest + 3) / 4, 0};
  std::sort(cand.begin(), cand.end(), std::greater<int>());
  cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
  GNG_HIP(hipSetDevice(b->device));
  // timed on the batch's own stream, which lives as long as the graphs it
  // launches (no stream of the tuner's own is left behind in a graph exec)
  hipStream_t s = b->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  int rc = he == hipSuccess ? GNG_OK : fail(GNG_ERR_HIP, "event creation failed");
  float best_ms = 1e30f;
  int best = b->box_split;
  for (int c : cand) {
    if (rc) break;
    if ((rc = gng_pm_batch_set_box_tables(b, c))) break;
    // the first solve captures the graph; the second warms it
    for (int w = 0; w < 2 && !rc; ++w)
      rc = gng_pm_batch_solve_device(b, d_x, d_y, d_z, d_mass, offsets, g, d_ax, d_ay, d_az, s);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = fail(GNG_ERR_HIP, "warm-up failed");
    float t_min = 1e30f;
    for (int r = 0; r < reps && !rc; ++r) {
      float ms = 0.0f;
      if (hipEventRecord(e0, s) != hipSuccess) rc = fail(GNG_ERR_HIP, "event record failed");
      if (!rc) rc = gng_pm_batch_solve_device(b, d_x, d_y, d_z, d_mass, offsets, g, d_ax, d_ay, d_az, s);
      if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                  hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = fail(GNG_ERR_HIP, "event timing failed");
      if (!rc) t_min = std::min(t_min, ms);
    }
    if (!rc && t_min < best_ms) {
      best_ms = t_min;
      best = c;
    }
  }
  (void)hipStreamSynchronize(s);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (chosen) *chosen = best;
  // the next solve captures a fresh graph for the chosen split
  std::lock_guard<std::mutex> lk(b->mu);
  b->box_split = best;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->key.clear();
  return GNG_OK;
}

// ---- device-memory helpers ------------------------------------------------

namespace {
__global__ void k_axpy(float* __restrict__ y, const float* __restrict__ x, long long n, float a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += a * x[i];
}
}  // namespace

int gng_dev_alloc(void** ptr, uint64_t bytes) {
  if (!ptr) return fail(GNG_ERR_INVALID, "null argument");
  *ptr = nullptr;
  GNG_HIP(hipMalloc(ptr, (size_t)bytes));
  return GNG_OK;
}

int gng_dev_free(void* ptr) {
  GNG_HIP(hipFree(ptr));
  return GNG_OK;
}

int gng_dev_copy(void* dst, const void* src, uint64_t bytes, int kind) {
  if (bytes == 0) return GNG_OK;
  if (!dst || !src) return fail(GNG_ERR_INVALID, "null argument");
  if (kind < 0 || kind > 2) return fail(GNG_ERR_INVALID, "bad copy kind");
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                        : hipMemcpyDeviceToDevice;
  GNG_HIP(hipMemcpy(dst, src, (size_t)bytes, k));
  return GNG_OK;
}

int gng_dev_synchronize(void) {
  GNG_HIP(hipDeviceSynchronize());
  return GNG_OK;
}

int gng_dev_axpy(float* d_y, const float* d_x, int64_t n, float a, void* stream) {
  if (n < 0) return fail(GNG_ERR_INVALID, "negative length");
  if (n == 0) return GNG_OK;
  if (!d_y || !d_x) return fail(GNG_ERR_INVALID, "null argument");
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                     d_y, d_x, (long long)n, a);
  GNG_HIP(hipGetLastError());
  return GNG_OK;
}

// ---- allocation-free solve ------------------------------------------------

}  // extern "C"

namespace {

// Every work buffer of gng_pm_solve_device_noalloc lives in this one static
// device array (one copy per GPU, held until the process ends).  Host code
// never passes its name to a kernel or a copy: the array's device address
// comes from hipGetSymbolAddress, once per device.
constexpr int kNoAllocMaxMesh = GNG_PM_NOALLOC_MAX_GRID;
constexpr long long kNoAllocMaxParticles = GNG_PM_NOALLOC_MAX_PARTICLES;
constexpr size_t kNoAllocBytes = GNG_PM_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_device(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(GNG_ERR_INVALID, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(GNG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return GNG_OK;
}

// Bump allocation of 256-byte aligned regions inside the scratch.
struct Carve {
  unsigned char* base;
  size_t used = 0;
  template <typename T>
  T* take(size_t count) {
    T* q = reinterpret_cast<T*>(base + used);
    used += (sizeof(T) * count + 255) / 256 * 256;
    return q;
  }
};

}  // namespace

extern "C" {

int gng_pm_solve_device_noalloc(const gng_pm_config* cfg, const float* d_x, const float* d_y, const float* d_z,
                                const float* d_mass, int64_t n, const int64_t* d_idx, int64_t k, double g,
                                float* d_ax, float* d_ay, float* d_az, void* stream) {
  if (!cfg) return fail(GNG_ERR_INVALID, "null config");
  if (!(cfg->box_size > 0.0) || !std::isfinite(cfg->box_size)) return fail(GNG_ERR_INVALID, "box_size must be > 0");
  const int nm = cfg->grid_size;
  const MeshOps* ops = find_ops(nm);
  if (!ops || nm > kNoAllocMaxMesh)
    return fail(GNG_ERR_INVALID, "the allocation-free solve takes a compiled grid_size of at most " +
                                     std::to_string(kNoAllocMaxMesh));
  if (n < 0 || k < 0) return fail(GNG_ERR_INVALID, "negative particle count");
  if (n > kNoAllocMaxParticles)
    return fail(GNG_ERR_INVALID, "the allocation-free solve takes at most " + std::to_string(kNoAllocMaxParticles) +
                                     " particles");
  if (!d_idx && k != n) return fail(GNG_ERR_INVALID, "idx is NULL but k != n");
  if (k >= (1LL << 31)) return fail(GNG_ERR_INVALID, "k of 2^31 or more is not supported");
  if (n > 0 && (!d_x || !d_y || !d_z || !d_mass)) return fail(GNG_ERR_INVALID, "null input array");
  if (k > 0 && (!d_ax || !d_ay || !d_az)) return fail(GNG_ERR_INVALID, "null output array");
  int dev = cfg->device;
  if (dev >= 0) {
    GNG_HIP(hipSetDevice(dev));
  } else {
    GNG_HIP(hipGetDevice(&dev));
  }
  unsigned char* base = nullptr;
  int rc = scratch_for_device(dev, &base);
  if (rc) return rc;
  // a plan that lives on the host stack and points into the scratch; its
  // work capacity is the particle limit, so the pipeline allocates nothing
  gng_pm_plan p;
  p.cfg = *cfg;
  p.ops = ops;
  p.device = dev;
  p.tune = default_tuning(nm);
  p.tm.T = ops->tile;
  p.tm.nt = nm / ops->tile;
  p.nb = (long long)p.tm.nt * p.tm.nt * p.tm.nt * 8;
  p.scan_blocks = (int)((p.nb + kScanBlock - 1) / kScanBlock);
  p.nm3 = (long long)nm * nm * nm;
  const size_t half = (size_t)nm * (size_t)mesh_plane(nm);
  Carve c{base};
  p.d_P = c.take<float2>(half);
  p.d_A = c.take<float2>(half);
  p.d_B = c.take<float2>(half);
  p.d_tw = c.take<float2>((size_t)nm);
  p.d_deconv = c.take<float>((size_t)nm);
  p.d_sc = c.take<SolveScalars>(1);
  p.d_count = c.take<unsigned>((size_t)((p.nb + 3) / 4 * 4));
  p.d_offset = c.take<unsigned>((size_t)(p.nb + 1));
  p.d_partial = c.take<unsigned>((size_t)p.scan_blocks);
  p.d_tile_rank = c.take<unsigned>((size_t)(p.nb / 8));
  p.d_tile_order = c.take<unsigned>((size_t)(p.nb / 8));
  p.d_tile_src = c.take<unsigned>((size_t)p.nb);
  p.d_tile_ranges = c.take<unsigned>((size_t)(p.nb / 8) * kRangeWords);
  p.d_sorted = c.take<PartRec>((size_t)kNoAllocMaxParticles);
  p.d_bucket = c.take<unsigned>((size_t)kNoAllocMaxParticles);
  p.d_rank = c.take<unsigned>((size_t)kNoAllocMaxParticles);
  if (c.used > kNoAllocBytes) return fail(GNG_ERR_INVALID, "allocation-free scratch layout exceeds its size");
  p.work_cap = kNoAllocMaxParticles;
  p.tm.rank = p.d_tile_rank;
  p.tm.order = p.d_tile_order;
  p.tm.src = p.d_tile_src;
  p.tm.ranges = p.d_tile_ranges;
  int bits = 0;
  while ((1 << bits) < p.tm.nt) ++bits;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // the tables are rebuilt on every call: the scratch keeps whatever the
  // last call on this GPU (any mesh size) left in it
  hipLaunchKernelGGL(k_plan_tables, dim3(1), dim3(256), 0, s, nm, p.tm.nt, bits, p.d_tw, p.d_deconv, p.d_tile_rank,
                     p.d_tile_order, p.d_tile_src);
  GNG_HIP(hipGetLastError());
  return run_pipeline(&p, soa32(d_x, d_y, d_z, d_mass), n, reinterpret_cast<const long long*>(d_idx), k, g,
                      acc32(d_ax, d_ay, d_az), s);
}

// ---- reference-compatible entry points ----------------------------------

void* hip_pm_create(int grid_size, float box_size) {
  gng_pm_config cfg;
  std::memset(&cfg, 0, sizeof(cfg));
  cfg.grid_size = grid_size;
  cfg.box_size = box_size;
  cfg.device = -1;
  gng_pm_plan* p = nullptr;
  if (gng_pm_create(&cfg, &p) != GNG_OK) {
    std::fprintf(stderr, "[gng_pm] hip_pm_create failed: %s\n", g_last_error.c_str());
    return nullptr;
  }
  return p;
}

void hip_pm_destroy(void* h) { gng_pm_destroy(static_cast<gng_pm_plan*>(h)); }

int hip_pm_solve(void* h, const float* x, const float* y, const float* z, const float* mass, float* ax, float* ay,
                 float* az, int n, float eps2, float g, float r_split) {
  (void)eps2;
  if (!h || n <= 0) return 0;
  gng_pm_plan* p = static_cast<gng_pm_plan*>(h);
  std::lock_guard<std::mutex> lk(p->mu);
  int rc = set_device(p);
  if (rc) return rc;
  // the reference passes r_split per call (pm_gravity.hip:363-367)
  const double keep = p->cfg.r_split;
  p->cfg.r_split = r_split > 0.0f ? (double)r_split : 0.0;
  rc = solve_host_soa(p, x, y, z, mass, n, nullptr, n, (double)g, ax, ay, az);
  p->cfg.r_split = keep;
  if (rc) std::fprintf(stderr, "[gng_pm] hip_pm_solve failed: %s\n", g_last_error.c_str());
  return rc;
}

}  // extern "C"
