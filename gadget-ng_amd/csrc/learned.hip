// learned.hip -- learned sub-cell force correction: a 3-layer network of
// periodic 3^3 convolutions on the PM mesh, trained and run on the MFMA units.
//
// Layers: conv(8->16) SiLU, conv(16->16) SiLU, conv(16->3).  Activations and
// weights are bfloat16, products accumulate in FP32 with
// v_mfma_f32_16x16x32_bf16.  The convolution is written as an implicit GEMM
// per 16-cell x row:  D[cell][o] = sum_k A[cell][k] B[k][o], k = tap*CIN + c,
// so one MFMA takes two taps of 16 channels (CIN = 16), four taps of 8
// channels or eight taps of 4.
//
// Input channels per cell (kIn = 8): the density contrast, the potential
// minus its 7^3 box mean, and, from the particles, the mass-weighted offsets
// of the particles from the cell centre (3 channels) and their mass-weighted
// spread about their own centre of mass, all per mean cell mass; the last
// two channels are zero.  A workgroup stages its 16x8x8
// output tile plus a one-cell periodic halo in LDS once and keeps every
// weight fragment of the layer in registers.  Inference skips the FP32
// pre-activations that only the backward pass reads.
//
// Backward: the data gradient is the same convolution with the taps mirrored
// (t -> 26 - t) and the channel roles swapped; the weight gradient is a GEMM
// whose K dimension runs over cells (dW[t][c][o] = sum_cells X[cell+off_t][c]
// dZ[cell][o]), accumulated per workgroup and reduced over workgroups in a
// fixed order, so gradients are deterministic.  The loss gradient reaches
// the mesh through the transpose of the CIC interpolation, with runs of
// particles that share a base cell combined inside the wavefront before the
// atomics.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/gng_learned.h"
#include "status.hpp"

namespace gng {
namespace net {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// forward / data-gradient tile: 16 x 8 x 8 output cells
constexpr int TX = 16, TY = 8, TZ = 8;
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2, HC = HX * HY * HZ;
// weight-gradient tile: 16 x 4 x 4 cells
constexpr int WX = 16, WY = 4, WZ = 4;
constexpr int WHX = WX + 2, WHY = WY + 2, WHZ = WZ + 2, WHC = WHX * WHY * WHZ;
constexpr int kWgradBlocks = 512;

// parameter layout (floats)
constexpr int kIn = 8;  // input channels (6 used)
constexpr int P_W1 = 0, P_B1 = P_W1 + 27 * kIn * 16, P_W2 = P_B1 + 16, P_B2 = P_W2 + 27 * 16 * 16,
              P_W3 = P_B2 + 16, P_B3 = P_W3 + 27 * 16 * 3, P_COUNT = P_B3 + 3;

// k-steps of 32 (forward and data-gradient convolutions)
template <int CIN>
struct KSteps {
  static constexpr int v = (27 * CIN + 31) / 32;
};

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// FP32 -> bfloat16, round to nearest even (gfx950's v_cvt_pk_bf16_f32)
__device__ __forceinline__ unsigned short f2bf(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }
// two values in one conversion: a in the low half, b in the high half
__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
  return __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2{a, b}, bf16x2));
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }
// sigma(z) = 1 / (1 + e^-z) with the hardware reciprocal (1 ulp); the
// activations are stored in bfloat16, 16 bits coarser
__device__ __forceinline__ float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

// halo-cell offset of tap t in a box of row length hx and plane hx*hy
__host__ __device__ constexpr int tap_off(int t, int hx, int hy) {
  return ((t / 9 - 1) * hy + ((t / 3) % 3 - 1)) * hx + (t % 3 - 1);
}

// Stage input cells (x0-1 .. x0+HXv-2, ...) mod nm, CIN bf16 channels each,
// into lds[((hz*HYv + hy)*HXv + hx)*CIN + c], moving 32-bit words.  With
// PLANAR (16 channels only) channels 0-7 and 8-15 go to two separate planes
// of 16-byte words, cell h at word h and HXv*HYv*HZv + h: sixteen lanes that
// read one channel half of sixteen consecutive cells then read 256
// consecutive bytes, which no two of them share a bank for.
template <int CIN>
struct HaloWord {  // 16-byte words for 8 and 16 channels, one word per cell for fewer
  using T = typename std::conditional<CIN >= 8, u32x4, typename std::conditional<CIN == 4, u32x2, unsigned>::type>::type;
  static constexpr int per_cell = CIN == 16 ? 2 : 1;
};

// All of a thread's global loads are issued before any is written to LDS,
// so the halo costs one memory latency rather than one per word.  NT is the
// workgroup size.
template <int CIN, int HXv, int HYv, int HZv, int NT, bool PLANAR = false>
__device__ __forceinline__ void load_halo(unsigned short* lds, const unsigned short* __restrict__ in, int x0,
                                          int y0, int z0, int nm) {
  using W = typename HaloWord<CIN>::T;
  constexpr int WPC = HaloWord<CIN>::per_cell;
  constexpr int NCELL = HXv * HYv * HZv, NW = NCELL * WPC, PER = (NW + NT - 1) / NT;
  static_assert(!PLANAR || CIN == 16, "the planar layout is for 16 channels");
  const W* src = reinterpret_cast<const W*>(in);
  W* dst = reinterpret_cast<W*>(lds);
  W v[PER];
  // Tiles away from the box faces need no periodic wrap: their halo is one
  // affine block of cells (cell indices inside a box stay below nm^3 <= 2^30).
  const bool interior = x0 >= 1 && x0 + HXv - 1 <= nm && y0 >= 1 && y0 + HYv - 1 <= nm && z0 >= 1 &&
                        z0 + HZv - 1 <= nm;
  if (interior) {
    const unsigned base = ((unsigned)(z0 - 1) * (unsigned)nm + (unsigned)(y0 - 1)) * (unsigned)nm + (unsigned)(x0 - 1);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = (int)threadIdx.x + k * NT;
      if (NW % NT == 0 || e < NW) {
        const int c = e / WPC, w = e - c * WPC;
        const int hx = c % HXv, hy = (c / HXv) % HYv, hz = c / (HXv * HYv);
        const unsigned gc = base + ((unsigned)hz * (unsigned)nm + (unsigned)hy) * (unsigned)nm + (unsigned)hx;
        v[k] = src[(size_t)gc * WPC + w];
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = (int)threadIdx.x + k * NT;
      if (NW % NT == 0 || e < NW) {
        const int c = e / WPC, w = e - c * WPC;
        const int hx = c % HXv, hy = (c / HXv) % HYv, hz = c / (HXv * HYv);
        int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
        gx += gx < 0 ? nm : 0;
        gx -= gx >= nm ? nm : 0;
        gy += gy < 0 ? nm : 0;
        gy -= gy >= nm ? nm : 0;
        gz += gz < 0 ? nm : 0;
        gz -= gz >= nm ? nm : 0;
        const unsigned gc = ((unsigned)gz * (unsigned)nm + (unsigned)gy) * (unsigned)nm + (unsigned)gx;
        v[k] = src[(size_t)gc * WPC + w];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int e = (int)threadIdx.x + k * NT;
    if (NW % NT == 0 || e < NW) {
      const int c = e / WPC, w = e - c * WPC;
      dst[PLANAR ? w * NCELL + c : e] = v[k];
    }
  }
}

// Fragment of k-step s for the 16-cell row whose lane-m cell sits at halo
// index `base`: lane (m, g) holds x[cell m + off_t][c] for k = 32s + 8g + j,
// t = k / CIN, c = k % CIN, j = 0..7 (zero for t >= 27).  For CIN = 16 the
// halo is in the planar layout of load_halo.
template <int CIN>
__device__ __forceinline__ bf16x8 a_frag(const unsigned short* tile, int base, int s, int g) {
  u32x4 v = {0u, 0u, 0u, 0u};
  if constexpr (CIN == 16) {
    // taps 2s and 2s + 1, channel halves by g & 1
    const int t = 2 * s + (g >> 1);
    const int off = (g >> 1) ? tap_off(2 * s + 1, HX, HY) : tap_off(2 * s, HX, HY);
    if (t < 27) v = reinterpret_cast<const u32x4*>(tile)[(g & 1) * HC + base + off];
  } else if constexpr (CIN == 4) {
    const int t0 = 8 * s + 2 * g;
    if (t0 < 27) {
      const u32x2 a = *reinterpret_cast<const u32x2*>(tile + (base + tap_off(t0, HX, HY)) * 4);
      v.x = a.x;
      v.y = a.y;
    }
    if (t0 + 1 < 27) {
      const u32x2 b = *reinterpret_cast<const u32x2*>(tile + (base + tap_off(t0 + 1, HX, HY)) * 4);
      v.z = b.x;
      v.w = b.y;
    }
  } else {
    static_assert(CIN == 8, "CIN must be 4, 8 or 16");
    // tap 4s + g, all 8 channels in one 16-byte word
    const int t = 4 * s + g;
    if (t < 27) v = reinterpret_cast<const u32x4*>(tile)[base + tap_off(t, HX, HY)];
  }
  return __builtin_bit_cast(bf16x8, v);
}

enum { MODE_HIDDEN = 0, MODE_OUT = 1, MODE_BWD = 2, MODE_INFER = 3 };


// One periodic 3^3 convolution over an nm^3 mesh (nm % 16 == 0), grid
// (nm/16, (nm/8)^2, boxes), kConvThreads threads (8 waves, 8 of the tile's 64
// rows each).  The MFMA computes the transposed product D^T[o][cell] = W^T X^T
// (the weight fragments as the first operand), so lane (m, g) ends with
// output channels 4g..4g+3 of cell m of the row and stores them as one
// vector: a wave writes one contiguous run of 16 cells.
//   MODE_HIDDEN  zout = conv + bias (f32, 16 ch), xout = bf16(silu(zout));
//   MODE_INFER   xout = bf16(silu(conv + bias)) only;
//   MODE_OUT     zout = conv + bias (f32, 4 ch: the 3 outputs and a zero);
//   MODE_BWD     xout = bf16(conv * silu'(zprev)) (16 ch).
constexpr int kConvThreads = 512;
constexpr int kRowsPerWave = TY * TZ / (kConvThreads / 64);
// rows of a wave that share one pipelined k-step loop (one accumulator each)
constexpr int kRowGroup = 2;
static_assert(kRowsPerWave % kRowGroup == 0, "row groups must tile a wave's rows");

template <int CIN, int MODE>
__global__ __launch_bounds__(kConvThreads, 4) void k_conv(const unsigned short* __restrict__ in,
                                              const u32x4* __restrict__ wfrag, const float* __restrict__ bias,
                                              float* __restrict__ zout, unsigned short* __restrict__ xout,
                                              const float* __restrict__ zprev, int nm) {
  constexpr int KS = KSteps<CIN>::v;
  __shared__ __attribute__((aligned(16))) unsigned short tile[HC * CIN];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int nty = nm / TY;
  const int x0 = blockIdx.x * TX, y0 = (blockIdx.y % nty) * TY, z0 = (blockIdx.y / nty) * TZ;
  const long long M = (long long)nm * nm * nm;
  const long long boff = (long long)blockIdx.z * M;

  bf16x8 b[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) b[s] = __builtin_bit_cast(bf16x8, wfrag[s * 64 + lane]);
  f32x4 bn = {0.0f, 0.0f, 0.0f, 0.0f};
  if constexpr (MODE != MODE_BWD) bn = *reinterpret_cast<const f32x4*>(bias + 4 * g);

  load_halo<CIN, HX, HY, HZ, kConvThreads, CIN == 16>(tile, in + boff * CIN, x0, y0, z0, nm);
  __syncthreads();

  // each wave: kRowsPerWave of the tile's rows, kRowGroup accumulators at a time
  constexpr int RG = kRowGroup;
#pragma unroll 1
  for (int q = 0; q < kRowsPerWave / RG; ++q) {
    f32x4 acc[RG];
    int base[RG];
#pragma unroll
    for (int a = 0; a < RG; ++a) {
      const int r = wave * kRowsPerWave + q * RG + a;
      const int ry = r % TY, rz = r / TY;
      base[a] = ((rz + 1) * HY + (ry + 1)) * HX + (m + 1);
      acc[a] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    // the fragments of k-step s + 1 are read while the MFMAs of step s run
    bf16x8 cur[RG];
#pragma unroll
    for (int a = 0; a < RG; ++a) cur[a] = a_frag<CIN>(tile, base[a], 0, g);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      bf16x8 nxt[RG];
      if (s + 1 < KS) {
#pragma unroll
        for (int a = 0; a < RG; ++a) nxt[a] = a_frag<CIN>(tile, base[a], s + 1, g);
      }
      // keep the scheduler from sinking those reads down to their uses
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int a = 0; a < RG; ++a) acc[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[s], cur[a], acc[a], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      if (s + 1 < KS) {
#pragma unroll
        for (int a = 0; a < RG; ++a) cur[a] = nxt[a];
      }
    }
    // lane (m, g) holds D^T[channel 4g + j][cell m]
#pragma unroll
    for (int a = 0; a < RG; ++a) {
      const int r = wave * kRowsPerWave + q * RG + a;
      const int ry = r % TY, rz = r / TY;
      const long long c = boff + ((long long)(z0 + rz) * nm + (y0 + ry)) * nm + x0 + m;
      if constexpr (MODE == MODE_OUT) {
        // channels 0-3: the 3 outputs and the zero pad column
        if (g == 0) *reinterpret_cast<f32x4*>(zout + c * 4) = acc[a] + bn;
      } else {
        f32x4 v;
        if constexpr (MODE == MODE_BWD) {
          const f32x4 zp = *reinterpret_cast<const f32x4*>(zprev + c * 16 + 4 * g);
#pragma unroll
          for (int j = 0; j < 4; ++j) v[j] = acc[a][j] * silu_grad(zp[j]);
        } else {
          const f32x4 z = acc[a] + bn;
          if constexpr (MODE == MODE_HIDDEN) *reinterpret_cast<f32x4*>(zout + c * 16 + 4 * g) = z;
#pragma unroll
          for (int j = 0; j < 4; ++j) v[j] = silu(z[j]);
        }
        u32x2 packed;
        packed.x = pack_bf16(v[0], v[1]);
        packed.y = pack_bf16(v[2], v[3]);
        *reinterpret_cast<u32x2*>(xout + c * 16 + 4 * g) = packed;
      }
    }
  }
}

// Weight gradient partials.  Columns n = t*CIN + c (27*CIN of them) plus a
// bias column of ones; rows o = output channel (< COS).  Each workgroup walks
// tiles in a grid-stride loop and writes its D[o][n] partial sums:
//   partial[block][n / 16][o * 16 + n % 16].
template <int CIN, int COS>
__global__ __launch_bounds__(256) void k_conv_wgrad(const unsigned short* __restrict__ X,
                                                    const unsigned short* __restrict__ dZ,
                                                    float* __restrict__ partial, int nm, int boxes) {
  constexpr int NCOL = 27 * CIN + 1, NB = (NCOL + 15) / 16, SPW = (NB + 3) / 4;
  __shared__ __attribute__((aligned(16))) unsigned short xt[WHC * CIN];
  __shared__ __attribute__((aligned(16))) unsigned short dt[WX * WY * WZ * COS];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;

  int coff[SPW], cch[SPW], ckind[SPW];  // kind 0: tap column, 1: ones, 2: zero
#pragma unroll
  for (int q = 0; q < SPW; ++q) {
    const int n = 16 * (wave + 4 * q) + m;
    coff[q] = 0;
    cch[q] = 0;
    if (n < 27 * CIN) {
      coff[q] = tap_off(n / CIN, WHX, WHY);
      cch[q] = n % CIN;
      ckind[q] = 0;
    } else {
      ckind[q] = n == 27 * CIN ? 1 : 2;
    }
  }
  f32x4 acc[SPW];
#pragma unroll
  for (int q = 0; q < SPW; ++q) acc[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

  const int ntx = nm / WX, nty = nm / WY, ntz = nm / WZ;
  const long long per_box = (long long)ntx * nty * ntz;
  const long long tiles = per_box * boxes;
  const long long M = (long long)nm * nm * nm;
  for (long long tt = blockIdx.x; tt < tiles; tt += gridDim.x) {
    const long long box = tt / per_box;
    const long long rem = tt - box * per_box;
    const int bx = (int)(rem % ntx), by = (int)((rem / ntx) % nty), bz = (int)(rem / ((long long)ntx * nty));
    const int x0 = bx * WX, y0 = by * WY, z0 = bz * WZ;
    __syncthreads();
    load_halo<CIN, WHX, WHY, WHZ, 256>(xt, X + box * M * CIN, x0, y0, z0, nm);
    {
      constexpr int WPC = COS / 2, NW = WX * WY * WZ * WPC, PER = NW / 256;
      static_assert(NW % 256 == 0, "gradient tile words must fill the workgroup");
      const unsigned* src = reinterpret_cast<const unsigned*>(dZ + box * M * COS);
      unsigned* dst = reinterpret_cast<unsigned*>(dt);
      unsigned v[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int e = (int)threadIdx.x + k * 256;
        const int c = e / WPC, w = e - c * WPC;
        const int cx = c % WX, cy = (c / WX) % WY, cz = c / (WX * WY);
        const long long gc = ((long long)(z0 + cz) * nm + (y0 + cy)) * nm + x0 + cx;
        v[k] = src[gc * WPC + w];
      }
#pragma unroll
      for (int k = 0; k < PER; ++k) dst[threadIdx.x + k * 256] = v[k];
    }
    __syncthreads();
#pragma unroll 1
    for (int r = 0; r < WY * WZ; ++r) {
      const int ry = r % WY, rz = r / WY;
      bf16x4 av;
#pragma unroll
      for (int j = 0; j < 4; ++j) av[j] = m < COS ? (short)dt[((rz * WY + ry) * WX + 4 * g + j) * COS + m] : (short)0;
      const int hb = ((rz + 1) * WHY + (ry + 1)) * WHX + 1 + 4 * g;
#pragma unroll
      for (int q = 0; q < SPW; ++q) {
        if (wave + 4 * q < NB) {
          bf16x4 bv;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            short v = 0;
            if (ckind[q] == 0) v = (short)xt[(hb + j + coff[q]) * CIN + cch[q]];
            else if (ckind[q] == 1) v = (short)0x3F80;
            bv[j] = v;
          }
          acc[q] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[q], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < SPW; ++q) {
    const int s = wave + 4 * q;
    if (s < NB) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        partial[((long long)blockIdx.x * NB + s) * 256 + (4 * g + j) * 16 + m] = acc[q][j];
    }
  }
}

// dW[t][c][o] (t*CIN + c = column) and db[o] from the partials, summed over
// workgroups in order in FP64.
template <int CIN, int COUT>
__global__ void k_wgrad_reduce(const float* __restrict__ partial, int nblocks, float* __restrict__ gw,
                               float* __restrict__ gb) {
  constexpr int NCOL = 27 * CIN + 1, NB = (NCOL + 15) / 16;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= NCOL * COUT) return;
  const int col = e / COUT, o = e % COUT;
  const int s = col / 16, nn = col % 16;
  double sum = 0.0;
  for (int b = 0; b < nblocks; ++b) sum += partial[((long long)b * NB + s) * 256 + o * 16 + nn];
  if (col < 27 * CIN) gw[col * COUT + o] = (float)sum;
  else gb[o] = (float)sum;
}

// Fragments of B[k][n] for one convolution from FP32 weights w[t][c][o]
// (cin x cout real channels, cin_pad channels in the k layout).  With
// mirror != 0 the fragments are those of the data-gradient convolution:
// B[k = t*cin_pad + o][n = c] = w[26 - t][c][o].
__global__ void k_pack(const float* __restrict__ w, int cin, int cout, int cin_pad, int ksteps, int mirror,
                       u32x4* __restrict__ frag) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ksteps * 64) return;
  const int s = e / 64, l = e % 64, g = l / 16, n = l % 16;
  unsigned short v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = 32 * s + 8 * g + j;
    const int t = k / cin_pad, c = k % cin_pad;
    float x = 0.0f;
    if (!mirror) {
      if (t < 27 && c < cin && n < cout) x = w[(t * cin + c) * cout + n];
    } else {
      // data gradient: input channels are the forward outputs (c < cout),
      // outputs the forward inputs (n < cin)
      if (t < 27 && c < cout && n < cin) x = w[((26 - t) * cin + n) * cout + c];
    }
    v[j] = f2bf(x);
  }
  u32x4 o;
#pragma unroll
  for (int q = 0; q < 4; ++q) o[q] = (unsigned)v[2 * q] | ((unsigned)v[2 * q + 1] << 16);
  frag[e] = o;
}

__global__ void k_bias16(const float* __restrict__ b, int cout, float* __restrict__ b16) {
  const int n = threadIdx.x;
  if (n < 16) b16[n] = n < cout ? b[n] : 0.0f;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m1,
                       float* __restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gv = gr[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gv;
  const float v = b2 * m2[i] + (1.0f - b2) * gv * gv;
  m1[i] = a;
  m2[i] = v;
  p[i] -= lr * (a * c1) / (sqrtf(v * c2) + eps);
}

__global__ void k_fill(float* __restrict__ p, long long n, float v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_d(double* p) { *p = 0.0; }

// dz3 = bf16(dL/dy), 4 channels per cell
__global__ void k_to_bf16(const float* __restrict__ a, unsigned short* __restrict__ b, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

// Periodic 7-point running mean along one axis (stride 1, nm or nm^2).
constexpr int kHalfWidth = 3;
__global__ void k_boxmean(const float* __restrict__ in, float* __restrict__ out, int nm, long long stride) {
  const long long M = (long long)nm * nm * nm;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  const int c = (int)((i / stride) % nm);
  const long long base = i - (long long)c * stride;
  float s = 0.0f;
#pragma unroll
  for (int d = -kHalfWidth; d <= kHalfWidth; ++d) {
    int j = c + d;
    j += j < 0 ? nm : 0;
    j -= j >= nm ? nm : 0;
    s += in[base + (long long)j * stride];
  }
  out[i] = s * (1.0f / (2 * kHalfWidth + 1));
}

// Particle moments per cell, as 64-bit fixed point (value * 2^shift) summed
// with integer atomics, so the sums do not depend on the order of the adds:
// mom[5 c + a] = sum m (f_a - 1/2) (a = x, y, z), mom[5 c + 3] =
// sum m |f - 1/2|^2, mom[5 c + 4] = sum m, over the particles whose base cell
// (the cell of cic_nodes) is c; f = in-cell fraction.  A workgroup first
// merges its particles by cell in an LDS hash table (open addressing, twice
// as many slots as particles, so every insert finds a slot), then adds one
// set of sums per distinct cell to the mesh: particles crowded into a few
// cells cost a few global atomics per workgroup instead of one set each.
constexpr int kMoments = 5;
constexpr int kMomThreads = 256, kMomSlots = 2 * kMomThreads;
__global__ __launch_bounds__(kMomThreads) void k_moments(const float* __restrict__ x, const float* __restrict__ y,
                                                         const float* __restrict__ z, const float* __restrict__ m,
                                                         long long n, float cs, int nm, double scale,
                                                         unsigned long long* __restrict__ mom) {
  constexpr unsigned kEmpty = 0xFFFFFFFFu;
  __shared__ unsigned key[kMomSlots];
  __shared__ unsigned long long acc[kMomSlots][kMoments];
  for (int i = threadIdx.x; i < kMomSlots; i += kMomThreads) {
    key[i] = kEmpty;
#pragma unroll
    for (int a = 0; a < kMoments; ++a) acc[i][a] = 0ull;
  }
  __syncthreads();
  const long long p = (long long)blockIdx.x * kMomThreads + threadIdx.x;
  if (p < n) {
    const float fx = x[p] * cs, fy = y[p] * cs, fz = z[p] * cs;
    const float ix = floorf(fx), iy = floorf(fy), iz = floorf(fz);
    const double dx = (double)(fx - ix) - 0.5, dy = (double)(fy - iy) - 0.5, dz = (double)(fz - iz) - 0.5;
    int i0 = (int)ix % nm, j0 = (int)iy % nm, k0 = (int)iz % nm;
    i0 += i0 < 0 ? nm : 0;
    j0 += j0 < 0 ? nm : 0;
    k0 += k0 < 0 ? nm : 0;
    const unsigned c = (unsigned)(((long long)k0 * nm + j0) * nm + i0);  // nm^3 <= 2^30
    unsigned h = (c * 2654435761u) >> 23;  // top 9 bits: a slot in [0, 512)
    for (;;) {
      const unsigned prev = atomicCAS(&key[h], kEmpty, c);
      if (prev == kEmpty || prev == c) break;
      h = (h + 1) & (kMomSlots - 1);
    }
    const double ms = (double)m[p] * scale;
    atomicAdd(&acc[h][0], (unsigned long long)__double2ll_rn(ms * dx));
    atomicAdd(&acc[h][1], (unsigned long long)__double2ll_rn(ms * dy));
    atomicAdd(&acc[h][2], (unsigned long long)__double2ll_rn(ms * dz));
    atomicAdd(&acc[h][3], (unsigned long long)__double2ll_rn(ms * (dx * dx + dy * dy + dz * dz)));
    atomicAdd(&acc[h][4], (unsigned long long)__double2ll_rn(ms));
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kMomSlots; i += kMomThreads) {
    const unsigned c = key[i];
    if (c == kEmpty) continue;
    unsigned long long* dst = mom + (long long)c * kMoments;
#pragma unroll
    for (int a = 0; a < kMoments; ++a) atomicAdd(dst + a, acc[i][a]);
  }
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// x0 channels: mass / rho_mean - 1, (phi - box7(phi)) / phi_scale, and from
// the moments (when mom is not null, else zeros) D_a / rho_mean for
// D = sum m (f - 1/2), and (sum m |f - 1/2|^2 - |D|^2 / sum m) / rho_mean,
// the spread about the particles' centre of mass; channels 6 and 7 are zero.
// The potential minus its 7^3 box mean keeps the part set by mass within a
// few cells and drops the box-scale modes, which carry no sub-cell
// information.
// The moments come as 64-bit fixed-point sums (mom, times mom_unit =
// 2^-shift / rho_mean) or as doubles (momd, times mom_unit = 1 / rho_mean),
// five per cell in the order of k_moments.
__global__ void k_input(const float* __restrict__ rho, const float* __restrict__ phi,
                        const float* __restrict__ phi_mean, const unsigned long long* __restrict__ mom,
                        const double* __restrict__ momd, double mom_unit, long long M, float inv_rho,
                        float inv_phi, u32x4* __restrict__ x0) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  float c2 = 0.0f, c3 = 0.0f, c4 = 0.0f, c5 = 0.0f;
  if (mom || momd) {
    double v[kMoments];
    if (mom) {
      const long long* q = reinterpret_cast<const long long*>(mom) + i * kMoments;
#pragma unroll
      for (int k = 0; k < kMoments; ++k) v[k] = (double)q[k] * mom_unit;
    } else {
#pragma unroll
      for (int k = 0; k < kMoments; ++k) v[k] = momd[i * kMoments + k] * mom_unit;
    }
    const double dx = v[0], dy = v[1], dz = v[2], m2 = v[3], m0 = v[4];
    const double spread = m0 > 0.0 ? m2 - (dx * dx + dy * dy + dz * dz) / m0 : 0.0;
    c2 = (float)dx;
    c3 = (float)dy;
    c4 = (float)dz;
    c5 = (float)(spread > 0.0 ? spread : 0.0);
  }
  u32x4 v;
  v.x = pack_bf16(rho[i] * inv_rho - 1.0f, (phi[i] - phi_mean[i]) * inv_phi);
  v.y = pack_bf16(c2, c3);
  v.z = pack_bf16(c4, c5);
  v.w = 0u;
  x0[i] = v;
}

struct Cic {
  long long node[8];
  float w[8];
};

// CIC nodes and weights of one position (cic.rs:100-170 weights; float cell
// coordinate, all indices wrapped).
__device__ __forceinline__ void cic_nodes(float x, float y, float z, float cs, int nm, long long boff, Cic& c) {
  float fx = x * cs, fy = y * cs, fz = z * cs;
  const float ix = floorf(fx), iy = floorf(fy), iz = floorf(fz);
  const float dx = fx - ix, dy = fy - iy, dz = fz - iz;
  int i0 = (int)ix % nm, j0 = (int)iy % nm, k0 = (int)iz % nm;
  i0 += i0 < 0 ? nm : 0;
  j0 += j0 < 0 ? nm : 0;
  k0 += k0 < 0 ? nm : 0;
  const int i1 = i0 + 1 == nm ? 0 : i0 + 1, j1 = j0 + 1 == nm ? 0 : j0 + 1, k1 = k0 + 1 == nm ? 0 : k0 + 1;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int kk = (a & 4) ? k1 : k0, jj = (a & 2) ? j1 : j0, ii = (a & 1) ? i1 : i0;
    c.node[a] = boff + ((long long)kk * nm + jj) * nm + ii;
    c.w[a] = ((a & 4) ? dz : 1.0f - dz) * ((a & 2) ? dy : 1.0f - dy) * ((a & 1) ? dx : 1.0f - dx);
  }
}

// Loss and dL/dy.  Lanes whose particles share a base cell (training sets are
// stored cell-sorted) form runs; a segmented suffix sum over each run leaves
// the run total in its first lane, which alone issues the atomics.
__global__ __launch_bounds__(256) void k_loss(const float4* __restrict__ pos, const float4* __restrict__ tgt,
                                              long long n, const float4* __restrict__ y, float* __restrict__ gy,
                                              double* __restrict__ loss, int nm, float cs, float out_scale,
                                              float gscale, int max_boxes) {
  __shared__ double red[4];
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const long long M = (long long)nm * nm * nm;
  float4 q = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
  if (p < n) q = pos[p];
  // records whose box index is not a valid slot are ignored
  const bool live = p < n && q.w >= 0.0f && q.w < (float)max_boxes;
  Cic c;
  float gx = 0.0f, gyv = 0.0f, gz = 0.0f;
  double lp = 0.0;
  long long key = -1 - lane;
  if (live) {
    const float4 t = tgt[p];
    cic_nodes(q.x, q.y, q.z, cs, nm, (long long)(int)q.w * M, c);
    float px = 0.0f, py = 0.0f, pz = 0.0f;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const float4 v = y[c.node[a]];
      px += c.w[a] * v.x;
      py += c.w[a] * v.y;
      pz += c.w[a] * v.z;
    }
    const float ex = out_scale * px - t.x, ey = out_scale * py - t.y, ez = out_scale * pz - t.z;
    lp = (double)t.w * ((double)ex * ex + (double)ey * ey + (double)ez * ez);
    gx = gscale * t.w * ex;
    gyv = gscale * t.w * ey;
    gz = gscale * t.w * ez;
    key = c.node[0];
  } else {
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      c.node[a] = 0;
      c.w[a] = 0.0f;
    }
  }
  // run boundaries
  const long long kprev = __shfl_up(key, 1);
  const bool head = lane == 0 || kprev != key;
  const unsigned long long heads = __ballot(head);
  const unsigned long long above = lane == 63 ? 0ull : (heads >> (lane + 1)) << (lane + 1);
  const int run_end = above ? __ffsll((long long)above) - 2 : 63;
  float v[24];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    v[3 * a] = c.w[a] * gx;
    v[3 * a + 1] = c.w[a] * gyv;
    v[3 * a + 2] = c.w[a] * gz;
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const bool take = lane + off <= run_end;
#pragma unroll
    for (int e = 0; e < 24; ++e) {
      const float o = __shfl_down(v[e], off);
      if (take) v[e] += o;
    }
  }
  if (live && head) {
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      float* dst = gy + c.node[a] * 4;
      atomicAdd(dst, v[3 * a]);
      atomicAdd(dst + 1, v[3 * a + 1]);
      atomicAdd(dst + 2, v[3 * a + 2]);
    }
  }
  // loss: wave then block reduction, one FP64 atomic per block
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lp += __shfl_down(lp, off);
  if (lane == 0) red[threadIdx.x >> 6] = lp;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, red[0] + red[1] + red[2] + red[3]);
}

// rho (when not null) is the mass per cell of the box; particles whose base
// cell holds more than rho_max keep their force unchanged.
__global__ void k_apply(const float* __restrict__ x, const float* __restrict__ yy, const float* __restrict__ z,
                        long long n, const float4* __restrict__ ymesh, int nm, float cs, float out_scale,
                        const float* __restrict__ rho, float rho_max, float* __restrict__ ax,
                        float* __restrict__ ay, float* __restrict__ az) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  Cic c;
  cic_nodes(x[p], yy[p], z[p], cs, nm, 0, c);
  if (rho && rho[c.node[0]] > rho_max) return;
  float px = 0.0f, py = 0.0f, pz = 0.0f;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const float4 v = ymesh[c.node[a]];
    px += c.w[a] * v.x;
    py += c.w[a] * v.y;
    pz += c.w[a] * v.z;
  }
  ax[p] += out_scale * px;
  ay[p] += out_scale * py;
  az[p] += out_scale * pz;
}

}  // namespace net
}  // namespace gng

using namespace gng::net;
using gng::fail;

struct gng_net {
  int nm = 0;
  int max_boxes = 0;
  int device = 0;
  long long M = 0;
  // weights
  float* d_params = nullptr;
  float* d_grad = nullptr;
  float* d_m1 = nullptr;
  float* d_m2 = nullptr;
  int adam_t = 0;
  u32x4* d_f1 = nullptr;   // forward fragments (8 bf16 per lane and k-step), layer 1..3
  u32x4* d_f2 = nullptr;
  u32x4* d_f3 = nullptr;
  u32x4* d_g3 = nullptr;   // data-gradient fragments of layers 3 and 2
  u32x4* d_g2 = nullptr;
  float* d_b16 = nullptr;   // 3 x 16 padded biases
  // activations
  unsigned short* d_x0 = nullptr;
  float* d_z1 = nullptr;
  unsigned short* d_x1 = nullptr;
  float* d_z2 = nullptr;
  unsigned short* d_x2 = nullptr;
  float* d_y = nullptr;
  float* d_gy = nullptr;
  unsigned short* d_dz3 = nullptr;
  unsigned short* d_dz2 = nullptr;
  unsigned short* d_dz1 = nullptr;
  float* d_partial = nullptr;
  double* d_loss = nullptr;
  long long loss_n = 0;
  float* d_tmp = nullptr;  // 2 x M floats for the input's box means
  unsigned long long* d_mom = nullptr;  // kMoments x M fixed-point particle moments (allocated on first use)
};

namespace {

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

void free_net(gng_net* t) {
  void* ptrs[] = {t->d_params, t->d_grad, t->d_m1, t->d_m2, t->d_f1, t->d_f2, t->d_f3, t->d_g3, t->d_g2,
                  t->d_b16, t->d_x0, t->d_z1, t->d_x1, t->d_z2, t->d_x2, t->d_y, t->d_gy, t->d_dz3,
                  t->d_dz2, t->d_dz1, t->d_partial, t->d_loss, t->d_tmp, t->d_mom};
  for (void* p : ptrs) (void)hipFree(p);
}

struct BufInfo {
  void* ptr;
  long long bytes;
};

BufInfo buffer(const gng_net* t, int which) {
  const long long cells = t->M * t->max_boxes;
  switch (which) {
    case GNG_NET_X0: return {t->d_x0, cells * kIn * 2};
    case GNG_NET_Z1: return {t->d_z1, cells * 16 * 4};
    case GNG_NET_X1: return {t->d_x1, cells * 16 * 2};
    case GNG_NET_Z2: return {t->d_z2, cells * 16 * 4};
    case GNG_NET_X2: return {t->d_x2, cells * 16 * 2};
    case GNG_NET_Y: return {t->d_y, cells * 4 * 4};
    case GNG_NET_GY: return {t->d_gy, cells * 4 * 4};
    case GNG_NET_DZ3: return {t->d_dz3, cells * 4 * 2};
    case GNG_NET_DZ2: return {t->d_dz2, cells * 16 * 2};
    case GNG_NET_DZ1: return {t->d_dz1, cells * 16 * 2};
    case GNG_NET_GRAD: return {t->d_grad, (long long)P_COUNT * 4};
    case GNG_NET_PARAMS: return {t->d_params, (long long)P_COUNT * 4};
    default: return {nullptr, -1};
  }
}

int repack(gng_net* t, hipStream_t s) {
  const float* p = t->d_params;
  constexpr int K1 = KSteps<kIn>::v, K4 = KSteps<4>::v, K16 = KSteps<16>::v;
  hipLaunchKernelGGL(k_pack, dim3(K1), dim3(64), 0, s, p + P_W1, kIn, 16, kIn, K1, 0, t->d_f1);
  hipLaunchKernelGGL(k_pack, dim3(K16), dim3(64), 0, s, p + P_W2, 16, 16, 16, K16, 0, t->d_f2);
  hipLaunchKernelGGL(k_pack, dim3(K16), dim3(64), 0, s, p + P_W3, 16, 3, 16, K16, 0, t->d_f3);
  hipLaunchKernelGGL(k_pack, dim3(K4), dim3(64), 0, s, p + P_W3, 16, 3, 4, K4, 1, t->d_g3);
  hipLaunchKernelGGL(k_pack, dim3(K16), dim3(64), 0, s, p + P_W2, 16, 16, 16, K16, 1, t->d_g2);
  hipLaunchKernelGGL(k_bias16, dim3(1), dim3(64), 0, s, p + P_B1, 16, t->d_b16);
  hipLaunchKernelGGL(k_bias16, dim3(1), dim3(64), 0, s, p + P_B2, 16, t->d_b16 + 16);
  hipLaunchKernelGGL(k_bias16, dim3(1), dim3(64), 0, s, p + P_B3, 3, t->d_b16 + 32);
  GNG_HIP(hipGetLastError());
  return GNG_OK;
}

int check_boxes(const gng_net* t, int boxes) {
  if (boxes < 1 || boxes > t->max_boxes) return fail(GNG_ERR_INVALID, "box count outside [1, max_boxes]");
  return GNG_OK;
}

dim3 conv_grid(const gng_net* t, int boxes) {
  return dim3((unsigned)(t->nm / TX), (unsigned)((t->nm / TY) * (t->nm / TZ)), (unsigned)boxes);
}

}  // namespace

namespace {

// The box means of the potential and the input kernel, moments from one of
// the two sources (or none).
int net_input(gng_net* net, int32_t box, const float* d_rho, const float* d_phi, float rho_mean, float phi_scale,
              const unsigned long long* mom, const double* momd, double mom_unit, hipStream_t s) {
  const dim3 g((unsigned)((net->M + 255) / 256));
  const int nm = net->nm;
  float* t0 = net->d_tmp;
  float* t1 = net->d_tmp + net->M;
  hipLaunchKernelGGL(k_boxmean, g, dim3(256), 0, s, d_phi, t0, nm, 1LL);
  hipLaunchKernelGGL(k_boxmean, g, dim3(256), 0, s, t0, t1, nm, (long long)nm);
  hipLaunchKernelGGL(k_boxmean, g, dim3(256), 0, s, t1, t0, nm, (long long)nm * nm);
  hipLaunchKernelGGL(k_input, g, dim3(256), 0, s, d_rho, d_phi, t0, mom, momd, mom_unit, net->M, 1.0f / rho_mean,
                     1.0f / phi_scale, reinterpret_cast<u32x4*>(net->d_x0) + (long long)box * net->M);
  GNG_HIP(hipGetLastError());
  return GNG_OK;
}

}  // namespace

extern "C" {

int32_t gng_net_param_count(void) { return P_COUNT; }

int gng_net_create(int32_t grid_size, int32_t max_boxes, int32_t device, gng_net** out) {
  if (!out) return fail(GNG_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (grid_size < 16 || grid_size % 16 != 0 || grid_size > 1024)
    return fail(GNG_ERR_INVALID, "network mesh must be a multiple of 16 in [16, 1024]");
  if (max_boxes < 1) return fail(GNG_ERR_INVALID, "max_boxes must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(GNG_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= ndev) return fail(GNG_ERR_INVALID, "device index out of range");
  GNG_HIP(hipSetDevice(device));
  gng_net* t = new (std::nothrow) gng_net();
  if (!t) return fail(GNG_ERR_OOM, "host allocation failed");
  t->nm = grid_size;
  t->max_boxes = max_boxes;
  t->device = device;
  t->M = (long long)grid_size * grid_size * grid_size;
  const long long cells = t->M * max_boxes;
  hipError_t e = hipSuccess;
  auto A = [&](auto** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(p), bytes);
  };
  A(&t->d_params, P_COUNT * 4);
  A(&t->d_grad, P_COUNT * 4);
  A(&t->d_m1, P_COUNT * 4);
  A(&t->d_m2, P_COUNT * 4);
  A(&t->d_f1, KSteps<kIn>::v * 64 * 16);
  A(&t->d_f2, KSteps<16>::v * 64 * 16);
  A(&t->d_f3, KSteps<16>::v * 64 * 16);
  A(&t->d_g3, KSteps<4>::v * 64 * 16);
  A(&t->d_g2, KSteps<16>::v * 64 * 16);
  A(&t->d_b16, 48 * 4);
  A(&t->d_x0, cells * kIn * 2);
  A(&t->d_z1, cells * 64);
  A(&t->d_x1, cells * 32);
  A(&t->d_z2, cells * 64);
  A(&t->d_x2, cells * 32);
  A(&t->d_y, cells * 16);
  A(&t->d_gy, cells * 16);
  A(&t->d_dz3, cells * 8);
  A(&t->d_dz2, cells * 32);
  A(&t->d_dz1, cells * 32);
  A(&t->d_partial, (size_t)kWgradBlocks * 28 * 256 * 4);
  A(&t->d_loss, 8);
  A(&t->d_tmp, t->M * 2 * 4);
  if (e != hipSuccess) {
    free_net(t);
    delete t;
    return fail(e == hipErrorOutOfMemory ? GNG_ERR_OOM : GNG_ERR_HIP,
                std::string("network buffers: ") + hipGetErrorString(e));
  }
  std::vector<float> zero(P_COUNT, 0.0f);
  e = hipMemcpy(t->d_params, zero.data(), P_COUNT * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t->d_m1, zero.data(), P_COUNT * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t->d_m2, zero.data(), P_COUNT * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t->d_grad, zero.data(), P_COUNT * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    free_net(t);
    delete t;
    return fail(GNG_ERR_HIP, std::string("network init: ") + hipGetErrorString(e));
  }
  int rc = repack(t, nullptr);
  if (rc == GNG_OK && hipDeviceSynchronize() != hipSuccess) rc = fail(GNG_ERR_HIP, "network init sync failed");
  if (rc != GNG_OK) {
    free_net(t);
    delete t;
    return rc;
  }
  *out = t;
  return GNG_OK;
}

void gng_net_destroy(gng_net* net) {
  if (!net) return;
  (void)hipSetDevice(net->device);
  (void)hipDeviceSynchronize();
  free_net(net);
  delete net;
}

int gng_net_set_params(gng_net* net, const float* host_params) {
  if (!net || !host_params) return fail(GNG_ERR_INVALID, "null argument");
  GNG_HIP(hipSetDevice(net->device));
  GNG_HIP(hipMemcpy(net->d_params, host_params, P_COUNT * 4, hipMemcpyHostToDevice));
  std::vector<float> zero(P_COUNT, 0.0f);
  GNG_HIP(hipMemcpy(net->d_m1, zero.data(), P_COUNT * 4, hipMemcpyHostToDevice));
  GNG_HIP(hipMemcpy(net->d_m2, zero.data(), P_COUNT * 4, hipMemcpyHostToDevice));
  net->adam_t = 0;
  int rc = repack(net, nullptr);
  if (rc != GNG_OK) return rc;
  GNG_HIP(hipDeviceSynchronize());
  return GNG_OK;
}

int gng_net_get_params(gng_net* net, float* host_params) {
  if (!net || !host_params) return fail(GNG_ERR_INVALID, "null argument");
  GNG_HIP(hipSetDevice(net->device));
  GNG_HIP(hipDeviceSynchronize());
  GNG_HIP(hipMemcpy(host_params, net->d_params, P_COUNT * 4, hipMemcpyDeviceToHost));
  return GNG_OK;
}

int64_t gng_net_buffer_bytes(const gng_net* net, int32_t which) {
  if (!net) return -1;
  return buffer(net, which).bytes;
}

int gng_net_device_buffer(gng_net* net, int32_t which, void** d_ptr) {
  if (!net || !d_ptr) return fail(GNG_ERR_INVALID, "null argument");
  const BufInfo b = buffer(net, which);
  if (!b.ptr) return fail(GNG_ERR_INVALID, "bad buffer id");
  *d_ptr = b.ptr;
  return GNG_OK;
}

int gng_net_read(gng_net* net, int32_t which, void* host, int64_t bytes) {
  if (!net || !host) return fail(GNG_ERR_INVALID, "null argument");
  const BufInfo b = buffer(net, which);
  if (!b.ptr || bytes < 0 || bytes > b.bytes) return fail(GNG_ERR_INVALID, "bad buffer id or size");
  GNG_HIP(hipSetDevice(net->device));
  GNG_HIP(hipDeviceSynchronize());
  GNG_HIP(hipMemcpy(host, b.ptr, (size_t)bytes, hipMemcpyDeviceToHost));
  return GNG_OK;
}

int gng_net_write(gng_net* net, int32_t which, const void* host, int64_t bytes) {
  if (!net || !host) return fail(GNG_ERR_INVALID, "null argument");
  const BufInfo b = buffer(net, which);
  if (!b.ptr || bytes < 0 || bytes > b.bytes) return fail(GNG_ERR_INVALID, "bad buffer id or size");
  if (which == GNG_NET_PARAMS) return gng_net_set_params(net, static_cast<const float*>(host));
  GNG_HIP(hipSetDevice(net->device));
  GNG_HIP(hipDeviceSynchronize());
  GNG_HIP(hipMemcpy(b.ptr, host, (size_t)bytes, hipMemcpyHostToDevice));
  return GNG_OK;
}


int gng_net_input_moments(gng_net* net, int32_t box, const float* d_rho, const float* d_phi, float rho_mean,
                          float phi_scale, const double* d_mom, void* stream) {
  if (!net || !d_rho || !d_phi || !d_mom) return fail(GNG_ERR_INVALID, "null argument");
  if (box < 0 || box >= net->max_boxes) return fail(GNG_ERR_INVALID, "box slot out of range");
  if (!(rho_mean > 0.0f) || !(phi_scale > 0.0f)) return fail(GNG_ERR_INVALID, "scales must be positive");
  GNG_HIP(hipSetDevice(net->device));
  return net_input(net, box, d_rho, d_phi, rho_mean, phi_scale, nullptr, d_mom, 1.0 / (double)rho_mean,
                   as_stream(stream));
}

int gng_net_input_particles(gng_net* net, int32_t box, const float* d_rho, const float* d_phi, float rho_mean,
                            float phi_scale, const float* d_x, const float* d_y, const float* d_z, const float* d_m,
                            int64_t n, double box_size, void* stream) {
  if (!net || !d_rho || !d_phi) return fail(GNG_ERR_INVALID, "null argument");
  if (box < 0 || box >= net->max_boxes) return fail(GNG_ERR_INVALID, "box slot out of range");
  if (!(rho_mean > 0.0f) || !(phi_scale > 0.0f)) return fail(GNG_ERR_INVALID, "scales must be positive");
  const bool particles = n > 0;
  if (n < 0) return fail(GNG_ERR_INVALID, "negative particle count");
  if (particles && (!d_x || !d_y || !d_z || !d_m)) return fail(GNG_ERR_INVALID, "null particle array");
  if (particles && !(box_size > 0.0)) return fail(GNG_ERR_INVALID, "box_size must be positive");
  GNG_HIP(hipSetDevice(net->device));
  hipStream_t s = as_stream(stream);
  const int nm = net->nm;
  double mom_unit = 0.0;
  if (particles) {
    if (!net->d_mom) GNG_HIP(hipMalloc(&net->d_mom, sizeof(unsigned long long) * kMoments * (size_t)net->M));
    // every per-cell sum is at most the total mass rho_mean * M (non-negative
    // masses), kept below 2^61
    const double total = (double)rho_mean * (double)net->M;
    const int shift = 61 - (int)std::ceil(std::log2(total));
    const long long words = kMoments * net->M;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, net->d_mom, words);
    hipLaunchKernelGGL(k_moments, dim3((unsigned)((n + kMomThreads - 1) / kMomThreads)), dim3(kMomThreads), 0, s,
                       d_x, d_y, d_z, d_m,
                       (long long)n, (float)((double)nm / box_size), nm, std::ldexp(1.0, shift), net->d_mom);
    mom_unit = std::ldexp(1.0, -shift) / (double)rho_mean;
  }
  return net_input(net, box, d_rho, d_phi, rho_mean, phi_scale,
                   particles ? (const unsigned long long*)net->d_mom : nullptr, nullptr, mom_unit, s);
}

int gng_net_input(gng_net* net, int32_t box, const float* d_rho, const float* d_phi, float rho_mean,
                  float phi_scale, void* stream) {
  return gng_net_input_particles(net, box, d_rho, d_phi, rho_mean, phi_scale, nullptr, nullptr, nullptr, nullptr, 0,
                                 0.0, stream);
}

int gng_net_forward(gng_net* net, int32_t boxes, void* stream) {
  if (!net) return fail(GNG_ERR_INVALID, "null network");
  int rc = check_boxes(net, boxes);
  if (rc != GNG_OK) return rc;
  GNG_HIP(hipSetDevice(net->device));
  hipStream_t s = as_stream(stream);
  const dim3 g = conv_grid(net, boxes);
  hipLaunchKernelGGL((k_conv<kIn, MODE_HIDDEN>), g, dim3(kConvThreads), 0, s, net->d_x0, net->d_f1, net->d_b16, net->d_z1,
                     net->d_x1, (const float*)nullptr, net->nm);
  hipLaunchKernelGGL((k_conv<16, MODE_HIDDEN>), g, dim3(kConvThreads), 0, s, net->d_x1, net->d_f2, net->d_b16 + 16,
                     net->d_z2, net->d_x2, (const float*)nullptr, net->nm);
  hipLaunchKernelGGL((k_conv<16, MODE_OUT>), g, dim3(kConvThreads), 0, s, net->d_x2, net->d_f3, net->d_b16 + 32, net->d_y,
                     (unsigned short*)nullptr, (const float*)nullptr, net->nm);
  GNG_HIP(hipGetLastError());
  return GNG_OK;
}

int gng_net_infer(gng_net* net, int32_t boxes, void* stream) {
  if (!net) return fail(GNG_ERR_INVALID, "null network");
  int rc = check_boxes(net, boxes);
  if (rc != GNG_OK) return rc;
  GNG_HIP(hipSetDevice(net->device));
  hipStream_t s = as_stream(stream);
  const dim3 g = conv_grid(net, boxes);
  hipLaunchKernelGGL((k_conv<kIn, MODE_INFER>), g, dim3(kConvThreads), 0, s, net->d_x0, net->d_f1, net->d_b16, (float*)nullptr,
                     net->d_x1, (const float*)nullptr, net->nm);
  hipLaunchKernelGGL((k_conv<16, MODE_INFER>), g, dim3(kConvThreads), 0, s, net->d_x1, net->d_f2, net->d_b16 + 16,
                     (float*)nullptr, net->d_x2, (const float*)nullptr, net->nm);
  hipLaunchKernelGGL((k_conv<16, MODE_OUT>), g, dim3(kConvThreads), 0, s, net->d_x2, net->d_f3, net->d_b16 + 32, net->d_y,
                     (unsigned short*)nullptr, (const float*)nullptr, net->nm);
  GNG_HIP(hipGetLastError());
  return GNG_OK;
}

int gng_net_loss_grad(gng_net* net, const float* d_pos4, const float* d_tgt4, int64_t n, double box_size,
                      float out_scale, void* stream) {
  if (!net || !d_pos4 || !d_tgt4) return fail(GNG_ERR_INVALID, "null argument");
  if (n <= 0) return fail(GNG_ERR_INVALID, "no training particles");
  if (!(box_size > 0.0)) return fail(GNG_ERR_INVALID, "box_size must be positive");
  GNG_HIP(hipSetDevice(net->device));
  hipStream_t s = as_stream(stream);
  const long long cells = net->M * net->max_boxes;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((cells * 4 + 255) / 256)), dim3(256), 0, s, net->d_gy, cells * 4, 0.0f);
  hipLaunchKernelGGL(k_fill_d, dim3(1), dim3(1), 0, s, net->d_loss);
  const float cs = (float)((double)net->nm / box_size);
  const float gscale = (float)(2.0 * out_scale / (double)n);
  hipLaunchKernelGGL(k_loss, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<const float4*>(d_pos4), reinterpret_cast<const float4*>(d_tgt4), (long long)n,
                     reinterpret_cast<const float4*>(net->d_y), net->d_gy, net->d_loss, net->nm, cs, out_scale, gscale,
                     net->max_boxes);
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((cells * 4 + 255) / 256)), dim3(256), 0, s, net->d_gy, net->d_dz3,
                     cells * 4);
  GNG_HIP(hipGetLastError());
  net->loss_n = n;
  return GNG_OK;
}

int gng_net_backward(gng_net* net, int32_t boxes, void* stream) {
  if (!net) return fail(GNG_ERR_INVALID, "null network");
  int rc = check_boxes(net, boxes);
  if (rc != GNG_OK) return rc;
  GNG_HIP(hipSetDevice(net->device));
  hipStream_t s = as_stream(stream);
  const int nm = net->nm;
  const dim3 g = conv_grid(net, boxes);
  const long long wtiles = (long long)(nm / WX) * (nm / WY) * (nm / WZ) * boxes;
  const int wb = (int)(wtiles < kWgradBlocks ? wtiles : kWgradBlocks);
  float* gr = net->d_grad;
  // layer 3: dW3 from x2 and dz3, then dz2 = conv^T(dz3) * silu'(z2)
  hipLaunchKernelGGL((k_conv_wgrad<16, 4>), dim3(wb), dim3(256), 0, s, net->d_x2, net->d_dz3, net->d_partial, nm,
                     boxes);
  hipLaunchKernelGGL((k_wgrad_reduce<16, 3>), dim3((433 * 3 + 255) / 256), dim3(256), 0, s, net->d_partial, wb,
                     gr + P_W3, gr + P_B3);
  hipLaunchKernelGGL((k_conv<4, MODE_BWD>), g, dim3(kConvThreads), 0, s, net->d_dz3, net->d_g3, (const float*)nullptr,
                     (float*)nullptr, net->d_dz2, net->d_z2, nm);
  // layer 2
  hipLaunchKernelGGL((k_conv_wgrad<16, 16>), dim3(wb), dim3(256), 0, s, net->d_x1, net->d_dz2, net->d_partial, nm,
                     boxes);
  hipLaunchKernelGGL((k_wgrad_reduce<16, 16>), dim3((433 * 16 + 255) / 256), dim3(256), 0, s, net->d_partial, wb,
                     gr + P_W2, gr + P_B2);
  hipLaunchKernelGGL((k_conv<16, MODE_BWD>), g, dim3(kConvThreads), 0, s, net->d_dz2, net->d_g2, (const float*)nullptr,
                     (float*)nullptr, net->d_dz1, net->d_z1, nm);
  // layer 1
  hipLaunchKernelGGL((k_conv_wgrad<kIn, 16>), dim3(wb), dim3(256), 0, s, net->d_x0, net->d_dz1, net->d_partial, nm,
                     boxes);
  hipLaunchKernelGGL((k_wgrad_reduce<kIn, 16>), dim3(((27 * kIn + 1) * 16 + 255) / 256), dim3(256), 0, s,
                     net->d_partial, wb, gr + P_W1, gr + P_B1);
  GNG_HIP(hipGetLastError());
  return GNG_OK;
}

int gng_net_adam(gng_net* net, float lr, float beta1, float beta2, float eps, void* stream) {
  if (!net) return fail(GNG_ERR_INVALID, "null network");
  if (!(lr >= 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
    return fail(GNG_ERR_INVALID, "bad Adam hyper-parameters");
  GNG_HIP(hipSetDevice(net->device));
  hipStream_t s = as_stream(stream);
  net->adam_t += 1;
  const float c1 = (float)(1.0 / (1.0 - std::pow((double)beta1, net->adam_t)));
  const float c2 = (float)(1.0 / (1.0 - std::pow((double)beta2, net->adam_t)));
  hipLaunchKernelGGL(k_adam, dim3((P_COUNT + 255) / 256), dim3(256), 0, s, net->d_params, net->d_grad, net->d_m1,
                     net->d_m2, P_COUNT, lr, beta1, beta2, eps, c1, c2);
  GNG_HIP(hipGetLastError());
  return repack(net, s);
}

int gng_net_train_step(gng_net* net, int32_t boxes, const float* d_pos4, const float* d_tgt4, int64_t n,
                       double box_size, float out_scale, float lr, void* stream) {
  int rc = gng_net_forward(net, boxes, stream);
  if (rc == GNG_OK) rc = gng_net_loss_grad(net, d_pos4, d_tgt4, n, box_size, out_scale, stream);
  if (rc == GNG_OK) rc = gng_net_backward(net, boxes, stream);
  if (rc == GNG_OK) rc = gng_net_adam(net, lr, 0.9f, 0.999f, 1e-8f, stream);
  return rc;
}

int gng_net_loss(gng_net* net, double* loss_out) {
  if (!net || !loss_out) return fail(GNG_ERR_INVALID, "null argument");
  if (net->loss_n <= 0) return fail(GNG_ERR_INVALID, "no loss has been computed");
  GNG_HIP(hipSetDevice(net->device));
  GNG_HIP(hipDeviceSynchronize());
  double v = 0.0;
  GNG_HIP(hipMemcpy(&v, net->d_loss, 8, hipMemcpyDeviceToHost));
  *loss_out = v / (double)net->loss_n;
  return GNG_OK;
}

int gng_net_apply_field(gng_net* net, int32_t box, const float* d_x, const float* d_y, const float* d_z,
                        int64_t n, double box_size, float out_scale, const float* d_rho, double rho_max,
                        float* d_ax, float* d_ay, float* d_az, void* stream) {
  if (!net || !d_x || !d_y || !d_z || !d_ax || !d_ay || !d_az) return fail(GNG_ERR_INVALID, "null argument");
  if (box < 0 || box >= net->max_boxes) return fail(GNG_ERR_INVALID, "box slot out of range");
  if (n < 0) return fail(GNG_ERR_INVALID, "negative particle count");
  if (!(box_size > 0.0)) return fail(GNG_ERR_INVALID, "box_size must be positive");
  if (d_rho && !(rho_max >= 0.0)) return fail(GNG_ERR_INVALID, "rho_max must be non-negative");
  if (n == 0) return GNG_OK;
  GNG_HIP(hipSetDevice(net->device));
  const float cs = (float)((double)net->nm / box_size);
  hipLaunchKernelGGL(k_apply, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), d_x, d_y, d_z,
                     (long long)n, reinterpret_cast<const float4*>(net->d_y) + (long long)box * net->M, net->nm, cs,
                     out_scale, d_rho, (float)rho_max, d_ax, d_ay, d_az);
  GNG_HIP(hipGetLastError());
  return GNG_OK;
}

int gng_net_apply(gng_net* net, int32_t box, const float* d_x, const float* d_y, const float* d_z, int64_t n,
                  double box_size, float out_scale, float* d_ax, float* d_ay, float* d_az, void* stream) {
  return gng_net_apply_field(net, box, d_x, d_y, d_z, n, box_size, out_scale, nullptr, 0.0, d_ax, d_ay, d_az,
                             stream);
}

}  // extern "C"
