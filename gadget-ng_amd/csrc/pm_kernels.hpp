// pm_kernels.hpp -- the periodic PM force pipeline as HIP kernels for gfx950.
//
// Stages (reference semantics: crates/gadget-ng-pm/src/solver.rs:74-140):
//   deposit    CIC mass per cell in 64-bit fixed point, converted in FP64 and
//              stored as FP32 (deposit.hpp; cic.rs:37-93 semantics)
//   fft_x_fwd  real-to-complex x transform; output is the packed half spectrum
//   fft_y_fwd  complex y transform
//   fft_z_green  z transform, Green's function x filters x (optional) CIC
//              deconvolution, then the inverse z transform of two fields:
//              A = phi, B = -i n_z phi (fft_poisson.rs:509-549)
//   fft_y_inv  inverse y transform: A -> A (for F_x), -i n_y A -> P (F_y),
//              B -> B (F_z)
//   fft_x_inv  -i n_x on the F_x field, complex-to-real x transform, in place
//   interp     CIC gather of the three force meshes (cic.rs:100-170)
//
// Packed half spectrum.  The x transform of a real row of length N has N/2+1
// values; the DC and Nyquist ones are real.  They are stored together as one
// complex number in slot 0 (re = DC, im = Nyquist), so every row is N/2
// complex values and the spectrum occupies exactly the bytes of the real
// mesh.  After the y and z transforms, slot 0 holds Z = R0 + i RN where R0,
// RN are the (Hermitian) spectra of the kx = 0 and kx = N/2 planes; the
// Green's function, which differs between the two planes, is applied as
//   out = (G0 + GN)/2 * Z(k) + (G0 - GN)/2 * conj(Z(-k)).
// The gradient factors -i n_alpha are the same on both planes (zero for x).
//
// Nyquist rule.  The reference takes Re() of a full complex inverse, which
// equals the Hermitian projection of F_alpha-hat; that projection is zero on
// the n_alpha = N/2 plane and the identity elsewhere.  Here the derivative
// uses n_alpha,eff = 0 on that plane while k^2 keeps n = +N/2
// (fft_poisson.rs:724-730), so every self-conjugate mode is exactly zero and
// the complex-to-real transform reproduces Re(Z2Z).
//
// Normalisation.  With M-hat the unnormalised DFT of the mass per cell and
// n the integer wavenumber, F_alpha-hat / N^3 = -i n_alpha phi,
//   phi = -(2 G / L^2) W(k) / |n|^2 * M-hat,
// which equals the reference's rho = M / V_cell, phi = -4 pi G W rho-hat/k^2,
// F = -i k phi, F = Re(IFFT(F-hat)) / N^3.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_device.hpp"

namespace gng {

// Positions to cells: c = (x * inv_cell) rem_euclid nm with inv_cell = nm / L
// in FP64, as cic.rs:50-52 and :262 form it.
struct CellMap {
  double inv_cell;
  int nm;
};

// Particle positions: x[i * stride], y[i * stride], z[i * stride] (stride 1:
// SoA arrays; stride 3 with y = x + 1, z = x + 2: the reference's Vec3 array).
template <typename T>
struct PosSrc {
  const T* x;
  const T* y;
  const T* z;
  long long stride;
};

// Accelerations out, with the same addressing.
template <typename T>
struct AccDst {
  T* x;
  T* y;
  T* z;
  long long stride;
};

constexpr int kFracBits = 43;
constexpr unsigned long long kFracOne = 1ull << kFracBits;

// x rounded to the nearest integer, ties to even (the value __double2ll_rn
// gives), for |x| < 2^51: adding 1.5 * 2^52 leaves that integer in the low
// mantissa bits.  One FP64 add and one integer subtraction instead of the
// emulated FP64 -> int64 conversion.
__device__ __forceinline__ long long rint_small(double x) {
  return __double_as_longlong(x + 6755399441055744.0) - 0x4338000000000000LL;
}

// Cell index and in-cell fraction of one coordinate, in FP64:
// c = (x * inv_cell) rem_euclid nm, i = floor(c) (cic.rs:50-56), and the
// fraction c - i (exact in FP64) rounded to kFracBits bits.  A fraction that
// rounds up to 1 moves to the next cell; index nm (rem_euclid rounding up
// for tiny negative x, or that carry) wraps to 0.  A non-finite coordinate
// lands in cell 0, so no index leaves the mesh.
__device__ __forceinline__ void cell_fix(double x, const CellMap cm, int& i, unsigned long long& frac) {
  const double nmf = (double)cm.nm;
  double c = x * cm.inv_cell;
  if (__builtin_expect(!(c >= 0.0 && c < nmf), 0)) {
    if (!isfinite(c)) {
      i = 0;
      frac = 0ull;
      return;
    }
    c = fmod(c, nmf);
    if (c < 0.0) c += nmf;
  }
  const double f = floor(c);
  int ii = (int)f;
  unsigned long long q = (unsigned long long)rint_small((c - f) * (double)kFracOne);
  if (q >= kFracOne) {
    q = 0ull;
    ++ii;
  }
  if ((unsigned)ii >= (unsigned)cm.nm) ii = 0;
  i = ii;
  frac = q;
}

// The same coordinate for the interpolation: index and FP32 fraction.
__device__ __forceinline__ void cell_frac32(double x, const CellMap cm, int& i, float& d) {
  unsigned long long q;
  cell_fix(x, cm, i, q);
  d = (float)q * (1.0f / (float)kFracOne);
}

// A particle in bucket order: 24 bytes.  Word a (axis a) holds the cell index
// (10 bits, nm <= 1024) at bit 43, the fraction in bits 0..42, and 11 bits of
// the FP32 mass's bit pattern at bit 53 (bits 0-10, 11-21, 22-31).
struct alignas(8) PartRec {
  unsigned long long a[3];
};

__device__ __forceinline__ PartRec encode_rec(const int (&c)[3], const unsigned long long (&f)[3], float m) {
  const unsigned mb = __float_as_uint(m);
  const unsigned mp[3] = {mb & 0x7FFu, (mb >> 11) & 0x7FFu, mb >> 22};
  PartRec r;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    r.a[k] = ((unsigned long long)mp[k] << 53) | ((unsigned long long)c[k] << kFracBits) | f[k];
  return r;
}
__device__ __forceinline__ int rec_cell(const PartRec& r, int k) { return (int)((r.a[k] >> kFracBits) & 0x3FFu); }
__device__ __forceinline__ double rec_frac(const PartRec& r, int k) {
  return (double)(r.a[k] & (kFracOne - 1ull)) * (1.0 / (double)kFracOne);
}
__device__ __forceinline__ float rec_mass(const PartRec& r) {
  return __uint_as_float((unsigned)(r.a[0] >> 53) | ((unsigned)(r.a[1] >> 53) << 11) |
                         ((unsigned)(r.a[2] >> 53) << 22));
}

// Complex values between consecutive z planes of the half-spectrum meshes:
// N*(N/2) plus a pad of 32 (512 <= N < 1024) or 64 (N >= 1024) values.  The
// z pass reads one 64- or 128-byte segment from each of N planes; at a
// power-of-two plane stride (1 MiB at N = 512, 4 MiB at N = 1024) these reads
// crowd onto few HBM channels.  On MI355X the pad raised a plain column copy
// with that pattern from 4.5 to 5.0 TB/s (N = 512, 128-byte segments) and
// from 2.3 to 3.3 TB/s (N = 1024, 64-byte segments).  A real field kept in the
// same buffer has the same bytes row by row: row (z, y) of N floats starts at
// float 2*z*plane + y*N, so the x transforms still work in place.
__host__ __device__ constexpr long long mesh_plane(int n) {
  return (long long)n * (n / 2) + (n >= 1024 ? 64 : (n >= 512 ? 32 : 0));
}

// Complex offset of row r = z*N + y of a half-spectrum mesh.
template <int N>
__device__ __forceinline__ long long mesh_row(long long r) {
  return (r / N) * mesh_plane(N) + (r % N) * (N / 2);
}

// Per-solve scalars kept on the device (no host round trip).
// The sum of round(m 2^s) over the particles is kept in kQSlots partial
// sums, one 128-byte line apart: blocks add to slot blockIdx % kQSlots, so
// tens of thousands of blocks do not serialise on one address.
constexpr int kQSlots = 64;
constexpr int kQStride = 16;  // 64-bit words between slots
struct SolveScalars {
  unsigned int mmax_bits;  // max |m| as float bits (non-negative floats order as integers)
  unsigned int split_budget;  // scratch tiles reserved by split deposit tiles
  unsigned long long qtot;    // sum of the qsum slots, once the scatter is done (k_tile_ranges)
  unsigned int pad[60];
  unsigned long long qsum[kQSlots * kQStride];
};

// Fixed-point exponent: contributions are w*m*2^s with N*max|m|*2^s < 2^62,
// so no cell sum can overflow int64 whatever the clustering, and with
// max|m|*2^s < 2^51 (binding only below about 2^11 particles), so every
// single contribution (w <= 1) converts with rint_small.
__device__ __forceinline__ int fixed_shift(const SolveScalars* sc, long long ntot) {
  const float mmax = __uint_as_float(sc->mmax_bits);
  int e = 0, em = 0;
  (void)frexp((double)ntot * (double)mmax, &e);
  (void)frexp((double)mmax, &em);
  return min(62 - e, 51 - em);
}

struct GreenParams {
  float c;        // 2 G / L^2
  float a;        // (2 pi / L)^2 (r_s^2 / 2 + eps^2): W = exp(-a |n|^2)
  double c64, a64;
  const float* deconv;  // 1/sinc^4(pi n / N) per axis index, or null
};

// One box of an ensemble: the `_box` form of each stage kernel takes a device
// array of these and works on box blockIdx.z, so one launch runs a stage for
// every box of one mesh side.  Meshes, sort buffers and scalars are the box's
// own plan's; the particles are [p0, p0 + n) of the ensemble's arrays.
struct BoxArgs {
  float2* P;
  float2* A;
  float2* B;
  GreenParams gp;
  CellMap cm;
  unsigned* count;
  unsigned* offset;
  unsigned* partial;
  SolveScalars* sc;
  PartRec* sorted;
  unsigned* bucket;
  unsigned* rank;
  long long n;
  long long p0;
};

__device__ __forceinline__ void reset_scalars_body(SolveScalars* sc) {
  if (threadIdx.x == 0) {
    sc->mmax_bits = 0u;
    sc->split_budget = 0u;
  }
  if (threadIdx.x < kQSlots) sc->qsum[threadIdx.x * kQStride] = 0ull;
}

__global__ void k_reset_scalars(SolveScalars* sc) {
  reset_scalars_body(sc);
}

__global__ void k_reset_scalars_box(const BoxArgs* __restrict__ bx) {
  const BoxArgs& b = bx[blockIdx.z];
  reset_scalars_body(b.sc);
}

// ---------------------------------------------------------------------------
// x transforms (rows are contiguous; a workgroup owns ROWS whole rows)
// ---------------------------------------------------------------------------

template <int N>
struct XCfg {
  static constexpr int H = N / 2;
  static constexpr int T = 256;
  static constexpr int ROWS_RAW = (2048 / H) > 0 ? 2048 / H : 1;
  static constexpr int LS = H + 1;  // odd stride against LDS bank conflicts
};

// R2C along x of the FP32 mass-per-cell rows.  rho and P may be the same
// buffer (a row of N reals has the bytes of N/2 complex values and every
// workgroup reads its rows completely before writing them).  The rows are
// read straight into the registers of the first FFT stage; the pairwise
// post-processing (X_k needs Z_k and Z_{H-k}) exchanges through LDS.
// rho_plane: complex-sized (float-pair) stride between z planes of rho, which
// is mesh_plane(N) when rho shares P's buffer and N*N/2 for a separate one.
template <int N, int ROWS>
__device__ __forceinline__ void fft_x_fwd_body(const float* rho, float2* P, const float2* __restrict__ tw_g,
                                                   long long rho_plane) {
  constexpr int H = N / 2, T = 256, LS = XCfg<N>::LS;
  using F = RegFft<H, false, false, ROWS, T, LS, 1, false, N>;
  __shared__ float2 buf[ROWS * LS];
  const int tid = threadIdx.x;
  // twiddles in LDS: the FFT stages then read them with LDS instructions, so
  // no global-load wait of a stage also waits for unrelated outstanding loads
  __shared__ float2 twl[N];
  for (int e = tid; e < N; e += (int)blockDim.x) twl[e] = tw_g[e];
  const float2* tw = twl;
  const long long row0 = (long long)blockIdx.x * ROWS;
  const float2* src = reinterpret_cast<const float2*>(rho);
  float2 vin[F::PER0][F::R0], vout[F::PERL][F::RL];
  for_edge<H, F::R0, ROWS, T, false>(tid, [&](int p, int r, int l, int i, bool ok) {
    const long long row = row0 + l;
    if (ok) vin[p][r] = src[(row / N) * rho_plane + (row % N) * H + i];
  });
  F::run(vin, vout, buf, tw, tid);
  __syncthreads();
  for_edge<H, F::RL, ROWS, T, false>(tid, [&](int p, int r, int l, int i, bool ok) {
    if (ok) buf[l * LS + i] = vout[p][r];
  });
  __syncthreads();
  constexpr int TOT = ROWS * H;
#pragma unroll
  for (int p = 0; p < (TOT + T - 1) / T; ++p) {
    const int e = tid + p * T;
    if ((TOT % T == 0) || e < TOT) {
      const int r = e / H, k = e % H;
      const float2 zk = buf[r * LS + k];
      float2 out;
      if (k == 0) {
        out = make_float2(zk.x + zk.y, zk.x - zk.y);
      } else {
        const float2 zc = cconj(buf[r * LS + (H - k)]);
        const float2 E = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
        // O = (zk - zc) / (2i)
        const float2 O = make_float2(0.5f * (zk.y - zc.y), -0.5f * (zk.x - zc.x));
        out = cadd(E, cmul(tw[k], O));
      }
      P[mesh_row<N>(row0 + r) + k] = out;
    }
  }
}

template <int N, int ROWS>
__global__ __launch_bounds__(256) void k_fft_x_fwd(const float* rho, float2* P, const float2* __restrict__ tw_g,
                                                   long long rho_plane) {
  fft_x_fwd_body<N, ROWS>(rho, P, tw_g, rho_plane);
}

template <int N, int ROWS>
__global__ __launch_bounds__(256) void k_fft_x_fwd_box(const BoxArgs* __restrict__ bx, const float2* __restrict__ tw_g) {
  const BoxArgs& b = bx[blockIdx.z];
  fft_x_fwd_body<N, ROWS>(reinterpret_cast<const float*>(b.P), b.P, tw_g, mesh_plane(N));
}

// Mode multipliers of the initial-condition passes (codes): 0 -> 1,
// 1 -> -i n, 2 -> n, 3 -> n^2.
__device__ __forceinline__ float2 ic_mul(int code, float n, float2 v) {
  switch (code) {
    case 1: return make_float2(n * v.y, -n * v.x);
    case 2: return make_float2(n * v.x, n * v.y);
    case 3: return make_float2(n * n * v.x, n * n * v.y);
    default: return v;
  }
}

// C2R along x, in place, for field blockIdx.y (0: A with -i n_x, 1: P, 2: B).
// With MUL the three fields take the multipliers codes.x, .y, .z instead (the
// initial-condition passes); the force solve instantiates MUL = false.
// The pairwise pre-processing reads the row from LDS and feeds the first FFT
// stage in registers; the last stage stores the real row directly.
template <int N, int ROWS, bool MUL = false>
__device__ __forceinline__ void fft_x_inv_body(float2* __restrict__ A, float2* __restrict__ P,
                                                   float2* __restrict__ B, float2* __restrict__ pot,
                                                   const float2* __restrict__ tw_g, int field0, float pot_scale,
                                                   int3 codes) {
  constexpr int H = N / 2, T = 256, LS = XCfg<N>::LS;
  using F = RegFft<H, true, false, ROWS, T, LS, 1, false, N>;
  __shared__ float2 buf[ROWS * LS];
  const int tid = threadIdx.x;
  // twiddles in LDS: the FFT stages then read them with LDS instructions, so
  // no global-load wait of a stage also waits for unrelated outstanding loads
  __shared__ float2 twl[N];
  for (int e = tid; e < N; e += (int)blockDim.x) twl[e] = tw_g[e];
  const float2* tw = twl;
  // fields 0..2: F_x from A (times -i n_x), F_y from P, F_z from B, in place;
  // field 3: the potential from A (no derivative) into pot, scaled
  const int field = blockIdx.y + field0;
  float2* data = field == 0 ? A : (field == 1 ? P : (field == 2 ? B : A));
  const long long row0 = (long long)blockIdx.x * ROWS;
  constexpr int TOT = ROWS * H;
#pragma unroll
  for (int p = 0; p < (TOT + T - 1) / T; ++p) {
    const int e = tid + p * T;
    if ((TOT % T == 0) || e < TOT) {
      float2 v = data[mesh_row<N>(row0 + e / H) + e % H];
      if constexpr (MUL) {
        // slot 0 (DC and the packed Nyquist plane) has n = 0
        v = ic_mul(field == 0 ? codes.x : (field == 1 ? codes.y : codes.z), (float)(e % H), v);
      } else if (field == 0) {
        // -i n_x; slot 0 holds DC (n=0) and Nyquist (n_eff=0): both vanish
        const float nx = (float)(e % H);
        v = make_float2(nx * v.y, -nx * v.x);
      }
      buf[(e / H) * LS + (e % H)] = v;
    }
  }
  __syncthreads();
  float2 vin[F::PER0][F::R0], vout[F::PERL][F::RL];
  for_edge<H, F::R0, ROWS, T, false>(tid, [&](int p, int r, int l, int k, bool ok) {
    if (!ok) return;
    const float2 y = buf[l * LS + k];
    if (k == 0) {
      vin[p][r] = make_float2(y.x + y.y, y.x - y.y);
    } else {
      const float2 xc = cconj(buf[l * LS + (H - k)]);
      const float2 sm = cadd(y, xc);
      // Z = (X_k + conj X_{H-k}) + i (X_k - conj X_{H-k}) e^{+2 pi i k/N}
      const float2 t = cmul(csub(y, xc), cconj(tw[k]));
      vin[p][r] = make_float2(sm.x - t.y, sm.y + t.x);
    }
  });
  F::run(vin, vout, buf, tw, tid);
  const float sc = field == 3 ? pot_scale : 1.0f;
  // the potential goes to its own unpadded mesh
  for_edge<H, F::RL, ROWS, T, false>(tid, [&](int p, int r, int l, int i, bool ok) {
    float2* orow = field == 3 ? pot + (row0 + l) * H : data + mesh_row<N>(row0 + l);
    if (ok) orow[i] = make_float2(vout[p][r].x * sc, vout[p][r].y * sc);
  });
}

template <int N, int ROWS, bool MUL = false>
__global__ __launch_bounds__(256) void k_fft_x_inv(float2* __restrict__ A, float2* __restrict__ P,
                                                   float2* __restrict__ B, float2* __restrict__ pot,
                                                   const float2* __restrict__ tw_g, int field0, float pot_scale,
                                                   int3 codes) {
  fft_x_inv_body<N, ROWS, MUL>(A, P, B, pot, tw_g, field0, pot_scale, codes);
}

template <int N, int ROWS>
__global__ __launch_bounds__(256) void k_fft_x_inv_box(const BoxArgs* __restrict__ bx, const float2* __restrict__ tw_g) {
  const BoxArgs& b = bx[blockIdx.z];
  fft_x_inv_body<N, ROWS, false>(b.A, b.P, b.B, nullptr, tw_g, 0, 1.0f, make_int3(0, 0, 0));
}

// ---------------------------------------------------------------------------
// y / z transforms: a workgroup owns C adjacent kx columns of one plane
// ---------------------------------------------------------------------------

// Column tile width: the largest divisor of N/2 that is at most 16, so the
// tiles cover every kx column exactly and each strided row segment is a whole
// 128-byte cache line (narrower tiles leave half of every line to another
// workgroup, which fetches it from HBM again).
__host__ __device__ constexpr int yz_cols(int h) {
  int c = h < 16 ? h : 16;
  while (h % c != 0) --c;
  return c;
}

template <int N>
struct YZCfg {
  static constexpr int H = N / 2;
  // 16 columns, or 8 at N = 1024 so the thread rule below stays within 1024
  static constexpr int C = N >= 1024 ? yz_cols(H < 8 ? H : 8) : yz_cols(H);
  // one radix-8 butterfly (8 complex values) per thread and register array:
  // C * N / 8 threads, rounded up to whole wavefronts, at most 1024.  More
  // values per thread push the register count past 128 and halve occupancy.
  static constexpr int T_RAW = C * N / 8;
  static constexpr int T = T_RAW >= 1024 ? 1024 : (T_RAW <= 64 ? 64 : (T_RAW + 63) / 64 * 64);
  static_assert(H % C == 0, "column tiles must cover the half spectrum");
  // the per-item z pass at N = 1024: full 128-byte tiles (16 values per thread)
  static constexpr int CZ = N >= 1024 ? yz_cols(H) : C;
  static constexpr int TZ_RAW = CZ * N / 8;
  static constexpr int TZ = TZ_RAW >= 1024 ? 1024 : (TZ_RAW <= 64 ? 64 : (TZ_RAW + 63) / 64 * 64);
  // the inverse y pass at N = 1024 as two launches (fields from A, then B) on
  // full 128-byte tiles, 16 values per thread
  static constexpr bool YSPLIT = N >= 1024;
  // The combined inverse y pass keeps three fields live (about 112 VGPRs at 8
  // values per thread), so a 1024-thread workgroup has a CU to itself and
  // nothing loads while it transforms.  It runs instead on tiles of 8 columns
  // and 512 threads, two workgroups to a CU (their line halves meet in one L2
  // through line_tile): 1.119 -> 1.094 ms for the pair at N = 512.
  static constexpr int CI = (!YSPLIT && T == 1024 && C > 8) ? 8 : C;
  static constexpr int TI = CI == C ? T : CI * N / 8;
};

// Column tiles narrower than a 128-byte line (C = 8 at nm = 1024) share each
// line with the neighbouring tile.  Workgroups are dealt to the 8 XCDs round
// robin, and each XCD has its own L2, so tile pairs (2k, 2k+1) are given to
// workgroup indices 8 apart: both halves of a line are then fetched through
// one L2 at about the same time instead of once through each of two.
template <int C, int TILES>
__device__ __forceinline__ int line_tile(int b) {
  if constexpr (C * 8 >= 128 || TILES % 16 != 0) {
    return b;
  } else {
    const int q = b >> 4, r = b & 15;
    return (q << 4) + ((r & 7) << 1) + (r >> 3);
  }
}

// Load / store a C-column tile of a plane-strided field into LDS columns
// [0, C) (element (i, c) at buf[i*IS + c]); `stride` is the distance in
// complex values between consecutive i.
template <int N, int C, int T>
__device__ __forceinline__ void fft_y_fwd_body(float2* __restrict__ P, const float2* __restrict__ tw_g) {
  constexpr int H = N / 2, IS = C + 1;
  using F = RegFft<N, false, false, C, T, 1, IS, true, N>;
  __shared__ float2 buf[N * IS];
  const int tid = threadIdx.x;
  // twiddles in LDS: the FFT stages then read them with LDS instructions, so
  // no global-load wait of a stage also waits for unrelated outstanding loads
  __shared__ float2 twl[N];
  for (int e = tid; e < N; e += (int)blockDim.x) twl[e] = tw_g[e];
  const float2* tw = twl;
  float2* base = P + (size_t)blockIdx.y * mesh_plane(N) + (size_t)line_tile<C, H / C>(blockIdx.x) * C;
  float2 vin[F::PER0][F::R0], vout[F::PERL][F::RL];
  for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
    if (ok) vin[p][r] = base[(unsigned)i * H + l];
  });
  F::run(vin, vout, buf, tw, tid);
  for_edge<N, F::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
    if (ok) base[(unsigned)i * H + l] = vout[p][r];
  });
}

template <int N, int C, int T>
__global__ __launch_bounds__(T) void k_fft_y_fwd(float2* __restrict__ P, const float2* __restrict__ tw_g) {
  fft_y_fwd_body<N, C, T>(P, tw_g);
}

template <int N, int C, int T>
__global__ __launch_bounds__(T) void k_fft_y_fwd_box(const BoxArgs* __restrict__ bx, const float2* __restrict__ tw_g) {
  const BoxArgs& b = bx[blockIdx.z];
  fft_y_fwd_body<N, C, T>(b.P, tw_g);
}


__device__ __forceinline__ int freq_of(int j, int n) { return j <= n / 2 ? j : j - n; }

// z transform, Green's function, then the inverse z transforms of A = phi and
// B = -i n_z phi, for kx >= 1.  The forward transform leaves rho-hat in
// registers, the Green's function is applied there, and the inverse runs on
// the reversed radix plan so it starts from those registers.  The packed
// kx = 0 column is done by k_fft_z_green_col0.
template <int N, int C, int T>
__device__ __forceinline__ void fft_z_green_body(const float2* __restrict__ P, float2* __restrict__ A,
                                                   float2* __restrict__ B, const float2* __restrict__ tw_g,
                                                   const GreenParams gp) {
  constexpr int H = N / 2, IS = C + 1;
  using FF = RegFft<N, false, false, C, T, 1, IS, true, N>;
  using FI = RegFft<N, true, true, C, T, 1, IS, true, N>;
  static_assert(FI::R0 == FF::RL && FI::PER0 == FF::PERL, "inverse must start in the forward output layout");
  __shared__ float2 buf[N * IS];
  const int tid = threadIdx.x;
  // twiddles in LDS: the FFT stages then read them with LDS instructions, so
  // no global-load wait of a stage also waits for unrelated outstanding loads
  __shared__ float2 twl[N];
  for (int e = tid; e < N; e += (int)blockDim.x) twl[e] = tw_g[e];
  const float2* tw = twl;
  const int tile = line_tile<C, H / C>(blockIdx.x), ky = blockIdx.y;
  const unsigned plane = (unsigned)mesh_plane(N);
  const size_t col0 = (size_t)ky * H + (size_t)tile * C;
  float2 vin[FF::PER0][FF::R0], ph[FF::PERL][FF::RL];
  for_edge<N, FF::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
    if (ok) vin[p][r] = P[col0 + (unsigned)i * plane + l];
  });
  FF::run(vin, ph, buf, tw, tid);
  const int ny = freq_of(ky, N);
  const float dy = gp.deconv ? gp.deconv[ky] : 1.0f;
  for_edge<N, FF::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
    if (!ok) return;
    const int kx = tile * C + l;
    const int nz = freq_of(i, N);
    const int n2 = kx * kx + ny * ny + nz * nz;
    const float fn2 = (float)(n2 > 0 ? n2 : 1);
    float g = kx == 0 ? 0.0f : -gp.c * expf(-gp.a * fn2) / fn2;
    if (gp.deconv) g *= gp.deconv[kx] * dy * gp.deconv[i];
    ph[p][r] = make_float2(g * ph[p][r].x, g * ph[p][r].y);
  });
  float2 w[FI::PER0][FI::R0];
#pragma unroll
  for (int p = 0; p < FI::PER0; ++p)
#pragma unroll
    for (int r = 0; r < FI::R0; ++r) w[p][r] = ph[p][r];
  {
    float2 out[FI::PERL][FI::RL];
    FI::run(w, out, buf, tw, tid);
    for_edge<N, FI::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      if (ok && (tile | l)) A[col0 + (unsigned)i * plane + l] = out[p][r];
    });
  }
  for_edge<N, FI::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
    const float nze = (i == N / 2) ? 0.0f : (float)freq_of(i, N);
    w[p][r] = make_float2(nze * ph[p][r].y, -nze * ph[p][r].x);
  });
  {
    float2 out[FI::PERL][FI::RL];
    FI::run(w, out, buf, tw, tid);
    for_edge<N, FI::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      if (ok && (tile | l)) B[col0 + (unsigned)i * plane + l] = out[p][r];
    });
  }
}

template <int N, int C, int T>
__global__ __launch_bounds__(T) void k_fft_z_green(const float2* __restrict__ P, float2* __restrict__ A,
                                                   float2* __restrict__ B, const float2* __restrict__ tw_g,
                                                   const GreenParams gp) {
  fft_z_green_body<N, C, T>(P, A, B, tw_g, gp);
}

template <int N, int C, int T>
__global__ __launch_bounds__(T) void k_fft_z_green_box(const BoxArgs* __restrict__ bx, const float2* __restrict__ tw_g) {
  const BoxArgs& b = bx[blockIdx.z];
  fft_z_green_body<N, C, T>(b.P, b.A, b.B, tw_g, b.gp);
}

// Persistent form of k_fft_z_green for large meshes: gridDim.x workgroups (one
// per CU) walk the (column tile, ky) items in a grid-stride loop.  Each
// workgroup loads the next item's column values into registers right after
// the forward transform of the current item has consumed its own, so that
// load is in flight while the Green's function and both inverse transforms
// run.  One workgroup fills a CU's registers, so without this the CU idles
// through every item's load latency.  To make room for the prefetched values,
// phi waits in LDS (each thread's own slots) while field A is transformed.
template <int N, int C, int T>
__global__ __launch_bounds__(T) void k_fft_z_green_persist(const float2* __restrict__ P, float2* __restrict__ A,
                                                           float2* __restrict__ B, const float2* __restrict__ tw_g,
                                                           const GreenParams gp) {
  constexpr int H = N / 2, IS = C + 1, TILES = H / C;
  using FF = RegFft<N, false, false, C, T, 1, IS, true, N>;
  using FI = RegFft<N, true, true, C, T, 1, IS, true, N>;
  static_assert(FI::R0 == FF::RL && FI::PER0 == FF::PERL, "inverse must start in the forward output layout");
  __shared__ float2 buf[N * IS];
  constexpr int KEEP = FF::PERL * FF::RL;
  __shared__ float2 keep[KEEP * T];
  const int tid = threadIdx.x;
  // twiddles in LDS: the FFT stages then read them with LDS instructions, so
  // no global-load wait of a stage also waits for unrelated outstanding loads
  __shared__ float2 twl[N];
  for (int e = tid; e < N; e += (int)blockDim.x) twl[e] = tw_g[e];
  const float2* tw = twl;
  const unsigned plane = (unsigned)mesh_plane(N);
  const int items = TILES * N;
  float2 vin[FF::PER0][FF::R0];
  int item = blockIdx.x;
  if (item < items) {
    const size_t c0 = (size_t)(item / TILES) * H + (size_t)line_tile<C, TILES>(item % TILES) * C;
    for_edge<N, FF::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      if (ok) vin[p][r] = P[c0 + (unsigned)i * plane + l];
    });
  }
  for (; item < items; item += gridDim.x) {
    const int tile = line_tile<C, TILES>(item % TILES), ky = item / TILES;
    const size_t col0 = (size_t)ky * H + (size_t)tile * C;
    const int ny = freq_of(ky, N);
    const float dy = gp.deconv ? gp.deconv[ky] : 1.0f;
    float2 ph[FF::PERL][FF::RL];
    FF::run(vin, ph, buf, tw, tid);
    const int next = item + (int)gridDim.x;
    if (next < items) {
      const size_t c1 = (size_t)(next / TILES) * H + (size_t)line_tile<C, TILES>(next % TILES) * C;
      for_edge<N, FF::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
        if (ok) vin[p][r] = P[c1 + (unsigned)i * plane + l];
      });
    }
    for_edge<N, FF::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      if (!ok) return;
      const int kx = tile * C + l;
      const int nz = freq_of(i, N);
      const int n2 = kx * kx + ny * ny + nz * nz;
      const float fn2 = (float)(n2 > 0 ? n2 : 1);
      float g = kx == 0 ? 0.0f : -gp.c * expf(-gp.a * fn2) / fn2;
      if (gp.deconv) g *= gp.deconv[kx] * dy * gp.deconv[i];
      ph[p][r] = make_float2(g * ph[p][r].x, g * ph[p][r].y);
    });
#pragma unroll
    for (int p = 0; p < FF::PERL; ++p)
#pragma unroll
      for (int r = 0; r < FF::RL; ++r) keep[(p * FF::RL + r) * T + tid] = ph[p][r];
    {
      float2 out[FI::PERL][FI::RL];
      FI::run(ph, out, buf, tw, tid);
      for_edge<N, FI::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
        if (ok && (tile | l)) A[col0 + (unsigned)i * plane + l] = out[p][r];
      });
    }
    float2 w[FI::PER0][FI::R0];
    for_edge<N, FI::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      const float nze = (i == N / 2) ? 0.0f : (float)freq_of(i, N);
      const float2 v = keep[(p * FF::RL + r) * T + tid];
      w[p][r] = make_float2(nze * v.y, -nze * v.x);
    });
    {
      float2 out[FI::PERL][FI::RL];
      FI::run(w, out, buf, tw, tid);
      for_edge<N, FI::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
        if (ok && (tile | l)) B[col0 + (unsigned)i * plane + l] = out[p][r];
      });
    }
  }
}

// The packed kx = 0 column: slot 0 holds Z = R0 + i RN, the spectra of the
// kx = 0 and kx = N/2 planes, whose Green's functions differ, so
//   phi(ky, kz) = (G0 + GN)/2 Z(ky, kz) + (G0 - GN)/2 conj(Z(-ky, -kz)).
// One workgroup per pair {ky, N - ky} transforms both z lines, combines them
// with FP64 Green's factors and writes A and B of both.
template <int N>
__device__ __forceinline__ void fft_z_green_col0_body(const float2* __restrict__ P, float2* __restrict__ A,
                                                          float2* __restrict__ B, const float2* __restrict__ tw,
                                                          const GreenParams gp) {
  constexpr int H = N / 2, T = 256, LS = N + 1;
  __shared__ float2 buf[4 * LS];
  const int tid = threadIdx.x;
  const int ky = blockIdx.x, kym = (N - ky) % N;
  const size_t plane = (size_t)mesh_plane(N);
  for (int e = tid; e < 2 * N; e += T) {
    const int l = e / N, z = e % N;
    buf[l * LS + z] = P[(size_t)z * plane + (size_t)(l ? kym : ky) * H];
  }
  __syncthreads();
  fft_lds<N, false, 2, T, LS, 1, false, N>(buf, tw, tid);
  constexpr int PER = (2 * N + T - 1) / T;
  float2 ph[PER];
  const int ny = freq_of(ky, N);  // |ny| is the same for both lines
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int e = tid + p * T;
    if (e < 2 * N) {
      const int l = e / N, i = e % N;
      const int nz = freq_of(i, N);
      const double n2a = (double)(ny * ny + nz * nz);
      const double n2b = n2a + (double)(N / 2) * (double)(N / 2);
      double dd = 1.0, dn = 1.0;
      if (gp.deconv) {
        const double dyz = (double)gp.deconv[l ? kym : ky] * (double)gp.deconv[i];
        dd = (double)gp.deconv[0] * dyz;
        dn = (double)gp.deconv[N / 2] * dyz;
      }
      const double g0 = n2a == 0.0 ? 0.0 : -gp.c64 * exp(-gp.a64 * n2a) / n2a * dd;
      const double gn = -gp.c64 * exp(-gp.a64 * n2b) / n2b * dn;
      const float sp = (float)(0.5 * (g0 + gn)), sm = (float)(0.5 * (g0 - gn));
      const float2 r = buf[l * LS + i];
      const float2 zm = cconj(buf[(1 - l) * LS + (N - i) % N]);
      ph[p] = make_float2(sp * r.x + sm * zm.x, sp * r.y + sm * zm.y);
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int e = tid + p * T;
    if (e < 2 * N) {
      const int l = e / N, i = e % N;
      const float nze = (i == N / 2) ? 0.0f : (float)freq_of(i, N);
      buf[l * LS + i] = ph[p];
      buf[(2 + l) * LS + i] = make_float2(nze * ph[p].y, -nze * ph[p].x);
    }
  }
  __syncthreads();
  fft_lds<N, true, 4, T, LS, 1, false, N>(buf, tw, tid);
  for (int e = tid; e < 4 * N; e += T) {
    const int l = e / N, z = e % N;
    if ((l & 1) && ky == kym) continue;  // self-paired row: written once
    float2* dst = l < 2 ? A : B;
    dst[(size_t)z * plane + (size_t)((l & 1) ? kym : ky) * H] = buf[l * LS + z];
  }
}

template <int N>
__global__ __launch_bounds__(256) void k_fft_z_green_col0(const float2* __restrict__ P, float2* __restrict__ A,
                                                          float2* __restrict__ B, const float2* __restrict__ tw,
                                                          const GreenParams gp) {
  fft_z_green_col0_body<N>(P, A, B, tw, gp);
}

template <int N>
__global__ __launch_bounds__(256) void k_fft_z_green_col0_box(const BoxArgs* __restrict__ bx, const float2* __restrict__ tw_g) {
  const BoxArgs& b = bx[blockIdx.z];
  fft_z_green_col0_body<N>(b.P, b.A, b.B, tw_g, b.gp);
}

// Inverse y transforms: A -> A, -i n_y A -> P, B -> B, one after the other,
// each starting from registers and storing from registers.  With MUL the
// three take the multipliers codes.x (A), .y (P, from A) and .z (B) instead
// (the initial-condition passes); the force solve instantiates MUL = false.
template <int N, int C, int T, bool MUL = false, int PART = 0>
__device__ __forceinline__ void fft_y_inv_body(float2* __restrict__ A, float2* __restrict__ B,
                                                 float2* __restrict__ P, const float2* __restrict__ tw_g,
                                                 int3 codes) {
  // PART 1: only the two fields from A (A and P); PART 2: only B.  Two such
  // launches read each input once, like the combined kernel, with fewer
  // live registers (for 16 values per thread at N = 1024).
  constexpr int H = N / 2, IS = C + 1;
  constexpr bool DO_A = PART != 2, DO_B = PART != 1;
  using F = RegFft<N, true, false, C, T, 1, IS, true, N>;
  __shared__ float2 buf[N * IS];
  const int tid = threadIdx.x;
  // twiddles in LDS: the FFT stages then read them with LDS instructions, so
  // no global-load wait of a stage also waits for unrelated outstanding loads
  __shared__ float2 twl[N];
  for (int e = tid; e < N; e += (int)blockDim.x) twl[e] = tw_g[e];
  const float2* tw = twl;
  const size_t off = (size_t)blockIdx.y * mesh_plane(N) + (size_t)line_tile<C, H / C>(blockIdx.x) * C;
  float2* a = A + off;
  float2* b = B + off;
  float2* pp = P + off;
  float2 out[F::PERL][F::RL];
  if constexpr (DO_A) {
    float2 av[F::PER0][F::R0], w[F::PER0][F::R0];
    for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      av[p][r] = ok ? a[(unsigned)i * H + l] : make_float2(0.f, 0.f);
    });
    float2 bv[F::PER0][F::R0];
    if constexpr (DO_B) {
      // field B is loaded now, so its load is in flight during both A transforms
      for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
        bv[p][r] = ok ? b[(unsigned)i * H + l] : make_float2(0.f, 0.f);
      });
    }
    if constexpr (MUL) {
      for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
        const float nye = (i == N / 2) ? 0.0f : (float)freq_of(i, N);
        w[p][r] = ic_mul(codes.x, nye, av[p][r]);
        if constexpr (DO_B) bv[p][r] = ic_mul(codes.z, nye, bv[p][r]);
      });
    } else {
#pragma unroll
      for (int p = 0; p < F::PER0; ++p)
#pragma unroll
        for (int r = 0; r < F::R0; ++r) w[p][r] = av[p][r];
    }
    F::run(w, out, buf, tw, tid);
    for_edge<N, F::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      if (ok) a[(unsigned)i * H + l] = out[p][r];
    });
    for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      const float nye = (i == N / 2) ? 0.0f : (float)freq_of(i, N);
      if constexpr (MUL)
        w[p][r] = ic_mul(codes.y, nye, av[p][r]);
      else
        w[p][r] = make_float2(nye * av[p][r].y, -nye * av[p][r].x);
    });
    F::run(w, out, buf, tw, tid);
    for_edge<N, F::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      if (ok) pp[(unsigned)i * H + l] = out[p][r];
    });
    if constexpr (DO_B) {
      F::run(bv, out, buf, tw, tid);
      for_edge<N, F::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
        if (ok) b[(unsigned)i * H + l] = out[p][r];
      });
    }
  } else {
    float2 bv[F::PER0][F::R0];
    for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      bv[p][r] = ok ? b[(unsigned)i * H + l] : make_float2(0.f, 0.f);
    });
    if constexpr (MUL) {
      for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
        const float nye = (i == N / 2) ? 0.0f : (float)freq_of(i, N);
        bv[p][r] = ic_mul(codes.z, nye, bv[p][r]);
      });
    }
    // no barrier before the first stage's LDS writes is needed: run begins with one
    F::run(bv, out, buf, tw, tid);
    for_edge<N, F::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      if (ok) b[(unsigned)i * H + l] = out[p][r];
    });
  }
}

template <int N, int C, int T, bool MUL = false, int PART = 0>
__global__ __launch_bounds__(T) void k_fft_y_inv(float2* __restrict__ A, float2* __restrict__ B,
                                                 float2* __restrict__ P, const float2* __restrict__ tw_g,
                                                 int3 codes) {
  fft_y_inv_body<N, C, T, MUL, PART>(A, B, P, tw_g, codes);
}

template <int N, int C, int T, int PART = 0>
__global__ __launch_bounds__(T) void k_fft_y_inv_box(const BoxArgs* __restrict__ bx, const float2* __restrict__ tw_g) {
  const BoxArgs& b = bx[blockIdx.z];
  fft_y_inv_body<N, C, T, false, PART>(b.A, b.B, b.P, tw_g, make_int3(0, 0, 0));
}

// Persistent forms of the y passes: gridDim.x workgroups walk the (column
// tile, z plane) items in a grid-stride loop and load the next item's field
// while the current item is transformed.  In the inverse pass the current
// item's A values wait in LDS (each thread's own slots) for the -i n_y field,
// which leaves registers for B and for the next item's A.
template <int N, int C, int T>
__global__ __launch_bounds__(T) void k_fft_y_fwd_persist(float2* __restrict__ P, const float2* __restrict__ tw_g) {
  constexpr int H = N / 2, IS = C + 1, TILES = H / C;
  using F = RegFft<N, false, false, C, T, 1, IS, true, N>;
  __shared__ float2 buf[N * IS];
  const int tid = threadIdx.x;
  __shared__ float2 twl[N];
  for (int e = tid; e < N; e += (int)blockDim.x) twl[e] = tw_g[e];
  const float2* tw = twl;
  const int items = TILES * N;
  float2 vin[F::PER0][F::R0];
  int item = blockIdx.x;
  if (item < items) {
    const float2* b0 = P + (size_t)(item / TILES) * mesh_plane(N) + (size_t)line_tile<C, TILES>(item % TILES) * C;
    for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      if (ok) vin[p][r] = b0[(unsigned)i * H + l];
    });
  }
  for (; item < items; item += gridDim.x) {
    float2* base = P + (size_t)(item / TILES) * mesh_plane(N) + (size_t)line_tile<C, TILES>(item % TILES) * C;
    float2 vout[F::PERL][F::RL];
    F::run(vin, vout, buf, tw, tid);
    const int next = item + (int)gridDim.x;
    if (next < items) {
      const float2* b1 = P + (size_t)(next / TILES) * mesh_plane(N) + (size_t)line_tile<C, TILES>(next % TILES) * C;
      for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
        if (ok) vin[p][r] = b1[(unsigned)i * H + l];
      });
    }
    for_edge<N, F::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      if (ok) base[(unsigned)i * H + l] = vout[p][r];
    });
  }
}

template <int N, int C, int T>
__global__ __launch_bounds__(T) void k_fft_y_inv_persist(float2* __restrict__ A, float2* __restrict__ B,
                                                         float2* __restrict__ P, const float2* __restrict__ tw_g) {
  constexpr int H = N / 2, IS = C + 1, TILES = H / C;
  using F = RegFft<N, true, false, C, T, 1, IS, true, N>;
  __shared__ float2 buf[N * IS];
  constexpr int KEEP = F::PER0 * F::R0;
  __shared__ float2 keep[KEEP * T];
  const int tid = threadIdx.x;
  __shared__ float2 twl[N];
  for (int e = tid; e < N; e += (int)blockDim.x) twl[e] = tw_g[e];
  const float2* tw = twl;
  const int items = TILES * N;
  float2 av[F::PER0][F::R0];
  int item = blockIdx.x;
  if (item < items) {
    const float2* a0 = A + (size_t)(item / TILES) * mesh_plane(N) + (size_t)line_tile<C, TILES>(item % TILES) * C;
    for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      av[p][r] = ok ? a0[(unsigned)i * H + l] : make_float2(0.f, 0.f);
    });
  }
  for (; item < items; item += gridDim.x) {
    const size_t off = (size_t)(item / TILES) * mesh_plane(N) + (size_t)line_tile<C, TILES>(item % TILES) * C;
    float2* a = A + off;
    float2* b = B + off;
    float2* pp = P + off;
    float2 bv[F::PER0][F::R0];
    for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
      bv[p][r] = ok ? b[(unsigned)i * H + l] : make_float2(0.f, 0.f);
    });
#pragma unroll
    for (int p = 0; p < F::PER0; ++p)
#pragma unroll
      for (int r = 0; r < F::R0; ++r) keep[(p * F::R0 + r) * T + tid] = av[p][r];
    {
      float2 out[F::PERL][F::RL];
      F::run(av, out, buf, tw, tid);
      const int next = item + (int)gridDim.x;
      if (next < items) {
        const float2* a1 = A + (size_t)(next / TILES) * mesh_plane(N) + (size_t)line_tile<C, TILES>(next % TILES) * C;
        for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
          av[p][r] = ok ? a1[(unsigned)i * H + l] : make_float2(0.f, 0.f);
        });
      }
      for_edge<N, F::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
        if (ok) a[(unsigned)i * H + l] = out[p][r];
      });
    }
    {
      float2 w[F::PER0][F::R0], out[F::PERL][F::RL];
      for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
        const float nye = (i == N / 2) ? 0.0f : (float)freq_of(i, N);
        const float2 v = keep[(p * F::R0 + r) * T + tid];
        w[p][r] = make_float2(nye * v.y, -nye * v.x);
      });
      F::run(w, out, buf, tw, tid);
      for_edge<N, F::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
        if (ok) pp[(unsigned)i * H + l] = out[p][r];
      });
    }
    {
      float2 out[F::PERL][F::RL];
      F::run(bv, out, buf, tw, tid);
      for_edge<N, F::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
        if (ok) b[(unsigned)i * H + l] = out[p][r];
      });
    }
  }
}

// CIC gather of (Fx, Fy, Fz) at base cell (ix, iy, iz) with in-cell offsets
// (dx, dy, dz) (fplane: the float stride between z planes of the force
// meshes), cic.rs:100-170.
__device__ __forceinline__ float3 cic_gather(int ix, int iy, int iz, float dx, float dy, float dz, int nm,
                                             long long fplane, const float* __restrict__ fx,
                                             const float* __restrict__ fy, const float* __restrict__ fz) {
  const float wx[2] = {1.0f - dx, dx}, wy[2] = {1.0f - dy, dy}, wz[2] = {1.0f - dz, dz};
  const int xs[2] = {ix, ix + 1 == nm ? 0 : ix + 1};
  const int ys[2] = {iy, iy + 1 == nm ? 0 : iy + 1};
  const int zs[2] = {iz, iz + 1 == nm ? 0 : iz + 1};
  float sx = 0.f, sy = 0.f, sz = 0.f;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const float w = wx[c] * wy[b] * wz[a];
        const size_t q = zs[a] * fplane + (size_t)ys[b] * nm + xs[c];
        sx = fmaf(fx[q], w, sx);
        sy = fmaf(fy[q], w, sy);
        sz = fmaf(fz[q], w, sz);
      }
  return make_float3(sx, sy, sz);
}

// The gather for every particle, walked in the deposit's sorted order (tile
// by tile along the Hilbert curve), so the 8 corners of consecutive
// particles come from the same few cache lines whatever order the caller's
// arrays are in; results go to the particles' own slots (perm).  Each
// particle's value is the same as the unsorted gather's.
template <typename O>
__global__ __launch_bounds__(256) void k_interp_sorted(const PartRec* __restrict__ sorted,
                                                       const unsigned* __restrict__ perm, long long n,
                                                       const CellMap cm, long long fplane,
                                                       const float* __restrict__ fx, const float* __restrict__ fy,
                                                       const float* __restrict__ fz, const AccDst<O> out) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const PartRec p = sorted[j];
  const float k = 1.0f / (float)kFracOne;
  const float3 a = cic_gather(rec_cell(p, 0), rec_cell(p, 1), rec_cell(p, 2),
                              (float)(p.a[0] & (kFracOne - 1ull)) * k, (float)(p.a[1] & (kFracOne - 1ull)) * k,
                              (float)(p.a[2] & (kFracOne - 1ull)) * k, cm.nm, fplane, fx, fy, fz);
  const long long o = (long long)perm[j] * out.stride;
  out.x[o] = (O)a.x;
  out.y[o] = (O)a.y;
  out.z[o] = (O)a.z;
}

// CIC gather of (Fx, Fy, Fz) at particle idx[j] (idx null: j itself) in the
// caller's order; the cell coordinate is formed in FP64 (cell_frac32).
template <typename T, typename O>
__device__ __forceinline__ void interp_body(const PosSrc<T> src, const long long* __restrict__ idx, long long k,
                                                const CellMap cm, long long fplane, const float* __restrict__ fx,
                                                const float* __restrict__ fy, const float* __restrict__ fz,
                                                const AccDst<O> out) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  const long long i = (idx ? idx[j] : j) * src.stride;
  const double px = (double)src.x[i], py = (double)src.y[i], pz = (double)src.z[i];
  int ix, iy, iz;
  float dx, dy, dz;
  cell_frac32(px, cm, ix, dx);
  cell_frac32(py, cm, iy, dy);
  cell_frac32(pz, cm, iz, dz);
  const float3 a = cic_gather(ix, iy, iz, dx, dy, dz, cm.nm, fplane, fx, fy, fz);
  const long long o = j * out.stride;
  out.x[o] = (O)a.x;
  out.y[o] = (O)a.y;
  out.z[o] = (O)a.z;
}

template <typename T, typename O>
__global__ __launch_bounds__(256) void k_interp(const PosSrc<T> src, const long long* __restrict__ idx, long long k,
                                                const CellMap cm, long long fplane, const float* __restrict__ fx,
                                                const float* __restrict__ fy, const float* __restrict__ fz,
                                                const AccDst<O> out) {
  interp_body<T, O>(src, idx, k, cm, fplane, fx, fy, fz, out);
}

__global__ __launch_bounds__(256) void k_interp_box(const BoxArgs* __restrict__ bx, const PosSrc<float> src, const AccDst<float> out) {
  const BoxArgs& b = bx[blockIdx.z];
  interp_body<float, float>(PosSrc<float>{src.x + b.p0, src.y + b.p0, src.z + b.p0, 1}, nullptr, b.n, b.cm,
                          2 * mesh_plane(b.cm.nm), reinterpret_cast<const float*>(b.A),
                          reinterpret_cast<const float*>(b.P), reinterpret_cast<const float*>(b.B),
                          AccDst<float>{out.x + b.p0, out.y + b.p0, out.z + b.p0, 1});
}

}  // namespace gng
