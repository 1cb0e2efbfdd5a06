// ic_kernels.hpp -- Zel'dovich initial conditions on the PM plan's mesh: the
// Gaussian field delta-hat generated mode by mode in the z pass, the three
// displacement components through the plan's inverse y and x passes, and
// the lattice particles displaced by them.
//
// Reference: crates/gadget-ng-core/src/ic_zeldovich.rs.
// * Each mode (ix, iy, iz) of the reference's flat order ix n^2 + iy n + iz
//   draws two Gaussians from its own generator, seeded by
//   hash(seed ^ hash(ix n^2 + iy n + iz)) and by hash(that ^ 0x9E37..15)
//   (:69-98, :320-324).  The generator is rand's StdRng: ChaCha12 keyed by
//   rand_core's seed_from_u64, eight PCG32 outputs (:93-98).
// * delta-hat = sigma(|n|) (g_r + i g_i); 0 at DC and at every mode with a
//   Nyquist component (:309-318); each pair {k, -k} keeps the value of its
//   lexicographically smaller index and the other gets the conjugate
//   (:333-366).
// * Psi_a = Re IFFT(i n_a delta-hat / |n|^2) d / n^3 (:541-603), lattice
//   points q = (i + 1/2) d, x = rem_euclid(q + s Psi, L), p = s Psi v
//   (:733-766).
//
// Axis mapping: the reference's slowest index ix is the mesh's z (planes),
// iy its y, and the fastest iz its x (the halved axis of the half spectrum).
// The reference's Psi_x is therefore the mesh's z gradient (field B after
// the inverse passes), Psi_y the y gradient (P) and Psi_z the x gradient (A).
// In the half spectrum the packed column 0 holds the kx = 0 plane plus i
// times the kx = N/2 plane; the latter is zero here (Nyquist), so column 0 is
// the kx = 0 plane itself and needs no unpacking.
#pragma once
#include "pm_kernels.hpp"

namespace gng {

struct IcParams {
  unsigned long long seed;
  const double* sigma;  // sigma[n2] = spectrum(sqrt(n2)), n2 = 0 .. 3 (N/2 - 1)^2
};

// murmur3 fmix64 (ic_zeldovich.rs:69-76)
__host__ __device__ __forceinline__ unsigned long long ic_hash(unsigned long long x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

__device__ __forceinline__ unsigned rotl32(unsigned x, int r) { return (x << r) | (x >> (32 - r)); }

#define GNG_CHACHA_QR(a, b, c, d) \
  a += b;                         \
  d = rotl32(d ^ a, 16);          \
  c += d;                         \
  b = rotl32(b ^ c, 12);          \
  a += b;                         \
  d = rotl32(d ^ a, 8);           \
  c += d;                         \
  b = rotl32(b ^ c, 7);

// The first two next_u64 of StdRng::seed_from_u64(seed): the key is eight
// PCG32 XSH-RR outputs, each taken after one LCG step (multiplier
// 6364136223846793005, increment 11634580027462260723); ChaCha12 block 0
// with counter and stream 0; u0 = words (0, 1), u1 = words (2, 3), low first.
__device__ __forceinline__ void stdrng_first_two(unsigned long long seed, unsigned long long& u0,
                                                 unsigned long long& u1) {
  unsigned k[8];
  unsigned long long st = seed;
#pragma unroll
  for (int w = 0; w < 8; ++w) {
    st = st * 6364136223846793005ull + 11634580027462260723ull;
    const unsigned xs = (unsigned)(((st >> 18) ^ st) >> 27);
    const unsigned rot = (unsigned)(st >> 59);
    k[w] = (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
  unsigned x0 = 0x61707865u, x1 = 0x3320646eu, x2 = 0x79622d32u, x3 = 0x6b206574u;
  unsigned x4 = k[0], x5 = k[1], x6 = k[2], x7 = k[3], x8 = k[4], x9 = k[5], x10 = k[6], x11 = k[7];
  unsigned x12 = 0u, x13 = 0u, x14 = 0u, x15 = 0u;
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    GNG_CHACHA_QR(x0, x4, x8, x12)
    GNG_CHACHA_QR(x1, x5, x9, x13)
    GNG_CHACHA_QR(x2, x6, x10, x14)
    GNG_CHACHA_QR(x3, x7, x11, x15)
    GNG_CHACHA_QR(x0, x5, x10, x15)
    GNG_CHACHA_QR(x1, x6, x11, x12)
    GNG_CHACHA_QR(x2, x7, x8, x13)
    GNG_CHACHA_QR(x3, x4, x9, x14)
  }
  x0 += 0x61707865u;
  x1 += 0x3320646eu;
  x2 += 0x79622d32u;
  x3 += 0x6b206574u;
  u0 = ((unsigned long long)x1 << 32) | x0;
  u1 = ((unsigned long long)x3 << 32) | x2;
}
#undef GNG_CHACHA_QR

// gaussian_std (ic_zeldovich.rs:88-98): two 53-bit uniforms, the first kept
// at or above 1e-300, then sqrt(-2 ln u1) cos(2 pi u2).
__device__ __forceinline__ double ic_gaussian(unsigned long long s) {
#pragma clang fp contract(off)
  unsigned long long a, b;
  stdrng_first_two(s, a, b);
  double u1 = (double)(a >> 11) / 9007199254740992.0;
  const double u2 = (double)(b >> 11) / 9007199254740992.0;
  u1 = u1 > 1e-300 ? u1 : 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// delta-hat of mode (ix, iy, iz) in the reference's index order, and |n|^2.
__device__ __forceinline__ double2 ic_mode(const IcParams& ip, int n, int ix, int iy, int iz, int& n2) {
#pragma clang fp contract(off)
  const int half = n / 2;
  const int nx = freq_of(ix, n), ny = freq_of(iy, n), nz = freq_of(iz, n);
  n2 = nx * nx + ny * ny + nz * nz;
  if (n2 == 0 || ix == half || iy == half || iz == half) return make_double2(0.0, 0.0);
  const int jx = ix ? n - ix : 0, jy = iy ? n - iy : 0, jz = iz ? n - iz : 0;
  // the pair's source: the smaller flat index
  const bool self = ix != jx ? ix < jx : (iy != jy ? iy < jy : iz < jz);
  const int sx = self ? ix : jx, sy = self ? iy : jy, sz = self ? iz : jz;
  const unsigned long long nn = (unsigned long long)n;
  const unsigned long long packed = (unsigned long long)sx * nn * nn + (unsigned long long)sy * nn + (unsigned long long)sz;
  const unsigned long long ms = ic_hash(ip.seed ^ ic_hash(packed));
  const double gr = ic_gaussian(ms);
  const double gi = ic_gaussian(ic_hash(ms ^ 0x9E3779B97F4A7C15ull));
  const double sg = ip.sigma[n2];
  return self ? make_double2(sg * gr, sg * gi) : make_double2(sg * gr, -(sg * gi));
}

// Every mode of the half spectrum iz = 0 .. n/2 in the reference's order,
// out[(ix n + iy) (n/2 + 1) + iz] (FP64, for the library's checks).
__global__ __launch_bounds__(256) void k_ic_delta(IcParams ip, int n, double2* __restrict__ out) {
  const int hz = n / 2 + 1;
  const long long total = (long long)n * n * hz;
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  if (e >= total) return;
  const int iz = (int)(e % hz);
  const long long r = e / hz;
  int n2;
  out[e] = ic_mode(ip, n, (int)(r / n), (int)(r % n), iz, n2);
}

// One (column tile, ky) item: the modes of C columns along all kz generated
// into LDS as a = -delta-hat / |n|^2, then the inverse z transforms of
// m_A(n_z) a and m_B(n_z) a (multiplier codes of ic_mul: zc.x, zc.y; the
// displacement takes 1 and -i n_z), written to the plan's A and B meshes in
// the layout the inverse y pass reads.  Mode (column kx, row ky, plane kz) is
// the reference's (ix, iy, iz) = (kz, ky, kx).
template <int N, int C, int T>
__global__ __launch_bounds__(T) void k_ic_z(const IcParams ip, float2* __restrict__ A, float2* __restrict__ B,
                                            const float2* __restrict__ tw_g, const int2 zc) {
  constexpr int H = N / 2, IS = C + 1;
  using F = RegFft<N, true, false, C, T, 1, IS, true, N>;
  __shared__ float2 buf[N * IS];
  __shared__ float2 twl[N];
  const int tid = threadIdx.x;
  for (int e = tid; e < N; e += T) twl[e] = tw_g[e];
  const int tile = line_tile<C, H / C>(blockIdx.x), ky = blockIdx.y;
  const unsigned plane = (unsigned)mesh_plane(N);
  const size_t col0 = (size_t)ky * H + (size_t)tile * C;
#pragma unroll 1
  for (int e = tid; e < C * N; e += T) {
    const int l = e % C, i = e / C;
    int n2;
    const double2 d = ic_mode(ip, N, i, ky, tile * C + l, n2);
    const double dn = n2 > 0 ? (double)n2 : 1.0;  // delta-hat is 0 at n2 = 0
    buf[i * IS + l] = make_float2((float)(-d.x / dn), (float)(-d.y / dn));
  }
  __syncthreads();
  float2 a[F::PER0][F::R0], w[F::PER0][F::R0], out[F::PERL][F::RL];
  for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
    a[p][r] = ok ? buf[i * IS + l] : make_float2(0.f, 0.f);
  });
  for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
    const float nze = (i == N / 2) ? 0.0f : (float)freq_of(i, N);
    w[p][r] = ic_mul(zc.x, nze, a[p][r]);
  });
  // begins with a barrier: every thread has read its values from LDS
  F::run(w, out, buf, twl, tid);
  for_edge<N, F::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
    if (ok) A[col0 + (unsigned)i * plane + l] = out[p][r];
  });
  for_edge<N, F::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
    const float nze = (i == N / 2) ? 0.0f : (float)freq_of(i, N);
    w[p][r] = ic_mul(zc.y, nze, a[p][r]);
  });
  F::run(w, out, buf, twl, tid);
  for_edge<N, F::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
    if (ok) B[col0 + (unsigned)i * plane + l] = out[p][r];
  });
}

// rem_euclid of f64 (Rust: r = x % l; r < 0 ? r + |l| : r); fmod is exact
__device__ __forceinline__ double ic_rem_euclid(double x, double l) {
#pragma clang fp contract(off)
  const double r = fmod(x, l);
  return r < 0.0 ? r + fabs(l) : r;
}

// Particles lo .. lo + cnt of the n^3 lattice (global id ix n^2 + iy n + iz):
// pos and vel as n x 3 interleaved f64 (Vec3).  psi_x, psi_y, psi_z are the
// reference's components (the mesh's B, P and A), real fields in the plan's
// layout (row (z, y) at float 2 z plane + y n); norm = d / n^3.
__global__ __launch_bounds__(256) void k_ic_particles(const float* __restrict__ psi_x, const float* __restrict__ psi_y,
                                                      const float* __restrict__ psi_z, long long fplane, int n,
                                                      double d, double box, double norm, double scale,
                                                      double vel_factor, long long lo, long long cnt,
                                                      double* __restrict__ pos, double* __restrict__ vel, int wrap) {
#pragma clang fp contract(off)
  // the workgroup's 256 particles are staged in LDS so that the interleaved
  // f64 outputs go out as contiguous, coalesced stores
  __shared__ double sp[3 * 256], sv[3 * 256];
  const long long base = (long long)blockIdx.x * 256;
  const long long t = base + threadIdx.x;
  const int nloc = (int)(cnt - base < 256 ? cnt - base : 256);
  if (t < cnt) {
    const long long gid = lo + t;
    const long long nn = (long long)n * n;
    const int ix = (int)(gid / nn);
    const long long rem = gid % nn;
    const int iy = (int)(rem / n), iz = (int)(rem % n);
    const long long off = (long long)ix * fplane + (long long)iy * n + iz;
    double p0 = (double)psi_x[off] * norm, p1 = (double)psi_y[off] * norm, p2 = (double)psi_z[off] * norm;
    if (scale != 1.0) {
      p0 *= scale;
      p1 *= scale;
      p2 *= scale;
    }
    const double q0 = ((double)ix + 0.5) * d, q1 = ((double)iy + 0.5) * d, q2 = ((double)iz + 0.5) * d;
    double* P = sp + 3 * threadIdx.x;
    double* V = sv + 3 * threadIdx.x;
    // without wrap: q + Psi, for the second-order kernel to complete
    P[0] = wrap ? ic_rem_euclid(q0 + p0, box) : q0 + p0;
    P[1] = wrap ? ic_rem_euclid(q1 + p1, box) : q1 + p1;
    P[2] = wrap ? ic_rem_euclid(q2 + p2, box) : q2 + p2;
    V[0] = p0 * vel_factor;
    V[1] = p1 * vel_factor;
    V[2] = p2 * vel_factor;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 3 * nloc; e += 256) {
    pos[3 * base + e] = sp[e];
    vel[3 * base + e] = sv[e];
  }
}

// Second order: the source S of the 2LPT Poisson equation from the six
// second derivatives of phi (ic_2lpt.rs:146-159), accumulated over two
// rounds of three fields: round 0 writes xx yy + yy zz + zz xx (f0..f2 =
// xx, yy, zz), round 1 subtracts xy^2 + yz^2 + xz^2 (f0..f2 in any order).
// The fields are real meshes in the plan's row layout (row (z, y) at float
// z fplane + y n), S an unpadded mesh (plane stride s_plane); FP64
// arithmetic, scaled by `scale` (1/n^6 undoes the unnormalised inverse
// transforms of both factors), S stored as f32.
__global__ __launch_bounds__(256) void k_ic_source(const float* __restrict__ f0, const float* __restrict__ f1,
                                                   const float* __restrict__ f2, float* __restrict__ S,
                                                   long long fplane, long long s_plane, int n, double scale,
                                                   int round) {
  const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long rows = (long long)n * n;
  if (e >= rows * n) return;
  const long long r = e / n, x = e % n;
  const long long off = (r / n) * fplane + (r % n) * n + x;
  const long long so = (r / n) * s_plane + (r % n) * n + x;
  const double a = f0[off], b = f1[off], c = f2[off];
  if (round == 0)
    S[so] = (float)((a * b + b * c + c * a) * scale);
  else
    S[so] = (float)((double)S[so] - (a * a + b * b + c * c) * scale);
}

// Second-order particles: adds D2/D1^2 s^2 Psi2 to the first-order positions
// q + s Psi1 that k_ic_particles left unwrapped (wrap = 0), wraps them, and
// adds s^2 Psi2 v2 to the first-order momenta (ic_2lpt.rs:536-550; the sums
// run in the reference's order).  psi2_* as in k_ic_particles.
__global__ __launch_bounds__(256) void k_ic_particles_2lpt(const float* __restrict__ psi_x,
                                                           const float* __restrict__ psi_y,
                                                           const float* __restrict__ psi_z, long long fplane, int n,
                                                           double box, double norm, double scale2, double d2,
                                                           double vel2, long long lo, long long cnt,
                                                           double* __restrict__ pos, double* __restrict__ vel) {
#pragma clang fp contract(off)
  // first-order values in and results out through LDS, coalesced
  __shared__ double sp[3 * 256], sv[3 * 256];
  const long long base = (long long)blockIdx.x * 256;
  const long long t = base + threadIdx.x;
  const int nloc = (int)(cnt - base < 256 ? cnt - base : 256);
  for (int e = threadIdx.x; e < 3 * nloc; e += 256) {
    sp[e] = pos[3 * base + e];
    sv[e] = vel[3 * base + e];
  }
  __syncthreads();
  if (t < cnt) {
    const long long gid = lo + t;
    const long long nn = (long long)n * n;
    const int ix = (int)(gid / nn);
    const long long rem = gid % nn;
    const int iy = (int)(rem / n), iz = (int)(rem % n);
    const long long off = (long long)ix * fplane + (long long)iy * n + iz;
    double p[3] = {(double)psi_x[off] * norm, (double)psi_y[off] * norm, (double)psi_z[off] * norm};
    double* P = sp + 3 * threadIdx.x;
    double* V = sv + 3 * threadIdx.x;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (scale2 != 1.0) p[c] *= scale2;
      P[c] = ic_rem_euclid(P[c] + d2 * p[c], box);
      V[c] = V[c] + p[c] * vel2;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 3 * nloc; e += 256) {
    pos[3 * base + e] = sp[e];
    vel[3 * base + e] = sv[e];
  }
}

}  // namespace gng
