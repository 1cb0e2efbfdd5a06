// pk_kernels.hpp -- power spectrum back end on the PM plan's spectrum: the z
// transform of the half spectrum after the x and y passes, and the binning of
// |M_k|^2 / W(k)^2 into spherical shells.
//
// Reference: crates/gadget-ng-analysis/src/power_spectrum.rs:35-127.  For the
// mass per cell M, delta = M / (rho_bar V_cell) - 1, so for k != 0
//   |delta_k|^2 (V/N^3)^2 = |M_k|^2 (V / M_tot)^2,
// and the kernels sum |M_k|^2 / W^2 per shell; the reduction applies
// (V / M_tot)^2.  The shell of integer |n| is floor(|n| - 1/2) for
// 1/2 <= |n| < N/2; W = prod_alpha sinc(n_alpha / N) (one sinc per axis, as
// the reference divides by (wx wy wz)^2).
//
// The half spectrum holds kx = 1 .. N/2-1 once for the pair {k, -k} (weight 2)
// and, packed in column 0, the kx = 0 and kx = N/2 planes (weight 1 each).
// Each workgroup bins into LDS and writes its own partial sums; the reduction
// adds the partials in a fixed order.  Only the order of the LDS additions
// inside a workgroup varies, so results repeat to about 1e-15 relative.
#pragma once
#include "pm_kernels.hpp"

namespace gng {

// Shell of integer |n|^2 = k2 > 0, floor(sqrt(k2) - 1/2), or -1 outside
// [1/2, N/2).  FP32 gives the same floor as FP64 for k2 < 2^24: sqrt(k2) is
// then within one FP32 ulp (< 6.2e-5 below 1024) of the exact root, while the
// root of an integer stays at least 0.25 / (2 sqrt(k2) + 1) > 1.4e-4 from any
// half-integer.
__device__ __forceinline__ int pk_shell(int k2, int nbins) {
  const float bf = sqrtf((float)k2) - 0.5f;
  return (bf >= 0.0f && bf < (float)nbins) ? (int)bf : -1;
}

// Columns kx >= 1 of one (column tile, ky): forward z transform in registers,
// then binning.  iw2[i] = 1 / sinc(freq(i) / N)^2.  Partials at part[s * nblk + block].
template <int N, int C, int T>
__global__ __launch_bounds__(T) void k_fft_z_pk(const float2* __restrict__ P, const float2* __restrict__ tw_g,
                                                const double* __restrict__ iw2, double* __restrict__ part,
                                                long long nblk, unsigned long long* __restrict__ modes) {
  constexpr int H = N / 2, IS = C + 1, NB = H;
  using FF = RegFft<N, false, false, C, T, 1, IS, true, N>;
  __shared__ float2 buf[N * IS];
  __shared__ float2 twl[N];
  __shared__ double bsum[NB];
  __shared__ unsigned bcnt[NB];
  const int tid = threadIdx.x;
  for (int e = tid; e < N; e += T) twl[e] = tw_g[e];
  for (int e = tid; e < NB; e += T) {
    bsum[e] = 0.0;
    bcnt[e] = 0u;
  }
  const int tile = line_tile<C, H / C>(blockIdx.x), ky = blockIdx.y;
  const unsigned plane = (unsigned)mesh_plane(N);
  const size_t col0 = (size_t)ky * H + (size_t)tile * C;
  float2 vin[FF::PER0][FF::R0], v[FF::PERL][FF::RL];
  for_edge<N, FF::R0, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
    if (ok) vin[p][r] = P[col0 + (unsigned)i * plane + l];
  });
  // begins with a barrier, so the LDS initialisation above is complete
  FF::run(vin, v, buf, twl, tid);
  const int ny = freq_of(ky, N);
  const double iwy = iw2[ky];
  // Back to LDS in natural order (mode (l, i) at buf[i*IS + l]); then each
  // thread walks RUN consecutive kz of one column, where the shell changes
  // only every few modes, and adds each run of one shell with one atomic.
  __syncthreads();
  for_edge<N, FF::RL, C, T, true>(tid, [&](int p, int r, int l, int i, bool ok) {
    if (ok) buf[i * IS + l] = v[p][r];
  });
  __syncthreads();
  constexpr int RUN = (C * N + T - 1) / T;
  int cur = -1;
  double acc = 0.0;
  unsigned cnt = 0;
#pragma unroll
  for (int q = 0; q < RUN; ++q) {
    const int e = tid * RUN + q;
    const int l = e / N, i = e % N;
    const int kx = tile * C + l;
    int sh = -1;
    double d2 = 0.0;
    if (e < C * N && kx != 0) {
      const int nz = freq_of(i, N);
      sh = pk_shell(kx * kx + ny * ny + nz * nz, NB);
      const float2 u = buf[i * IS + l];
      const double re = (double)u.x, im = (double)u.y;
      d2 = 2.0 * ((re * re + im * im) * (iw2[kx] * iwy * iw2[i]));  // k and -k
    }
    if (sh != cur) {
      if (cur >= 0) {
        atomicAdd(&bsum[cur], acc);
        atomicAdd(&bcnt[cur], cnt);
      }
      cur = sh;
      acc = 0.0;
      cnt = 0u;
    }
    if (sh >= 0) {
      acc += d2;
      cnt += 2u;
    }
  }
  if (cur >= 0) {
    atomicAdd(&bsum[cur], acc);
    atomicAdd(&bcnt[cur], cnt);
  }
  __syncthreads();
  const size_t blk = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
  for (int e = tid; e < NB; e += T) {
    part[(size_t)e * nblk + blk] = bsum[e];
    if (bcnt[e]) atomicAdd(&modes[e], (unsigned long long)bcnt[e]);
  }
}

// The packed column kx = 0 for the pair {ky, N - ky}: Z = R0 + i RN with R0,
// RN the spectra of the kx = 0 and kx = N/2 planes, so
//   R0(k) = (Z(k) + conj Z(-k)) / 2,  RN(k) = (Z(k) - conj Z(-k)) / (2i).
// Partials at part[s * nblk + base + blockIdx.x].
template <int N>
__global__ __launch_bounds__(256) void k_pk_col0(const float2* __restrict__ P, const float2* __restrict__ tw,
                                                 const double* __restrict__ iw2, double* __restrict__ part,
                                                 long long nblk, long long base,
                                                 unsigned long long* __restrict__ modes) {
  constexpr int H = N / 2, T = 256, LS = N + 1, NB = H;
  __shared__ float2 buf[2 * LS];
  __shared__ double bsum[NB];
  __shared__ unsigned bcnt[NB];
  const int tid = threadIdx.x;
  const int ky = blockIdx.x, kym = (N - ky) % N;
  const size_t plane = (size_t)mesh_plane(N);
  for (int e = tid; e < NB; e += T) {
    bsum[e] = 0.0;
    bcnt[e] = 0u;
  }
  for (int e = tid; e < 2 * N; e += T) {
    const int l = e / N, z = e % N;
    buf[l * LS + z] = P[(size_t)z * plane + (size_t)(l ? kym : ky) * H];
  }
  __syncthreads();
  fft_lds<N, false, 2, T, LS, 1, false, N>(buf, tw, tid);
  // lines ky and N - ky; for ky = 0 and ky = N/2 they are one line
  const int lines = kym == ky ? 1 : 2;
  for (int e = tid; e < lines * N; e += T) {
    const int l = e / N, i = e % N;
    const int yi = l ? kym : ky;
    const int ny = freq_of(yi, N), nz = freq_of(i, N);
    const float2 zk = buf[l * LS + i];
    const float2 zm = cconj(buf[(1 - l) * LS + (N - i) % N]);
    const double r0x = 0.5 * ((double)zk.x + (double)zm.x), r0y = 0.5 * ((double)zk.y + (double)zm.y);
    const double rnx = 0.5 * ((double)zk.y - (double)zm.y), rny = -0.5 * ((double)zk.x - (double)zm.x);
    const double iwyz = iw2[yi] * iw2[i];
    const int k2 = ny * ny + nz * nz;
    if (k2 > 0) {
      const int s = pk_shell(k2, NB);
      if (s >= 0) {
        atomicAdd(&bsum[s], (r0x * r0x + r0y * r0y) * (iw2[0] * iwyz));
        atomicAdd(&bcnt[s], 1u);
      }
    }
    const int s = pk_shell(H * H + k2, NB);
    if (s >= 0) {
      atomicAdd(&bsum[s], (rnx * rnx + rny * rny) * (iw2[H] * iwyz));
      atomicAdd(&bcnt[s], 1u);
    }
  }
  __syncthreads();
  for (int e = tid; e < NB; e += T) {
    part[(size_t)e * nblk + base + blockIdx.x] = bsum[e];
    if (bcnt[e]) atomicAdd(&modes[e], (unsigned long long)bcnt[e]);
  }
}

// Per-block sums of the masses, each block in a fixed order: msum[blockIdx.x].
template <typename M>
__global__ __launch_bounds__(256) void k_mass_partials(const M* __restrict__ m, long long n,
                                                       double* __restrict__ msum) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  // four independent chains, so four loads are in flight; added in a fixed order
  const long long stride = (long long)gridDim.x * 256;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  long long i = (long long)blockIdx.x * 256 + tid;
  for (; i + 3 * stride < n; i += 4 * stride) {
    s0 += (double)m[i];
    s1 += (double)m[i + stride];
    s2 += (double)m[i + 2 * stride];
    s3 += (double)m[i + 3 * stride];
  }
  for (; i < n; i += stride) s0 += (double)m[i];
  red[tid] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  if (tid == 0) msum[blockIdx.x] = red[0];
}

// msum[nms] = the sum of msum[0 .. nms), in a fixed order (one workgroup of
// 1024 threads, nms <= 1024).
__global__ __launch_bounds__(1024) void k_mass_total(double* __restrict__ msum, int nms) {
  __shared__ double red[1024];
  const int tid = threadIdx.x;
  red[tid] = tid < nms ? msum[tid] : 0.0;
  __syncthreads();
  for (int h = 512; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  if (tid == 0) msum[nms] = red[0];
}

// One workgroup per shell: sum of the shell's nblk partials in a fixed order,
// times (V / M_tot)^2.
__global__ __launch_bounds__(256) void k_pk_reduce(const double* __restrict__ part, long long nblk,
                                                   const double* __restrict__ mtot, double vol,
                                                   double* __restrict__ out) {
  __shared__ double red[256];
  const int tid = threadIdx.x, s = blockIdx.x;
  const double* row = part + (size_t)s * nblk;
  double acc = 0.0;
  for (long long q = tid; q < nblk; q += 256) acc += row[q];
  red[tid] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) red[tid] += red[tid + h];
    __syncthreads();
  }
  if (tid == 0) {
    const double sc = vol / *mtot;
    out[s] = red[0] * sc * sc;
  }
}

__global__ void k_zero_u64(unsigned long long* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

}  // namespace gng
