// pm64.hip -- the reference PM algorithm moved to the GPU in FP64.
//
// This is the "FP64 GPU form" the FP32 product is measured against, not the
// product path.  It follows the reference step by step:
//   cic::assign            crates/gadget-ng-pm/src/cic.rs:37-93 (mass per
//                          cell, weights (1-d, d), every index wrapped), here
//                          one FP64 atomic per particle and node;
//   solve_forces_impl      fft_poisson.rs:306-476: rho = mass / V_cell into a
//                          complex mesh, full complex-to-complex 3-D FFT as
//                          three 1-D passes (x, then y and z strided), the
//                          spectral kernel fft_poisson.rs:509-549 with the
//                          Nyquist index mapped to +nm/2 (freq_index,
//                          :724-730), three inverse 3-D FFTs and Re() / nm^3;
//   cic::interpolate       cic.rs:100-170 in FP64.
// The host entry point copies the f64 particles in and the accelerations out
// on every call, as the reference's GPU crates do.
//
// FFT: radix-2 decimation in time inside LDS, one workgroup per group of C
// adjacent lines (so strided passes read C consecutive complex values per
// row), twiddles from a long-double table.  Mesh sides: powers of two.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/gng_pm.h"
#include "status.hpp"

namespace gng {
namespace pm64 {

constexpr int kThreads = 256;

// lines per workgroup: C * N * 16 bytes of LDS (32 KiB at most), C <= N
__host__ __device__ constexpr int lines_per_block(int n) { return n <= 8 ? n : (n <= 256 ? 8 : (n == 512 ? 4 : 2)); }

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ void k_zero(double2* __restrict__ a, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = make_double2(0.0, 0.0);
}

__device__ __forceinline__ void cell_of(double x, double cs, int nm, int& i0, int& i1, double& d) {
  double c = std::fmod(x * cs, (double)nm);
  if (c < 0.0) c += (double)nm;
  const double f = std::floor(c);
  d = c - f;
  i0 = (int)f % nm;
  i1 = i0 + 1 == nm ? 0 : i0 + 1;
}

__global__ void k_deposit(const double* __restrict__ pos, const double* __restrict__ mass, long long n, int nm,
                          double cs, double inv_vcell, double2* __restrict__ rho) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int ix[2], iy[2], iz[2];
  double dx, dy, dz;
  cell_of(pos[3 * p], cs, nm, ix[0], ix[1], dx);
  cell_of(pos[3 * p + 1], cs, nm, iy[0], iy[1], dy);
  cell_of(pos[3 * p + 2], cs, nm, iz[0], iz[1], dz);
  const double q = mass[p] * inv_vcell;
  const double wx[2] = {1.0 - dx, dx}, wy[2] = {1.0 - dy, dy}, wz[2] = {1.0 - dz, dz};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const long long e = ((long long)iz[a] * nm + iy[b]) * nm + ix[c];
        atomicAdd(&rho[e].x, q * wz[a] * wy[b] * wx[c]);
      }
}

// One 1-D pass of N-point complex FFTs along AXIS (0 = x, contiguous; 1 = y,
// stride N; 2 = z, stride N^2) over the whole mesh; unnormalised, sign -1
// (forward) or +1 (INV).  tw[j] = exp(-2 pi i j / N), j < N/2.
template <int N, int AXIS, bool INV>
__global__ __launch_bounds__(kThreads) void k_fft(double2* __restrict__ data, const double2* __restrict__ tw) {
  constexpr int C = lines_per_block(N);
  constexpr int LOG = __builtin_ctz(N);
  __shared__ double2 buf[C * N];
  const int tid = threadIdx.x;
  long long base;
  long long cstride, jstride;  // element (c, j) at base + c*cstride + j*jstride
  if constexpr (AXIS == 0) {
    base = (long long)blockIdx.x * C * N;
    cstride = N;
    jstride = 1;
  } else {
    const int xg = blockIdx.x % (N / C), o = blockIdx.x / (N / C);
    const long long plane = (long long)N * N;
    base = (AXIS == 1 ? (long long)o * plane : (long long)o * N) + (long long)xg * C;
    cstride = 1;
    jstride = AXIS == 1 ? N : plane;
  }
  // load, bit-reversed within each line
  for (int e = tid; e < C * N; e += kThreads) {
    int c, j;
    if constexpr (AXIS == 0) {
      c = e / N;
      j = e % N;
    } else {
      c = e % C;
      j = e / C;
    }
    const int r = (int)(__brev((unsigned)j) >> (32 - LOG));
    buf[c * N + r] = data[base + c * cstride + j * jstride];
  }
  __syncthreads();
  for (int s = 0; s < LOG; ++s) {
    const int h = 1 << s;
    for (int b = tid; b < C * (N / 2); b += kThreads) {
      const int line = b / (N / 2), k = b % (N / 2);
      const int grp = k >> s, ps = k & (h - 1);
      const int i0 = line * N + grp * 2 * h + ps, i1 = i0 + h;
      double2 w = tw[ps * (N / (2 * h))];
      if (INV) w.y = -w.y;
      const double2 t = cmul(buf[i1], w);
      const double2 u = buf[i0];
      buf[i0] = make_double2(u.x + t.x, u.y + t.y);
      buf[i1] = make_double2(u.x - t.x, u.y - t.y);
    }
    __syncthreads();
  }
  for (int e = tid; e < C * N; e += kThreads) {
    int c, j;
    if constexpr (AXIS == 0) {
      c = e / N;
      j = e % N;
    } else {
      c = e % C;
      j = e / C;
    }
    data[base + c * cstride + j * jstride] = buf[c * N + j];
  }
}

struct Spectral {
  int nm;
  double two_pi_over_l;
  double g4pi;  // -4 pi G
  double rs2;   // r_split^2 (0 = off)
  double eps2;  // plummer eps^2 (0 = off)
  const double* deconv;  // 1/sinc^4(pi n / nm) per index, or null
};

// F_hat_alpha = -i k_alpha Phi_hat, Phi_hat = -4 pi G W rho_hat / k^2, with
// W = exp(-k^2 r_s^2 / 2) exp(-k^2 eps^2) [/ W_CIC^2]; k^2 < 1e-30 -> 0.
// The x component overwrites rho_hat in place.
__global__ void k_spectral(double2* __restrict__ fx, double2* __restrict__ fy, double2* __restrict__ fz, Spectral sp) {
  const long long M = (long long)sp.nm * sp.nm * sp.nm;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= M) return;
  const int nm = sp.nm;
  const int ix = (int)(e % nm), iy = (int)((e / nm) % nm), iz = (int)(e / ((long long)nm * nm));
  // freq_index: indices above nm/2 are negative frequencies; nm/2 stays +nm/2
  const int nx = ix <= nm / 2 ? ix : ix - nm, ny = iy <= nm / 2 ? iy : iy - nm, nz = iz <= nm / 2 ? iz : iz - nm;
  const double kx = sp.two_pi_over_l * nx, ky = sp.two_pi_over_l * ny, kz = sp.two_pi_over_l * nz;
  const double k2 = kx * kx + ky * ky + kz * kz;
  if (k2 < 1e-30) {
    fx[e] = fy[e] = fz[e] = make_double2(0.0, 0.0);
    return;
  }
  double w = 1.0;
  if (sp.rs2 > 0.0) w *= std::exp(-0.5 * k2 * sp.rs2);
  if (sp.eps2 > 0.0) w *= std::exp(-k2 * sp.eps2);
  if (sp.deconv) w *= sp.deconv[ix] * sp.deconv[iy] * sp.deconv[iz];
  const double2 r = fx[e];
  const double f = sp.g4pi * w / k2;
  const double2 phi = make_double2(f * r.x, f * r.y);
  // -i k phi = (k phi.im, -k phi.re)
  fx[e] = make_double2(kx * phi.y, -kx * phi.x);
  fy[e] = make_double2(ky * phi.y, -ky * phi.x);
  fz[e] = make_double2(kz * phi.y, -kz * phi.x);
}

__global__ void k_interp(const double* __restrict__ pos, const long long* __restrict__ idx, long long k, int nm,
                         double cs, double inv_n3, const double2* __restrict__ fx, const double2* __restrict__ fy,
                         const double2* __restrict__ fz, double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= k) return;
  const long long p = idx ? idx[t] : t;
  int ix[2], iy[2], iz[2];
  double dx, dy, dz;
  cell_of(pos[3 * p], cs, nm, ix[0], ix[1], dx);
  cell_of(pos[3 * p + 1], cs, nm, iy[0], iy[1], dy);
  cell_of(pos[3 * p + 2], cs, nm, iz[0], iz[1], dz);
  const double wx[2] = {1.0 - dx, dx}, wy[2] = {1.0 - dy, dy}, wz[2] = {1.0 - dz, dz};
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int c = 0; c < 2; ++c) {
        const long long e = ((long long)iz[a] * nm + iy[b]) * nm + ix[c];
        const double w = wz[a] * wy[b] * wx[c];
        ax += w * fx[e].x;
        ay += w * fy[e].x;
        az += w * fz[e].x;
      }
  out[3 * t] = ax * inv_n3;
  out[3 * t + 1] = ay * inv_n3;
  out[3 * t + 2] = az * inv_n3;
}

struct Ops {
  int nm;
  void (*fft)(double2*, const double2*, int axis, bool inv, hipStream_t);
};

template <int N>
Ops make_ops() {
  Ops o;
  o.nm = N;
  o.fft = [](double2* d, const double2* tw, int axis, bool inv, hipStream_t s) {
    constexpr int C = lines_per_block(N);
    const dim3 grid((unsigned)((long long)N * N / C));
    if (axis == 0 && inv) hipLaunchKernelGGL((k_fft<N, 0, true>), grid, dim3(kThreads), 0, s, d, tw);
    if (axis == 0 && !inv) hipLaunchKernelGGL((k_fft<N, 0, false>), grid, dim3(kThreads), 0, s, d, tw);
    if (axis == 1 && inv) hipLaunchKernelGGL((k_fft<N, 1, true>), grid, dim3(kThreads), 0, s, d, tw);
    if (axis == 1 && !inv) hipLaunchKernelGGL((k_fft<N, 1, false>), grid, dim3(kThreads), 0, s, d, tw);
    if (axis == 2 && inv) hipLaunchKernelGGL((k_fft<N, 2, true>), grid, dim3(kThreads), 0, s, d, tw);
    if (axis == 2 && !inv) hipLaunchKernelGGL((k_fft<N, 2, false>), grid, dim3(kThreads), 0, s, d, tw);
  };
  return o;
}

const Ops* find_ops(int nm) {
  static const Ops table[] = {make_ops<4>(),   make_ops<8>(),   make_ops<16>(),  make_ops<32>(),
                              make_ops<64>(),  make_ops<128>(), make_ops<256>(), make_ops<512>(),
                              make_ops<1024>()};
  for (const Ops& o : table)
    if (o.nm == nm) return &o;
  return nullptr;
}

}  // namespace pm64
}  // namespace gng

using namespace gng::pm64;
using gng::fail;

struct gng_pm64_plan {
  gng_pm_config cfg{};
  const Ops* ops = nullptr;
  int device = 0;
  long long M = 0;
  double2* d_f[3] = {nullptr, nullptr, nullptr};
  double2* d_tw = nullptr;
  double* d_deconv = nullptr;
  long long cap = 0;
  double* d_pos = nullptr;
  double* d_mass = nullptr;
  double* d_out = nullptr;
  long long idx_cap = 0;
  long long* d_idx = nullptr;
  hipStream_t stream = nullptr;
  std::mutex mu;
};

namespace {

void free64(gng_pm64_plan* p) {
  for (double2* f : p->d_f) (void)hipFree(f);
  (void)hipFree(p->d_tw);
  (void)hipFree(p->d_deconv);
  (void)hipFree(p->d_pos);
  (void)hipFree(p->d_mass);
  (void)hipFree(p->d_out);
  (void)hipFree(p->d_idx);
  if (p->stream) (void)hipStreamDestroy(p->stream);
}

int run64(gng_pm64_plan* p, const double* pos, const double* mass, long long n, const long long* idx, long long k,
          double g, double* out, hipStream_t s) {
  const int nm = p->cfg.grid_size;
  const double L = p->cfg.box_size;
  const long long M = p->M;
  const unsigned mb = (unsigned)((M + 255) / 256);
  for (double2* f : p->d_f) hipLaunchKernelGGL(k_zero, dim3(mb), dim3(256), 0, s, f, M);
  const double cs = nm / L;
  const double cell = L / nm;
  if (n > 0)
    hipLaunchKernelGGL(k_deposit, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pos, mass, n, nm, cs,
                       1.0 / (cell * cell * cell), p->d_f[0]);
  for (int ax = 0; ax < 3; ++ax) p->ops->fft(p->d_f[0], p->d_tw, ax, false, s);
  Spectral sp;
  sp.nm = nm;
  sp.two_pi_over_l = 2.0 * M_PI / L;
  sp.g4pi = -4.0 * M_PI * g;
  sp.rs2 = p->cfg.r_split > 0.0 ? p->cfg.r_split * p->cfg.r_split : 0.0;
  sp.eps2 = p->cfg.plummer_eps > 0.0 ? p->cfg.plummer_eps * p->cfg.plummer_eps : 0.0;
  sp.deconv = p->cfg.deconvolve ? p->d_deconv : nullptr;
  hipLaunchKernelGGL(k_spectral, dim3(mb), dim3(256), 0, s, p->d_f[0], p->d_f[1], p->d_f[2], sp);
  for (int f = 0; f < 3; ++f)
    for (int ax = 0; ax < 3; ++ax) p->ops->fft(p->d_f[f], p->d_tw, ax, true, s);
  if (k > 0)
    hipLaunchKernelGGL(k_interp, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, s, pos, idx, k, nm, cs,
                       1.0 / ((double)nm * nm * nm), p->d_f[0], p->d_f[1], p->d_f[2], out);
  GNG_HIP(hipGetLastError());
  return GNG_OK;
}

}  // namespace

extern "C" {

int gng_pm64_create(const gng_pm_config* cfg, gng_pm64_plan** out) {
  if (!cfg || !out) return fail(GNG_ERR_INVALID, "null argument");
  *out = nullptr;
  const Ops* ops = find_ops(cfg->grid_size);
  if (!ops) return fail(GNG_ERR_UNSUPPORTED, "FP64 form: mesh side must be a power of two in [4, 1024]");
  if (!(cfg->box_size > 0.0)) return fail(GNG_ERR_INVALID, "box_size must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(GNG_ERR_NO_DEVICE, "no HIP device");
  int dev = cfg->device;
  if (dev < 0) GNG_HIP(hipGetDevice(&dev));
  if (dev >= ndev) return fail(GNG_ERR_INVALID, "device index out of range");
  GNG_HIP(hipSetDevice(dev));
  gng_pm64_plan* p = new (std::nothrow) gng_pm64_plan();
  if (!p) return fail(GNG_ERR_OOM, "host allocation failed");
  p->cfg = *cfg;
  p->ops = ops;
  p->device = dev;
  const int nm = cfg->grid_size;
  p->M = (long long)nm * nm * nm;
  hipError_t e = hipSuccess;
  for (int f = 0; f < 3 && e == hipSuccess; ++f) e = hipMalloc(&p->d_f[f], sizeof(double2) * (size_t)p->M);
  if (e == hipSuccess) e = hipMalloc(&p->d_tw, sizeof(double2) * (size_t)(nm / 2));
  if (e == hipSuccess) e = hipMalloc(&p->d_deconv, sizeof(double) * (size_t)nm);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
  if (e == hipSuccess) {
    std::vector<double2> tw(nm / 2);
    for (int j = 0; j < nm / 2; ++j) {
      const long double a = -2.0L * 3.141592653589793238462643383279502884L * j / nm;
      tw[j] = make_double2((double)cosl(a), (double)sinl(a));
    }
    std::vector<double> dc(nm);
    for (int i = 0; i < nm; ++i) {
      const int n = i <= nm / 2 ? i : i - nm;
      const long double x = 3.141592653589793238462643383279502884L * n / nm;
      const long double s = n == 0 ? 1.0L : sinl(x) / x;
      dc[i] = (double)(1.0L / (s * s * s * s));
    }
    e = hipMemcpy(p->d_tw, tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_deconv, dc.data(), sizeof(double) * dc.size(), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    free64(p);
    delete p;
    return fail(e == hipErrorOutOfMemory ? GNG_ERR_OOM : GNG_ERR_HIP,
                std::string("FP64 form buffers: ") + hipGetErrorString(e));
  }
  *out = p;
  return GNG_OK;
}

void gng_pm64_destroy(gng_pm64_plan* p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  (void)hipDeviceSynchronize();
  free64(p);
  delete p;
}

int gng_pm64_accelerations(gng_pm64_plan* p, const double* pos, const double* mass, int64_t n, const int64_t* idx,
                           int64_t k, double g, double* out) {
  if (!p) return fail(GNG_ERR_INVALID, "null plan");
  if (n < 0 || k < 0) return fail(GNG_ERR_INVALID, "negative particle count");
  if (!idx && k != n) return fail(GNG_ERR_INVALID, "idx is NULL but k != n");
  if ((n > 0 && (!pos || !mass)) || (k > 0 && !out)) return fail(GNG_ERR_INVALID, "null array");
  std::lock_guard<std::mutex> lk(p->mu);
  GNG_HIP(hipSetDevice(p->device));
  const long long need = n > k ? n : k;
  if (need > p->cap) {
    (void)hipFree(p->d_pos);
    (void)hipFree(p->d_mass);
    (void)hipFree(p->d_out);
    p->d_pos = p->d_mass = p->d_out = nullptr;
    p->cap = 0;
    GNG_HIP(hipMalloc(&p->d_pos, sizeof(double) * 3 * (size_t)need));
    GNG_HIP(hipMalloc(&p->d_mass, sizeof(double) * (size_t)need));
    GNG_HIP(hipMalloc(&p->d_out, sizeof(double) * 3 * (size_t)need));
    p->cap = need;
  }
  if (idx && k > p->idx_cap) {
    (void)hipFree(p->d_idx);
    p->d_idx = nullptr;
    p->idx_cap = 0;
    GNG_HIP(hipMalloc(&p->d_idx, sizeof(long long) * (size_t)k));
    p->idx_cap = k;
  }
  if (idx)
    for (int64_t t = 0; t < k; ++t)
      if (idx[t] < 0 || idx[t] >= n) return fail(GNG_ERR_INVALID, "index out of range");
  hipStream_t s = p->stream;
  if (n > 0) {
    GNG_HIP(hipMemcpyAsync(p->d_pos, pos, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, s));
    GNG_HIP(hipMemcpyAsync(p->d_mass, mass, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, s));
  }
  if (idx && k > 0) GNG_HIP(hipMemcpyAsync(p->d_idx, idx, sizeof(long long) * (size_t)k, hipMemcpyHostToDevice, s));
  int rc = run64(p, p->d_pos, p->d_mass, n, idx ? p->d_idx : nullptr, k, g, p->d_out, s);
  if (rc) return rc;
  if (k > 0) GNG_HIP(hipMemcpyAsync(out, p->d_out, sizeof(double) * 3 * (size_t)k, hipMemcpyDeviceToHost, s));
  GNG_HIP(hipStreamSynchronize(s));
  return GNG_OK;
}

int gng_pm64_solve_device(gng_pm64_plan* p, const double* d_pos, const double* d_mass, int64_t n,
                          const int64_t* d_idx, int64_t k, double g, double* d_out, void* stream) {
  if (!p) return fail(GNG_ERR_INVALID, "null plan");
  if (n < 0 || k < 0) return fail(GNG_ERR_INVALID, "negative particle count");
  if (!d_idx && k != n) return fail(GNG_ERR_INVALID, "idx is NULL but k != n");
  std::lock_guard<std::mutex> lk(p->mu);
  GNG_HIP(hipSetDevice(p->device));
  return run64(p, d_pos, d_mass, n, reinterpret_cast<const long long*>(d_idx), k, g, d_out,
               reinterpret_cast<hipStream_t>(stream));
}

}  // extern "C"
