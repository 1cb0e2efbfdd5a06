// integrate.hip -- device-resident KDK leapfrog over the PM solver, and the
// host-side cosmological drift/kick factors.
//
// Reference: crates/gadget-ng-integrators/src/leapfrog.rs:5-95 (the step),
// crates/gadget-ng-core/src/cosmology.rs:141-221 (a(t), drift/kick factors),
// :354-356 (G a^3 coupling), :759-778 (periodic wrap), and the periodic PM
// loop of crates/gadget-ng-cli/src/engine/stepping/mod.rs:1076-1576.
//
// The state lives in HBM as SoA f64 (x, y, z, vx, vy, vz), f32 masses and the
// f32 accelerations of the last force call.  The PM force reads the f64
// positions (cells formed in FP64); an f32 copy of the positions feeds the
// optional short-range sum.  Each step is: PM force, kick+drift
// kernel, PM force, kick(+wrap) kernel -- all on one stream, no host sync.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/gng_integrate.h"
#include "../../include/gng_learned.h"
#include "status.hpp"

using gng::fail;

namespace {

// ---------------------------------------------------------------------------
// Kernels.  Contraction is off so every product and sum rounds separately,
// as the reference's Vec3 arithmetic does (v + a*k, x + v*d).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_kick_drift(double* __restrict__ x, double* __restrict__ y,
                                                    double* __restrict__ z, double* __restrict__ vx,
                                                    double* __restrict__ vy, double* __restrict__ vz,
                                                    const float* __restrict__ ax, const float* __restrict__ ay,
                                                    const float* __restrict__ az, long long n, double kick,
                                                    double drift, float* __restrict__ xf, float* __restrict__ yf,
                                                    float* __restrict__ zf) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ux = vx[i] + (double)ax[i] * kick;
  const double uy = vy[i] + (double)ay[i] * kick;
  const double uz = vz[i] + (double)az[i] * kick;
  vx[i] = ux;
  vy[i] = uy;
  vz[i] = uz;
  const double px = x[i] + ux * drift;
  const double py = y[i] + uy * drift;
  const double pz = z[i] + uz * drift;
  x[i] = px;
  y[i] = py;
  z[i] = pz;
  xf[i] = (float)px;
  yf[i] = (float)py;
  zf[i] = (float)pz;
}

// rem_euclid of f64 (Rust: r = x % l; r < 0 ? r + |l| : r); fmod is exact
__device__ __forceinline__ double rem_euclid(double x, double l) {
#pragma clang fp contract(off)
  const double r = fmod(x, l);
  return r < 0.0 ? r + fabs(l) : r;
}

__global__ __launch_bounds__(256) void k_kick_wrap(double* __restrict__ x, double* __restrict__ y,
                                                   double* __restrict__ z, double* __restrict__ vx,
                                                   double* __restrict__ vy, double* __restrict__ vz,
                                                   const float* __restrict__ ax, const float* __restrict__ ay,
                                                   const float* __restrict__ az, long long n, double kick,
                                                   int wrap, double box, float* __restrict__ xf,
                                                   float* __restrict__ yf, float* __restrict__ zf) {
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  vx[i] = vx[i] + (double)ax[i] * kick;
  vy[i] = vy[i] + (double)ay[i] * kick;
  vz[i] = vz[i] + (double)az[i] * kick;
  if (wrap) {
    const double px = rem_euclid(x[i], box), py = rem_euclid(y[i], box), pz = rem_euclid(z[i], box);
    x[i] = px;
    y[i] = py;
    z[i] = pz;
    xf[i] = (float)px;
    yf[i] = (float)py;
    zf[i] = (float)pz;
  }
}

// Device-side state load: n x 3 interleaved f64 (Vec3) positions and
// velocities into the SoA state, the f32 copies the force reads, masses
// (per particle, or one value when m is null) and zeroed accelerations.
__global__ __launch_bounds__(256) void k_load_state(const double* __restrict__ pos, const double* __restrict__ vel,
                                                    const double* __restrict__ m, double m_all, long long n,
                                                    double* __restrict__ soa, float* __restrict__ f) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double p = pos[3 * i + c];
    soa[c * n + i] = p;
    soa[(3 + c) * n + i] = vel[3 * i + c];
    f[(1 + c) * n + i] = (float)p;
    f[(4 + c) * n + i] = 0.0f;
  }
  f[i] = (float)(m ? m[i] : m_all);
}

__global__ __launch_bounds__(256) void k_add3(float* __restrict__ ax, float* __restrict__ ay, float* __restrict__ az,
                                              const float* __restrict__ bx, const float* __restrict__ by,
                                              const float* __restrict__ bz, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ax[i] += bx[i];
  ay[i] += by[i];
  az[i] += bz[i];
}

// ---------------------------------------------------------------------------
// Cosmology (host).  Same operations in the same order as the reference, so
// the factors agree to the last bit with the same libm.
// ---------------------------------------------------------------------------

double da_dt(const gng_cosmology& c, double a) {
  double omega_de;
  if (std::fabs(c.w0 + 1.0) < 1e-10 && std::fabs(c.wa) < 1e-10) {
    omega_de = c.omega_lambda;
  } else {
    const double exp_factor = std::exp(3.0 * c.wa * (a - 1.0));
    const double power = -3.0 * (1.0 + c.w0 + c.wa);
    omega_de = c.omega_lambda * std::pow(a, power) * exp_factor;
  }
  const double h_sq = (c.omega_m + c.omega_nu) / (a * a * a) + omega_de;
  return a * c.h0 * std::sqrt(h_sq > 0.0 ? h_sq : 0.0);
}

double advance_a(const gng_cosmology& c, double a, double dt) {
  const double k1 = da_dt(c, a);
  const double k2 = da_dt(c, a + 0.5 * dt * k1);
  const double k3 = da_dt(c, a + 0.5 * dt * k2);
  const double k4 = da_dt(c, a + dt * k3);
  return a + dt / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4);
}

// H(a) (cosmology.rs:509-520)
double hubble_param(const gng_cosmology& c, double a) {
  double omega_de;
  if (std::fabs(c.w0 + 1.0) < 1e-10 && std::fabs(c.wa) < 1e-10) {
    omega_de = c.omega_lambda;
  } else {
    const double exp_factor = std::exp(3.0 * c.wa * (a - 1.0));
    const double power = -3.0 * (1.0 + c.w0 + c.wa);
    omega_de = c.omega_lambda * std::pow(a, power) * exp_factor;
  }
  const double h_sq = (c.omega_m + c.omega_nu) / (a * a * a) + omega_de;
  return c.h0 * std::sqrt(h_sq > 0.0 ? h_sq : 0.0);
}

// f(a) = Omega_m(a)^0.55 (cosmology.rs:598-610)
double growth_rate_f(const gng_cosmology& c, double a) {
  if (a <= 0.0) return 1.0;
  const double h = hubble_param(c, a);
  if (h <= 0.0 || c.h0 <= 0.0) return 1.0;
  const double h_ratio_sq = (h / c.h0) * (h / c.h0);
  const double omega_m_a = c.omega_m / (a * a * a) / h_ratio_sq;
  return std::pow(omega_m_a > 0.0 ? omega_m_a : 0.0, 0.55);
}

// Carroll-Press-Turner 1992 growing mode (cosmology.rs:638-654)
double growth_cpt92(const gng_cosmology& c, double a) {
  const double a3 = a * a * a;
  const double denom = c.omega_m + c.omega_lambda * a3;
  if (denom <= 0.0) return a;
  const double om_a = c.omega_m / denom;
  const double ol_a = c.omega_lambda * a3 / denom;
  const double bracket =
      std::pow(om_a > 0.0 ? om_a : 0.0, 4.0 / 7.0) - ol_a + (1.0 + om_a / 2.0) * (1.0 + ol_a / 70.0);
  if (bracket <= 0.0) return a;
  return a * (2.5 * om_a / bracket);
}

// (cosmology.rs:656-666)
double omega_matter_fraction(const gng_cosmology& c, double a) {
  const double h = hubble_param(c, a);
  if (h <= 0.0 || c.h0 <= 0.0) return 0.0;
  const double h_ratio_sq = (h / c.h0) * (h / c.h0);
  const double rho_m = (c.omega_m + c.omega_nu) / (a * a * a);
  return rho_m / (h_ratio_sq > 1e-300 ? h_ratio_sq : 1e-300);
}

// (cosmology.rs:668-678)
double d_ln_h_d_ln_a(const gng_cosmology& c, double a) {
  const double eps = std::max(1e-7, a * 1e-9);
  const double hp = hubble_param(c, std::min(a + eps, 1e9));
  const double hm = hubble_param(c, std::max(a - eps, 1e-18));
  const double h = hubble_param(c, a);
  if (h <= 0.0) return 0.0;
  return (a / h) * (hp - hm) / (2.0 * eps);
}

// RK4 on D'' + (2 + d ln H / d ln a) D' - 3/2 Omega_m(a) D = 0 in ln a, 768
// steps from a = 1e-5 with D = D' = 1e-5 (cosmology.rs:680-711)
double growth_ode(const gng_cosmology& c, double a) {
  constexpr double kStart = 1e-5;
  constexpr int kSteps = 768;
  if (a <= kStart) return a > 0.0 ? a : 0.0;
  const double ln_end = std::log(a), ln_start = std::log(kStart);
  const double h = (ln_end - ln_start) / (double)kSteps;
  double ln_a = ln_start, d = kStart, dp = kStart;
  auto deriv = [&](double la, double y, double yp, double& o0, double& o1) {
    const double aa = std::exp(la);
    const double om = omega_matter_fraction(c, aa);
    const double q = d_ln_h_d_ln_a(c, aa);
    o0 = yp;
    o1 = -(2.0 + q) * yp + 1.5 * om * y;
  };
  for (int i = 0; i < kSteps; ++i) {
    double k10, k11, k20, k21, k30, k31, k40, k41;
    deriv(ln_a, d, dp, k10, k11);
    deriv(ln_a + 0.5 * h, d + 0.5 * h * k10, dp + 0.5 * h * k11, k20, k21);
    deriv(ln_a + 0.5 * h, d + 0.5 * h * k20, dp + 0.5 * h * k21, k30, k31);
    deriv(ln_a + h, d + h * k30, dp + h * k31, k40, k41);
    d += h / 6.0 * (k10 + 2.0 * k20 + 2.0 * k30 + k40);
    dp += h / 6.0 * (k11 + 2.0 * k21 + 2.0 * k31 + k41);
    ln_a += h;
  }
  return d;
}

// D(a), not normalised (cosmology.rs:715-724)
double growth_factor_d(const gng_cosmology& c, double a) {
  if (a <= 0.0) return 0.0;
  if (std::fabs(c.w0 + 1.0) < 1e-10 && std::fabs(c.wa) < 1e-10) return growth_cpt92(c, a);
  return growth_ode(c, a);
}

void drift_kick_factors(const gng_cosmology& c, double a0, double dt, double out[3]) {
  constexpr int kSub = 16;
  const double sub_dt = dt / (double)kSub;
  double a = a0, drift = 0.0, kick_half = 0.0, kick_half2 = 0.0;
  for (int sub = 0; sub < kSub; ++sub) {
    const double a_left = a;
    const double a_mid = advance_a(c, a, sub_dt * 0.5);
    const double a_right = advance_a(c, a, sub_dt);
    drift += sub_dt / 6.0 * (1.0 / (a_left * a_left) + 4.0 / (a_mid * a_mid) + 1.0 / (a_right * a_right));
    const double k_sub = sub_dt / 6.0 * (1.0 / a_left + 4.0 / a_mid + 1.0 / a_right);
    if (sub < kSub / 2)
      kick_half += k_sub;
    else
      kick_half2 += k_sub;
    a = a_right;
  }
  out[0] = drift;
  out[1] = kick_half;
  out[2] = kick_half2;
}

}  // namespace

struct gng_kdk {
  gng_pm_plan* plan = nullptr;
  int device = 0;
  long long n = 0;
  double box = 0.0;
  // optional short-range half (TreePM): force = PM long range + short range
  gng_sr* sr = nullptr;
  double sr_r_split = 0.0, sr_r_cut = 0.0, sr_eps2 = 0.0;
  int sr_kernel = 0;
  // state
  double* d_pos = nullptr;  // x, y, z, vx, vy, vz: 6 n doubles
  float* d_f = nullptr;     // m, xf, yf, zf, ax, ay, az (+ 3 n short-range scratch): 7 or 10 n floats
  bool have_state = false;
  std::mutex mu;

  double* x() { return d_pos; }
  double* y() { return d_pos + n; }
  double* z() { return d_pos + 2 * n; }
  double* vx() { return d_pos + 3 * n; }
  double* vy() { return d_pos + 4 * n; }
  double* vz() { return d_pos + 5 * n; }
  float* m() { return d_f; }
  float* xf() { return d_f + n; }
  float* yf() { return d_f + 2 * n; }
  float* zf() { return d_f + 3 * n; }
  float* ax() { return d_f + 4 * n; }
  float* ay() { return d_f + 5 * n; }
  float* az() { return d_f + 6 * n; }
  float* sx() { return d_f + 7 * n; }
};

namespace {

void free_kdk(gng_kdk* s) {
  (void)hipFree(s->d_pos);
  (void)hipFree(s->d_f);
  s->d_pos = nullptr;
  s->d_f = nullptr;
}

// a = F(x): PM on the f64 positions (long range when the plan has r_split)
// plus the optional short range on their f32 copy
int force(gng_kdk* s, double g, hipStream_t st) {
  int rc = gng_pm_solve_device_soa64(s->plan, s->x(), s->y(), s->z(), s->m(), s->n, nullptr, s->n, g, s->ax(),
                                     s->ay(), s->az(), st);
  if (rc) return rc;
  if (s->sr) {
    rc = gng_sr_accelerations(s->sr, s->xf(), s->yf(), s->zf(), s->m(), s->n, s->box, s->sr_r_split, s->sr_r_cut,
                              s->sr_eps2, g, s->sr_kernel, s->sx(), s->sx() + s->n, s->sx() + 2 * s->n, st);
    if (rc) return rc;
    GNG_HIP(hipSetDevice(s->device));
    hipLaunchKernelGGL(k_add3, dim3((unsigned)((s->n + 255) / 256)), dim3(256), 0, st, s->ax(), s->ay(), s->az(),
                       s->sx(), s->sx() + s->n, s->sx() + 2 * s->n, s->n);
    GNG_HIP(hipGetLastError());
  }
  return GNG_OK;
}

int step_locked(gng_kdk* s, double drift, double kick_half, double kick_half2, double g, int wrap, hipStream_t st) {
  if (!std::isfinite(drift) || !std::isfinite(kick_half) || !std::isfinite(kick_half2) || !std::isfinite(g))
    return fail(GNG_ERR_INVALID, "non-finite step factor or coupling");
  if (s->n == 0) return GNG_OK;
  const dim3 grid((unsigned)((s->n + 255) / 256)), block(256);
  int rc = force(s, g, st);
  if (rc) return rc;
  GNG_HIP(hipSetDevice(s->device));
  hipLaunchKernelGGL(k_kick_drift, grid, block, 0, st, s->x(), s->y(), s->z(), s->vx(), s->vy(), s->vz(), s->ax(),
                     s->ay(), s->az(), s->n, kick_half, drift, s->xf(), s->yf(), s->zf());
  GNG_HIP(hipGetLastError());
  if ((rc = force(s, g, st))) return rc;
  GNG_HIP(hipSetDevice(s->device));
  hipLaunchKernelGGL(k_kick_wrap, grid, block, 0, st, s->x(), s->y(), s->z(), s->vx(), s->vy(), s->vz(), s->ax(),
                     s->ay(), s->az(), s->n, kick_half2, wrap, s->box, s->xf(), s->yf(), s->zf());
  GNG_HIP(hipGetLastError());
  return GNG_OK;
}

}  // namespace

extern "C" {

double gng_cosmo_advance_a(const gng_cosmology* c, double a, double dt) {
  if (!c) return std::nan("");
  return advance_a(*c, a, dt);
}

double gng_cosmo_hubble(const gng_cosmology* c, double a) {
  if (!c) return std::nan("");
  return hubble_param(*c, a);
}

double gng_cosmo_growth_rate(const gng_cosmology* c, double a) {
  if (!c) return std::nan("");
  return growth_rate_f(*c, a);
}

double gng_cosmo_growth_factor(const gng_cosmology* c, double a) {
  if (!c) return std::nan("");
  return growth_factor_d(*c, a);
}

int gng_cosmo_drift_kick_factors(const gng_cosmology* c, double a0, double dt, double* out) {
  if (!c || !out) return fail(GNG_ERR_INVALID, "null argument");
  drift_kick_factors(*c, a0, dt, out);
  return GNG_OK;
}

int gng_kdk_create(gng_pm_plan* plan, int64_t n, gng_kdk** out) {
  if (!out) return fail(GNG_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (!plan) return fail(GNG_ERR_INVALID, "null plan");
  if (n < 0) return fail(GNG_ERR_INVALID, "negative particle count");
  if (n >= (1LL << 31)) return fail(GNG_ERR_INVALID, "n of 2^31 or more is not supported");
  gng_pm_config cfg;
  int rc = gng_pm_get_config(plan, &cfg);
  if (rc) return rc;
  gng_kdk* s = new (std::nothrow) gng_kdk();
  if (!s) return fail(GNG_ERR_OOM, "host allocation failed");
  s->plan = plan;
  s->device = cfg.device;
  s->n = n;
  s->box = cfg.box_size;
  if (hipSetDevice(s->device) != hipSuccess) {
    delete s;
    return fail(GNG_ERR_HIP, "hipSetDevice failed");
  }
  const size_t cnt = n > 0 ? (size_t)n : 1;
  if (hipMalloc(&s->d_pos, sizeof(double) * 6 * cnt) != hipSuccess ||
      hipMalloc(&s->d_f, sizeof(float) * 7 * cnt) != hipSuccess ||
      hipMemset(s->d_f, 0, sizeof(float) * 7 * cnt) != hipSuccess) {
    free_kdk(s);
    delete s;
    return fail(GNG_ERR_OOM, "integrator state allocation failed");
  }
  *out = s;
  return GNG_OK;
}

void gng_kdk_destroy(gng_kdk* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  (void)hipDeviceSynchronize();
  free_kdk(s);
  delete s;
}

int gng_kdk_set_short_range(gng_kdk* s, gng_sr* sr, double r_split, double r_cut, double eps2, int32_t kernel) {
  if (!s) return fail(GNG_ERR_INVALID, "null integrator");
  std::lock_guard<std::mutex> lk(s->mu);
  if (sr && !(r_split > 0.0 && r_cut > 0.0)) return fail(GNG_ERR_INVALID, "short range needs r_split > 0 and r_cut > 0");
  if (sr && s->sr == nullptr && s->n > 0) {
    // room for the short-range accelerations next to the state
    GNG_HIP(hipSetDevice(s->device));
    GNG_HIP(hipDeviceSynchronize());
    float* f = nullptr;
    GNG_HIP(hipMalloc(&f, sizeof(float) * 10 * (size_t)s->n));
    GNG_HIP(hipMemcpy(f, s->d_f, sizeof(float) * 7 * (size_t)s->n, hipMemcpyDeviceToDevice));
    (void)hipFree(s->d_f);
    s->d_f = f;
  }
  s->sr = sr;
  s->sr_r_split = r_split;
  s->sr_r_cut = r_cut;
  s->sr_eps2 = eps2;
  s->sr_kernel = kernel;
  return GNG_OK;
}

int gng_kdk_set_state(gng_kdk* s, const double* pos, const double* vel, const double* mass) {
  if (!s) return fail(GNG_ERR_INVALID, "null integrator");
  if (s->n > 0 && (!pos || !vel || !mass)) return fail(GNG_ERR_INVALID, "null state array");
  std::lock_guard<std::mutex> lk(s->mu);
  GNG_HIP(hipSetDevice(s->device));
  const size_t n = (size_t)s->n;
  if (n == 0) {
    s->have_state = true;
    return GNG_OK;
  }
  std::vector<double> soa(6 * n);
  std::vector<float> f(4 * n);
  for (size_t i = 0; i < n; ++i) {
    for (int c = 0; c < 3; ++c) {
      soa[c * n + i] = pos[3 * i + c];
      soa[(3 + c) * n + i] = vel[3 * i + c];
      f[(1 + c) * n + i] = (float)pos[3 * i + c];
    }
    f[i] = (float)mass[i];
  }
  GNG_HIP(hipDeviceSynchronize());
  GNG_HIP(hipMemcpy(s->d_pos, soa.data(), sizeof(double) * 6 * n, hipMemcpyHostToDevice));
  GNG_HIP(hipMemcpy(s->d_f, f.data(), sizeof(float) * 4 * n, hipMemcpyHostToDevice));
  GNG_HIP(hipMemset(s->ax(), 0, sizeof(float) * 3 * n));
  s->have_state = true;
  return GNG_OK;
}

int gng_kdk_set_state_device(gng_kdk* s, const double* d_pos, const double* d_vel, const double* d_mass,
                             double uniform_mass, void* stream) {
  if (!s) return fail(GNG_ERR_INVALID, "null integrator");
  if (s->n > 0 && (!d_pos || !d_vel)) return fail(GNG_ERR_INVALID, "null state array");
  if (!d_mass && !(std::isfinite(uniform_mass) && uniform_mass > 0.0))
    return fail(GNG_ERR_INVALID, "without d_mass, uniform_mass must be finite and > 0");
  std::lock_guard<std::mutex> lk(s->mu);
  GNG_HIP(hipSetDevice(s->device));
  const long long n = s->n;
  if (n > 0)
    hipLaunchKernelGGL(k_load_state, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), d_pos, d_vel, d_mass, uniform_mass, n, s->d_pos,
                       s->d_f);
  GNG_HIP(hipGetLastError());
  s->have_state = true;
  return GNG_OK;
}

int gng_kdk_get_state(gng_kdk* s, double* pos, double* vel, double* acc) {
  if (!s) return fail(GNG_ERR_INVALID, "null integrator");
  std::lock_guard<std::mutex> lk(s->mu);
  GNG_HIP(hipSetDevice(s->device));
  GNG_HIP(hipDeviceSynchronize());
  const size_t n = (size_t)s->n;
  if (n == 0) return GNG_OK;
  if (pos || vel) {
    std::vector<double> soa(6 * n);
    GNG_HIP(hipMemcpy(soa.data(), s->d_pos, sizeof(double) * 6 * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) {
        if (pos) pos[3 * i + c] = soa[c * n + i];
        if (vel) vel[3 * i + c] = soa[(3 + c) * n + i];
      }
  }
  if (acc) {
    std::vector<float> a(3 * n);
    GNG_HIP(hipMemcpy(a.data(), s->ax(), sizeof(float) * 3 * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) acc[3 * i + c] = (double)a[c * n + i];
  }
  return GNG_OK;
}

int gng_kdk_step(gng_kdk* s, double drift, double kick_half, double kick_half2, double g, int wrap, void* stream) {
  if (!s) return fail(GNG_ERR_INVALID, "null integrator");
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->have_state) return fail(GNG_ERR_INVALID, "no state: call gng_kdk_set_state first");
  return step_locked(s, drift, kick_half, kick_half2, g, wrap, reinterpret_cast<hipStream_t>(stream));
}

int gng_kdk_run(gng_kdk* s, const gng_cosmology* c, double* a, double dt, double g, int nsteps, int wrap,
                void* stream) {
  if (!s) return fail(GNG_ERR_INVALID, "null integrator");
  if (nsteps < 0) return fail(GNG_ERR_INVALID, "negative step count");
  if (c && !a) return fail(GNG_ERR_INVALID, "a cosmological run needs the scale factor");
  if (c && !(*a > 0.0)) return fail(GNG_ERR_INVALID, "the scale factor must be positive");
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->have_state) return fail(GNG_ERR_INVALID, "no state: call gng_kdk_set_state first");
  const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  for (int k = 0; k < nsteps; ++k) {
    double f[3] = {dt, 0.5 * dt, 0.5 * dt};
    double gk = g;
    if (c) {
      gk = g * (*a) * (*a) * (*a);
      drift_kick_factors(*c, *a, dt, f);
      *a = advance_a(*c, *a, dt);
    }
    const int rc = step_locked(s, f[0], f[1], f[2], gk, wrap, st);
    if (rc) return rc;
  }
  return GNG_OK;
}

}  // extern "C"
