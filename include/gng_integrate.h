/*
 * gng_integrate.h -- device-resident KDK leapfrog over the PM solver, and the
 * cosmological drift/kick factors it takes.
 *
 * The reference integrates on the host: each force call copies f64 positions
 * to f32, pushes them over PCIe, solves, and copies the accelerations back
 * (crates/gadget-ng-hip/src/pm_solver.rs:171-214).  Here the particle state
 * (f64 positions and velocities, as in gadget_ng_core::Particle) stays in HBM
 * between force calls; only the step's scalars cross PCIe.
 *
 * Step semantics (crates/gadget-ng-integrators/src/leapfrog.rs:76-95,
 * leapfrog_cosmo_kdk_step; leapfrog_kdk_step at :5-23 is the flat case):
 *   a  = F(x)                 (PM with coupling g)
 *   v += a * kick_half
 *   x += v * drift
 *   a  = F(x)
 *   v += a * kick_half2
 * and, with wrap != 0, x = rem_euclid(x, L) per component afterwards, as the
 * periodic cosmological engine does after each step
 * (crates/gadget-ng-cli/src/engine/stepping/mod.rs:1572-1576,
 * crates/gadget-ng-core/src/cosmology.rs:759-778).  The PM force reads the
 * f64 positions (gng_pm_solve_device_soa64: cells formed in FP64, as the
 * reference's CPU PM does); an attached short-range sum sees them rounded to
 * f32.  The kick and drift are f64 with separately rounded products and sums
 * (no fused multiply-add), so given the same f32 accelerations a step
 * reproduces the reference's f64 update bit for bit.
 */
#ifndef GNG_INTEGRATE_H
#define GNG_INTEGRATE_H

#include <stdint.h>

#include "gng_learned.h"
#include "gng_pm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Background cosmology: mirrors CosmologyParams
 * (crates/gadget-ng-core/src/cosmology.rs:37-44).  h0 is in internal time
 * units; w0 = -1, wa = 0 is LambdaCDM. */
typedef struct gng_cosmology {
  double omega_m, omega_lambda, h0, w0, wa, omega_nu;
} gng_cosmology;

/* a(t + dt): one RK4 step of da/dt = a h0 sqrt((omega_m + omega_nu)/a^3 +
 * Omega_DE(a)) (cosmology.rs:141-158 da_dt, :160-167 advance_a).  Host-only
 * arithmetic; returns NaN for a NULL c. */
double gng_cosmo_advance_a(const gng_cosmology *c, double a, double dt);

/* out[0] = drift = int_t^{t+dt} dt'/a^2, out[1] = kick_half =
 * int_t^{t+dt/2} dt'/a, out[2] = kick_half2 = int_{t+dt/2}^{t+dt} dt'/a, by
 * 16 Simpson sub-steps along advance_a (cosmology.rs:188-221,
 * drift_kick_factors).  Host-only arithmetic. */
int gng_cosmo_drift_kick_factors(const gng_cosmology *c, double a0, double dt, double *out);

/* H(a) = h0 sqrt(max((omega_m + omega_nu)/a^3 + Omega_DE(a), 0))
 * (cosmology.rs:509-520).  Host-only; NaN for a NULL c. */
double gng_cosmo_hubble(const gng_cosmology *c, double a);

/* Linear growth rate f(a) = Omega_m(a)^0.55 (Linder 2005), 1 for a <= 0
 * (cosmology.rs:598-610).  Host-only; NaN for a NULL c. */
double gng_cosmo_growth_rate(const gng_cosmology *c, double a);

/* Linear growth factor D(a), not normalised: Carroll-Press-Turner 1992 for
 * w0 = -1, wa = 0, otherwise RK4 on the growth equation in ln a, 768 steps
 * from a = 1e-5 (cosmology.rs:638-724).  0 for a <= 0.  Host-only; NaN for a
 * NULL c.  Zel'dovich momenta use a^2 f(a) H(a) and the optional rescaling
 * D(a_init) / D(1) (ic_zeldovich.rs:685-703). */
double gng_cosmo_growth_factor(const gng_cosmology *c, double a);

/* Particle state of one simulation on the plan's device. */
typedef struct gng_kdk gng_kdk;

/* State for n particles integrated with `plan` (borrowed: it must outlive the
 * state; the plan's box size is the wrap period).  Allocates 76 bytes per
 * particle of device memory. */
int gng_kdk_create(gng_pm_plan *plan, int64_t n, gng_kdk **out);
void gng_kdk_destroy(gng_kdk *s);

/* Adds the TreePM short range (gng_sr_accelerations with these arguments) to
 * every force evaluation, so that with a plan created with the same r_split
 * the force is TreePM's long + short range
 * (crates/gadget-ng-treepm/src/solver.rs:65-107).  sr NULL removes it; sr is
 * borrowed.  Grows the state by 12 bytes per particle. */
int gng_kdk_set_short_range(gng_kdk *s, gng_sr *sr, double r_split, double r_cut, double eps2, int32_t kernel);

/* Host f64 arrays, n x 3 interleaved as Vec3 (pos, vel) and n masses.
 * Synchronous.  The masses enter the force as f32. */
int gng_kdk_set_state(gng_kdk *s, const double *pos, const double *vel, const double *mass);

/* Same from device memory (d_pos, d_vel: n x 3 interleaved f64; d_mass: n
 * f64, or NULL for uniform_mass on every particle, which must then be
 * finite and > 0), enqueued on `stream` (a hipStream_t; NULL = default).  No
 * host copy: initial conditions made with gng_ic_zeldovich_device /
 * gng_ic_2lpt_device go straight into the state.  The load is not ordered
 * against steps in flight on other streams: the caller orders it after them
 * (as after the initial-condition kernels that wrote d_pos and d_vel).  No
 * reference counterpart (the reference's state lives on the host). */
int gng_kdk_set_state_device(gng_kdk *s, const double *d_pos, const double *d_vel, const double *d_mass,
                             double uniform_mass, void *stream);

/* Any output may be NULL.  acc is the last force evaluated (the
 * Particle::acceleration the reference stores after a step); it is zero
 * before the first step.  Synchronises with the state's work. */
int gng_kdk_get_state(gng_kdk *s, double *pos, double *vel, double *acc);

/* One KDK step with the given factors and coupling g for both force
 * evaluations.  Enqueued on `stream` (a hipStream_t; NULL = default); does not
 * synchronise. */
int gng_kdk_step(gng_kdk *s, double drift, double kick_half, double kick_half2, double g, int wrap,
                 void *stream);

/* nsteps steps of length dt from scale factor *a, which is advanced.  With a
 * cosmology c: each step uses g_cosmo = g a^3 at the start of the step
 * (gravity_coupling_qksl, cosmology.rs:354-356), its factors from
 * drift_kick_factors(a, dt), then a = advance_a(a, dt)
 * (stepping/mod.rs:1076-1081, 1535-1545).  With c NULL: flat factors (dt,
 * dt/2, dt/2) and a fixed coupling g; *a is left unchanged (a may be NULL).
 * Enqueued on `stream`; does not synchronise. */
int gng_kdk_run(gng_kdk *s, const gng_cosmology *c, double *a, double dt, double g, int nsteps, int wrap,
                void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GNG_INTEGRATE_H */
