/* gng_ic.h -- Zel'dovich initial conditions on the GPU, on a PM plan's mesh.
 *
 * Replaces the CPU generator of crates/gadget-ng-core/src/ic_zeldovich.rs:
 * generate_delta_kspace (:287-369), delta_to_displacement (:541-603) and the
 * particle loop of zeldovich_ics (:621-766), and the second-order generator
 * of crates/gadget-ng-core/src/ic_2lpt.rs.  The plan's grid_size n is both
 * the lattice side (n^3 particles, global id ix n^2 + iy n + iz) and the
 * side of the Gaussian field, as in the reference; its box_size is L.  The
 * plan's meshes are used as work space (a later gng_pm_density or
 * gng_pm_potential has nothing to return until the next solve).
 *
 * The spectrum is passed as a table, sigma[n2] = spectrum_fn(sqrt(n2)) for
 * n2 = 0 .. 3 (n/2 - 1)^2 (at least that many entries): the reference calls
 * its spectrum closure with sqrt(n2) of each mode's integer |n|^2
 * (:315-316), so the table holds exactly the values the closure would give.
 *
 * Each mode draws its Gaussians from rand's StdRng seeded per mode
 * (:80-98): ChaCha12 keyed by rand_core's seed_from_u64.  The FP64 modes
 * (gng_ic_delta_kspace) match the reference's up to the last-ulp rounding
 * of log and cos; the displacements go through the plan's FP32 transforms.
 */
#ifndef GNG_IC_H
#define GNG_IC_H

#include <stdint.h>

#include "gng_pm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The Gaussian field delta-hat of generate_delta_kspace
 * (ic_zeldovich.rs:287-369) for the modes iz = 0 .. n/2 of every (ix, iy):
 * out[2 ((ix n + iy) (n/2 + 1) + iz) + {0, 1}] = (re, im), host memory of
 * 2 n^2 (n/2 + 1) doubles.  The modes iz > n/2 are the conjugates of these
 * (delta-hat(-k) = conj delta-hat(k)). */
int gng_ic_delta_kspace(gng_pm_plan *p, uint64_t seed, const double *sigma, int64_t sigma_len, double *out);

/* Particles lo .. hi - 1 of zeldovich_ics (ic_zeldovich.rs:733-766):
 *   Psi = s * (the displacement of delta_to_displacement),
 *   pos = rem_euclid(q + Psi, L) with q = ((ix, iy, iz) + 1/2) L / n,
 *   vel = Psi * vel_factor,
 * as (hi - lo) x 3 interleaved f64 (Vec3) in host memory.  vel_factor is
 * a^2 f(a) H(a) for the reference's canonical momenta (:688-690) and s the
 * optional growth rescaling D(a_init) / D(1) (1 for none, :699-703,
 * :720-730).  0 <= lo <= hi <= n^3. */
int gng_ic_zeldovich(gng_pm_plan *p, uint64_t seed, const double *sigma, int64_t sigma_len, double vel_factor,
                     double scale, int64_t lo, int64_t hi, double *pos, double *vel);

/* Same into device memory (d_pos, d_vel: (hi - lo) x 3 f64 each), enqueued on
 * `stream` (a hipStream_t; NULL = default).  sigma is host memory; it is
 * copied before this returns. */
int gng_ic_zeldovich_device(gng_pm_plan *p, uint64_t seed, const double *sigma, int64_t sigma_len,
                            double vel_factor, double scale, int64_t lo, int64_t hi, double *d_pos, double *d_vel,
                            void *stream);

/* Second-order (2LPT) particles lo .. hi - 1 of zeldovich_2lpt_ics
 * (crates/gadget-ng-core/src/ic_2lpt.rs:348-553):
 *   Psi1 = s * (the first-order displacement), Psi2 = s^2 * Re IFFT(-i n
 *   S-hat / |n|^2) d with S = sum over axis pairs of (phi,aa phi,bb -
 *   phi,ab^2) and phi,ab-hat = -n_a n_b delta-hat / |n|^2 (:86-244),
 *   pos = rem_euclid(q + Psi1 + d2_over_d1sq Psi2, L),
 *   vel = Psi1 vel1 + Psi2 vel2,
 * as (hi - lo) x 3 interleaved f64 in host memory.  vel1 = a^2 f1 H,
 * vel2 = a^2 f2 H d2_over_d1sq, d2_over_d1sq = -3/7 Omega_m(a)^(-1/143)
 * (:414-443), s the optional rescaling D(a_init) / D(1).  Needs one extra
 * n^3 f32 buffer while it runs. */
int gng_ic_2lpt(gng_pm_plan *p, uint64_t seed, const double *sigma, int64_t sigma_len, double vel1, double vel2,
                double d2_over_d1sq, double scale, int64_t lo, int64_t hi, double *pos, double *vel);

/* Same into device memory, enqueued on `stream`; returns after the work
 * completes (the source buffer is freed). */
int gng_ic_2lpt_device(gng_pm_plan *p, uint64_t seed, const double *sigma, int64_t sigma_len, double vel1,
                       double vel2, double d2_over_d1sq, double scale, int64_t lo, int64_t hi, double *d_pos,
                       double *d_vel, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GNG_IC_H */
