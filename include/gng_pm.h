/*
 * gng_pm.h -- C ABI of the MI355X-native periodic PM gravity solver.
 *
 * One shared library, libgng_pm.so, built for gfx950 only.  All pointers are
 * plain C pointers; no HIP or torch types appear in any signature (streams
 * are passed as void*).
 *
 * Two families of entry points:
 *
 *  1. hip_pm_*   -- binary drop-ins for the symbols the reference's Rust crate
 *                   links against (crates/gadget-ng-hip/src/ffi.rs:14-33,
 *                   declared in crates/gadget-ng-hip/hip/pm_gravity.h:24-51),
 *                   so HipPmSolver links unchanged.
 *  2. gng_pm_*   -- the versioned API: configuration struct, integer error
 *                   codes with gng_last_error_string(), an f64 Vec3 entry point
 *                   that matches GravitySolver::accelerations_for_indices
 *                   (crates/gadget-ng-core/src/gravity.rs:14-24), a device-
 *                   resident entry point on a caller stream, the density mesh
 *                   readback and per-stage timings.
 *
 * Semantics common to every solve (reference: crates/gadget-ng-pm/src/solver.rs:74-140):
 *   - the density uses all n particles; accelerations are produced for the
 *     k requested indices (or all n);
 *   - positions outside [0, L) are wrapped periodically (cic.rs:50-52);
 *   - eps2 is accepted and ignored, as PmSolver ignores it (solver.rs:70-72);
 *   - the output is linear in g;
 *   - r_split > 0 applies the long-range filter exp(-k^2 r_split^2 / 2)
 *     (fft_poisson.rs:87-103, TreePM long range: treepm/src/solver.rs:83-90);
 *   - plummer_eps > 0 applies exp(-k^2 eps^2) (fft_poisson.rs:106-122);
 *   - deconvolve != 0 divides the Green's function by W_CIC(k)^2,
 *     W_CIC = prod_alpha sinc^2(pi n_alpha / nm) (an extension; off = reference);
 *   - the cell coordinate c = (x nm / L) rem_euclid nm is formed in FP64 on
 *     the device (cic.rs:50-52) for FP32 and FP64 positions alike; the
 *     CIC weights are FP64 and accumulate in 64-bit fixed point, so the
 *     density is bit-identical from run to run; the FP32 mesh holds the mass
 *     per cell minus the mean (FFT, Green's function and interpolation FP32);
 *   - a non-finite coordinate puts its particle on node 0 of that axis (the
 *     reference's weights would turn the whole mesh into NaN); no index
 *     leaves the mesh whatever the input.
 * A handle is not reentrant: calls on one handle must not overlap.  The GPU
 * work of successive calls on one handle is ordered even across streams: an
 * entry point that runs on another stream than the previous call's first
 * waits for that call's work.
 */
#ifndef GNG_PM_H
#define GNG_PM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Raised whenever a struct's layout or an entry point's signature changes,
 * so a binding can refuse a library it does not match (2: gng_pm_tuning_t
 * has the fields through deposit_pair). */
#define GNG_PM_ABI_VERSION 2

/* Error codes (0 = success). */
enum gng_status {
  GNG_OK = 0,
  GNG_ERR_INVALID = 1,     /* bad argument or unsupported size            */
  GNG_ERR_HIP = 2,         /* HIP runtime error (see gng_last_error_string) */
  GNG_ERR_NO_DEVICE = 3,   /* no usable GPU                                 */
  GNG_ERR_OOM = 4,         /* device allocation failed                      */
  GNG_ERR_UNSUPPORTED = 5  /* mesh size not compiled in                     */
};

/* ---- 1. reference-compatible entry points -------------------------------- */

/* Replaces hip_pm_create (crates/gadget-ng-hip/hip/pm_gravity.h:24,
 * pm_gravity.hip:278-334).  NULL on any failure. */
void *hip_pm_create(int grid_size, float box_size);

/* Replaces hip_pm_destroy (pm_gravity.h:29).  NULL-safe. */
void hip_pm_destroy(void *h);

/* Replaces hip_pm_solve (pm_gravity.h:45-51, pm_gravity.hip:363-484).
 * Host SoA f32 arrays of length n.  Synchronous.  n <= 0 or h == NULL is a
 * no-op returning 0.  Returns 0, or a nonzero error code. */
int hip_pm_solve(void *h, const float *x, const float *y, const float *z, const float *mass, float *ax,
                 float *ay, float *az, int n, float eps2, float g, float r_split);

/* ---- 2. versioned API ----------------------------------------------------- */

typedef struct gng_pm_config {
  int32_t grid_size;    /* mesh side nm; 2^a 3^b, 4 <= nm <= 1024 (see gng_pm_mesh_supported) */
  int32_t deconvolve;   /* 0 = reference semantics                              */
  double box_size;      /* L, same units as positions                           */
  double r_split;       /* <= 0: none.  TreePM long-range split radius           */
  double plummer_eps;   /* <= 0: none                                           */
  int32_t device;       /* HIP device ordinal, -1 = current                     */
  int32_t reserved[7];
} gng_pm_config;

typedef struct gng_pm_plan gng_pm_plan;

int gng_pm_abi_version(void);

/* Message of the last error on the calling thread ("" if none). */
const char *gng_last_error_string(void);

/* 1 if nm is compiled in, else 0. */
int gng_pm_mesh_supported(int grid_size);

/* Number of visible GPUs (0 when none, never an error). */
int gng_pm_device_count(void);

/* Allocates the meshes (about 3.5 * 4 * nm^3 bytes) and the tables. */
int gng_pm_create(const gng_pm_config *cfg, gng_pm_plan **out);

void gng_pm_destroy(gng_pm_plan *p);

/* The configuration the plan was created with, device resolved to the
 * ordinal in use. */
int gng_pm_get_config(const gng_pm_plan *p, gng_pm_config *out);

/* Mirrors GravitySolver::accelerations_for_indices
 * (crates/gadget-ng-core/src/gravity.rs:16-24, PmSolver at
 * crates/gadget-ng-pm/src/solver.rs:74-140): pos is n x 3 f64 interleaved
 * (the layout of Vec3), mass n f64, idx k int64 indices into [0, n) or NULL
 * (then k must equal n), out k x 3 f64.  Host memory; synchronous.  The f64
 * positions go to the device unrounded; cells and weights are formed in
 * FP64 there (masses enter the fixed-point deposit as FP32). */
int gng_pm_accelerations_f64(gng_pm_plan *p, const double *pos, const double *mass, int64_t n,
                             const int64_t *idx, int64_t k, double eps2, double g, double *out);

/* Same on host SoA f32 arrays (the hip_pm_solve data layout); idx may be NULL. */
int gng_pm_accelerations_f32(gng_pm_plan *p, const float *x, const float *y, const float *z,
                             const float *mass, int64_t n, const int64_t *idx, int64_t k, double g,
                             float *ax, float *ay, float *az);

/* Device-resident solve: every pointer is device memory (SoA f32, idx int64
 * or NULL), enqueued on `stream` (a hipStream_t, NULL = default stream); the
 * call returns without synchronising.  Inputs must stay valid until the
 * stream reaches the end of the solve. */
int gng_pm_solve_device(gng_pm_plan *p, const float *d_x, const float *d_y, const float *d_z,
                        const float *d_mass, int64_t n, const int64_t *d_idx, int64_t k, double g,
                        float *d_ax, float *d_ay, float *d_az, void *stream);

/* gng_pm_accelerations_f64 on device memory (GravitySolver::accelerations_for_indices,
 * gravity.rs:16-24, with the Vec3 layouts of solver.rs:74-140): d_pos n x 3
 * f64 interleaved, d_mass n f64, d_idx k int64 or NULL (k == n), d_out k x 3
 * f64; enqueued on `stream`, returns without synchronising. */
int gng_pm_solve_device_f64(gng_pm_plan *p, const double *d_pos, const double *d_mass, int64_t n,
                            const int64_t *d_idx, int64_t k, double g, double *d_out, void *stream);

/* The order in which the deposit visits n device Vec3 particles (f64 positions
 * and masses, as gng_pm_solve_device_f64): d_order[j] (uint32) is the
 * particle in sorted slot j.  Slots are grouped by bucket = Hilbert rank of
 * the particle's 16^3 (or the mesh's) deposit tile times 8 plus its class
 * (bit a set when the base cell is on the tile's last layer along axis a),
 * buckets ascending; the order inside a bucket varies from run to run (no
 * result depends on it).  Runs no deposit or FFT; the plan's meshes are left
 * as they were.  No counterpart in the reference; it exposes the spatial
 * order of the deposit (the role sfc.rs:118-183 plays for the reference's
 * domain decomposition) to callers that lay particles out for locality. */
int gng_pm_deposit_order(gng_pm_plan *p, const double *d_pos, const double *d_mass, int64_t n, uint32_t *d_order,
                         void *stream);

/* Device-resident solve on SoA f64 positions (the integrator's particle state,
 * crates/gadget-ng-integrators/src/leapfrog.rs:5-90 keeps positions in f64),
 * f32 masses and f32 accelerations; otherwise as gng_pm_solve_device. */
int gng_pm_solve_device_soa64(gng_pm_plan *p, const double *d_x, const double *d_y, const double *d_z,
                              const float *d_mass, int64_t n, const int64_t *d_idx, int64_t k, double g,
                              float *d_ax, float *d_ay, float *d_az, void *stream);

/* Power spectrum of the particles' CIC density on the plan's mesh: mirrors
 * power_spectrum (crates/gadget-ng-analysis/src/power_spectrum.rs:35-127).
 * delta = rho / rho_bar - 1; P(k) = <|delta_k|^2 / W(k)^2> (V / nm^3)^2 over
 * the shell floor(|n| - 1/2) of integer wavenumber |n| in [1/2, nm/2), with
 * W = prod_alpha sinc(n_alpha / nm); k = (shell + 1) 2 pi / L.  Writes the
 * non-empty shells in increasing k (n_bins of them, at most nm/2; capacity
 * must allow nm/2) to k, pk and n_modes (any may be NULL).  The plan's
 * filters and G play no part; the mass mesh is the force solve's (64-bit
 * fixed point, f32 per cell) and the transform FP32.  Device pointers on
 * `stream`; the call synchronises it.  n must be positive; zero total mass
 * gives non-finite P as in the reference. */
int gng_pm_power_spectrum(gng_pm_plan *p, const float *d_x, const float *d_y, const float *d_z,
                          const float *d_mass, int64_t n, int32_t capacity, double *k, double *pk,
                          uint64_t *n_modes, int32_t *n_bins, void *stream);

/* Same from host f64 arrays: pos n x 3 interleaved (Vec3), mass n; the
 * positions reach the device unrounded (cells formed in FP64). */
int gng_pm_power_spectrum_f64(gng_pm_plan *p, const double *pos, const double *mass, int64_t n,
                              int32_t capacity, double *k, double *pk, uint64_t *n_modes,
                              int32_t *n_bins);

/* Keep the density of the following solves for gng_pm_density (costs one
 * extra nm^3 f32 buffer; off by default, when the density is written straight
 * into the spectrum buffer). */
int gng_pm_keep_density(gng_pm_plan *p, int enable);

/* Density of the last solve as mass per cell (nm^3 values widened to f64,
 * flat iz*nm^2+iy*nm+ix: the output of cic::assign,
 * crates/gadget-ng-pm/src/cic.rs:260-292).  The deposit sums 64-bit fixed
 * point and rounds once to f32, so the mesh is bit-identical for identical
 * inputs in any particle order.  Needs gng_pm_keep_density(p, 1) before the
 * solve.  Synchronous. */
int gng_pm_density(gng_pm_plan *p, double *out);

/* Force meshes of the last solve (3 x nm^3 f32: F_x, F_y, F_z, flat
 * iz*nm^2+iy*nm+ix), the output of fft_poisson::solve_forces
 * (crates/gadget-ng-pm/src/fft_poisson.rs:66-122).  Synchronous. */
int gng_pm_force_meshes(gng_pm_plan *p, float *out);

/* Launch-choice autotuner.  Some stages have more than one compiled variant:
 * the y forward, z + Green and y inverse passes run either one workgroup per
 * column item or one persistent workgroup per CU that prefetches its next
 * item; the deposit runs 256- or 512-thread tile workgroups.
 * gng_pm_autotune times whole solves of nm^3/8 synthetic uniform particles on
 * the plan's own buffers (best of `reps` after a warm-up), flips one choice at
 * a time from the defaults and keeps each flip that is faster.  The result
 * applies to this plan and to plans created later in the process with the
 * same device architecture and mesh side; with GNG_PM_TUNE_FILE set it is also
 * read from and written to that file.  Untuned plans use the persistent z pass
 * at nm = 512 only and the 512-thread deposit.  No reference counterpart.
 * Synchronous; the density and potential of earlier solves are invalidated. */
typedef struct gng_pm_tuning_t {
  int32_t z_persistent;      /* z + Green pass: persistent workgroups          */
  int32_t deposit_wide;      /* deposit: 512-thread (1) or 256-thread (0)       */
  int32_t y_fwd_persistent;  /* y forward pass: persistent workgroups          */
  int32_t y_inv_persistent;  /* y inverse pass: persistent workgroups          */
  int32_t tuned;             /* 1 when measured or read from the cache (get)  */
  int32_t interp_sorted;     /* full-set interpolation in the deposit's sorted
                                order (1) instead of the caller's (0, default).
                                Same values either way.  Sorted order keeps the
                                mesh reads local whatever the caller's order but
                                scatters the writes: at 256^3 on 512^3 it takes
                                2.6 instead of 4.7 ms for shuffled particles and
                                1.1 instead of 0.44 ms for lattice-ordered ones.
                                Not autotuned: it depends on the caller's order */
  int32_t plane_chunk_fwd;   /* the forward x/y pass pair runs in chunks of
                                this many z planes (0: over the whole mesh), so
                                the y pass reads the x pass's output from the
                                last-level cache */
  int32_t plane_chunk_inv;   /* same for the inverse y/x pair (three fields).
                                set: 0 or a proper divisor of grid_size; the
                                autotuner chooses between 0 and a size-based
                                value (fwd: 256 MiB of planes, 256 at 512^3 and
                                64 at 1024^3; inv: 64 MiB, 64 at 512^3 and 16
                                at 1024^3, where the default is 0); none at
                                256^3 and below */
  int32_t deposit_split;     /* deposit: tiles holding more than 32768 particles
                                (dense clumps) are shared by several workgroups
                                (1, default) or each run by one (0).  Same
                                density bit for bit either way */
  int32_t y_inv_narrow;      /* inverse y pass on 8-column tiles of 512 threads
                                (1, default) instead of 16 columns of 1024, where
                                both are compiled (nm = 512; elsewhere ignored):
                                the three live fields leave a 1024-thread
                                workgroup alone on its CU.  Same values either
                                way */
  int32_t deposit_pair;      /* deposit without split tiles: two tiles per
                                workgroup (1), whose records load while the
                                other tile is written, or one (0, default).
                                Same density bit for bit either way */
} gng_pm_tuning_t;

int gng_pm_autotune(gng_pm_plan *p, int reps);
/* The same coordinate descent, timed on the caller's own particles (device
 * memory in the layout of gng_pm_solve_device_f64; d_out receives the
 * accelerations of the last trial), so choices that depend on how the
 * particles are distributed (the deposit's workgroup size, the split of dense
 * tiles) are made for that distribution. */
int gng_pm_autotune_f64(gng_pm_plan *p, const double *d_pos, const double *d_mass, int64_t n, double *d_out,
                        int reps);
int gng_pm_get_tuning(const gng_pm_plan *p, gng_pm_tuning_t *out);
/* Sets the choices explicitly (reproducible runs, testing every variant);
 * the plan then reports tuned = 0. */
int gng_pm_set_tuning(gng_pm_plan *p, const gng_pm_tuning_t *choices);

/* Keep the mesh potential of the following solves (one extra nm^3 f32 buffer
 * and one extra x pass).  The potential is the field whose gradient the
 * solve returns, Phi_hat = -4 pi G W rho_hat / k^2 with the same filters and
 * deconvolution (fft_poisson.rs:509-549 before the -i k factor); k = 0 is
 * dropped.  No reference counterpart: the learned correction reads it. */
int gng_pm_keep_potential(gng_pm_plan *p, int enable);

/* Potential of the last solve (nm^3 values widened to f64, flat
 * iz*nm^2+iy*nm+ix).  Synchronous. */
int gng_pm_potential(gng_pm_plan *p, double *out);

/* Device pointers of the kept mass-per-cell and potential meshes of the last
 * solve (either output pointer may be NULL).  They stay valid until the plan
 * is destroyed and are overwritten by the next solve. */
int gng_pm_device_meshes(gng_pm_plan *p, const float **d_mass, const float **d_potential);

/* Particle moments per cell of the particles of the last solve (or
 * gng_pm_deposit_order query), on `stream`: d_mom (device, 5 doubles per
 * cell, cell = (iz nm + iy) nm + ix) receives sum m (f_a - 1/2) for the
 * three axes, sum m |f - 1/2|^2 and sum m over the particles whose base cell
 * it is, f being the in-cell fraction of the FP64 cell coordinate the
 * deposit uses.  The sums are 64-bit fixed point inside each deposit tile
 * (no atomics leave the tile), so they do not depend on the particle order.
 * No counterpart in the reference; the learned correction's network input
 * (gng_net_input_moments) reads them. */
int gng_pm_cell_moments(gng_pm_plan *p, double *d_mom, void *stream);

/* Per-stage GPU times in milliseconds, from HIP events recorded on the solve's
 * stream between the stages of every solve issued while timing is enabled
 * (gng_pm_set_timing resets the record), averaged over the most recent 64
 * such solves.  Recording does not synchronise; gng_pm_stage_times waits for
 * the recorded events.  Stages, in order: 0 bucket count (tile/class of
 * every particle, max mass), 1 sort (scan + scatter), 2 deposit, 3 x forward,
 * 4 y forward, 5 z + Green, 6 y inverse, 7 x inverse, 8 interpolation;
 * ms[9], when capacity > 9, receives the mean whole-solve time.  Returns the
 * stage count (negative on error). */
int gng_pm_set_timing(gng_pm_plan *p, int enable);
int gng_pm_stage_times(gng_pm_plan *p, float *ms, int capacity);

/* ---- ensembles of independent boxes --------------------------------------
 * No single reference function: the reference solves one box per call
 * (PmSolver::accelerations_for_indices); an ensemble is a set of such calls
 * run together.  Each box has its own configuration (mesh size, box size,
 * filters).  The first solve with a given set of pointers, offsets and g
 * captures the ensemble into one HIP graph; later solves with the same
 * arguments replay it.  By default the boxes are grouped by mesh side and
 * every stage (bucket count, scan, scatter, deposit, each FFT pass, the
 * interpolation) is one launch over all boxes of a side, box b on
 * blockIdx.z of the launch with its pointers and constants from a device
 * table; the groups run on parallel graph branches.  With box tables off
 * (gng_pm_batch_set_box_tables(b, 0)) each box runs its own pipeline and the
 * boxes are spread over 8 branches.  Every split gives the same accelerations
 * bit for bit. */
typedef struct gng_pm_batch gng_pm_batch;

int gng_pm_batch_create(int n_boxes, const gng_pm_config *cfgs, gng_pm_batch **out);
void gng_pm_batch_destroy(gng_pm_batch *b);
int gng_pm_batch_size(const gng_pm_batch *b);
/* The plan of box i (for per-box readback such as gng_pm_force_meshes). */
int gng_pm_batch_plan(gng_pm_batch *b, int i, gng_pm_plan **out);

/* Device-resident ensemble solve.  The particles of box i are entries
 * [offsets[i], offsets[i+1]) of the concatenated SoA f32 device arrays
 * (offsets is a host array of n_boxes + 1 values); accelerations are written
 * at the same entries.  Enqueued on `stream`; returns without synchronising. */
/* Autotunes one plan per distinct mesh side and applies the choice to every
 * box of that side; the next solve captures a new graph. */
int gng_pm_batch_autotune(gng_pm_batch *b, int reps);
int gng_pm_batch_solve_device(gng_pm_batch *b, const float *d_x, const float *d_y, const float *d_z,
                              const float *d_mass, const int64_t *offsets, double g, float *d_ax,
                              float *d_ay, float *d_az, void *stream);
/* Launch split of the ensemble graph: with k >= 1 every stage is one launch
 * over at most k boxes of one mesh side (the default, a very large k, is one
 * launch per stage per mesh side); k = 0 runs one pipeline per box.  Runs go
 * to 8 graph branches round robin.  The next solve captures a new graph. */
int gng_pm_batch_set_box_tables(gng_pm_batch *b, int boxes_per_launch);
int gng_pm_batch_get_box_tables(const gng_pm_batch *b);
/* Times whole solves on the caller's particles (best of reps) for k = the
 * largest mesh side's box count, its half and quarter, and 0, and keeps the
 * fastest (written to *chosen when not NULL). */
int gng_pm_batch_tune_launches(gng_pm_batch *b, const float *d_x, const float *d_y, const float *d_z,
                               const float *d_mass, const int64_t *offsets, double g, float *d_ax,
                               float *d_ay, float *d_az, int reps, int *chosen);

/* ---- allocation-free solve ------------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  gng_pm_solve_device without a plan: the meshes, tables
 * and sort buffers live in one static device array of
 * GNG_PM_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends); the call
 * makes no hipMalloc, hipFree or other allocating call.  Limits: a compiled
 * grid_size of at most GNG_PM_NOALLOC_MAX_GRID and at most
 * GNG_PM_NOALLOC_MAX_PARTICLES particles (larger inputs return
 * GNG_ERR_INVALID).  Same semantics, inputs and results as
 * gng_pm_solve_device on a plan of *cfg with default launch choices; the
 * device is cfg->device (-1: current).  Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, other threads)
 * race, so the caller serialises them (e.g. one stream). */
#define GNG_PM_NOALLOC_MAX_GRID 64
#define GNG_PM_NOALLOC_MAX_PARTICLES (1LL << 19)
#define GNG_PM_NOALLOC_SCRATCH_BYTES (20u << 20)
int gng_pm_solve_device_noalloc(const gng_pm_config *cfg, const float *d_x, const float *d_y, const float *d_z,
                                const float *d_mass, int64_t n, const int64_t *d_idx, int64_t k, double g,
                                float *d_ax, float *d_ay, float *d_az, void *stream);

/* 3-D Hilbert key of (x, y, z), `bits` bits per axis: Skilling's
 * AxesToTranspose (J. Skilling, AIP Conf. Proc. 707, 381 (2004)) with the axes
 * visited in ascending order and the transposed bits packed from the most
 * significant level down, x first in each level.  Consecutive keys are face
 * neighbours at every order.  The deposit orders its tiles, and hence the
 * sorted particles, by this key.  Pure host function; needs no device.
 * Analogue of coords_to_hilbert (crates/gadget-ng-parallel/src/sfc.rs:118-183,
 * used there for MPI domains), which agrees at bits = 1 (the order listed at
 * sfc.rs:113-116) but not beyond: it visits the axes in reverse (:134) and
 * packs the least significant level first (:177-181), so its keys at
 * bits >= 2 are not a face-adjacent curve.  This key does not reproduce it. */
uint64_t gng_hilbert_key3(uint32_t x, uint32_t y, uint32_t z, int32_t bits);

/* ---- FP64 GPU form --------------------------------------------------------
 * The reference algorithm moved to the GPU in FP64, kept as the baseline the
 * FP32 solver is measured against (not a product path): per-particle FP64
 * atomic deposit (cic.rs:37-93), full complex-to-complex 3-D FFT in three 1-D
 * passes, the spectral kernel with the Nyquist index at +nm/2
 * (fft_poisson.rs:509-549, 724-730), three inverse FFTs, Re()/nm^3 and FP64
 * interpolation (cic.rs:100-170).  Mesh sides: powers of two, 4..1024; cfg
 * filters and deconvolve as above.  About 48 nm^3 bytes of device memory. */
typedef struct gng_pm64_plan gng_pm64_plan;
int gng_pm64_create(const gng_pm_config *cfg, gng_pm64_plan **out);
void gng_pm64_destroy(gng_pm64_plan *p);
/* Host f64 in/out (pos n x 3 interleaved, idx k or NULL), with the H2D and
 * D2H copies of every call.  Synchronous. */
int gng_pm64_accelerations(gng_pm64_plan *p, const double *pos, const double *mass, int64_t n,
                           const int64_t *idx, int64_t k, double g, double *out);
/* Device-resident variant on `stream` (asynchronous). */
int gng_pm64_solve_device(gng_pm64_plan *p, const double *d_pos, const double *d_mass, int64_t n,
                          const int64_t *d_idx, int64_t k, double g, double *d_out, void *stream);

/* ---- 3. device-memory helpers --------------------------------------------
 * No counterpart in the reference.  For callers that hold no GPU runtime of
 * their own (test and benchmark harnesses, FFI hosts) and need device-resident
 * inputs for gng_pm_solve_device; they run on the HIP runtime this library
 * links, so a process never mixes two runtimes.  kind: 0 host->device,
 * 1 device->host, 2 device->device.  Synchronous. */
int gng_dev_alloc(void **ptr, uint64_t bytes);
int gng_dev_free(void *ptr);
int gng_dev_copy(void *dst, const void *src, uint64_t bytes, int kind);
int gng_dev_synchronize(void);
/* d_y[i] += a * d_x[i] for n floats on `stream` (asynchronous). */
int gng_dev_axpy(float *d_y, const float *d_x, int64_t n, float a, void *stream);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* GNG_PM_H */
