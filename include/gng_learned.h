/*
 * gng_learned.h -- TreePM short range and the learned sub-cell correction
 * (part of libgng_pm.so, gfx950 only).
 *
 * Short range.  gng_sr_accelerations() replaces the short-range half of
 * TreePmSolver::accelerations_for_indices (crates/gadget-ng-treepm/src/
 * solver.rs:92-107 -> short_range.rs:120-208 and the periodic walk at
 * short_range.rs:232-330).  The reference walks an octree with a monopole
 * shortcut for small nodes; this library sums every pair inside r_cut over a
 * chaining mesh (cells of side >= r_cut, 27 neighbour cells), which is the
 * "direct sum" the reference's walk approximates.  Pair law: short_range.rs:
 * 384-398 (softened r, minimum image, bare-distance cut).  kernel selects the
 * split factor S(x), x = r / (sqrt(2) r_split):
 *   GNG_SR_EXACT      S = erfc(x) + (2x/sqrt(pi)) exp(-x^2), the exact
 *                     complement of the long-range filter exp(-k^2 r_s^2/2)
 *                     (fft_poisson.rs:529-531): long + short = Newton;
 *   GNG_SR_REFERENCE  S = erfc(x), the reference's factor (short_range.rs:
 *                     103-110), which misses the second term.
 *
 * Learned correction (no reference counterpart).  A 3-layer network of 3^3
 * periodic convolutions, conv(8->16) SiLU, conv(16->16) SiLU, conv(16->3),
 * on the CIC density/potential mesh of a PM solve and the particles' moments
 * per cell (gng_net_input_particles); its 3-channel output mesh
 * is CIC-interpolated to the particles and added to the PM acceleration.
 * Weights and activations are bfloat16, products accumulate in FP32 on the
 * MFMA units (v_mfma_f32_16x16x32_bf16); the master weights, gradients and Adam moments are FP32.  Mesh
 * layout: [box][iz][iy][ix][channel], channels last.  A network object holds
 * the weights and the work buffers for up to max_boxes meshes of one size
 * (grid_size a multiple of 16).
 *
 * Parameters (gng_net_param_count() floats, in this order):
 *   w1[27][8][16] b1[16] w2[27][16][16] b2[16] w3[27][16][3] b3[3]
 * with tap t = (dz+1)*9 + (dy+1)*3 + (dx+1) reading input cell c + (dz,dy,dx).
 *
 * Training particles are two float4 arrays: pos4 = (x, y, z, box index) and
 * tgt4 = (target residual x, y, z, weight w_p); the loss is
 *   L = (1/n) sum_p w_p |out_scale * CIC(y)(x_p) - target_p|^2.
 * Every call on one object must be ordered (one stream, or synchronised).
 */
#ifndef GNG_LEARNED_H
#define GNG_LEARNED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- short range --------------------------------------------------------- */

enum gng_sr_kernel { GNG_SR_EXACT = 0, GNG_SR_REFERENCE = 1 };

typedef struct gng_sr gng_sr;

int gng_sr_create(int32_t device, gng_sr** out);
void gng_sr_destroy(gng_sr* sr);

/* Device-resident short-range accelerations of all n particles (float32 SoA
 * device arrays), written (not added) to d_ax/d_ay/d_az on `stream`.
 * Requires box_size >= 3 r_cut.  Synchronises `stream` once (the work-list
 * size is read back). */
int gng_sr_accelerations(gng_sr* sr, const float* d_x, const float* d_y, const float* d_z, const float* d_m,
                         int64_t n, double box_size, double r_split, double r_cut, double eps2, double g,
                         int32_t kernel, float* d_ax, float* d_ay, float* d_az, void* stream);

/* Pair interactions evaluated by the last call (inside and outside r_cut). */
int64_t gng_sr_last_pair_count(const gng_sr* sr);

/* The order in which gng_sr_accelerations() visits the particles: d_order[j]
 * (uint32, n entries) is the index of the j-th particle sorted by chaining-
 * mesh cell (nc = min(256, floor(box_size / r_cut)) cells per side, cell
 * (cz * nc + cy) * nc + cx of floor(x nc / L) mod nc), equal cells in
 * ascending index.  Stable and identical from run to run.  The reference has
 * no counterpart; the order fixes every pair sum's order, and callers can
 * use it to lay particles out in cell order. */
int gng_sr_cell_order(gng_sr* sr, const float* d_x, const float* d_y, const float* d_z, int64_t n,
                      double box_size, double r_cut, uint32_t* d_order, void* stream);

/* ---- learned correction -------------------------------------------------- */

enum gng_net_buffer {
  GNG_NET_X0 = 0,    /* bf16 [box][M][8]  network input                      */
  GNG_NET_Z1 = 1,    /* f32  [box][M][16] layer-1 pre-activation             */
  GNG_NET_X1 = 2,    /* bf16 [box][M][16] layer-1 activation                 */
  GNG_NET_Z2 = 3,    /* f32  [box][M][16]                                     */
  GNG_NET_X2 = 4,    /* bf16 [box][M][16]                                     */
  GNG_NET_Y = 5,     /* f32  [box][M][4]  output mesh (channel 3 = 0)         */
  GNG_NET_GY = 6,    /* f32  [box][M][4]  dL/dy                               */
  GNG_NET_DZ3 = 7,   /* bf16 [box][M][4]  dL/dy in bf16 (backward input)      */
  GNG_NET_DZ2 = 8,   /* bf16 [box][M][16] dL/dz2                              */
  GNG_NET_DZ1 = 9,   /* bf16 [box][M][16] dL/dz1                              */
  GNG_NET_GRAD = 10, /* f32  [param_count] gradient of the last backward      */
  GNG_NET_PARAMS = 11
};

typedef struct gng_net gng_net;

int32_t gng_net_param_count(void);
int gng_net_create(int32_t grid_size, int32_t max_boxes, int32_t device, gng_net** out);
void gng_net_destroy(gng_net* net);
int gng_net_set_params(gng_net* net, const float* host_params);
int gng_net_get_params(gng_net* net, float* host_params);

/* Buffer access for inspection and tests (host copies, synchronous). */
int64_t gng_net_buffer_bytes(const gng_net* net, int32_t which);
int gng_net_read(gng_net* net, int32_t which, void* host, int64_t bytes);
int gng_net_write(gng_net* net, int32_t which, const void* host, int64_t bytes);
/* Device address of a buffer (e.g. to stage many input meshes and copy a
 * batch of them into GNG_NET_X0 between training steps). */
int gng_net_device_buffer(gng_net* net, int32_t which, void** d_ptr);

/* Input slot `box` from a PM solve's device meshes (FP32 mass per cell and
 * potential, grid_size^3 each, gng_pm_device_meshes) and its n particles
 * (float32 SoA device arrays, non-negative masses, box of side box_size):
 *   x0[0] = mass/rho_mean - 1,
 *   x0[1] = (phi - box7(phi))/phi_scale,
 *   x0[2..4] = D/rho_mean, D = sum m (f - 1/2),
 *   x0[5] = (sum m |f - 1/2|^2 - |D|^2 / sum m)/rho_mean,
 *   x0[6..7] = 0,
 * with the sums over the particles whose base cell floor(x nm / L) mod nm is
 * the cell and f their in-cell fraction; rho_mean = mean mass per cell,
 * box7 = the periodic 7^3 box mean.  Every channel is linear in the mass, as
 * the PM force error is, and independent of the box size.  The moments are
 * 64-bit fixed-point sums, so the input does not depend on the order of the
 * particles.  n = 0 leaves channels 2-5 at zero. */
int gng_net_input_particles(gng_net* net, int32_t box, const float* d_rho, const float* d_phi, float rho_mean,
                            float phi_scale, const float* d_x, const float* d_y, const float* d_z, const float* d_m,
                            int64_t n, double box_size, void* stream);
/* gng_net_input_particles with the moments given per cell (d_mom: 5 doubles
 * per cell in that order, sum m (f - 1/2) per axis, sum m |f - 1/2|^2,
 * sum m), as gng_pm_cell_moments writes them for the particles of a PM
 * solve: no atomics, the FP64 cell coordinate of the deposit. */
int gng_net_input_moments(gng_net* net, int32_t box, const float* d_rho, const float* d_phi, float rho_mean,
                          float phi_scale, const double* d_mom, void* stream);
/* gng_net_input_particles without particles (channels 2-7 zero). */
int gng_net_input(gng_net* net, int32_t box, const float* d_rho, const float* d_phi, float rho_mean,
                  float phi_scale, void* stream);

/* Forward pass for training: also keeps the FP32 pre-activations that the
 * backward pass reads. */
int gng_net_forward(gng_net* net, int32_t boxes, void* stream);
/* Forward pass for inference: writes only the activations and the output
 * mesh (GNG_NET_Z1/Z2 are left untouched). */
int gng_net_infer(gng_net* net, int32_t boxes, void* stream);
/* Loss and dL/dy for n training particles (uses the output of the last
 * forward); the loss is kept on the device until gng_net_loss() reads it. */
int gng_net_loss_grad(gng_net* net, const float* d_pos4, const float* d_tgt4, int64_t n, double box_size,
                      float out_scale, void* stream);
int gng_net_backward(gng_net* net, int32_t boxes, void* stream);
int gng_net_adam(gng_net* net, float lr, float beta1, float beta2, float eps, void* stream);
/* forward + loss_grad + backward + adam */
int gng_net_train_step(gng_net* net, int32_t boxes, const float* d_pos4, const float* d_tgt4, int64_t n,
                       double box_size, float out_scale, float lr, void* stream);
int gng_net_loss(gng_net* net, double* loss_out);

/* Adds out_scale * CIC(y of slot `box`) to the accelerations of n particles
 * (float32 SoA device arrays). */
int gng_net_apply(gng_net* net, int32_t box, const float* d_x, const float* d_y, const float* d_z, int64_t n,
                  double box_size, float out_scale, float* d_ax, float* d_ay, float* d_az, void* stream);
/* gng_net_apply for the field only: particles whose base cell holds more
 * than rho_max (d_rho = the box's mass per cell, nm^3 float32, the mesh kept
 * by a plan created with keep_density) keep their force, as they carry no
 * weight in the training loss.  d_rho = NULL applies to every particle. */
int gng_net_apply_field(gng_net* net, int32_t box, const float* d_x, const float* d_y, const float* d_z,
                        int64_t n, double box_size, float out_scale, const float* d_rho, double rho_max,
                        float* d_ax, float* d_ay, float* d_az, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GNG_LEARNED_H */
