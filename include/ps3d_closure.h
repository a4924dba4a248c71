/*
 * ps3d_closure.h -- learned subgrid closure for a coarse ps3d grid (no counterpart
 * in the reference; ps3d has no subgrid model).
 *
 * The net: four 3x3x3 convolutions 6 -> 32 -> 32 -> 32 -> 3 channels with GELU
 * between them, on the nx x ny x (nz+1) points of a solver grid (x, y periodic,
 * zero beyond the lids in z).  Inputs per point: u, v, w, xi, eta, zeta (scaled);
 * outputs: the correction to the three vorticity tendency components (scaled).
 * Matrix products run on the MI355X matrix cores in bf16 with fp32 accumulation.
 *
 * Host tensors are float32, channels last, points in the solver's order
 * p = (ix*ny + iy)*(nz+1) + iz.  Parameters are packed layer by layer as
 * W[27][cout][cin] (tap t = (dx+1)*9 + (dy+1)*3 + (dz+1)) followed by b[cout].
 * Loss: mean over points and the 3 output channels of (y - target)^2.  Optimiser:
 * Adam (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected).
 * Errors: nonzero ps3d_status, message from ps3d_cnn_last_error().
 */
#ifndef PS3D_CLOSURE_H
#define PS3D_CLOSURE_H

#include "ps3d_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ps3d_cnn ps3d_cnn;

const char *ps3d_cnn_last_error(void);
/* number of packed parameters (63171) */
int ps3d_cnn_nparams(int *n);
/* a net for the grid nx x ny x (nz+1); weights start at zero */
int ps3d_cnn_create(int nx, int ny, int nz, ps3d_cnn **net);
int ps3d_cnn_destroy(ps3d_cnn *net);
int ps3d_cnn_set_weights(ps3d_cnn *net, const float *packed);   /* also resets Adam */
int ps3d_cnn_get_weights(ps3d_cnn *net, float *packed);
/* x: [npts][6], y: [npts][3] */
int ps3d_cnn_forward(ps3d_cnn *net, const float *x, float *y);
/* loss, packed parameter gradient (may be NULL) and input gradient [npts][6] (may be NULL) */
int ps3d_cnn_grad(ps3d_cnn *net, const float *x, const float *target, float *loss, float *grad, float *gx);
/* forward + backward + one Adam step with learning rate lr */
int ps3d_cnn_train_step(ps3d_cnn *net, const float *x, const float *target, float lr, float *loss);
/* device-resident training set: nsamples (x, target) pairs, then epochs over an order */
int ps3d_cnn_dataset_reserve(ps3d_cnn *net, int nsamples);
int ps3d_cnn_dataset_set(ps3d_cnn *net, int i, const float *x, const float *target);
int ps3d_cnn_train_epoch(ps3d_cnn *net, const int *order, int n, float lr, float *mean_loss);
/* mean kernel time (HIP events) of one forward and of one forward + backward */
int ps3d_cnn_timing(ps3d_cnn *net, int reps, float *fwd_ms, float *train_ms);

/* Attach a net to a solver handle of the same grid (net = NULL detaches): every RK
 * stage of ps3d_gpu_advance_n then adds out_scale[c] * net(in_scale * (u, v, w, xi,
 * eta, zeta))_c, decomposed like a physical field, to the vorticity tendency before
 * the substep.  The net stays owned by the caller and must outlive the attachment. */
int ps3d_gpu_set_closure(ps3d_handle *h, ps3d_cnn *net, const float *in_scale, const float *out_scale);

#ifdef __cplusplus
}
#endif
#endif /* PS3D_CLOSURE_H */
