/*
 * spimlearn.h -- C ABI of libspimlearn.so: learned reconstruction of a full-resolution
 * max projection from a coarse one, on gfx950 matrix cores (BF16 operands, FP32
 * accumulation), with the training step on the device as well.
 *
 * No counterpart in the reference.  Its nearest analogue is progressive multipass
 * rendering (numParts / currentPart, spimagine/volumerender/kernels/volume_kernel.cl:100-106,
 * 172-183; driven from spimagine/gui/glwidget.py:633-635, 688-694): a cheaper frame first,
 * the exact one later.  Here the cheap frame is a max projection rendered at half the
 * resolution with a quarter of the ray samples (max_steps / 4), from the volume and from a
 * max-filtered copy of it (spr_volume_add_dilated); the network maps those two planes to
 * the full-resolution frame.  The exact renderer stays the reference and the fallback.
 *
 * The network (all 3x3 convolutions, zero padding 1, C = 32 channels):
 *   a0  = relu(conv_in(x))             x: 2 planes h x w (coarse MIP, coarse dilated MIP)
 *   e0  = relu(conv(a0)), e0b = relu(conv(e0))                 h   x w
 *   e1  = relu(conv(avgpool2(e0b))), e1b = relu(conv(e1))      h/2 x w/2
 *   e2  = relu(conv(relu(conv(avgpool2(e1b)))))                h/4 x w/4
 *   u1  = relu(conv(up2(e2) + e1b))                            h/2 x w/2  (up2: nearest x2)
 *   u0  = relu(conv(up2(u1) + e0b)), u0b = relu(conv(u0))      h   x w
 *   y   = pixel_shuffle2(head(u0b)) + bilinear_up2(x plane 0)  2h x 2w
 * Parameters: one float32 array of spl_param_count() values, in this order:
 *   conv_in W[9][2][32], b[32];  nine 32->32 layers (e0, e0b, e1, e1b, e2a, e2, u1, u0,
 *   u0b) W[9][32][32], b[32];  head W[9][32][4], b[4];  tap index t = 3 (dy + 1) + (dx + 1),
 *   out(y, x, co) = b[co] + sum_t sum_ci in(y + dy, x + dx, ci) W[t][ci][co].
 * Activations are stored as bf16 (channels fastest), gradients as float32; the master
 * weights and the Adam state are float32.
 */
#ifndef SPIMLEARN_H
#define SPIMLEARN_H

#ifdef __cplusplus
extern "C" {
#endif

#define SPL_ABI_VERSION 2

#define SPL_OK 0
#define SPL_ERR_INVALID_ARG (-1)
#define SPL_ERR_OUT_OF_MEMORY (-4)
#define SPL_ERR_HIP (-5)
#define SPL_ERR_NO_DEVICE (-6)

typedef struct spl_net spl_net;

int spl_abi_version(void);
/* Number of float32 parameters (84,996). */
int spl_param_count(void);

/* A network on a device, parameters zero until spl_set_params. */
int spl_create(int device, spl_net **net);
int spl_destroy(spl_net *net);
/* Message of the calling thread's last failing call on `net` ("" if none). */
const char *spl_last_error(const spl_net *net);

/* Copy all parameters in / out (host float32, n == spl_param_count()).  Setting them
   resets the Adam state. */
int spl_set_params(spl_net *net, const float *host, int n);
int spl_get_params(spl_net *net, float *host, int n);

/* Inference: d_in = n images of 2 float32 planes (n, 2, h, w); d_out = (n, 2h, 2w)
   float32.  h and w must be multiples of 4.  Enqueued on `stream` (hipStream_t; NULL =
   the network's own stream); no host synchronisation. */
int spl_forward(spl_net *net, const float *d_in, int n, int h, int w, float *d_out,
                void *stream);

/* Gradient of the summed squared error sum (y - target)^2 / (n 4 h w) with respect to
   every parameter, into host_grad (spl_param_count() floats); *sse = the summed squared
   error.  Blocks until done.  (The training step's gradient, exposed for testing.) */
int spl_gradients(spl_net *net, const float *d_in, const float *d_target, int n, int h, int w,
                  float *host_grad, double *sse, void *stream);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, learning rate lr) on the mean squared
   error of the batch; *sse = the batch's summed squared error before the step (may be
   NULL).  Blocks until the loss is read back when sse != NULL. */
int spl_train_step(spl_net *net, const float *d_in, const float *d_target, int n, int h, int w,
                   float lr, double *sse, void *stream);

/* The per-frame judge's error against gx x gy exact patches of patch x patch pixels (an
   exact render of those pixels): patch (ty, tx) covers the frame pixels from row
   ty (h / gy) + oy and column tx (w / gx) + ox on; d_patches holds the patches one after
   another (tx fastest), rows of `patch` floats (patch 1: a (w / gx) x (h / gy)-strided
   sparse frame, row-major).  *d_sse (one device float) = (accumulate ? *d_sse : 0) + the
   summed squared difference, summed in a fixed order.  Enqueued on `stream`. */
int spl_judge_sse(spl_net *net, const float *d_frame, int w, int h, const float *d_patches,
                  int patch, int gx, int gy, int ox, int oy, float *d_sse, int accumulate,
                  void *stream);

/* The judge's error of n frames at once against their sparse exact frames (the patch-1
   form of spl_judge_sse): frames = a host array of n device pointers to h x w frames,
   d_sparse = n consecutive gx x gy sparse frames (sparse pixel (tx, ty) of frame k is
   frame pixel (tx w/gx + ox, ty h/gy + oy)); d_sse[k] = frame k's summed squared
   difference and, when d_sse4 is not NULL, d_sse4[k] = its summed fourth power, each
   summed in a fixed order over many workgroups.  Enqueued on `stream`; one stream at a
   time per network (the partial sums live in the network). */
int spl_judge_sse_frames(spl_net *net, const float *const *frames, int n, int w, int h,
                         const float *d_sparse, int gx, int gy, int ox, int oy, float *d_sse,
                         float *d_sse4, void *stream);

/* The judge's decision with a confidence bound, on the device: with m2 = d_sse[k] / npix
   and m4 = d_sse4[k] / npix, d_flags[k] = 0 (keep the learned frame) only when
   m2 + z sqrt((m4 - m2^2) / npix) <= mse_max -- the mean squared error of the judge's
   npix pixels plus z standard errors of that mean -- else 1 (not-a-number included).
   Frames whose error sits in few pixels get a wider bound than frames with an even error
   of the same mean. */
int spl_judge_flags_bound(spl_net *net, const float *d_sse, const float *d_sse4, int n, int npix,
                          float mse_max, float z, unsigned int *d_flags, void *stream);

/* The judge's decision on the device: d_flags[k] = 1 when d_sse[k] / npix > mse_max (the
   frame's estimated PSNR below the acceptance level; not-a-number included), else 0, for
   k < n, on `stream`.  spr_render_slot_if reads a flag to render the exact fallback without
   a host round trip. */
int spl_judge_flags(spl_net *net, const float *d_sse, int n, int npix, float mse_max,
                    unsigned int *d_flags, void *stream);

/* d_dst (2h x 2w) = nearest-neighbour x2 upsampling of d_src (h x w), on `stream` (the
   learned frame's alpha from the coarse pass). */
int spl_upsample2(spl_net *net, const float *d_src, int w, int h, float *d_dst, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SPIMLEARN_H */
