/*
 * spimrender.h -- C ABI of libspimrender.so, the MI355X (gfx950) volume ray caster
 * behind spimagine's VolumeRenderer.
 *
 * Every entry point returns an int status (SPR_OK == 0, negative on error); no C++
 * exception crosses the boundary.  spr_last_error(ctx) returns the message of the calling
 * thread's last failing call if that call concerned ctx, else "" (spr_last_error(NULL):
 * last failure of spr_create or of a call that received a NULL context).
 *
 * A context owns one HIP device, one HIP stream (default priority), the resident volumes
 * and the device-side output buffers.  The stateful calls (spr_set_*, spr_render,
 * spr_render_device, spr_volume_minmax, spr_count_samples, the graph calls) are used by one
 * host thread at a time.  spr_render_batch, spr_render_slot and spr_volume_info may run on
 * several threads at once, on different streams, concurrently with spr_volume_add and
 * spr_volume_release of other slots.
 *
 * Reference interfaces replaced (paths relative to the spimagine checkout):
 *   VolumeRenderer.__init__ / resize / reset_buffer   spimagine/volumerender/volumerender.py:72-135,181-197
 *   VolumeRenderer.set_data / update_data             spimagine/volumerender/volumerender.py:235-294
 *   VolumeRenderer.update_matrices + setters          spimagine/volumerender/volumerender.py:214-233,296-325
 *   VolumeRenderer.rebuild_program (interpolation)    spimagine/volumerender/volumerender.py:137-165
 *   VolumeRenderer.render / _render_max_project       spimagine/volumerender/volumerender.py:327-390,508-547
 *   VolumeRenderer._render_isosurface                 spimagine/volumerender/volumerender.py:446-506
 *   GLWidget._get_min_max (GPU min/max of the volume) spimagine/gui/glwidget.py:328-344
 */
#ifndef SPIMRENDER_H
#define SPIMRENDER_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPR_ABI_VERSION 1

/* status codes */
#define SPR_OK 0
#define SPR_ERR_INVALID_ARG (-1)
#define SPR_ERR_NO_VOLUME (-2)
#define SPR_ERR_UNSUPPORTED_DTYPE (-3)
#define SPR_ERR_OUT_OF_MEMORY (-4)
#define SPR_ERR_HIP (-5)
#define SPR_ERR_NO_DEVICE (-6)

/* voxel types: VolumeRenderer.dtypes = [float32, uint16, uint8] (volumerender.py:68) */
#define SPR_DTYPE_F32 0
#define SPR_DTYPE_U16 1
#define SPR_DTYPE_U8 2
/* Two uint16 channels per voxel, channel 0 in the low half (4 bytes a voxel): only
   spr_volume_add, spr_render_pair, spr_volume_info, spr_volume_read and spr_volume_release
   accept such a slot (spr_volume_add_pooled_pair makes one). */
#define SPR_DTYPE_U16X2 3

/* sampler filter: interpolation_defines "linear" / "nearest" (volumerender.py:69-70) */
#define SPR_FILTER_LINEAR 0
#define SPR_FILTER_NEAREST 1

/* render(method=...) (volumerender.py:543-547) */
#define SPR_METHOD_MAX_PROJECT 0
#define SPR_METHOD_ISO_SURFACE 1

/* device layout of the resident volume (no counterpart in the reference, whose OpenCL
   runtime owns the image layout):
   PADDED  a (nz+2, ny+2, nx+2) copy with replicated edges (clamp-to-edge without
           per-sample clamping), addressed with 32-bit offsets: up to 4 GiB padded
   LINEAR  the (nz, ny, nx) volume as uploaded, any size up to 2^32-1 voxels
   AUTO    PADDED when it fits (default); above 4 GiB, LINEAR plus a padded copy that only
           the staged max projection reads (addressed per z-plane); such a slot takes both
           copies' bytes and is not batchable */
#define SPR_LAYOUT_AUTO 0
#define SPR_LAYOUT_LINEAR 1
#define SPR_LAYOUT_PADDED 2

/* launch tuning keys (spr_set_tuning; no counterpart in the reference).  Results do not
   depend on them.
   SPR_TUNE_TILE_MAP  how 16x16 pixel tiles are dealt to workgroups and so to the 8 XCDs:
                      0 = tile rows interleaved over XCDs (default), 1 = one contiguous
                      band per XCD, 2 = row-major, 3 = as 0 with each XCD's tiles in
                      descending estimated cost (per-wave staged max projection and
                      opacity mode; two small kernels order the tiles before the frame,
                      on its stream)
   SPR_TUNE_SLAB_KB   per-wave LDS slab (KiB) of the staged max projection: 8, 12, 16 or
                      19 (default: 20,448 B, or 20,464 B with one wave per workgroup;
                      per-wave staging only, 16 for workgroup staging); 0 = sample every
                      chunk from global memory
   SPR_TUNE_STAGING   0 = each wave stages its own box (default), 1 = the four waves of a
                      workgroup stage one union box into a slab of 4 x SLAB_KB
   SPR_TUNE_CHUNK     samples per staged box: 0 = from the ray step (default: 8 samples
                      when 15 steps span more than 40 voxels, else boxes of two 16-sample
                      groups; the plain max projection: 8 samples beyond 36 voxels, two
                      groups only while 31 steps span at most 40 voxels, else one), 4, 8,
                      16, 32 or 64 (boxes of 2 or 4 groups of 16 samples).  A box of
                      several groups that does not fit the slab is staged one group at a
                      time.
   SPR_TUNE_ROW_PAD   1 = pad staged rows to an odd number of 16-byte units (spreads the
                      LDS banks rows a few apart fall on), 0 = no padding (default)
   SPR_TUNE_WG_WAVES  waves per workgroup of the per-wave staged max projection (slab 16
                      or 19, and volumes above 4 GiB) and opacity mode: 1 (an 8x8 quarter
                      of a tile: a CU starts new work as soon as any of its waves ends),
                      4 (a 16x16 tile), 0 = default (1: faster at 1024^2 and 2048^2)
   SPR_TUNE_SKIP      1 = empty-space skipping and early ray termination by the volume's
                      brick bounds (default), 0 = off.  Every padded one-channel volume
                      keeps the minimum and maximum of each 8x8x8 brick (built on upload,
                      8 bytes a brick).  With the default per-wave slabs (slab_kb 19) the
                      max projection skips a 16-sample box
                      whose bricks' maximum cannot exceed any of the wave's running maxima,
                      the opacity mode stops a wave once no later sample can raise a ray's
                      colour (cumsum x the volume's normalised maximum <= colour, with a
                      transparency factor provably in [0, 1]), and the iso march skips a
                      box whose bricks lie wholly on each marching ray's starting side of
                      the iso value (after a skipped box, the next four tested as one
                      while 63 steps span at most 40 voxels).  Only skipped samples'
                      positions are still advanced, by the same additions: results are
                      bit-identical either way. */
#define SPR_TUNE_TILE_MAP 0
#define SPR_TUNE_SLAB_KB 1
#define SPR_TUNE_STAGING 2
#define SPR_TUNE_CHUNK 3
#define SPR_TUNE_ROW_PAD 4
#define SPR_TUNE_WG_WAVES 5
#define SPR_TUNE_SKIP 6

typedef struct spr_ctx spr_ctx;

/* Everything VolumeRenderer passes to its kernels besides the volume
   (volumerender.py:366-386, 451-497).  inv_p / inv_m are the row-major float32
   inverses inv(projection) and inv(modelView . mScale) computed by update_matrices
   (volumerender.py:310-316).  box = [x0,x1,y0,y1,z0,z1] (set_box_boundaries).
   max_steps is the config value max_steps (config.py:27, default 200), a runtime
   argument here instead of a compile-time define. */
typedef struct spr_params {
    float inv_p[16];
    float inv_m[16];
    float box[6];
    float min_val;      /* set_min_val */
    float max_val;      /* set_max_val; iso level is max_val/2 (volumerender.py:463) */
    float gamma;        /* set_gamma */
    float alpha_pow;    /* set_alpha_pow; 0 = plain MIP */
    float occ_strength; /* set_occ_strength */
    int occ_radius;     /* set_occ_radius */
    int occ_n_points;   /* set_occ_n_points */
    int max_steps;
    int filter;         /* SPR_FILTER_* */
} spr_params;

/* ABI version (SPR_ABI_VERSION of the built library). */
int spr_abi_version(void);

/* Number of visible HIP devices (0 without a GPU). */
int spr_device_count(int *count);

/* Free and total bytes of device memory (the reference sizes its volume budget as
   0.7 x MAX_MEM_ALLOC_SIZE, volumerender.py:102). */
int spr_device_memory(int device, size_t *free_bytes, size_t *total_bytes);

/* VolumeRenderer.__init__ (volumerender.py:72-135): bind a device, create the stream. */
int spr_create(int device, spr_ctx **ctx);
int spr_destroy(spr_ctx *ctx);
const char *spr_last_error(const spr_ctx *ctx);

/* Layout (SPR_LAYOUT_*) used by the next spr_set_volume*; spr_get_layout reports the
   resident volume's.  Both layouts give identical render results. */
int spr_set_layout(spr_ctx *ctx, int layout);
int spr_get_layout(spr_ctx *ctx, int *layout);

/* Set / read a launch tuning value (SPR_TUNE_*). */
int spr_set_tuning(spr_ctx *ctx, int key, int value);
int spr_get_tuning(spr_ctx *ctx, int key, int *value);

/* set_data / update_data (volumerender.py:235-294): copy a (nz,ny,nx) C-order volume,
   x fastest, from host memory into the context's device layout.  The caller keeps
   ownership of `host`.  dtype is SPR_DTYPE_*. */
int spr_set_volume(spr_ctx *ctx, const void *host, int dtype, int nx, int ny, int nz);

/* The same from a device pointer on the context's device (device-to-device copy).  Waits
   for all work already submitted to the device (any stream) before copying, so a source
   still being produced on another stream is complete. */
int spr_set_volume_device(spr_ctx *ctx, const void *dev, int dtype, int nx, int ny, int nz);

/* GPU min/max of the resident volume (glwidget.py:328-344 _get_min_max). */
int spr_volume_minmax(spr_ctx *ctx, float *min_out, float *max_out);

/* update_matrices + scalar setters (volumerender.py:214-233, 296-316). */
int spr_set_params(spr_ctx *ctx, const spr_params *params);

/* render(method, numParts, currentPart) (volumerender.py:508-547) followed by the
   blocking .get() copies (:388-390, :502-506).  The context keeps w x h device output
   buffers between calls (reallocated, zero-filled, when w or h changes), so parts > 0
   of a multipass render merge into the previous contents with fmax
   (volume_kernel.cl:172-183).  max_project writes out/alpha and leaves the depth buffer
   untouched (volume_kernel.cl never writes d_depth_output); iso_surface writes all five.
   Any host pointer may be NULL to skip that copy.  normals is (h,w,3). */
int spr_render(spr_ctx *ctx, int method, int num_parts, int current_part, int w, int h,
               float *out, float *alpha, float *depth, float *normals, float *occlusion);

/* The same render into caller-owned device buffers, enqueued on `stream` (a hipStream_t;
   NULL = the context's stream) without host synchronisation.  depth, normals and
   occlusion may be NULL for SPR_METHOD_MAX_PROJECT; iso_surface needs all of them. */
int spr_render_device(spr_ctx *ctx, int method, int num_parts, int current_part, int w, int h,
                      float *d_out, float *d_alpha, float *d_depth, float *d_normals,
                      float *d_occlusion, void *stream);

/* ---- Resident volume table and batched frames (no counterpart in the reference, whose
   renderer holds one volume and draws one frame per render(); its many-frame paths are
   the keyframe record loop, spimagine/gui/keyframe_view.py:641-653, and the timepoint
   change, spimagine/gui/glwidget.py:372-374, which re-uploads the volume each time).
   spr_set_volume* fill slot 0, the volume spr_render / spr_render_device draw; slots >= 1
   are added and released here and drawn by spr_render_batch. */

/* Upload a (nz,ny,nx) volume (host or device source) into a new slot. */
int spr_volume_add(spr_ctx *ctx, const void *src, int src_on_device, int dtype, int nx, int ny,
                   int nz, int *slot);
/* A new slot holding the max filter of slot src_slot over (2 radius + 1)^3 voxels (edges
   clamped), computed on the device: the "dilated" volume whose coarse projection bounds
   the samples a coarse ray step skips (input of the learned reconstruction, spimlearn.h).
   radius in [1, 16]. */
int spr_volume_add_dilated(spr_ctx *ctx, int src_slot, int radius, int *slot);
/* A new slot holding the volume of slot src_slot pooled by 2 along each axis
   (ceil(n/2) voxels): mode 0 = mean of each 2x2x2 block (rounded to the voxel type),
   1 = max (edges clamped); one level of a mean / max pyramid, computed on the device (the
   coarse inputs of the learned reconstruction, spimlearn.h). */
int spr_volume_add_pooled(spr_ctx *ctx, int src_slot, int mode, int *slot);
/* A new channel-pair slot (SPR_DTYPE_U16X2, padded layout) holding both 2x poolings of
   the uint16 volume of slot src_slot: channel 0 = mode 0 (mean), channel 1 = mode 1 (max),
   each equal to spr_volume_add_pooled's volume.  spr_render_pair draws both at once. */
int spr_volume_add_pooled_pair(spr_ctx *ctx, int src_slot, int *slot);
/* Copy a slot's volume back to host memory as a (nz,ny,nx) C-order array (bytes must be
   nx ny nz x voxel size); the readback of update_data's image (volumerender.py:294). */
int spr_volume_read(spr_ctx *ctx, int slot, void *host, size_t bytes);
/* Free a slot >= 1 (waits for the device to finish work that may read it). */
int spr_volume_release(spr_ctx *ctx, int slot);
/* Voxel type, size and device bytes of a slot (0 = the renderer's volume). */
int spr_volume_info(spr_ctx *ctx, int slot, int *dtype, int *nx, int *ny, int *nz,
                    size_t *device_bytes);

/* One frame of a batch: a resident volume, the frame's inverse matrices (as update_matrices
   computes them for that volume and pose) and two w x h float32 device planes. */
typedef struct spr_job {
    int slot;
    float inv_p[16];
    float inv_m[16];
    float *d_out;
    float *d_alpha;
} spr_job;

/* Max projection (plain or opacity mode) of n frames in one persistent launch, with
   `params` (NULL: the spr_set_params state; its matrices are ignored, each job carries
   its own).  Frames share w, h and every parameter except the matrices; all their slots
   must use the padded layout and one voxel type.  Results equal n separate
   spr_render_device calls.  Returns once the batch is enqueued on `stream` (its job table
   is copied through pinned host memory the context keeps); synchronise the stream before
   reading the frames.  Calls from several host threads on different streams run
   concurrently on the device. */
int spr_render_batch(spr_ctx *ctx, const spr_params *params, const spr_job *jobs, int n, int w,
                     int h, int num_parts, int current_part, void *stream);

/* One frame of a mixed batch (spr_render_frames): any padded slot, its own frame size,
   box, transfer range, gamma and max_steps, and two w x h float32 device planes. */
typedef struct spr_frame {
    int slot;
    int w, h;
    int max_steps;
    float inv_p[16];
    float inv_m[16];
    float box[6];
    float min_val, max_val, gamma;
    float *d_out;
    float *d_alpha;
} spr_frame;

/* Max projections (alpha_pow 0) or opacity-mode projections (alpha_pow > 0) of n frames of
   any sizes and settings in one persistent launch: the continuous batching of a render
   server, which takes whatever requests are waiting.  The frames share the voxel type,
   alpha_pow and the filter; every slot uses the padded layout.  Each frame equals
   spr_render_slot with its own parameters (num_parts 1).  Returns once the batch is
   enqueued on `stream`, as spr_render_batch; calls from several host threads on different
   streams run concurrently. */
int spr_render_frames(spr_ctx *ctx, const spr_frame *frames, int n, float alpha_pow, int filter,
                      void *stream);

/* Plain max projections (alpha_pow 0, linear filter) of n frames whose rays lie far apart,
   such as a sparse grid of a frame's pixels (no counterpart in the reference, whose frames
   are always full pixel grids; here for the learned renderer's exact judge): each ray's
   samples are split into segments of `seg` samples (32, 64 or 128; 0 = 64) marched by
   different waves, so few rays still fill the device; a segment is staged through LDS
   slabs as in the single-frame kernel (SPR_TUNE_SLAB_KB > 0, 16-sample boxes) or gathered
   from global memory (SPR_TUNE_SLAB_KB 0).  Sample positions follow the
   reference's running sum exactly (volume_kernel.cl:121-131, 293-298), so each frame equals
   spr_render_frames / spr_render_slot of the same frame up to the sign of zero outputs.
   Slots as spr_render_frames (padded layout, one voxel type).  The frames' d_out planes
   are cleared and used as the accumulators of their rays.  d_gates: NULL, or n device
   flags: frame k is rendered only when d_gates[k] != 0 (read on the device when the
   launch runs; a closed frame's planes are left untouched), so a caller can render
   frames that an earlier kernel on the stream selects without a host round trip.
   Returns once enqueued on `stream`. */
int spr_render_frames_split(spr_ctx *ctx, const spr_frame *frames, int n, int seg,
                            const unsigned int *d_gates, void *stream);

/* One frame of any method from any slot into device buffers on `stream`, with explicit
   `params` (matrices included): the stateless form of spr_render_device for callers that
   keep many volumes resident and render with per-request settings from several threads. */
int spr_render_slot(spr_ctx *ctx, int slot, const spr_params *params, int method, int num_parts,
                    int current_part, int w, int h, float *d_out, float *d_alpha, float *d_depth,
                    float *d_normals, float *d_occlusion, void *stream);

/* spr_render_slot's max projection (num_parts 1) performed only if *d_flag (a device
   word) is nonzero when its kernels run, else leaving d_out and d_alpha untouched: a
   decision taken on the device by earlier work on `stream`, without a host round trip
   (the learned reconstruction's exact fallback, include/spimlearn.h). */
int spr_render_slot_if(spr_ctx *ctx, int slot, const spr_params *params, int w, int h,
                       float *d_out, float *d_alpha, const unsigned int *d_flag, void *stream);

/* Max projections (linear filter, alpha_pow 0) of both channels of a channel-pair slot
   with one set of rays: d_out holds two w x h planes (channel 0's, then channel 1's),
   d_alpha the alpha plane.  Each plane is bit-identical to spr_render_slot's max
   projection of a one-channel volume holding that channel (same staged sampling, same
   operation order); one box, one address and one LDS read per corner pair serve both.
   Enqueued on `stream`.  The coarse input of the learned reconstruction (spimlearn.h). */
int spr_render_pair(spr_ctx *ctx, int slot, const spr_params *params, int num_parts,
                    int current_part, int w, int h, float *d_out, float *d_alpha, void *stream);

/* Coarse max projections of two slots of one voxel type and shape (a volume and its
   spr_volume_add_dilated copy) with the same params, in one launch: d_out holds two w x h
   planes (slot_a's, then slot_b's), d_alpha slot_a's alpha.  Not the reference's sample
   sequence: sample k of a ray is at p0 + k dt d (one fma instead of a running sum), which
   lets four waves share a ray's samples; the last bits of a sample position can differ
   from spr_render_slot's.  The input of the learned reconstruction (spimlearn.h). */
int spr_render_coarse_pair(spr_ctx *ctx, int slot_a, int slot_b, const spr_params *params, int w,
                           int h, float *d_out, float *d_alpha, void *stream);

/* spr_render_batch without any device memory allocation (no counterpart in the reference;
   for callers that cannot allocate: code between hipMalloc-free checkpoints, tight serving
   loops, harnesses that forbid hipMalloc).  Same arguments, checks and results as
   spr_render_batch, which takes its job table and queue counters from the stream-ordered
   allocator; here they live in a fixed device scratch array (one per device, held until
   the process ends) of 256 + SPR_NOALLOC_MAX_JOBS x 256 bytes (~1 MiB), so n is limited
   to SPR_NOALLOC_MAX_JOBS (SPR_ERR_INVALID_ARG beyond).  The scratch keeps the last call's
   job table.  Calls share one host staging table and each device's scratch: the library
   runs them one at a time (concurrent callers wait), and each call blocks until its batch
   completes. */
#define SPR_NOALLOC_MAX_JOBS 4096
int spr_render_batch_noalloc(spr_ctx *ctx, const spr_params *params, const spr_job *jobs, int n,
                             int w, int h, int num_parts, int current_part, void *stream);

/* The same batch (num_parts = 1, the spr_set_params state) captured once in a HIP graph: replay it any number of
   times on any stream, rewrite its jobs (same count; waits for the last replay first) and
   destroy it.  Launches are asynchronous.  A graph records the slots its jobs read:
   releasing or refilling one of them (spr_volume_release, a spr_volume_add that reuses the
   slot, a spr_set_volume* that reallocates slot 0 or changes its shape, type or layout)
   makes spr_batch_graph_launch fail with SPR_ERR_INVALID_ARG until spr_batch_graph_update
   names live slots again.  An update must keep the voxel type and whether every volume's
   dimensions are powers of two (the captured kernel's texel rule). */
typedef struct spr_batch_graph spr_batch_graph;
int spr_batch_graph_create(spr_ctx *ctx, const spr_job *jobs, int n, int w, int h,
                           spr_batch_graph **graph);
int spr_batch_graph_update(spr_ctx *ctx, spr_batch_graph *graph, const spr_job *jobs, int n);
int spr_batch_graph_launch(spr_ctx *ctx, spr_batch_graph *graph, void *stream);
int spr_batch_graph_destroy(spr_ctx *ctx, spr_batch_graph *graph);

/* Wait for all work on the context's stream. */
int spr_synchronize(spr_ctx *ctx);

/* The context's hipStream_t (for callers that order their own work after a render). */
int spr_get_stream(spr_ctx *ctx, void **stream);

/* Number of reference samples a max_project frame takes with the current volume and
   params: (#rays hitting the box) x ((S/numParts)/16 + 1) x 16.  Used to report
   Gsamples/s; computed on the device with the kernels' own ray set-up. */
int spr_count_samples(spr_ctx *ctx, int num_parts, int w, int h, double *samples);

#ifdef __cplusplus
}
#endif

#endif /* SPIMRENDER_H */
