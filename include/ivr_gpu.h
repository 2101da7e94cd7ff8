/*
 * ivr_gpu.h -- C-ABI of the MI355X render path (libivr.so).
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes
 * (0 = ok, negative = error; ivr_last_error() gives the message of the
 * calling thread's last failure).  The caller owns host buffers; the library
 * owns device memory behind the handles.  All GPU work of one context is
 * ordered on that context's stream; ivr_frame_download() and ivr_ctx_sync()
 * are the synchronisation points.  One context is used by one thread at a
 * time.
 *
 * Reference interfaces replaced (file:line in the reference tree):
 *   ivr_render             createlocalimage(int)            main.c:937-1026
 *                          raycast/getsamplesonray/getlinearsample/
 *                          getrayintersect                   ray.c:59-318
 *                          transferfunc (CHEM, SEISMIC2, ...) transferfuncs.c:24-225
 *                          compositetiles + RGBA8 quantiser  main.c:1203-1218, 1404
 *   ivr_volume_upload      readvoldata/readfromfile + brick   main.c:629-706, 1220-1269,
 *                          decomposition (getlocalbrick)     utils.c:332-388
 *   ivr_tf_create          the colour LUT (lut, lutoffset)   main.c:1075-1200
 *   ivr_volume_histogram   histogram                         mathutil.c:444-470
 *   ivr_tf_colormap        the colormap builders             colormap.c:21-205
 *   ivr_volume_luts        initvolluts (histogram, colormap,  main.c:1075-1200
 *                          lutoffset, maxlutindx)
 *   ivr_state_orbit        the client's STATE producer       ivrclient.c:536-569
 *   ivr_volume_gradient    gradient3d (GRAD)                 mathutil.c:242-338,
 *                                                            main.c:1181-1196
 */
#ifndef IVR_GPU_H
#define IVR_GPU_H

#include <stddef.h>
#include <stdint.h>

#include "ivr_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define IVR_OK 0
#define IVR_ERR_ARG (-1)
#define IVR_ERR_HIP (-2)
#define IVR_ERR_NOMEM (-3)
#define IVR_ERR_UNSUPPORTED (-4)
#define IVR_ERR_BUDGET (-5)
#define IVR_ERR_IO (-6)

/* voxel storage types */
#define IVR_F32 0
#define IVR_F16 1

/* render modes */
#define IVR_MODE_COMPAT 0 /* exact ray.c semantics (dt = DELTASAMPLE t-units, 256-sample
                             cap, reference sampler, back-to-front transfer function,
                             projected-bbox pixel rectangle) */
#define IVR_MODE_CONFIG 1 /* step in voxels, no cap, true trilinear, front-to-back
                             compositing with early ray termination */

/* compositing of bricks */
#define IVR_FUSED 0           /* each ray walks all bricks front to back in one launch */
#define IVR_SORT_LAST 1       /* the reference's sort-last scheme: one partial RGBA per
                                 brick (front to back from zero, early termination
                                 inside the brick), combined far to near with "over";
                                 one workgroup per pixel block and brick run writes the
                                 partials to planes a second pass combines (IVR_SL_SPLIT=0
                                 in the environment: one launch, partials in LDS) */
#define IVR_SORT_LAST_TILES 2 /* the same images with the partials materialised as
                                 RGBA32F tiles in HBM (one launch per brick) and a
                                 separate "over" pass: the exchange form of a
                                 multi-GPU brick decomposition */

/* synthetic volumes */
#define IVR_GEN_BLOBS 0       /* sum of Gaussian blobs, min-max normalised to [0,1] */
#define IVR_GEN_NOISE_BLOBS 1 /* Perlin fBm + blobs, zeroed below a quantile, [0,1] */

typedef struct ivr_ctx ivr_ctx;
typedef struct ivr_volume ivr_volume;
typedef struct ivr_tf ivr_tf;
typedef struct ivr_frame ivr_frame;

typedef struct ivr_gen_params {
    int kind;          /* IVR_GEN_*                                          */
    int nblobs;        /* Gaussian blobs: centres U[0.2,0.8]^3, sigma U[0.05,0.15],
                          amplitude U[0.5,1]                                  */
    int octaves;       /* fBm octaves                                        */
    float lacunarity;  /* frequency ratio between octaves                    */
    float gain;        /* amplitude ratio between octaves                    */
    float base_freq;   /* octave-0 frequency, periods per volume edge        */
    float empty_frac;  /* quantile below which voxels are set to 0           */
    float value_scale; /* final values are multiplied by this (1 or 255)      */
    uint64_t seed;
} ivr_gen_params;

typedef struct ivr_frame_desc {
    ivr_state state;   /* camera, view volume, cut planes, alpha, width, height */
    int mode;          /* IVR_MODE_*                                          */
    int xfer;          /* compat: transfer function id (transferfuncs.h:8-13) */
    float step_vox;    /* config: sample spacing in voxels                    */
    float ert;         /* config: stop once accumulated alpha >= ert (>1: off) */
    float lutscale;    /* LUT index = (int)(s * lutscale) - lutoffset          */
    int lutoffset;
    int maxlutindx;    /* SEISMIC2 upper clamp                                 */
    int skip;          /* config: macrocell empty-space skipping on/off        */
    int composite;     /* IVR_FUSED, IVR_SORT_LAST or IVR_SORT_LAST_TILES       */
} ivr_frame_desc;

typedef struct ivr_frame_stats {
    unsigned long long samples; /* sample positions marched, those skipped in
                                   transparent macrocells included              */
    unsigned long long visited; /* samples classified: voxels fetched (after skip) */
    float render_ms;            /* device time of the last render's kernel(s), after
                                   the macrocell pass (events)                  */
    float composite_ms;         /* device time of its sort-last combine pass    */
    unsigned long long partials; /* IVR_SORT_LAST: per-run partials written to HBM */
    unsigned long long evaluations; /* surrogate frames: network lane slots evaluated
                                       (64 per wavefront step that ran the network,
                                       idle lanes included); visited / evaluations is
                                       the network's lane utilisation              */
} ivr_frame_stats;

/* ---- context ---- */
int ivr_ctx_create(int device, ivr_ctx **out);
int ivr_ctx_destroy(ivr_ctx *ctx);
int ivr_ctx_sync(ivr_ctx *ctx);
/* record event `slot` (0..15) on the context stream; elapsed ms between two */
int ivr_ctx_record(ivr_ctx *ctx, int slot);
int ivr_ctx_elapsed(ivr_ctx *ctx, int slot0, int slot1, float *ms);
/* bytes of device memory held by the context's objects */
unsigned long long ivr_ctx_device_bytes(const ivr_ctx *ctx);
const char *ivr_last_error(void);

/* ---- volumes: nx*ny*nz scalars, stored as bricks of brick^3 voxels plus a
 * one-voxel ghost layer on every side, bricks in Morton order ---- */
int ivr_volume_create(ivr_ctx *ctx, int nx, int ny, int nz, int dtype, int brick,
                      ivr_volume **out);
int ivr_volume_upload(ivr_ctx *ctx, ivr_volume *vol, const float *host_xfastest);
int ivr_volume_generate(ivr_ctx *ctx, ivr_volume *vol, const ivr_gen_params *gp);
int ivr_volume_download(ivr_ctx *ctx, const ivr_volume *vol, float *host_xfastest);
int ivr_volume_destroy(ivr_ctx *ctx, ivr_volume *vol);
/* Gradient field by central differences (replaces gradient3d, reference
 * mathutil.c:242-328, and the GRAD block of initvolluts, main.c:1181-1196,
 * whose volume.gradient it fills): per voxel (d/dx, d/dy, d/dz, magnitude)
 * as 4 floats, x fastest, nx*ny*nz*16 bytes.  (f[+1] - f[-1]) / 2 inside,
 * (f[1] - f[0]) / 2 and (f[n-1] - f[n-2]) / 2 on the faces, per axis.  dst is
 * device memory when dst_on_device (stream-ordered, no sync), else host
 * memory (returns when filled). */
int ivr_volume_gradient(ivr_ctx *ctx, const ivr_volume *vol, float *dst, int dst_on_device);
/* dims[3], nbricks[3], brick edge, dtype, device bytes */
int ivr_volume_info(const ivr_volume *vol, int dims[3], int nbricks[3], int *brick,
                    int *dtype, unsigned long long *bytes);

/* ---- transfer function LUT (n RGBA entries, n <= 4096) ---- */
int ivr_tf_create(ivr_ctx *ctx, const float *rgba, int n, ivr_tf **out);
int ivr_tf_update(ivr_ctx *ctx, ivr_tf *tf, const float *rgba, int n);
int ivr_tf_destroy(ivr_ctx *ctx, ivr_tf *tf);

/* ---- colour tables built on the device from the volume ---- */
/* histogram (reference mathutil.c:444-470) of the volume's values: vmin and
 * vmax (NaN ignored, as fmin/fmax do), hsize = (int)((vmax - vmin) + 1) and,
 * when hist != NULL and hsize <= cap, hist[i] = the number of voxels v with
 * (int)(v - vmin) == i */
int ivr_volume_histogram(ivr_ctx *ctx, const ivr_volume *vol, float *vmin, float *vmax, int *hsize, int *hist,
                         int cap);
/* colormap builder `id` (colormap.c:21-181, colormap.h:9-14: BASE 0,
 * RAINBOW 1, MEDICAL 2, SEISMIC 3, JET 4; NOMAP -1 leaves the entries as
 * they are) of hsize <= 4096 entries over the value range [hmin, hmax] into
 * tf, computed on the device with the reference's arithmetic; tf then holds
 * hsize entries, those the builder leaves unwritten zero */
int ivr_tf_colormap(ivr_ctx *ctx, ivr_tf *tf, int id, int hsize, float hmin, float hmax);
/* the LUT's first n entries as RGBA floats into host memory */
int ivr_tf_download(ivr_ctx *ctx, const ivr_tf *tf, float *rgba, int n);
/* initvolluts (main.c:1075-1200) as the render server runs it: the
 * histogram of the volume, colormap `id` over its [min, max] into tf (hsize
 * entries), *lutoffset = (int)(min + 0.5), *maxlutindx = hsize - 1 (main.c:1158);
 * hsize_out may be NULL.  IVR_ERR_UNSUPPORTED when hsize exceeds 4096. */
int ivr_volume_luts(ivr_ctx *ctx, const ivr_volume *vol, int colormap_id, ivr_tf *tf, int *lutoffset,
                    int *maxlutindx, int *hsize_out);

/* ---- frames: device RGBA8 image (+ optional RGBA32F), row 0 = bottom ---- */
int ivr_frame_create(ivr_ctx *ctx, int width, int height, int keep_rgba32, ivr_frame **out);
int ivr_frame_destroy(ivr_ctx *ctx, ivr_frame *fr);
/* waits for the frame's work; either pointer may be NULL */
int ivr_frame_download(ivr_ctx *ctx, ivr_frame *fr, unsigned char *rgba8, float *rgba32);
int ivr_frame_stats_get(ivr_ctx *ctx, ivr_frame *fr, ivr_frame_stats *st);

/* ---- render: enqueue one frame on the context stream ---- */
int ivr_render(ivr_ctx *ctx, const ivr_volume *vol, const ivr_tf *tf,
               const ivr_frame_desc *desc, ivr_frame *fr);

/* ivr_render into caller device memory, allocating nothing.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * d_rgba8 (W*H*4 bytes) and d_rgba32 (W*H*16 bytes, or NULL) are device
 * buffers; W, H come from desc->state.  The frame counters and the sort-last
 * block maxima, run counts and partial planes live in a static device array
 * of 72 MiB per GPU that the library holds until the process ends: IVR_FUSED
 * frames of any size fit, IVR_SORT_LAST frames need
 * runs*W*H*16 + W*H*4 bytes plus a few KiB (runs = sum of bricks per axis - 2;
 * 1024x1024 with 2x2x2 bricks fits), larger ones fail with IVR_ERR_ARG, and
 * IVR_SORT_LAST_TILES is IVR_ERR_UNSUPPORTED.  The counters are cleared by a
 * kernel, no event is recorded and no statistics are kept.  Calls on one GPU
 * share the array: calls that overlap on one GPU (other contexts' streams,
 * other threads) race and must be serialised by the caller. */
int ivr_render_noalloc(ivr_ctx *ctx, const ivr_volume *vol, const ivr_tf *tf,
                       const ivr_frame_desc *desc, unsigned char *d_rgba8, float *d_rgba32);

/* ---- learned surrogate: per-brick multiresolution hash grids (16 levels
 * x 2 features) and one MLP 32 -> 64 -> 64 -> 1 on bf16 MFMA, trained on
 * samples of a plain volume and rendered in place of its trilinear fetch.
 * No counterpart in the reference. ---- */
typedef struct ivr_surrogate ivr_surrogate;
typedef struct ivr_surrogate_params {
    int log2_table;   /* hash-table entries per level and brick = 2^log2_table */
    int base_res;     /* level-0 resolution per brick edge                      */
    float growth;     /* resolution ratio between levels                        */
    float lr;         /* Adam step size of the hash tables                      */
    float lr_mlp;     /* Adam step size of the MLP                              */
    float beta1, beta2, eps;
    float init_scale; /* tables start U[-init_scale, init_scale]                */
    uint64_t seed;
    /* a surrogate brick (one hash grid) spans 2^log2_group volume bricks per
     * axis, in [0, 4]: 0 gives one grid per volume brick, larger values a
     * representation 8^log2_group times smaller over the same volume */
    int log2_group;
} ivr_surrogate_params;

void ivr_surrogate_defaults(ivr_surrogate_params *sp);
int ivr_surrogate_create(ivr_ctx *ctx, const ivr_volume *vol, const ivr_surrogate_params *sp,
                         ivr_surrogate **out);
int ivr_surrogate_destroy(ivr_ctx *ctx, ivr_surrogate *s);
/* `steps` Adam steps of `samples` samples whose targets are the trilinear
 * values of `vol`; loss_out = last MSE.  The samples of a step go to the
 * bricks in equal contiguous ranges, each uniform in its brick (stratified
 * in tiles of 64), and are processed in chunks of up to 2^22 on the
 * context's stream and one more stream of the surrogate's own; the first
 * call allocates two chunk buffers of encoding gradients (128 B per sample
 * of a chunk each), which ivr_surrogate_freeze releases. */
int ivr_surrogate_train(ivr_ctx *ctx, ivr_surrogate *s, const ivr_volume *vol, int steps,
                        long long samples, float *loss_out);
/* gradients of one batch of caller samples (local xyz in [0,1]^3, brick id,
 * target), without an update: outputs y, MSE, MLP gradient (6337 floats in
 * the W1 b1 W2 b2 W3 b3 order) and one brick's table gradient */
int ivr_surrogate_grad(ivr_ctx *ctx, ivr_surrogate *s, int n, const float *xyz, const int *bricks,
                       const float *targets, float *y_out, float *loss_out, float *gmlp_out,
                       float *gtable_out, int gtable_brick);
/* one Adam update from the gradients of the last ivr_surrogate_grad */
int ivr_surrogate_adam(ivr_ctx *ctx, ivr_surrogate *s);
int ivr_surrogate_get(ivr_ctx *ctx, const ivr_surrogate *s, float *mlp, float *table, int brick);
int ivr_surrogate_set(ivr_ctx *ctx, ivr_surrogate *s, const float *mlp, const float *table, int brick);
int ivr_surrogate_info(const ivr_surrogate *s, int *nbricks, unsigned long long *table_entries,
                       int *step, int res[16]);
/* render a config-mode frame from the surrogate; bricks flagged for
 * fallback fetch `vol` trilinearly instead (vol may be NULL if none is) */
int ivr_surrogate_render(ivr_ctx *ctx, const ivr_surrogate *s, const ivr_volume *vol,
                         const ivr_tf *tf, const ivr_frame_desc *desc, ivr_frame *fr);
/* per-brick image PSNR of the surrogate against the plain render of each
 * brick alone (RGBA8, dB); bricks below threshold_db fall back */
int ivr_surrogate_evaluate(ivr_ctx *ctx, ivr_surrogate *s, const ivr_volume *vol, const ivr_tf *tf,
                           const ivr_frame_desc *desc, float threshold_db, float *psnr_out,
                           int *nfallback);
int ivr_surrogate_set_fallback(ivr_ctx *ctx, ivr_surrogate *s, const unsigned char *flags);
/* the network value the renderer composites (encoding from the half-precision
 * tables, bf16 MFMA MLP: the device function of ivr_surrogate_render) at n
 * caller points: local xyz in [0,1]^3 of brick bricks[i]; y_out host, n floats */
int ivr_surrogate_eval_points(ivr_ctx *ctx, const ivr_surrogate *s, int n, const float *xyz, const int *bricks,
                              float *y_out);
/* versioned binary checkpoint of parameters, Adam state and the 4^3-voxel
 * macrocell ranges (format 5; files of formats 1-4 load too, their ranges of
 * 16^3-voxel cells are dropped for the ones computed at create) */
int ivr_surrogate_save(ivr_ctx *ctx, const ivr_surrogate *s, const char *path);
int ivr_surrogate_load(ivr_ctx *ctx, ivr_surrogate *s, const char *path);
/* keep only what rendering reads: the half-precision tables, the MLP and the
 * 4^3-voxel macrocell ranges (2 bytes per table parameter, 4 per 64
 * voxels); the fp32 tables, their
 * gradients and Adam moments (16 bytes per parameter) and the training
 * buffers and stream are released.  Render,
 * evaluate, eval_points, set_fallback and info work as before; train, grad,
 * adam, get, set, save and load then fail with IVR_ERR_UNSUPPORTED. */
int ivr_surrogate_freeze(ivr_ctx *ctx, ivr_surrogate *s);

/* ---- steering-request batcher: continuous batching of render requests of
 * any size, camera, transfer function and resident volume into persistent
 * launches.  Requests are admitted into slots while a slot is free and the
 * HBM budget (resident volumes + per-request frame, LUT and macrocell
 * bytes) allows; up to two batches are in flight on two streams, so work
 * admitted while a batch runs overlaps it on the device.  Replaces the
 * one-frame-at-a-time loop of the render server (main.c:206-250). ---- */
typedef struct ivr_batcher ivr_batcher;
typedef struct ivr_request {
    unsigned long long id;  /* caller's request id                              */
    int volume;             /* index from ivr_batcher_add_volume / _add_surrogate */
    ivr_frame_desc desc;    /* config-mode frame (state carries camera and size) */
    const float *lut;       /* RGBA LUT of lutsize <= 256 entries, copied at submit */
    int lutsize;
    unsigned char *rgba8;   /* host destination, winwidth*winheight*4 bytes      */
    double t_submit;        /* ivr_now() clock, seconds                          */
} ivr_request;
typedef struct ivr_completion {
    unsigned long long id;
    int status;             /* 0 ok, negative: rejected (error code)             */
    double t_submit, t_admit, t_done;
    unsigned long long samples;
} ivr_completion;
typedef struct ivr_batcher_stats {
    unsigned long long submitted, admitted, completed, rejected, batches;
    unsigned long long bytes_in_use, budget;
    int pending, in_flight;
} ivr_batcher_stats;

int ivr_batcher_create(ivr_ctx *ctx, unsigned long long hbm_budget, int max_slots, int max_width,
                       int max_height, ivr_batcher **out);
int ivr_batcher_destroy(ivr_batcher *b);
/* make a resident volume servable; returns its index (>= 0) or an error */
int ivr_batcher_add_volume(ivr_batcher *b, const ivr_volume *vol);
/* make a trained surrogate servable: requests with the returned index are
 * rendered from the network (the surrogate requests of a batch share one
 * launch on its stream, beside the batch kernel), bit-identical to
 * ivr_surrogate_render.  The
 * per-brick fallback flags are copied now; vol (kept alive by the caller,
 * and charged to the budget only when it is registered itself) must be
 * given when a flag is set.  No counterpart in the reference. */
int ivr_batcher_add_surrogate(ivr_batcher *b, const ivr_surrogate *sg, const ivr_volume *vol);
int ivr_batcher_submit(ivr_batcher *b, const ivr_request *req);
/* retire finished batches, admit pending requests, launch a batch; with
 * block != 0 waits for the oldest batch when nothing else can proceed;
 * returns the number of completions ready to poll */
int ivr_batcher_progress(ivr_batcher *b, int block);
int ivr_batcher_poll(ivr_batcher *b, ivr_completion *out, int max);
int ivr_batcher_stats_get(const ivr_batcher *b, ivr_batcher_stats *st);
/* monotonic clock of the batcher timestamps, seconds */
double ivr_now(void);

/* ---- host-side helpers (no GPU) ---- */
/* Defaults for a config-mode frame of the given state. */
void ivr_desc_config(ivr_frame_desc *d);
/* Orbit camera: STATE of a client looking at the origin from distance 5 after
 * rotating the scene by azimuth/elevation drawn from `seed` (radians). */
void ivr_state_orbit(uint64_t seed, int width, int height, ivr_state *st,
                     float *az_out, float *el_out);
void ivr_state_from_angles(float az, float el, int width, int height, ivr_state *st);
/* 256-entry piecewise-linear RGBA LUT from 5 control points drawn from `seed`;
 * alpha is 0 at value 0 and U[0, alpha_max] elsewhere.  When jitter > 0 the
 * control points are perturbed by U[-jitter, jitter] drawn from jitter_seed. */
void ivr_tf_build(uint64_t seed, float alpha_max, float jitter, uint64_t jitter_seed,
                  float *lut256);
/* CPU evaluation of the synthetic generator, bit-identical to the device one
 * (same code, same operation order); out is x-fastest. */
int ivr_generate_host(const ivr_gen_params *gp, int nx, int ny, int nz, float *out);
/* STATE <-> its 184 wire bytes: the three ints in network byte order, the
 * floats native little-endian (connection.c:167-200) */
void ivr_state_pack(const ivr_state *st, unsigned char out[184]);
void ivr_state_unpack(const unsigned char in[184], ivr_state *st);
/* byte-swap the STATE's floats, as recvstate does in a build with BIGENDIAN
 * (connection.c:191-197) for a peer of the other byte order */
void ivr_state_swap_floats(ivr_state *st);
/* writetotiff (utils.c:18-154): a 512-byte big-endian TIFF header and
 * w x h x planes 8-bit samples (planes 1: red; 3: RGB) of an RGBA32F image
 * in its row order, each value * 255 truncated, into out (cap bytes);
 * returns the bytes written (512 + w*h*planes) or -1 */
long long ivr_tiff_encode(const float *rgba32, int w, int h, int planes, unsigned char *out, size_t cap);
/* deterministic PRNG used by every generator: SplitMix64-seeded PCG32 */
uint32_t ivr_rng_u32(uint64_t *state);

#ifdef __cplusplus
}
#endif
#endif
