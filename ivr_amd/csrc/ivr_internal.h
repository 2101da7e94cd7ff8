/*
 * ivr_internal.h -- objects and kernel arguments shared by the HIP sources
 * of libivr.so (internal header).
 */
#ifndef IVR_INTERNAL_H
#define IVR_INTERNAL_H

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <string>
#include <vector>

#include "ivr_host.h"

#define IVR_CELL_SHIFT 4 /* macrocells of 16^3 voxels */
#define IVR_MAX_LUT 4096
#define IVR_MAX_BRICKS 4096
#define IVR_MAX_STEPS (1 << 22) /* per-ray sample bound */
#define IVR_BLOCK 256 /* 4 wavefronts = 2x2 tiles of 8x8 pixels */

int ivr_fail(int code, const char *fmt, ...);
#define IVR_HIP(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ivr_fail(IVR_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                               \
    } while (0)

struct ivr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[16] = {};
    unsigned long long bytes = 0;
};

int ivr_dalloc(ivr_ctx *ctx, void **p, size_t bytes);
void ivr_dfree(ivr_ctx *ctx, void *p, size_t bytes);

struct ivr_volume {
    int n[3] = {0, 0, 0};
    int dtype = IVR_F32;
    int B = 0, bshift = 0, P = 0;      /* brick edge, log2, padded edge B+2   */
    int nb[3] = {0, 0, 0};              /* bricks per axis owning samples      */
    int nbricks = 0;
    size_t stride = 0;                  /* elements per brick                  */
    size_t bytes = 0;
    void *data = nullptr;               /* all bricks, Morton order            */
    std::vector<unsigned long long> h_boff; /* linear brick id -> element offset */
    unsigned long long *d_boff = nullptr;
    int nc[3] = {0, 0, 0};              /* macrocells per axis                  */
    float2 *d_cellrange = nullptr;      /* per-cell value min/max (with footprint) */
    unsigned char *d_cellempty = nullptr; /* per-cell "all samples transparent"  */
    size_t ncells = 0;
};

struct ivr_tf {
    int n = 0;
    float4 *d_lut = nullptr;
};

struct ivr_frame {
    int W = 0, H = 0;
    uchar4 *d_rgba8 = nullptr;
    float4 *d_rgba32 = nullptr;
    unsigned long long *d_counters = nullptr; /* [0] composited, [1] fetched */
    float4 *d_tiles = nullptr;                /* sort-last per-brick tiles    */
    size_t tiles_cap = 0;                     /* float4 elements              */
    int *d_tilemeta = nullptr;                /* sort-last tile table (device) */
    void *h_tilemeta = nullptr;               /* pinned staging of that table  */
    hipEvent_t e[3] = {};
    int timed = 0;
    int composited = 0;
};

/* per-brick tile of the sort-last scheme */
struct TileInfo {
    int x0, y0, w, h;     /* pixel rectangle               */
    int bx, by, bz;       /* brick                         */
    int order;            /* rank in front-to-back order   */
    long long offset;     /* float4 offset in d_tiles      */
};

struct RenderArgs {
    int W, H;
    int rect[4];
    int mode, xfer;
    float o[3], M[9], m5[3];
    float l, b, du, dv;
    float d0[3], di[3], dj[3];
    float bmin[3], bmax[3], ext[3], s[3], nm1f[3];
    float vox;
    float dconst[6];
    int n[3], nb[3];
    int bshift, bmask, P;
    int nbricks;
    const void *vox_data;
    int dtype_f16;
    const unsigned long long *boff;
    const unsigned char *cellempty;
    int nc[3];
    const float4 *lut;
    int lutsize, lutoffset, maxlutindx;
    float lutscale, alpha, step_vox, ert;
    uchar4 *out8;
    float4 *out32;
    unsigned long long *counters;
    int nblk_x, nblk, nblk_pad;
    /* sort-last tile pass */
    int tb[3];
    int trect[4]; /* x0, y0, w, h */
    float4 *tile;
};

/* kernel arguments of one frame (api.hip); f may be null for per-ray calls */
void ivr_fill_args(RenderArgs &a, const ivr_cam &cam, const ivr_volume *v, const ivr_tf *tf,
                   const ivr_frame_desc *d, ivr_frame *f);

/* launchers (render.hip) */
int ivr_launch_render(ivr_ctx *ctx, const ivr_volume *vol, RenderArgs &a, int skip);
int ivr_launch_cells(ivr_ctx *ctx, const ivr_volume *vol, const RenderArgs &a);
int ivr_launch_cells_on(hipStream_t stream, const ivr_volume *vol, const RenderArgs &a, unsigned char *out);
int ivr_launch_tile(ivr_ctx *ctx, const ivr_volume *vol, RenderArgs &a);
int ivr_launch_composite(ivr_ctx *ctx, ivr_frame *fr, int ntiles);

/* volume kernels (volume.hip) */
int ivr_volume_cells_build(ivr_ctx *ctx, ivr_volume *vol);

#endif
