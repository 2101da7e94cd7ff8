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
#define IVR_MIP_LEVELS 3 /* coarser transparent nodes: 2^3, 4^3, 8^3 macrocells */
#define IVR_MAX_LUT 4096
#define IVR_MAX_BRICKS 4096
#define IVR_MAX_STEPS (1 << 22) /* per-ray sample bound */
#define IVR_BLOCK 256 /* threads of the batch and generation kernels */
#ifndef IVR_FRAME_BLOCK
#define IVR_FRAME_BLOCK 256 /* threads of a frame kernel's workgroup (one pixel block) */
#endif
#define IVR_SL_LDS (64 * IVR_FRAME_BLOCK) /* LDS for sort-last partials: 4 runs */
/* frame counters: IVR_CTR_SLOTS stripes of IVR_CTRS u64 (samples, visited,
 * partials, network lane slots evaluated), a workgroup adding to stripe blockIdx % IVR_CTR_SLOTS;
 * one address taking every wavefront's atomic was measured to slow a frame
 * by up to 40 % */
#define IVR_CTR_SLOTS 256
#define IVR_CTRS 4
#define IVR_CTR_BYTES (IVR_CTR_SLOTS * IVR_CTRS * 8)

int ivr_fail(int code, const char *fmt, ...);
#define IVR_HIP(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return ivr_fail(IVR_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                               \
    } while (0)

/* workgroup -> XCD chunking of the frame kernels (ivr_ctx::xcd_chunk) */
#define IVR_CHUNK_ROW (-1)  /* one row of 16x16 blocks per chunk            */
#define IVR_CHUNK_BAND (-2) /* one contiguous eighth of the blocks per XCD  */

struct ivr_ctx {
    int device = 0;
    /* launch-shape tuning, read from IVR_XCD_CHUNK / IVR_WAVE_W / IVR_BLOCK_W at creation
     * (results are identical for every setting) */
    int xcd_chunk = IVR_CHUNK_ROW;
    /* 0: by the volume's storage type, 16 and 16 for fp32, 32 and 64 for fp16 */
    int wave_w = 0;
    int block_w = 0; /* pixel block of a workgroup: block_w x (256 / block_w) */
    int sl_split = 1; /* sort-last as one workgroup per (pixel block, run) */
    int sg_helpers = 1; /* surrogate frames: lanes of ended rays evaluate live rays' next samples */
    /* this device's copy of the library's static scratch array (ivr_render_noalloc),
     * found by hipGetSymbolAddress when the context is created */
    unsigned char *scratch = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev[16] = {};
    unsigned long long bytes = 0;
};

int ivr_dalloc(ivr_ctx *ctx, void **p, size_t bytes);
void ivr_dfree(ivr_ctx *ctx, void *p, size_t bytes);

/* a temporary device buffer, freed when the scope ends (on error returns too) */
struct ivr_dscope {
    ivr_ctx *ctx;
    void *p = nullptr;
    size_t bytes = 0;
    explicit ivr_dscope(ivr_ctx *c) : ctx(c) {}
    ivr_dscope(const ivr_dscope &) = delete;
    ivr_dscope &operator=(const ivr_dscope &) = delete;
    int alloc(size_t n)
    {
        const int rc = ivr_dalloc(ctx, &p, n);
        if (!rc) bytes = n;
        return rc;
    }
    ~ivr_dscope()
    {
        if (p) ivr_dfree(ctx, p, bytes);
    }
};

struct ivr_volume {
    int n[3] = {0, 0, 0};
    int dtype = IVR_F32;
    int B = 0, bshift = 0, P = 0;      /* brick edge, log2, padded edge B+2   */
    int nb[3] = {0, 0, 0};              /* bricks per axis owning samples      */
    int nbricks = 0;
    size_t stride = 0;                  /* elements per brick, brick_stride(P) */
    size_t bytes = 0;
    void *data = nullptr;               /* all bricks, Morton order            */
    void *pairs = nullptr;              /* their row-pair copy (pair_elem), 2 x bytes */
    std::vector<unsigned long long> h_boff; /* linear brick id -> element offset */
    unsigned long long *d_boff = nullptr;
    int nc[3] = {0, 0, 0};              /* macrocells per axis                  */
    float2 *d_cellrange = nullptr;      /* per-cell value min/max (with footprint) */
    /* per-cell transparency code of the current frame: 0 = some sample may
     * be visible, else 1 + the largest level L <= IVR_MIP_LEVELS whose node
     * of (2^L)^3 cells around the cell is all transparent */
    unsigned char *d_cellempty = nullptr;
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
    unsigned long long *d_counters = nullptr; /* IVR_CTR_SLOTS stripes of IVR_CTRS */
    float4 *d_tiles = nullptr;                /* sort-last per-brick tiles    */
    size_t tiles_cap = 0;                     /* float4 elements              */
    unsigned *d_runcount = nullptr;           /* per-run sort-last: runs per ray, W*H */
    int *d_blockmax = nullptr;                /* and the most runs of a ray per pixel block */
    size_t blockmax_cap = 0;                  /* ints */
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
    const void *pair_data; /* the row-pair copy of the bricks (pair_elem) */
    unsigned pair_bytes;   /* its bytes when below 4 GiB - 1 MiB (buffer loads), else 0 */
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
    int xcd_chunk;  /* blocks per XCD chunk, 0 = hardware order */
    int wave_shift; /* log2 of a wavefront's pixel width          */
    int blk_shift;  /* log2 of a workgroup block's pixel width    */
    /* sort-last tile pass */
    int tb[3];
    int trect[4]; /* x0, y0, w, h */
    float4 *tile;
    unsigned *runcount; /* k_runs: runs per ray (W x H), or null: every run plane written */
    int *blockmax;      /* with run counts: the most runs of a ray of each pixel block */
};

/* Voxel order inside a brick: storage coordinates (x,y,z) in [0, P), P =
 * B+2, are the brick's floor indices plus one with the ghost layer at 0 and
 * B+1, in rows of P elements, x fastest; a brick holds P^3 elements rounded
 * up to 64, the tail being padding (zero, never read).  A 4x4x4-blocked
 * order halves the HBM bytes per frame but was measured slower for the
 * render kernels on gfx950, whose cost is in the L1 gathers, not in HBM. */
__host__ __device__ static inline size_t brick_stride(int P)
{
    return ((size_t)P * P * P + 63) & ~(size_t)63;
}

__host__ __device__ static inline unsigned brick_elem(int P, int x, int y, int z)
{
    return ((unsigned)z * (unsigned)P + (unsigned)y) * (unsigned)P + (unsigned)x;
}

/* Row-pair copy of the bricks, read by the config-mode marches: brick b
 * starts at element 2 * boff[b] (twice the brick storage, so no table of
 * its own), and its element 2 (x + P (y + (P-1) z)) + dy holds the voxel at
 * storage coordinates (x, y + dy, z), y < P-1.  The 8 voxels of a footprint
 * (x..x+1, y..y+1, z..z+1) are then 4 consecutive elements in each of two
 * slices: two loads instead of the four row loads of the brick storage. */
__host__ __device__ static inline unsigned pair_elem(int P, int x, int y, int z)
{
    return 2u * (((unsigned)z * (unsigned)(P - 1) + (unsigned)y) * (unsigned)P + (unsigned)x);
}

/* inverse of brick_elem (padding elements decode to z >= P) */
__host__ __device__ static inline void brick_coords(int P, unsigned e, int l[3])
{
    l[0] = (int)(e % P);
    l[1] = (int)((e / P) % P);
    l[2] = (int)(e / ((unsigned)P * P));
}

/* kernel arguments of one frame (api.hip); f may be null for per-ray calls */
void ivr_fill_args(RenderArgs &a, const ivr_cam &cam, const ivr_volume *v, const ivr_tf *tf,
                   const ivr_frame_desc *d, ivr_frame *f);

/* grid of pixel blocks (block_w x 256/block_w) over a w x h frame with the context's chunking;
 * a context left at the default shape takes the measured best one of the kernel kind: for the
 * plain renderer by a.dtype_f16, for the surrogate 16 x 16 blocks of 8 x 8 */
#define IVR_GRID_RENDER 0
#define IVR_GRID_SURROGATE 1
void ivr_set_grid(RenderArgs &a, const ivr_ctx *ctx, int w, int h, int kind = IVR_GRID_RENDER);

/* launchers (render.hip) */
int ivr_launch_render(ivr_ctx *ctx, const ivr_volume *vol, RenderArgs &a, int skip, int sort_last);
int ivr_launch_cells(ivr_ctx *ctx, const ivr_volume *vol, const RenderArgs &a);
/* transparency codes (see ivr_volume::d_cellempty) of the frame's LUT into out, ncells bytes */
int ivr_launch_cells_on(hipStream_t stream, const ivr_volume *vol, const RenderArgs &a, unsigned char *out);
/* the same from half-precision ranges of nc[0] x nc[1] x nc[2] cells */
int ivr_launch_cells_half(hipStream_t stream, const __half2 *range, const int nc[3], const RenderArgs &a,
                          unsigned char *out);
int ivr_launch_tile(ivr_ctx *ctx, const ivr_volume *vol, RenderArgs &a);
/* sort-last with one workgroup per (pixel block, run): partial planes into a.tile, then combined */
int ivr_launch_runs(ivr_ctx *ctx, const ivr_volume *vol, RenderArgs &a, int skip);
/* a.blockmax of every pixel block (a.nblk_pad ints) ahead of k_runs with run counts */
int ivr_launch_blockruns(ivr_ctx *ctx, RenderArgs &a);
int ivr_launch_combine_runs(ivr_ctx *ctx, const RenderArgs &a, int runs);
int ivr_launch_composite(ivr_ctx *ctx, ivr_frame *fr, int ntiles);
/* n u64 words set to zero by a kernel (no memset node, so callers may capture it) */
int ivr_launch_zero_u64(hipStream_t stream, unsigned long long *p, size_t n);

/* colormap builder `id` (luts.hip) over the current n entries of a device LUT */
int ivr_colormap_apply(hipStream_t st, float4 *d_lut, int id, int n, float hmin, float hmax);

/* volume kernels (volume.hip) */
int ivr_volume_cells_build(ivr_ctx *ctx, ivr_volume *vol);
/* value range (with the trilinear footprint) of every cell of 2^shift voxels
 * per axis, nc cells per axis ((n - 2) >> shift) + 1, into out (device) */
int ivr_volume_cellrange(ivr_ctx *ctx, const ivr_volume *vol, int shift, const int nc[3], float2 *out);

#endif
