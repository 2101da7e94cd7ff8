/*
 * render.hip -- gfx950 render kernels of libivr.so.
 *
 * One lane per ray; a 256-thread workgroup is 2x2 wavefront tiles of 8x8
 * pixels, so the 64 rays of a wavefront are spatial neighbours and their
 * 8-voxel gathers hit the same cache lines.  Workgroups are remapped so that
 * each of the 8 XCDs (which receive workgroups round-robin) renders one
 * contiguous band of the image and its L2 holds that band's bricks.  The
 * transfer-function LUT and the brick offset table are staged in LDS.
 *
 * Kernels and the reference code they replace:
 *   k_render<COMPAT>  createlocalimage pixel loop + raycast + getsamplesonray +
 *                     getlinearsample + transfer function, bit-exact with the
 *                     reference's float order (main.c:983-1009, ray.c:59-318,
 *                     transferfuncs.c:24-199)
 *   k_render<CONFIG>  the same path with step in voxels, true trilinear, FTB
 *                     compositing, early ray termination and 16^3 macrocell
 *                     empty-space skipping; all bricks in one launch
 *   k_tile            one brick's partial image (sort-last, main.c:708-934)
 *   k_composite       compositetiles + RGBA8 quantiser (main.c:1203-1218, 1404)
 *   k_cells           per-frame macrocell transparency from the LUT
 *
 * This file is compiled with -ffp-contract=off: every fused multiply-add is
 * an explicit fmaf, as in the CPU restatement, so results match it bit for
 * bit.
 */
#include <algorithm>

#include "render_dev.h"

namespace {

using namespace ivr_dev;

/* SL: 0 fused, 1 sort-last with the partials in LDS, 2 sort-last walking
 * the runs from the far end */
template <int MODE, typename VT, bool SKIP, int SL>
__global__ __launch_bounds__(IVR_FRAME_BLOCK) void k_render(const RenderArgs a)
{
    extern __shared__ float4 smem[];
    float4 *s_lut = smem;
    unsigned long long *s_boff = (unsigned long long *)(smem + a.lutsize);
    float4 *s_part = smem + a.lutsize + (a.nbricks + 1) / 2;
    int i, j;
    if (!block_pixel(a, blockIdx.x, i, j)) return; /* uniform per workgroup */
    stage(a, s_lut, s_boff);
    const bool inside = (i < a.H) && (j < a.W);
    Acc acc = {0.0f, 0.0f, 0.0f, 0.0f};
    unsigned cnt = 0, fetched = 0;
    if (inside) {
        if (MODE == IVR_MODE_CONFIG) {
            float d[3];
            cfg_dir(a, i, j, d);
            if (SL == 1) march_sort_last<VT, SKIP>(a, s_lut, s_boff, s_part, a.o, d, acc, cnt, fetched);
            else if (SL == 2) march_sort_last_rev<VT, SKIP>(a, s_lut, s_boff, a.o, d, acc, cnt, fetched);
            else march_config<VT, SKIP>(a, s_lut, s_boff, a.o, d, acc, cnt, fetched);
        } else if (i >= a.rect[0] && i < a.rect[1] && j >= a.rect[2] && j < a.rect[3]) {
            /* rows by multiply, columns by repeated addition (main.c:987-999) */
            const float fi = (float)(i - a.rect[0]);
            float d[3] = {a.d0[0] + a.di[0] * fi, a.d0[1] + a.di[1] * fi, a.d0[2] + a.di[2] * fi};
            for (int c = a.rect[2]; c <= j; c++) {
                d[0] += a.dj[0];
                d[1] += a.dj[1];
                d[2] += a.dj[2];
            }
            march_compat<VT>(a, s_lut, s_boff, a.o, d, acc, cnt);
            fetched = cnt;
        }
        const size_t px = (size_t)i * a.W + j;
        a.out8[px] = make_uchar4(quant8(acc.r), quant8(acc.g), quant8(acc.b), quant8(acc.a));
        if (a.out32) a.out32[px] = make_float4(acc.r, acc.g, acc.b, acc.a);
    }
    count_add(a.counters, 0, cnt);
    count_add(a.counters, 1, fetched);
}

/* Sort-last with every run of every ray as its own work: workgroup g
 * renders run g / nblk_pad (counted front to back along each ray) of the
 * rays of pixel block g % nblk_pad, composited from zero with early
 * termination inside the run, into plane `run` of a.tile (W x H float4 per
 * plane).  With COUNTS, a ray's run count goes to a.runcount from the
 * workgroup of its last run (from run 0's for a ray without samples) and
 * planes past it are neither written nor read; without, rays write zero
 * partials for the runs they do not have.  Splitting the rays at the brick
 * faces shortens the longest chains of dependent samples, which otherwise
 * leave most of the GPU idle in the second half of a frame. */
template <typename VT, bool SKIP, bool COUNTS>
__global__ __launch_bounds__(IVR_FRAME_BLOCK) void k_runs(const RenderArgs a)
{
    extern __shared__ float4 smem[];
    float4 *s_lut = smem;
    unsigned long long *s_boff = (unsigned long long *)(smem + a.lutsize);
    const int run = blockIdx.x / a.nblk_pad;
    /* with run counts, a run no ray of the block has is skipped whole (its rays write nothing) */
    if (COUNTS && run > 0 && run >= a.blockmax[blockIdx.x - run * a.nblk_pad]) return;
    int i, j;
    if (!block_pixel(a, blockIdx.x - run * a.nblk_pad, i, j)) return; /* uniform per workgroup */
    stage(a, s_lut, s_boff);
    unsigned cnt = 0, fetched = 0, wrote = 0;
    if (i < a.H && j < a.W) {
        float d[3];
        cfg_dir(a, i, j, d);
        Acc acc = {0.0f, 0.0f, 0.0f, 0.0f};
        const int st = march_run<VT, SKIP>(a, s_lut, s_boff, a.o, d, run, acc, cnt, fetched);
        wrote = (st || !COUNTS) ? 1u : 0u;
        const size_t px = (size_t)i * a.W + j;
        if (st || !COUNTS) a.tile[(size_t)run * a.W * a.H + px] = make_float4(acc.r, acc.g, acc.b, acc.a);
        if (COUNTS && (st == 2 || (st == 0 && run == 0))) a.runcount[px] = (unsigned)(st ? run + 1 : 0);
    }
    count_add(a.counters, 0, cnt);
    count_add(a.counters, 1, fetched);
    count_add(a.counters, 2, wrote);
}

/* an upper bound of the most runs of a ray of each pixel block: each axis
 * brick index is monotone along a ray, so the bricks its samples visit
 * number at most 1 + the sum over axes of |last brick - first brick| (fewer
 * where it crosses an edge or corner); k_runs skips only runs past it */
__global__ __launch_bounds__(IVR_FRAME_BLOCK) void k_blockruns(const RenderArgs a)
{
    int i, j;
    if (!block_pixel(a, blockIdx.x, i, j)) { /* padding: uniform per workgroup */
        if (threadIdx.x == 0) a.blockmax[blockIdx.x] = 0;
        return;
    }
    int cnt = 0;
    if (i < a.H && j < a.W) {
        float d[3];
        cfg_dir(a, i, j, d);
        Ray r;
        if (ray_setup(a, a.o, d, r) && r.n > 0) {
            const Pos p0 = cfg_pos(a, r.go, r.gd, r.tin, r.dt, 0);
            const Pos p1 = cfg_pos(a, r.go, r.gd, r.tin, r.dt, r.n - 1);
            cnt = 1 + abs((p1.j >> a.bshift) - (p0.j >> a.bshift)) + abs((p1.i >> a.bshift) - (p0.i >> a.bshift)) +
                  abs((p1.k >> a.bshift) - (p0.k >> a.bshift));
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cnt = max(cnt, __shfl_xor(cnt, off, 64));
    __shared__ int s_m[IVR_FRAME_BLOCK / 64];
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = 0;
        for (int w = 0; w < IVR_FRAME_BLOCK / 64; w++) m = max(m, s_m[w]);
        a.blockmax[blockIdx.x] = m;
    }
}

/* the planes of k_runs combined far to near per pixel, then quantised;
 * with run counts only a ray's own planes are read (its missing runs would
 * be zero partials, which `over` leaves without effect) */
__global__ __launch_bounds__(256) void k_combine_runs(const float4 *part, const unsigned *runcount, int runs, int W,
                                                      int H, uchar4 *out8, float4 *out32)
{
    const size_t np = (size_t)W * H, px = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (px >= np) return;
    Acc d = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int u = (runcount ? (int)runcount[px] : runs) - 1; u >= 0; u--) {
        const float4 t = part[u * np + px];
        const Acc s = {t.x, t.y, t.z, t.w};
        over(d, s);
    }
    out8[px] = make_uchar4(quant8(d.r), quant8(d.g), quant8(d.b), quant8(d.a));
    if (out32) out32[px] = make_float4(d.r, d.g, d.b, d.a);
}

/* Sort-last pass for brick a.tb: the brick's owned samples of each ray of
 * its pixel rectangle, composited front to back from zero with ERT inside
 * the brick, into a premultiplied RGBA32F tile. */
template <typename VT>
__global__ __launch_bounds__(IVR_FRAME_BLOCK) void k_tile(const RenderArgs a)
{
    extern __shared__ float4 smem[];
    float4 *s_lut = smem;
    unsigned long long *s_boff = (unsigned long long *)(smem + a.lutsize);
    int ti, tj;
    if (!block_pixel(a, blockIdx.x, ti, tj)) return; /* uniform per workgroup */
    stage(a, s_lut, s_boff);
    const bool inside = (ti < a.trect[3]) && (tj < a.trect[2]);
    Acc acc = {0.0f, 0.0f, 0.0f, 0.0f};
    unsigned cnt = 0, fetched = 0;
    if (inside) {
        const int i = a.trect[1] + ti, j = a.trect[0] + tj;
        float d[3];
        cfg_dir(a, i, j, d);
        float tin, tout;
        if (slab(a.dconst, a.o, d, tin, tout)) {
            tin = fmaxf(tin, 0.0f);
            const float dlen = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            const float dt = (a.vox * a.step_vox) / dlen;
            const float nf = 1.0f + (tout - tin) / dt + 0.5f;
            const int n = nf < (float)IVR_MAX_STEPS ? (int)nf : IVR_MAX_STEPS;
            float gor[3], gdr[3];
            ray_voxel(a, a.o, d, gor, gdr);
            /* k range of the brick's ownership box (open towards the volume
             * faces, where clamped samples belong to the boundary brick);
             * grazing rays scan the whole grid */
            int k0 = 0, k1 = n;
            float lo = tin, hi = tin + (float)n * dt;
            bool graze = false;
            for (int r = 0; r < 3; r++) {
                const float gd = gdr[r], go = gor[r];
                if (fabsf(gd) * 20.0f < dlen * a.s[r]) { graze = true; continue; }
                const float blo = a.tb[r] == 0 ? -FLT_MAX : (float)(a.tb[r] << a.bshift);
                const float bhi = a.tb[r] == a.nb[r] - 1 ? FLT_MAX : (float)((a.tb[r] + 1) << a.bshift);
                float t0 = (blo - go) / gd, t1 = (bhi - go) / gd;
                if (t0 > t1) { const float tt = t0; t0 = t1; t1 = tt; }
                lo = fmaxf(lo, t0);
                hi = fminf(hi, t1);
            }
            if (!graze) {
                if (lo > hi) { k0 = 0; k1 = 0; }
                else {
                    k0 = max(0, (int)floorf((lo - tin) / dt) - 2);
                    k1 = min(n, (int)ceilf((hi - tin) / dt) + 3);
                }
            }
            for (int k = k0; k < k1; k++) {
                const Pos p = cfg_pos(a, gor, gdr, tin, dt, k);
                if ((p.j >> a.bshift) != a.tb[0] || (p.i >> a.bshift) != a.tb[1] ||
                    (p.k >> a.bshift) != a.tb[2])
                    continue;
                float v[8];
                fetch8<VT>(a, s_boff, p.j, p.i, p.k, v);
                const float s = trilinear(v, p.fx, p.fy, p.fz);
                const float4 e = s_lut[lut_index(a, s)];
                const float al = e.w * a.alpha;
                const float w = (1.0f - acc.a) * al;
                acc.r = fmaf(w, e.x, acc.r);
                acc.g = fmaf(w, e.y, acc.g);
                acc.b = fmaf(w, e.z, acc.b);
                acc.a = acc.a + w;
                cnt++;
                fetched++;
                if (acc.a >= a.ert) break;
            }
        }
        a.tile[(size_t)ti * a.trect[2] + tj] = make_float4(acc.r, acc.g, acc.b, acc.a);
    }
    count_add(a.counters, 0, cnt);
    count_add(a.counters, 1, fetched);
}

/* compositetiles (main.c:1203-1218): tiles sorted near -> far, composited
 * far -> near with d = s + (1 - s.a) d, then quantised (main.c:1404) */
__global__ __launch_bounds__(256) void k_composite(const float4 *tiles, const TileInfo *ti, int nt,
                                                   uchar4 *out8, float4 *out32, int W, int H)
{
    const int j = blockIdx.x * 16 + (threadIdx.x & 15);
    const int i = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= H || j >= W) return;
    float4 d = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int k = nt - 1; k >= 0; k--) {
        const TileInfo t = ti[k];
        const int x = j - t.x0, y = i - t.y0;
        if (x < 0 || y < 0 || x >= t.w || y >= t.h) continue;
        const float4 s = tiles[t.offset + (long long)y * t.w + x];
        d.z = s.z + (1.0f - s.w) * d.z;
        d.y = s.y + (1.0f - s.w) * d.y;
        d.x = s.x + (1.0f - s.w) * d.x;
        d.w = s.w + (1.0f - s.w) * d.w;
    }
    const size_t px = (size_t)i * W + j;
    out8[px] = make_uchar4(quant8(d.x), quant8(d.y), quant8(d.z), quant8(d.w));
    if (out32) out32[px] = d;
}

/* Macrocell transparency for this frame's LUT: a cell is transparent when
 * every LUT entry its samples can index has zero opacity.  A sample is a
 * chain of three lerps of the cell's voxels, so it lies in [min, max] up to
 * a few ulps; the range is widened by 2^-20 relative (8 ulps) plus 2^-20
 * absolute before the monotone index map (int)(s * lutscale) - lutoffset.
 *
 * One workgroup per node of 8^3 cells: the cells' flags go to LDS and are
 * reduced to the nodes of 2^3, 4^3 and 8^3 cells containing them (cells
 * past the volume's extent hold no samples and count as transparent).  The
 * code written per cell is 0 when it may hold a visible sample, else 1 +
 * the largest level whose node around it is transparent. */
__device__ __forceinline__ bool cell_transparent(float2 r, const float4 *lut, int lutsize, float lutscale,
                                                 int lutoffset, float alpha)
{
    const float tol = 9.5367431640625e-07f; /* 2^-20 */
    float lo_s = r.x * lutscale, hi_s = r.y * lutscale;
    lo_s = lo_s - fabsf(lo_s) * tol - tol;
    hi_s = hi_s + fabsf(hi_s) * tol + tol;
    lo_s = fminf(fmaxf(lo_s, -1.0e9f), 1.0e9f);
    hi_s = fminf(fmaxf(hi_s, -1.0e9f), 1.0e9f);
    int lo = (int)lo_s - lutoffset, hi = (int)hi_s - lutoffset;
    lo = min(max(lo, 0), lutsize - 1);
    hi = min(max(hi, 0), lutsize - 1);
    for (int x = lo; x <= hi; x++)
        if (lut[x].w * alpha != 0.0f) return false;
    return true;
}

__device__ __forceinline__ float2 range_of(float2 r) { return r; }
__device__ __forceinline__ float2 range_of(__half2 r) { return __half22float2(r); }

/* RT: float2 ranges of a volume's cells, or __half2 ranges (rounded outward
 * when stored) of a surrogate's cells */
template <typename RT>
__global__ __launch_bounds__(512) void k_cells(const RT *range, unsigned char *code, int ncx, int ncy, int ncz,
                                               const float4 *lut, int lutsize, float lutscale, int lutoffset,
                                               float alpha)
{
    __shared__ unsigned char s0[512], s1[64], s2[8], s3;
    const int t = threadIdx.x;
    const int tx = t & 7, ty = (t >> 3) & 7, tz = t >> 6;
    const int x = blockIdx.x * 8 + tx, y = blockIdx.y * 8 + ty, z = blockIdx.z * 8 + tz;
    const bool in = x < ncx && y < ncy && z < ncz;
    const size_t c = (size_t)x + (size_t)ncx * ((size_t)y + (size_t)ncy * z);
    const bool e = !in || cell_transparent(range_of(range[in ? c : 0]), lut, lutsize, lutscale, lutoffset, alpha);
    s0[t] = e;
    __syncthreads();
    if (t < 64) {
        const int ux = t & 3, uy = (t >> 2) & 3, uz = t >> 4;
        unsigned char m = 1;
        for (int q = 0; q < 8; q++)
            m &= s0[(2 * ux + (q & 1)) + 8 * ((2 * uy + ((q >> 1) & 1)) + 8 * (2 * uz + (q >> 2)))];
        s1[t] = m;
    }
    __syncthreads();
    if (t < 8) {
        const int ux = t & 1, uy = (t >> 1) & 1, uz = t >> 2;
        unsigned char m = 1;
        for (int q = 0; q < 8; q++)
            m &= s1[(2 * ux + (q & 1)) + 4 * ((2 * uy + ((q >> 1) & 1)) + 4 * (2 * uz + (q >> 2)))];
        s2[t] = m;
    }
    __syncthreads();
    if (t == 0) {
        unsigned char m = 1;
        for (int q = 0; q < 8; q++) m &= s2[q];
        s3 = m;
    }
    __syncthreads();
    if (!in) return;
    unsigned char v = 0;
    if (e) {
        v = 1;
        if (s1[(tx >> 1) + 4 * ((ty >> 1) + 4 * (tz >> 1))]) {
            v = 2;
            if (s2[(tx >> 2) + 2 * ((ty >> 2) + 2 * (tz >> 2))]) v = s3 ? 4 : 3;
        }
    }
    code[c] = v;
}

__global__ __launch_bounds__(256) void k_zero_u64(unsigned long long *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0ull;
}

size_t smem_bytes(const RenderArgs &a, int runs = 0)
{
    return (size_t)a.lutsize * sizeof(float4) + (size_t)((a.nbricks + 1) / 2) * sizeof(float4) +
           (size_t)runs * IVR_FRAME_BLOCK * sizeof(float4);
}

template <int MODE, typename VT, bool SKIP, int SL = 0>
void launch_one(const RenderArgs &a, hipStream_t s, int runs = 0)
{
    hipLaunchKernelGGL((k_render<MODE, VT, SKIP, SL>), dim3(a.nblk_pad), dim3(IVR_FRAME_BLOCK), smem_bytes(a, runs),
                       s, a);
}

} // namespace

void ivr_set_grid(RenderArgs &a, const ivr_ctx *ctx, int w, int h, int kind)
{
    /* fp16 volumes: 64 x 4 blocks of 32 x 2 wavefronts were measured faster on C3 (sort-last
     * 1.19-1.20 vs 1.24-1.26 ms with skipping, 2.00 vs 2.30 ms without; fused 1.56 vs 1.60 ms),
     * fp32 volumes keep 16 x 16 blocks of 16 x 4 (C2 1.05 vs 1.20-1.25 ms); the surrogate's
     * hash-grid gathers share more lines between the 8 x 8 pixels of a wavefront (C4 49.5 vs
     * 52.2 ms with 16 x 4) */
    const bool sg = kind == IVR_GRID_SURROGATE, wide = !sg && a.dtype_f16;
    const int bw = ctx->block_w ? ctx->block_w : (wide ? 64 : 16), bh = IVR_FRAME_BLOCK / bw;
    const int ww = ctx->wave_w ? ctx->wave_w : (wide ? 32 : sg ? 8 : 16);
    a.blk_shift = bw == 64 ? 6 : bw == 32 ? 5 : 4;
    a.nblk_x = (w + bw - 1) / bw;
    a.nblk = a.nblk_x * ((h + bh - 1) / bh);
    int c = ctx->xcd_chunk;
    if (c == IVR_CHUNK_ROW) c = a.nblk_x;
    else if (c == IVR_CHUNK_BAND) c = (a.nblk + 7) / 8;
    a.xcd_chunk = c > 0 ? c : 0;
    const int pad = 8 * (c > 0 ? c : 1);
    a.nblk_pad = (a.nblk + pad - 1) / pad * pad;
    a.wave_shift = ww == 4 ? 2 : ww == 16 ? 4 : ww == 32 ? 5 : ww == 64 ? 6 : 3;
    if (a.wave_shift > a.blk_shift) a.wave_shift = a.blk_shift;
}

int ivr_launch_render(ivr_ctx *ctx, const ivr_volume *vol, RenderArgs &a, int skip, int sort_last)
{
    ivr_set_grid(a, ctx, a.W, a.H);
    const bool f16 = vol->dtype == IVR_F16;
    hipStream_t s = ctx->stream;
    if (a.mode == IVR_MODE_COMPAT) {
        if (f16) launch_one<IVR_MODE_COMPAT, h16, false>(a, s);
        else launch_one<IVR_MODE_COMPAT, float, false>(a, s);
    } else if (sort_last) {
        /* partials in LDS while they fit in IVR_SL_LDS bytes per workgroup */
        const int runs = a.nb[0] + a.nb[1] + a.nb[2] - 2;
        if ((size_t)runs * IVR_FRAME_BLOCK * sizeof(float4) <= IVR_SL_LDS && smem_bytes(a, runs) <= 65536) {
            if (f16) {
                if (skip) launch_one<IVR_MODE_CONFIG, h16, true, 1>(a, s, runs);
                else launch_one<IVR_MODE_CONFIG, h16, false, 1>(a, s, runs);
            } else {
                if (skip) launch_one<IVR_MODE_CONFIG, float, true, 1>(a, s, runs);
                else launch_one<IVR_MODE_CONFIG, float, false, 1>(a, s, runs);
            }
        } else if (f16) {
            if (skip) launch_one<IVR_MODE_CONFIG, h16, true, 2>(a, s);
            else launch_one<IVR_MODE_CONFIG, h16, false, 2>(a, s);
        } else {
            if (skip) launch_one<IVR_MODE_CONFIG, float, true, 2>(a, s);
            else launch_one<IVR_MODE_CONFIG, float, false, 2>(a, s);
        }
    } else if (skip) {
        if (f16) launch_one<IVR_MODE_CONFIG, h16, true>(a, s);
        else launch_one<IVR_MODE_CONFIG, float, true>(a, s);
    } else {
        if (f16) launch_one<IVR_MODE_CONFIG, h16, false>(a, s);
        else launch_one<IVR_MODE_CONFIG, float, false>(a, s);
    }
    IVR_HIP(hipGetLastError());
    return IVR_OK;
}

int ivr_launch_runs(ivr_ctx *ctx, const ivr_volume *vol, RenderArgs &a, int skip)
{
    ivr_set_grid(a, ctx, a.W, a.H);
    const int runs = a.nb[0] + a.nb[1] + a.nb[2] - 2;
    const dim3 g((unsigned)a.nblk_pad * (unsigned)runs);
    const size_t sm = smem_bytes(a);
    hipStream_t s = ctx->stream;
    const dim3 b(IVR_FRAME_BLOCK);
    if (a.runcount) {
        if (vol->dtype == IVR_F16) {
            if (skip) hipLaunchKernelGGL((k_runs<h16, true, true>), g, b, sm, s, a);
            else hipLaunchKernelGGL((k_runs<h16, false, true>), g, b, sm, s, a);
        } else {
            if (skip) hipLaunchKernelGGL((k_runs<float, true, true>), g, b, sm, s, a);
            else hipLaunchKernelGGL((k_runs<float, false, true>), g, b, sm, s, a);
        }
    } else if (vol->dtype == IVR_F16) {
        if (skip) hipLaunchKernelGGL((k_runs<h16, true, false>), g, b, sm, s, a);
        else hipLaunchKernelGGL((k_runs<h16, false, false>), g, b, sm, s, a);
    } else {
        if (skip) hipLaunchKernelGGL((k_runs<float, true, false>), g, b, sm, s, a);
        else hipLaunchKernelGGL((k_runs<float, false, false>), g, b, sm, s, a);
    }
    IVR_HIP(hipGetLastError());
    return IVR_OK;
}

int ivr_launch_blockruns(ivr_ctx *ctx, RenderArgs &a)
{
    ivr_set_grid(a, ctx, a.W, a.H);
    hipLaunchKernelGGL(k_blockruns, dim3((unsigned)a.nblk_pad), dim3(IVR_FRAME_BLOCK), 0, ctx->stream, a);
    IVR_HIP(hipGetLastError());
    return IVR_OK;
}

int ivr_launch_combine_runs(ivr_ctx *ctx, const RenderArgs &a, int runs)
{
    const size_t np = (size_t)a.W * a.H;
    hipLaunchKernelGGL(k_combine_runs, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ctx->stream, a.tile,
                       a.runcount, runs, a.W, a.H, a.out8, a.out32);
    IVR_HIP(hipGetLastError());
    return IVR_OK;
}

int ivr_launch_tile(ivr_ctx *ctx, const ivr_volume *vol, RenderArgs &a)
{
    ivr_set_grid(a, ctx, a.trect[2], a.trect[3]);
    if (vol->dtype == IVR_F16)
        hipLaunchKernelGGL(k_tile<h16>, dim3(a.nblk_pad), dim3(IVR_FRAME_BLOCK), smem_bytes(a), ctx->stream, a);
    else
        hipLaunchKernelGGL(k_tile<float>, dim3(a.nblk_pad), dim3(IVR_FRAME_BLOCK), smem_bytes(a), ctx->stream, a);
    IVR_HIP(hipGetLastError());
    return IVR_OK;
}

int ivr_launch_cells(ivr_ctx *ctx, const ivr_volume *vol, const RenderArgs &a)
{
    return ivr_launch_cells_on(ctx->stream, vol, a, vol->d_cellempty);
}

namespace {

template <typename RT>
int launch_cells(hipStream_t stream, const RT *range, const int nc[3], const RenderArgs &a, unsigned char *out)
{
    static_assert(IVR_MIP_LEVELS == 3, "k_cells reduces nodes of 8^3 cells");
    const dim3 g((nc[0] + 7) / 8, (nc[1] + 7) / 8, (nc[2] + 7) / 8);
    hipLaunchKernelGGL(k_cells<RT>, g, dim3(512), 0, stream, range, out, nc[0], nc[1], nc[2], a.lut, a.lutsize,
                       a.lutscale, a.lutoffset, a.alpha);
    IVR_HIP(hipGetLastError());
    return IVR_OK;
}

} // namespace

int ivr_launch_cells_on(hipStream_t stream, const ivr_volume *vol, const RenderArgs &a, unsigned char *out)
{
    return launch_cells(stream, vol->d_cellrange, vol->nc, a, out);
}

int ivr_launch_cells_half(hipStream_t stream, const __half2 *range, const int nc[3], const RenderArgs &a,
                          unsigned char *out)
{
    return launch_cells(stream, range, nc, a, out);
}

int ivr_launch_composite(ivr_ctx *ctx, ivr_frame *fr, int ntiles)
{
    dim3 g((fr->W + 15) / 16, (fr->H + 15) / 16);
    hipLaunchKernelGGL(k_composite, g, dim3(256), 0, ctx->stream, fr->d_tiles,
                       (const TileInfo *)fr->d_tilemeta, ntiles, fr->d_rgba8, fr->d_rgba32, fr->W,
                       fr->H);
    IVR_HIP(hipGetLastError());
    return IVR_OK;
}

int ivr_launch_zero_u64(hipStream_t stream, unsigned long long *p, size_t n)
{
    if (!n) return IVR_OK;
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_zero_u64, dim3(g), dim3(256), 0, stream, p, n);
    IVR_HIP(hipGetLastError());
    return IVR_OK;
}
