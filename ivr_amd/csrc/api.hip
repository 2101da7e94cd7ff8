/*
 * api.hip -- C-ABI entry points of libivr.so: contexts, transfer functions,
 * frames and the per-frame render orchestration (fused or sort-last).
 *
 * ivr_render replaces the reference frame step of the render server
 * (main.c:226-244): the STATE-derived camera set-up runs on the host
 * (ivr_host.c), the ray casting of all bricks runs as one device launch,
 * and the sort-last tile exchange of compositelocalimage (main.c:708-934),
 * whose MPI send/receive becomes per-brick tiles resident in HBM, is a
 * second launch that composites them in visibility order (compositetiles,
 * main.c:1203-1218) fused with the RGBA8 quantiser (main.c:1404).
 */
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "ivr_internal.h"

/* Scratch of ivr_render_noalloc: the frame counters, the sort-last block
 * maxima, run counts and partial planes of a frame rendered into caller
 * buffers.  One copy per GPU, held while the library is loaded; host code
 * uses only the device address hipGetSymbolAddress gives for it (the name
 * itself is a host-side stand-in).  Calls on one GPU share it. */
#define IVR_SCRATCH_BYTES (72ull << 20)
static __device__ __align__(256) unsigned char scratch[IVR_SCRATCH_BYTES];

namespace {
thread_local char g_err[512] = "";

std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[64] = {}; /* per device */

/* the current device's address of `scratch` (device made current by the caller) */
int scratch_for(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= 64) return ivr_fail(IVR_ERR_ARG, "device %d", device);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return ivr_fail(IVR_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[device] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[device];
    return IVR_OK;
}
}

int ivr_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *ivr_last_error(void) { return g_err; }

int ivr_dalloc(ivr_ctx *ctx, void **p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 1;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return ivr_fail(IVR_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    ctx->bytes += bytes;
    return IVR_OK;
}

void ivr_dfree(ivr_ctx *ctx, void *p, size_t bytes)
{
    if (!p) return;
    if (bytes == 0) bytes = 1;
    (void)hipFree(p);
    ctx->bytes -= std::min<unsigned long long>(ctx->bytes, bytes);
}

/* ---------------- context ---------------- */

int ivr_ctx_create(int device, ivr_ctx **out)
{
    if (!out) return ivr_fail(IVR_ERR_ARG, "null out");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0)
        return ivr_fail(IVR_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= ndev) return ivr_fail(IVR_ERR_ARG, "device %d of %d", device, ndev);
    IVR_HIP(hipSetDevice(device));
    ivr_ctx *c = new ivr_ctx();
    c->device = device;
    if (const char *e = getenv("IVR_XCD_CHUNK")) {
        if (!strcmp(e, "row")) c->xcd_chunk = IVR_CHUNK_ROW;
        else if (!strcmp(e, "band")) c->xcd_chunk = IVR_CHUNK_BAND;
        else c->xcd_chunk = atoi(e) > 0 ? atoi(e) : 0;
    }
    if (const char *e = getenv("IVR_WAVE_W")) {
        const int w = atoi(e);
        c->wave_w = (w == 4 || w == 8 || w == 16 || w == 32 || w == 64) ? w : 0;
    }
    if (const char *e = getenv("IVR_BLOCK_W")) {
        const int w = atoi(e);
        c->block_w = (w == 16 || w == 32 || w == 64) ? w : 0;
    }
    if (const char *e = getenv("IVR_SL_SPLIT")) c->sl_split = atoi(e) != 0;
    if (const char *e = getenv("IVR_SG_HELPERS")) c->sg_helpers = atoi(e) != 0;
    int rc = scratch_for(device, &c->scratch);
    if (rc) {
        delete c;
        return rc;
    }
    IVR_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &ev : c->ev) IVR_HIP(hipEventCreate(&ev));
    *out = c;
    return IVR_OK;
}

int ivr_ctx_destroy(ivr_ctx *c)
{
    if (!c) return IVR_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &ev : c->ev) (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return IVR_OK;
}

int ivr_ctx_sync(ivr_ctx *c)
{
    if (!c) return ivr_fail(IVR_ERR_ARG, "null ctx");
    IVR_HIP(hipStreamSynchronize(c->stream));
    return IVR_OK;
}

int ivr_ctx_record(ivr_ctx *c, int slot)
{
    if (!c || slot < 0 || slot >= 16) return ivr_fail(IVR_ERR_ARG, "bad event slot %d", slot);
    IVR_HIP(hipEventRecord(c->ev[slot], c->stream));
    return IVR_OK;
}

int ivr_ctx_elapsed(ivr_ctx *c, int s0, int s1, float *ms)
{
    if (!c || !ms || s0 < 0 || s0 >= 16 || s1 < 0 || s1 >= 16) return ivr_fail(IVR_ERR_ARG, "bad slots");
    IVR_HIP(hipEventSynchronize(c->ev[s1]));
    IVR_HIP(hipEventElapsedTime(ms, c->ev[s0], c->ev[s1]));
    return IVR_OK;
}

unsigned long long ivr_ctx_device_bytes(const ivr_ctx *c) { return c ? c->bytes : 0; }

/* ---------------- transfer functions ---------------- */

int ivr_tf_create(ivr_ctx *ctx, const float *rgba, int n, ivr_tf **out)
{
    if (!ctx || !rgba || !out) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (n < 1 || n > IVR_MAX_LUT) return ivr_fail(IVR_ERR_ARG, "LUT size %d outside [1,%d]", n, IVR_MAX_LUT);
    ivr_tf *t = new ivr_tf();
    int rc = ivr_dalloc(ctx, (void **)&t->d_lut, sizeof(float4) * IVR_MAX_LUT);
    if (rc) {
        delete t;
        return rc;
    }
    *out = t;
    return ivr_tf_update(ctx, t, rgba, n);
}

int ivr_tf_update(ivr_ctx *ctx, ivr_tf *t, const float *rgba, int n)
{
    if (!ctx || !t || !rgba) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (n < 1 || n > IVR_MAX_LUT) return ivr_fail(IVR_ERR_ARG, "LUT size %d outside [1,%d]", n, IVR_MAX_LUT);
    t->n = n;
    IVR_HIP(hipMemcpyAsync(t->d_lut, rgba, sizeof(float4) * n, hipMemcpyHostToDevice, ctx->stream));
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    return IVR_OK;
}

int ivr_tf_destroy(ivr_ctx *ctx, ivr_tf *t)
{
    if (!t) return IVR_OK;
    if (ctx) {
        (void)hipStreamSynchronize(ctx->stream);
        ivr_dfree(ctx, t->d_lut, sizeof(float4) * IVR_MAX_LUT);
    }
    delete t;
    return IVR_OK;
}

/* ---------------- frames ---------------- */

int ivr_frame_create(ivr_ctx *ctx, int W, int H, int keep32, ivr_frame **out)
{
    if (!ctx || !out) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (W < 1 || H < 1 || W > 16384 || H > 16384) return ivr_fail(IVR_ERR_ARG, "frame %dx%d", W, H);
    ivr_frame *f = new ivr_frame();
    f->W = W;
    f->H = H;
    const size_t np = (size_t)W * H;
    int rc;
    if ((rc = ivr_dalloc(ctx, (void **)&f->d_rgba8, np * sizeof(uchar4))) ||
        (keep32 && (rc = ivr_dalloc(ctx, (void **)&f->d_rgba32, np * sizeof(float4)))) ||
        (rc = ivr_dalloc(ctx, (void **)&f->d_counters, IVR_CTR_BYTES)) ||
        (rc = ivr_dalloc(ctx, (void **)&f->d_tilemeta, sizeof(TileInfo) * IVR_MAX_BRICKS))) {
        ivr_frame_destroy(ctx, f);
        return rc;
    }
    for (auto &e : f->e) IVR_HIP(hipEventCreate(&e));
    IVR_HIP(hipHostMalloc(&f->h_tilemeta, sizeof(TileInfo) * IVR_MAX_BRICKS, hipHostMallocDefault));
    IVR_HIP(hipEventRecord(f->e[2], ctx->stream));
    IVR_HIP(hipMemsetAsync(f->d_counters, 0, IVR_CTR_BYTES, ctx->stream));
    *out = f;
    return IVR_OK;
}

int ivr_frame_destroy(ivr_ctx *ctx, ivr_frame *f)
{
    if (!f) return IVR_OK;
    if (ctx) {
        (void)hipStreamSynchronize(ctx->stream);
        const size_t np = (size_t)f->W * f->H;
        ivr_dfree(ctx, f->d_rgba8, np * sizeof(uchar4));
        ivr_dfree(ctx, f->d_rgba32, np * sizeof(float4));
        ivr_dfree(ctx, f->d_counters, IVR_CTR_BYTES);
        ivr_dfree(ctx, f->d_tilemeta, sizeof(TileInfo) * IVR_MAX_BRICKS);
        ivr_dfree(ctx, f->d_tiles, f->tiles_cap * sizeof(float4));
        ivr_dfree(ctx, f->d_runcount, np * 4);
        ivr_dfree(ctx, f->d_blockmax, f->blockmax_cap * sizeof(int));
    }
    for (auto &e : f->e)
        if (e) (void)hipEventDestroy(e);
    if (f->h_tilemeta) (void)hipHostFree(f->h_tilemeta);
    delete f;
    return IVR_OK;
}

int ivr_frame_download(ivr_ctx *ctx, ivr_frame *f, unsigned char *rgba8, float *rgba32)
{
    if (!ctx || !f) return ivr_fail(IVR_ERR_ARG, "null argument");
    const size_t np = (size_t)f->W * f->H;
    if (rgba8)
        IVR_HIP(hipMemcpyAsync(rgba8, f->d_rgba8, np * sizeof(uchar4), hipMemcpyDeviceToHost, ctx->stream));
    if (rgba32) {
        if (!f->d_rgba32) return ivr_fail(IVR_ERR_ARG, "frame was created without RGBA32F");
        IVR_HIP(hipMemcpyAsync(rgba32, f->d_rgba32, np * sizeof(float4), hipMemcpyDeviceToHost, ctx->stream));
    }
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    return IVR_OK;
}

int ivr_frame_stats_get(ivr_ctx *ctx, ivr_frame *f, ivr_frame_stats *st)
{
    if (!ctx || !f || !st) return ivr_fail(IVR_ERR_ARG, "null argument");
    std::vector<unsigned long long> c(IVR_CTR_SLOTS * IVR_CTRS);
    IVR_HIP(hipMemcpyAsync(c.data(), f->d_counters, IVR_CTR_BYTES, hipMemcpyDeviceToHost, ctx->stream));
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    unsigned long long sum[IVR_CTRS] = {};
    for (int q = 0; q < IVR_CTR_SLOTS; q++)
        for (int k = 0; k < IVR_CTRS; k++) sum[k] += c[q * IVR_CTRS + k];
    st->samples = sum[0];
    st->visited = sum[1];
    st->partials = sum[2];
    st->evaluations = sum[3];
    st->render_ms = 0.0f;
    st->composite_ms = 0.0f;
    if (f->timed) {
        IVR_HIP(hipEventElapsedTime(&st->render_ms, f->e[0], f->e[1]));
        if (f->composited) IVR_HIP(hipEventElapsedTime(&st->composite_ms, f->e[1], f->e[2]));
    }
    return IVR_OK;
}

/* ---------------- render ---------------- */

void ivr_fill_args(RenderArgs &a, const ivr_cam &cam, const ivr_volume *v, const ivr_tf *tf,
                   const ivr_frame_desc *d, ivr_frame *f)
{
    memset(&a, 0, sizeof a);
    a.W = f ? f->W : 0;
    a.H = f ? f->H : 0;
    for (int r = 0; r < 4; r++) a.rect[r] = cam.rect[r];
    a.mode = d->mode;
    a.xfer = d->xfer;
    for (int r = 0; r < 3; r++) {
        a.o[r] = cam.o[r];
        a.m5[r] = cam.m5[r];
        a.d0[r] = cam.d0[r];
        a.di[r] = cam.di[r];
        a.dj[r] = cam.dj[r];
        a.bmin[r] = cam.bmin[r];
        a.bmax[r] = cam.bmax[r];
        a.ext[r] = cam.bmax[r] - cam.bmin[r];
        a.s[r] = cam.s[r];
        a.n[r] = v->n[r];
        a.nm1f[r] = (float)(v->n[r] - 1);
        a.nb[r] = v->nb[r];
        a.nc[r] = v->nc[r];
    }
    for (int r = 0; r < 9; r++) a.M[r] = cam.M[r];
    for (int r = 0; r < 6; r++) a.dconst[r] = cam.dconst[r];
    a.l = cam.l;
    a.b = cam.b;
    a.du = cam.du;
    a.dv = cam.dv;
    a.vox = cam.vox;
    a.bshift = v->bshift;
    a.bmask = v->B - 1;
    a.P = v->P;
    a.nbricks = v->nbricks;
    a.vox_data = v->data;
    a.pair_data = v->pairs;
    a.pair_bytes = 2ull * v->bytes < (1ull << 32) - (1ull << 20) ? (unsigned)(2 * v->bytes) : 0u;
    a.dtype_f16 = v->dtype == IVR_F16;
    a.boff = v->d_boff;
    a.cellempty = v->d_cellempty;
    a.lut = tf->d_lut;
    a.lutsize = tf->n;
    a.lutoffset = d->lutoffset;
    a.maxlutindx = d->maxlutindx;
    a.lutscale = d->lutscale;
    a.alpha = d->state.alpha;
    a.step_vox = d->step_vox;
    a.ert = d->ert;
    a.out8 = f ? f->d_rgba8 : nullptr;
    a.out32 = f ? f->d_rgba32 : nullptr;
    a.counters = f ? f->d_counters : nullptr;
}

namespace {

/* pixel rectangle of a brick's ownership box grown by one voxel, projected
 * with pm * mvm like updatevolbbox (mathutil.c:948-1018); whole image when
 * a corner is behind the eye */
void brick_rect(const ivr_state &st, const ivr_cam &cam, const ivr_volume *v, const int b[3], int W,
                int H, int rect[4])
{
    float lo[3], hi[3];
    for (int r = 0; r < 3; r++) {
        const float vlo = b[r] == 0 ? -1.0f : (float)(b[r] * v->B) - 1.0f;
        const float vhi = b[r] == v->nb[r] - 1 ? (float)v->n[r] + 1.0f : (float)((b[r] + 1) * v->B) + 1.0f;
        lo[r] = cam.bmin[r] + vlo / cam.s[r];
        hi[r] = cam.bmin[r] + vhi / cam.s[r];
    }
    double pmv[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double acc = 0.0;
            for (int k = 0; k < 4; k++) acc += (double)st.pm[k * 4 + i] * (double)st.mvm[j * 4 + k];
            pmv[i][j] = acc;
        }
    double x0 = 1e30, y0 = 1e30, x1 = -1e30, y1 = -1e30;
    for (int c = 0; c < 8; c++) {
        const double p[4] = {(c & 1) ? hi[0] : lo[0], (c & 2) ? hi[1] : lo[1], (c & 4) ? hi[2] : lo[2], 1.0};
        double q[4];
        for (int i = 0; i < 4; i++) q[i] = pmv[i][0] * p[0] + pmv[i][1] * p[1] + pmv[i][2] * p[2] + pmv[i][3] * p[3];
        if (q[3] <= 1e-6) {
            rect[0] = 0; rect[1] = 0; rect[2] = W; rect[3] = H;
            return;
        }
        const double wx = (q[0] / q[3] + 1.0) * 0.5 * W, wy = (q[1] / q[3] + 1.0) * 0.5 * H;
        x0 = std::min(x0, wx); x1 = std::max(x1, wx);
        y0 = std::min(y0, wy); y1 = std::max(y1, wy);
    }
    int ix0 = std::max(0, (int)std::floor(x0) - 2), iy0 = std::max(0, (int)std::floor(y0) - 2);
    int ix1 = std::min(W, (int)std::ceil(x1) + 2), iy1 = std::min(H, (int)std::ceil(y1) + 2);
    rect[0] = ix0;
    rect[1] = iy0;
    rect[2] = std::max(0, ix1 - ix0);
    rect[3] = std::max(0, iy1 - iy0);
}

int render_sort_last(ivr_ctx *ctx, const ivr_volume *v, const ivr_frame_desc *d, const ivr_cam &cam,
                     RenderArgs &base, ivr_frame *f)
{
    /* visibility order: the sum over axes of |brick - eye brick| grows
     * strictly along every ray, so ascending sums are front to back */
    int eb[3];
    for (int r = 0; r < 3; r++) {
        const float ev = (cam.o[r] - cam.bmin[r]) * cam.s[r];
        eb[r] = (int)std::floor(ev / (float)v->B);
    }
    std::vector<std::pair<long long, int>> order;
    for (int z = 0; z < v->nb[2]; z++)
        for (int y = 0; y < v->nb[1]; y++)
            for (int x = 0; x < v->nb[0]; x++) {
                const long long key = (long long)std::abs(x - eb[0]) + std::abs(y - eb[1]) + std::abs(z - eb[2]);
                order.push_back({key, x + v->nb[0] * (y + v->nb[1] * z)});
            }
    std::sort(order.begin(), order.end());
    std::vector<TileInfo> tiles;
    long long total = 0;
    for (size_t q = 0; q < order.size(); q++) {
        const int id = order[q].second;
        const int b[3] = {id % v->nb[0], (id / v->nb[0]) % v->nb[1], id / (v->nb[0] * v->nb[1])};
        int rc4[4];
        brick_rect(d->state, cam, v, b, f->W, f->H, rc4);
        if (rc4[2] <= 0 || rc4[3] <= 0) continue;
        TileInfo t;
        t.x0 = rc4[0]; t.y0 = rc4[1]; t.w = rc4[2]; t.h = rc4[3];
        t.bx = b[0]; t.by = b[1]; t.bz = b[2];
        t.order = (int)tiles.size();
        t.offset = total;
        total += (long long)t.w * t.h;
        tiles.push_back(t);
    }
    if ((size_t)total > f->tiles_cap) {
        ivr_dfree(ctx, f->d_tiles, f->tiles_cap * sizeof(float4));
        f->tiles_cap = 0;
        int rc = ivr_dalloc(ctx, (void **)&f->d_tiles, (size_t)total * sizeof(float4));
        if (rc) return rc;
        f->tiles_cap = (size_t)total;
    }
    for (const TileInfo &t : tiles) {
        RenderArgs a = base;
        a.tb[0] = t.bx; a.tb[1] = t.by; a.tb[2] = t.bz;
        a.trect[0] = t.x0; a.trect[1] = t.y0; a.trect[2] = t.w; a.trect[3] = t.h;
        a.tile = f->d_tiles + t.offset;
        int rc = ivr_launch_tile(ctx, v, a);
        if (rc) return rc;
    }
    if (f->timed) IVR_HIP(hipEventRecord(f->e[1], ctx->stream));
    /* stage the table in the frame's pinned buffer once the previous
     * frame's copy out of it has completed */
    IVR_HIP(hipEventSynchronize(f->e[2]));
    memcpy(f->h_tilemeta, tiles.data(), sizeof(TileInfo) * tiles.size());
    IVR_HIP(hipMemcpyAsync(f->d_tilemeta, f->h_tilemeta, sizeof(TileInfo) * tiles.size(),
                           hipMemcpyHostToDevice, ctx->stream));
    f->composited = 1;
    return ivr_launch_composite(ctx, f, (int)tiles.size());
}

} // namespace

namespace {

/* device buffer of `need` elements for a frame: grown (freed and allocated)
 * for a library frame, checked against its capacity for a frame in scratch */
template <typename T>
int frame_buffer(ivr_ctx *ctx, bool noalloc, T **p, size_t *cap, size_t need)
{
    if (need <= *cap) return IVR_OK;
    if (noalloc)
        return ivr_fail(IVR_ERR_ARG, "frame needs %zu bytes of scratch beyond the %llu the library holds", need * sizeof(T),
                        IVR_SCRATCH_BYTES);
    ivr_dfree(ctx, *p, *cap * sizeof(T));
    *cap = 0;
    int rc = ivr_dalloc(ctx, (void **)p, need * sizeof(T));
    if (!rc) *cap = need;
    return rc;
}

int check_desc(const ivr_volume *v, const ivr_tf *tf, const ivr_frame_desc *d, int W, int H)
{
    if (d->state.winwidth != W || d->state.winheight != H)
        return ivr_fail(IVR_ERR_ARG, "state %dx%d does not match frame %dx%d", d->state.winwidth,
                        d->state.winheight, W, H);
    if (d->mode != IVR_MODE_COMPAT && d->mode != IVR_MODE_CONFIG)
        return ivr_fail(IVR_ERR_ARG, "unknown mode %d", d->mode);
    if (d->mode == IVR_MODE_CONFIG && !(d->step_vox > 0.0f))
        return ivr_fail(IVR_ERR_ARG, "step_vox must be > 0");
    if (d->composite != IVR_FUSED && d->composite != IVR_SORT_LAST && d->composite != IVR_SORT_LAST_TILES)
        return ivr_fail(IVR_ERR_ARG, "unknown composite %d", d->composite);
    if (d->mode == IVR_MODE_COMPAT && d->composite != IVR_FUSED)
        return ivr_fail(IVR_ERR_UNSUPPORTED, "sort-last compositing is a config-mode path");
    if ((size_t)tf->n * sizeof(float4) + (size_t)v->nbricks * 8 > 65536)
        return ivr_fail(IVR_ERR_UNSUPPORTED, "LUT of %d entries and %d bricks exceed 64 KiB of LDS", tf->n,
                        v->nbricks);
    return IVR_OK;
}

/* One frame on the context stream into f's buffers.  noalloc: f's buffers
 * are the caller's image and the scratch array (capacities set by the
 * caller); nothing is allocated, counters are cleared by a kernel, no event
 * is recorded. */
int render_into(ivr_ctx *ctx, const ivr_volume *v, const ivr_tf *tf, const ivr_frame_desc *d, ivr_frame *f,
                bool noalloc)
{
    ivr_cam cam;
    ivr_cam_setup(&d->state, v->n[0], v->n[1], v->n[2], &cam);
    RenderArgs a;
    ivr_fill_args(a, cam, v, tf, d, f);
    int rc;
    if (noalloc) {
        if ((rc = ivr_launch_zero_u64(ctx->stream, f->d_counters, IVR_CTR_BYTES / 8))) return rc;
    } else {
        IVR_HIP(hipMemsetAsync(f->d_counters, 0, IVR_CTR_BYTES, ctx->stream));
    }
    f->timed = !noalloc;
    f->composited = 0;
    const int skip = d->mode == IVR_MODE_CONFIG && d->skip;
    if (skip && (rc = ivr_launch_cells(ctx, v, a))) return rc;
    if (f->timed) IVR_HIP(hipEventRecord(f->e[0], ctx->stream)); /* render_ms: the render kernel(s) alone */
    if (d->mode == IVR_MODE_CONFIG && d->composite == IVR_SORT_LAST_TILES) {
        if (noalloc) return ivr_fail(IVR_ERR_UNSUPPORTED, "materialised tiles need a library frame");
        if ((rc = render_sort_last(ctx, v, d, cam, a, f))) return rc;
        IVR_HIP(hipEventRecord(f->e[2], ctx->stream));
        return IVR_OK;
    }
    if (d->mode == IVR_MODE_CONFIG && d->composite == IVR_SORT_LAST && ctx->sl_split) {
        const int runs = v->nb[0] + v->nb[1] + v->nb[2] - 2;
        /* Per-ray run counts spare the zero planes of rays that have few of
         * the frame's run slots (4x4x4 bricks: 10 slots, C3 1.38 -> 1.29 ms);
         * with at most 4 slots (2x2x2 bricks) writing the zero planes was
         * measured faster (C2 1.08 vs 1.12 ms). */
        const bool counts = runs > 4;
        size_t rc_cap = f->d_runcount ? (size_t)f->W * f->H : 0;
        if ((rc = frame_buffer(ctx, noalloc, &f->d_tiles, &f->tiles_cap, (size_t)runs * f->W * f->H)) ||
            (counts && (rc = frame_buffer(ctx, noalloc, &f->d_runcount, &rc_cap, (size_t)f->W * f->H))))
            return rc;
        a.tile = f->d_tiles;
        a.runcount = counts ? f->d_runcount : nullptr;
        if (counts) {
            /* the workgroups of runs that no ray of their pixel block has are skipped
             * (C3 1.22 -> 1.18 ms per frame); the bound pass is not part of render_ms */
            ivr_set_grid(a, ctx, f->W, f->H);
            if ((rc = frame_buffer(ctx, noalloc, &f->d_blockmax, &f->blockmax_cap, (size_t)a.nblk_pad))) return rc;
            a.blockmax = f->d_blockmax;
            if ((rc = ivr_launch_blockruns(ctx, a))) return rc;
            if (f->timed) IVR_HIP(hipEventRecord(f->e[0], ctx->stream));
        }
        if ((rc = ivr_launch_runs(ctx, v, a, skip))) return rc;
        if (f->timed) IVR_HIP(hipEventRecord(f->e[1], ctx->stream));
        f->composited = 1;
        if ((rc = ivr_launch_combine_runs(ctx, a, runs))) return rc;
        if (f->timed) IVR_HIP(hipEventRecord(f->e[2], ctx->stream));
        return IVR_OK;
    }
    if ((rc = ivr_launch_render(ctx, v, a, skip, d->mode == IVR_MODE_CONFIG && d->composite == IVR_SORT_LAST)))
        return rc;
    if (f->timed) IVR_HIP(hipEventRecord(f->e[1], ctx->stream));
    return IVR_OK;
}

} // namespace

int ivr_render(ivr_ctx *ctx, const ivr_volume *v, const ivr_tf *tf, const ivr_frame_desc *d, ivr_frame *f)
{
    if (!ctx || !v || !tf || !d || !f) return ivr_fail(IVR_ERR_ARG, "null argument");
    int rc = check_desc(v, tf, d, f->W, f->H);
    if (rc) return rc;
    return render_into(ctx, v, tf, d, f, false);
}

int ivr_render_noalloc(ivr_ctx *ctx, const ivr_volume *v, const ivr_tf *tf, const ivr_frame_desc *d,
                       unsigned char *d_rgba8, float *d_rgba32)
{
    if (!ctx || !v || !tf || !d || !d_rgba8) return ivr_fail(IVR_ERR_ARG, "null argument");
    const int W = d->state.winwidth, H = d->state.winheight;
    if (W < 1 || H < 1 || W > 16384 || H > 16384) return ivr_fail(IVR_ERR_ARG, "frame %dx%d", W, H);
    int rc = check_desc(v, tf, d, W, H);
    if (rc) return rc;
    if (!ctx->scratch) return ivr_fail(IVR_ERR_HIP, "no scratch address for device %d", ctx->device);
    /* scratch layout: counters | block maxima | run counts | partial planes, 256-byte aligned */
    const size_t np = (size_t)W * H, nblk_max = (np + IVR_FRAME_BLOCK - 1) / IVR_FRAME_BLOCK + 8 * 1024;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_blk = up(IVR_CTR_BYTES), o_cnt = o_blk + up(nblk_max * sizeof(int)), o_part = o_cnt + up(np * 4);
    if (o_part > IVR_SCRATCH_BYTES) return ivr_fail(IVR_ERR_ARG, "frame %dx%d exceeds the scratch array", W, H);
    ivr_frame f; /* a frame over the caller's image and the scratch array */
    f.W = W;
    f.H = H;
    f.d_rgba8 = (uchar4 *)d_rgba8;
    f.d_rgba32 = (float4 *)d_rgba32;
    f.d_counters = (unsigned long long *)ctx->scratch;
    f.d_blockmax = (int *)(ctx->scratch + o_blk);
    f.blockmax_cap = nblk_max;
    f.d_runcount = (unsigned *)(ctx->scratch + o_cnt);
    f.d_tiles = (float4 *)(ctx->scratch + o_part);
    f.tiles_cap = (IVR_SCRATCH_BYTES - o_part) / sizeof(float4);
    return render_into(ctx, v, tf, d, &f, true);
}
