/*
 * compat.hip -- libivr_compat.so: the reference's ray.c / transferfuncs.c
 * entry points under their own names and signatures, run on the GPU.
 *
 * A host built from the reference's main.c links this library in place of
 * ray.o and transferfuncs.o (colormap.o is replaced by compat_cmap.c).  The
 * library defines the globals those files define (dconst, ray.c:38; the
 * function pointers of ray.h:25 and transferfuncs.h:22) and reads the host's
 * globals (main.c:63-105).  Every per-ray or per-sample call is one launch of
 * k_ops on the library's stream, computing exactly what the reference
 * function computes (same float order), raycast calls of concurrent threads
 * combined into one launch; ivr_compat_createlocalimage runs the
 * whole pixel loop of createlocalimage (main.c:983-1009) as one launch of the
 * compat-mode frame kernel; the colormap builders (colormap.c:21-181, exported
 * by compat_cmap.c) run as one launch over the caller's table.
 */
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <vector>

#include "render_dev.h"
#include "../../include/ivr_compat.h"

extern "C" {
/* host globals (main.c:63-105) */
extern ivr_ref_volume volume;
extern ivr_ref_aabb aabb;
extern ivr_color4f *lut;
extern int lutoffset;
extern int maxlutindx;
extern ivr_state mystate;
extern ivr_color4f *colorbuf;
extern ivr_ref_window viewport;
extern ivr_ref_frustum viewvolume;
extern float mvm[4][4];
extern float pm[4][4];

/* globals defined by ray.c:38, ray.h:25 and transferfuncs.h:22 */
float dconst[6];
float (*getsamplefunc)(float, float, float);
ivr_color4f (*transferfunc)(ivr_color4f *, ivr_color4f, float *, int, float);
}

namespace {

using namespace ivr_dev;

enum {
    OP_RAYCAST = 0, OP_SAMPLES = 1, OP_INTERSECT = 2, OP_LINEAR = 3, OP_NEAREST = 4, OP_TF = 5,
    OP_FIRSTVOXEL = 6, OP_VOXELS = 7
};
/* the voxel-walk ops read the host's aabb box and localdims from the first
 * DDA_HEAD floats of their input, the rays after them */
#define DDA_HEAD 12

struct Compat {
    std::mutex mu;
    ivr_ctx *ctx = nullptr;
    ivr_volume *vol = nullptr;
    const float *vol_src = nullptr;
    int dims[3] = {0, 0, 0};
    int hsize = 0;
    ivr_tf *tf = nullptr;
    ivr_frame *fr = nullptr;
    int xfer = 0;
    int sampler = 1;
    float newconst[6] = {1e30f, 1e30f, 1e30f, 1e30f, 1e30f, 1e30f};
    float *d_io = nullptr;
    size_t io_cap = 0;
    /* pinned, device-mapped staging of the per-ray ops: the kernel reads its
     * rays from and writes its results to host memory, so a call is one
     * launch and one synchronisation, with no copies queued */
    float *h_io = nullptr, *h_io_dev = nullptr;
    size_t h_cap = 0;
    std::vector<ivr_color4f> lut_copy; /* the host LUT last uploaded */
    const ivr_color4f *lut_src = nullptr;
    unsigned long long ray_batches = 0, rays_cast = 0; /* combined raycast launches */
};

Compat g;

int ensure_ctx()
{
    if (g.ctx) return IVR_OK;
    return ivr_ctx_create(0, &g.ctx);
}

/* mirror the host volume (volume.data, volume.localdims) on the device and
 * size the LUT as the reference's histogram does, (int)(max - min + 1)
 * (mathutil.c:444-470), from the device min and max of the uploaded copy */
int ensure_volume()
{
    int rc = ensure_ctx();
    if (rc) return rc;
    const int d[3] = {volume.localdims.x, volume.localdims.y, volume.localdims.z};
    if (g.vol && g.vol_src == volume.data && !memcmp(d, g.dims, sizeof d)) return IVR_OK;
    if (!volume.data) return ivr_fail(IVR_ERR_ARG, "host volume.data is NULL");
    if (g.vol) ivr_volume_destroy(g.ctx, g.vol);
    g.vol = nullptr;
    int md = std::max(d[0], std::max(d[1], d[2]));
    int brick = 16;
    while (brick < 256 && brick < md - 1) brick <<= 1;
    if ((rc = ivr_volume_create(g.ctx, d[0], d[1], d[2], IVR_F32, brick, &g.vol))) return rc;
    if ((rc = ivr_volume_upload(g.ctx, g.vol, volume.data))) return rc;
    float lo, hi;
    int hs;
    if ((rc = ivr_volume_histogram(g.ctx, g.vol, &lo, &hi, &hs, nullptr, 0))) return rc;
    g.hsize = std::max(1, std::min(hs, IVR_MAX_LUT));
    g.vol_src = volume.data;
    memcpy(g.dims, d, sizeof d);
    return IVR_OK;
}

/* the host LUT to the device, skipped while its address, size and contents
 * equal those of the last upload (the per-ray calls check it every time) */
int upload_lut(const ivr_color4f *l, int n)
{
    if (!l) return ivr_fail(IVR_ERR_ARG, "LUT pointer is NULL");
    n = std::max(1, std::min(n, IVR_MAX_LUT));
    if (g.tf && g.lut_src == l && (int)g.lut_copy.size() == n && g.tf->n == n &&
        !memcmp(g.lut_copy.data(), l, sizeof(ivr_color4f) * (size_t)n))
        return IVR_OK;
    int rc = g.tf ? ivr_tf_update(g.ctx, g.tf, (const float *)l, n) : ivr_tf_create(g.ctx, (const float *)l, n, &g.tf);
    if (rc) return rc;
    g.lut_src = l;
    g.lut_copy.assign(l, l + n);
    return IVR_OK;
}

/* the pinned mapped staging buffer, at least `floats` long */
int host_io(size_t floats)
{
    if (floats <= g.h_cap) return IVR_OK;
    if (g.h_io) (void)hipHostFree(g.h_io);
    g.h_io = g.h_io_dev = nullptr;
    g.h_cap = 0;
    const size_t cap = std::max<size_t>(floats, 1 << 16);
    IVR_HIP(hipHostMalloc((void **)&g.h_io, cap * sizeof(float), hipHostMallocMapped | hipHostMallocCoherent));
    IVR_HIP(hipHostGetDevicePointer((void **)&g.h_io_dev, g.h_io, 0));
    g.h_cap = cap;
    return IVR_OK;
}

int io_buffer(size_t floats)
{
    if (floats <= g.io_cap) return IVR_OK;
    ivr_dfree(g.ctx, g.d_io, g.io_cap * sizeof(float));
    g.io_cap = 0;
    int rc = ivr_dalloc(g.ctx, (void **)&g.d_io, floats * sizeof(float));
    if (!rc) g.io_cap = floats;
    return rc;
}

/* the STATE of the host's current frame (inverse of setmvm/setpm,
 * mathutil.h:211-233; setfrustum mathutil.h:178-181) */
void host_state(ivr_state *st)
{
    memset(st, 0, sizeof *st);
    st->winwidth = (int)viewport.w;
    st->winheight = (int)viewport.h;
    st->alpha = mystate.alpha;
    memcpy(st->newconst, g.newconst, sizeof st->newconst);
    st->vv[0] = viewvolume.left;
    st->vv[1] = viewvolume.right;
    st->vv[2] = viewvolume.bottom;
    st->vv[3] = viewvolume.top;
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
            st->mvm[c * 4 + r] = mvm[r][c];
            st->pm[c * 4 + r] = pm[r][c];
        }
}

void op_args(RenderArgs &a, int lutsize, float alpha)
{
    ivr_state st;
    host_state(&st);
    st.alpha = alpha;
    ivr_cam cam;
    ivr_cam_setup(&st, g.vol->n[0], g.vol->n[1], g.vol->n[2], &cam);
    ivr_frame_desc d;
    ivr_desc_config(&d);
    d.state = st;
    d.mode = IVR_MODE_COMPAT;
    d.xfer = g.xfer;
    d.lutscale = 1.0f;
    d.lutoffset = lutoffset;
    d.maxlutindx = maxlutindx;
    ivr_fill_args(a, cam, g.vol, g.tf, &d, nullptr);
    for (int i = 0; i < 6; i++) a.dconst[i] = dconst[i];
    a.lutsize = lutsize;
}

/* Entry voxel of getfirstvoxel (ray.c:333-384).  box = aabb min xyz, max
 * xyz.  The candidate parameters are box plane / direction per axis with a
 * non-zero direction (the origin is not subtracted there), visited min x,
 * max x, min y, ...; the first strictly smallest positive one is kept, and
 * the entry point d*t + o is rounded half up per axis.  -1 when no candidate
 * is positive (v untouched). */
__device__ int first_voxel(const float *box, const float *o, const float *d, int v[3])
{
    float t = 0.0f;
    int found = 0;
#pragma unroll
    for (int c = 0; c < 6; c++) {
        const int ax = c >> 1;
        if (d[ax] == 0.0f) continue;
        const float tc = box[(c & 1) * 3 + ax] / d[ax];
        if (tc > 0.0f && (t == 0.0f || tc < t)) {
            t = tc;
            found = 1;
        }
    }
    if (!found) return -1;
#pragma unroll
    for (int ax = 0; ax < 3; ax++) v[ax] = (int)((d[ax] * t + o[ax]) + 0.5f);
    return 0;
}

/* getvoxelsonray (ray.c:387-483): Amanatides-Woo walk of maxs voxels from
 * first_voxel, writing maptoffset3d(x, y, z, w, h, d) = ((z*h + x)*w + y
 * (mathutil.h:139-143; wrapping 32-bit) as floats.  Face parameters
 * (boundary - o)/dir and steps sign/dir turn NaN into +inf (ray.c:436-455);
 * the x face wins only when strictly nearest, then y over z (ray.c:462-481). */
__device__ int voxel_walk(const float *box, const int *dims, const float *o, const float *d, int maxs,
                          float *res)
{
    int v[3];
    if (first_voxel(box, o, d, v) < 0) return -1;
    int st[3];
    float tm[3], dl[3];
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
        st[ax] = d[ax] > 0.0f ? 1 : (d[ax] < 0.0f ? -1 : 0);
        tm[ax] = ((float)(v[ax] + (st[ax] > 0)) - o[ax]) / d[ax];
        dl[ax] = (float)st[ax] / d[ax];
        if (tm[ax] != tm[ax]) tm[ax] = __builtin_huge_valf();
        if (dl[ax] != dl[ax]) dl[ax] = __builtin_huge_valf();
    }
    const unsigned w = (unsigned)dims[0], h = (unsigned)dims[1];
    unsigned x = (unsigned)v[0], y = (unsigned)v[1], z = (unsigned)v[2];
    for (int i = 0; i < maxs; i++) {
        res[i] = (float)(int)((z * h + x) * w + y);
        if (tm[0] < tm[1] && tm[0] < tm[2]) {
            x += (unsigned)st[0];
            tm[0] += dl[0];
        } else if (tm[1] < tm[2]) {
            y += (unsigned)st[1];
            tm[1] += dl[1];
        } else {
            z += (unsigned)st[2];
            tm[2] += dl[2];
        }
    }
    return 0;
}

/* per-ray / per-sample reference functions; in/out layouts per op */
__global__ void k_ops(const RenderArgs a, int op, int n, const float *in, float *out, int maxs)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (op == OP_RAYCAST) { /* raycast (ray.c:59-85) */
        Acc c = {0.0f, 0.0f, 0.0f, 0.0f};
        unsigned cnt = 0;
        march_compat<float>(a, a.lut, a.boff, in + t * 6, in + t * 6 + 3, c, cnt);
        out[t * 4 + 0] = c.r;
        out[t * 4 + 1] = c.g;
        out[t * 4 + 2] = c.b;
        out[t * 4 + 3] = c.a;
    } else if (op == OP_SAMPLES || op == OP_INTERSECT) { /* ray.c:239-318 */
        const float *o = in + t * 6, *d = o + 3;
        float tin, tout;
        const bool hit = slab(a.dconst, o, d, tin, tout);
        if (op == OP_INTERSECT) {
            out[t * 3 + 0] = hit ? 1.0f : 0.0f;
            out[t * 3 + 1] = tin;
            out[t * 3 + 2] = tout;
            return;
        }
        float *res = out + (size_t)t * (maxs + 1);
        if (!hit) {
            res[0] = -1.0f;
            return;
        }
        const float dt = 0.001f;
        int iend = 1.0f + (tout - tin) / dt + 0.5f;
        iend = iend < maxs ? iend : maxs;
        for (int i = 0; i < iend; i++) {
            const float tt = tin + ((float)i * dt);
            res[1 + i] = sample_ref<float>(a, a.boff, d[0] * tt + o[0], d[1] * tt + o[1], d[2] * tt + o[2]);
        }
        res[0] = (float)(iend > 0 ? iend : 0);
    } else if (op == OP_LINEAR) { /* getlinearsample (ray.c:87-181) */
        out[t] = sample_ref<float>(a, a.boff, in[t * 3], in[t * 3 + 1], in[t * 3 + 2]);
    } else if (op == OP_NEAREST) { /* getnearestsample (ray.c:182-238): the voxel's
                                      bits read as an unsigned int, converted */
        const float x = in[t * 3], y = in[t * 3 + 1], z = in[t * 3 + 2];
        int j = (int)((float)a.n[0] * ((x - a.bmin[0]) / a.ext[0]));
        int i = (int)((float)a.n[1] * ((y - a.bmin[1]) / a.ext[1]));
        int k = (int)((float)a.n[2] * ((z - a.bmin[2]) / a.ext[2]));
        i = max((i >= a.n[1] - 2) ? a.n[1] - 2 : i, 0);
        j = max((j >= a.n[0] - 2) ? a.n[0] - 2 : j, 0);
        k = max((k >= a.n[2] - 2) ? a.n[2] - 2 : k, 0);
        float v[8];
        fetch8<float>(a, a.boff, j, i, k, v);
        out[t] = (float)__float_as_uint(v[0]);
    } else if (op == OP_FIRSTVOXEL) { /* out: rc, x, y, z (integers as bits) */
        const float *r = in + DDA_HEAD + t * 6;
        int v[3] = {0, 0, 0};
        const int rc = first_voxel(in, r, r + 3, v);
        out[t * 4 + 0] = (float)rc;
        for (int ax = 0; ax < 3; ax++) out[t * 4 + 1 + ax] = __int_as_float(v[ax]);
    } else if (op == OP_VOXELS) { /* out: rc, then maxs offsets */
        const float *r = in + DDA_HEAD + t * 6;
        float *res = out + (size_t)t * (maxs + 1);
        res[0] = (float)voxel_walk(in, (const int *)(in + 6), r, r + 3, maxs, res + 1);
    } else if (op == OP_TF) { /* one transfer function over in[0..n_s) from colour out[0..3] */
        Acc c = {out[0], out[1], out[2], out[3]};
        const int ns = maxs;
        const bool back = tf_backward(a.xfer);
        for (int q = 0; q < ns; q++) tf_step(a, a.lut, a.xfer, in[back ? ns - 1 - q : q], c);
        out[0] = c.r;
        out[1] = c.g;
        out[2] = c.b;
        out[3] = c.a;
    }
}

/* raycast (ray.c:59-85) of one ray per wavefront: the lanes sample the
 * ray's (at most 256) positions and fetch their LUT entries in parallel into
 * LDS, then one lane applies the transfer function over them in the
 * reference's order (back to front for SEISMIC2, CHEM and RGBBLEND), so a
 * call waits for one round of gathers instead of 256 dependent ones.  Same
 * float operations as march_compat. */
__global__ __launch_bounds__(64) void k_raycast_rays(const RenderArgs a, int n, const float *in, float *out)
{
    __shared__ float s_val[256];
    __shared__ float4 s_ent[256];
    const int r = blockIdx.x, lane = threadIdx.x;
    if (r >= n) return;
    const float *o = in + r * 6, *d = o + 3;
    float tin, tout;
    int iend = 0;
    if (slab(a.dconst, o, d, tin, tout)) {
        const float dt = 0.001f; /* DELTASAMPLE (types.h:17) */
        iend = 1.0f + (tout - tin) / dt + 0.5f;
        iend = iend < 256 ? iend : 256;
        for (int q = lane; q < iend; q += 64) {
            const float t = tin + ((float)q * dt);
            const float v = sample_ref<float>(a, a.boff, d[0] * t + o[0], d[1] * t + o[1], d[2] * t + o[2]);
            const int idx = tf_lut_index(a, a.xfer, v);
            s_val[q] = v;
            s_ent[q] = idx >= 0 ? a.lut[idx] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    __syncthreads();
    if (lane) return;
    Acc c = {0.0f, 0.0f, 0.0f, 0.0f};
    const bool backward = tf_backward(a.xfer);
    for (int q = 0; q < iend; q++) {
        const int si = backward ? iend - 1 - q : q;
        tf_apply(a, a.xfer, s_val[si], s_ent[si], c);
    }
    out[r * 4 + 0] = c.r;
    out[r * 4 + 1] = c.g;
    out[r * 4 + 2] = c.b;
    out[r * 4 + 3] = c.a;
}

/* compositetiles (main.c:1203-1218) into the existing tilebuf contents */
__global__ void k_over(float4 *dst, const float4 *src, int num, int npix)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    float4 d = dst[p];
    for (int k = num - 1; k >= 0; k--) {
        const float4 s = src[(size_t)k * npix + p];
        d.z = s.z + (1.0f - s.w) * d.z;
        d.y = s.y + (1.0f - s.w) * d.y;
        d.x = s.x + (1.0f - s.w) * d.x;
        d.w = s.w + (1.0f - s.w) * d.w;
    }
    dst[p] = d;
}

/* run op over n items: upload `in`, launch, download `out` */
int run_op(int op, int n, const float *in, size_t in_floats, float *out, size_t out_floats, int maxs,
           int lutsize, float alpha)
{
    int rc;
    RenderArgs a;
    if (op == OP_FIRSTVOXEL || op == OP_VOXELS)
        memset(&a, 0, sizeof a); /* the walk reads only its input */
    else
        op_args(a, lutsize, alpha);
    if (in_floats + out_floats <= (1u << 20)) { /* per-ray sizes: through mapped host memory */
        if ((rc = host_io(in_floats + out_floats))) return rc;
        float *h_in = g.h_io, *h_out = g.h_io + in_floats;
        memcpy(h_in, in, in_floats * sizeof(float));
        if (op == OP_TF) memcpy(h_out, out, 4 * sizeof(float));
        if (op == OP_RAYCAST)
            hipLaunchKernelGGL(k_raycast_rays, dim3(n), dim3(64), 0, g.ctx->stream, a, n, g.h_io_dev,
                               g.h_io_dev + in_floats);
        else
            hipLaunchKernelGGL(k_ops, dim3((n + 63) / 64), dim3(64), 0, g.ctx->stream, a, op, n, g.h_io_dev,
                               g.h_io_dev + in_floats, maxs);
        IVR_HIP(hipGetLastError());
        IVR_HIP(hipStreamSynchronize(g.ctx->stream));
        memcpy(out, h_out, out_floats * sizeof(float));
        return IVR_OK;
    }
    if ((rc = io_buffer(in_floats + out_floats))) return rc;
    float *d_in = g.d_io, *d_out = g.d_io + in_floats;
    if (in_floats) IVR_HIP(hipMemcpyAsync(d_in, in, in_floats * sizeof(float), hipMemcpyHostToDevice, g.ctx->stream));
    if (op == OP_TF)
        IVR_HIP(hipMemcpyAsync(d_out, out, 4 * sizeof(float), hipMemcpyHostToDevice, g.ctx->stream));
    hipLaunchKernelGGL(k_ops, dim3((n + 63) / 64), dim3(64), 0, g.ctx->stream, a, op, n, d_in, d_out, maxs);
    IVR_HIP(hipGetLastError());
    IVR_HIP(hipMemcpyAsync(out, d_out, out_floats * sizeof(float), hipMemcpyDeviceToHost, g.ctx->stream));
    IVR_HIP(hipStreamSynchronize(g.ctx->stream));
    return IVR_OK;
}

int prepare(int lut_entries)
{
    int rc = ensure_volume();
    if (rc) return rc;
    return upload_lut(lut, lut_entries > 0 ? lut_entries : g.hsize);
}

void report(int rc, const char *fn)
{
    if (rc) fprintf(stderr, "libivr_compat: %s failed: %s\n", fn, ivr_last_error());
}

/* input head of the voxel-walk ops: aabb min, max, localdims, then the ray */
void dda_input(const ivr_ray3f &ray, float *in)
{
    memset(in, 0, sizeof(float) * (DDA_HEAD + 6));
    const float box[6] = {aabb.min.x, aabb.min.y, aabb.min.z, aabb.max.x, aabb.max.y, aabb.max.z};
    const int dims[3] = {volume.localdims.x, volume.localdims.y, volume.localdims.z};
    memcpy(in, box, sizeof box);
    memcpy(in + 6, dims, sizeof dims);
    const float r[6] = {ray.orig.x, ray.orig.y, ray.orig.z, ray.dir.x, ray.dir.y, ray.dir.z};
    memcpy(in + DDA_HEAD, r, sizeof r);
}

/* LUT entries a transfer function touches for these samples */
int lut_extent(int x, const float *s, int n)
{
    int hi = 0;
    for (int i = 0; i < n; i++) {
        int idx = (int)s[i] - (x == 4 ? 0 : lutoffset);
        if (x == 2 && idx > maxlutindx) idx = maxlutindx;
        hi = std::max(hi, idx);
    }
    return hi + 1;
}

/* raycast calls waiting to be combined into one launch */
struct RayCall {
    float ray[6];
    float out[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    bool done = false;
};
struct RayQueue {
    std::mutex mu;
    std::condition_variable cv;
    std::vector<RayCall *> queue;
    bool busy = false; /* a caller is casting a batch */
} g_rays;

/* one launch of OP_RAYCAST over the batch's rays */
void cast_batch(const std::vector<RayCall *> &batch)
{
    std::lock_guard<std::mutex> lk(g.mu);
    const int n = (int)batch.size();
    std::vector<float> in((size_t)n * 6), out((size_t)n * 4, 0.0f);
    for (int k = 0; k < n; k++) memcpy(&in[(size_t)k * 6], batch[k]->ray, 6 * sizeof(float));
    int rc = prepare(0);
    if (!rc) rc = run_op(OP_RAYCAST, n, in.data(), in.size(), out.data(), out.size(), 0, g.hsize, mystate.alpha);
    report(rc, "raycast");
    for (int k = 0; k < n; k++) memcpy(batch[k]->out, &out[(size_t)k * 4], 4 * sizeof(float));
    g.ray_batches++;
    g.rays_cast += (unsigned long long)n;
}

ivr_color4f tf_call(int x, ivr_color4f *l, ivr_color4f c, float *s, int n, float alpha)
{
    std::lock_guard<std::mutex> lk(g.mu);
    int rc = ensure_volume();
    const int ext = x == 5 ? 1 : lut_extent(x, s, n);
    if (!rc) rc = upload_lut(x == 5 ? &c : l, ext);
    if (!rc && n > 0) {
        const int save_x = g.xfer;
        g.xfer = x;
        float o[4] = {c.r, c.g, c.b, c.a};
        rc = run_op(OP_TF, 1, s, (size_t)n, o, 4, n, ext, alpha);
        g.xfer = save_x;
        if (!rc) c = ivr_color4f{o[0], o[1], o[2], o[3]};
    }
    report(rc, "transfer function");
    return c;
}

} // namespace

extern "C" {

void setplaneconst(float *planeconst, float *newconst)
{
    for (int i = 0; i < 6; i++) {
        dconst[i] = (newconst[i] < planeconst[i]) ? newconst[i] : planeconst[i];
        g.newconst[i] = newconst[i];
    }
}

void initsamplefunc(int type)
{
    g.sampler = type;
    getsamplefunc = type == 0 ? &getnearestsample : &getlinearsample;
}

void inittransferfunc(int xfer)
{
    g.xfer = (xfer >= 0 && xfer <= 5) ? xfer : 0;
    static ivr_color4f (*const fns[6])(ivr_color4f *, ivr_color4f, float *, int, float) = {
        seismictransferfunc, medtransferfunc, seismictransferfunc2,
        chemistrytransferfunc, blendtransferfunc, rgbtransferfunc};
    transferfunc = fns[g.xfer];
}

void raycast(ivr_ray3f ray, ivr_color4f *color)
{
    /* Calls from concurrent threads (the OpenMP pixel loop of createlocalimage,
     * main.c:983-1009) are combined: the first caller to find no launch in
     * progress takes every queued ray (its own included) and casts them in one
     * launch; the others wait for their results.  Each ray's result is the one
     * a launch of it alone gives. */
    RayCall call;
    const float r[6] = {ray.orig.x, ray.orig.y, ray.orig.z, ray.dir.x, ray.dir.y, ray.dir.z};
    memcpy(call.ray, r, sizeof r);
    std::unique_lock<std::mutex> lk(g_rays.mu);
    g_rays.queue.push_back(&call);
    while (!call.done) {
        if (g_rays.busy) {
            g_rays.cv.wait(lk);
            continue;
        }
        g_rays.busy = true;
        std::vector<RayCall *> batch;
        batch.swap(g_rays.queue);
        lk.unlock();
        cast_batch(batch);
        lk.lock();
        for (RayCall *c : batch) c->done = true;
        g_rays.busy = false;
        g_rays.cv.notify_all();
    }
    lk.unlock();
    color->r = call.out[0];
    color->g = call.out[1];
    color->b = call.out[2];
    color->a = call.out[3];
}

int getsamplesonray(ivr_ray3f ray, int maxsamples, float *samples)
{
    std::lock_guard<std::mutex> lk(g.mu);
    float r[6] = {ray.orig.x, ray.orig.y, ray.orig.z, ray.dir.x, ray.dir.y, ray.dir.z};
    std::vector<float> o((size_t)maxsamples + 1, 0.0f);
    int rc = prepare(0);
    if (!rc) rc = run_op(OP_SAMPLES, 1, r, 6, o.data(), o.size(), maxsamples, g.hsize, mystate.alpha);
    report(rc, "getsamplesonray");
    if (rc) return -1;
    const int n = (int)o[0];
    if (n > 0) memcpy(samples, o.data() + 1, sizeof(float) * (size_t)n);
    return n;
}

bool getrayintersect(ivr_ray3f ray, float *tin, float *tout)
{
    std::lock_guard<std::mutex> lk(g.mu);
    float r[6] = {ray.orig.x, ray.orig.y, ray.orig.z, ray.dir.x, ray.dir.y, ray.dir.z};
    float o[3] = {0, 0, 0};
    int rc = prepare(0);
    if (!rc) rc = run_op(OP_INTERSECT, 1, r, 6, o, 3, 0, g.hsize, mystate.alpha);
    report(rc, "getrayintersect");
    *tin = o[1];
    *tout = o[2];
    return o[0] != 0.0f;
}

int getfirstvoxel(ivr_ray3f ray, ivr_ref_vector3i *first)
{
    std::lock_guard<std::mutex> lk(g.mu);
    float in[DDA_HEAD + 6], o[4] = {-1.0f, 0.0f, 0.0f, 0.0f};
    dda_input(ray, in);
    int rc = ensure_ctx();
    if (!rc) rc = run_op(OP_FIRSTVOXEL, 1, in, DDA_HEAD + 6, o, 4, 0, 0, 0.0f);
    report(rc, "getfirstvoxel");
    if (rc || o[0] != 0.0f) return -1;
    memcpy(&first->x, &o[1], sizeof(int));
    memcpy(&first->y, &o[2], sizeof(int));
    memcpy(&first->z, &o[3], sizeof(int));
    return 0;
}

void getvoxelsonray(ivr_ray3f ray, int maxdepth, float *samples)
{
    if (maxdepth <= 0) return;
    std::lock_guard<std::mutex> lk(g.mu);
    float in[DDA_HEAD + 6];
    dda_input(ray, in);
    std::vector<float> o((size_t)maxdepth + 1, -1.0f);
    int rc = ensure_ctx();
    if (!rc) rc = run_op(OP_VOXELS, 1, in, DDA_HEAD + 6, o.data(), o.size(), maxdepth, 0, 0.0f);
    report(rc, "getvoxelsonray");
    if (!rc && o[0] == 0.0f) memcpy(samples, o.data() + 1, sizeof(float) * (size_t)maxdepth);
}

float getlinearsample(float x, float y, float z)
{
    std::lock_guard<std::mutex> lk(g.mu);
    float p[3] = {x, y, z}, o = 0.0f;
    int rc = prepare(0);
    if (!rc) rc = run_op(OP_LINEAR, 1, p, 3, &o, 1, 0, g.hsize, mystate.alpha);
    report(rc, "getlinearsample");
    return o;
}

float getnearestsample(float x, float y, float z)
{
    std::lock_guard<std::mutex> lk(g.mu);
    float p[3] = {x, y, z}, o = 0.0f;
    int rc = prepare(0);
    if (!rc) rc = run_op(OP_NEAREST, 1, p, 3, &o, 1, 0, g.hsize, mystate.alpha);
    report(rc, "getnearestsample");
    return o;
}

ivr_color4f seismictransferfunc(ivr_color4f *l, ivr_color4f c, float *s, int n, float a) { return tf_call(0, l, c, s, n, a); }
ivr_color4f medtransferfunc(ivr_color4f *l, ivr_color4f c, float *s, int n, float a) { return tf_call(1, l, c, s, n, a); }
ivr_color4f seismictransferfunc2(ivr_color4f *l, ivr_color4f c, float *s, int n, float a) { return tf_call(2, l, c, s, n, a); }
ivr_color4f chemistrytransferfunc(ivr_color4f *l, ivr_color4f c, float *s, int n, float a) { return tf_call(3, l, c, s, n, a); }
ivr_color4f blendtransferfunc(ivr_color4f *l, ivr_color4f c, float *s, int n, float a) { return tf_call(4, l, c, s, n, a); }
ivr_color4f rgbtransferfunc(ivr_color4f *l, ivr_color4f c, float *s, int n, float a) { return tf_call(5, l, c, s, n, a); }

int ivr_compat_createlocalimage(int rank)
{
    (void)rank;
    std::lock_guard<std::mutex> lk(g.mu);
    int rc = prepare(0);
    if (rc) return rc;
    ivr_state st;
    host_state(&st);
    const int W = st.winwidth, H = st.winheight;
    if (!colorbuf) return ivr_fail(IVR_ERR_ARG, "host colorbuf is NULL");
    if (g.fr && (g.fr->W != W || g.fr->H != H)) {
        ivr_frame_destroy(g.ctx, g.fr);
        g.fr = nullptr;
    }
    if (!g.fr && (rc = ivr_frame_create(g.ctx, W, H, 1, &g.fr))) return rc;
    ivr_frame_desc d;
    ivr_desc_config(&d);
    d.state = st;
    d.mode = IVR_MODE_COMPAT;
    d.xfer = g.xfer;
    d.lutscale = 1.0f;
    d.lutoffset = lutoffset;
    d.maxlutindx = maxlutindx;
    if ((rc = ivr_render(g.ctx, g.vol, g.tf, &d, g.fr))) return rc;
    return ivr_frame_download(g.ctx, g.fr, nullptr, (float *)colorbuf);
}

int ivr_compat_colormap(int id, ivr_color4f *l, int hsize, float hmin, float hmax)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (!l || hsize < 1) return ivr_fail(IVR_ERR_ARG, "colormap of %d entries into %p", hsize, (void *)l);
    int rc = ensure_ctx();
    if (rc) return rc;
    if ((rc = io_buffer((size_t)hsize * 4))) return rc;
    IVR_HIP(hipMemcpyAsync(g.d_io, l, (size_t)hsize * 16, hipMemcpyHostToDevice, g.ctx->stream));
    if ((rc = ivr_colormap_apply(g.ctx->stream, (float4 *)g.d_io, id, hsize, hmin, hmax))) return rc;
    IVR_HIP(hipMemcpyAsync(l, g.d_io, (size_t)hsize * 16, hipMemcpyDeviceToHost, g.ctx->stream));
    IVR_HIP(hipStreamSynchronize(g.ctx->stream));
    return IVR_OK;
}

int ivr_compat_compositetiles(float *tilebuf, float *rcolorbuf, int num, int w, int h)
{
    std::lock_guard<std::mutex> lk(g.mu);
    int rc = ensure_ctx();
    if (rc) return rc;
    const size_t npix = (size_t)w * h;
    if ((rc = io_buffer(npix * 4 * (size_t)(num + 1)))) return rc;
    float *d_dst = g.d_io, *d_src = g.d_io + npix * 4;
    IVR_HIP(hipMemcpyAsync(d_dst, tilebuf, npix * 16, hipMemcpyHostToDevice, g.ctx->stream));
    IVR_HIP(hipMemcpyAsync(d_src, rcolorbuf, npix * 16 * (size_t)num, hipMemcpyHostToDevice, g.ctx->stream));
    hipLaunchKernelGGL(k_over, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, g.ctx->stream,
                       (float4 *)d_dst, (const float4 *)d_src, num, (int)npix);
    IVR_HIP(hipGetLastError());
    IVR_HIP(hipMemcpyAsync(tilebuf, d_dst, npix * 16, hipMemcpyDeviceToHost, g.ctx->stream));
    IVR_HIP(hipStreamSynchronize(g.ctx->stream));
    return IVR_OK;
}

void ivr_compat_raycast_stats(unsigned long long *launches, unsigned long long *rays)
{
    std::lock_guard<std::mutex> lk(g.mu);
    if (launches) *launches = g.ray_batches;
    if (rays) *rays = g.rays_cast;
}

void ivr_compat_invalidate(void)
{
    std::lock_guard<std::mutex> lk(g.mu);
    g.vol_src = nullptr;
}

} // extern "C"
