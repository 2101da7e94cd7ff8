/*
 * luts.hip -- the reference's colour-table construction on the device:
 * the value histogram of a volume (histogram, mathutil.c:444-470), the five
 * colormap builders (colormap.c:21-181) and initvolluts' LUT set-up
 * (main.c:1075-1200: hsize = (int)(max - min + 1), lutoffset =
 * (int)(min + 0.5), maxlutindx = hsize - 1 as the MPI render server sets it).
 *
 * The histogram reads every voxel once from its owning brick (x fastest over
 * the volume, like k_gather); min and max are reduced as order-preserving
 * integer keys with one atomic pair per workgroup, then the counts of
 * (int)(v - min) go to LDS bins per workgroup (global atomics beyond 8192
 * bins).  The colormaps compute one LUT entry per thread with the
 * reference's arithmetic types (float where it computes in float, double
 * where its literals promote), so the tables are bit-identical to it; entries
 * a builder leaves unwritten are zero, as in a zero-filled table.
 */
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <vector>

#include "ivr_internal.h"

namespace {

#define HIST_LDS_BINS 8192

/* value of voxel idx (x fastest over the volume) from its owning brick */
struct VolRead {
    int n[3], nb[3], B, bshift, P;
    const unsigned long long *boff;
    const void *data;
    int f16;
    __device__ __forceinline__ float at(size_t idx) const
    {
        const int x = (int)(idx % n[0]), y = (int)((idx / n[0]) % n[1]), z = (int)(idx / ((size_t)n[0] * n[1]));
        const int bx = min(x >> bshift, nb[0] - 1), by = min(y >> bshift, nb[1] - 1), bz = min(z >> bshift, nb[2] - 1);
        const int bid = bx + nb[0] * (by + nb[1] * bz);
        const size_t e = boff[bid] + brick_elem(P, x - bx * B + 1, y - by * B + 1, z - bz * B + 1);
        return f16 ? (float)((const _Float16 *)data)[e] : ((const float *)data)[e];
    }
};

VolRead reader(const ivr_volume *v)
{
    VolRead r;
    for (int q = 0; q < 3; q++) {
        r.n[q] = v->n[q];
        r.nb[q] = v->nb[q];
    }
    r.B = v->B;
    r.bshift = v->bshift;
    r.P = v->P;
    r.boff = v->d_boff;
    r.data = v->data;
    r.f16 = v->dtype == IVR_F16;
    return r;
}

/* order-preserving unsigned key of a float */
__device__ __forceinline__ unsigned fkey(float f)
{
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float fkey_inv(unsigned k)
{
    const unsigned u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* fmin / fmax over every voxel (NaN ignored, as fmin does) */
__global__ __launch_bounds__(256) void k_minmax_keys(const VolRead r, size_t total, unsigned *mm)
{
    unsigned lo = 0xffffffffu, hi = 0u;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const float v = r.at(i);
        if (v != v) continue;
        const unsigned k = fkey(v);
        lo = min(lo, k);
        hi = max(hi, k);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (unsigned)__shfl_xor((int)lo, off, 64));
        hi = max(hi, (unsigned)__shfl_xor((int)hi, off, 64));
    }
    __shared__ unsigned s_lo[4], s_hi[4];
    if ((threadIdx.x & 63) == 0) {
        s_lo[threadIdx.x >> 6] = lo;
        s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicMin(&mm[0], min(min(s_lo[0], s_lo[1]), min(s_lo[2], s_lo[3])));
        atomicMax(&mm[1], max(max(s_hi[0], s_hi[1]), max(s_hi[2], s_hi[3])));
    }
}

/* counts of (int)(v - vmin) over every voxel (index clamped into [0, hsize)
 * where float rounding of a huge range would step past the last bin) */
__global__ __launch_bounds__(256) void k_histogram(const VolRead r, size_t total, float vmin, int hsize, int *hist)
{
    __shared__ int s_bin[HIST_LDS_BINS];
    const bool lds = hsize <= HIST_LDS_BINS;
    if (lds) {
        for (int t = threadIdx.x; t < hsize; t += 256) s_bin[t] = 0;
        __syncthreads();
    }
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const float v = r.at(i);
        if (v != v) continue;
        const int b = min(max((int)(v - vmin), 0), hsize - 1);
        if (lds) atomicAdd(&s_bin[b], 1);
        else atomicAdd(&hist[b], 1);
    }
    if (lds) {
        __syncthreads();
        for (int t = threadIdx.x; t < hsize; t += 256)
            if (s_bin[t]) atomicAdd(&hist[t], s_bin[t]);
    }
}

/* one tissue-class blend of medcolormap: weights (1 - a) p0 and a p1 */
__device__ __forceinline__ void blend2(float4 &e, float a, float p0, float p1, double r0, double g0, double b0,
                                       double a0, double r1, double g1, double b1, double a1)
{
    const float y1 = (1.0 - a) * p0;
    const float y2 = a * p1;
    e.x = (r0 * y1 + r1 * y2);
    e.y = (g0 * y1 + g1 * y2);
    e.z = (b0 * y1 + b1 * y2);
    e.w = (a0 * y1 + a1 * y2);
}

/* entry i of colormap `id` (BASE 0, RAINBOW 1, MEDICAL 2, SEISMIC 3, JET 4)
 * written over the table's current entry, field by field as the reference
 * writes it; float4 = (r, g, b, a) */
__global__ __launch_bounds__(256) void k_colormap(float4 *lut, int id, int hsize, float hmin, float hmax)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= hsize) return;
    float4 e = lut[i];
    if (id == 0) { /* basecolormap (colormap.c:24-57) */
        const float c = (float)(i + hmin) / hmax;
        e.x = e.y = e.z = c;
        if (c > 0.8) { e.x = 0.7f; e.y = 0.75f; e.z = 0.46; e.w = 0.07; }
        if (c < 0.3) { e.z = 0.6; e.w = 0.01; }
        if (c > 0.3 && c < 0.6) { e.x = 0.6; e.w = 0.01; }
        if (c > 0.6 && c < 0.8) { e.x = 0.6; e.y = 0.6; e.w = 0.005; }
    } else if (id == 1) { /* rainbowcolormap (colormap.c:59-74): += on the table's entries */
        const float dr = 1.0f / hsize;
        if (i == 0) {
            e.z = 1.0f;
            e.y = 0.0f;
            e.x = 0.0f;
        }
        e.x += dr;
        e.z += -dr;
        e.w = 1.0f;
    } else if (id == 2) { /* medcolormap (colormap.c:76-139), RANGE 4096 (types.h:43) */
        const float parm[4] = {0.04, 0.9, 0.9, 0.8};
        const float x = i / (float)(hsize - 1) * 4096.0f;
        if (0 <= x && x <= 484) {
            e.x = 0.1; e.y = 0.1; e.z = 0.1; e.w = 0.1 * parm[0];
        } else if (484 < x && x <= 545) {
            blend2(e, (x - 484.0) / (545.0 - 484.0), parm[0], parm[1], 0.6, 0.6, 0.8, 0.1, 0.8, 0.7, 0.0, 0.03125);
        } else if (545 < x && x <= 726) {
            e.x = 0.8; e.y = 0.7; e.z = 0.0; e.w = 0.03125 * parm[1];
        } else if (726 < x && x <= 819) {
            blend2(e, (x - 726.0) / (819.0 - 726.0), parm[1], parm[2], 0.8, 0.7, 0.0, 0.03125, 0.2, 0.0, 0.0, 0.1875);
        } else if (819 < x && x <= 883) {
            e.x = 0.2; e.y = 0.0; e.z = 0.0; e.w = 0.1875 * parm[2];
        } else if (883 < x && x <= 1535) {
            blend2(e, (x - 883.0) / (1535.0 - 883.0), parm[2], parm[3], 0.2, 0.0, 0.0, 0.1875, 1.0, 1.0, 0.9, 0.75);
        } else if (1535 < x) {
            e.x = 1.0; e.y = 1.0; e.z = 0.9; e.w = 0.75 * parm[3];
        }
    } else if (id == 3) { /* seismiccolormap (colormap.c:140-168): ramps through (1, h/2, h) */
        const float k0 = 1.0f, k1 = 0.5f * hsize, k2 = (float)hsize;
        const float rv[3] = {1.0f, 1.0f, 0.0f}, gv[3] = {0.0f, 1.0f, 0.0f}, bv[3] = {0.0f, 1.0f, 1.0f};
        const int p = i + 1; /* the ramp's integer position */
        float out[3] = {e.x, e.y, e.z};
        const float *v1s[3] = {rv, gv, bv};
#pragma unroll
        for (int ch = 0; ch < 3; ch++) {
            const float *v1 = v1s[ch];
            const float s1 = (v1[1] - v1[0]) / (k1 - k0);
            const float s2 = (v1[2] - v1[1]) / (k2 - k1);
            if (p >= k0 && p <= k1) out[ch] = v1[0] + p * s1;
            else if (p >= k1 && p <= k2) out[ch] = v1[1] + (p - k1) * s2;
        }
        e.x = out[0];
        e.y = out[1];
        e.z = out[2];
        e.w = 1.0f;
    } else if (id == 4) { /* jetcolormap (colormap.c:170-181) */
        const float q = (4.0f * i) / hsize;
        e.x = fmin(fmax(fmin(q - 1.5, -q + 4.5), 0.0f), 1.0f);
        e.y = fmin(fmax(fmin(q - 0.5, -q + 3.5), 0.0f), 1.0f);
        e.z = fmin(fmax(fmin(q + 0.5, -q + 2.5), 0.0f), 1.0f);
        e.w = 1.0f;
    }
    lut[i] = e;
}

} // namespace

int ivr_volume_histogram(ivr_ctx *ctx, const ivr_volume *vol, float *vmin, float *vmax, int *hsize, int *hist,
                         int cap)
{
    if (!ctx || !vol || !vmin || !vmax || !hsize) return ivr_fail(IVR_ERR_ARG, "null argument");
    const size_t total = (size_t)vol->n[0] * vol->n[1] * vol->n[2];
    const VolRead r = reader(vol);
    ivr_dscope mm_buf(ctx);
    int rc = mm_buf.alloc(2 * sizeof(unsigned));
    if (rc) return rc;
    unsigned *d_mm = (unsigned *)mm_buf.p;
    const unsigned init[2] = {0xffffffffu, 0u};
    IVR_HIP(hipMemcpyAsync(d_mm, init, sizeof init, hipMemcpyHostToDevice, ctx->stream));
    const unsigned grid = (unsigned)std::min<size_t>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_minmax_keys, dim3(grid), dim3(256), 0, ctx->stream, r, total, d_mm);
    IVR_HIP(hipGetLastError());
    unsigned mm[2];
    IVR_HIP(hipMemcpyAsync(mm, d_mm, sizeof mm, hipMemcpyDeviceToHost, ctx->stream));
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    if (mm[0] > mm[1]) return ivr_fail(IVR_ERR_ARG, "volume has no finite values");
    const float lo = fkey_inv(mm[0]), hi = fkey_inv(mm[1]);
    const float range = (hi - lo) + 1;
    if (!(range < 2147483647.0f)) return ivr_fail(IVR_ERR_ARG, "value range %g .. %g has no histogram", lo, hi);
    const int hs = (int)range;
    *vmin = lo;
    *vmax = hi;
    *hsize = hs;
    if (!hist || hs > cap) return IVR_OK;
    ivr_dscope h_buf(ctx);
    if ((rc = h_buf.alloc((size_t)hs * sizeof(int)))) return rc;
    IVR_HIP(hipMemsetAsync(h_buf.p, 0, (size_t)hs * sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_histogram, dim3(grid), dim3(256), 0, ctx->stream, r, total, lo, hs, (int *)h_buf.p);
    IVR_HIP(hipGetLastError());
    IVR_HIP(hipMemcpyAsync(hist, h_buf.p, (size_t)hs * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    return IVR_OK;
}

int ivr_colormap_apply(hipStream_t st, float4 *d_lut, int id, int hsize, float hmin, float hmax)
{
    if (id < 0 || id > 4) return IVR_OK; /* nomap: the entries stay as they are */
    hipLaunchKernelGGL(k_colormap, dim3((unsigned)((hsize + 255) / 256)), dim3(256), 0, st, d_lut, id, hsize, hmin,
                       hmax);
    IVR_HIP(hipGetLastError());
    return IVR_OK;
}

int ivr_tf_colormap(ivr_ctx *ctx, ivr_tf *tf, int id, int hsize, float hmin, float hmax)
{
    if (!ctx || !tf) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (hsize < 1 || hsize > IVR_MAX_LUT) return ivr_fail(IVR_ERR_ARG, "LUT size %d outside [1,%d]", hsize, IVR_MAX_LUT);
    tf->n = hsize;
    if (id < 0 || id > 4) return IVR_OK;
    IVR_HIP(hipMemsetAsync(tf->d_lut, 0, (size_t)hsize * sizeof(float4), ctx->stream));
    int rc = ivr_colormap_apply(ctx->stream, tf->d_lut, id, hsize, hmin, hmax);
    if (rc) return rc;
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    return IVR_OK;
}

int ivr_tf_download(ivr_ctx *ctx, const ivr_tf *tf, float *rgba, int n)
{
    if (!ctx || !tf || !rgba) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (n < 1 || n > tf->n) return ivr_fail(IVR_ERR_ARG, "%d entries of a %d-entry LUT", n, tf->n);
    IVR_HIP(hipMemcpyAsync(rgba, tf->d_lut, (size_t)n * sizeof(float4), hipMemcpyDeviceToHost, ctx->stream));
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    return IVR_OK;
}

int ivr_volume_luts(ivr_ctx *ctx, const ivr_volume *vol, int colormap_id, ivr_tf *tf, int *lutoffset,
                    int *maxlutindx, int *hsize_out)
{
    if (!ctx || !vol || !tf || !lutoffset || !maxlutindx) return ivr_fail(IVR_ERR_ARG, "null argument");
    float lo, hi;
    int hs;
    int rc = ivr_volume_histogram(ctx, vol, &lo, &hi, &hs, nullptr, 0);
    if (rc) return rc;
    if (hs < 1 || hs > IVR_MAX_LUT)
        return ivr_fail(IVR_ERR_UNSUPPORTED, "histogram of %d bins (values %g .. %g) exceeds %d LUT entries", hs, lo,
                        hi, IVR_MAX_LUT);
    if ((rc = ivr_tf_colormap(ctx, tf, colormap_id, hs, lo, hi))) return rc;
    *lutoffset = (int)(lo + 0.5);
    *maxlutindx = hs - 1;
    if (hsize_out) *hsize_out = hs;
    return IVR_OK;
}
