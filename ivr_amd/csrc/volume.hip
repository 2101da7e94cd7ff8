/*
 * volume.hip -- device volumes of libivr.so: Morton-ordered bricks with a
 * one-voxel ghost layer, synthetic generation, upload/download, exact
 * quantile selection and 16^3 macrocell value ranges.
 *
 * Replaces the reference's volume ingest and decomposition (readvoldata /
 * readfromfile main.c:629-706, 1220-1269; getlocalbrick utils.c:332-388;
 * histogram mathutil.c:444-470): instead of one brick per MPI rank with a
 * one-sided overlap, every brick of the volume lives in HBM with a ghost
 * layer on all six sides, so a sample's 8-voxel footprint is always inside
 * the brick that owns its floor index.
 *
 * Brick layout: brick (bx,by,bz) owns floor indices [b*B, b*B+B) per axis
 * (the last brick also owns the clamped tail); its storage coordinates are
 * local = global - b*B + 1 in [0, B+2), with out-of-volume ghosts clamped to
 * the edge voxel, in rows (brick_elem); padding elements are zero and
 * never read.  Bricks are stored in Morton order of
 * their coordinates; d_boff maps the linear brick id to its element offset.
 */
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <vector>

#include "ivr_internal.h"
#include "volgen.h"

namespace {

typedef _Float16 h16;

constexpr int MINMAX_PER = 16;

struct BrickGeom {
    int n[3], nb[3], B, bshift, P;
    unsigned long long stride;
};

/* storage coordinates l of element e of brick bid (the inverse of
 * brick_elem), its brick coordinates b and the clamped global voxel gl;
 * returns false for a padding element */
__device__ __forceinline__ bool elem_coords(const BrickGeom &g, int bid, int e, int l[3], int gl[3],
                                            int b[3])
{
    brick_coords(g.P, (unsigned)e, l);
    b[0] = bid % g.nb[0];
    b[1] = (bid / g.nb[0]) % g.nb[1];
    b[2] = bid / (g.nb[0] * g.nb[1]);
    for (int r = 0; r < 3; r++) {
        int x = b[r] * g.B + l[r] - 1;
        gl[r] = min(max(x, 0), g.n[r] - 1);
    }
    return l[0] < g.P && l[1] < g.P && l[2] < g.P;
}

/* is element (brick b, local l) the one copy of its voxel that is counted */
__device__ __forceinline__ bool counted(const BrickGeom &g, const int l[3], const int b[3])
{
    for (int r = 0; r < 3; r++) {
        const int x = b[r] * g.B + l[r] - 1;
        if (x < 0 || x >= g.n[r]) return false;
        const bool interior = l[r] >= 1 && l[r] <= g.B;
        const bool tail = l[r] == g.B + 1 && b[r] == g.nb[r] - 1;
        if (!interior && !tail) return false;
    }
    return true;
}

/* linear fp32 (x fastest) -> bricks */
template <typename VT>
__global__ void k_scatter(BrickGeom g, const unsigned long long *boff, const float *lin, VT *dst)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int bid = blockIdx.y;
    if (e >= (int)g.stride) return;
    int l[3], gl[3], b[3];
    const bool real = elem_coords(g, bid, e, l, gl, b);
    const float v = real ? lin[((size_t)gl[2] * g.n[1] + gl[1]) * g.n[0] + gl[0]] : 0.0f;
    dst[boff[bid] + e] = (VT)v;
}

/* the row-pair copy of brick blockIdx.y (pair_elem): one thread per
 * (x, y, z), y < P-1, reading voxels (x, y, z) and (x, y+1, z) of the brick
 * storage and writing them side by side */
template <typename VT>
__global__ void k_pairs(BrickGeom g, const unsigned long long *boff, const VT *src, VT *dst)
{
    const int P = g.P;
    const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (unsigned)P * (unsigned)(P - 1) * (unsigned)P) return;
    const int x = (int)(e % (unsigned)P), y = (int)((e / (unsigned)P) % (unsigned)(P - 1)),
              z = (int)(e / ((unsigned)P * (unsigned)(P - 1)));
    const VT *b = src + boff[blockIdx.y];
    VT *d = dst + 2 * boff[blockIdx.y] + pair_elem(P, x, y, z);
    d[0] = b[brick_elem(P, x, y, z)];
    d[1] = b[brick_elem(P, x, y + 1, z)];
}

/* bricks -> linear fp32 */
template <typename VT>
__global__ void k_gather(BrickGeom g, const unsigned long long *boff, const VT *src, float *lin)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)g.n[0] * g.n[1] * g.n[2];
    if (idx >= total) return;
    int gl[3];
    gl[0] = (int)(idx % g.n[0]);
    gl[1] = (int)((idx / g.n[0]) % g.n[1]);
    gl[2] = (int)(idx / ((size_t)g.n[0] * g.n[1]));
    int b[3], l[3];
    for (int r = 0; r < 3; r++) {
        b[r] = min(gl[r] >> g.bshift, g.nb[r] - 1);
        l[r] = gl[r] - b[r] * g.B + 1;
    }
    const int bid = b[0] + g.nb[0] * (b[1] + g.nb[1] * b[2]);
    lin[idx] = (float)src[boff[bid] + brick_elem(g.P, l[0], l[1], l[2])];
}

/* Central-difference gradient field (the reference's volume.gradient,
 * gradient3d mathutil.c:242-328, with its store offset i*j*k replaced by the
 * voxel's own offset).  Per axis: (f[+1] - f[-1]) / 2 inside the volume,
 * (f[1] - f[0]) / 2 and (f[n-1] - f[n-2]) / 2 on its two faces, each axis by
 * its own position (the reference fills face interiors only; its last-z face
 * takes f[n-2] - f[n-3], read here as f[n-1] - f[n-2] like the other faces).
 * w = sqrtf(x*x + y*y + z*z) (mag3f, mathutil.c:233-236).  Output x fastest.
 *
 * A workgroup covers a 64 x 8 (x, y) tile, one wavefront per row, and
 * marches GRAD_ZC slices along z keeping f(z-1), f(z), f(z+1) in registers,
 * so a voxel row is fetched from HBM about 1.5 times (its x / y neighbours
 * mostly come from the caches); the output is written with non-temporal
 * stores (6 % faster than ordinary ones).  Dealing each XCD a contiguous
 * range of tiles cut the fetched bytes from 0.80 to 0.59 GB for 512^3 but
 * was 5 % slower, so tiles keep the hardware order.  A chunk of GRAD_ZC slices never crosses a brick face (B
 * is a power of two >= 16); every read lies in the voxel's own brick, whose
 * storage coordinates run over [0, B+2) while a voxel's are in [1, B+1]. */
#define GRAD_ZC 16
template <typename VT>
__global__ __launch_bounds__(512) void k_gradient(BrickGeom g, const unsigned long long *boff,
                                                  const VT *src, float4 *out)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 8 + (threadIdx.x >> 6);
    const int z0 = blockIdx.z * GRAD_ZC;
    if (x >= g.n[0] || y >= g.n[1]) return;
    const int z1 = min(z0 + GRAD_ZC, g.n[2]);
    const int gl[3] = {x, y, z0};
    int b[3], l[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        b[r] = min(gl[r] >> g.bshift, g.nb[r] - 1);
        l[r] = gl[r] - b[r] * g.B + 1;
    }
    const int bid = b[0] + g.nb[0] * (b[1] + g.nb[1] * b[2]);
    const VT *p = src + boff[bid] + brick_elem(g.P, l[0], l[1], l[2]);
    const int PP = g.P * g.P;
    const int ux = x < g.n[0] - 1 ? 1 : 0, dx = x > 0 ? 1 : 0;
    const int uy = y < g.n[1] - 1 ? g.P : 0, dy = y > 0 ? g.P : 0;
    float fz = (float)p[0];
    float fm = z0 > 0 ? (float)p[-PP] : fz;
    float4 *o = out + ((size_t)z0 * g.n[1] + y) * g.n[0] + x;
    const size_t ostep = (size_t)g.n[0] * g.n[1];
    typedef float f4v __attribute__((ext_vector_type(4)));
#pragma unroll
    for (int k = 0; k < GRAD_ZC; k++) {
        const int z = z0 + k;
        if (z >= z1) break;
        const float fp = z < g.n[2] - 1 ? (float)p[PP] : fz;
        const float gx = ((float)p[ux] - (float)p[-dx]) / 2.0f;
        const float gy = ((float)p[uy] - (float)p[-dy]) / 2.0f;
        const float gz = (fp - fm) / 2.0f;
        const f4v val = {gx, gy, gz, sqrtf(gx * gx + gy * gy + gz * gz)};
        __builtin_nontemporal_store(val, (f4v *)o); /* streamed out, not re-read here */
        o += ostep;
        p += PP;
        fm = fz;
        fz = fp;
    }
}

/* raw generator field into fp32 bricks */
__global__ void k_gen(BrickGeom g, const unsigned long long *boff, const GenTable *tab, float *dst)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int bid = blockIdx.y;
    if (e >= (int)g.stride) return;
    int l[3], gl[3], b[3];
    const bool real = elem_coords(g, bid, e, l, gl, b);
    dst[boff[bid] + e] = real ? gen_field(*tab, gl[0], gl[1], gl[2], g.n[0], g.n[1], g.n[2]) : 0.0f;
}

/* min/max key over counted voxels */
__global__ void k_minmax(BrickGeom g, const unsigned long long *boff, const float *src,
                         unsigned *mm)
{
    /* MINMAX_PER elements per thread, then one atomic pair per workgroup */
    __shared__ unsigned s_lo[4], s_hi[4];
    const int bid = blockIdx.y;
    unsigned lo = 0xffffffffu, hi = 0u;
    for (int q = 0; q < MINMAX_PER; q++) {
        const int e = (blockIdx.x * MINMAX_PER + q) * 256 + threadIdx.x;
        int l[3], gl[3], b[3];
        if (e < (int)g.stride && elem_coords(g, bid, e, l, gl, b) && counted(g, l, b)) {
            const unsigned k = gen_key(src[boff[bid] + e]);
            lo = min(lo, k);
            hi = max(hi, k);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (unsigned)__shfl_xor((int)lo, off, 64));
        hi = max(hi, (unsigned)__shfl_xor((int)hi, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        s_lo[threadIdx.x >> 6] = lo;
        s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        lo = min(min(s_lo[0], s_lo[1]), min(s_lo[2], s_lo[3]));
        hi = max(max(s_hi[0], s_hi[1]), max(s_hi[2], s_hi[3]));
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

/* one radix-select pass: histogram of key bits [shift, shift+bits) over
 * counted voxels whose higher bits equal `prefix` */
__global__ void k_radix(BrickGeom g, const unsigned long long *boff, const float *src, unsigned prefix,
                        unsigned pmask, int shift, unsigned dmask, unsigned *hist)
{
    __shared__ unsigned sh[2048];
    for (int t = threadIdx.x; t < 2048; t += blockDim.x) sh[t] = 0;
    __syncthreads();
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int bid = blockIdx.y;
    if (e < (int)g.stride) {
        int l[3], gl[3], b[3];
        if (elem_coords(g, bid, e, l, gl, b) && counted(g, l, b)) {
            const unsigned k = gen_key(src[boff[bid] + e]);
            if ((k & pmask) == prefix) atomicAdd(&sh[(k >> shift) & dmask], 1u);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t <= (int)dmask; t += blockDim.x)
        if (sh[t]) atomicAdd(&hist[t], sh[t]);
}

template <typename VT>
__global__ void k_normalise(BrickGeom g, const float *src, VT *dst, size_t total, int kind, float lo,
                            float hi, float scale)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    dst[e] = (VT)gen_normalise(src[e], kind, lo, hi, scale);
}

/* per-cell min/max over the footprint voxels [c 2^shift, min((c+1) 2^shift, n-1)] per axis */
template <typename VT>
__global__ void k_cellrange(BrickGeom g, const unsigned long long *boff, const VT *src, int nc0,
                            int nc1, int nc2, int shift, float2 *out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc0 * nc1 * nc2) return;
    const int cx = c % nc0, cy = (c / nc0) % nc1, cz = c / (nc0 * nc1);
    const int c3[3] = {cx, cy, cz};
    int lo[3], hi[3];
    for (int r = 0; r < 3; r++) {
        lo[r] = c3[r] << shift;
        hi[r] = min(lo[r] + (1 << shift), g.n[r] - 1);
    }
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (int z = lo[2]; z <= hi[2]; z++)
        for (int y = lo[1]; y <= hi[1]; y++)
            for (int x = lo[0]; x <= hi[0]; x++) {
                const int b0 = min(x >> g.bshift, g.nb[0] - 1), b1 = min(y >> g.bshift, g.nb[1] - 1),
                          b2 = min(z >> g.bshift, g.nb[2] - 1);
                const int bid = b0 + g.nb[0] * (b1 + g.nb[1] * b2);
                const float v = (float)src[boff[bid] + brick_elem(g.P, x - b0 * g.B + 1, y - b1 * g.B + 1,
                                                                  z - b2 * g.B + 1)];
                mn = fminf(mn, v);
                mx = fmaxf(mx, v);
            }
    out[c] = make_float2(mn, mx);
}

BrickGeom geom_of(const ivr_volume *v)
{
    BrickGeom g;
    for (int r = 0; r < 3; r++) {
        g.n[r] = v->n[r];
        g.nb[r] = v->nb[r];
    }
    g.B = v->B;
    g.bshift = v->bshift;
    g.P = v->P;
    g.stride = v->stride;
    return g;
}

unsigned morton3(unsigned x, unsigned y, unsigned z)
{
    unsigned m = 0;
    for (int b = 0; b < 10; b++)
        m |= ((x >> b) & 1u) << (3 * b) | ((y >> b) & 1u) << (3 * b + 1) | ((z >> b) & 1u) << (3 * b + 2);
    return m;
}

/* host tables of the generator: permutation shuffle first, then
 * cx, cy, cz, sigma, amplitude per blob, all from one PCG32 stream */
int make_table(const ivr_gen_params *gp, GenTable *t)
{
    if (gp->nblobs < 0 || gp->nblobs > IVR_MAX_BLOBS)
        return ivr_fail(IVR_ERR_ARG, "nblobs %d outside [0,%d]", gp->nblobs, IVR_MAX_BLOBS);
    if (gp->kind != IVR_GEN_BLOBS && gp->kind != IVR_GEN_NOISE_BLOBS)
        return ivr_fail(IVR_ERR_ARG, "unknown generator kind %d", gp->kind);
    memset(t, 0, sizeof(*t));
    t->kind = gp->kind;
    t->nblobs = gp->nblobs;
    t->octaves = gp->octaves;
    t->lacunarity = gp->lacunarity;
    t->gain = gp->gain;
    t->base_freq = gp->base_freq;
    uint64_t st = gp->seed;
    /* seed mixing identical to ivr_host.c's rng_seed */
    uint64_t z = st + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    st = z ^ (z >> 31);
    unsigned char p[256];
    for (int i = 0; i < 256; i++) p[i] = (unsigned char)i;
    for (int i = 255; i > 0; i--) {
        const int j = (int)(ivr_rng_u32(&st) % (uint32_t)(i + 1));
        const unsigned char tmp = p[i];
        p[i] = p[j];
        p[j] = tmp;
    }
    for (int i = 0; i < 512; i++) t->perm[i] = p[i & 255];
    auto unit = [&]() { return (float)(ivr_rng_u32(&st) >> 8) * (1.0f / 16777216.0f); };
    for (int b = 0; b < gp->nblobs; b++) {
        const float cx = 0.2f + 0.6f * unit();
        const float cy = 0.2f + 0.6f * unit();
        const float cz = 0.2f + 0.6f * unit();
        const float sg = 0.05f + 0.10f * unit();
        const float am = 0.5f + 0.5f * unit();
        t->blob[b][0] = cx;
        t->blob[b][1] = cy;
        t->blob[b][2] = cz;
        t->blob[b][3] = 1.0f / (2.0f * sg * sg);
        t->blob[b][4] = am;
    }
    return IVR_OK;
}

long long quantile_rank(const ivr_gen_params *gp, long long total)
{
    long long k = (long long)((double)gp->empty_frac * (double)total);
    if (k < 0) k = 0;
    if (k > total - 1) k = total - 1;
    return k;
}

} // namespace

int ivr_volume_create(ivr_ctx *ctx, int nx, int ny, int nz, int dtype, int brick, ivr_volume **out)
{
    if (!ctx || !out) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (nx < 2 || ny < 2 || nz < 2) return ivr_fail(IVR_ERR_ARG, "volume dims must be >= 2");
    if (dtype != IVR_F32 && dtype != IVR_F16) return ivr_fail(IVR_ERR_ARG, "dtype %d", dtype);
    if (brick < 16 || brick > 512 || (brick & (brick - 1)))
        return ivr_fail(IVR_ERR_ARG, "brick edge %d must be a power of two in [16,512]", brick);
    ivr_volume *v = new ivr_volume();
    v->n[0] = nx;
    v->n[1] = ny;
    v->n[2] = nz;
    v->dtype = dtype;
    v->B = brick;
    v->bshift = __builtin_ctz((unsigned)brick);
    v->P = brick + 2;
    for (int r = 0; r < 3; r++) v->nb[r] = ((v->n[r] - 2) >> v->bshift) + 1;
    v->nbricks = v->nb[0] * v->nb[1] * v->nb[2];
    if (v->nbricks > IVR_MAX_BRICKS) {
        delete v;
        return ivr_fail(IVR_ERR_ARG, "%d bricks exceed %d", v->nb[0] * v->nb[1] * v->nb[2], IVR_MAX_BRICKS);
    }
    v->stride = brick_stride(v->P);
    const size_t esz = dtype == IVR_F16 ? 2 : 4;
    v->bytes = v->stride * v->nbricks * esz;
    /* Morton order of brick coordinates -> storage slot */
    std::vector<std::pair<unsigned, int>> order;
    for (int z = 0; z < v->nb[2]; z++)
        for (int y = 0; y < v->nb[1]; y++)
            for (int x = 0; x < v->nb[0]; x++)
                order.push_back({morton3(x, y, z), x + v->nb[0] * (y + v->nb[1] * z)});
    std::sort(order.begin(), order.end());
    v->h_boff.assign(v->nbricks, 0);
    for (size_t s = 0; s < order.size(); s++) v->h_boff[order[s].second] = (unsigned long long)(s * v->stride);
    for (int r = 0; r < 3; r++) v->nc[r] = ((v->n[r] - 2) >> IVR_CELL_SHIFT) + 1;
    v->ncells = (size_t)v->nc[0] * v->nc[1] * v->nc[2];
    int rc;
    if ((rc = ivr_dalloc(ctx, &v->data, v->bytes)) || (rc = ivr_dalloc(ctx, &v->pairs, 2 * v->bytes)) ||
        (rc = ivr_dalloc(ctx, (void **)&v->d_boff, sizeof(unsigned long long) * v->nbricks)) ||
        (rc = ivr_dalloc(ctx, (void **)&v->d_cellrange, sizeof(float2) * v->ncells)) ||
        (rc = ivr_dalloc(ctx, (void **)&v->d_cellempty, v->ncells))) {
        ivr_volume_destroy(ctx, v);
        return rc;
    }
    IVR_HIP(hipMemcpyAsync(v->d_boff, v->h_boff.data(), sizeof(unsigned long long) * v->nbricks,
                           hipMemcpyHostToDevice, ctx->stream));
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    *out = v;
    return IVR_OK;
}

int ivr_volume_destroy(ivr_ctx *ctx, ivr_volume *v)
{
    if (!v) return IVR_OK;
    if (ctx) {
        (void)hipStreamSynchronize(ctx->stream);
        const size_t esz = v->dtype == IVR_F16 ? 2 : 4;
        ivr_dfree(ctx, v->data, v->stride * v->nbricks * esz);
        ivr_dfree(ctx, v->pairs, 2 * v->stride * v->nbricks * esz);
        ivr_dfree(ctx, v->d_boff, sizeof(unsigned long long) * v->nbricks);
        ivr_dfree(ctx, v->d_cellrange, sizeof(float2) * v->ncells);
        ivr_dfree(ctx, v->d_cellempty, v->ncells);
    }
    delete v;
    return IVR_OK;
}

int ivr_volume_info(const ivr_volume *v, int dims[3], int nbricks[3], int *brick, int *dtype,
                    unsigned long long *bytes)
{
    if (!v) return ivr_fail(IVR_ERR_ARG, "null volume");
    for (int r = 0; r < 3; r++) {
        if (dims) dims[r] = v->n[r];
        if (nbricks) nbricks[r] = v->nb[r];
    }
    if (brick) *brick = v->B;
    if (dtype) *dtype = v->dtype;
    if (bytes) *bytes = v->bytes;
    return IVR_OK;
}

int ivr_volume_cellrange(ivr_ctx *ctx, const ivr_volume *v, int shift, const int nc[3], float2 *out)
{
    const BrickGeom g = geom_of(v);
    const size_t n = (size_t)nc[0] * nc[1] * nc[2];
    const unsigned nb = (unsigned)((n + 127) / 128);
    if (v->dtype == IVR_F16)
        hipLaunchKernelGGL(k_cellrange<h16>, dim3(nb), dim3(128), 0, ctx->stream, g, v->d_boff,
                           (const h16 *)v->data, nc[0], nc[1], nc[2], shift, out);
    else
        hipLaunchKernelGGL(k_cellrange<float>, dim3(nb), dim3(128), 0, ctx->stream, g, v->d_boff,
                           (const float *)v->data, nc[0], nc[1], nc[2], shift, out);
    IVR_HIP(hipGetLastError());
    return IVR_OK;
}

int ivr_volume_cells_build(ivr_ctx *ctx, ivr_volume *v)
{
    return ivr_volume_cellrange(ctx, v, IVR_CELL_SHIFT, v->nc, v->d_cellrange);
}

/* the row-pair copy of every brick, after the bricks were written */
int ivr_volume_pairs_build(ivr_ctx *ctx, ivr_volume *v)
{
    const BrickGeom g = geom_of(v);
    const unsigned n = (unsigned)v->P * (unsigned)(v->P - 1) * (unsigned)v->P;
    const dim3 grid((n + 255) / 256, (unsigned)v->nbricks);
    if (v->dtype == IVR_F16)
        hipLaunchKernelGGL(k_pairs<h16>, grid, dim3(256), 0, ctx->stream, g, v->d_boff, (const h16 *)v->data,
                           (h16 *)v->pairs);
    else
        hipLaunchKernelGGL(k_pairs<float>, grid, dim3(256), 0, ctx->stream, g, v->d_boff, (const float *)v->data,
                           (float *)v->pairs);
    IVR_HIP(hipGetLastError());
    return IVR_OK;
}

int ivr_volume_upload(ivr_ctx *ctx, ivr_volume *v, const float *host)
{
    if (!ctx || !v || !host) return ivr_fail(IVR_ERR_ARG, "null argument");
    const size_t total = (size_t)v->n[0] * v->n[1] * v->n[2];
    ivr_dscope lin_buf(ctx);
    int rc = lin_buf.alloc(total * sizeof(float));
    if (rc) return rc;
    float *lin = (float *)lin_buf.p;
    IVR_HIP(hipMemcpyAsync(lin, host, total * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    const BrickGeom g = geom_of(v);
    dim3 grid((unsigned)((g.stride + 255) / 256), (unsigned)v->nbricks);
    if (v->dtype == IVR_F16)
        hipLaunchKernelGGL(k_scatter<h16>, grid, dim3(256), 0, ctx->stream, g, v->d_boff, lin, (h16 *)v->data);
    else
        hipLaunchKernelGGL(k_scatter<float>, grid, dim3(256), 0, ctx->stream, g, v->d_boff, lin,
                           (float *)v->data);
    IVR_HIP(hipGetLastError());
    if ((rc = ivr_volume_cells_build(ctx, v)) || (rc = ivr_volume_pairs_build(ctx, v))) return rc;
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    return IVR_OK;
}

int ivr_volume_download(ivr_ctx *ctx, const ivr_volume *v, float *host)
{
    if (!ctx || !v || !host) return ivr_fail(IVR_ERR_ARG, "null argument");
    const size_t total = (size_t)v->n[0] * v->n[1] * v->n[2];
    ivr_dscope lin_buf(ctx);
    int rc = lin_buf.alloc(total * sizeof(float));
    if (rc) return rc;
    float *lin = (float *)lin_buf.p;
    const BrickGeom g = geom_of(v);
    const unsigned nb = (unsigned)((total + 255) / 256);
    if (v->dtype == IVR_F16)
        hipLaunchKernelGGL(k_gather<h16>, dim3(nb), dim3(256), 0, ctx->stream, g, v->d_boff,
                           (const h16 *)v->data, lin);
    else
        hipLaunchKernelGGL(k_gather<float>, dim3(nb), dim3(256), 0, ctx->stream, g, v->d_boff,
                           (const float *)v->data, lin);
    IVR_HIP(hipGetLastError());
    IVR_HIP(hipMemcpyAsync(host, lin, total * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    return IVR_OK;
}

int ivr_volume_gradient(ivr_ctx *ctx, const ivr_volume *v, float *dst, int dst_on_device)
{
    if (!ctx || !v || !dst) return ivr_fail(IVR_ERR_ARG, "null argument");
    const size_t total = (size_t)v->n[0] * v->n[1] * v->n[2];
    float4 *out = (float4 *)dst;
    ivr_dscope staging(ctx);
    if (!dst_on_device) {
        const int rc = staging.alloc(total * sizeof(float4));
        if (rc) return rc;
        out = (float4 *)staging.p;
    }
    const BrickGeom g = geom_of(v);
    const dim3 grid((unsigned)((g.n[0] + 63) / 64), (unsigned)((g.n[1] + 7) / 8),
                    (unsigned)((g.n[2] + GRAD_ZC - 1) / GRAD_ZC));
    if (v->dtype == IVR_F16)
        hipLaunchKernelGGL(k_gradient<h16>, grid, dim3(512), 0, ctx->stream, g, v->d_boff,
                           (const h16 *)v->data, out);
    else
        hipLaunchKernelGGL(k_gradient<float>, grid, dim3(512), 0, ctx->stream, g, v->d_boff,
                           (const float *)v->data, out);
    IVR_HIP(hipGetLastError());
    if (!dst_on_device) {
        IVR_HIP(hipMemcpyAsync(dst, out, total * sizeof(float4), hipMemcpyDeviceToHost, ctx->stream));
        IVR_HIP(hipStreamSynchronize(ctx->stream));
    }
    return IVR_OK;
}

int ivr_volume_generate(ivr_ctx *ctx, ivr_volume *v, const ivr_gen_params *gp)
{
    if (!ctx || !v || !gp) return ivr_fail(IVR_ERR_ARG, "null argument");
    GenTable tab;
    int rc = make_table(gp, &tab);
    if (rc) return rc;
    const bool inplace = v->dtype == IVR_F32;
    const size_t total_e = v->stride * v->nbricks;
    ivr_dscope tab_buf(ctx), aux_buf(ctx), raw_buf(ctx);
    if ((rc = tab_buf.alloc(sizeof(GenTable))) || (rc = aux_buf.alloc(sizeof(unsigned) * 2050)) ||
        (!inplace && (rc = raw_buf.alloc(total_e * sizeof(float)))))
        return rc;
    GenTable *d_tab = (GenTable *)tab_buf.p;
    unsigned *d_aux = (unsigned *)aux_buf.p; /* [0..1] min/max keys, [2..2049] histogram */
    float *raw = inplace ? (float *)v->data : (float *)raw_buf.p;
    IVR_HIP(hipMemcpyAsync(d_tab, &tab, sizeof(GenTable), hipMemcpyHostToDevice, ctx->stream));
    const BrickGeom g = geom_of(v);
    dim3 grid((unsigned)((g.stride + 255) / 256), (unsigned)v->nbricks);
    hipLaunchKernelGGL(k_gen, grid, dim3(256), 0, ctx->stream, g, v->d_boff, d_tab, raw);
    IVR_HIP(hipGetLastError());
    const unsigned init_mm[2] = {0xffffffffu, 0u};
    IVR_HIP(hipMemcpyAsync(d_aux, init_mm, sizeof init_mm, hipMemcpyHostToDevice, ctx->stream));
    const dim3 grid_mm((unsigned)((g.stride + 256 * MINMAX_PER - 1) / (256 * MINMAX_PER)), (unsigned)v->nbricks);
    hipLaunchKernelGGL(k_minmax, grid_mm, dim3(256), 0, ctx->stream, g, v->d_boff, raw, d_aux);
    IVR_HIP(hipGetLastError());
    unsigned mm[2];
    IVR_HIP(hipMemcpyAsync(mm, d_aux, sizeof mm, hipMemcpyDeviceToHost, ctx->stream));
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    float lo = gen_unkey(mm[0]), hi = gen_unkey(mm[1]);
    if (gp->kind == IVR_GEN_NOISE_BLOBS) {
        /* exact rank-k value by 11/11/10-bit radix selection on the keys */
        const long long total = (long long)v->n[0] * v->n[1] * v->n[2];
        long long k = quantile_rank(gp, total);
        unsigned prefix = 0, pmask = 0;
        const int shifts[3] = {21, 10, 0}, bits[3] = {11, 11, 10};
        std::vector<unsigned> h(2048);
        for (int pass = 0; pass < 3; pass++) {
            const unsigned dmask = (1u << bits[pass]) - 1u;
            IVR_HIP(hipMemsetAsync(d_aux + 2, 0, sizeof(unsigned) * 2048, ctx->stream));
            hipLaunchKernelGGL(k_radix, grid, dim3(256), 0, ctx->stream, g, v->d_boff, raw, prefix, pmask,
                               shifts[pass], dmask, d_aux + 2);
            IVR_HIP(hipGetLastError());
            IVR_HIP(hipMemcpyAsync(h.data(), d_aux + 2, sizeof(unsigned) * 2048, hipMemcpyDeviceToHost,
                                   ctx->stream));
            IVR_HIP(hipStreamSynchronize(ctx->stream));
            unsigned dsel = 0;
            for (unsigned dgt = 0; dgt <= dmask; dgt++) {
                if (k < (long long)h[dgt]) { dsel = dgt; break; }
                k -= h[dgt];
            }
            prefix |= dsel << shifts[pass];
            pmask |= dmask << shifts[pass];
        }
        lo = gen_unkey(prefix);
    }
    const float scale = gp->value_scale > 0.0f ? gp->value_scale : 1.0f;
    const unsigned nbn = (unsigned)((total_e + 255) / 256);
    if (inplace)
        hipLaunchKernelGGL(k_normalise<float>, dim3(nbn), dim3(256), 0, ctx->stream, g, raw,
                           (float *)v->data, total_e, gp->kind, lo, hi, scale);
    else
        hipLaunchKernelGGL(k_normalise<h16>, dim3(nbn), dim3(256), 0, ctx->stream, g, raw, (h16 *)v->data,
                           total_e, gp->kind, lo, hi, scale);
    IVR_HIP(hipGetLastError());
    if ((rc = ivr_volume_cells_build(ctx, v)) || (rc = ivr_volume_pairs_build(ctx, v))) return rc;
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    return IVR_OK;
}

/* CPU evaluation of the same generator (same source functions) */
int ivr_generate_host(const ivr_gen_params *gp, int nx, int ny, int nz, float *out)
{
    if (!gp || !out || nx < 1 || ny < 1 || nz < 1) return ivr_fail(IVR_ERR_ARG, "bad argument");
    GenTable tab;
    int rc = make_table(gp, &tab);
    if (rc) return rc;
    const size_t total = (size_t)nx * ny * nz;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++)
                out[((size_t)z * ny + y) * nx + x] = gen_field(tab, x, y, z, nx, ny, nz);
    float lo = out[0], hi = out[0];
    for (size_t i = 1; i < total; i++) {
        lo = std::min(lo, out[i]);
        hi = std::max(hi, out[i]);
    }
    if (gp->kind == IVR_GEN_NOISE_BLOBS) {
        std::vector<float> tmp(out, out + total);
        const long long k = quantile_rank(gp, (long long)total);
        std::nth_element(tmp.begin(), tmp.begin() + k, tmp.end());
        lo = tmp[k];
    }
    const float scale = gp->value_scale > 0.0f ? gp->value_scale : 1.0f;
    for (size_t i = 0; i < total; i++) out[i] = gen_normalise(out[i], gp->kind, lo, hi, scale);
    return IVR_OK;
}
