/*
 * render_dev.h -- device helpers of the render kernels, shared by render.hip
 * (frame kernels) and compat.hip (reference entry points run per ray or per
 * sample).  Every helper restates one piece of the reference render path in
 * its float operation order; contraction is off for the whole library, so
 * the only fused multiply-adds are the explicit fmaf calls.
 */
#ifndef IVR_RENDER_DEV_H
#define IVR_RENDER_DEV_H

#include <float.h>

#include "ivr_internal.h"

namespace ivr_dev {

typedef _Float16 h16;

template <typename VT>
__device__ __forceinline__ float ldv(const VT *p) { return (float)(*p); }

/* 8-voxel footprint at global floor index (j,i,k), j along x.  Order:
 * (j,i,k) (j+1,i,k) (j,i+1,k) (j+1,i+1,k) then the same at k+1. */
template <typename VT>
__device__ __forceinline__ void fetch8(const RenderArgs &a, const unsigned long long *s_boff, int j,
                                       int i, int k, float v[8])
{
    const int bid = (j >> a.bshift) + a.nb[0] * ((i >> a.bshift) + a.nb[1] * (k >> a.bshift));
    const VT *p = (const VT *)a.vox_data + s_boff[bid] +
                  brick_elem(a.P, (j & a.bmask) + 1, (i & a.bmask) + 1, (k & a.bmask) + 1);
    const int dy = a.P, dz = a.P * a.P;
    v[0] = ldv(p);
    v[1] = ldv(p + 1);
    v[2] = ldv(p + dy);
    v[3] = ldv(p + dy + 1);
    v[4] = ldv(p + dz);
    v[5] = ldv(p + dz + 1);
    v[6] = ldv(p + dz + dy);
    v[7] = ldv(p + dz + dy + 1);
}

__device__ __forceinline__ float dot3n(float n0, float n1, float n2, const float *v)
{
    return n0 * v[0] + n1 * v[1] + n2 * v[2];
}

/* getrayintersect (ray.c:276-318): planes -x,+x,-y,+y,-z,+z */
__device__ __forceinline__ bool slab(const float *dc, const float *o, const float *d, float &tin,
                                     float &tout)
{
    float ti = -FLT_MAX, to = FLT_MAX;
#pragma unroll
    for (int f = 0; f < 6; f++) {
        if (!(ti < to)) break;
        const float s = (f & 1) ? 1.0f : -1.0f;
        const float n0 = (f >> 1) == 0 ? s : 0.0f;
        const float n1 = (f >> 1) == 1 ? s : 0.0f;
        const float n2 = (f >> 1) == 2 ? s : 0.0f;
        float numer = dc[f] - dot3n(n0, n1, n2, o);
        float denom = dot3n(n0, n1, n2, d);
        if (denom == 0.0f) {
            if (numer < 0.0f) ti = to + 1.0f;
        } else {
            float th = numer / denom;
            if (denom > 0.0f) to = fminf(to, th);
            else ti = fmaxf(ti, th);
        }
    }
    tin = ti;
    tout = to;
    return (ti < to) && (ti > 0.0f || to > 0.0f);
}

__device__ __forceinline__ float lerp3(float a, float b, float t) { return fmaf(t, b - a, a); }

__device__ __forceinline__ float trilinear(const float v[8], float fx, float fy, float fz)
{
    float c00 = lerp3(v[0], v[1], fx);
    float c10 = lerp3(v[2], v[3], fx);
    float c01 = lerp3(v[4], v[5], fx);
    float c11 = lerp3(v[6], v[7], fx);
    float c0 = lerp3(c00, c10, fy);
    float c1 = lerp3(c01, c11, fy);
    return lerp3(c0, c1, fz);
}

__device__ __forceinline__ int lut_index(const RenderArgs &a, float s)
{
    int idx = (int)(s * a.lutscale) - a.lutoffset;
    return min(max(idx, 0), a.lutsize - 1);
}

/* config-mode sample position: global grid t_k = tin + k dt, voxel coords
 * clamped to [0, n-1], floor index clamped to n-2 */
struct Pos {
    int j, i, k;
    float fx, fy, fz;
};

/* sample k of a ray in voxel coordinates: g = gd t + go with
 * gd = dir * N/ext, go = (eye - min) * N/ext precomputed per ray */
__device__ __forceinline__ Pos cfg_pos(const RenderArgs &a, const float *go, const float *gd, float tin,
                                       float dt, int kk)
{
    const float t = fmaf((float)kk, dt, tin);
    /* clamp to [0, N-1]: v_med3 is min(max()) for the finite values here */
    const float gx = __builtin_amdgcn_fmed3f(fmaf(gd[0], t, go[0]), 0.0f, a.nm1f[0]);
    const float gy = __builtin_amdgcn_fmed3f(fmaf(gd[1], t, go[1]), 0.0f, a.nm1f[1]);
    const float gz = __builtin_amdgcn_fmed3f(fmaf(gd[2], t, go[2]), 0.0f, a.nm1f[2]);
    Pos p;
    p.j = min((int)gx, a.n[0] - 2);
    p.i = min((int)gy, a.n[1] - 2);
    p.k = min((int)gz, a.n[2] - 2);
    p.fx = gx - (float)p.j;
    p.fy = gy - (float)p.i;
    p.fz = gz - (float)p.k;
    return p;
}

/* per-ray voxel-space origin and direction of cfg_pos */
__device__ __forceinline__ void ray_voxel(const RenderArgs &a, const float *o, const float *d, float *go,
                                          float *gd)
{
#pragma unroll
    for (int r = 0; r < 3; r++) {
        go[r] = (o[r] - a.bmin[r]) * a.s[r];
        gd[r] = d[r] * a.s[r];
    }
}

/* CS: log2 of the macrocell edge in voxels (IVR_CELL_SHIFT for the plain
 * marches; the surrogate renderer uses finer cells, SG_CELL_SHIFT) */
template <int CS = IVR_CELL_SHIFT>
__device__ __forceinline__ int cell_of(const RenderArgs &a, const Pos &p)
{
    /* on the full-rate 24-bit multiplier: the cell counts per axis and the
     * partial sums are below 2^24, the index below 2^32 */
    return (int)((unsigned)(p.j >> CS) +
                 __umul24((unsigned)a.nc[0], (unsigned)(p.i >> CS) + __umul24((unsigned)a.nc[1], (unsigned)(p.k >> CS))));
}

/* First sample index past k on the grid t_k = tin + k dt whose cell may be
 * visible, given that sample k at p has the nonzero transparency code e
 * (its node of 2^(e-1) cells per axis is transparent): the sample just
 * before the ray's exit from that node, estimated from the exit faces (with
 * rdt ~ 1/dt), is checked to lie in the node, else the walk advances by one
 * sample.  Each axis index of the samples is monotone in k, so all samples
 * between lie in the node.  The result is at most kmax.  Only samples of
 * transparent cells are passed, whatever the estimate: it decides how many
 * steps a skip takes, never which samples are composited. */
template <int CS = IVR_CELL_SHIFT>
__device__ __forceinline__ int skip_empty(const RenderArgs &a, const float *go, const float *gd,
                                          const float *rgd, float tin, float dt, float rdt, const Pos &p,
                                          int e, int k, int kmax)
{
    const int sh = CS - 1 + e;
    const int cc[3] = {p.j >> sh, p.i >> sh, p.k >> sh};
    float te = FLT_MAX;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        if (gd[q] != 0.0f) {
            const float face = (float)((cc[q] + (gd[q] > 0.0f ? 1 : 0)) << sh);
            te = fminf(te, (face - go[q]) * rgd[q]);
        }
    }
    int kc = k + 1;
    const float qk = (te - tin) * rdt;
    if (qk > (float)(k + 1)) {
        const int cand = qk < (float)kmax ? (int)qk : kmax;
        if (cand > k + 1) {
            const Pos pv = cfg_pos(a, go, gd, tin, dt, cand - 1);
            if ((((pv.j ^ p.j) | (pv.i ^ p.i) | (pv.k ^ p.k)) >> sh) == 0) kc = cand;
        }
    }
    return kc;
}

/* per-pixel config-mode ray direction: M (u, v, 5) */
__device__ __forceinline__ void cfg_dir(const RenderArgs &a, int i, int j, float *d)
{
    const float uu = fmaf((float)j + 0.5f, a.du, a.l);
    const float vv = fmaf((float)i + 0.5f, a.dv, a.b);
#pragma unroll
    for (int r = 0; r < 3; r++) d[r] = fmaf(a.M[r * 3 + 0], uu, fmaf(a.M[r * 3 + 1], vv, a.m5[r]));
}

struct Acc {
    float r, g, b, a;
};

/* one composited sample (front to back, config mode) */
__device__ __forceinline__ void ftb_sample(const RenderArgs &a, const float4 *s_lut, const float v[8],
                                           const Pos &p, Acc &acc)
{
    const float s = trilinear(v, p.fx, p.fy, p.fz);
    const float4 e = s_lut[lut_index(a, s)];
    const float al = e.w * a.alpha;
    const float w = (1.0f - acc.a) * al;
    acc.r = fmaf(w, e.x, acc.r);
    acc.g = fmaf(w, e.y, acc.g);
    acc.b = fmaf(w, e.z, acc.b);
    acc.a = acc.a + w;
}

/* One config-mode ray on the global sample grid t_k = tin + k dt, k < n. */
struct Ray {
    float tin, dt, rdt; /* rdt = 1/dt, for estimates that are checked exactly */
    int n;
    float go[3], gd[3], rgd[3]; /* voxel-space ray g(t) = go + gd t */
};

/* slab test and sample count of the ray through pixel direction d; false
 * when it misses the volume */
__device__ __forceinline__ bool ray_setup(const RenderArgs &a, const float *o, const float *d, Ray &r)
{
    float tin, tout;
    if (!slab(a.dconst, o, d, tin, tout)) return false;
    r.tin = fmaxf(tin, 0.0f);
    const float dlen = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    r.dt = (a.vox * a.step_vox) / dlen;
    r.rdt = 1.0f / r.dt;
    const float nf = 1.0f + (tout - r.tin) / r.dt + 0.5f;
    /* guard for degenerate directions; never reached by a finite ray */
    r.n = nf < (float)IVR_MAX_STEPS ? (int)nf : IVR_MAX_STEPS;
    ray_voxel(a, o, d, r.go, r.gd);
#pragma unroll
    for (int q = 0; q < 3; q++) r.rgd[q] = 1.0f / r.gd[q];
    return true;
}

/* The 8-voxel footprint (fetch8's order) from the row-pair copy
 * (pair_elem): 4 consecutive elements (x,y) (x,y+1) (x+1,y) (x+1,y+1) at
 * byte offset o and the same at o + czp (the next slice), through a raw
 * buffer resource over the copy (offsets at or past its size return zeros
 * without a memory access), or from a pointer p. */
template <typename VT>
__device__ __forceinline__ void pair_unpack(const unsigned q0[4], const unsigned q1[4], float v[8])
{
    if (sizeof(VT) == 4) {
        v[0] = __uint_as_float(q0[0]); v[1] = __uint_as_float(q0[2]);
        v[2] = __uint_as_float(q0[1]); v[3] = __uint_as_float(q0[3]);
        v[4] = __uint_as_float(q1[0]); v[5] = __uint_as_float(q1[2]);
        v[6] = __uint_as_float(q1[1]); v[7] = __uint_as_float(q1[3]);
    } else {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        const h2 a = __builtin_bit_cast(h2, q0[0]), b = __builtin_bit_cast(h2, q0[1]);
        const h2 c = __builtin_bit_cast(h2, q1[0]), d = __builtin_bit_cast(h2, q1[1]);
        v[0] = (float)a.x; v[1] = (float)b.x; v[2] = (float)a.y; v[3] = (float)b.y;
        v[4] = (float)c.x; v[5] = (float)d.x; v[6] = (float)c.y; v[7] = (float)d.y;
    }
}

template <typename VT>
__device__ __forceinline__ void fetch8_pair_buf(__amdgpu_buffer_rsrc_t rs, unsigned o, unsigned czp, float v[8])
{
    unsigned q0[4], q1[4];
    if (sizeof(VT) == 4) {
        typedef unsigned u4 __attribute__((ext_vector_type(4)));
        const u4 x = __builtin_bit_cast(u4, __builtin_amdgcn_raw_buffer_load_b128(rs, o, 0, 0));
        const u4 y = __builtin_bit_cast(u4, __builtin_amdgcn_raw_buffer_load_b128(rs, o + czp, 0, 0));
        q0[0] = x.x; q0[1] = x.y; q0[2] = x.z; q0[3] = x.w;
        q1[0] = y.x; q1[1] = y.y; q1[2] = y.z; q1[3] = y.w;
    } else {
        typedef unsigned u2 __attribute__((ext_vector_type(2)));
        const u2 x = __builtin_bit_cast(u2, __builtin_amdgcn_raw_buffer_load_b64(rs, o, 0, 0));
        const u2 y = __builtin_bit_cast(u2, __builtin_amdgcn_raw_buffer_load_b64(rs, o + czp, 0, 0));
        q0[0] = x.x; q0[1] = x.y; q1[0] = y.x; q1[1] = y.y;
    }
    pair_unpack<VT>(q0, q1, v);
}

template <typename VT>
__device__ __forceinline__ void fetch8_pair_ptr(const char *p, unsigned czp, float v[8])
{
    unsigned q0[4], q1[4];
    __builtin_memcpy(q0, p, 4 * sizeof(VT));
    __builtin_memcpy(q1, p + czp, 4 * sizeof(VT));
    pair_unpack<VT>(q0, q1, v);
}

/* byte offset in the row-pair copy of the footprint of sample position q
 * (from the brick offsets: brick b's copy starts at element 2 boff[b]);
 * 64-bit for copies past 4 GiB */
template <typename VT>
__device__ __forceinline__ unsigned long long pair_offset(const RenderArgs &a, const unsigned long long *s_boff,
                                                          const Pos &q)
{
    const int bid = (q.j >> a.bshift) + a.nb[0] * ((q.i >> a.bshift) + a.nb[1] * (q.k >> a.bshift));
    return (2ull * s_boff[bid] + pair_elem(a.P, (q.j & a.bmask) + 1, (q.i & a.bmask) + 1, (q.k & a.bmask) + 1)) *
           sizeof(VT);
}

/* For the footprints of one brick (a sort-last run, first sample p0): the
 * brick's byte offset in the row-pair copy, and loc0 such that the offset
 * of (j,i,k) inside it is loc0 + 2 esz j + cyp i + czp k modulo 2^32 (two
 * 24-bit multiply-adds: cyp = 2 esz P and czp = 2 esz P (P-1) are below
 * 2^24 for bricks of at most 512 voxels).  The true offset inside a brick is
 * below 2^32, so the wrapped sum is exact. */
__device__ __forceinline__ void pair_run(const RenderArgs &a, const unsigned long long *s_boff, const Pos &p0,
                                         unsigned esz, unsigned long long &brick, unsigned &loc0)
{
    const int bx = p0.j >> a.bshift, by = p0.i >> a.bshift, bz = p0.k >> a.bshift;
    const int b = bx + (int)__umul24((unsigned)a.nb[0], (unsigned)(by + (int)__umul24((unsigned)a.nb[1], (unsigned)bz)));
    brick = 2ull * s_boff[b] * esz;
    const unsigned ox = 1u - ((unsigned)bx << a.bshift), oy = 1u - ((unsigned)by << a.bshift),
                   oz = 1u - ((unsigned)bz << a.bshift);
    loc0 = 2u * esz * (ox + (unsigned)a.P * (oy + (unsigned)(a.P - 1) * oz));
}


/* march_segment without macrocell skipping, software-pipelined: the next
 * sample's footprint is requested before the current sample is composited,
 * and a lane whose next sample shares its floor index requests one past the
 * end of the buffer (zeros, no memory traffic) and keeps its voxels, so the
 * reuse is kept with no branch around the loads (a branch makes the
 * compiler wait for every load at its join).  Same float operations and
 * counts as march_segment. */
template <typename VT>
__device__ __forceinline__ void march_segment_pipe(const RenderArgs &a, const float4 *s_lut,
                                                   const unsigned long long *s_boff, const Ray &r, int k0,
                                                   int k1, Acc &acc, unsigned &cnt, unsigned &fetched)
{
    if (k0 >= k1) return;
    /* footprints from the row-pair copy: two loads each instead of four
     * (C2 sort-last k_runs 1.05 -> 0.79 ms in a probe issuing two) */
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)a.pair_data, (short)0, (int)a.pair_bytes, 0x00020000);
    const unsigned esz = sizeof(VT), czp = 2u * esz * (unsigned)a.P * (unsigned)(a.P - 1);
    auto offset = [&](const Pos &q) { return (unsigned)pair_offset<VT>(a, s_boff, q); };
    int k = k0;
    Pos p = cfg_pos(a, r.go, r.gd, r.tin, r.dt, k);
    float v[8];
    fetch8_pair_buf<VT>(rs, offset(p), czp, v);
    for (;;) {
        const bool more = k + 1 < k1;
        const Pos pn = cfg_pos(a, r.go, r.gd, r.tin, r.dt, more ? k + 1 : k);
        const bool need = more && (pn.j != p.j || pn.i != p.i || pn.k != p.k);
        float vn[8];
        fetch8_pair_buf<VT>(rs, need ? offset(pn) : a.pair_bytes, czp, vn);
        __builtin_amdgcn_sched_barrier(0); /* keep the requests ahead of the compositing */
        ftb_sample(a, s_lut, v, p, acc);
        cnt++;
        fetched++;
        k++;
        if (acc.a >= a.ert || k >= k1) break;
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = need ? vn[q] : v[q];
        p = pn;
    }
}

/* Front-to-back march over samples [k0, k1) from acc, with early ray
 * termination.  One loop for skipping and sampling, so that within a
 * wavefront the rays crossing empty cells and the rays sampling advance
 * together.  cnt counts composited and skipped samples, fetched the
 * samples that were classified.  RUNB: every sample of the segment lies in
 * one brick (a sort-last run), whose footprints are addressed inside it
 * (pair_run). */
template <typename VT, bool SKIP, bool RUNB = false>
__device__ __forceinline__ void march_segment(const RenderArgs &a, const float4 *s_lut,
                                              const unsigned long long *s_boff, const Ray &r, int k0,
                                              int k1, Acc &acc, unsigned &cnt, unsigned &fetched)
{
    if (!SKIP && a.pair_bytes) {
        march_segment_pipe<VT>(a, s_lut, s_boff, r, k0, k1, acc, cnt, fetched);
        return;
    }
    /* footprints from the row-pair copy, two loads each, through the buffer
     * resource when the copy is below 4 GiB (uniform per launch), else through
     * pointers; a run's footprints addressed inside its brick (pair_run) */
    const bool buf = a.pair_bytes != 0;
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)a.pair_data, (short)0, (int)a.pair_bytes, 0x00020000);
    const unsigned esz = sizeof(VT), cyp = 2u * esz * (unsigned)a.P, czp = cyp * (unsigned)(a.P - 1);
    unsigned long long rbrick = 0;
    unsigned rloc0 = 0;
    if (RUNB && k0 < k1) pair_run(a, s_boff, cfg_pos(a, r.go, r.gd, r.tin, r.dt, k0), esz, rbrick, rloc0);
    int k = k0;
    int cj = -1, ci = -1, ck = -1; /* floor index of the voxels held in v */
    int cc = -1, e = 0;            /* macrocell of the code held in e      */
    float v[8];
    while (k < k1) {
        const Pos p = cfg_pos(a, r.go, r.gd, r.tin, r.dt, k);
        if (SKIP) {
            const int cell = cell_of(a, p);
            if (cell != cc) { /* the code is loaded once per macrocell visited */
                cc = cell;
                e = a.cellempty[cell];
            }
            if (e) { /* transparent node: jump past it */
                const int kc = skip_empty(a, r.go, r.gd, r.rgd, r.tin, r.dt, r.rdt, p, e, k, k1);
                cnt += (unsigned)(kc - k);
                k = kc;
                continue;
            }
        }
        if (p.j != cj || p.i != ci || p.k != ck) { /* consecutive samples often share their 8 voxels */
            if (RUNB) {
                const unsigned o = __umul24((unsigned)p.k, czp) + (__umul24((unsigned)p.i, cyp) + ((unsigned)p.j * (2u * esz) + rloc0));
                if (buf) fetch8_pair_buf<VT>(rs, (unsigned)rbrick + o, czp, v);
                else fetch8_pair_ptr<VT>((const char *)a.pair_data + rbrick + o, czp, v);
            } else {
                const unsigned long long o = pair_offset<VT>(a, s_boff, p);
                if (buf) fetch8_pair_buf<VT>(rs, (unsigned)o, czp, v);
                else fetch8_pair_ptr<VT>((const char *)a.pair_data + o, czp, v);
            }
            cj = p.j;
            ci = p.i;
            ck = p.k;
        }
        ftb_sample(a, s_lut, v, p, acc);
        cnt++;
        fetched++;
        k++;
        if (acc.a >= a.ert) break;
    }
}

/* front-to-back march of one ray over the global sample grid */
template <typename VT, bool SKIP>
__device__ __forceinline__ void march_config(const RenderArgs &a, const float4 *s_lut,
                                             const unsigned long long *s_boff, const float *o,
                                             const float *d, Acc &acc, unsigned &cnt,
                                             unsigned &fetched)
{
    Ray r;
    if (!ray_setup(a, o, d, r)) return;
    march_segment<VT, SKIP>(a, s_lut, s_boff, r, 0, r.n, acc, cnt, fetched);
}

__device__ __forceinline__ int brick_of(const RenderArgs &a, const Pos &p)
{
    return (p.j >> a.bshift) + a.nb[0] * ((p.i >> a.bshift) + a.nb[1] * (p.k >> a.bshift));
}

/* First sample index after k (k < r.n, owned by brick b) whose floor index
 * lies outside brick b, or r.n.  Each axis index of the samples is monotone
 * in k, so a brick owns one contiguous run of k: the run end is estimated
 * from the ray's crossing of the brick faces and then settled exactly by
 * galloping and bisection on brick_of. */
__device__ __forceinline__ int run_end(const RenderArgs &a, const Ray &r, int k, int b)
{
    const int bc[3] = {b % a.nb[0], (b / a.nb[0]) % a.nb[1], b / (a.nb[0] * a.nb[1])};
    float te = FLT_MAX;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        /* the first and last bricks own the clamped samples beyond the faces */
        const bool up = r.gd[q] > 0.0f;
        if (r.gd[q] == 0.0f || (up && bc[q] == a.nb[q] - 1) || (!up && bc[q] == 0)) continue;
        const float face = (float)((bc[q] + (up ? 1 : 0)) << a.bshift);
        te = fminf(te, (face - r.go[q]) * r.rgd[q]);
    }
    int e = r.n;
    const float qk = ceilf((te - r.tin) * r.rdt);
    if (qk < (float)r.n) e = max(k + 1, (int)qk);
    auto owned = [&](int kk) { return brick_of(a, cfg_pos(a, r.go, r.gd, r.tin, r.dt, kk)) == b; };
    int lo, hi; /* owned(lo), and hi == r.n or !owned(hi) */
    if (e < r.n && owned(e)) {
        lo = e; /* estimate short: gallop forward */
        for (int step = 1;; step <<= 1) {
            hi = min(lo + step, r.n);
            if (hi == r.n || !owned(hi)) break;
            lo = hi;
        }
    } else {
        hi = e; /* estimate at or past the end: gallop backward */
        lo = hi - 1;
        for (int step = 1; lo > k && !owned(lo); step <<= 1) {
            hi = lo;
            lo = max(hi - step, k);
        }
    }
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (owned(mid)) lo = mid;
        else hi = mid;
    }
    return hi;
}

/* Sort-last compositing of one ray without materialised tiles.  A brick's
 * partial is the front-to-back composite (from zero, early termination
 * inside the brick) of the samples whose floor index it owns; the partials
 * are combined far to near with compositetiles' d = s + (1 - s.a) d
 * (main.c:1206-1215, component order as there).  Each axis index of the
 * samples is monotone in k, so a brick owns one contiguous run of k and the
 * runs come in visibility order: the per-ray run order is the materialised
 * scheme's tile order restricted to the bricks the ray crosses, and bricks
 * it misses contribute exact zeros there. */

__device__ __forceinline__ void over(Acc &d, const Acc &s)
{
    d.b = s.b + (1.0f - s.a) * d.b;
    d.g = s.g + (1.0f - s.a) * d.g;
    d.r = s.r + (1.0f - s.a) * d.r;
    d.a = s.a + (1.0f - s.a) * d.a;
}

/* Front-to-back walk in one pass: the march goes through the runs in
 * order, each run's end found once by run_end (a saturated run is left at
 * its end), and the partials wait in LDS (s_part[run * blockDim.x +
 * thread], at most nb0+nb1+nb2-2 runs) until the far-to-near combination. */
template <typename VT, bool SKIP>
__device__ __forceinline__ void march_sort_last(const RenderArgs &a, const float4 *s_lut,
                                                const unsigned long long *s_boff, float4 *s_part,
                                                const float *o, const float *d, Acc &out, unsigned &cnt,
                                                unsigned &fetched)
{
    Ray r;
    if (!ray_setup(a, o, d, r)) return;
    const int stride = blockDim.x, tid = threadIdx.x;
    int nrun = 0, run_b = -1, run_e = 0; /* current run: brick, first index past it */
    Acc acc = {0.0f, 0.0f, 0.0f, 0.0f};
    int k = 0;
    int cj = -1, ci = -1, ck = -1; /* floor index of the voxels held in v */
    int cc = -1, e = 0;            /* macrocell of the code held in e      */
    float v[8];
    while (k < r.n) {
        const Pos p = cfg_pos(a, r.go, r.gd, r.tin, r.dt, k);
        if (k >= run_e) { /* first sample of the next run */
            if (run_b >= 0) s_part[(nrun++) * stride + tid] = make_float4(acc.r, acc.g, acc.b, acc.a);
            acc.r = acc.g = acc.b = acc.a = 0.0f;
            run_b = brick_of(a, p);
            run_e = run_end(a, r, k, run_b);
        }
        if (SKIP) {
            const int cell = cell_of(a, p);
            if (cell != cc) {
                cc = cell;
                e = a.cellempty[cell];
            }
            if (e) {
                /* transparent node: jump past it.  Its samples add exact
                 * zeros to the run's partial, and a run skipped whole has an
                 * all-zero partial, which `over` leaves without effect. */
                const int kc = skip_empty(a, r.go, r.gd, r.rgd, r.tin, r.dt, r.rdt, p, e, k, r.n);
                cnt += (unsigned)(kc - k);
                k = kc;
                continue;
            }
        }
        if (p.j != cj || p.i != ci || p.k != ck) {
            fetch8<VT>(a, s_boff, p.j, p.i, p.k, v);
            cj = p.j;
            ci = p.i;
            ck = p.k;
        }
        ftb_sample(a, s_lut, v, p, acc);
        cnt++;
        fetched++;
        /* early termination inside the brick: continue at the next run */
        k = acc.a >= a.ert ? run_e : k + 1;
    }
    if (run_b >= 0) s_part[(nrun++) * stride + tid] = make_float4(acc.r, acc.g, acc.b, acc.a);
    for (int q = nrun - 1; q >= 0; q--) {
        const float4 t = s_part[q * stride + tid];
        const Acc s = {t.x, t.y, t.z, t.w};
        over(out, s);
    }
}

/* The same images for brick grids whose run count would not fit in LDS:
 * the runs are walked from the far end, each run's start found by
 * bisection, so only the running combination is kept. */
template <typename VT, bool SKIP>
__device__ __forceinline__ void march_sort_last_rev(const RenderArgs &a, const float4 *s_lut,
                                                    const unsigned long long *s_boff, const float *o,
                                                    const float *d, Acc &out, unsigned &cnt,
                                                    unsigned &fetched)
{
    Ray r;
    if (!ray_setup(a, o, d, r)) return;
    int kend = r.n;
    while (kend > 0) {
        const int b = brick_of(a, cfg_pos(a, r.go, r.gd, r.tin, r.dt, kend - 1));
        int lo = 0, hi = kend - 1; /* first k of the run that ends at kend - 1 */
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (brick_of(a, cfg_pos(a, r.go, r.gd, r.tin, r.dt, mid)) == b) hi = mid;
            else lo = mid + 1;
        }
        Acc s = {0.0f, 0.0f, 0.0f, 0.0f};
        march_segment<VT, SKIP>(a, s_lut, s_boff, r, lo, kend, s, cnt, fetched);
        over(out, s);
        kend = lo;
    }
}

/* Run `run` (counted front to back) of one ray, composited from zero with
 * early termination inside the run; its start is found by walking the
 * boundaries before it.  Returns 0 when the ray has no such run (nothing
 * composited), 1 for a run followed by others, 2 for the ray's last run. */
template <typename VT, bool SKIP>
__device__ __forceinline__ int march_run(const RenderArgs &a, const float4 *s_lut,
                                         const unsigned long long *s_boff, const float *o, const float *d,
                                         int run, Acc &acc, unsigned &cnt, unsigned &fetched)
{
    Ray r;
    if (!ray_setup(a, o, d, r)) return 0;
    int k = 0;
    for (int t = 0; t < run && k < r.n; t++)
        k = run_end(a, r, k, brick_of(a, cfg_pos(a, r.go, r.gd, r.tin, r.dt, k)));
    if (k >= r.n) return 0;
    const int e = run_end(a, r, k, brick_of(a, cfg_pos(a, r.go, r.gd, r.tin, r.dt, k)));
    const int st = e >= r.n ? 2 : 1; /* known before the march */
    march_segment<VT, SKIP, true>(a, s_lut, s_boff, r, k, e, acc, cnt, fetched);
    return st;
}

/* getlinearsample, non-BGQ (ray.c:122-178): reference weights and axes */
template <typename VT>
__device__ __forceinline__ float sample_ref(const RenderArgs &a, const unsigned long long *s_boff,
                                            float x, float y, float z)
{
    float ratio = (x - a.bmin[0]) / a.ext[0];
    int j = (int)((float)a.n[0] * ratio);
    ratio = (y - a.bmin[1]) / a.ext[1];
    int i = (int)((float)a.n[1] * ratio);
    ratio = (z - a.bmin[2]) / a.ext[2];
    int k = (int)((float)a.n[2] * ratio);
    i = max(min(i, a.n[1] - 2), 0);
    j = max(min(j, a.n[0] - 2), 0);
    k = max(min(k, a.n[2] - 2), 0);
    float v[8];
    fetch8<VT>(a, s_boff, j, i, k, v);
    const float xp = x + fabsf(a.bmin[0]);
    const float yp = y + fabsf(a.bmin[1]);
    const float zp = z + fabsf(a.bmin[2]);
    const float x0 = ((float)j * a.ext[0]) / (float)a.n[0];
    const float y0 = ((float)i * a.ext[1]) / (float)a.n[1];
    const float z0 = ((float)k * a.ext[2]) / (float)a.n[2];
    const float xd = (fabsf(xp - x0) * a.ext[0]) / (float)a.n[0];
    const float yd = (fabsf(yp - y0) * a.ext[1]) / (float)a.n[1];
    const float zd = (fabsf(zp - z0) * a.ext[2]) / (float)a.n[2];
    /* v000=v[0], v100=(j,i+1)=v[2], v010=(j+1,i)=v[1], v110=v[3], ... */
    const float c00 = v[0] * (1.0f - xd) + v[2] * xd;
    const float c10 = v[1] * (1.0f - xd) + v[3] * xd;
    const float c01 = v[4] * (1.0f - xd) + v[6] * xd;
    const float c11 = v[5] * (1.0f - xd) + v[7] * xd;
    const float c0 = c00 * (1.0f - yd) + c10 * yd;
    const float c1 = c01 * (1.0f - yd) + c11 * yd;
    return c0 * (1.0f - zd) + c1 * zd;
}

/* the LUT entry one step of transfer function `x` (transferfuncs.c:24-168)
 * reads for sample value s; RGBBLEND (5) reads none (-1) */
__device__ __forceinline__ int tf_lut_index(const RenderArgs &a, int x, float s)
{
    if (x == 5) return -1;
    if (x == 2 || x == 3) { /* SEISMIC2 / CHEM */
        int idx = (int)(s * a.lutscale) - a.lutoffset;
        if (idx < 0) idx = 0;
        if (x == 2 && idx > a.maxlutindx) idx = a.maxlutindx;
        return min(idx, a.lutsize - 1);
    }
    if (x == 4) return min(max((int)(s * a.lutscale), 0), a.lutsize - 1); /* BLEND: no lutoffset */
    return min(max((int)(s * a.lutscale) - a.lutoffset, 0), a.lutsize - 1);
}

/* one step of transfer function `x` (transferfuncs.c:24-199) with sample
 * value s and its LUT entry e (tf_lut_index) onto colour c */
__device__ __forceinline__ void tf_apply(const RenderArgs &a, int x, float s, float4 e, Acc &c)
{
    if (x == 2 || x == 3) { /* SEISMIC2 / CHEM (transferfuncs.c:109-168) */
        const float al = e.w * a.alpha;
        const float om = 1.0f - al;
        c.b = om * c.b + al * e.z;
        c.g = om * c.g + al * e.y;
        c.r = om * c.r + al * e.x;
        c.a = al + om * c.a;
        if (c.a > 1.0f) c.a = 1.0f;
    } else if (x == 5) { /* RGBBLEND (transferfuncs.c:170-199) */
        const unsigned u = __float_as_uint(s);
        const float r = 1.0f / 255.0f * (float)((u >> 24) & 0xff);
        const float g = 1.0f / 255.0f * (float)((u >> 16) & 0xff);
        const float b = 1.0f / 255.0f * (float)((u >> 8) & 0xff);
        const float al = a.alpha, om = 1.0f - al;
        c.b = om * c.b + al * b;
        c.g = om * c.g + al * g;
        c.r = om * c.r + al * r;
        c.a = al + om * c.a;
        if (c.a > 1.0f) c.a = 1.0f;
    } else if (x == 1) { /* MEDXFER (transferfuncs.c:24-68) */
        c.b = ((double)e.z < 0.8) ? fmaxf(e.z, c.b) : c.b;
        c.g = ((double)e.y < 0.8) ? fmaxf(e.y, c.g) : c.g;
        c.r = ((double)e.x < 0.8) ? fmaxf(e.x, c.r) : c.r;
        c.a = e.w * a.alpha;
    } else if (x == 4) { /* BLEND (transferfuncs.c:70-84) */
        const float al = e.w;
        c.b = al * e.z + (1.0f - al) * c.b;
        c.g = al * e.y + (1.0f - al) * c.g;
        c.r = al * e.x + (1.0f - al) * c.r;
        c.a = al * a.alpha;
    } else { /* SEISMIC (transferfuncs.c:86-107) */
        const double lb = e.z, lg = e.y, lr = e.x;
        c.b = ((lb > 0.65) && (e.z < 1.0f)) ? 0.6f : c.b;
        c.g = ((lg > 0.65) && (e.y < 1.0f)) ? 0.0f : c.g;
        c.r = ((lr > 0.65) && (e.x < 1.0f)) ? 0.0f : c.r;
        c.b = ((lb > 0.5) && (e.z < 0.65f)) ? 0.9f : c.b;
        c.g = ((lg > 0.5) && (e.y < 0.65f)) ? 0.9f : c.g;
        c.r = ((lr > 0.5) && (e.x < 0.65f)) ? 0.9f : c.r;
        c.b = ((lb > 0.3) && (e.z < 0.5f)) ? 0.0f : c.b;
        c.g = ((lg > 0.3) && (e.y < 0.5f)) ? 0.0f : c.g;
        c.r = ((lr > 0.3) && (e.x < 0.5f)) ? 0.6f : c.r;
        c.a = 1.0f * a.alpha;
    }
}

/* one step of transfer function `x` with sample value s onto colour c; lut
 * is the RGBA table (LDS or global) */
__device__ __forceinline__ void tf_step(const RenderArgs &a, const float4 *s_lut, int x, float s, Acc &c)
{
    const int idx = tf_lut_index(a, x, s);
    const float4 e = idx >= 0 ? s_lut[idx] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    tf_apply(a, x, s, e, c);
}

/* transfer functions that walk the samples from the far end (BTF) */
__device__ __forceinline__ bool tf_backward(int x) { return x == 2 || x == 3 || x == 5; }

/* one compat-mode ray: getsamplesonray + the selected transfer function */
template <typename VT>
__device__ void march_compat(const RenderArgs &a, const float4 *s_lut, const unsigned long long *s_boff,
                             const float *o, const float *d, Acc &c, unsigned &cnt)
{
    float tin, tout;
    if (!slab(a.dconst, o, d, tin, tout)) return;
    const float dt = 0.001f; /* DELTASAMPLE (types.h:17) */
    int iend = 1.0f + (tout - tin) / dt + 0.5f;
    iend = iend < 256 ? iend : 256;
    if (iend <= 0) return;
    cnt += (unsigned)iend;
    const int x = a.xfer;
    const bool backward = tf_backward(x);
    for (int q = 0; q < iend; q++) {
        const int s_i = backward ? iend - 1 - q : q;
        const float t = tin + ((float)s_i * dt);
        const float px = d[0] * t + o[0];
        const float py = d[1] * t + o[1];
        const float pz = d[2] * t + o[2];
        tf_step(a, s_lut, x, sample_ref<VT>(a, s_boff, px, py, pz), c);
    }
}

__device__ __forceinline__ unsigned char quant8(float f)
{
    f = fminf(fmaxf(f, 0.0f), 1.0f);
    return (unsigned char)((double)f * 255.0 + 0.5);
}

__device__ __forceinline__ void stage(const RenderArgs &a, float4 *s_lut, unsigned long long *s_boff)
{
    for (int t = threadIdx.x; t < a.lutsize; t += blockDim.x) s_lut[t] = a.lut[t];
    for (int t = threadIdx.x; t < a.nbricks; t += blockDim.x) s_boff[t] = a.boff[t];
    __syncthreads();
}

/* Workgroup b -> pixel block (2^blk_shift wide, IVR_FRAME_BLOCK pixels);
 * returns false for padding.  Workgroups
 * are dealt to the 8 XCDs round robin (XCD = blockIdx % 8).  With a chunk
 * C > 0 each XCD renders runs of C consecutive blocks, so one L2 sees
 * neighbouring blocks, while the runs stay interleaved over the frame so
 * every XCD gets a share of the expensive centre.  Inside the block each
 * wavefront covers 2^s x 2^(6-s) pixels, s = wave_shift. */
__device__ __forceinline__ bool block_pixel(const RenderArgs &a, int b, int &i, int &j)
{
    int L = b;
    if (a.xcd_chunk > 0) {
        const int g = b >> 3, q = g / a.xcd_chunk, r = g - q * a.xcd_chunk;
        L = ((q << 3) + (b & 7)) * a.xcd_chunk + r;
    }
    if (L >= a.nblk) return false;
    const int bx = L % a.nblk_x, by = L / a.nblk_x;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s = a.wave_shift, ww = 1 << s, per_row = (1 << a.blk_shift) >> s;
    j = (bx << a.blk_shift) + (w % per_row) * ww + (lane & (ww - 1));
    i = by * (IVR_FRAME_BLOCK >> a.blk_shift) + (w / per_row) * (64 >> s) + (lane >> s);
    return true;
}

/* adds the wavefront's sum of v to counter k of this workgroup's stripe */
__device__ __forceinline__ void count_add(unsigned long long *ctrs, int k, unsigned v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v)
        atomicAdd(&ctrs[(blockIdx.x & (IVR_CTR_SLOTS - 1)) * IVR_CTRS + k], (unsigned long long)v);
}

} // namespace ivr_dev

#endif
