/*
 * surrogate.hip -- learned volume surrogate: per-brick multiresolution hash
 * grids (T. Mueller et al., SIGGRAPH 2022) feeding one shared MLP
 * 32 -> 64 -> 64 -> 1 (ReLU, linear output) whose two 64-wide layers run on
 * gfx950 bf16 MFMA (v_mfma_f32_16x16x32_bf16) with fp32 accumulation.
 * Trained with the kernels below (forward, backward, Adam) on samples drawn
 * from the plain volume; rendered by k_render_sg in place of the 8-voxel
 * trilinear fetch, brick by brick, with a per-brick fallback to the plain
 * volume.  No reference counterpart: this is an addition beside the render
 * path of the reference.  The hashed levels hash 4x4x4 blocks of vertices
 * to 64-entry slots (corner_index) rather than single vertices, so the 8
 * corners of a cell share cache lines.
 *
 * Wave-level MLP.  A wavefront evaluates 64 samples.  Activations are kept
 * transposed (features x samples): with the MFMA C layout (lane l holds rows
 * 4*(l/16)+0..3 of column l%16 of a 16x16 tile) a layer's output tile lands
 * in exactly the lanes that need it as the next layer's B operand (column =
 * sample l%16), once the K order of the next product is permuted to the
 * rows each lane group already holds.  The sum over K does not depend on its
 * order, so the weights (A operand) are stored once per step with the same
 * permutation and no activation ever moves between lanes.  Only the weight
 * gradients (a product over samples) stage their operands through LDS.
 *
 * Encoding work is spread the same way: lane l computes levels 4*(l/16) ..
 * 4*(l/16)+3 of the samples l%16, 16+l%16, 32+l%16, 48+l%16, which is the
 * B-operand fragment of the first layer; the gathers of all four sample
 * columns are in flight before any MLP work.
 *
 * Training runs in chunks of samples, two kernels per chunk: k_train_net
 * (encoding, forward, backward, weight gradients, and the gradient of the
 * loss with respect to every sample's encoding, written to a chunk buffer)
 * and k_train_scatter (that gradient added to the 8 corners of each cell at
 * every level, two fp32 atomics per corner, level and sample, by 16 lanes
 * per (sample, level) so that the 64 atomics of an instruction fall on the
 * few cache lines of four cells; coarse levels whose vertices under a
 * wavefront's samples form a small box are summed in LDS first).  A chunk's
 * scatter runs on a second stream while the next chunk's network pass runs.
 */
#include <cstddef>
#include <cstring>
#include <cmath>
#include <vector>

#include "render_dev.h"
#include "surrogate.h"

namespace {

using namespace ivr_dev;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma(const bf16x8 &a, const bf16x8 &b, const f32x4 &c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

/* K order of the permuted products: slot j of lane group g in k-step ks */
__host__ __device__ __forceinline__ int kperm(int ks, int g, int j) { return (2 * ks + (j >> 2)) * 16 + g * 4 + (j & 3); }

/* ---------------------------------------------------------------- encoding */

struct Enc {
    const __half2 *table16; /* [nbricks][L][T] feature pairs, the half-precision copy of the tables */
    unsigned T;
    int log2T;
    int res[SG_L];
    unsigned n1[SG_L];  /* N+1 for dense levels, 0 for hashed */
};

/* the level a lane group works on: l = a + step * g for lane group g, with
 * its resolution and dense edge picked from the four candidates by g, so
 * the level tables (kernel arguments) are only ever indexed uniformly */
struct Lvl {
    int l, res;
    unsigned n1;
};

__device__ __forceinline__ int sel4(int g, int v0, int v1, int v2, int v3)
{
    return g == 0 ? v0 : (g == 1 ? v1 : (g == 2 ? v2 : v3));
}

__device__ __forceinline__ Lvl lvl_of(const Enc &e, int g, int a, int step)
{
    Lvl L;
    L.l = a + step * g;
    L.res = sel4(g, e.res[a], e.res[a + step], e.res[a + 2 * step], e.res[a + 3 * step]);
    L.n1 = (unsigned)sel4(g, (int)e.n1[a], (int)e.n1[a + step], (int)e.n1[a + 2 * step], (int)e.n1[a + 3 * step]);
    return L;
}

__device__ __forceinline__ unsigned corner_index(unsigned T, const Lvl &L, int x, int y, int z)
{
    const unsigned n1 = L.n1;
    /* both forms computed and one selected: the lane groups of a wavefront
     * work on levels of both kinds, and a branch here would split the
     * gathers into separately scheduled blocks */
    const unsigned dense = (unsigned)x + n1 * ((unsigned)y + n1 * (unsigned)z);
    /* hashed levels: the 4x4x4 block of corners hashes to a slot of 64
     * entries (the spatial hash of Mueller et al. 2022 on block coordinates)
     * holding its corners at x1 x0 | y0 | z0 | y1 | z1 (low bits first): rows
     * of 4 along x are contiguous, so a cell's x-neighbour corners are one
     * 8-byte load unless x is the block's last, and a 64-byte group is a
     * 4x2x2 brick of corners, so a cell's 8 corners take about 2.3 cache
     * lines instead of about 4 */
    const unsigned bx = (unsigned)x >> 2, by = (unsigned)y >> 2, bz = (unsigned)z >> 2;
    const unsigned blk = (bx * 1u ^ by * 2654435761u ^ bz * 805459861u) & ((T >> 6) - 1u);
    const unsigned local = (x & 3) | (y & 1) << 2 | (z & 1) << 3 | (y & 2) << 3 | (z & 2) << 4;
    return n1 ? dense : (blk << 6 | local);
}

/* corner 0 and fractional weights of level L at local position x in [0,1] */
__device__ __forceinline__ void level_cell(const Lvl &L, const float *x, int *v0, float *w)
{
    const float N = (float)L.res;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float u = x[a] * N;
        int v = (int)floorf(u);
        v = min(max(v, 0), L.res - 1);
        v0[a] = v;
        w[a] = u - (float)v;
    }
}

/* level L of the encoding at local position x: trilinear blend of the 8
 * corner features of the half-precision copy of the tables (training and
 * rendering both read it: half the bytes per corner of the fp32 master
 * tables, which only the optimiser reads) */
__device__ __forceinline__ float2 encode_level(const Enc &e, int brick, const Lvl &L, const float *x)
{
    int v0[3];
    float w[3];
    level_cell(L, x, v0, w);
    const __half2 *tb = e.table16 + ((size_t)(brick * SG_L + L.l) << e.log2T);
    /* The entries of corner rows (y + dy, z + dz), r = dy + 2 dz, as
     * corner_index gives them, from per-level terms: dense levels
     * x + n1 y + n1^2 z plus n1 dy + n1^2 dz (24-bit products: n1^3 <= T);
     * hashed levels hash the block coordinates once, a row that crosses into
     * the next block adding the axis prime, and place the corner in its slot
     * from the low two bits of each coordinate.  Both are computed and one
     * selected per row, with no multiply per corner. */
    const unsigned X = (unsigned)v0[0], Y = (unsigned)v0[1], Z = (unsigned)v0[2];
    const unsigned n1 = L.n1, n2 = __umul24(n1, n1);
    const unsigned d0 = X + __umul24(n1, Y) + __umul24(n2, Z);
    const unsigned mask = (e.T >> 6) - 1u, bx = X >> 2;
    const unsigned hy0 = (Y >> 2) * 2654435761u, hz0 = (Z >> 2) * 805459861u;
    const unsigned hy[2] = {hy0, (Y & 3u) == 3u ? hy0 + 2654435761u : hy0};
    const unsigned hz[2] = {hz0, (Z & 3u) == 3u ? hz0 + 805459861u : hz0};
    const unsigned ly[2] = {(Y & 1u) << 2 | (Y & 2u) << 3, ((Y + 1u) & 1u) << 2 | ((Y + 1u) & 2u) << 3};
    const unsigned lz[2] = {(Z & 1u) << 3 | (Z & 2u) << 4, ((Z + 1u) & 1u) << 3 | ((Z + 1u) & 2u) << 4};
    const bool dense = n1 != 0u;
    /* the corners (x, y, z) and (x+1, y, z) are consecutive entries on dense
     * levels (x < N) and on hashed levels unless x is the last of its block:
     * one 8-byte load per x-pair, and for that case a second load of the
     * (x+1) corner, which lies at place 0 of the next block's slot, into
     * registers of its own (the half-precision copy has a 16-byte tail, so
     * the pair load of the very last entry stays inside the allocation) */
    const bool split = !dense && (X & 3u) == 3u;
    unsigned pr[4][2], ex[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int dy = r & 1, dz = r >> 1;
        const unsigned hashed = ((bx ^ hy[dy] ^ hz[dz]) & mask) << 6 | (X & 3u) | ly[dy] | lz[dz];
        const unsigned dn = d0 + (dy ? n1 : 0u) + (dz ? n2 : 0u);
        __builtin_memcpy(pr[r], tb + (dense ? dn : hashed), 8);
    }
    if (split) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int dy = r & 1, dz = r >> 1;
            __builtin_memcpy(&ex[r], tb + ((((bx + 1u) ^ hy[dy] ^ hz[dz]) & mask) << 6 | ly[dy] | lz[dz]), 4);
        }
    }
    /* corners in the order c = dx + 2 dy + 4 dz, weights multiplied x, y, z */
    float2 f = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int dy = r & 1, dz = r >> 1;
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
            const float wc = (dx ? w[0] : 1.0f - w[0]) * (dy ? w[1] : 1.0f - w[1]) * (dz ? w[2] : 1.0f - w[2]);
            const unsigned u = dx ? (split ? ex[r] : pr[r][1]) : pr[r][0];
            const float2 t = __half22float2(__builtin_bit_cast(__half2, u));
            f.x = fmaf(wc, t.x, f.x);
            f.y = fmaf(wc, t.y, f.y);
        }
    }
    return f;
}

/* the levels lane group l/16 encodes: lvl_of(e, l/16, q, 4) for q = 0..3 */
__device__ __forceinline__ void lane_levels(const Enc &e, Lvl lv[4])
{
    const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int q = 0; q < 4; q++) lv[q] = lvl_of(e, g, q, 4);
}

/* The first layer's B fragments of the wavefront's 64 samples (lane l's
 * own sample at local position x of brick b): E[nt] holds, for sample
 * nt*16 + l%16, the features 8g..8g+7 (levels lv[0..3] of lane group g),
 * rounded to bf16.  The gathers of all four columns are issued before any
 * is blended: measured on C4 and the 1024^3 surrogate at 1080p, one
 * wavefront per SIMD with every gather of the wavefront in flight (41.8 and
 * 31.0 ms per frame) beats two wavefronts each with two columns in flight
 * (48.7, 34.2 ms) and three wavefronts gathering one column at a time (53.9,
 * 36.4 ms). */
__device__ __forceinline__ void encode_cols(const Enc &enc, const Lvl lv[4], const float x[3], int b, bf16x8 E[4])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        const int src = nt * 16 + (lane & 15);
        const float xs[3] = {__shfl(x[0], src, 64), __shfl(x[1], src, 64), __shfl(x[2], src, 64)};
        const int bs = __shfl(b, src, 64);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float2 f = encode_level(enc, bs, lv[q], xs);
            E[nt][2 * q] = (__bf16)f.x;
            E[nt][2 * q + 1] = (__bf16)f.y;
        }
    }
}

/* ---------------------------------------------------------------- MLP */

struct Mlp {
    const bf16x8 *frag; /* [SG_NFRAG][64] */
    const float *p;     /* fp32 parameters */
};

/* What inference reads of the network, staged in LDS by each workgroup: the
 * A fragments of W1 and W2 (12 sets of 64 lanes) and the fp32 b1, b2, W3, b3.
 * Read from LDS at each evaluation, they are not held in registers across a
 * march (the render kernel's registers bound its occupancy). */
#define SG_INFER_FRAGS (SG_FRAG_G2 - SG_FRAG_F1)
#define SG_INFER_LDS (SG_INFER_FRAGS * 64 * 16 + (3 * SG_H + 1) * 4)
struct MlpLds {
    const bf16x8 *frag; /* [SG_INFER_FRAGS][64], sets numbered as in Mlp::frag */
    const float *b1, *b2, *w3, *b3;
};

/* stage m into lds (SG_INFER_LDS bytes, 16-byte aligned); the caller synchronises */
__device__ __forceinline__ MlpLds stage_mlp(const Mlp &m, void *lds)
{
    bf16x8 *f = (bf16x8 *)lds;
    float *q = (float *)(f + SG_INFER_FRAGS * 64);
    for (int t = threadIdx.x; t < SG_INFER_FRAGS * 64; t += blockDim.x) f[t] = m.frag[SG_FRAG_F1 * 64 + t];
    for (int t = threadIdx.x; t < 3 * SG_H + 1; t += blockDim.x)
        q[t] = t < SG_H ? m.p[SG_P_B1 + t] : t < 2 * SG_H ? m.p[SG_P_B2 + t - SG_H] : t < 3 * SG_H ? m.p[SG_P_W3 + t - 2 * SG_H] : m.p[SG_P_B3];
    MlpLds v;
    v.frag = f;
    v.b1 = q;
    v.b2 = q + SG_H;
    v.w3 = q + 2 * SG_H;
    v.b3 = q + 3 * SG_H;
    return v;
}

/* forward of 64 samples: returns y for samples nt*16+l%16 in y[nt]; keeps
 * the first-layer activations (bf16 fragments) and second-layer
 * pre-activations for the backward pass */
struct Fwd {
    bf16x8 H[2][4];
    f32x4 Z2[4][4];
    float y[4];
};

__device__ __forceinline__ void mlp_forward(const Mlp &m, const bf16x8 E[4], Fwd &f)
{
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    /* layer 1: Z1T[out][sample] = W1 E^T */
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        const bf16x8 a = m.frag[(SG_FRAG_F1 + mt) * 64 + lane];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 c = mfma(a, E[nt], zero);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float z = c[i] + m.p[SG_P_B1 + mt * 16 + g * 4 + i];
                f.H[mt >> 1][nt][(mt & 1) * 4 + i] = (__bf16)(z > 0.0f ? z : 0.0f);
            }
        }
    }
    /* layer 2: Z2T = W2 H1T with the permuted K order */
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        const bf16x8 a0 = m.frag[(SG_FRAG_F2 + mt * 2 + 0) * 64 + lane];
        const bf16x8 a1 = m.frag[(SG_FRAG_F2 + mt * 2 + 1) * 64 + lane];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 c = mfma(a0, f.H[0][nt], zero);
            c = mfma(a1, f.H[1][nt], c);
#pragma unroll
            for (int i = 0; i < 4; i++) c[i] += m.p[SG_P_B2 + mt * 16 + g * 4 + i];
            f.Z2[mt][nt] = c;
        }
    }
    /* layer 3 (fp32): y = W3 relu(Z2) + b3, reduced over the lane groups */
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        float acc = 0.0f;
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float z = f.Z2[mt][nt][i];
                acc = fmaf(m.p[SG_P_W3 + mt * 16 + g * 4 + i], z > 0.0f ? z : 0.0f, acc);
            }
        acc += __shfl_xor(acc, 16, 64);
        acc += __shfl_xor(acc, 32, 64);
        f.y[nt] = acc + m.p[SG_P_B3];
    }
}

/* The training forward of 64 samples: the first layer's activations H
 * (bf16 fragments, kept for the backward pass) and y for samples nt*16+l%16,
 * with the float operations of mlp_forward; the second layer's
 * pre-activations are folded into y tile by tile and recomputed by the
 * backward pass (a few dozen MFMAs per tile) instead of being held in 64
 * registers across it. */
__device__ __forceinline__ void mlp_forward_h(const Mlp &m, const bf16x8 E[4], bf16x8 H[2][4], float y[4])
{
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        const bf16x8 a = m.frag[(SG_FRAG_F1 + mt) * 64 + lane];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const f32x4 c = mfma(a, E[nt], zero);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float z = c[i] + m.p[SG_P_B1 + mt * 16 + g * 4 + i];
                H[mt >> 1][nt][(mt & 1) * 4 + i] = (__bf16)(z > 0.0f ? z : 0.0f);
            }
        }
    }
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        const bf16x8 a0 = m.frag[(SG_FRAG_F2 + mt * 2 + 0) * 64 + lane];
        const bf16x8 a1 = m.frag[(SG_FRAG_F2 + mt * 2 + 1) * 64 + lane];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            f32x4 c = mfma(a0, H[0][nt], zero);
            c = mfma(a1, H[1][nt], c);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float z = c[i] + m.p[SG_P_B2 + mt * 16 + g * 4 + i];
                acc[nt] = fmaf(m.p[SG_P_W3 + mt * 16 + g * 4 + i], z > 0.0f ? z : 0.0f, acc[nt]);
            }
        }
    }
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        float v = acc[nt];
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        y[nt] = v + m.p[SG_P_B3];
    }
}

/* The forward pass of inference for one column of 16 samples (the B
 * fragment E of samples nt*16 + l%16): y of mlp_forward for that column with
 * the same float operations in the same order (layer 3's sum runs over mt,
 * then i), each layer-2 output tile folded into y as soon as it is made. */
__device__ __forceinline__ float mlp_infer_col(const MlpLds &m, const bf16x8 &E)
{
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    bf16x8 H[2];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        const bf16x8 a = m.frag[mt * 64 + lane]; /* the staged sets start at SG_FRAG_F1 */
        const f32x4 c = mfma(a, E, zero);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float z = c[i] + m.b1[mt * 16 + g * 4 + i];
            H[mt >> 1][(mt & 1) * 4 + i] = (__bf16)(z > 0.0f ? z : 0.0f);
        }
    }
    float acc = 0.0f;
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        const bf16x8 a0 = m.frag[(SG_FRAG_F2 - SG_FRAG_F1 + mt * 2 + 0) * 64 + lane];
        const bf16x8 a1 = m.frag[(SG_FRAG_F2 - SG_FRAG_F1 + mt * 2 + 1) * 64 + lane];
        f32x4 c = mfma(a0, H[0], zero);
        c = mfma(a1, H[1], c);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float z = c[i] + m.b2[mt * 16 + g * 4 + i];
            acc = fmaf(m.w3[mt * 16 + g * 4 + i], z > 0.0f ? z : 0.0f, acc);
        }
    }
    acc += __shfl_xor(acc, 16, 64);
    acc += __shfl_xor(acc, 32, 64);
    return acc + *m.b3;
}

/* own sample's value from the y[nt] a lane holds (its sample is nt = l/16) */
__device__ __forceinline__ float own(const float v[4])
{
    const int g = (threadIdx.x & 63) >> 4;
    return g == 0 ? v[0] : (g == 1 ? v[1] : (g == 2 ? v[2] : v[3]));
}

/* The network value at each lane's sample (local position x in [0,1]^3 of
 * brick b; lv from lane_levels) as the renderer evaluates it: encoding from
 * the half-precision tables, rounded to bf16 for the first layer, MLP on
 * MFMA.  Every lane of the wavefront must call it (lanes without a sample
 * pass any valid point). */
__device__ __forceinline__ float sg_infer(const Enc &enc, const Lvl lv[4], const MlpLds &mlp, const float x[3], int b)
{
    const int g = (threadIdx.x & 63) >> 4;
    bf16x8 E[4];
    encode_cols(enc, lv, x, b, E);
    float yown = 0.0f;
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        const float y = mlp_infer_col(mlp, E[nt]);
        if (g == nt) yown = y; /* a lane's own sample is column l/16 */
    }
    return yown;
}

/* ---------------------------------------------------------------- training */

struct TrainArgs {
    RenderArgs va;        /* plain volume (brick table, dims) for targets */
    int snb[3], sbshift;  /* surrogate bricks: per axis, log2 of the edge */
    int snbricks;
    Enc enc;
    Mlp mlp;
    float *gtable;
    float *gparam;        /* fp32 [SG_NPARAM] */
    float *loss;          /* sum of squared errors */
    long long nsamples;   /* samples of this step */
    long long s0, nchunk; /* this launch: samples s0 .. s0 + nchunk - 1 of the step */
    float *dE;            /* [nchunk][SG_IN] fp32: d(loss)/d(encoding) of the chunk's samples */
    float inv_n;
    unsigned seed, step;
    const float *xs;      /* explicit samples: local x,y,z [n][3] (or null) */
    const int *bs;        /* explicit bricks [n] */
    const float *ts;      /* explicit targets [n] */
    float *ys;            /* explicit: outputs [n] (or null) */
};

#define SG_LDS_STRIDE 72 /* bf16 per staged row: 64 samples + pad */

/* Jarzynski & Olano, "Hash Functions for GPU Rendering" (JCGT 2020): pcg hash */
__device__ __forceinline__ unsigned pcg_hash(unsigned v)
{
    const unsigned state = v * 747796405u + 2891336453u;
    const unsigned word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

__device__ __forceinline__ float field_at(const RenderArgs &a, const unsigned long long *boff, float gx,
                                          float gy, float gz)
{
    gx = fminf(fmaxf(gx, 0.0f), a.nm1f[0]);
    gy = fminf(fmaxf(gy, 0.0f), a.nm1f[1]);
    gz = fminf(fmaxf(gz, 0.0f), a.nm1f[2]);
    const int j = min((int)gx, a.n[0] - 2), i = min((int)gy, a.n[1] - 2), k = min((int)gz, a.n[2] - 2);
    float v[8];
    if (a.dtype_f16) fetch8<_Float16>(a, boff, j, i, k, v);
    else fetch8<float>(a, boff, j, i, k, v);
    return trilinear(v, gx - (float)j, gy - (float)i, gz - (float)k);
}

/* p, with the compiler unable to see that it is the same value: address
 * arithmetic on it cannot move ahead of this point.  Keeps a persistent
 * kernel from computing its loop's (or its epilogue's) addresses once before
 * the loop and holding them, spilled, across it. */
template <typename T>
__device__ __forceinline__ T *opaque(T *p)
{
    asm volatile("" : "+s"(p));
    return p;
}

/* Sample sid of a training step: its surrogate brick b, local position x
 * in [0,1]^3 and, with t, its target (the plain volume's trilinear value).
 * Drawn samples go to the bricks in contiguous ranges of the step (brick
 * floor(sid * nbricks / n)), so a chunk of the step trains one or two
 * bricks, whose tables and gradients then stay in the caches.  Inside a
 * brick the samples are stratified: with K whole tiles of 64 samples in the
 * brick and m = floor(cbrt(K)), tile k < m^3 draws its samples uniformly in
 * cell k of an m^3 grid over the brick (the rest uniformly in the brick), so
 * each sample is uniform in its brick and the 64 samples of a wavefront lie
 * close together (their table gradients at the coarse levels are summed in
 * LDS by k_train_scatter).  Positions come from the pcg hash of (seed,
 * step, sid).  Explicit samples are read from A.xs, A.bs, A.ts. */
__device__ __forceinline__ void train_sample(const TrainArgs &A, long long sid, float x[3], int &b, float *t)
{
    if (A.xs) {
        x[0] = A.xs[sid * 3];
        x[1] = A.xs[sid * 3 + 1];
        x[2] = A.xs[sid * 3 + 2];
        b = A.bs[sid];
        if (t) *t = A.ts[sid];
        return;
    }
    const RenderArgs &va = A.va;
    const long long nb = A.snbricks, n = A.nsamples;
    b = (int)min(sid * nb / n, nb - 1);
    const long long first = ((long long)b * n + nb - 1) / nb, last = ((long long)(b + 1) * n + nb - 1) / nb;
    const long long k = (sid - first) >> 6, K = (last - first) >> 6;
    int m = (int)cbrtf((float)K);
    while ((long long)(m + 1) * (m + 1) * (m + 1) <= K) m++;
    while (m > 0 && (long long)m * m * m > K) m--;
    const bool strat = k < (long long)m * m * m;
    const int cell[3] = {strat ? (int)(k % m) : 0, strat ? (int)((k / m) % m) : 0, strat ? (int)(k / ((long long)m * m)) : 0};
    const float inv_m = strat ? 1.0f / (float)m : 1.0f;
    const int bc[3] = {b % A.snb[0], (b / A.snb[0]) % A.snb[1], b / (A.snb[0] * A.snb[1])};
    float gv[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const unsigned h = pcg_hash(A.seed ^ pcg_hash(A.step * 3u + (unsigned)r) ^ pcg_hash((unsigned)sid * 3u + r));
        const float v = (float)(h >> 8) * (1.0f / 16777216.0f);
        const float u = fminf(((float)cell[r] + v) * inv_m, 0x1.fffffep-1f); /* in [0, 1) */
        const int lo = bc[r] << A.sbshift;
        const float ext = (float)min(1 << A.sbshift, (va.n[r] - 1) - lo);
        gv[r] = (float)lo + u * ext;
        x[r] = u * ext / (float)(1 << A.sbshift);
    }
    if (t) *t = field_at(va, va.boff, gv[0], gv[1], gv[2]);
}

/* The network part of a training step over one chunk of samples: encoding
 * (all gathers of a wavefront in flight, as in rendering), MLP forward and
 * backward on MFMA, the weight gradients (products over samples, operands
 * staged in LDS), the loss, and d(loss)/d(encoding) of every sample written
 * to A.dE for k_train_scatter.  A workgroup is two wavefronts of 64 samples;
 * wavefront w keeps the output row tiles 2w and 2w+1 of both weight-gradient
 * products, summed over the samples of both wavefronts. */
__global__ __launch_bounds__(128) void k_train_net(const TrainArgs A)
{
    __shared__ __attribute__((aligned(16))) __bf16 lds[2][64 * SG_LDS_STRIDE * 3 + 32 * SG_LDS_STRIDE];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
    __bf16 *sDZ2 = lds[wave], *sH1 = sDZ2 + 64 * SG_LDS_STRIDE, *sDZ1 = sH1 + 64 * SG_LDS_STRIDE,
           *sE = sDZ1 + 64 * SG_LDS_STRIDE;
    const f32x4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    f32x4 gW2[2][4], gW1[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++) {
#pragma unroll
        for (int b = 0; b < 4; b++) gW2[a][b] = zero;
        gW1[a][0] = zero;
        gW1[a][1] = zero;
    }
    float gb1[16], gb2[16], gw3[16], gb3 = 0.0f, lsum = 0.0f;
#pragma unroll
    for (int q = 0; q < 16; q++) gb1[q] = gb2[q] = gw3[q] = 0.0f;
    Lvl lv[4];
    lane_levels(A.enc, lv);

    const long long ngroups = ((A.nchunk + 63) / 64 + 1) / 2;
    for (long long grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
        const Mlp mlp = {opaque(A.mlp.frag), opaque(A.mlp.p)};
        const long long tile = (grp * 2 + wave) * 64; /* first sample of the wavefront, in the chunk */
        const bool valid = tile + lane < A.nchunk;
        float x[3], t = 0.0f;
        int b = 0;
        train_sample(A, A.s0 + (valid ? tile + lane : 0), x, b, &t);
        bf16x8 E[4];
        encode_cols(A.enc, lv, x, b, E);
        /* encodings as ET[feature][sample] for the first layer's weight gradient,
         * staged now so that E dies with the first layer (the previous tile's
         * readers of this buffer finished before its last barrier) */
#pragma unroll
        for (int nt = 0; nt < 4; nt++)
#pragma unroll
            for (int j = 0; j < 8; j++) sE[(g * 8 + j) * SG_LDS_STRIDE + nt * 16 + (lane & 15)] = E[nt][j];
        bf16x8 H[2][4];
        float yv[4];
        mlp_forward_h(mlp, E, H, yv);
        /* loss gradient for the samples nt*16+l%16 */
        float dy[4];
#pragma unroll
        for (int nt = 0; nt < 4; nt++) {
            const int src = nt * 16 + (lane & 15);
            const float tn = __shfl(t, src, 64);
            const bool vn = __shfl((int)valid, src, 64) != 0;
            const float r = yv[nt] - tn;
            dy[nt] = vn ? 2.0f * r * A.inv_n : 0.0f;
            if (g == 0 && vn) lsum = fmaf(r, r, lsum);
            if (g == 0) gb3 += dy[nt];
        }
        if (A.ys && valid) A.ys[A.s0 + tile + lane] = own(yv);
        /* output layer and second hidden layer: dZ2 = relu'(Z2) * W3 dy, with each
         * Z2 tile recomputed as the forward pass made it */
        bf16x8 DZ[2][4];
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const bf16x8 a0 = mlp.frag[(SG_FRAG_F2 + mt * 2 + 0) * 64 + lane];
            const bf16x8 a1 = mlp.frag[(SG_FRAG_F2 + mt * 2 + 1) * 64 + lane];
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                f32x4 c = mfma(a0, H[0][nt], zero);
                c = mfma(a1, H[1][nt], c);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int o = mt * 16 + g * 4 + i;
                    const float z = c[i] + mlp.p[SG_P_B2 + o];
                    gw3[mt * 4 + i] = fmaf(dy[nt], z > 0.0f ? z : 0.0f, gw3[mt * 4 + i]);
                    const float dz = z > 0.0f ? mlp.p[SG_P_W3 + o] * dy[nt] : 0.0f;
                    const __bf16 dzb = (__bf16)dz;
                    gb2[mt * 4 + i] += (float)dzb;
                    DZ[mt >> 1][nt][(mt & 1) * 4 + i] = dzb;
                    sDZ2[o * SG_LDS_STRIDE + nt * 16 + (lane & 15)] = dzb;
                    sH1[o * SG_LDS_STRIDE + nt * 16 + (lane & 15)] = H[mt >> 1][nt][(mt & 1) * 4 + i];
                }
            }
        }
        /* dH1T = W2^T dZ2T, masked by the first layer's ReLU */
        bf16x8 D1[2][4];
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            const bf16x8 a0 = mlp.frag[(SG_FRAG_G2 + mt * 2 + 0) * 64 + lane];
            const bf16x8 a1 = mlp.frag[(SG_FRAG_G2 + mt * 2 + 1) * 64 + lane];
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                f32x4 c = mfma(a0, DZ[0][nt], zero);
                c = mfma(a1, DZ[1][nt], c);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int o = mt * 16 + g * 4 + i;
                    const float h1 = (float)H[mt >> 1][nt][(mt & 1) * 4 + i];
                    const __bf16 d = (__bf16)(h1 > 0.0f ? c[i] : 0.0f);
                    gb1[mt * 4 + i] += (float)d;
                    D1[mt >> 1][nt][(mt & 1) * 4 + i] = d;
                    sDZ1[o * SG_LDS_STRIDE + nt * 16 + (lane & 15)] = d;
                }
            }
        }
        /* dET = W1^T dZ1T: the encoding gradient of samples nt*16+l%16, features
         * mt*16+4g..+3, to the chunk buffer for the table scatter */
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            const bf16x8 a0 = mlp.frag[(SG_FRAG_G1 + mt * 2 + 0) * 64 + lane];
            const bf16x8 a1 = mlp.frag[(SG_FRAG_G1 + mt * 2 + 1) * 64 + lane];
#pragma unroll
            for (int nt = 0; nt < 4; nt++) {
                f32x4 c = mfma(a0, D1[0][nt], zero);
                c = mfma(a1, D1[1][nt], c);
                const long long sn = tile + nt * 16 + (lane & 15);
                if (sn < A.nchunk) *(f32x4 *)&A.dE[sn * SG_IN + mt * 16 + g * 4] = c;
            }
        }
        __syncthreads();
        /* weight gradients, products over the 64 samples of each wavefront's tile */
#pragma unroll
        for (int src = 0; src < 2; src++) {
            const __bf16 *tDZ2 = lds[src], *tH1 = tDZ2 + 64 * SG_LDS_STRIDE, *tDZ1 = tH1 + 64 * SG_LDS_STRIDE,
                         *tE = tDZ1 + 64 * SG_LDS_STRIDE;
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                const int kc = ks * 32 + g * 8;
#pragma unroll
                for (int mm = 0; mm < 2; mm++) {
                    const int mo = 2 * wave + mm;
                    const bf16x8 a2 = *(const bf16x8 *)&tDZ2[(mo * 16 + (lane & 15)) * SG_LDS_STRIDE + kc];
                    const bf16x8 a1 = *(const bf16x8 *)&tDZ1[(mo * 16 + (lane & 15)) * SG_LDS_STRIDE + kc];
#pragma unroll
                    for (int ni = 0; ni < 4; ni++) {
                        const bf16x8 bh = *(const bf16x8 *)&tH1[(ni * 16 + (lane & 15)) * SG_LDS_STRIDE + kc];
                        gW2[mm][ni] = mfma(a2, bh, gW2[mm][ni]);
                    }
#pragma unroll
                    for (int ni = 0; ni < 2; ni++) {
                        const bf16x8 be = *(const bf16x8 *)&tE[(ni * 16 + (lane & 15)) * SG_LDS_STRIDE + kc];
                        gW1[mm][ni] = mfma(a1, be, gW1[mm][ni]);
                    }
                }
            }
        }
        __syncthreads();
    }
    /* flush: weight tiles are unique per lane; per-output sums reduce over l%16 */
    float *const gp = opaque(A.gparam);
#pragma unroll
    for (int mm = 0; mm < 2; mm++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int o = (2 * wave + mm) * 16 + g * 4 + i;
#pragma unroll
            for (int ni = 0; ni < 4; ni++)
                unsafeAtomicAdd(&gp[SG_P_W2 + o * SG_H + ni * 16 + (lane & 15)], gW2[mm][ni][i]);
#pragma unroll
            for (int ni = 0; ni < 2; ni++)
                unsafeAtomicAdd(&gp[SG_P_W1 + o * SG_IN + ni * 16 + (lane & 15)], gW1[mm][ni][i]);
        }
#pragma unroll
    for (int q = 0; q < 16; q++) {
        float v1 = gb1[q], v2 = gb2[q], v3 = gw3[q];
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            v1 += __shfl_xor(v1, off, 64);
            v2 += __shfl_xor(v2, off, 64);
            v3 += __shfl_xor(v3, off, 64);
        }
        if ((lane & 15) == 0) {
            const int o = (q >> 2) * 16 + g * 4 + (q & 3);
            unsafeAtomicAdd(&gp[SG_P_B1 + o], v1);
            unsafeAtomicAdd(&gp[SG_P_B2 + o], v2);
            unsafeAtomicAdd(&gp[SG_P_W3 + o], v3);
        }
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        gb3 += __shfl_xor(gb3, off, 64);
        lsum += __shfl_xor(lsum, off, 64);
    }
    if (lane == 0) {
        unsafeAtomicAdd(&gp[SG_P_B3], gb3);
        unsafeAtomicAdd(A.loss, lsum);
    }
}

/* The hash-table gradient of one chunk: d(loss)/d(encoding) of each sample
 * (k_train_net's A.dE) added to the 8 corners of its cell at every level.
 * Each wavefront takes 64 samples, staged in LDS with their positions; 16
 * lanes work on one (sample, level), one lane per corner and feature, so
 * the atomics of one instruction fall on the few cache lines of four cells
 * instead of on 64 unrelated lines.  Levels whose vertices under the
 * wavefront's samples form a box of at most SG_POOL vertices in all (the
 * coarse levels when the samples lie close together, as drawn samples do)
 * are summed in LDS first and added to the table once per touched vertex;
 * the other levels go straight to the table with hardware fp32 atomics. */
#define SG_SC_WAVES 4
#define SG_POOL 1024 /* vertices (2 features each) of a wavefront's LDS sums */

/* the box of level l's vertices under samples with local positions in
 * [xmin, xmax]: first vertex lo, edge d per axis (level_cell's clamping) */
__device__ __forceinline__ void level_box(const Lvl &L, const float xmin[3], const float xmax[3], int lo[3], int d[3])
{
    const float N = (float)L.res;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const int c0 = min(max((int)floorf(xmin[a] * N), 0), L.res - 1);
        const int c1 = min(max((int)floorf(xmax[a] * N), 0), L.res - 1);
        lo[a] = c0;
        d[a] = c1 - c0 + 2;
    }
}

__global__ __launch_bounds__(64 * SG_SC_WAVES) void k_train_scatter(const TrainArgs A)
{
    __shared__ float s_g[SG_SC_WAVES][64 * SG_IN];  /* [sample][feature] */
    __shared__ float4 s_x[SG_SC_WAVES][64];         /* x, y, z, brick (-1: past the chunk) */
    __shared__ __attribute__((aligned(16))) float s_pool[SG_SC_WAVES][2 * SG_POOL]; /* LDS sums of the pooled levels */
    __shared__ int s_plan[SG_SC_WAVES][SG_L][4];    /* per level: first pool vertex (-1: not pooled), lo */
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4;
    const int corner = lane & 7, feat = (lane >> 3) & 1;
    const int dx = corner & 1, dy = (corner >> 1) & 1, dz = corner >> 2;
    float *pool = s_pool[wave];
    for (int k = lane; k < 2 * SG_POOL; k += 64) pool[k] = 0.0f;
    /* lane group g works on level 4q + g in pass q */
    Lvl lv[4];
#pragma unroll
    for (int q = 0; q < 4; q++) lv[q] = lvl_of(A.enc, g, 4 * q, 1);
    /* every wavefront of a workgroup runs the same number of tiles (the loop
     * has barriers): tiles past the chunk have no samples */
    const long long ntiles = ((A.nchunk + 63) / 64 + SG_SC_WAVES - 1) / SG_SC_WAVES * SG_SC_WAVES;
    for (long long tl = (long long)blockIdx.x * SG_SC_WAVES + wave; tl < ntiles; tl += (long long)gridDim.x * SG_SC_WAVES) {
        const long long tile = tl * 64;
        const bool valid = tile + lane < A.nchunk;
        float x[3];
        int b = 0;
        train_sample(A, A.s0 + (valid ? tile + lane : 0), x, b, nullptr);
        s_x[wave][lane] = make_float4(x[0], x[1], x[2], __int_as_float(valid ? b : -1));
        const f32x4 *src = (const f32x4 *)(A.dE + tile * SG_IN);
        f32x4 *dst = (f32x4 *)s_g[wave];
#pragma unroll
        for (int k = 0; k < 64 * SG_IN / 4 / 64; k++) {
            const long long e = (long long)k * 64 + lane; /* f32x4 index in the tile */
            if (tile + e / (SG_IN / 4) < A.nchunk) dst[e] = src[e];
        }
        /* the wavefront's position box and brick; its levels pooled in LDS in
         * order while their vertex boxes fit */
        float xmin[3], xmax[3]; /* over every lane: a lane past the chunk only widens the box */
        int bmin = valid ? b : INT_MAX, bmax = valid ? b : -1;
#pragma unroll
        for (int a = 0; a < 3; a++) xmin[a] = xmax[a] = x[a];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
            for (int a = 0; a < 3; a++) {
                xmin[a] = fminf(xmin[a], __shfl_xor(xmin[a], off, 64));
                xmax[a] = fmaxf(xmax[a], __shfl_xor(xmax[a], off, 64));
            }
            bmin = min(bmin, __shfl_xor(bmin, off, 64));
            bmax = max(bmax, __shfl_xor(bmax, off, 64));
        }
        if (lane < SG_L) s_plan[wave][lane][0] = -1;
        if (lane == 0 && bmin == bmax) {
            int used = 0;
            for (int l = 0; l < SG_L; l++) {
                const Lvl L = {l, A.enc.res[l], A.enc.n1[l]};
                int lo[3], d[3];
                level_box(L, xmin, xmax, lo, d);
                const int vol = d[0] * d[1] * d[2];
                if (used + vol > SG_POOL) break;
                s_plan[wave][l][0] = used;
                s_plan[wave][l][1] = lo[0] | d[0] << 16;
                s_plan[wave][l][2] = lo[1] | d[1] << 16;
                s_plan[wave][l][3] = lo[2] | d[2] << 16;
                used += vol;
            }
        }
        __syncthreads();
        for (int s = 0; s < 64; s++) {
            const float4 xs = s_x[wave][s];
            const int bb = __float_as_int(xs.w);
            if (bb < 0) continue; /* uniform over the wavefront */
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const Lvl &L = lv[q];
                const float xp[3] = {xs.x, xs.y, xs.z};
                int v0[3];
                float w[3];
                level_cell(L, xp, v0, w);
                const float wc = (dx ? w[0] : 1.0f - w[0]) * (dy ? w[1] : 1.0f - w[1]) * (dz ? w[2] : 1.0f - w[2]);
                const float gr = s_g[wave][s * SG_IN + L.l * 2 + feat];
                const int base = s_plan[wave][L.l][0];
                if (base >= 0) {
                    const int px = s_plan[wave][L.l][1], py = s_plan[wave][L.l][2], pz = s_plan[wave][L.l][3];
                    const int e = base + (v0[0] + dx - (px & 0xffff)) +
                                  (px >> 16) * ((v0[1] + dy - (py & 0xffff)) + (py >> 16) * (v0[2] + dz - (pz & 0xffff)));
                    atomicAdd(&pool[2 * e + feat], wc * gr); /* LDS */
                } else {
                    const unsigned idx = corner_index(A.enc.T, L, v0[0] + dx, v0[1] + dy, v0[2] + dz);
                    float *tb = A.gtable + ((size_t)bb * SG_L + L.l) * (size_t)A.enc.T * SG_F;
                    unsafeAtomicAdd(tb + (size_t)idx * 2 + feat, wc * gr);
                }
            }
        }
        /* the pooled sums to the table, one atomic per touched vertex and
         * feature, and the pool cleared for the next tile */
        if (bmin == bmax && bmin >= 0) {
            float *tb0 = A.gtable + (size_t)bmin * SG_L * (size_t)A.enc.T * SG_F;
            for (int l = 0; l < SG_L; l++) {
                const int base = s_plan[wave][l][0];
                if (base < 0) break;
                const int px = s_plan[wave][l][1], py = s_plan[wave][l][2], pz = s_plan[wave][l][3];
                const int ex = px >> 16, ey = py >> 16, ez = pz >> 16;
                const Lvl L = {l, A.enc.res[l], A.enc.n1[l]};
                float *tb = tb0 + (size_t)l * A.enc.T * SG_F;
                for (int k = lane; k < ex * ey * ez; k += 64) {
                    const float2 v = *(const float2 *)&pool[2 * (base + k)];
                    if (v.x != 0.0f || v.y != 0.0f) {
                        const int vx = (px & 0xffff) + k % ex, vy = (py & 0xffff) + (k / ex) % ey,
                                  vz = (pz & 0xffff) + k / (ex * ey);
                        const unsigned idx = corner_index(A.enc.T, L, vx, vy, vz);
                        unsafeAtomicAdd(tb + (size_t)idx * 2, v.x);
                        unsafeAtomicAdd(tb + (size_t)idx * 2 + 1, v.y);
                        *(float2 *)&pool[2 * (base + k)] = make_float2(0.0f, 0.0f);
                    }
                }
            }
        }
        __syncthreads();
    }
}

/* bf16 A-operand fragments of the weights for all four products */
__global__ void k_frags(const float *p, bf16x8 *frag)
{
    const int lane = threadIdx.x & 63, set = blockIdx.x, g = lane >> 4, r = lane & 15;
    bf16x8 v;
    for (int j = 0; j < 8; j++) {
        float w;
        if (set < SG_FRAG_F2) { /* W1, natural K */
            const int mt = set - SG_FRAG_F1;
            w = p[SG_P_W1 + (mt * 16 + r) * SG_IN + g * 8 + j];
        } else if (set < SG_FRAG_G2) { /* W2, permuted K */
            const int mt = (set - SG_FRAG_F2) >> 1, ks = (set - SG_FRAG_F2) & 1;
            w = p[SG_P_W2 + (mt * 16 + r) * SG_H + kperm(ks, g, j)];
        } else if (set < SG_FRAG_G1) { /* W2^T, permuted K */
            const int mt = (set - SG_FRAG_G2) >> 1, ks = (set - SG_FRAG_G2) & 1;
            w = p[SG_P_W2 + kperm(ks, g, j) * SG_H + mt * 16 + r];
        } else { /* W1^T, permuted K */
            const int mt = (set - SG_FRAG_G1) >> 1, ks = (set - SG_FRAG_G1) & 1;
            w = p[SG_P_W1 + kperm(ks, g, j) * SG_IN + mt * 16 + r];
        }
        v[j] = (__bf16)w;
    }
    frag[set * 64 + lane] = v;
}

/* Adam with bias correction (Kingma & Ba 2015) */
__global__ void k_adam(float *p, const float *gr, float *m, float *v, size_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = gr[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

/* uniform init: tables U[-s, s], MLP weights U[-1/sqrt(fan_in), 1/sqrt(fan_in)], biases 0 */
__global__ void k_init(float *p, size_t n, float scale, unsigned seed)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned h = pcg_hash(seed ^ pcg_hash((unsigned)i) ^ pcg_hash((unsigned)(i >> 32) + 0x9e3779b9u));
    p[i] = ((float)(h >> 8) * (1.0f / 16777216.0f) * 2.0f - 1.0f) * scale;
}

/* fp32 feature pairs -> half2 (round to nearest even) */
__global__ __launch_bounds__(256) void k_half(const float2 *src, __half2 *dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = __float22half2_rn(src[i]);
}

/* the half at or below x (at or above with up): the nearest half, moved one
 * step outward when rounding went the other way (its bit pattern is a
 * sign-magnitude integer, so one step is +-1 on the pattern) */
__device__ __forceinline__ __half half_outward(float x, bool up)
{
    const __half h = __float2half_rn(x);
    const float back = __half2float(h);
    if (up ? !(back < x) : !(back > x)) return h;
    const unsigned short b = __half_as_ushort(h);
    const bool neg = (b & 0x8000u) != 0;
    unsigned short m;
    if (up) m = neg ? (b == 0x8000u ? (unsigned short)0x0001u : (unsigned short)(b - 1u)) : (unsigned short)(b + 1u);
    else m = neg ? (unsigned short)(b + 1u) : (b == 0u ? (unsigned short)0x8001u : (unsigned short)(b - 1u));
    return __ushort_as_half(m);
}

/* cell ranges (min, max) -> half pairs that still contain them */
__global__ __launch_bounds__(256) void k_range_half(const float2 *src, __half2 *dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = __halves2half2(half_outward(src[i].x, false), half_outward(src[i].y, true));
}

} // namespace

/* ================================================================= host API */

namespace {

Enc make_enc(const ivr_surrogate *s)
{
    Enc e;
    e.table16 = (const __half2 *)s->tables16;
    e.T = s->T;
    e.log2T = s->sp.log2_table;
    for (int l = 0; l < SG_L; l++) {
        e.res[l] = s->res[l];
        e.n1[l] = s->dense_n1[l];
    }
    return e;
}

void volume_args(RenderArgs &a, const ivr_volume *v)
{
    memset(&a, 0, sizeof a);
    for (int r = 0; r < 3; r++) {
        a.n[r] = v->n[r];
        a.nb[r] = v->nb[r];
        a.nm1f[r] = (float)(v->n[r] - 1);
    }
    a.bshift = v->bshift;
    a.bmask = v->B - 1;
    a.P = v->P;
    a.nbricks = v->nbricks;
    a.vox_data = v->data;
    a.boff = v->d_boff;
    a.dtype_f16 = v->dtype == IVR_F16;
}

/* after every change of the weights: the MLP's bf16 operand fragments and
 * the half-precision copy of the hash tables that rendering reads */
int refresh_frags(ivr_ctx *ctx, ivr_surrogate *s)
{
    hipLaunchKernelGGL(k_frags, dim3(SG_NFRAG), dim3(64), 0, ctx->stream, s->params, (bf16x8 *)s->frag);
    const size_t n2 = s->ntable / 2;
    hipLaunchKernelGGL(k_half, dim3((unsigned)std::min<size_t>((n2 + 255) / 256, 65536)), dim3(256), 0, ctx->stream,
                       (const float2 *)s->tables, (__half2 *)s->tables16, n2);
    IVR_HIP(hipGetLastError());
    return IVR_OK;
}

} // namespace

namespace {

/* the side stream and events of the training pipeline, made on first use */
int train_streams(ivr_surrogate *s)
{
    if (s->side) return IVR_OK;
    IVR_HIP(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    for (int q = 0; q < 2; q++) {
        IVR_HIP(hipEventCreateWithFlags(&s->ev_net[q], hipEventDisableTiming));
        IVR_HIP(hipEventCreateWithFlags(&s->ev_sc[q], hipEventDisableTiming));
    }
    return IVR_OK;
}

void train_streams_release(ivr_surrogate *s)
{
    for (int q = 0; q < 2; q++) {
        if (s->ev_net[q]) (void)hipEventDestroy(s->ev_net[q]);
        if (s->ev_sc[q]) (void)hipEventDestroy(s->ev_sc[q]);
        s->ev_net[q] = s->ev_sc[q] = nullptr;
    }
    if (s->side) (void)hipStreamDestroy(s->side);
    s->side = nullptr;
}

} // namespace

bool ivr_surrogate_matches(const ivr_surrogate *s, const ivr_volume *v)
{
    return v->n[0] == s->n[0] && v->n[1] == s->n[1] && v->n[2] == s->n[2] &&
           v->bshift + s->sp.log2_group == s->bshift;
}

int ivr_surrogate_create(ivr_ctx *ctx, const ivr_volume *vol, const ivr_surrogate_params *sp, ivr_surrogate **out)
{
    if (!ctx || !vol || !sp || !out) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (sp->log2_table < 10 || sp->log2_table > 24) return ivr_fail(IVR_ERR_ARG, "log2_table %d", sp->log2_table);
    if (sp->log2_group < 0 || sp->log2_group > 4) return ivr_fail(IVR_ERR_ARG, "log2_group %d", sp->log2_group);
    ivr_surrogate *s = new ivr_surrogate();
    s->sp = *sp;
    s->T = 1u << sp->log2_table;
    s->bshift = vol->bshift + sp->log2_group;
    s->B = 1 << s->bshift;
    s->nbricks = 1;
    for (int r = 0; r < 3; r++) {
        s->n[r] = vol->n[r];
        s->nb[r] = ((vol->n[r] - 2) >> s->bshift) + 1; /* bricks owning floor indices 0 .. n-2 */
        s->nbricks *= s->nb[r];
    }
    for (int l = 0; l < SG_L; l++) {
        s->res[l] = (int)std::floor(sp->base_res * std::pow((double)sp->growth, l));
        const unsigned long long n1 = (unsigned long long)s->res[l] + 1;
        s->dense_n1[l] = n1 * n1 * n1 <= s->T ? (unsigned)n1 : 0u;
    }
    s->ntable = (size_t)s->nbricks * SG_L * s->T * SG_F;
    s->ncells = 1;
    for (int r = 0; r < 3; r++) {
        s->nc[r] = ((vol->n[r] - 2) >> SG_CELL_SHIFT) + 1;
        s->ncells *= (size_t)s->nc[r];
    }
    int rc;
    if ((rc = ivr_dalloc(ctx, (void **)&s->tables, s->ntable * 4)) ||
        (rc = ivr_dalloc(ctx, (void **)&s->tables16, s->ntable * 2 + SG_T16_TAIL)) ||
        (rc = ivr_dalloc(ctx, (void **)&s->gtables, s->ntable * 4)) ||
        (rc = ivr_dalloc(ctx, (void **)&s->mt, s->ntable * 4)) ||
        (rc = ivr_dalloc(ctx, (void **)&s->vt, s->ntable * 4)) ||
        (rc = ivr_dalloc(ctx, (void **)&s->params, SG_NPARAM * 4)) ||
        (rc = ivr_dalloc(ctx, (void **)&s->gparams, SG_NPARAM * 4)) ||
        (rc = ivr_dalloc(ctx, (void **)&s->mp, SG_NPARAM * 4)) ||
        (rc = ivr_dalloc(ctx, (void **)&s->vp, SG_NPARAM * 4)) ||
        (rc = ivr_dalloc(ctx, (void **)&s->frag, SG_NFRAG * 64 * 16)) ||
        (rc = ivr_dalloc(ctx, (void **)&s->d_loss, 4)) ||
        (rc = ivr_dalloc(ctx, (void **)&s->d_fallback, IVR_MAX_BRICKS)) ||
        (rc = ivr_dalloc(ctx, (void **)&s->d_cellrange, sizeof(__half2) * s->ncells)) ||
        (rc = ivr_dalloc(ctx, (void **)&s->d_cellempty, s->ncells))) {
        ivr_surrogate_destroy(ctx, s);
        return rc;
    }
    {
        ivr_dscope range32(ctx);
        if ((rc = range32.alloc(sizeof(float2) * s->ncells)) ||
            (rc = ivr_volume_cellrange(ctx, vol, SG_CELL_SHIFT, s->nc, (float2 *)range32.p))) {
            ivr_surrogate_destroy(ctx, s);
            return rc;
        }
        hipLaunchKernelGGL(k_range_half, dim3((unsigned)std::min<size_t>((s->ncells + 255) / 256, 4096)), dim3(256),
                           0, ctx->stream, (const float2 *)range32.p, s->d_cellrange, s->ncells);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
            ivr_surrogate_destroy(ctx, s);
            return ivr_fail(IVR_ERR_HIP, "macrocell ranges of the surrogate failed");
        }
    }
    const unsigned seed = (unsigned)(sp->seed ^ (sp->seed >> 32));
    hipLaunchKernelGGL(k_init, dim3((unsigned)((s->ntable + 255) / 256)), dim3(256), 0, ctx->stream, s->tables,
                       s->ntable, sp->init_scale, seed);
    /* MLP: W1 fan-in 32, W2 fan-in 64, W3 fan-in 64; biases zero */
    hipLaunchKernelGGL(k_init, dim3((SG_H * SG_IN + 255) / 256), dim3(256), 0, ctx->stream, s->params + SG_P_W1,
                       (size_t)SG_H * SG_IN, 1.0f / std::sqrt((float)SG_IN), seed + 1u);
    hipLaunchKernelGGL(k_init, dim3((SG_H * SG_H + 255) / 256), dim3(256), 0, ctx->stream, s->params + SG_P_W2,
                       (size_t)SG_H * SG_H, 1.0f / std::sqrt((float)SG_H), seed + 2u);
    hipLaunchKernelGGL(k_init, dim3(1), dim3(256), 0, ctx->stream, s->params + SG_P_W3, (size_t)SG_H,
                       1.0f / std::sqrt((float)SG_H), seed + 3u);
    IVR_HIP(hipGetLastError());
    IVR_HIP(hipMemsetAsync(s->params + SG_P_B1, 0, SG_H * 4, ctx->stream));
    IVR_HIP(hipMemsetAsync(s->params + SG_P_B2, 0, SG_H * 4, ctx->stream));
    IVR_HIP(hipMemsetAsync(s->params + SG_P_B3, 0, 4, ctx->stream));
    IVR_HIP(hipMemsetAsync(s->mt, 0, s->ntable * 4, ctx->stream));
    IVR_HIP(hipMemsetAsync(s->vt, 0, s->ntable * 4, ctx->stream));
    IVR_HIP(hipMemsetAsync(s->mp, 0, SG_NPARAM * 4, ctx->stream));
    IVR_HIP(hipMemsetAsync(s->vp, 0, SG_NPARAM * 4, ctx->stream));
    IVR_HIP(hipMemsetAsync(s->d_fallback, 0, IVR_MAX_BRICKS, ctx->stream));
    if ((rc = refresh_frags(ctx, s))) return rc;
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    *out = s;
    return IVR_OK;
}

int ivr_surrogate_destroy(ivr_ctx *ctx, ivr_surrogate *s)
{
    if (!s) return IVR_OK;
    if (ctx) {
        (void)hipStreamSynchronize(ctx->stream);
        ivr_dfree(ctx, s->tables, s->ntable * 4);
        ivr_dfree(ctx, s->tables16, s->ntable * 2 + SG_T16_TAIL);
        ivr_dfree(ctx, s->gtables, s->ntable * 4);
        ivr_dfree(ctx, s->mt, s->ntable * 4);
        ivr_dfree(ctx, s->vt, s->ntable * 4);
        ivr_dfree(ctx, s->params, SG_NPARAM * 4);
        ivr_dfree(ctx, s->gparams, SG_NPARAM * 4);
        ivr_dfree(ctx, s->mp, SG_NPARAM * 4);
        ivr_dfree(ctx, s->vp, SG_NPARAM * 4);
        ivr_dfree(ctx, s->frag, SG_NFRAG * 64 * 16);
        ivr_dfree(ctx, s->d_loss, 4);
        ivr_dfree(ctx, s->d_fallback, IVR_MAX_BRICKS);
        ivr_dfree(ctx, s->d_cellrange, sizeof(__half2) * s->ncells);
        ivr_dfree(ctx, s->d_cellempty, s->ncells);
        ivr_dfree(ctx, s->dE, 2 * s->dE_n * SG_IN * 4);
        train_streams_release(s);
    }
    delete s;
    return IVR_OK;
}

namespace {

/* samples per k_train_net / k_train_scatter launch pair: the two chunk
 * buffers of encoding gradients are 2 x SG_CHUNK x 128 bytes (1 GiB),
 * allocated at the first training call (a step of 2^24 samples: 2^22 per
 * chunk 41.1 ms, 2^21 42.1 ms, 2^20 44.9 ms) */
#define SG_CHUNK (1ll << 22)

int launch_train(ivr_ctx *ctx, ivr_surrogate *s, const ivr_volume *vol, long long n, const float *xs, const int *bs,
                 const float *ts, float *ys)
{
    const long long chunk = std::min<long long>(SG_CHUNK, (n + 63) / 64 * 64);
    int rc;
    if ((rc = train_streams(s))) return rc;
    if ((size_t)chunk > s->dE_n) {
        IVR_HIP(hipStreamSynchronize(s->side)); /* a previous step's scatter may still read dE */
        ivr_dfree(ctx, s->dE, 2 * s->dE_n * SG_IN * 4);
        s->dE = nullptr;
        s->dE_n = 0;
        if ((rc = ivr_dalloc(ctx, (void **)&s->dE, 2 * (size_t)chunk * SG_IN * 4))) return rc;
        s->dE_n = (size_t)chunk;
    }
    TrainArgs A;
    memset(&A, 0, sizeof A);
    if (vol) volume_args(A.va, vol);
    for (int r = 0; r < 3; r++) A.snb[r] = s->nb[r];
    A.sbshift = s->bshift;
    A.snbricks = s->nbricks;
    A.enc = make_enc(s);
    A.mlp.frag = (const bf16x8 *)s->frag;
    A.mlp.p = s->params;
    A.gtable = s->gtables;
    A.gparam = s->gparams;
    A.loss = s->d_loss;
    A.nsamples = n;
    A.dE = s->dE;
    A.inv_n = 1.0f / (float)n;
    A.seed = (unsigned)(s->sp.seed * 0x9e3779b1ull);
    A.step = (unsigned)s->step;
    A.xs = xs;
    A.bs = bs;
    A.ts = ts;
    A.ys = ys;
    IVR_HIP(hipMemsetAsync(s->gtables, 0, s->ntable * 4, ctx->stream));
    IVR_HIP(hipMemsetAsync(s->gparams, 0, SG_NPARAM * 4, ctx->stream));
    IVR_HIP(hipMemsetAsync(s->d_loss, 0, 4, ctx->stream));
    /* chunk i: its network pass on the context's stream into dE buffer i % 2
     * (after the scatter of chunk i - 2 has read that buffer), its scatter on
     * the side stream once the network pass is done; the context's stream
     * then waits for the last scatters */
    int i = 0;
    for (long long s0 = 0; s0 < n; s0 += chunk, i++) {
        const int q = i & 1;
        A.s0 = s0;
        A.nchunk = std::min(chunk, n - s0);
        A.dE = s->dE + (size_t)q * s->dE_n * SG_IN;
        if (i >= 2) IVR_HIP(hipStreamWaitEvent(ctx->stream, s->ev_sc[q], 0));
        const long long groups = ((A.nchunk + 63) / 64 + 1) / 2;
        hipLaunchKernelGGL(k_train_net, dim3((unsigned)std::min<long long>(groups, 2048)), dim3(128), 0, ctx->stream, A);
        IVR_HIP(hipGetLastError());
        IVR_HIP(hipEventRecord(s->ev_net[q], ctx->stream));
        IVR_HIP(hipStreamWaitEvent(s->side, s->ev_net[q], 0));
        const long long tiles = (A.nchunk + 63) / 64;
        hipLaunchKernelGGL(k_train_scatter, dim3((unsigned)std::min<long long>((tiles + SG_SC_WAVES - 1) / SG_SC_WAVES, 2048)),
                           dim3(64 * SG_SC_WAVES), 0, s->side, A);
        IVR_HIP(hipGetLastError());
        IVR_HIP(hipEventRecord(s->ev_sc[q], s->side));
    }
    for (int q = 0; q < std::min(i, 2); q++) IVR_HIP(hipStreamWaitEvent(ctx->stream, s->ev_sc[q], 0));
    return IVR_OK;
}

int adam_all(ivr_ctx *ctx, ivr_surrogate *s)
{
    s->step++;
    const ivr_surrogate_params &p = s->sp;
    const float c1 = (float)(1.0 - std::pow((double)p.beta1, s->step));
    const float c2 = (float)(1.0 - std::pow((double)p.beta2, s->step));
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((s->ntable + 255) / 256)), dim3(256), 0, ctx->stream, s->tables,
                       s->gtables, s->mt, s->vt, s->ntable, p.lr, p.beta1, p.beta2, p.eps, c1, c2);
    hipLaunchKernelGGL(k_adam, dim3((SG_NPARAM + 255) / 256), dim3(256), 0, ctx->stream, s->params, s->gparams,
                       s->mp, s->vp, (size_t)SG_NPARAM, p.lr_mlp, p.beta1, p.beta2, p.eps, c1, c2);
    IVR_HIP(hipGetLastError());
    return refresh_frags(ctx, s);
}

} // namespace

int ivr_surrogate_train(ivr_ctx *ctx, ivr_surrogate *s, const ivr_volume *vol, int steps, long long samples,
                        float *loss_out)
{
    if (!ctx || !s || !vol) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (s->frozen) return ivr_fail(IVR_ERR_UNSUPPORTED, "the surrogate is frozen (no training state)");
    if (!ivr_surrogate_matches(s, vol)) return ivr_fail(IVR_ERR_ARG, "volume does not match surrogate");
    if (samples < 64 || steps < 0) return ivr_fail(IVR_ERR_ARG, "samples %lld steps %d", samples, steps);
    int rc;
    for (int k = 0; k < steps; k++) {
        if ((rc = launch_train(ctx, s, vol, samples, nullptr, nullptr, nullptr, nullptr))) return rc;
        if ((rc = adam_all(ctx, s))) return rc;
    }
    if (loss_out) {
        float l = 0.0f;
        IVR_HIP(hipMemcpyAsync(&l, s->d_loss, 4, hipMemcpyDeviceToHost, ctx->stream));
        IVR_HIP(hipStreamSynchronize(ctx->stream));
        *loss_out = l / (float)samples;
    }
    return IVR_OK;
}

int ivr_surrogate_grad(ivr_ctx *ctx, ivr_surrogate *s, int n, const float *xyz, const int *bricks,
                       const float *targets, float *y_out, float *loss_out, float *gmlp_out,
                       float *gtable_out, int gtable_brick)
{
    if (!ctx || !s || n < 1 || !xyz || !bricks || !targets) return ivr_fail(IVR_ERR_ARG, "bad argument");
    if (s->frozen) return ivr_fail(IVR_ERR_UNSUPPORTED, "the surrogate is frozen (no training state)");
    for (int i = 0; i < n; i++)
        if (bricks[i] < 0 || bricks[i] >= s->nbricks) return ivr_fail(IVR_ERR_ARG, "brick %d", bricks[i]);
    const size_t bytes = (size_t)n * 4 * 6;
    ivr_dscope io(ctx);
    int rc = io.alloc(bytes);
    if (rc) return rc;
    float *d = (float *)io.p;
    float *dx = d, *dt = d + 3 * (size_t)n, *dy = d + 4 * (size_t)n;
    int *db = (int *)(d + 5 * (size_t)n);
    IVR_HIP(hipMemcpyAsync(dx, xyz, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
    IVR_HIP(hipMemcpyAsync(dt, targets, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    IVR_HIP(hipMemcpyAsync(db, bricks, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = launch_train(ctx, s, nullptr, n, dx, db, dt, dy))) return rc;
    if (y_out) IVR_HIP(hipMemcpyAsync(y_out, dy, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (gmlp_out) IVR_HIP(hipMemcpyAsync(gmlp_out, s->gparams, SG_NPARAM * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (gtable_out) {
        if (gtable_brick < 0 || gtable_brick >= s->nbricks) return ivr_fail(IVR_ERR_ARG, "gtable brick");
        const size_t per = (size_t)SG_L * s->T * SG_F;
        IVR_HIP(hipMemcpyAsync(gtable_out, s->gtables + per * gtable_brick, per * 4, hipMemcpyDeviceToHost,
                               ctx->stream));
    }
    float l = 0.0f;
    IVR_HIP(hipMemcpyAsync(&l, s->d_loss, 4, hipMemcpyDeviceToHost, ctx->stream));
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    if (loss_out) *loss_out = l / (float)n;
    return IVR_OK;
}

int ivr_surrogate_adam(ivr_ctx *ctx, ivr_surrogate *s)
{
    if (!ctx || !s) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (s->frozen) return ivr_fail(IVR_ERR_UNSUPPORTED, "the surrogate is frozen (no training state)");
    int rc = adam_all(ctx, s);
    if (rc) return rc;
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    return IVR_OK;
}

int ivr_surrogate_get(ivr_ctx *ctx, const ivr_surrogate *s, float *mlp, float *table, int brick)
{
    if (!ctx || !s) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (s->frozen) return ivr_fail(IVR_ERR_UNSUPPORTED, "the surrogate is frozen (no training state)");
    if (mlp) IVR_HIP(hipMemcpyAsync(mlp, s->params, SG_NPARAM * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (table) {
        if (brick < 0 || brick >= s->nbricks) return ivr_fail(IVR_ERR_ARG, "brick %d", brick);
        const size_t per = (size_t)SG_L * s->T * SG_F;
        IVR_HIP(hipMemcpyAsync(table, s->tables + per * brick, per * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    return IVR_OK;
}

int ivr_surrogate_set(ivr_ctx *ctx, ivr_surrogate *s, const float *mlp, const float *table, int brick)
{
    if (!ctx || !s) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (s->frozen) return ivr_fail(IVR_ERR_UNSUPPORTED, "the surrogate is frozen (no training state)");
    if (mlp) IVR_HIP(hipMemcpyAsync(s->params, mlp, SG_NPARAM * 4, hipMemcpyHostToDevice, ctx->stream));
    if (table) {
        if (brick < 0 || brick >= s->nbricks) return ivr_fail(IVR_ERR_ARG, "brick %d", brick);
        const size_t per = (size_t)SG_L * s->T * SG_F;
        IVR_HIP(hipMemcpyAsync(s->tables + per * brick, table, per * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    int rc = refresh_frags(ctx, s);
    if (rc) return rc;
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    return IVR_OK;
}

int ivr_surrogate_info(const ivr_surrogate *s, int *nbricks, unsigned long long *table_entries, int *step,
                       int res[16])
{
    if (!s) return ivr_fail(IVR_ERR_ARG, "null surrogate");
    if (nbricks) *nbricks = s->nbricks;
    if (table_entries) *table_entries = s->T;
    if (step) *step = s->step;
    if (res)
        for (int l = 0; l < SG_L; l++) res[l] = s->res[l];
    return IVR_OK;
}

/* ================================================================= render */

namespace {

struct SgRender {
    RenderArgs a;
    Enc enc;
    Mlp mlp;
    const unsigned char *fallback; /* per surrogate brick: 1 = fetch the plain volume */
    int only_brick;                /* >= 0: composite only that surrogate brick's samples */
    int snb[3], sbshift;           /* surrogate bricks per axis, log2 of their edge */
    int skip;                      /* macrocell skipping (k_render_sg_items) */
    int helpers;                   /* lanes of ended rays evaluate live rays' next samples */
    int nblocks;                   /* pixel-block workgroups of the frame */
};

/* index of the r-th set bit (from 0) of m, which has more than r set bits */
__device__ __forceinline__ int nth_bit(unsigned long long m, int r)
{
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const unsigned long long low = m & ((1ull << w) - 1ull);
        const int c = __popcll(low);
        if (r >= c) {
            r -= c;
            m >>= w;
            pos += w;
        } else {
            m = low;
        }
    }
    return pos;
}

/* From sample index k of the ray (go, gd, tin, dt), the first index below
 * kmax whose macrocell may hold a visible sample (kmax if none), with its
 * position in p; the indices passed are added to skipped.  cc/e cache the
 * code of the last cell looked up (a property of the cell, whatever the
 * ray). */
template <bool SKIP>
__device__ __forceinline__ int sg_walk(const RenderArgs &a, const float *go, const float *gd, const float *rgd,
                                       float tin, float dt, float rdt, int k, int kmax, int &cc, int &e, Pos &p,
                                       unsigned &skipped)
{
    while (k < kmax) {
        p = cfg_pos(a, go, gd, tin, dt, k);
        if (!SKIP) break;
        const int cell = cell_of<SG_CELL_SHIFT>(a, p);
        if (cell != cc) {
            cc = cell;
            e = a.cellempty[cell];
        }
        if (!e) break;
        const int kc = skip_empty<SG_CELL_SHIFT>(a, go, gd, rgd, tin, dt, rdt, p, e, k, kmax);
        skipped += (unsigned)(kc - k);
        k = kc;
    }
    return k;
}

/* Config-mode frame with the surrogate in place of the trilinear fetch.
 * The march is wave-synchronous: every lane evaluates one sample per
 * iteration (lanes whose ray ended feed a dummy sample) so the MLP's MFMAs
 * see all 64 lanes; the loop ends when a ballot finds no live ray.  The 64
 * lanes are neighbouring pixels at the same step, so their hash-grid
 * gathers share cache lines at the coarser levels.  (Refilling finished
 * lanes with new pixels keeps all columns live but loses that coherence:
 * measured 2.1x slower on C4.)  With SKIP a lane first passes the samples
 * of transparent macrocells (the same exact jumps as the plain renderer),
 * so only samples that can contribute reach the network.
 *
 * Helpers: at each step the r-th lane whose ray has ended takes the r-th
 * live ray, walks it from where its owner starts to the sample after the
 * owner's (the second visible one of the step) and evaluates that sample in
 * the same network pass; the owner composites its own sample, then, unless
 * the ray terminated, its helper's, with the indices between counted as
 * passed.  Each sample's value depends on its position only, so the frame
 * and its counts are those of one sample per ray and step, in fewer steps. */
template <bool SKIP>
__device__ __forceinline__ void sg_frame_block(const SgRender &S, int blk, float4 *smem)
{
    const RenderArgs &a = S.a;
    float4 *s_lut = smem;
    unsigned long long *s_boff = (unsigned long long *)(smem + a.lutsize);
    int i, j;
    if (!block_pixel(a, blk, i, j)) return; /* uniform per workgroup */
    const MlpLds mlp = stage_mlp(S.mlp, smem + a.lutsize + (a.nbricks + 1) / 2);
    stage(a, s_lut, s_boff); /* its barrier covers the MLP staging too */
    const bool inside = (i < a.H) && (j < a.W);
    float d[3] = {0.0f, 0.0f, 1.0f};
    float tin = 0.0f, dt = 1.0f;
    int n = 0;
    if (inside) {
        cfg_dir(a, i, j, d);
        float tout;
        if (slab(a.dconst, a.o, d, tin, tout)) {
            tin = fmaxf(tin, 0.0f);
            const float dlen = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            dt = (a.vox * a.step_vox) / dlen;
            const float nf = 1.0f + (tout - tin) / dt + 0.5f;
            n = nf < (float)IVR_MAX_STEPS ? (int)nf : IVR_MAX_STEPS;
        }
    }
    Acc acc = {0.0f, 0.0f, 0.0f, 0.0f};
    unsigned cnt = 0, fetched = 0, evals = 0;
    bool live = n > 0;
    int k = 0;
    const float invB = 1.0f / (float)(1 << S.sbshift);
    float go[3], gd[3], rgd[3], rdt = 0.0f;
    ray_voxel(a, a.o, d, go, gd);
    if (SKIP) {
#pragma unroll
        for (int q = 0; q < 3; q++) rgd[q] = 1.0f / gd[q];
        rdt = 1.0f / dt;
    }
    int cc = -1, e = 0; /* macrocell of the transparency code held in e */
    Lvl lv[4];
    lane_levels(S.enc, lv);
    const int lane = threadIdx.x & 63;
    /* the fallback flags as a mask when the surrogate has at most 64 bricks,
     * so the march does not wait on a load for them at every step */
    const int nsb = S.snb[0] * S.snb[1] * S.snb[2];
    const bool fbmask = nsb <= 64;
    const unsigned long long fbm = fbmask ? __ballot(lane < nsb && S.fallback[lane] != 0) : 0ull;
    auto falls_back = [&](int bb) { return fbmask ? ((fbm >> bb) & 1ull) != 0ull : S.fallback[bb] != 0; };
    while (__ballot(live) != 0ull) {
        /* pair the r-th ended lane with the r-th live one (every lane of the
         * wavefront runs this loop, so the ballots see all 64) */
        const unsigned long long lm = __ballot(live);
        bool helper = false;
        int mate = lane; /* a helper's owner, an owner's helper, else the lane itself */
        if (S.helpers) {
            const unsigned long long dm = ~lm, below = (1ull << lane) - 1ull;
            const int m = min(__popcll(lm), __popcll(dm));
            if (live) {
                const int r = __popcll(lm & below);
                if (r < m) mate = nth_bit(dm, r);
            } else {
                const int r = __popcll(dm & below);
                if (r < m) {
                    mate = nth_bit(lm, r);
                    helper = true;
                }
            }
        }
        /* the ray a lane walks: its own, or its owner's from where the owner starts */
        const int src = helper ? mate : lane;
        float wgo[3], wgd[3], wrgd[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int q = 0; q < 3; q++) {
            wgo[q] = __shfl(go[q], src, 64);
            wgd[q] = __shfl(gd[q], src, 64);
            if (SKIP) wrgd[q] = __shfl(rgd[q], src, 64);
        }
        const float wtin = __shfl(tin, src, 64), wdt = __shfl(dt, src, 64), wrdt = SKIP ? __shfl(rdt, src, 64) : 0.0f;
        const int wn = __shfl(n, src, 64), wk = __shfl(k, src, 64);
        Pos p;
        int ks = wn;
        if (live || helper) {
            unsigned skipped = 0, passed = 0;
            ks = sg_walk<SKIP>(a, wgo, wgd, wrgd, wtin, wdt, wrdt, wk, wn, cc, e, p, skipped);
            if (helper) {
                if (ks < wn) ks = sg_walk<SKIP>(a, wgo, wgd, wrgd, wtin, wdt, wrdt, ks + 1, wn, cc, e, p, passed);
            } else {
                cnt += skipped;
                k = ks;
                if (k >= n) live = false;
            }
        }
        const bool sampled = live || (helper && ks < wn); /* the lane has a sample at p */
        float x[3] = {0.5f, 0.5f, 0.5f};
        int b = 0;
        if (sampled) {
            const int sh = S.sbshift;
            const int bx = p.j >> sh, by = p.i >> sh, bz = p.k >> sh;
            b = bx + S.snb[0] * (by + S.snb[1] * bz);
            x[0] = ((float)p.j + p.fx - (float)(bx << sh)) * invB;
            x[1] = ((float)p.i + p.fy - (float)(by << sh)) * invB;
            x[2] = ((float)p.k + p.fz - (float)(bz << sh)) * invB;
        }
        /* The network runs only when a lane of the wavefront takes its value:
         * wavefronts whose samples all lie in fallback bricks (or outside
         * only_brick) skip the encoding and the MLP.  Lanes that discard the value
         * evaluate the centre of their brick, which keeps their gathers off the
         * cache lines of samples that are not needed.  A lane's value depends on
         * its own sample only, so this changes no result. */
        const bool net = sampled && !falls_back(b) && (S.only_brick < 0 || S.only_brick == b);
        if (!net) x[0] = x[1] = x[2] = 0.5f;
        float yv = 0.0f;
        if (__ballot(net) != 0ull) {
            yv = sg_infer(S.enc, lv, mlp, x, b);
            evals++;
        }
        /* an owner's second sample: its helper's index (n if none) and value */
        const int hl = live ? mate : lane;
        const float y2 = __shfl(yv, hl, 64);
        const int k2 = __shfl(ks, hl, 64);
        if (live) {
            const bool two = hl != lane;
#pragma unroll
            for (int u = 0; u < 2; u++) {
                if (u == 1) {
                    if (!live || !two) break;
                    if (k2 >= n) { /* the helper's walk passed every remaining index */
                        cnt += (unsigned)(n - k);
                        k = n;
                        live = false;
                        break;
                    }
                    cnt += (unsigned)(k2 - k);
                    k = k2;
                    p = cfg_pos(a, go, gd, tin, dt, k);
                    const int sh = S.sbshift;
                    b = (p.j >> sh) + S.snb[0] * ((p.i >> sh) + S.snb[1] * (p.k >> sh));
                    yv = y2;
                }
                if (S.only_brick < 0 || S.only_brick == b) {
                    float s = yv;
                    if (falls_back(b)) {
                        float v[8];
                        if (a.dtype_f16) fetch8<_Float16>(a, s_boff, p.j, p.i, p.k, v);
                        else fetch8<float>(a, s_boff, p.j, p.i, p.k, v);
                        s = trilinear(v, p.fx, p.fy, p.fz);
                    }
                    fetched++; /* classified: by the network or the fallback */
                    const float4 e = s_lut[lut_index(a, s)];
                    const float al = e.w * a.alpha;
                    const float w = (1.0f - acc.a) * al;
                    acc.r = fmaf(w, e.x, acc.r);
                    acc.g = fmaf(w, e.y, acc.g);
                    acc.b = fmaf(w, e.z, acc.b);
                    acc.a = acc.a + w;
                    cnt++;
                }
                k++;
                if (acc.a >= a.ert || k >= n) live = false;
            }
        }
    }
    if (inside) {
        const size_t px = (size_t)i * a.W + j;
        a.out8[px] = make_uchar4(quant8(acc.r), quant8(acc.g), quant8(acc.b), quant8(acc.a));
        if (a.out32) a.out32[px] = make_float4(acc.r, acc.g, acc.b, acc.a);
    }
    count_add(a.counters, 0, cnt);
    count_add(a.counters, 1, fetched);
    count_add(a.counters, 3, evals);
}

template <bool SKIP>
__global__ __launch_bounds__(IVR_FRAME_BLOCK) void k_render_sg(const SgRender S)
{
    extern __shared__ float4 smem[];
    sg_frame_block<SKIP>(S, blockIdx.x, smem);
}

/* the frames of several requests in one grid: workgroup g renders block
 * g - prefix[q] of item q, prefix[q] <= g < prefix[q+1]; requests of any
 * size then share the GPU instead of running one after another */
__global__ __launch_bounds__(IVR_FRAME_BLOCK) void k_render_sg_items(const SgRender *__restrict__ items,
                                                                     const int *__restrict__ prefix, int nitems)
{
    extern __shared__ float4 smem[];
    int q = 0;
    while (q + 1 < nitems && prefix[q + 1] <= (int)blockIdx.x) q++;
    const SgRender &S = items[q];
    const int blk = (int)blockIdx.x - prefix[q];
    if (S.skip) sg_frame_block<true>(S, blk, smem);
    else sg_frame_block<false>(S, blk, smem);
}

/* the renderer's network value (sg_infer) at caller points: one lane per
 * point, padding lanes evaluate a dummy point */
__global__ __launch_bounds__(256) void k_sg_points(const Enc enc, const Mlp mlp, int n, const float *__restrict__ xs,
                                                   const int *__restrict__ bs, float *__restrict__ ys)
{
    __shared__ __attribute__((aligned(16))) unsigned char s_mlp[SG_INFER_LDS];
    const MlpLds m = stage_mlp(mlp, s_mlp);
    __syncthreads();
    const int t = blockIdx.x * 256 + threadIdx.x;
    float x[3] = {0.5f, 0.5f, 0.5f};
    int b = 0;
    if (t < n) {
        x[0] = xs[3 * t];
        x[1] = xs[3 * t + 1];
        x[2] = xs[3 * t + 2];
        b = bs[t];
    }
    Lvl lv[4];
    lane_levels(enc, lv);
    const float y = sg_infer(enc, lv, m, x, b);
    if (t < n) ys[t] = y;
}

/* sum of squared RGBA8 differences */
__global__ void k_sqerr(const uchar4 *x, const uchar4 *y, size_t n, unsigned long long *out)
{
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned v = 0;
    if (p < n) {
        const uchar4 u = x[p], w = y[p];
        const int d0 = (int)u.x - w.x, d1 = (int)u.y - w.y, d2 = (int)u.z - w.z, d3 = (int)u.w - w.w;
        v = (unsigned)(d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(out, (unsigned long long)v);
}

/* the surrogate frame's kernels on stream st: macrocell codes of the frame's
 * LUT into `cells` (with skipping), then k_render_sg; fr's counters must be
 * cleared by the caller */
int sg_prepare(ivr_ctx *ctx, hipStream_t st, const ivr_surrogate *s, const ivr_volume *vol, const ivr_tf *tf,
               const ivr_frame_desc *d, ivr_frame *fr, unsigned char *cells, const unsigned char *fallback,
               int only_brick, SgRender &S, size_t &lds)
{
    if (d->mode != IVR_MODE_CONFIG) return ivr_fail(IVR_ERR_UNSUPPORTED, "surrogate renders config mode only");
    if (d->state.winwidth != fr->W || d->state.winheight != fr->H)
        return ivr_fail(IVR_ERR_ARG, "state %dx%d does not match frame %dx%d", d->state.winwidth,
                        d->state.winheight, fr->W, fr->H);
    if (vol && !ivr_surrogate_matches(s, vol)) return ivr_fail(IVR_ERR_ARG, "volume does not match surrogate");
    /* the geometry of the trained volume; a stand-in volume object carries it
     * when no plain data is resident */
    ivr_volume geo;
    const ivr_volume *v = vol;
    if (!v) {
        for (int r = 0; r < 3; r++) {
            geo.n[r] = s->n[r];
            geo.nb[r] = s->nb[r];
        }
        geo.B = s->B;
        geo.bshift = s->bshift;
        geo.P = s->B + 2;
        geo.nbricks = s->nbricks;
        geo.d_boff = nullptr;
        v = &geo;
    }
    ivr_cam cam;
    ivr_cam_setup(&d->state, s->n[0], s->n[1], s->n[2], &cam);
    ivr_fill_args(S.a, cam, v, tf, d, fr);
    if (!vol) S.a.nbricks = 0; /* no brick table to stage */
    S.enc = make_enc(s);
    S.mlp.frag = (const bf16x8 *)s->frag;
    S.mlp.p = s->params;
    S.fallback = fallback;
    S.only_brick = only_brick;
    for (int r = 0; r < 3; r++) S.snb[r] = s->nb[r];
    S.sbshift = s->bshift;
    for (int r = 0; r < 3; r++) S.a.nc[r] = s->nc[r];
    S.a.cellempty = cells;
    ivr_set_grid(S.a, ctx, fr->W, fr->H, IVR_GRID_SURROGATE);
    const bool skip = d->skip && s->ncells > 0;
    if (skip) {
        int rc = ivr_launch_cells_half(st, s->d_cellrange, s->nc, S.a, cells); /* the surrogate's own cells */
        if (rc) return rc;
    }
    /* LDS: the LUT, the brick table (rounded to 16 bytes), the inference weights */
    lds = (size_t)S.a.lutsize * sizeof(float4) + (size_t)((S.a.nbricks + 1) / 2) * 16 + SG_INFER_LDS;
    if (lds > 65536)
        return ivr_fail(IVR_ERR_UNSUPPORTED, "LUT of %d entries, %d bricks and the network exceed 64 KiB of LDS",
                        S.a.lutsize, S.a.nbricks);
    S.skip = skip ? 1 : 0;
    S.helpers = ctx->sg_helpers;
    S.nblocks = S.a.nblk_pad;
    return IVR_OK;
}

int sg_launch_on(ivr_ctx *ctx, hipStream_t st, const ivr_surrogate *s, const ivr_volume *vol,
                 const ivr_tf *tf, const ivr_frame_desc *d, ivr_frame *fr, unsigned char *cells,
                 const unsigned char *fallback, int only_brick)
{
    SgRender S;
    size_t lds = 0;
    int rc = sg_prepare(ctx, st, s, vol, tf, d, fr, cells, fallback, only_brick, S, lds);
    if (rc) return rc;
    if (S.skip) hipLaunchKernelGGL(k_render_sg<true>, dim3(S.nblocks), dim3(IVR_FRAME_BLOCK), lds, st, S);
    else hipLaunchKernelGGL(k_render_sg<false>, dim3(S.nblocks), dim3(IVR_FRAME_BLOCK), lds, st, S);
    IVR_HIP(hipGetLastError());
    return IVR_OK;
}

int sg_launch(ivr_ctx *ctx, const ivr_surrogate *s, const ivr_volume *vol, const ivr_tf *tf,
              const ivr_frame_desc *d, ivr_frame *fr, const unsigned char *fallback, int only_brick)
{
    IVR_HIP(hipMemsetAsync(fr->d_counters, 0, IVR_CTR_BYTES, ctx->stream));
    fr->timed = 1;
    fr->composited = 0;
    IVR_HIP(hipEventRecord(fr->e[0], ctx->stream));
    int rc = sg_launch_on(ctx, ctx->stream, s, vol, tf, d, fr, s->d_cellempty, fallback, only_brick);
    if (rc) return rc;
    IVR_HIP(hipEventRecord(fr->e[1], ctx->stream));
    return IVR_OK;
}

} // namespace

int ivr_surrogate_launch_on(ivr_ctx *ctx, hipStream_t st, const ivr_surrogate *s, const ivr_volume *vol,
                            const ivr_tf *tf, const ivr_frame_desc *d, ivr_frame *fr, unsigned char *cells,
                            const unsigned char *fallback)
{
    return sg_launch_on(ctx, st, s, vol, tf, d, fr, cells, fallback, -1);
}

size_t ivr_surrogate_item_bytes(void) { return sizeof(SgRender); }

int ivr_surrogate_prepare(ivr_ctx *ctx, hipStream_t st, const ivr_surrogate *s, const ivr_volume *vol,
                          const ivr_tf *tf, const ivr_frame_desc *d, ivr_frame *fr, unsigned char *cells,
                          const unsigned char *fallback, void *item, int *nblocks, size_t *lds)
{
    SgRender S;
    int rc = sg_prepare(ctx, st, s, vol, tf, d, fr, cells, fallback, -1, S, *lds);
    if (rc) return rc;
    memcpy(item, &S, sizeof S);
    *nblocks = S.nblocks;
    return IVR_OK;
}

int ivr_surrogate_launch_items(hipStream_t st, const void *d_items, const int *d_prefix, int nitems, int total,
                               size_t lds)
{
    if (nitems < 1 || total < 1) return IVR_OK;
    hipLaunchKernelGGL(k_render_sg_items, dim3(total), dim3(IVR_FRAME_BLOCK), lds, st, (const SgRender *)d_items,
                       d_prefix, nitems);
    IVR_HIP(hipGetLastError());
    return IVR_OK;
}

namespace {

} // namespace

void ivr_surrogate_defaults(ivr_surrogate_params *sp)
{
    memset(sp, 0, sizeof *sp);
    sp->log2_table = 19;
    sp->base_res = 16;
    sp->growth = 1.5f;
    sp->lr = 1e-2f;
    sp->lr_mlp = 1e-3f;
    sp->beta1 = 0.9f;
    sp->beta2 = 0.99f;
    sp->eps = 1e-15f;
    sp->init_scale = 1e-4f;
    sp->seed = 4004;
    sp->log2_group = 0;
}

int ivr_surrogate_render(ivr_ctx *ctx, const ivr_surrogate *s, const ivr_volume *vol, const ivr_tf *tf,
                         const ivr_frame_desc *d, ivr_frame *fr)
{
    if (!ctx || !s || !tf || !d || !fr) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (!vol) {
        std::vector<unsigned char> h(s->nbricks);
        IVR_HIP(hipMemcpyAsync(h.data(), s->d_fallback, s->nbricks, hipMemcpyDeviceToHost, ctx->stream));
        IVR_HIP(hipStreamSynchronize(ctx->stream));
        for (unsigned char c : h)
            if (c) return ivr_fail(IVR_ERR_ARG, "a brick falls back to the plain volume, which was not given");
    }
    return sg_launch(ctx, s, vol, tf, d, fr, s->d_fallback, -1);
}

int ivr_surrogate_eval_points(ivr_ctx *ctx, const ivr_surrogate *s, int n, const float *xyz, const int *bricks,
                              float *y_out)
{
    if (!ctx || !s || n < 1 || !xyz || !bricks || !y_out) return ivr_fail(IVR_ERR_ARG, "bad argument");
    for (int i = 0; i < n; i++)
        if (bricks[i] < 0 || bricks[i] >= s->nbricks) return ivr_fail(IVR_ERR_ARG, "brick %d", bricks[i]);
    ivr_dscope io(ctx);
    int rc = io.alloc((size_t)n * 20);
    if (rc) return rc;
    float *dx = (float *)io.p, *dy = dx + 3 * (size_t)n;
    int *db = (int *)(dy + n);
    IVR_HIP(hipMemcpyAsync(dx, xyz, (size_t)n * 12, hipMemcpyHostToDevice, ctx->stream));
    IVR_HIP(hipMemcpyAsync(db, bricks, (size_t)n * 4, hipMemcpyHostToDevice, ctx->stream));
    Mlp m;
    m.frag = (const bf16x8 *)s->frag;
    m.p = s->params;
    hipLaunchKernelGGL(k_sg_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, make_enc(s), m, n,
                       (const float *)dx, (const int *)db, dy);
    IVR_HIP(hipGetLastError());
    IVR_HIP(hipMemcpyAsync(y_out, dy, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    return IVR_OK;
}

int ivr_surrogate_set_fallback(ivr_ctx *ctx, ivr_surrogate *s, const unsigned char *flags)
{
    if (!ctx || !s || !flags) return ivr_fail(IVR_ERR_ARG, "null argument");
    IVR_HIP(hipMemcpyAsync(s->d_fallback, flags, s->nbricks, hipMemcpyHostToDevice, ctx->stream));
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    return IVR_OK;
}

int ivr_surrogate_freeze(ivr_ctx *ctx, ivr_surrogate *s)
{
    if (!ctx || !s) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (s->frozen) return IVR_OK;
    IVR_HIP(hipStreamSynchronize(ctx->stream)); /* queued work may still read the training state */
    for (float **p : {&s->tables, &s->gtables, &s->mt, &s->vt}) {
        ivr_dfree(ctx, *p, s->ntable * 4);
        *p = nullptr;
    }
    for (float **p : {&s->gparams, &s->mp, &s->vp}) {
        ivr_dfree(ctx, *p, SG_NPARAM * 4);
        *p = nullptr;
    }
    ivr_dfree(ctx, s->dE, 2 * s->dE_n * SG_IN * 4);
    s->dE = nullptr;
    s->dE_n = 0;
    train_streams_release(s);
    s->frozen = true;
    return IVR_OK;
}

int ivr_surrogate_evaluate(ivr_ctx *ctx, ivr_surrogate *s, const ivr_volume *vol, const ivr_tf *tf,
                           const ivr_frame_desc *d, float threshold_db, float *psnr_out, int *nfallback)
{
    if (!ctx || !s || !vol || !tf || !d) return ivr_fail(IVR_ERR_ARG, "null argument");
    const int W = d->state.winwidth, H = d->state.winheight;
    /* frames and buffers of the test renders, released on every return */
    struct Frames {
        ivr_ctx *ctx;
        ivr_frame *f[2] = {nullptr, nullptr};
        ~Frames()
        {
            for (ivr_frame *x : f)
                if (x) ivr_frame_destroy(ctx, x);
        }
    } frames{ctx};
    ivr_dscope flag_buf(ctx), sum_buf(ctx);
    int rc;
    if ((rc = ivr_frame_create(ctx, W, H, 0, &frames.f[0])) || (rc = ivr_frame_create(ctx, W, H, 0, &frames.f[1])) ||
        (rc = flag_buf.alloc(2 * IVR_MAX_BRICKS)) || (rc = sum_buf.alloc(8)))
        return rc;
    ivr_frame *fp = frames.f[0], *fs = frames.f[1];
    unsigned char *dflags = (unsigned char *)flag_buf.p;
    unsigned long long *dsum = (unsigned long long *)sum_buf.p;
    IVR_HIP(hipMemsetAsync(dflags, 1, IVR_MAX_BRICKS, ctx->stream));
    IVR_HIP(hipMemsetAsync(dflags + IVR_MAX_BRICKS, 0, IVR_MAX_BRICKS, ctx->stream));
    std::vector<unsigned char> flags(s->nbricks, 0);
    int nf = 0;
    const size_t np = (size_t)W * H;
    for (int b = 0; b < s->nbricks; b++) {
        if ((rc = sg_launch(ctx, s, vol, tf, d, fp, dflags, b))) return rc;
        if ((rc = sg_launch(ctx, s, vol, tf, d, fs, dflags + IVR_MAX_BRICKS, b))) return rc;
        IVR_HIP(hipMemsetAsync(dsum, 0, 8, ctx->stream));
        hipLaunchKernelGGL(k_sqerr, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ctx->stream, fp->d_rgba8,
                           fs->d_rgba8, np, dsum);
        IVR_HIP(hipGetLastError());
        unsigned long long sq = 0;
        IVR_HIP(hipMemcpyAsync(&sq, dsum, 8, hipMemcpyDeviceToHost, ctx->stream));
        IVR_HIP(hipStreamSynchronize(ctx->stream));
        const double mse = (double)sq / (double)(np * 4);
        const float psnr = mse > 0.0 ? (float)(10.0 * std::log10(255.0 * 255.0 / mse)) : 99.0f;
        if (psnr_out) psnr_out[b] = psnr;
        flags[b] = psnr < threshold_db;
        nf += flags[b];
    }
    if (nfallback) *nfallback = nf;
    return ivr_surrogate_set_fallback(ctx, s, flags.data());
}

/* checkpoint: "IVRSG005", params struct, dims, step, then MLP (params, m, v),
 * tables (values, m, v) as little-endian float32, and the 4^3-voxel
 * macrocell ranges as half pairs (4 bytes a cell). Older files still load:
 * "IVRSG004" and before hold float32 ranges of 16^3-voxel cells, which are
 * read and dropped (the ranges computed from the volume at create stay); "IVRSG003" and before stored a hashed level's
 * 4x4x4 block Morton-ordered in its slot (x0 y0 z0 x1 y1 z1), which is
 * permuted to the current order on load; "IVRSG002" files have a params
 * struct without log2_group (then 0), "IVRSG001" files no ranges either. */
namespace {

/* a table section [brick][level][T][2] written with the slots of hashed
 * levels in Morton order (place x0 y0 z0 x1 y1 z1 of a block's corner),
 * rearranged in place to corner_index's order */
void morton_slots_to_rows(const ivr_surrogate *s, float *t)
{
    unsigned to[64];
    for (unsigned m = 0; m < 64; m++) {
        const unsigned x = (m & 1) | (m >> 2 & 2), y = (m >> 1 & 1) | (m >> 3 & 2), z = (m >> 2 & 1) | (m >> 4 & 2);
        to[m] = x | (y & 1) << 2 | (z & 1) << 3 | (y & 2) << 3 | (z & 2) << 4;
    }
    float slot[128];
    for (int b = 0; b < s->nbricks; b++)
        for (int l = 0; l < SG_L; l++) {
            if (s->dense_n1[l]) continue;
            float *lv = t + ((size_t)b * SG_L + l) * (size_t)s->T * SG_F;
            for (size_t q = 0; q < s->T; q += 64) {
                float *p = lv + q * SG_F;
                for (unsigned m = 0; m < 64; m++) {
                    slot[to[m] * 2] = p[m * 2];
                    slot[to[m] * 2 + 1] = p[m * 2 + 1];
                }
                memcpy(p, slot, sizeof slot);
            }
        }
}

} // namespace

int ivr_surrogate_save(ivr_ctx *ctx, const ivr_surrogate *s, const char *path)
{
    if (!ctx || !s || !path) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (s->frozen) return ivr_fail(IVR_ERR_UNSUPPORTED, "the surrogate is frozen (no training state)");
    FILE *f = fopen(path, "wb");
    if (!f) return ivr_fail(IVR_ERR_IO, "cannot open %s", path);
    const char magic[8] = {'I', 'V', 'R', 'S', 'G', '0', '0', '5'};
    int hdr[8] = {s->nbricks, s->B, s->n[0], s->n[1], s->n[2], s->step, (int)s->T, SG_NPARAM};
    bool ok = fwrite(magic, 8, 1, f) == 1 && fwrite(&s->sp, sizeof s->sp, 1, f) == 1 &&
              fwrite(hdr, sizeof hdr, 1, f) == 1;
    std::vector<float> buf;
    const float *dev[7] = {s->params, s->mp, s->vp, s->tables, s->mt, s->vt, (const float *)s->d_cellrange};
    const size_t cnt[7] = {SG_NPARAM, SG_NPARAM, SG_NPARAM, s->ntable, s->ntable, s->ntable, s->ncells};
    for (int k = 0; k < 7 && ok; k++) {
        buf.resize(cnt[k]);
        if (hipMemcpy(buf.data(), dev[k], cnt[k] * 4, hipMemcpyDeviceToHost) != hipSuccess) ok = false;
        else ok = fwrite(buf.data(), 4, cnt[k], f) == cnt[k];
    }
    fclose(f);
    return ok ? IVR_OK : ivr_fail(IVR_ERR_IO, "writing %s failed", path);
}

int ivr_surrogate_load(ivr_ctx *ctx, ivr_surrogate *s, const char *path)
{
    if (!ctx || !s || !path) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (s->frozen) return ivr_fail(IVR_ERR_UNSUPPORTED, "the surrogate is frozen (no training state)");
    FILE *f = fopen(path, "rb");
    if (!f) return ivr_fail(IVR_ERR_IO, "cannot open %s", path);
    char magic[8];
    ivr_surrogate_params sp;
    int hdr[8];
    bool ok = fread(magic, 8, 1, f) == 1 && !memcmp(magic, "IVRSG00", 7) && magic[7] >= '1' && magic[7] <= '5';
    /* versions 1 and 2 stored the struct up to seed */
    const size_t sp_bytes = ok && magic[7] < '3' ? offsetof(ivr_surrogate_params, log2_group) : sizeof sp;
    memset(&sp, 0, sizeof sp);
    ok = ok && fread(&sp, sp_bytes, 1, f) == 1 && fread(hdr, sizeof hdr, 1, f) == 1;
    if (!ok) {
        fclose(f);
        return ivr_fail(IVR_ERR_IO, "%s is not a surrogate checkpoint", path);
    }
    const int nsec = magic[7] >= '2' ? 7 : 6;
    /* the table layout (level resolutions, dense/hashed split, slots) is
     * fixed at create by log2_table, base_res and growth: a checkpoint made
     * with other values would load without error and render wrong values */
    if (hdr[0] != s->nbricks || hdr[1] != s->B || hdr[2] != s->n[0] || hdr[3] != s->n[1] || hdr[4] != s->n[2] ||
        hdr[6] != (int)s->T || hdr[7] != SG_NPARAM || sp.log2_table != s->sp.log2_table ||
        sp.base_res != s->sp.base_res || sp.growth != s->sp.growth || sp.log2_group != s->sp.log2_group) {
        fclose(f);
        return ivr_fail(IVR_ERR_ARG, "checkpoint %s does not match this surrogate (bricks, dims, table size or "
                                     "level resolutions differ)", path);
    }
    /* every section is read into host memory before the device state is touched,
     * so a short or failed read leaves the surrogate as it was */
    float *dev[7] = {s->params, s->mp, s->vp, s->tables, s->mt, s->vt, (float *)s->d_cellrange};
    size_t cnt[7] = {SG_NPARAM, SG_NPARAM, SG_NPARAM, s->ntable, s->ntable, s->ntable, s->ncells};
    /* files before "IVRSG005" hold ranges of 16^3-voxel cells */
    const bool old_ranges = nsec == 7 && magic[7] < '5';
    if (old_ranges) {
        size_t n16 = 1;
        for (int r = 0; r < 3; r++) n16 *= (size_t)(((s->n[r] - 2) >> 4) + 1);
        cnt[6] = 2 * n16;
    }
    std::vector<std::vector<float>> buf(nsec);
    for (int k = 0; k < nsec && ok; k++) {
        buf[k].resize(cnt[k]);
        ok = fread(buf[k].data(), 4, cnt[k], f) == cnt[k];
    }
    fclose(f);
    if (!ok) return ivr_fail(IVR_ERR_IO, "reading %s failed", path);
    if (magic[7] < '4') /* Morton-ordered slots of the hashed levels to the current order */
        for (int k = 3; k < 6; k++) morton_slots_to_rows(s, buf[k].data());
    for (int k = 0; k < (old_ranges ? 6 : nsec); k++)
        IVR_HIP(hipMemcpy(dev[k], buf[k].data(), cnt[k] * 4, hipMemcpyHostToDevice));
    s->sp = sp;
    s->step = hdr[5];
    int rc = refresh_frags(ctx, s);
    if (rc) return rc;
    IVR_HIP(hipStreamSynchronize(ctx->stream));
    return IVR_OK;
}
