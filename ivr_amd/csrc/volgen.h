/*
 * volgen.h -- synthetic volume generator, compiled for both the host and the
 * device passes of hipcc from this one source so that CPU and GPU volumes
 * are bit-identical (contraction is off for the whole library and every
 * fused multiply-add is an explicit fmaf).
 *
 *  - Perlin "improved noise" (K. Perlin, "Improving Noise", SIGGRAPH 2002):
 *    quintic fade 6t^5-15t^4+10t^3, 12 edge gradients selected by hash&15,
 *    permutation table shuffled from the seed.
 *  - fBm: sum over octaves of gain^o * noise(freq * lacunarity^o * x).
 *  - Gaussian blobs: A * exp(-|x-c|^2 / (2 sigma^2)), with exp evaluated by
 *    Cody-Waite range reduction (x = k ln2 + r) and the degree-7 Taylor
 *    polynomial of e^r, so both passes compute the same bits.
 * Voxel (ix,iy,iz) of an n^3 grid sits at ((ix+0.5)/nx, (iy+0.5)/ny, (iz+0.5)/nz).
 */
#ifndef IVR_VOLGEN_H
#define IVR_VOLGEN_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define IVR_MAX_BLOBS 64

struct GenTable {
    int kind;
    int nblobs;
    int octaves;
    float lacunarity, gain, base_freq;
    float blob[IVR_MAX_BLOBS][5]; /* cx, cy, cz, 1/(2 sigma^2), amplitude */
    unsigned char perm[512];
};

__host__ __device__ inline float gen_exp(float x)
{
    if (x < -86.0f) return 0.0f;
    if (x > 88.0f) x = 88.0f;
    float kf = floorf(x * 1.44269504f + 0.5f);
    float r = fmaf(kf, -0.693145751953125f, x);    /* ln2 high part (exact product) */
    r = fmaf(kf, -1.42860682030941723e-06f, r);    /* ln2 low part                 */
    float p = 1.0f / 5040.0f;
    p = fmaf(p, r, 1.0f / 720.0f);
    p = fmaf(p, r, 1.0f / 120.0f);
    p = fmaf(p, r, 1.0f / 24.0f);
    p = fmaf(p, r, 1.0f / 6.0f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int k = (int)kf;
    union { uint32_t u; float f; } sc;
    sc.u = (uint32_t)(k + 127) << 23;
    return p * sc.f;
}

__host__ __device__ inline float gen_fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }

__host__ __device__ inline float gen_lerp(float t, float a, float b) { return a + t * (b - a); }

__host__ __device__ inline float gen_grad(int h, float x, float y, float z)
{
    h &= 15;
    float u = h < 8 ? x : y;
    float v = h < 4 ? y : ((h == 12 || h == 14) ? x : z);
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
}

__host__ __device__ inline float gen_noise(const unsigned char *p, float x, float y, float z)
{
    float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    int X = (int)fx & 255, Y = (int)fy & 255, Z = (int)fz & 255;
    x -= fx;
    y -= fy;
    z -= fz;
    float u = gen_fade(x), v = gen_fade(y), w = gen_fade(z);
    int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z;
    int B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
    float x1 = x - 1.0f, y1 = y - 1.0f, z1 = z - 1.0f;
    float l0 = gen_lerp(u, gen_grad(p[AA], x, y, z), gen_grad(p[BA], x1, y, z));
    float l1 = gen_lerp(u, gen_grad(p[AB], x, y1, z), gen_grad(p[BB], x1, y1, z));
    float l2 = gen_lerp(u, gen_grad(p[AA + 1], x, y, z1), gen_grad(p[BA + 1], x1, y, z1));
    float l3 = gen_lerp(u, gen_grad(p[AB + 1], x, y1, z1), gen_grad(p[BB + 1], x1, y1, z1));
    return gen_lerp(w, gen_lerp(v, l0, l1), gen_lerp(v, l2, l3));
}

/* raw (un-normalised) field at voxel (ix,iy,iz) */
__host__ __device__ inline float gen_field(const GenTable &g, int ix, int iy, int iz, int nx, int ny,
                                           int nz)
{
    float x = ((float)ix + 0.5f) / (float)nx;
    float y = ((float)iy + 0.5f) / (float)ny;
    float z = ((float)iz + 0.5f) / (float)nz;
    float f = 0.0f;
    if (g.kind == 1) {
        float amp = 1.0f, freq = g.base_freq;
        for (int o = 0; o < g.octaves; o++) {
            f = f + amp * gen_noise(g.perm, x * freq, y * freq, z * freq);
            amp = amp * g.gain;
            freq = freq * g.lacunarity;
        }
    }
    for (int b = 0; b < g.nblobs; b++) {
        float dx = x - g.blob[b][0], dy = y - g.blob[b][1], dz = z - g.blob[b][2];
        float r2 = dx * dx + dy * dy + dz * dz;
        f = f + g.blob[b][4] * gen_exp(-r2 * g.blob[b][3]);
    }
    return f;
}

/* final value: blobs min-max normalised; noise zeroed at or below the
 * threshold and normalised above it; then scaled */
__host__ __device__ inline float gen_normalise(float v, int kind, float lo, float hi, float scale)
{
    float o;
    if (kind == 1) o = v > lo ? (v - lo) / (hi - lo) : 0.0f;
    else o = (v - lo) / (hi - lo);
    return o * scale;
}

/* order-preserving map of a float to uint32 (for radix selection) */
__host__ __device__ inline uint32_t gen_key(float v)
{
    union { float f; uint32_t u; } c;
    c.f = v;
    return (c.u & 0x80000000u) ? ~c.u : (c.u | 0x80000000u);
}

__host__ __device__ inline float gen_unkey(uint32_t k)
{
    union { float f; uint32_t u; } c;
    c.u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return c.f;
}

#endif
