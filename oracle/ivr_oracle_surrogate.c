/*
 * ivr_oracle_surrogate.c -- CPU oracle of the learned volume surrogate.
 *
 * TEST INFRASTRUCTURE ONLY (see ivr_oracle.h).  The surrogate has no
 * counterpart in the reference; this file restates its published algorithms
 * in double precision so the GPU kernels can be checked:
 *   - multiresolution hash encoding (T. Mueller et al., "Instant Neural
 *     Graphics Primitives with a Multiresolution Hash Encoding", SIGGRAPH
 *     2022): level l has resolution floor(N_min * b^l); a level whose
 *     (N+1)^3 vertices fit in T entries is indexed densely, otherwise the
 *     4x4x4 block of vertices holding the vertex is placed by the spatial
 *     hash (bx*1 ^ by*2654435761 ^ bz*805459861) mod T/64 of its block
 *     coordinates into a slot of 64 entries, at place x1 x0 | y0 | z0 | y1 |
 *     z1 (low bits first) inside it (a locality-keeping variant of the
 *     paper's per-vertex hash); features of
 *     the 8 surrounding vertices are blended trilinearly;
 *   - MLP 32 -> 64 -> 64 -> 1 with ReLU hidden layers and a linear output,
 *     mean-squared-error loss, analytic backward pass;
 *   - Adam (D. Kingma, J. Ba, ICLR 2015) with bias correction.
 * With bf16 != 0 the operands of the two 64-wide layers' matrix products
 * (encoding, first hidden layer, both layers' weights and back-propagated
 * pre-activation gradients) are rounded to bfloat16 (round to nearest even)
 * where the GPU rounds them; the 64 -> 1 output layer and all accumulation
 * stay unrounded.  With bf16 == 0 everything is double (for
 * finite-difference checks).
 *
 * Pinned by: finite differences of the double path (gradients of MLP
 * weights, biases and hash-table entries), hash-function known answers and
 * the dense/hash level split, and the Adam update on known vectors.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define SG_L 16
#define SG_F 2
#define SG_IN (SG_L * SG_F)
#define SG_H 64

static double bf16r(double x, int on)
{
    if (!on) return x;
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    uint32_t r = 0x7fffu + ((u >> 16) & 1u);
    u = (u + r) & 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

int ivros_level_res(int l, int nmin, double growth) { return (int)floor(nmin * pow(growth, l)); }

uint32_t ivros_index(int l, int nmin, double growth, uint32_t T, int x, int y, int z)
{
    const uint32_t n1 = (uint32_t)ivros_level_res(l, nmin, growth) + 1u;
    if ((uint64_t)n1 * n1 * n1 <= T) return (uint32_t)x + n1 * ((uint32_t)y + n1 * (uint32_t)z);
    /* hashed: block (x/4, y/4, z/4) -> slot of 64 entries by the primes
     * 1, 2654435761, 805459861; the corner's place in the slot is, from the
     * lowest bit up, x mod 4 (two bits), then y0, z0, y1, z1 of its low
     * coordinate bits */
    const uint32_t bx = (uint32_t)x / 4u, by = (uint32_t)y / 4u, bz = (uint32_t)z / 4u;
    const uint32_t slot = (bx * 1u ^ by * 2654435761u ^ bz * 805459861u) % (T / 64u);
    const uint32_t lx = (uint32_t)x % 4u, ly = (uint32_t)y % 4u, lz = (uint32_t)z % 4u;
    const uint32_t local = lx + 4u * (ly % 2u) + 8u * (lz % 2u) + 16u * (ly / 2u) + 32u * (lz / 2u);
    return slot * 64u + local;
}

/* encode one point x in [0,1]^3 with table[L][T][F] (float), double result */
void ivros_encode(const float *table, uint32_t T, int nmin, double growth, const double x[3], double *enc)
{
    for (int l = 0; l < SG_L; l++) {
        const int N = ivros_level_res(l, nmin, growth);
        double u[3], w[3];
        int v0[3];
        for (int a = 0; a < 3; a++) {
            u[a] = x[a] * N;
            v0[a] = (int)floor(u[a]);
            if (v0[a] > N - 1) v0[a] = N - 1;
            if (v0[a] < 0) v0[a] = 0;
            w[a] = u[a] - v0[a];
        }
        double f[SG_F] = {0, 0};
        for (int c = 0; c < 8; c++) {
            const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
            const double wc = (dx ? w[0] : 1 - w[0]) * (dy ? w[1] : 1 - w[1]) * (dz ? w[2] : 1 - w[2]);
            const uint32_t idx = ivros_index(l, nmin, growth, T, v0[0] + dx, v0[1] + dy, v0[2] + dz);
            for (int k = 0; k < SG_F; k++) f[k] += wc * table[((size_t)l * T + idx) * SG_F + k];
        }
        for (int k = 0; k < SG_F; k++) enc[l * SG_F + k] = f[k];
    }
}

/* scatter d(loss)/d(enc) of one point into the table gradient */
void ivros_encode_backward(uint32_t T, int nmin, double growth, const double x[3], const double *denc,
                           double *dtable)
{
    for (int l = 0; l < SG_L; l++) {
        const int N = ivros_level_res(l, nmin, growth);
        double u[3], w[3];
        int v0[3];
        for (int a = 0; a < 3; a++) {
            u[a] = x[a] * N;
            v0[a] = (int)floor(u[a]);
            if (v0[a] > N - 1) v0[a] = N - 1;
            if (v0[a] < 0) v0[a] = 0;
            w[a] = u[a] - v0[a];
        }
        for (int c = 0; c < 8; c++) {
            const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
            const double wc = (dx ? w[0] : 1 - w[0]) * (dy ? w[1] : 1 - w[1]) * (dz ? w[2] : 1 - w[2]);
            const uint32_t idx = ivros_index(l, nmin, growth, T, v0[0] + dx, v0[1] + dy, v0[2] + dz);
            for (int k = 0; k < SG_F; k++) dtable[((size_t)l * T + idx) * SG_F + k] += wc * denc[l * SG_F + k];
        }
    }
}

/* Parameter block (float), in this order:
 *   W1[64][32] b1[64] W2[64][64] b2[64] W3[64] b3[1]   (6337 values)
 * Forward/backward over n encoded points enc[n][32] with targets t[n].
 * Outputs: y[n]; grads (same layout, double) of L = mean((y - t)^2);
 * denc[n][32]; returns L. */
double ivros_mlp(const float *P, const double *enc, const float *t, int n, int bf16, double *y,
                 double *grads, double *denc)
{
    const float *W1 = P, *b1 = W1 + SG_H * SG_IN, *W2 = b1 + SG_H, *b2 = W2 + SG_H * SG_H,
                *W3 = b2 + SG_H, *b3 = W3 + SG_H;
    double *gW1 = grads, *gb1 = gW1 + SG_H * SG_IN, *gW2 = gb1 + SG_H, *gb2 = gW2 + SG_H * SG_H,
           *gW3 = gb2 + SG_H, *gb3 = gW3 + SG_H;
    memset(grads, 0, sizeof(double) * (SG_H * SG_IN + SG_H + SG_H * SG_H + SG_H + SG_H + 1));
    double loss = 0.0;
    for (int s = 0; s < n; s++) {
        double e[SG_IN], z1[SG_H], h1[SG_H], z2[SG_H], h2[SG_H];
        for (int k = 0; k < SG_IN; k++) e[k] = bf16r(enc[(size_t)s * SG_IN + k], bf16);
        for (int o = 0; o < SG_H; o++) {
            double acc = 0.0;
            for (int k = 0; k < SG_IN; k++) acc += bf16r(W1[o * SG_IN + k], bf16) * e[k];
            z1[o] = acc + b1[o];
            h1[o] = bf16r(z1[o] > 0 ? z1[o] : 0.0, bf16);
        }
        for (int o = 0; o < SG_H; o++) {
            double acc = 0.0;
            for (int k = 0; k < SG_H; k++) acc += bf16r(W2[o * SG_H + k], bf16) * h1[k];
            z2[o] = acc + b2[o];
            h2[o] = z2[o] > 0 ? z2[o] : 0.0; /* layer 3 runs in fp32 on the GPU */
        }
        double acc = 0.0;
        for (int k = 0; k < SG_H; k++) acc += (double)W3[k] * h2[k];
        const double out = acc + b3[0];
        y[s] = out;
        const double r = out - t[s];
        loss += r * r;
        const double dy = 2.0 * r / n;
        /* backward; activations and weights enter as the forward used them */
        double dz2[SG_H], dz1[SG_H];
        *gb3 += dy;
        for (int k = 0; k < SG_H; k++) {
            gW3[k] += dy * h2[k];
            const double dh = (double)W3[k] * dy;
            dz2[k] = z2[k] > 0 ? bf16r(dh, bf16) : 0.0;
        }
        for (int o = 0; o < SG_H; o++) {
            gb2[o] += dz2[o];
            for (int k = 0; k < SG_H; k++) gW2[o * SG_H + k] += dz2[o] * h1[k];
        }
        for (int k = 0; k < SG_H; k++) {
            double dh = 0.0;
            for (int o = 0; o < SG_H; o++) dh += bf16r(W2[o * SG_H + k], bf16) * dz2[o];
            dz1[k] = z1[k] > 0 ? bf16r(dh, bf16) : 0.0;
        }
        for (int o = 0; o < SG_H; o++) {
            gb1[o] += dz1[o];
            for (int k = 0; k < SG_IN; k++) gW1[o * SG_IN + k] += dz1[o] * e[k];
        }
        for (int k = 0; k < SG_IN; k++) {
            double de = 0.0;
            for (int o = 0; o < SG_H; o++) de += bf16r(W1[o * SG_IN + k], bf16) * dz1[o];
            denc[(size_t)s * SG_IN + k] = de;
        }
    }
    return loss / n;
}

/* one Adam step over n parameters (float state), step t >= 1 */
void ivros_adam(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2,
                float eps, int t)
{
    const double c1 = 1.0 - pow(b1, t), c2 = 1.0 - pow(b2, t);
    for (long long i = 0; i < n; i++) {
        const double mi = b1 * (double)m[i] + (1.0 - b1) * g[i];
        const double vi = b2 * (double)v[i] + (1.0 - b2) * (double)g[i] * g[i];
        m[i] = (float)mi;
        v[i] = (float)vi;
        p[i] = (float)(p[i] - lr * (mi / c1) / (sqrt(vi / c2) + eps));
    }
}

/* Forward only: the network's value for n points of one brick (encoding
 * then the MLP with the same bf16 rounding points as ivros_mlp's forward).
 * The CPU baseline of the surrogate's per-sample evaluation; OpenMP over
 * the points. */
void ivros_eval(const float *table, uint32_t T, int nmin, double growth, const float *P, const double *xyz,
                int n, int bf16, double *y)
{
    const float *W1 = P, *b1 = W1 + SG_H * SG_IN, *W2 = b1 + SG_H, *b2 = W2 + SG_H * SG_H,
                *W3 = b2 + SG_H, *b3 = W3 + SG_H;
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; s++) {
        double enc[SG_IN], e[SG_IN], h1[SG_H];
        ivros_encode(table, T, nmin, growth, xyz + 3 * (size_t)s, enc);
        for (int k = 0; k < SG_IN; k++) e[k] = bf16r(enc[k], bf16);
        for (int o = 0; o < SG_H; o++) {
            double acc = 0.0;
            for (int k = 0; k < SG_IN; k++) acc += bf16r(W1[o * SG_IN + k], bf16) * e[k];
            const double z = acc + b1[o];
            h1[o] = bf16r(z > 0 ? z : 0.0, bf16);
        }
        double out = 0.0;
        for (int o = 0; o < SG_H; o++) {
            double acc = 0.0;
            for (int k = 0; k < SG_H; k++) acc += bf16r(W2[o * SG_H + k], bf16) * h1[k];
            const double z = acc + b2[o];
            out += (double)W3[o] * (z > 0 ? z : 0.0);
        }
        y[s] = out + b3[0];
    }
}
