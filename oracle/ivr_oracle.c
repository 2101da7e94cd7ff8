/*
 * ivr_oracle.c -- CPU oracle for the iVR render path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, as the checker (and the timed CPU baseline);
 * never linked into or called by the product library.
 *
 * A clean-room restatement of the reference render path: ray set-up
 * (mathutil.c:1115-1154, main.c:937-1009), slab test (ray.c:276-318),
 * sampling (ray.c:87-273), the voxel walk (ray.c:333-483), the gradient
 * field (mathutil.c:242-338), transfer functions (transferfuncs.c),
 * colormaps (colormap.c), tile compositing (main.c:1203-1218) and the RGBA8
 * quantiser (main.c:1404).  Compat mode reproduces the reference's float
 * operation order exactly (this file is compiled with -ffp-contract=off, so
 * fmaf() is the only fused multiply-add and compat mode uses none); config
 * mode is the benchmark contract (see ivr_oracle.h).
 *
 * Pinning (tests/test_oracle.py): compat mode, samplers (linear and
 * nearest), slab test, sample marching, the voxel walk, transfer functions,
 * colormaps and tile compositing are compared bit-for-bit with the reference
 * compiled from its own sources (oracle/_ref, golden fixtures under
 * tests/golden); config mode against closed forms (affine fields under true
 * trilinear, constant-slab opacity, ERT bound, BTF==FTB) and the FP64 path;
 * the gradient field against numpy.gradient and an affine field.
 */
#include "ivr_oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef struct { float r, g, b, a; } col4;

/* ------------------------------------------------------------------ */
/* small float helpers with the reference's evaluation order           */
/* ------------------------------------------------------------------ */

/* dot3f (mathutil.c:28-33): ((a.x*b.x + a.y*b.y) + a.z*b.z) */
static inline float dot3(const float a[3], const float b[3])
{
    float p = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    return p;
}

/* dot4f (mathutil.c:34-39) */
static inline float dot4(const float a[4], const float b[4])
{
    float p = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
    return p;
}

/* mat3fxvec3f (mathutil.c:129-149): out[r] = dot3f(row r, v) */
static void m3v(const float m[3][3], const float v[3], float out[3])
{
    float t[3];
    for (int r = 0; r < 3; r++) t[r] = dot3(m[r], v);
    out[0] = t[0]; out[1] = t[1]; out[2] = t[2];
}

/* mat3fxmat3f (mathutil.c:150-171): inout = inout * in */
static void m3m(float io[3][3], const float in[3][3])
{
    float t[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            float col[3] = {in[0][j], in[1][j], in[2][j]};
            t[i][j] = dot3(io[i], col);
        }
    memcpy(io, t, sizeof t);
}

/* mat4fxmat4f (mathutil.c:172-196): inout = inout * in */
static void m4m(float io[4][4], const float in[4][4])
{
    float t[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float col[4] = {in[0][j], in[1][j], in[2][j], in[3][j]};
            t[i][j] = dot4(io[i], col);
        }
    memcpy(io, t, sizeof t);
}

/* mat4fxvec4f (mathutil.c:197-218) */
static void m4v(const float m[4][4], const float v[4], float out[4])
{
    float t[4];
    for (int r = 0; r < 4; r++) t[r] = dot4(m[r], v);
    memcpy(out, t, sizeof t);
}

/* setmvm / setpm (mathutil.h:211-233): GL column-major -> row-major */
static void from_gl(const float cm[16], float m[4][4])
{
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) m[r][c] = cm[c * 4 + r];
}

/* ------------------------------------------------------------------ */
/* volume box                                                         */
/* ------------------------------------------------------------------ */

/* initvolbbox + getlocalbrick with one brick (main.c:509-598,
 * utils.c:332-376): scalefactor = dim * (1/maxdim); box [-s, -s + 2s]. */
void ivro_box(int nx, int ny, int nz, float bmin[3], float bmax[3])
{
    int dims[3] = {nx, ny, nz};
    int md = nx > ny ? nx : ny;
    md = md > nz ? md : nz;
    float invdim = 1.0f / (float)md;
    for (int a = 0; a < 3; a++) {
        float s = (float)dims[a] * invdim;
        float range = s * 2.0f;
        float w = 2.0f * s * 1.0f;            /* brick->w, ratio 1 */
        float x = -s + range * 0.0f;          /* brick->x, start 0 */
        bmin[a] = x;
        bmax[a] = x + w;
    }
}

/* pzero (main.c:615-617) combined with newconst (setplaneconst, ray.c:44-51) */
static void plane_consts(const float bmin[3], const float bmax[3], const float nc[6],
                         float dconst[6])
{
    float pz[6] = {bmin[0] * -1.0f, bmax[0] * 1.0f, bmin[1] * -1.0f,
                   bmax[1] * 1.0f,  bmin[2] * -1.0f, bmax[2] * 1.0f};
    for (int i = 0; i < 6; i++) dconst[i] = (nc[i] < pz[i]) ? nc[i] : pz[i];
}

/* face normals (main.c:611-613) */
static const float kNormals[6][3] = {
    {-1.0f, 0.0f, 0.0f}, {1.0f, 0.0f, 0.0f}, {0.0f, -1.0f, 0.0f},
    {0.0f, 1.0f, 0.0f},  {0.0f, 0.0f, -1.0f}, {0.0f, 0.0f, 1.0f}};

/* getrayintersect (ray.c:276-318) */
static int slab(const float dconst[6], const float o[3], const float d[3], float *tin,
                float *tout)
{
    float ti = -FLT_MAX, to = FLT_MAX;
    int i = 0;
    while (i < 6 && ti < to) {
        float numer = dconst[i] - dot3(kNormals[i], o);
        float denom = dot3(kNormals[i], d);
        if (denom == 0.0f) {
            if (numer < 0.0f) ti = to + 1.0f;
        } else {
            float thit = numer / denom;
            if (denom > 0.0f) to = (float)fmin(to, thit);
            else ti = (float)fmax(ti, thit);
        }
        i++;
    }
    *tin = ti;
    *tout = to;
    if (ti < to) {
        if (ti > 0.0f) return 1;
        if (to > 0.0f) return 1;
    }
    return 0;
}

int ivro_rayintersect(const float bmin[3], const float bmax[3], const float nc[6],
                      const float ray[6], float *tin, float *tout)
{
    float dc[6];
    plane_consts(bmin, bmax, nc, dc);
    return slab(dc, ray, ray + 3, tin, tout);
}

/* ------------------------------------------------------------------ */
/* voxel walk (Amanatides & Woo, as in ray.c:333-483)                 */
/* ------------------------------------------------------------------ */

/* getfirstvoxel (ray.c:333-384).  Each axis with a non-zero direction
 * offers the parameters bmin/d and bmax/d, in the order min x, max x, min y,
 * ... (the ray origin does not enter them, ray.c:339-361); the first
 * strictly smallest positive one is taken (ray.c:362-369) and the entry
 * point d*t + o is rounded half up to an index per axis (ray.c:371-378).
 * The hitface() tests there feed an array nothing reads, so they are not
 * restated.  No positive parameter: -1 with first[] untouched, where the
 * reference reads one element past its times[] array. */
int ivro_firstvoxel(const float bmin[3], const float bmax[3], const float ray[6], int first[3])
{
    const float *o = ray, *d = ray + 3;
    float best = 0.0f;
    int found = 0;
    for (int a = 0; a < 3; a++) {
        if (d[a] == 0.0f) continue; /* both parameters are 0 there: never taken */
        const float cand[2] = {bmin[a] / d[a], bmax[a] / d[a]};
        for (int s = 0; s < 2; s++)
            if (cand[s] > 0.0f && (cand[s] < best || best == 0.0f)) {
                best = cand[s];
                found = 1;
            }
    }
    if (!found) return -1;
    for (int a = 0; a < 3; a++) {
        const float v = d[a] * best + o[a];
        first[a] = (int)(v + 0.5f);
    }
    return 0;
}

/* getvoxelsonray (ray.c:387-483): from the entry voxel, step to the
 * neighbour across the nearest cell face, maxdepth times, recording each
 * voxel's maptoffset3d(x, y, z, w, h, d) = ((z*h + x)*w + y (mathutil.h:139-143,
 * x and y in that order) converted to float.  Face parameters are
 * (boundary - o)/d with NaN replaced by +inf (ray.c:436-455); ties go to z,
 * then y (strict comparisons, ray.c:462-481).  The offset is formed in
 * wrapping 32-bit arithmetic. */
int ivro_voxelsonray(const float bmin[3], const float bmax[3], const int dims[3], const float ray[6],
                     int maxdepth, float *samples)
{
    const float *o = ray, *d = ray + 3;
    int v[3], step[3];
    float tmax[3], delta[3];
    if (ivro_firstvoxel(bmin, bmax, ray, v) < 0) return -1;
    for (int a = 0; a < 3; a++) {
        step[a] = d[a] > 0.0f ? 1 : (d[a] < 0.0f ? -1 : 0);
        const float boundary = (float)(v[a] + (step[a] > 0 ? 1 : 0));
        tmax[a] = (boundary - o[a]) / d[a];
        if (isnan(tmax[a])) tmax[a] = INFINITY;
        delta[a] = (float)step[a] / d[a];
        if (isnan(delta[a])) delta[a] = INFINITY;
    }
    for (int i = 0; i < maxdepth; i++) {
        const unsigned off = ((unsigned)v[2] * (unsigned)dims[1] + (unsigned)v[0]) * (unsigned)dims[0] +
                             (unsigned)v[1];
        samples[i] = (float)(int)off;
        int a;
        if (tmax[0] < tmax[1] && tmax[0] < tmax[2]) a = 0;
        else if (tmax[1] < tmax[2]) a = 1;
        else a = 2;
        v[a] += step[a];
        tmax[a] += delta[a];
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* gradient field                                                     */
/* ------------------------------------------------------------------ */

/* gradient3d (mathutil.c:242-328) with grad() = the voxel at
 * maptoffset3d(i, j, k) (mathutil.c:329-338): component x differences along
 * j (x), y along i (y), z along k (z), each (a - b) / 2.0f * 1.0f; the
 * interior takes the two neighbours, the faces i = 0 / m-1 (and j, k) the
 * voxel and its inner neighbour (mathutil.c:265-326), here for every voxel
 * and axis independently, stored at the voxel's own offset instead of the
 * reference's i*j*k; magnitude mag3f (mathutil.c:233-236). */
void ivro_gradient(const float *vol, int nx, int ny, int nz, float *out)
{
    const long long sx = 1, sy = nx, sz = (long long)nx * ny;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const long long o = z * sz + y * sy + x;
                const int pos[3] = {x, y, z}, n[3] = {nx, ny, nz};
                const long long st[3] = {sx, sy, sz};
                float g[3];
                for (int a = 0; a < 3; a++) {
                    long long hi = o + st[a], lo = o - st[a];
                    if (pos[a] == 0) lo = o;
                    if (pos[a] == n[a] - 1) hi = o;
                    if (n[a] == 1) hi = lo = o;
                    g[a] = (vol[hi] - vol[lo]) / 2.0f * 1.0f;
                }
                float *q = out + 4 * o;
                q[0] = g[0];
                q[1] = g[1];
                q[2] = g[2];
                q[3] = sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
            }
}

/* ------------------------------------------------------------------ */
/* samplers                                                           */
/* ------------------------------------------------------------------ */

#define VOX(v, x, y, z, nx, ny) ((v)[((size_t)(z) * (size_t)(ny) + (size_t)(y)) * (size_t)(nx) + (size_t)(x)])

/* getnearestsample, non-BGQ branch (ray.c:225-237): index = (int)(n * ((p - min) / extent))
 * per axis, clamped to n-2 above (and 0 below, where the reference reads before the array);
 * the voxel's bit pattern as an unsigned int (volume.data.i, types.h:139-141) converted to
 * float */
float ivro_sample_nearest(const float *vol, int nx, int ny, int nz, const float bmin[3],
                          const float bmax[3], float x, float y, float z)
{
    const float w = bmax[0] - bmin[0], h = bmax[1] - bmin[1], d = bmax[2] - bmin[2];
    int j = (int)((float)nx * ((x - bmin[0]) / w));
    int i = (int)((float)ny * ((y - bmin[1]) / h));
    int k = (int)((float)nz * ((z - bmin[2]) / d));
    i = i >= ny - 2 ? ny - 2 : (i < 0 ? 0 : i);
    j = j >= nx - 2 ? nx - 2 : (j < 0 ? 0 : j);
    k = k >= nz - 2 ? nz - 2 : (k < 0 ? 0 : k);
    uint32_t bits;
    memcpy(&bits, &VOX(vol, j, i, k, nx, ny), sizeof bits);
    return (float)bits;
}

/* getlinearsample, non-BGQ branch (ray.c:122-178), including its weight
 * scale (ext/N instead of N/ext) and the x-weight blending the +y
 * neighbour.  Deviation: indices are clamped below at 0, where the
 * reference would read before the array. */
float ivro_sample_ref(const float *vol, int nx, int ny, int nz, const float bmin[3],
                      const float bmax[3], float x, float y, float z)
{
    float aw = bmax[0] - bmin[0], ah = bmax[1] - bmin[1], ad = bmax[2] - bmin[2];
    float ratio = (x - bmin[0]) / aw;
    int j = (int)((float)nx * ratio);
    ratio = (y - bmin[1]) / ah;
    int i = (int)((float)ny * ratio);
    ratio = (z - bmin[2]) / ad;
    int k = (int)((float)nz * ratio);
    i = (i > ny - 2) ? ny - 2 : i;
    j = (j > nx - 2) ? nx - 2 : j;
    k = (k > nz - 2) ? nz - 2 : k;
    i = i < 0 ? 0 : i;
    j = j < 0 ? 0 : j;
    k = k < 0 ? 0 : k;
    /* maptoffset3d(i,j,k) = ((k*h)+i)*w + j : i is the row (y), j the column (x) */
    float v000 = VOX(vol, j, i, k, nx, ny);
    float v100 = VOX(vol, j, i + 1, k, nx, ny);
    float v010 = VOX(vol, j + 1, i, k, nx, ny);
    float v001 = VOX(vol, j, i, k + 1, nx, ny);
    float v101 = VOX(vol, j, i + 1, k + 1, nx, ny);
    float v011 = VOX(vol, j + 1, i, k + 1, nx, ny);
    float v110 = VOX(vol, j + 1, i + 1, k, nx, ny);
    float v111 = VOX(vol, j + 1, i + 1, k + 1, nx, ny);
    /* xp = x + fabs(min.x): a double sum of two floats rounded to float,
     * identical to the float sum */
    float xp = (float)((double)x + fabs((double)bmin[0]));
    float yp = (float)((double)y + fabs((double)bmin[1]));
    float zp = (float)((double)z + fabs((double)bmin[2]));
    float x0 = ((float)j * aw) / (float)nx;
    float y0 = ((float)i * ah) / (float)ny;
    float z0 = ((float)k * ad) / (float)nz;
    float xd = (fabsf(xp - x0) * aw) / (float)nx;
    float yd = (fabsf(yp - y0) * ah) / (float)ny;
    float zd = (fabsf(zp - z0) * ad) / (float)nz;
    float c00 = v000 * (1.0f - xd) + v100 * xd;
    float c10 = v010 * (1.0f - xd) + v110 * xd;
    float c01 = v001 * (1.0f - xd) + v101 * xd;
    float c11 = v011 * (1.0f - xd) + v111 * xd;
    float c0 = c00 * (1.0f - yd) + c10 * yd;
    float c1 = c01 * (1.0f - yd) + c11 * yd;
    float c = c0 * (1.0f - zd) + c1 * zd;
    return c;
}

static inline float lerpf_(float a, float b, float t) { return fmaf(t, b - a, a); }

/* Config-mode sampler: the reference's index mapping (ray.c:126-135) with the
 * weights it intends, f = u - floor(u), each on its own axis, clamp-to-edge.
 * u = voxel coordinate (the caller's fmaf(dir*N/ext, t, (eye-min)*N/ext)),
 * clamped to [0, N-1]; j = min((int)u, N-2). */
static inline float tri_fetch(const float *vol, int nx, int ny, int nz, float gx, float gy,
                              float gz)
{
    gx = fminf(fmaxf(gx, 0.0f), (float)(nx - 1));
    gy = fminf(fmaxf(gy, 0.0f), (float)(ny - 1));
    gz = fminf(fmaxf(gz, 0.0f), (float)(nz - 1));
    int j = (int)gx, i = (int)gy, k = (int)gz;
    j = j > nx - 2 ? nx - 2 : j;
    i = i > ny - 2 ? ny - 2 : i;
    k = k > nz - 2 ? nz - 2 : k;
    float fx = gx - (float)j, fy = gy - (float)i, fz = gz - (float)k;
    const float *b = &VOX(vol, j, i, k, nx, ny);
    size_t sy = (size_t)nx, sz = (size_t)nx * (size_t)ny;
    float c00 = lerpf_(b[0], b[1], fx);
    float c10 = lerpf_(b[sy], b[sy + 1], fx);
    float c01 = lerpf_(b[sz], b[sz + 1], fx);
    float c11 = lerpf_(b[sz + sy], b[sz + sy + 1], fx);
    float c0 = lerpf_(c00, c10, fy);
    float c1 = lerpf_(c01, c11, fy);
    return lerpf_(c0, c1, fz);
}

float ivro_sample_tri(const float *vol, int nx, int ny, int nz, const float bmin[3],
                      const float bmax[3], float x, float y, float z)
{
    float sx = (float)nx / (bmax[0] - bmin[0]);
    float sy = (float)ny / (bmax[1] - bmin[1]);
    float sz = (float)nz / (bmax[2] - bmin[2]);
    return tri_fetch(vol, nx, ny, nz, (x - bmin[0]) * sx, (y - bmin[1]) * sy,
                     (z - bmin[2]) * sz);
}

static inline double tri_fetch64(const float *vol, int nx, int ny, int nz, double gx,
                                 double gy, double gz)
{
    gx = fmin(fmax(gx, 0.0), (double)(nx - 1));
    gy = fmin(fmax(gy, 0.0), (double)(ny - 1));
    gz = fmin(fmax(gz, 0.0), (double)(nz - 1));
    int j = (int)gx, i = (int)gy, k = (int)gz;
    j = j > nx - 2 ? nx - 2 : j;
    i = i > ny - 2 ? ny - 2 : i;
    k = k > nz - 2 ? nz - 2 : k;
    double fx = gx - j, fy = gy - i, fz = gz - k;
    const float *b = &VOX(vol, j, i, k, nx, ny);
    size_t sy = (size_t)nx, sz = (size_t)nx * (size_t)ny;
    double c00 = b[0] + fx * ((double)b[1] - b[0]);
    double c10 = b[sy] + fx * ((double)b[sy + 1] - b[sy]);
    double c01 = b[sz] + fx * ((double)b[sz + 1] - b[sz]);
    double c11 = b[sz + sy] + fx * ((double)b[sz + sy + 1] - b[sz + sy]);
    double c0 = c00 + fy * (c10 - c00);
    double c1 = c01 + fy * (c11 - c01);
    return c0 + fz * (c1 - c0);
}

double ivro_sample_tri64(const float *vol, int nx, int ny, int nz, const double bmin[3],
                         const double bmax[3], double x, double y, double z)
{
    return tri_fetch64(vol, nx, ny, nz, (x - bmin[0]) * (nx / (bmax[0] - bmin[0])),
                       (y - bmin[1]) * (ny / (bmax[1] - bmin[1])),
                       (z - bmin[2]) * (nz / (bmax[2] - bmin[2])));
}

/* ------------------------------------------------------------------ */
/* transfer functions (transferfuncs.c:24-199)                        */
/* ------------------------------------------------------------------ */

static col4 tf_apply(int xfer, const col4 *lut, int lutoffset, int maxlutindx,
                     const float *s, int n, float myalpha)
{
    col4 c = {0.0f, 0.0f, 0.0f, 0.0f};
    switch (xfer) {
    case 1: /* MEDXFER (transferfuncs.c:24-68) */
        for (int i = 0; i < n; ++i) {
            int idx = (int)s[i] - lutoffset;
            c.b = (lut[idx].b < 0.8) ? (float)fmax(lut[idx].b, c.b) : c.b;
            c.g = (lut[idx].g < 0.8) ? (float)fmax(lut[idx].g, c.g) : c.g;
            c.r = (lut[idx].r < 0.8) ? (float)fmax(lut[idx].r, c.r) : c.r;
            c.a = lut[idx].a * myalpha;
        }
        break;
    case 4: /* BLEND (transferfuncs.c:70-84): no lutoffset */
        for (int i = 0; i < n; ++i) {
            int idx = (int)s[i];
            float al = lut[idx].a;
            c.b = al * lut[idx].b + (1.0f - al) * c.b;
            c.g = al * lut[idx].g + (1.0f - al) * c.g;
            c.r = al * lut[idx].r + (1.0f - al) * c.r;
            c.a = al * myalpha;
        }
        break;
    case 2: /* SEISMIC2 (transferfuncs.c:109-136) */
    case 3: /* CHEM (transferfuncs.c:137-168) */
        for (int i = n - 1; i >= 0; i--) {
            int idx = (int)s[i] - lutoffset;
            if (idx < 0) idx = 0;
            if (xfer == 2 && idx > maxlutindx) idx = maxlutindx;
            const col4 *src = &lut[idx];
            float al = src->a * myalpha;
            float om = 1.0f - al;
            c.b = om * c.b + al * src->b;
            c.g = om * c.g + al * src->g;
            c.r = om * c.r + al * src->r;
            c.a = al + om * c.a;
            if (c.a > 1.0f) c.a = 1.0f;
        }
        break;
    case 5: /* RGBBLEND (transferfuncs.c:170-199): colour packed in the float bits */
        for (int i = n - 1; i >= 0; i--) {
            uint32_t u;
            memcpy(&u, &s[i], 4);
            float r = 1.0f / 255.0f * (float)((u >> 24) & 0xff);
            float g = 1.0f / 255.0f * (float)((u >> 16) & 0xff);
            float b = 1.0f / 255.0f * (float)((u >> 8) & 0xff);
            float al = myalpha;
            float om = 1.0f - al;
            c.b = om * c.b + al * b;
            c.g = om * c.g + al * g;
            c.r = om * c.r + al * r;
            c.a = al + om * c.a;
            if (c.a > 1.0f) c.a = 1.0f;
        }
        break;
    default: /* SEISMIC (transferfuncs.c:86-107); also the default id */
        for (int i = 0; i < n; ++i) {
            int idx = (int)s[i] - lutoffset;
            const col4 *l = &lut[idx];
            c.b = ((l->b > 0.65) && (l->b < 1.0f)) ? 0.6f : c.b;
            c.g = ((l->g > 0.65) && (l->g < 1.0f)) ? 0.0f : c.g;
            c.r = ((l->r > 0.65) && (l->r < 1.0f)) ? 0.0f : c.r;
            c.b = ((l->b > 0.5) && (l->b < 0.65f)) ? 0.9f : c.b;
            c.g = ((l->g > 0.5) && (l->g < 0.65f)) ? 0.9f : c.g;
            c.r = ((l->r > 0.5) && (l->r < 0.65f)) ? 0.9f : c.r;
            c.b = ((l->b > 0.3) && (l->b < 0.5f)) ? 0.0f : c.b;
            c.g = ((l->g > 0.3) && (l->g < 0.5f)) ? 0.0f : c.g;
            c.r = ((l->r > 0.3) && (l->r < 0.5f)) ? 0.6f : c.r;
            c.a = 1.0f * myalpha;
        }
        break;
    }
    return c;
}

void ivro_transferfunc(int xfer, const float *lut, int lutoffset, int maxlutindx,
                       const float *samples, int n, float alpha, float out[4])
{
    col4 c = tf_apply(xfer, (const col4 *)lut, lutoffset, maxlutindx, samples, n, alpha);
    out[0] = c.r; out[1] = c.g; out[2] = c.b; out[3] = c.a;
}

/* ------------------------------------------------------------------ */
/* colormaps (colormap.c:21-181); out is zero-initialised first, which */
/* is what the reference's entries that a map leaves unset hold when   */
/* the LUT comes from calloc                                           */
/* ------------------------------------------------------------------ */

static void lerp_map(float *out, const float *v0, const float *v1, int maplen)
{
    /* lerp (mathutil.h:42-59) */
    float incr1 = (v1[1] - v1[0]) / (v0[1] - v0[0]);
    float incr2 = (v1[2] - v1[1]) / (v0[2] - v0[1]);
    for (int i = 1; i <= maplen; i++) {
        if ((i >= v0[0]) && (i <= v0[1])) out[i - 1] = v1[0] + i * incr1;
        else if ((i >= v0[1]) && (i <= v0[2])) out[i - 1] = v1[1] + (i - v0[1]) * incr2;
    }
}

void ivro_colormap(int id, int hsize, float hmin, float hmax, float *out)
{
    col4 *l = (col4 *)out;
    memset(out, 0, sizeof(col4) * (size_t)hsize);
    if (id == 0) { /* BASE */
        for (int i = 0; i < hsize; i++) {
            float c = (float)(i + hmin) / hmax;
            l[i].b = l[i].g = l[i].r = c;
            if (c > 0.8) { l[i].r = 0.7f; l[i].g = 0.75f; l[i].b = 0.46; l[i].a = 0.07; }
            if (c < 0.3) { l[i].b = 0.6; l[i].a = 0.01; }
            if (c > 0.3 && c < 0.6) { l[i].r = 0.6; l[i].a = 0.01; }
            if (c > 0.6 && c < 0.8) { l[i].r = 0.6; l[i].g = 0.6; l[i].a = 0.005; }
        }
    } else if (id == 1) { /* RAINBOW: per-entry increments of an unset LUT */
        l[0].b = 1.0f; l[0].g = 0.0f; l[0].r = 0.0f;
        float dr = 1.0f / hsize, db = -dr;
        for (int i = 0; i < hsize; i++) { l[i].r += dr; l[i].b += db; l[i].a = 1.0f; }
    } else if (id == 2) { /* MEDICAL: tissue classes over RANGE=4096 */
        float a, x, y1, y2;
        const float parm[4] = {0.04, 0.9, 0.9, 0.8};
        for (int i = 0; i < hsize; i++) {
            x = i / (float)(hsize - 1) * 4096.0f;
            if ((0 <= x) && (x <= 484)) {
                l[i].r = 0.1; l[i].g = 0.1; l[i].b = 0.1; l[i].a = 0.1 * parm[0];
            } else if ((484 < x) && (x <= 545)) {
                a = (x - 484.0) / (545.0 - 484.0);
                y1 = (1.0 - a) * parm[0];
                y2 = a * parm[1];
                l[i].r = (0.6 * y1 + 0.8 * y2);
                l[i].g = (0.6 * y1 + 0.7 * y2);
                l[i].b = (0.8 * y1 + 0.0 * y2);
                l[i].a = (0.1 * y1 + 0.03125 * y2);
            } else if ((545 < x) && (x <= 726)) {
                l[i].r = 0.8; l[i].g = 0.7; l[i].b = 0.0; l[i].a = 0.03125 * parm[1];
            } else if ((726 < x) && (x <= 819)) {
                a = (x - 726.0) / (819.0 - 726.0);
                y1 = (1.0 - a) * parm[1];
                y2 = a * parm[2];
                l[i].r = (0.8 * y1 + 0.2 * y2);
                l[i].g = (0.7 * y1 + 0.0 * y2);
                l[i].b = (0.0 * y1 + 0.0 * y2);
                l[i].a = (0.03125 * y1 + 0.1875 * y2);
            } else if ((819 < x) && (x <= 883)) {
                l[i].r = 0.2; l[i].g = 0.0; l[i].b = 0.0; l[i].a = 0.1875 * parm[2];
            } else if ((883 < x) && (x <= 1535)) {
                a = (x - 883.0) / (1535.0 - 883.0);
                y1 = (1.0 - a) * parm[2];
                y2 = a * parm[3];
                l[i].r = (0.2 * y1 + 1.0 * y2);
                l[i].g = (0.0 * y1 + 1.0 * y2);
                l[i].b = (0.0 * y1 + 0.9 * y2);
                l[i].a = (0.1875 * y1 + 0.75 * y2);
            } else if (1535 < x) {
                l[i].r = 1.0; l[i].g = 1.0; l[i].b = 0.9; l[i].a = 0.75 * parm[3];
            }
        }
    } else if (id == 3) { /* SEISMIC: two-segment ramps */
        float er[3] = {1.0f, 1.0f, 0.0f}, eg[3] = {0.0f, 1.0f, 0.0f}, eb[3] = {0.0f, 1.0f, 1.0f};
        float sr[3] = {1.0f, 0.5f * hsize, (float)hsize};
        float *r = (float *)calloc((size_t)hsize, sizeof(float));
        float *g = (float *)calloc((size_t)hsize, sizeof(float));
        float *b = (float *)calloc((size_t)hsize, sizeof(float));
        lerp_map(r, sr, er, hsize);
        lerp_map(g, sr, eg, hsize);
        lerp_map(b, sr, eb, hsize);
        for (int i = 0; i < hsize; i++) { l[i].r = r[i]; l[i].g = g[i]; l[i].b = b[i]; l[i].a = 1.0f; }
        free(r); free(g); free(b);
    } else if (id == 4) { /* JET */
        for (int i = 0; i < hsize; i++) {
            float i4 = (4.0f * i) / hsize;
            l[i].r = (float)fmin(fmax(fmin(i4 - 1.5, -i4 + 4.5), 0.0f), 1.0f);
            l[i].g = (float)fmin(fmax(fmin(i4 - 0.5, -i4 + 3.5), 0.0f), 1.0f);
            l[i].b = (float)fmin(fmax(fmin(i4 + 0.5, -i4 + 2.5), 0.0f), 1.0f);
            l[i].a = 1.0f;
        }
    }
    /* NOMAP (-1) leaves the LUT untouched */
}

/* ------------------------------------------------------------------ */
/* camera                                                             */
/* ------------------------------------------------------------------ */

/* initcamera (main.c:428-500): n=(0,0,-1), v=(0,1,0), u=n x v, e=(0,0,-5),
 * vwm = [u v n] as columns (setvwm, mathutil.c:382-388). */
static void view_to_world(float vwm[3][3], float e[3])
{
    float n[3] = {0.0f, 0.0f, -1.0f}, v[3] = {0.0f, 1.0f, 0.0f}, u[3];
    float mag = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] = v[0] / mag; v[1] = v[1] / mag; v[2] = v[2] / mag;
    u[0] = n[1] * v[2] - n[2] * v[1];
    u[1] = n[2] * v[0] - n[0] * v[2];
    u[2] = n[0] * v[1] - n[1] * v[0];
    for (int r = 0; r < 3; r++) { vwm[r][0] = u[r]; vwm[r][1] = v[r]; vwm[r][2] = n[r]; }
    e[0] = 0.0f; e[1] = 0.0f; e[2] = -5.0f;
}

void ivro_camera(const float bmin[3], const float bmax[3], const ivro_view *vw,
                 float cam[23], int rect[4])
{
    float mvm[4][4], pm[4][4];
    from_gl(vw->mvm, mvm);
    from_gl(vw->pm, pm);
    float W = (float)vw->width, H = (float)vw->height;
    float l = vw->vv[0], r = vw->vv[1], b = vw->vv[2], t = vw->vv[3];

    /* updatevolbbox (mathutil.c:948-1018) with vpm from setviewport
     * (mathutil.h:147-163) at x=y=0 */
    float vpm[4][4] = {{vw->width / 2.0f, 0.0f, 0.0f, (0 + vw->width) / 2.0f},
                       {0.0f, vw->height / 2.0f, 0.0f, (0 + vw->height) / 2.0f},
                       {0.0f, 0.0f, 0.5f, 0.5f},
                       {0.0f, 0.0f, 0.0f, 1.0f}};
    float ctm4[4][4];
    memcpy(ctm4, pm, sizeof ctm4);
    m4m(ctm4, mvm);
    float wminx = FLT_MAX, wminy = FLT_MAX, wmaxx = -FLT_MAX, wmaxy = -FLT_MAX;
    for (int c = 0; c < 8; c++) {
        float v4[4] = {(c & 1) ? bmax[0] : bmin[0], (c & 2) ? bmax[1] : bmin[1],
                       (c & 4) ? bmax[2] : bmin[2], 1.0f};
        float pc[4], win[4];
        m4v(ctm4, v4, pc);
        float w = pc[3];
        float ndc[4] = {pc[0] / w, pc[1] / w, pc[2] / w, 1.0f};
        m4v(vpm, ndc, win);
        wminx = (float)fmin(wminx, win[0]);
        wminy = (float)fmin(wminy, win[1]);
        wmaxx = (float)fmax(wmaxx, win[0]);
        wmaxy = (float)fmax(wmaxy, win[1]);
    }
    float vx = 0.0f, vy = 0.0f;
    wminx = (wminx < vx) ? vx : wminx;
    wmaxx = (wmaxx > vx + W) ? vx + W : wmaxx;
    wminx = (wminx > vx + W) ? vx + W : wminx;
    wmaxx = (wmaxx < vx) ? vx : wmaxx;
    wminy = (wminy < vy) ? vy : wminy;
    wmaxy = (wmaxy > vy + H) ? vy + H : wmaxy;
    wminy = (wminy > vy + H) ? vy + H : wminy;
    wmaxy = (wmaxy < vy) ? vy : wmaxy;
    float wcx = wminx, wcy = wminy, wcw = wmaxx - wminx, wch = wmaxy - wminy;
    rect[0] = (int)(wcy - 0.5f);
    rect[1] = (int)(wcy + wch + 1.0f);
    rect[2] = (int)(wcx - 0.5f);
    rect[3] = (int)(wcx + wcw + 1.0f);

    /* createlocalimage (main.c:950-981) */
    float du = (r - l) / W;
    float dv = (t - b) / H;
    float v = b + rect[0] * dv + dv / 2.0f;
    float u = l + rect[2] * du - du / 2.0f;

    /* buildray (mathutil.c:1115-1154) */
    float vwm[3][3], e[3];
    view_to_world(vwm, e);
    float dir[3] = {u, v, -e[2]};
    float ii[3] = {0.0f, dv, 0.0f}, jj[3] = {du, 0.0f, 0.0f};
    float trans[3] = {mvm[0][3], mvm[1][3], 0.0f};
    float eye[3];
    m3v(vwm, e, eye);
    eye[0] = eye[0] + 0.0f; eye[1] = eye[1] + 0.0f; eye[2] = eye[2] + 0.0f;
    float o[3] = {eye[0] - trans[0], eye[1] - trans[1], eye[2] - trans[2]};
    float ctm[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) ctm[i][j] = mvm[j][i];
    m3v(ctm, o, o);
    m3m(ctm, vwm);
    m3v(ctm, dir, dir);
    m3v(ctm, ii, ii);
    m3v(ctm, jj, jj);

    cam[0] = o[0]; cam[1] = o[1]; cam[2] = o[2];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) cam[3 + i * 3 + j] = ctm[i][j];
    cam[12] = du; cam[13] = dv;
    cam[14] = dir[0]; cam[15] = dir[1]; cam[16] = dir[2];
    cam[17] = ii[0]; cam[18] = ii[1]; cam[19] = ii[2];
    cam[20] = jj[0]; cam[21] = jj[1]; cam[22] = jj[2];
}

/* ------------------------------------------------------------------ */
/* per-ray renderers                                                  */
/* ------------------------------------------------------------------ */

typedef struct {
    const float *vol;
    int nx, ny, nz;
    float bmin[3], bmax[3], dconst[6];
    const col4 *lut;
    int lutsize;
    const ivro_params *p;
    float alpha;
} rctx;

/* raycast (ray.c:59-85) + getsamplesonray (ray.c:239-273), compat mode */
static col4 ray_compat(const rctx *c, const float o[3], const float d[3], float *buf, int *ns)
{
    col4 z = {0.0f, 0.0f, 0.0f, 0.0f};
    float tin, tout;
    *ns = 0;
    if (!slab(c->dconst, o, d, &tin, &tout)) return z;
    const float dt = 0.001f; /* DELTASAMPLE, types.h:17 */
    int iend = 1.0f + (tout - tin) / dt + 0.5f;
    iend = (iend < 256) ? iend : 256; /* MAXVOLUMESAMPLES*4, ray.c:61 */
    int i;
    for (i = 0; i < iend; i++) {
        float t = tin + (i * dt);
        float x = d[0] * t + o[0];
        float y = d[1] * t + o[1];
        float zz = d[2] * t + o[2];
        buf[i] = ivro_sample_ref(c->vol, c->nx, c->ny, c->nz, c->bmin, c->bmax, x, y, zz);
    }
    *ns = i > 0 ? i : 0;
    if (i <= 0) return z;
    return tf_apply(c->p->xfer, c->lut, c->p->lutoffset, c->p->maxlutindx, buf, i, c->alpha);
}

int ivro_samplesonray_compat(const float *vol, int nx, int ny, int nz, const float newconst[6],
                             const float ray[6], int maxs, float *samples)
{
    float bmin[3], bmax[3], dc[6], tin, tout;
    ivro_box(nx, ny, nz, bmin, bmax);
    plane_consts(bmin, bmax, newconst, dc);
    if (!slab(dc, ray, ray + 3, &tin, &tout)) return -1;
    const float dt = 0.001f;
    int iend = 1.0f + (tout - tin) / dt + 0.5f;
    iend = (iend < maxs) ? iend : maxs;
    int i;
    for (i = 0; i < iend; i++) {
        float t = tin + (i * dt);
        samples[i] = ivro_sample_ref(vol, nx, ny, nz, bmin, bmax, ray[3] * t + ray[0],
                                     ray[4] * t + ray[1], ray[5] * t + ray[2]);
    }
    return i;
}

/* Config mode: global sample grid t_k = tin + k*dt, dt = step voxels / |dir|,
 * n = (int)(1 + (tout - tin)/dt + 0.5) (the count rule of ray.c:253),
 * front-to-back C += (1-A) a c, A += (1-A) a, stop at A >= ert.  In tile mode
 * only samples whose voxel index lies in brick `brick_id` are composited. */
static col4 ray_config(const rctx *c, const float o[3], const float d[3], long long *ns)
{
    col4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    float tin, tout;
    if (!slab(c->dconst, o, d, &tin, &tout)) return acc;
    tin = fmaxf(tin, 0.0f);
    const ivro_params *p = c->p;
    float vox = (c->bmax[0] - c->bmin[0]) / (float)c->nx;
    float dlen = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    float dt = (vox * p->step_vox) / dlen;
    int n = (int)(1.0f + (tout - tin) / dt + 0.5f);
    float sx = (float)c->nx / (c->bmax[0] - c->bmin[0]);
    float sy = (float)c->ny / (c->bmax[1] - c->bmin[1]);
    float sz = (float)c->nz / (c->bmax[2] - c->bmin[2]);
    /* the ray in voxel coordinates: g(t) = gd t + go */
    const float gox = (o[0] - c->bmin[0]) * sx, goy = (o[1] - c->bmin[1]) * sy, goz = (o[2] - c->bmin[2]) * sz;
    const float gdx = d[0] * sx, gdy = d[1] * sy, gdz = d[2] * sz;
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    long long cnt = 0;
    for (int k = 0; k < n; k++) {
        float t = fmaf((float)k, dt, tin);
        float gx = fmaf(gdx, t, gox);
        float gy = fmaf(gdy, t, goy);
        float gz = fmaf(gdz, t, goz);
        if (p->brick > 0) {
            float cx = fminf(fmaxf(gx, 0.0f), (float)(nx - 1));
            float cy = fminf(fmaxf(gy, 0.0f), (float)(ny - 1));
            float cz = fminf(fmaxf(gz, 0.0f), (float)(nz - 1));
            int j = (int)cx, i = (int)cy, kk = (int)cz;
            j = j > nx - 2 ? nx - 2 : j;
            i = i > ny - 2 ? ny - 2 : i;
            kk = kk > nz - 2 ? nz - 2 : kk;
            if (j / p->brick != p->brick_id[0] || i / p->brick != p->brick_id[1] ||
                kk / p->brick != p->brick_id[2])
                continue;
        }
        float s = tri_fetch(c->vol, nx, ny, nz, gx, gy, gz);
        int idx = (int)(s * p->lutscale) - p->lutoffset;
        idx = idx < 0 ? 0 : (idx > c->lutsize - 1 ? c->lutsize - 1 : idx);
        const col4 *e = &c->lut[idx];
        float a = e->a * c->alpha;
        float w = (1.0f - acc.a) * a;
        acc.r = fmaf(w, e->r, acc.r);
        acc.g = fmaf(w, e->g, acc.g);
        acc.b = fmaf(w, e->b, acc.b);
        acc.a = acc.a + w;
        cnt++;
        if (acc.a >= p->ert) break;
    }
    *ns += cnt;
    return acc;
}

/* Sort-last in one pass per ray: the samples a brick owns (floor index in
 * the brick, as in tile mode) form one contiguous run of k, since every axis
 * index is monotone in k, and the runs come in visibility order.  Each run is
 * composited from zero with ERT inside it (the remaining samples of a
 * saturated run are passed over without a fetch), and the partials are
 * combined far to near with compositetiles' d = s + (1 - s.a) d
 * (main.c:1206-1215): the same image as the per-brick tiles composited in
 * visibility order, where the bricks a ray misses add exact zeros. */
#define IVRO_MAX_RUNS 64
static col4 ray_sort_last(const rctx *c, const float o[3], const float d[3], long long *ns)
{
    col4 out = {0.0f, 0.0f, 0.0f, 0.0f};
    float tin, tout;
    if (!slab(c->dconst, o, d, &tin, &tout)) return out;
    tin = fmaxf(tin, 0.0f);
    const ivro_params *p = c->p;
    float vox = (c->bmax[0] - c->bmin[0]) / (float)c->nx;
    float dlen = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    float dt = (vox * p->step_vox) / dlen;
    int n = (int)(1.0f + (tout - tin) / dt + 0.5f);
    float sx = (float)c->nx / (c->bmax[0] - c->bmin[0]);
    float sy = (float)c->ny / (c->bmax[1] - c->bmin[1]);
    float sz = (float)c->nz / (c->bmax[2] - c->bmin[2]);
    const float gox = (o[0] - c->bmin[0]) * sx, goy = (o[1] - c->bmin[1]) * sy, goz = (o[2] - c->bmin[2]) * sz;
    const float gdx = d[0] * sx, gdy = d[1] * sy, gdz = d[2] * sz;
    const int nx = c->nx, ny = c->ny, nz = c->nz, B = p->brick;
    col4 part[IVRO_MAX_RUNS];
    int nrun = 0, cur = -1, done = 0;
    col4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    long long cnt = 0;
    for (int k = 0; k < n; k++) {
        float t = fmaf((float)k, dt, tin);
        float gx = fmaf(gdx, t, gox);
        float gy = fmaf(gdy, t, goy);
        float gz = fmaf(gdz, t, goz);
        float cx = fminf(fmaxf(gx, 0.0f), (float)(nx - 1));
        float cy = fminf(fmaxf(gy, 0.0f), (float)(ny - 1));
        float cz = fminf(fmaxf(gz, 0.0f), (float)(nz - 1));
        int j = (int)cx, i = (int)cy, kk = (int)cz;
        j = j > nx - 2 ? nx - 2 : j;
        i = i > ny - 2 ? ny - 2 : i;
        kk = kk > nz - 2 ? nz - 2 : kk;
        const int b = j / B + (nx / B + 1) * (i / B + (ny / B + 1) * (kk / B));
        if (b != cur) { /* first sample of the next run */
            if (cur >= 0 && nrun < IVRO_MAX_RUNS) part[nrun++] = acc;
            acc.r = acc.g = acc.b = acc.a = 0.0f;
            cur = b;
            done = 0;
        }
        if (done) continue;
        float s = tri_fetch(c->vol, nx, ny, nz, gx, gy, gz);
        int idx = (int)(s * p->lutscale) - p->lutoffset;
        idx = idx < 0 ? 0 : (idx > c->lutsize - 1 ? c->lutsize - 1 : idx);
        const col4 *e = &c->lut[idx];
        float a = e->a * c->alpha;
        float w = (1.0f - acc.a) * a;
        acc.r = fmaf(w, e->r, acc.r);
        acc.g = fmaf(w, e->g, acc.g);
        acc.b = fmaf(w, e->b, acc.b);
        acc.a = acc.a + w;
        cnt++;
        if (acc.a >= p->ert) done = 1;
    }
    if (cur >= 0 && nrun < IVRO_MAX_RUNS) part[nrun++] = acc;
    for (int q = nrun - 1; q >= 0; q--) {
        const col4 *s = &part[q];
        out.b = s->b + (1.0f - s->a) * out.b;
        out.g = s->g + (1.0f - s->a) * out.g;
        out.r = s->r + (1.0f - s->a) * out.r;
        out.a = s->a + (1.0f - s->a) * out.a;
    }
    *ns += cnt;
    return out;
}

static void setup_ctx(rctx *c, const float *vol, int nx, int ny, int nz, const ivro_view *v,
                      const float *lut, int lutsize, const ivro_params *p)
{
    c->vol = vol;
    c->nx = nx; c->ny = ny; c->nz = nz;
    ivro_box(nx, ny, nz, c->bmin, c->bmax);
    plane_consts(c->bmin, c->bmax, v->newconst, c->dconst);
    c->lut = (const col4 *)lut;
    c->lutsize = lutsize;
    c->p = p;
    c->alpha = v->alpha;
}

long long ivro_render(const float *vol, int nx, int ny, int nz, const ivro_view *v,
                      const float *lut, int lutsize, const ivro_params *p, float *rgba,
                      int row0, int row1, int nthreads)
{
    rctx c;
    setup_ctx(&c, vol, nx, ny, nz, v, lut, lutsize, p);
    float cam[23];
    int rect[4];
    ivro_camera(c.bmin, c.bmax, v, cam, rect);
    const int W = v->width, H = v->height;
    long long total = 0;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    if (row1 > H) row1 = H;
    if (p->mode == IVRO_COMPAT) {
        const int istrt = rect[0], istop = rect[1], jstrt = rect[2], jstop = rect[3];
        const float *d0 = cam + 14, *ii = cam + 17, *jj = cam + 20;
        /* main.c:983-1009: static OpenMP over rows, columns accumulated */
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+ : total)
        for (int i = istrt; i < istop; i++) {
            float buf[256];
            float d[3];
            d[0] = d0[0] + ii[0] * (i - istrt);
            d[1] = d0[1] + ii[1] * (i - istrt);
            d[2] = d0[2] + ii[2] * (i - istrt);
            for (int j = jstrt; j < jstop; j++) {
                d[0] += jj[0];
                d[1] += jj[1];
                d[2] += jj[2];
                if (i < row0 || i >= row1 || i < 0 || j < 0 || j >= W) continue;
                int ns;
                col4 col = ray_compat(&c, cam, d, buf, &ns);
                total += ns;
                float *px = rgba + ((size_t)i * W + j) * 4;
                px[0] = col.r; px[1] = col.g; px[2] = col.b; px[3] = col.a;
            }
        }
        return total;
    }
    const float du = cam[12], dv = cam[13];
    const float l = v->vv[0], b = v->vv[2];
    const float(*M)[3] = (const float(*)[3])(cam + 3);
    const float m5[3] = {M[0][2] * 5.0f, M[1][2] * 5.0f, M[2][2] * 5.0f};
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+ : total)
    for (int i = row0; i < row1; i++) {
        float vv = fmaf((float)i + 0.5f, dv, b);
        for (int j = 0; j < W; j++) {
            float uu = fmaf((float)j + 0.5f, du, l);
            float d[3];
            for (int r = 0; r < 3; r++) d[r] = fmaf(M[r][0], uu, fmaf(M[r][1], vv, m5[r]));
            long long ns = 0;
            col4 col = (p->brick > 0 && p->sort_last) ? ray_sort_last(&c, cam, d, &ns) : ray_config(&c, cam, d, &ns);
            total += ns;
            float *px = rgba + ((size_t)i * W + j) * 4;
            px[0] = col.r; px[1] = col.g; px[2] = col.b; px[3] = col.a;
        }
    }
    return total;
}

/* ------------------------------------------------------------------ */
/* FP64 config mode                                                   */
/* ------------------------------------------------------------------ */

long long ivro_render_f64(const float *vol, int nx, int ny, int nz, const ivro_view *v,
                          const float *lut, int lutsize, const ivro_params *p, double *rgba,
                          int row0, int row1, int nthreads)
{
    float bminf[3], bmaxf[3];
    ivro_box(nx, ny, nz, bminf, bmaxf);
    double bmin[3], bmax[3], dconst[6];
    for (int a = 0; a < 3; a++) { bmin[a] = bminf[a]; bmax[a] = bmaxf[a]; }
    double pz[6] = {-bmin[0], bmax[0], -bmin[1], bmax[1], -bmin[2], bmax[2]};
    for (int i = 0; i < 6; i++) dconst[i] = (v->newconst[i] < pz[i]) ? v->newconst[i] : pz[i];
    /* camera in double: eye (0,0,5) minus pan, rotated by mvm^T; view
     * directions (u, v, 5) -> (u, v, -5) -> mvm^T */
    double m[4][4];
    for (int r = 0; r < 4; r++)
        for (int cc = 0; cc < 4; cc++) m[r][cc] = v->mvm[cc * 4 + r];
    double eo[3] = {0.0 - m[0][3], 0.0 - m[1][3], 5.0};
    double o[3], M[3][3];
    for (int r = 0; r < 3; r++) {
        o[r] = m[0][r] * eo[0] + m[1][r] * eo[1] + m[2][r] * eo[2];
        M[r][0] = m[0][r];
        M[r][1] = m[1][r];
        M[r][2] = -m[2][r];
    }
    const int W = v->width, H = v->height;
    double du = ((double)v->vv[1] - v->vv[0]) / W, dv = ((double)v->vv[3] - v->vv[2]) / H;
    double vox = (bmax[0] - bmin[0]) / nx;
    double sx = nx / (bmax[0] - bmin[0]), sy = ny / (bmax[1] - bmin[1]),
           sz = nz / (bmax[2] - bmin[2]);
    const col4 *L = (const col4 *)lut;
    long long total = 0;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    if (row1 > H) row1 = H;
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+ : total)
    for (int i = row0; i < row1; i++) {
        double vv = v->vv[2] + (i + 0.5) * dv;
        for (int j = 0; j < W; j++) {
            double uu = v->vv[0] + (j + 0.5) * du;
            double d[3];
            for (int r = 0; r < 3; r++) d[r] = M[r][0] * uu + M[r][1] * vv + M[r][2] * 5.0;
            double ti = -DBL_MAX, to = DBL_MAX;
            for (int f = 0; f < 6 && ti < to; f++) {
                int ax = f >> 1;
                double sgn = (f & 1) ? 1.0 : -1.0;
                double numer = dconst[f] - sgn * o[ax];
                double denom = sgn * d[ax];
                if (denom == 0.0) {
                    if (numer < 0.0) ti = to + 1.0;
                } else {
                    double th = numer / denom;
                    if (denom > 0.0) to = fmin(to, th);
                    else ti = fmax(ti, th);
                }
            }
            double acc[4] = {0, 0, 0, 0};
            if (ti < to && (ti > 0.0 || to > 0.0)) {
                ti = fmax(ti, 0.0);
                double dlen = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                double dt = vox * p->step_vox / dlen;
                int n = (int)(1.0 + (to - ti) / dt + 0.5);
                for (int k = 0; k < n; k++) {
                    double t = ti + k * dt;
                    double gx = (o[0] + d[0] * t - bmin[0]) * sx;
                    double gy = (o[1] + d[1] * t - bmin[1]) * sy;
                    double gz = (o[2] + d[2] * t - bmin[2]) * sz;
                    if (p->brick > 0) { /* tile mode: the brick's owned samples only */
                        int j = (int)fmin(fmax(gx, 0.0), nx - 1.0), i = (int)fmin(fmax(gy, 0.0), ny - 1.0),
                            kk = (int)fmin(fmax(gz, 0.0), nz - 1.0);
                        j = j > nx - 2 ? nx - 2 : j;
                        i = i > ny - 2 ? ny - 2 : i;
                        kk = kk > nz - 2 ? nz - 2 : kk;
                        if (j / p->brick != p->brick_id[0] || i / p->brick != p->brick_id[1] ||
                            kk / p->brick != p->brick_id[2])
                            continue;
                    }
                    double s = tri_fetch64(vol, nx, ny, nz, gx, gy, gz);
                    int idx = (int)(s * p->lutscale) - p->lutoffset;
                    idx = idx < 0 ? 0 : (idx > lutsize - 1 ? lutsize - 1 : idx);
                    double a = (double)L[idx].a * v->alpha;
                    double w = (1.0 - acc[3]) * a;
                    acc[0] += w * L[idx].r;
                    acc[1] += w * L[idx].g;
                    acc[2] += w * L[idx].b;
                    acc[3] += w;
                    total++;
                    if (acc[3] >= p->ert) break;
                }
            }
            double *px = rgba + ((size_t)i * W + j) * 4;
            px[0] = acc[0]; px[1] = acc[1]; px[2] = acc[2]; px[3] = acc[3];
        }
    }
    return total;
}

/* ------------------------------------------------------------------ */
/* sort-last tile compositing and quantisation                         */
/* ------------------------------------------------------------------ */

void ivro_composite_tiles(float *dst, const float *tiles, int num, int npix)
{
    for (int k = num - 1; k >= 0; k--) {
        col4 *d = (col4 *)dst;
        const col4 *s = (const col4 *)(tiles + (size_t)k * npix * 4);
        for (int i = 0; i < npix; i++, d++, s++) {
            d->b = s->b + (1.0f - s->a) * d->b;
            d->g = s->g + (1.0f - s->a) * d->g;
            d->r = s->r + (1.0f - s->a) * d->r;
            d->a = s->a + (1.0f - s->a) * d->a;
        }
    }
}

void ivro_quantize(const float *rgba, long long n, unsigned char *out, int clamp)
{
    for (long long i = 0; i < n; i++) {
        float f = rgba[i];
        if (clamp) f = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
        out[i] = (unsigned char)(f * 255.0 + 0.5);
    }
}
