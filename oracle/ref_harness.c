/*
 * ref_harness.c -- TEST INFRASTRUCTURE ONLY (never linked into the product).
 *
 * Drives the reference renderer, compiled from its own C sources under
 * /root/reference by oracle/Makefile, headless.  The reference has no test
 * driver: its main() only initialises unless built with MPI or LOCAL
 * (main.c:171-258), so this file replays the per-frame sequence of the
 * reference server loop (main.c:226-233) and exposes the per-ray, per-sample,
 * transfer-function and colormap entry points through plain C functions that
 * the golden-vector script (tests/golden/make_golden.py) calls with ctypes.
 *
 * Built only into oracle/_ref/libivrref.so.  Nothing here is shipped.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "types.h"
#include "connection.h"
#include "mathutil.h"
#include "transferfuncs.h"
#include "colormap.h"

/* reference globals (main.c:63-105, ray.c:38) */
extern VOLUME volume;
extern AABB aabb;
extern STATE mystate;
extern color4f *lut;
extern color4f *colorbuf;
extern int lutoffset;
extern int maxlutindx;
extern int winwidth, winheight;
extern float pzero[6];
extern float dconst[6];

/* reference functions (main.c, ray.c, mathutil.c) */
int init(CMDARGS, int);
void initcolbufs(int);
void createlocalimage(int);
void setplaneconst(float *, float *);
int compositetiles(float *, float *, int, int, int);
void raycast(ray3f, color4f *);
int getsamplesonray(ray3f, int, float *);
boolean_t getrayintersect(ray3f, float *, float *);
float getlinearsample(float, float, float);
float getnearestsample(float, float, float);
void initsamplefunc(int);
int getfirstvoxel(ray3f, vector3i *);
void getvoxelsonray(ray3f, int, float *);

/* present when linked against the drop-in libivr_compat.so */
extern int ivr_compat_createlocalimage(int) __attribute__((weak));

static color4f *g_user_lut = NULL;
static int g_frame_gpu = 0;

/* 1: render frames through ivr_compat_createlocalimage (one GPU launch)
 * instead of the reference createlocalimage loop; returns whether available */
int ivrref_frame_on_gpu(int on)
{
    g_frame_gpu = on && ivr_compat_createlocalimage != NULL;
    return g_frame_gpu;
}

/* OpenMP threads of the reference's parallel loops (createlocalimage,
 * main.c:983) for the following calls; returns the count in effect */
int ivrref_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

/* init(): volume file of nx*ny*nz little-endian float32, x fastest. */
int ivrref_init(const char *volfile, int nx, int ny, int nz, int cmap, int xfer,
                int tile)
{
    CMDARGS a;
    memset(&a, 0, sizeof(a));
    strncpy(a.filename, volfile, sizeof(a.filename) - 1);
    a.width = nx;
    a.height = ny;
    a.depth = nz;
    a.dataendian = 1; /* 1 = little endian, sysutil.h:110-115 */
    a.colormap = cmap;
    a.xfer = xfer;
    a.procdims[0] = a.procdims[1] = a.procdims[2] = 1;
    a.tilewid = a.tilehgt = tile;
    return init(a, 0);
}

/* Replace the histogram LUT by a caller LUT (n entries of RGBA). */
void ivrref_set_lut(const float *rgba, int n, int offset, int maxidx)
{
    free(g_user_lut);
    g_user_lut = (color4f *)malloc(sizeof(color4f) * (size_t)n);
    memcpy(g_user_lut, rgba, sizeof(color4f) * (size_t)n);
    lut = g_user_lut;
    lutoffset = offset;
    maxlutindx = maxidx;
}

/* select transfer function `xfer` for following renders (transferfuncs.c:201-225) */
void ivrref_set_xfer(int xfer) { inittransferfunc(xfer); }

int ivrref_lutoffset(void) { return lutoffset; }
void ivrref_get_lut(float *out, int n) { memcpy(out, lut, sizeof(color4f) * (size_t)n); }

void ivrref_box(float out[6])
{
    out[0] = aabb.min.x; out[1] = aabb.min.y; out[2] = aabb.min.z;
    out[3] = aabb.max.x; out[4] = aabb.max.y; out[5] = aabb.max.z;
}

/* One frame as the server loop renders it (main.c:226-233, 239).  `out`
 * receives W*H RGBA32F in colorbuf order (row 0 = bottom).  rect receives
 * the cast pixel rectangle istrt, istop, jstrt, jstop (main.c:959-962). */
int ivrref_render(const float mvm_cm[16], const float pm_cm[16], const float vv[4],
                  const float newconst[6], float alpha, int W, int H,
                  float *out, int rect[4])
{
    float planeconst[6];
    float nc[6];
    float m[16], p[16], v4[4];
    memcpy(m, mvm_cm, sizeof m);
    memcpy(p, pm_cm, sizeof p);
    memcpy(v4, vv, sizeof v4);
    memcpy(nc, newconst, sizeof nc);
    memcpy(planeconst, pzero, sizeof planeconst);
    winwidth = W;
    winheight = H;
    mystate.alpha = alpha;
    setviewport(&viewport, 0, 0, W, H);
    setplaneconst(planeconst, nc);
    setmvm(mvm, m);
    setpm(pm, p);
    setfrustum(&viewvolume, v4);
    updatevolbbox();
    initcolbufs(0);
    /* createlocalimage may write one row past the viewport when the box
     * touches the top edge (main.c:960); give it room. */
    free(colorbuf);
    colorbuf = (color4f *)calloc((size_t)W * (size_t)(H + 2), sizeof(color4f));
    if (g_frame_gpu) {
        memset(colorbuf, 0, sizeof(color4f) * (size_t)W * (size_t)H);
        if (ivr_compat_createlocalimage(0) != 0) return -1;
    } else {
        createlocalimage(0);
    }
    memcpy(out, colorbuf, sizeof(color4f) * (size_t)W * (size_t)H);
    rect[0] = (int)(aabb.wincoord.y - 0.5f);
    rect[1] = (int)(aabb.wincoord.y + aabb.wincoord.h + 1.0f);
    rect[2] = (int)(aabb.wincoord.x - 0.5f);
    rect[3] = (int)(aabb.wincoord.x + aabb.wincoord.w + 1.0f);
    return 0;
}

static ray3f mkray(const float r[6])
{
    ray3f ray;
    ray.orig.x = r[0]; ray.orig.y = r[1]; ray.orig.z = r[2];
    ray.dir.x = r[3]; ray.dir.y = r[4]; ray.dir.z = r[5];
    return ray;
}

float ivrref_getlinearsample(float x, float y, float z) { return getlinearsample(x, y, z); }
float ivrref_getnearestsample(float x, float y, float z) { return getnearestsample(x, y, z); }

int ivrref_getrayintersect(const float r[6], float *tin, float *tout)
{
    return getrayintersect(mkray(r), tin, tout) ? 1 : 0;
}

int ivrref_getsamplesonray(const float r[6], int maxs, float *samples)
{
    initsamplefunc(1);
    return getsamplesonray(mkray(r), maxs, samples);
}

/* getfirstvoxel; out[] written only on a hit (the reference's output on a
 * miss comes from a read past its times[] array).  Callers pass rays with a
 * non-zero direction component. */
int ivrref_getfirstvoxel(const float r[6], int out[3])
{
    vector3i v;
    int rc = getfirstvoxel(mkray(r), &v);
    if (rc == 0) {
        out[0] = v.x; out[1] = v.y; out[2] = v.z;
    }
    return rc;
}

void ivrref_getvoxelsonray(const float r[6], int maxdepth, float *samples)
{
    getvoxelsonray(mkray(r), maxdepth, samples);
}

/* volume.localdims, read or replaced (getvoxelsonray only forms offsets
 * from them, so other dims exercise its index order) */
void ivrref_localdims(int out[3])
{
    out[0] = volume.localdims.x; out[1] = volume.localdims.y; out[2] = volume.localdims.z;
}

void ivrref_set_localdims(const int d[3])
{
    volume.localdims.x = d[0]; volume.localdims.y = d[1]; volume.localdims.z = d[2];
}

void ivrref_raycast(const float r[6], float alpha, float out[4])
{
    color4f c;
    mystate.alpha = alpha;
    raycast(mkray(r), &c);
    out[0] = c.r; out[1] = c.g; out[2] = c.b; out[3] = c.a;
}

/* transfer function `xfer` (transferfuncs.h:8-13) over a caller sample
 * array with the current lut/lutoffset/maxlutindx. */
void ivrref_transferfunc(int xfer, float *samples, int ns, float alpha, float out[4])
{
    color4f c;
    c.r = c.g = c.b = c.a = 0.0f;
    inittransferfunc(xfer);
    c = (*transferfunc)(lut, c, samples, ns, alpha);
    out[0] = c.r; out[1] = c.g; out[2] = c.b; out[3] = c.a;
}

/* colormap `id` (colormap.h:9-14) into out[hsize*4]; hist is all ones. */
void ivrref_colormap(int id, int hsize, float hmin, float hmax, float *out)
{
    int *hist = (int *)malloc(sizeof(int) * (size_t)hsize);
    color4f *l = (color4f *)calloc((size_t)hsize, sizeof(color4f));
    for (int i = 0; i < hsize; i++) hist[i] = 1;
    initcolormap(id);
    (*colormap)(l, hist, hsize, hmin, hmax);
    memcpy(out, l, sizeof(color4f) * (size_t)hsize);
    free(hist);
    free(l);
}

int ivrref_compositetiles(float *dst, float *tiles, int num, int w, int h)
{
    return compositetiles(dst, tiles, num, w, h);
}
