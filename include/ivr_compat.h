/*
 * ivr_compat.h -- drop-in replacement of the reference's ray.c,
 * transferfuncs.c and colormap.c (libivr_compat.so).
 *
 * Reference host code (main.c, mathutil.c, utils.c, glui.c) links against
 * libivr_compat.so instead of ray.o, transferfuncs.o and colormap.o.  The
 * library exports every function of ray.h, transferfuncs.h and colormap.h
 * with the reference name and signature, defines the globals those files
 * define (dconst, ray.c:38; the function pointers of ray.h:25,
 * transferfuncs.h:22 and colormap.h:22), and reads the globals the host
 * defines (volume, aabb, lut, lutoffset, maxlutindx, mystate, cnt,
 * colorbuf, viewport, viewvolume, mvm, pm; main.c:63-105).
 *
 * The per-ray and per-sample functions run on the GPU, bit-exact with the
 * reference: one launch per call, except that raycast calls made at the same
 * time by several threads (the reference's OpenMP pixel loop) share a launch;
 * the frame loop of createlocalimage
 * (main.c:937-1026) becomes one launch through
 * ivr_compat_createlocalimage(rank), which a host replaces its pixel loop
 * with.  The LUT builders of colormap.c run on the device over the caller's
 * table (ivr_compat_colormap).  The host's volume data is mirrored to the
 * device on first use (its value range, which sizes the LUT, is reduced
 * there too); call ivr_compat_invalidate() after changing volume.data or
 * dims.
 */
#ifndef IVR_COMPAT_H
#define IVR_COMPAT_H

#include <stdbool.h>
#include <stddef.h>

#include "ivr_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference layouts (types.h), asserted against the reference ---- */
typedef struct { int x, y, z; } ivr_ref_vector3i;     /* vector3i types.h:72-83 */
typedef struct { float x, y, z, w; } ivr_ref_vector4f; /* vector4f types.h:103-105 */
typedef struct {                                       /* VOLUME types.h:143-156 */
    float *data;
    ivr_ref_vector4f *gradient;
    ivr_ref_vector3i globaldims, localdims, procdims, startindex;
    ivr_vector3f celldims, fvoldims, rfvoldims;
    int endianness;
    int byteswap;
} ivr_ref_volume;
typedef struct { float x, y, z, w, h, d, cx, cy; } ivr_ref_rect; /* RECT types.h:158-163 */
typedef struct {                                                  /* AABB types.h:171-185 */
    int rank;
    float w, h, d;
    ivr_ref_rect wincoord;
    ivr_vector3f mvmin, mvmax, min, max;
    ivr_vector3f facenormals[6];
    ivr_ref_vector4f verts[8], everts[8], pverts[8], mverts[8];
} ivr_ref_aabb;
typedef struct { float left, right, top, bottom, near_, far_; } ivr_ref_frustum; /* types.h:196-203 */
typedef struct { int tilex, tiley, tilewid, tileheight; float x, y, w, h; } ivr_ref_window; /* types.h:205-212 */

/* sizes and offsets measured on the reference's types.h (x86-64, gcc) */
#ifndef __cplusplus
_Static_assert(sizeof(ivr_ref_volume) == 112, "VOLUME layout");
_Static_assert(offsetof(ivr_ref_volume, localdims) == 28, "VOLUME.localdims");
_Static_assert(offsetof(ivr_ref_volume, rfvoldims) == 88, "VOLUME.rfvoldims");
_Static_assert(sizeof(ivr_ref_aabb) == 680, "AABB layout");
_Static_assert(offsetof(ivr_ref_aabb, wincoord) == 16, "AABB.wincoord");
_Static_assert(offsetof(ivr_ref_aabb, min) == 72, "AABB.min");
_Static_assert(offsetof(ivr_ref_aabb, facenormals) == 96, "AABB.facenormals");
_Static_assert(offsetof(ivr_ref_aabb, mverts) == 552, "AABB.mverts");
_Static_assert(sizeof(ivr_ref_window) == 32, "WINDOW layout");
_Static_assert(sizeof(ivr_ref_frustum) == 24, "FRUSTUM layout");
#endif

/* ---- ray.h:16-26 ---- */
void raycast(ivr_ray3f ray, ivr_color4f *color);                       /* ray.c:59-85   */
int getsamplesonray(ivr_ray3f ray, int maxsamples, float *samples);     /* ray.c:239-273 */
bool getrayintersect(ivr_ray3f ray, float *tin, float *tout);           /* ray.c:276-318 */
float getlinearsample(float x, float y, float z);                       /* ray.c:87-181  */
float getnearestsample(float x, float y, float z);                      /* ray.c:182-238 */
void setplaneconst(float *planeconst, float *newconst);                 /* ray.c:44-51   */
void initsamplefunc(int type);                                          /* ray.c:485-501 */
/* voxel walk (reference ray.c:333-483, which nothing in the reference calls):
 * entry voxel from aabb, then maxdepth offsets over volume.localdims.  On a
 * miss (-1; only an all-zero direction) first / samples are left as they
 * are, where the reference reads past its times[] array. */
int getfirstvoxel(ivr_ray3f ray, ivr_ref_vector3i *first);              /* ray.c:333-384 */
void getvoxelsonray(ivr_ray3f ray, int maxdepth, float *samples);       /* ray.c:387-483 */

/* ---- transferfuncs.h:15-22 ---- */
void inittransferfunc(int xfer);                                         /* transferfuncs.c:201-225 */
ivr_color4f medtransferfunc(ivr_color4f *lut, ivr_color4f c, float *s, int n, float alpha);
ivr_color4f blendtransferfunc(ivr_color4f *lut, ivr_color4f c, float *s, int n, float alpha);
ivr_color4f seismictransferfunc(ivr_color4f *lut, ivr_color4f c, float *s, int n, float alpha);
ivr_color4f seismictransferfunc2(ivr_color4f *lut, ivr_color4f c, float *s, int n, float alpha);
ivr_color4f chemistrytransferfunc(ivr_color4f *lut, ivr_color4f c, float *s, int n, float alpha);
ivr_color4f rgbtransferfunc(ivr_color4f *lut, ivr_color4f c, float *s, int n, float alpha);

/* ---- colormap.h:15-22 ---- */
void initcolormap(int type);                                             /* colormap.c:183-205 */
void basecolormap(ivr_color4f *lut, int *hist, int hsize, float hmin, float hmax);
void rainbowcolormap(ivr_color4f *lut, int *hist, int hsize, float hmin, float hmax);
void medcolormap(ivr_color4f *lut, int *hist, int hsize, float hmin, float hmax);
void seismiccolormap(ivr_color4f *lut, int *hist, int hsize, float hmin, float hmax);
void jetcolormap(ivr_color4f *lut, int *hist, int hsize, float hmin, float hmax);
void nomap(ivr_color4f *lut, int *hist, int hsize, float hmin, float hmax);

/* ---- frame level ---- */
/* createlocalimage(rank) (main.c:937-1026) as one GPU launch: fills the
 * host's colorbuf (viewport.w x viewport.h RGBA32F) exactly as the
 * reference pixel loop does. */
int ivr_compat_createlocalimage(int rank);
/* compositetiles (main.c:1203-1218) on the GPU */
int ivr_compat_compositetiles(float *tilebuf, float *rcolorbuf, int num, int w, int h);
/* colormap builder `id` (colormap.h:9-14; -1 nomap) applied on the device to
 * the caller's host table of hsize entries, as the builders of colormap.c:21-181
 * write it (fields they leave alone keep the caller's values); the exported
 * builders above call it */
int ivr_compat_colormap(int id, ivr_color4f *lut, int hsize, float hmin, float hmax);
/* launches of the combined raycast calls so far and the rays they cast */
void ivr_compat_raycast_stats(unsigned long long *launches, unsigned long long *rays);
/* forget the device mirror of the host volume */
void ivr_compat_invalidate(void);

#ifdef __cplusplus
}
#endif
#endif
