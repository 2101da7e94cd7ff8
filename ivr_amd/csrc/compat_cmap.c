/*
 * compat_cmap.c -- libivr_compat.so: the reference's colour-map LUT builders
 * (colormap.c:21-205) under their own names, with the colormap function
 * pointer colormap.h:22 defines.  Each builder runs on the device over the
 * caller's table (ivr_compat_colormap, one launch: the k_colormap kernel of
 * luts.hip, bit-identical to the reference's arithmetic); the histogram
 * argument is unused, as in the reference builders.
 */
#include <stdio.h>

#include "../../include/ivr_compat.h"
#include "../../include/ivr_gpu.h"

void (*colormap)(ivr_color4f *, int *, int, float, float);

static void build(int id, ivr_color4f *l, int hsize, float hmin, float hmax)
{
    if (ivr_compat_colormap(id, l, hsize, hmin, hmax))
        fprintf(stderr, "libivr_compat: colormap %d failed: %s\n", id, ivr_last_error());
}

void nomap(ivr_color4f *l, int *hist, int hsize, float hmin, float hmax)
{
    (void)l; (void)hist; (void)hsize; (void)hmin; (void)hmax;
}

void basecolormap(ivr_color4f *l, int *hist, int hsize, float hmin, float hmax)
{
    (void)hist;
    build(0, l, hsize, hmin, hmax);
}

void rainbowcolormap(ivr_color4f *l, int *hist, int hsize, float hmin, float hmax)
{
    (void)hist;
    build(1, l, hsize, hmin, hmax);
}

void medcolormap(ivr_color4f *l, int *hist, int hsize, float hmin, float hmax)
{
    (void)hist;
    build(2, l, hsize, hmin, hmax);
}

void seismiccolormap(ivr_color4f *l, int *hist, int hsize, float hmin, float hmax)
{
    (void)hist;
    build(3, l, hsize, hmin, hmax);
}

void jetcolormap(ivr_color4f *l, int *hist, int hsize, float hmin, float hmax)
{
    (void)hist;
    build(4, l, hsize, hmin, hmax);
}

void initcolormap(int type)
{
    switch (type) {
    case 0: colormap = &basecolormap; break;
    case 1: colormap = &rainbowcolormap; break;
    case 2: colormap = &medcolormap; break;
    case 3: colormap = &seismiccolormap; break;
    case 4: colormap = &jetcolormap; break;
    default: colormap = &nomap; break;
    }
}
