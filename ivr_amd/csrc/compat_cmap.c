/*
 * compat_cmap.c -- libivr_compat.so: the reference's colour-map LUT builders
 * (colormap.c:21-205) under their own names, with the colormap function
 * pointer colormap.h:22 defines.  They are host initialisation code in the
 * reference (called once from initvolluts, main.c:1075-1200) and stay host
 * code here; their arithmetic types follow the reference (double literals,
 * float stores) so the LUT entries are bit-identical.  The tissue-class
 * breakpoints and colours of medcolormap are the reference's table.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/ivr_compat.h"

void (*colormap)(ivr_color4f *, int *, int, float, float);

void nomap(ivr_color4f *l, int *hist, int hsize, float hmin, float hmax)
{
    (void)l; (void)hist; (void)hsize; (void)hmin; (void)hmax;
}

void basecolormap(ivr_color4f *l, int *hist, int hsize, float hmin, float hmax)
{
    (void)hist;
    for (int i = 0; i < hsize; i++) {
        const float c = (float)(i + hmin) / hmax;
        l[i].r = l[i].g = l[i].b = c;
        if (c > 0.8) { l[i].r = 0.7f; l[i].g = 0.75f; l[i].b = 0.46; l[i].a = 0.07; }
        if (c < 0.3) { l[i].b = 0.6; l[i].a = 0.01; }
        if (c > 0.3 && c < 0.6) { l[i].r = 0.6; l[i].a = 0.01; }
        if (c > 0.6 && c < 0.8) { l[i].r = 0.6; l[i].g = 0.6; l[i].a = 0.005; }
    }
}

void rainbowcolormap(ivr_color4f *l, int *hist, int hsize, float hmin, float hmax)
{
    (void)hist; (void)hmin; (void)hmax;
    const float dr = 1.0f / hsize;
    l[0].b = 1.0f;
    l[0].g = 0.0f;
    l[0].r = 0.0f;
    for (int i = 0; i < hsize; i++) { /* increments the caller's entries */
        l[i].r += dr;
        l[i].b += -dr;
        l[i].a = 1.0f;
    }
}

/* one blend of two tissue classes: weights (1-a) p0 and a p1 (double math) */
static void blend2(ivr_color4f *e, float a, float p0, float p1, const double c0[4], const double c1[4])
{
    const float y1 = (1.0 - a) * p0;
    const float y2 = a * p1;
    e->r = (c0[0] * y1 + c1[0] * y2);
    e->g = (c0[1] * y1 + c1[1] * y2);
    e->b = (c0[2] * y1 + c1[2] * y2);
    e->a = (c0[3] * y1 + c1[3] * y2);
}

void medcolormap(ivr_color4f *l, int *hist, int hsize, float hmin, float hmax)
{
    (void)hist; (void)hmin; (void)hmax;
    const float parm[4] = {0.04, 0.9, 0.9, 0.8};
    static const double air_fat[2][4] = {{0.6, 0.6, 0.8, 0.1}, {0.8, 0.7, 0.0, 0.03125}};
    static const double fat_tis[2][4] = {{0.8, 0.7, 0.0, 0.03125}, {0.2, 0.0, 0.0, 0.1875}};
    static const double tis_bone[2][4] = {{0.2, 0.0, 0.0, 0.1875}, {1.0, 1.0, 0.9, 0.75}};
    for (int i = 0; i < hsize; i++) {
        const float x = i / (float)(hsize - 1) * 4096.0f; /* RANGE, types.h:43 */
        ivr_color4f *e = &l[i];
        if (0 <= x && x <= 484) {
            e->r = 0.1; e->g = 0.1; e->b = 0.1; e->a = 0.1 * parm[0];
        } else if (484 < x && x <= 545) {
            blend2(e, (x - 484.0) / (545.0 - 484.0), parm[0], parm[1], air_fat[0], air_fat[1]);
        } else if (545 < x && x <= 726) {
            e->r = 0.8; e->g = 0.7; e->b = 0.0; e->a = 0.03125 * parm[1];
        } else if (726 < x && x <= 819) {
            blend2(e, (x - 726.0) / (819.0 - 726.0), parm[1], parm[2], fat_tis[0], fat_tis[1]);
        } else if (819 < x && x <= 883) {
            e->r = 0.2; e->g = 0.0; e->b = 0.0; e->a = 0.1875 * parm[2];
        } else if (883 < x && x <= 1535) {
            blend2(e, (x - 883.0) / (1535.0 - 883.0), parm[2], parm[3], tis_bone[0], tis_bone[1]);
        } else if (1535 < x) {
            e->r = 1.0; e->g = 1.0; e->b = 0.9; e->a = 0.75 * parm[3];
        }
    }
}

/* two-segment ramp through (v0[k], v1[k]) at integer positions 1..maplen
 * (the lerp helper of mathutil.h:42-59) */
static void ramp(float *out, const float *v0, const float *v1, int maplen)
{
    const float s1 = (v1[1] - v1[0]) / (v0[1] - v0[0]);
    const float s2 = (v1[2] - v1[1]) / (v0[2] - v0[1]);
    for (int i = 1; i <= maplen; i++) {
        if (i >= v0[0] && i <= v0[1]) out[i - 1] = v1[0] + i * s1;
        else if (i >= v0[1] && i <= v0[2]) out[i - 1] = v1[1] + (i - v0[1]) * s2;
    }
}

void seismiccolormap(ivr_color4f *l, int *hist, int hsize, float hmin, float hmax)
{
    (void)hist; (void)hmin; (void)hmax;
    const float knots[3] = {1.0f, 0.5f * hsize, (float)hsize};
    const float rv[3] = {1.0f, 1.0f, 0.0f}, gv[3] = {0.0f, 1.0f, 0.0f}, bv[3] = {0.0f, 1.0f, 1.0f};
    float *ch = (float *)calloc((size_t)hsize * 3, sizeof(float));
    ramp(ch, knots, rv, hsize);
    ramp(ch + hsize, knots, gv, hsize);
    ramp(ch + 2 * (size_t)hsize, knots, bv, hsize);
    for (int i = 0; i < hsize; i++) {
        l[i].r = ch[i];
        l[i].g = ch[hsize + i];
        l[i].b = ch[2 * hsize + i];
        l[i].a = 1.0f;
    }
    free(ch);
}

void jetcolormap(ivr_color4f *l, int *hist, int hsize, float hmin, float hmax)
{
    (void)hist; (void)hmin; (void)hmax;
    for (int i = 0; i < hsize; i++) {
        const float q = (4.0f * i) / hsize;
        l[i].r = fmin(fmax(fmin(q - 1.5, -q + 4.5), 0.0f), 1.0f);
        l[i].g = fmin(fmax(fmin(q - 0.5, -q + 3.5), 0.0f), 1.0f);
        l[i].b = fmin(fmax(fmin(q + 0.5, -q + 2.5), 0.0f), 1.0f);
        l[i].a = 1.0f;
    }
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
