/*
 * ivr_host.c -- host-side C of libivr.so: frame set-up, orbit steering
 * states, transfer-function builder and the PRNG every generator uses.
 *
 * Frame set-up restates the reference's per-frame host math in its float
 * operation order (this file is compiled with -ffp-contract=off), so the
 * GPU receives the same origin, rotation and pixel pitch the reference
 * computes:
 *   box           initvolbbox/getlocalbrick  main.c:509-598, utils.c:332-376
 *   planes        setplaneconst              ray.c:44-51, main.c:611-617
 *   camera        initcamera + buildray      main.c:428-500, mathutil.c:1115-1154
 *   pixel rect    updatevolbbox + createlocalimage  mathutil.c:948-1018, main.c:950-981
 */
#include "ivr_host.h"

#include <float.h>
#include <math.h>
#include <string.h>

/* ---------------- PRNG: SplitMix64 seeding, PCG32 (XSH-RR) draws ---------------- */

static uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

uint32_t ivr_rng_u32(uint64_t *state)
{
    uint64_t old = *state;
    *state = old * 6364136223846793005ULL + 1442695040888963407ULL;
    uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
    uint32_t rot = (uint32_t)(old >> 59);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

static uint64_t rng_seed(uint64_t seed) { return splitmix64(seed); }
static float rng_unit(uint64_t *s) { return (float)(ivr_rng_u32(s) >> 8) * (1.0f / 16777216.0f); }
static float rng_range(uint64_t *s, float lo, float hi) { return lo + (hi - lo) * rng_unit(s); }

/* ---------------- frame set-up ---------------- */

/* dot product in the reference order ((x + y) + z) */
static float dot3f_(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static float dot4f_(const float *a, const float *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

void ivr_box_of(int nx, int ny, int nz, float bmin[3], float bmax[3])
{
    int n[3] = {nx, ny, nz};
    int maxdim = nx;
    if (ny > maxdim) maxdim = ny;
    if (nz > maxdim) maxdim = nz;
    float inv = 1.0f / (float)maxdim;
    for (int a = 0; a < 3; a++) {
        float sf = (float)n[a] * inv;            /* scalefactor          */
        float lo = -sf + (sf * 2.0f) * 0.0f;     /* brick origin, start 0 */
        bmin[a] = lo;
        bmax[a] = lo + 2.0f * sf * 1.0f;         /* + brick extent, ratio 1 */
    }
}

/* row-major 4x4 from GL column-major (setmvm/setpm, mathutil.h:211-233) */
static void gl_to_rows(const float *cm, float m[16])
{
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) m[r * 4 + c] = cm[c * 4 + r];
}

void ivr_cam_setup(const ivr_state *st, int nx, int ny, int nz, ivr_cam *cam)
{
    memset(cam, 0, sizeof(*cam));
    ivr_box_of(nx, ny, nz, cam->bmin, cam->bmax);
    const float pz[6] = {cam->bmin[0] * -1.0f, cam->bmax[0] * 1.0f, cam->bmin[1] * -1.0f,
                         cam->bmax[1] * 1.0f,  cam->bmin[2] * -1.0f, cam->bmax[2] * 1.0f};
    for (int i = 0; i < 6; i++) cam->dconst[i] = st->newconst[i] < pz[i] ? st->newconst[i] : pz[i];

    float mv[16], pr[16];
    gl_to_rows(st->mvm, mv);
    gl_to_rows(st->pm, pr);
    const int W = st->winwidth, H = st->winheight;

    /* projected box rectangle: ctm = pm * mvm, NDC, viewport transform */
    float ctm[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float col[4] = {mv[0 * 4 + j], mv[1 * 4 + j], mv[2 * 4 + j], mv[3 * 4 + j]};
            ctm[i * 4 + j] = dot4f_(&pr[i * 4], col);
        }
    const float vpm[16] = {W / 2.0f, 0.0f, 0.0f, (0 + W) / 2.0f, 0.0f, H / 2.0f, 0.0f, (0 + H) / 2.0f,
                           0.0f,     0.0f, 0.5f, 0.5f,           0.0f, 0.0f,     0.0f, 1.0f};
    float lox = FLT_MAX, loy = FLT_MAX, hix = -FLT_MAX, hiy = -FLT_MAX;
    for (int c = 0; c < 8; c++) {
        float v[4] = {(c & 1) ? cam->bmax[0] : cam->bmin[0], (c & 2) ? cam->bmax[1] : cam->bmin[1],
                      (c & 4) ? cam->bmax[2] : cam->bmin[2], 1.0f};
        float p[4], ndc[4], win[4];
        for (int r = 0; r < 4; r++) p[r] = dot4f_(&ctm[r * 4], v);
        ndc[0] = p[0] / p[3]; ndc[1] = p[1] / p[3]; ndc[2] = p[2] / p[3]; ndc[3] = 1.0f;
        for (int r = 0; r < 2; r++) win[r] = dot4f_(&vpm[r * 4], ndc);
        lox = win[0] < lox ? win[0] : lox;
        loy = win[1] < loy ? win[1] : loy;
        hix = win[0] > hix ? win[0] : hix;
        hiy = win[1] > hiy ? win[1] : hiy;
    }
    const float fw = (float)W, fh = (float)H;
    lox = lox < 0.0f ? 0.0f : lox;  hix = hix > fw ? fw : hix;
    lox = lox > fw ? fw : lox;      hix = hix < 0.0f ? 0.0f : hix;
    loy = loy < 0.0f ? 0.0f : loy;  hiy = hiy > fh ? fh : hiy;
    loy = loy > fh ? fh : loy;      hiy = hiy < 0.0f ? 0.0f : hiy;
    float rw = hix - lox, rh = hiy - loy;
    cam->rect[0] = (int)(loy - 0.5f);
    cam->rect[1] = (int)(loy + rh + 1.0f);
    cam->rect[2] = (int)(lox - 0.5f);
    cam->rect[3] = (int)(lox + rw + 1.0f);

    /* pixel pitch and first ray (main.c:950-981) */
    const float l = st->vv[0], r = st->vv[1], b = st->vv[2], t = st->vv[3];
    cam->l = l;
    cam->b = b;
    cam->du = (r - l) / fw;
    cam->dv = (t - b) / fh;
    float v0 = b + cam->rect[0] * cam->dv + cam->dv / 2.0f;
    float u0 = l + cam->rect[2] * cam->du - cam->du / 2.0f;

    /* camera frame: n = (0,0,-1), v = (0,1,0)/|v|, u = n x v, eye (0,0,-5)
     * in u,v,n; vwm has u, v, n as columns */
    const float n3[3] = {0.0f, 0.0f, -1.0f};
    float up[3] = {0.0f, 1.0f, 0.0f};
    float mag = sqrtf(up[0] * up[0] + up[1] * up[1] + up[2] * up[2]);
    up[0] = up[0] / mag; up[1] = up[1] / mag; up[2] = up[2] / mag;
    float u3[3] = {n3[1] * up[2] - n3[2] * up[1], n3[2] * up[0] - n3[0] * up[2],
                   n3[0] * up[1] - n3[1] * up[0]};
    float vwm[9];
    for (int k = 0; k < 3; k++) { vwm[k * 3 + 0] = u3[k]; vwm[k * 3 + 1] = up[k]; vwm[k * 3 + 2] = n3[k]; }
    const float e[3] = {0.0f, 0.0f, -5.0f};

    /* buildray: eye = vwm e + r, minus the x/y pan, rotated by mvm^T */
    float eye[3], o[3], tmp[3];
    for (int k = 0; k < 3; k++) eye[k] = dot3f_(&vwm[k * 3], e);
    for (int k = 0; k < 3; k++) eye[k] = eye[k] + 0.0f;
    o[0] = eye[0] - mv[0 * 4 + 3];
    o[1] = eye[1] - mv[1 * 4 + 3];
    o[2] = eye[2] - 0.0f;
    float rt[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) rt[i * 3 + j] = mv[j * 4 + i];
    for (int k = 0; k < 3; k++) tmp[k] = dot3f_(&rt[k * 3], o);
    memcpy(cam->o, tmp, sizeof tmp);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            float col[3] = {vwm[0 * 3 + j], vwm[1 * 3 + j], vwm[2 * 3 + j]};
            cam->M[i * 3 + j] = dot3f_(&rt[i * 3], col);
        }
    for (int k = 0; k < 3; k++) cam->m5[k] = cam->M[k * 3 + 2] * 5.0f;
    const float dv3[3] = {u0, v0, -e[2]};
    const float iv[3] = {0.0f, cam->dv, 0.0f};
    const float jv[3] = {cam->du, 0.0f, 0.0f};
    for (int k = 0; k < 3; k++) {
        cam->d0[k] = dot3f_(&cam->M[k * 3], dv3);
        cam->di[k] = dot3f_(&cam->M[k * 3], iv);
        cam->dj[k] = dot3f_(&cam->M[k * 3], jv);
    }
    int n[3] = {nx, ny, nz};
    cam->vox = (cam->bmax[0] - cam->bmin[0]) / (float)nx;
    for (int a = 0; a < 3; a++) cam->s[a] = (float)n[a] / (cam->bmax[a] - cam->bmin[a]);
}

/* ---------------- steering states ---------------- */

static void rows_to_gl(const double m[4][4], float *cm)
{
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) cm[c * 4 + r] = (float)m[r][c];
}

void ivr_state_from_angles(float az, float el, int width, int height, ivr_state *st)
{
    memset(st, 0, sizeof(*st));
    st->finish = 0;
    st->winwidth = width;
    st->winheight = height;
    st->alpha = 1.0f;
    for (int i = 0; i < 6; i++) st->newconst[i] = 10.0f; /* cut planes outside the box */
    /* modelview as the client builds it (ivrclient.c:545-553):
     * lookAt((0,0,5) -> origin) = translate(0,0,-5), times the rotation
     * Rx(el) * Ry(az) */
    double ca = cos(az), sa = sin(az), ce = cos(el), se = sin(el);
    double R[4][4] = {{ca, 0, sa, 0}, {se * sa, ce, -se * ca, 0}, {-ce * sa, se, ce * ca, 0}, {0, 0, 0, 1}};
    R[2][3] = -5.0;
    rows_to_gl(R, st->mvm);
    /* gluPerspective(43.6028, aspect, 0.1, 100) and getfrustum
     * (ivrclient.c:562-563, mathutil.h:166-175) */
    double aspect = (double)width / (double)height;
    double fovy = 43.6028 * M_PI / 180.0;
    double f = 1.0 / tan(fovy / 2.0), zn = 0.1, zf = 100.0;
    double P[4][4] = {{f / aspect, 0, 0, 0}, {0, f, 0, 0},
                      {0, 0, (zf + zn) / (zn - zf), 2.0 * zf * zn / (zn - zf)}, {0, 0, -1, 0}};
    rows_to_gl(P, st->pm);
    float t = (float)(tan(fovy / 2.0) * 5.0);
    float asp = (float)aspect;
    st->vv[0] = asp * -t;
    st->vv[1] = asp * t;
    st->vv[2] = -t;
    st->vv[3] = t;
}

void ivr_state_orbit(uint64_t seed, int width, int height, ivr_state *st, float *az_out,
                     float *el_out)
{
    uint64_t s = rng_seed(seed);
    float az = rng_range(&s, 0.0f, 6.2831853f);
    float el = rng_range(&s, -0.6f, 0.6f);
    ivr_state_from_angles(az, el, width, height, st);
    if (az_out) *az_out = az;
    if (el_out) *el_out = el;
}

void ivr_desc_config(ivr_frame_desc *d)
{
    memset(d, 0, sizeof(*d));
    d->mode = IVR_MODE_CONFIG;
    d->xfer = 3;
    d->step_vox = 0.5f;
    d->ert = 0.99f;
    d->lutscale = 255.0f;
    d->lutoffset = 0;
    d->maxlutindx = 255;
    d->skip = 0;
    d->composite = IVR_FUSED;
}

/* ---------------- wire format of STATE ---------------- */

static uint32_t be32(uint32_t v)
{
    const unsigned char b[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16), (unsigned char)(v >> 8),
                                (unsigned char)v};
    uint32_t r;
    memcpy(&r, b, 4);
    return r;
}

void ivr_state_pack(const ivr_state *st, unsigned char out[184])
{
    ivr_state w = *st;
    w.finish = (int)be32((uint32_t)st->finish);
    w.winwidth = (int)be32((uint32_t)st->winwidth);
    w.winheight = (int)be32((uint32_t)st->winheight);
    memcpy(out, &w, sizeof w);
}

void ivr_state_unpack(const unsigned char in[184], ivr_state *st)
{
    memcpy(st, in, sizeof *st);
    uint32_t v[3];
    memcpy(v, in, 12);
    st->finish = (int)be32(v[0]);
    st->winwidth = (int)be32(v[1]);
    st->winheight = (int)be32(v[2]);
}

static float bswap_f(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u = (u >> 24) | ((u >> 8) & 0xff00u) | ((u << 8) & 0xff0000u) | (u << 24);
    memcpy(&f, &u, 4);
    return f;
}

void ivr_state_swap_floats(ivr_state *st)
{
    st->alpha = bswap_f(st->alpha);
    for (int i = 0; i < 6; i++) st->newconst[i] = bswap_f(st->newconst[i]);
    for (int i = 0; i < 4; i++) st->vv[i] = bswap_f(st->vv[i]);
    for (int i = 0; i < 16; i++) st->mvm[i] = bswap_f(st->mvm[i]);
    for (int i = 0; i < 16; i++) st->pm[i] = bswap_f(st->pm[i]);
}

/* ---------------- TIFF image output ---------------- */

/* one 12-byte directory entry, big-endian (make_tag, utils.c:135-154) */
static unsigned char *tiff_tag(unsigned char *b, int tag, int type, long lng, long fld, int *cnt)
{
    const unsigned long v[4] = {(unsigned long)tag, (unsigned long)type, (unsigned long)lng, (unsigned long)fld};
    const int width[4] = {2, 2, 4, 4};
    for (int k = 0; k < 4; k++)
        for (int byte = width[k] - 1; byte >= 0; byte--) *b++ = (unsigned char)((v[k] >> (8 * byte)) & 0xffu);
    (*cnt)++;
    return b;
}

long long ivr_tiff_encode(const float *rgba, int w, int h, int planes, unsigned char *out, size_t cap)
{
    if (!rgba || !out || w < 1 || h < 1 || w > 65535 || h > 65535 || (planes != 1 && planes != 3)) return -1;
    const size_t need = 512 + (size_t)w * h * planes;
    if (cap < need) return -1;
    /* header (writetiffheader, utils.c:24-100): "MM", 42, the directory at
     * byte 10 with its entry count at byte 11, SHORT values in the upper
     * half of their field; 8 bits per sample written at byte 256 for RGB */
    unsigned char hdr[512];
    memset(hdr, 0, sizeof hdr);
    hdr[0] = 0x4d;
    hdr[1] = 0x4d;
    hdr[3] = 42;
    hdr[7] = 10;
    int cnt = 0;
    const int SHORT_T = 3, LONG_T = 4;
    unsigned char *b = hdr + 12;
    b = tiff_tag(b, 254, LONG_T, 1, 0, &cnt);
    b = tiff_tag(b, 256, SHORT_T, 1, (long)w << 16, &cnt);
    b = tiff_tag(b, 257, SHORT_T, 1, (long)h << 16, &cnt);
    b = planes == 3 ? tiff_tag(b, 258, SHORT_T, 3, 256, &cnt) : tiff_tag(b, 258, SHORT_T, 1, 8L << 16, &cnt);
    b = tiff_tag(b, 259, SHORT_T, 1, 1L << 16, &cnt);
    b = tiff_tag(b, 262, SHORT_T, 1, (planes == 3 ? 2L : 1L) << 16, &cnt);
    b = tiff_tag(b, 273, LONG_T, 1, 512, &cnt);
    b = tiff_tag(b, 277, SHORT_T, 1, (long)planes << 16, &cnt);
    b = tiff_tag(b, 278, SHORT_T, 1, (long)h << 16, &cnt);
    b = tiff_tag(b, 279, LONG_T, 1, (long)h * w * planes, &cnt);
    b = tiff_tag(b, 284, SHORT_T, 1, 1L << 16, &cnt);
    b = tiff_tag(b, 0, 0, 0, 0, &cnt);
    hdr[11] = (unsigned char)(cnt - 1);
    if (planes == 3) hdr[257] = hdr[259] = hdr[261] = 8;
    memcpy(out, hdr, 512);
    /* samples (writetiffdata, utils.c:102-133): rows in the image's order,
     * each channel value * 255 truncated to a byte */
    unsigned char *d = out + 512;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const float *px = rgba + ((size_t)y * w + x) * 4;
            for (int c = 0; c < planes; c++) *d++ = (unsigned char)(px[c] * 255);
        }
    return (long long)need;
}

/* ---------------- transfer-function builder ---------------- */

static float clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

void ivr_tf_build(uint64_t seed, float alpha_max, float jitter, uint64_t jitter_seed, float *lut)
{
    uint64_t s = rng_seed(seed);
    float pos[5], col[5][4];
    pos[0] = 0.0f;
    pos[4] = 1.0f;
    for (int i = 1; i < 4; i++) pos[i] = rng_unit(&s);
    for (int i = 1; i < 4; i++) /* insertion sort of the three interior positions */
        for (int j = i; j > 1 && pos[j] < pos[j - 1]; j--) {
            float tt = pos[j]; pos[j] = pos[j - 1]; pos[j - 1] = tt;
        }
    for (int i = 0; i < 5; i++) {
        col[i][0] = rng_unit(&s);
        col[i][1] = rng_unit(&s);
        col[i][2] = rng_unit(&s);
        col[i][3] = i == 0 ? 0.0f : alpha_max * rng_unit(&s);
    }
    if (jitter > 0.0f) {
        uint64_t js = rng_seed(jitter_seed);
        for (int i = 1; i < 4; i++) pos[i] = clamp01(pos[i] + rng_range(&js, -jitter, jitter));
        for (int i = 1; i < 4; i++)
            for (int j = i; j > 1 && pos[j] < pos[j - 1]; j--) {
                float tt = pos[j]; pos[j] = pos[j - 1]; pos[j - 1] = tt;
            }
        for (int i = 0; i < 5; i++)
            for (int c = 0; c < 4; c++) {
                if (i == 0 && c == 3) continue; /* value 0 stays transparent */
                float hi = c == 3 ? alpha_max : 1.0f;
                float v = col[i][c] + rng_range(&js, -jitter, jitter);
                col[i][c] = v < 0.0f ? 0.0f : (v > hi ? hi : v);
            }
    }
    for (int e = 0; e < 256; e++) {
        float x = (float)e / 255.0f;
        int sgm = 0;
        while (sgm < 3 && x > pos[sgm + 1]) sgm++;
        float w = pos[sgm + 1] - pos[sgm];
        float t = w > 0.0f ? (x - pos[sgm]) / w : 0.0f;
        for (int c = 0; c < 4; c++) lut[e * 4 + c] = col[sgm][c] + t * (col[sgm + 1][c] - col[sgm][c]);
    }
}
