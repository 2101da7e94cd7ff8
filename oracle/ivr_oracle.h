/*
 * ivr_oracle.h -- CPU oracle for the iVR render path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the checker
 * or the timed CPU baseline; the product (ivr_amd/libivr.so) never links or
 * calls it.
 *
 * Three renderers over the same inputs:
 *   IVRO_COMPAT  exact ray.c semantics: t-space dt=DELTASAMPLE, 256-sample
 *                cap, the reference "trilinear" formula, back-to-front
 *                transfer function, projected-bbox pixel rectangle and the
 *                per-row column accumulation of ray directions.
 *   IVRO_CONFIG  ray.c structure with the benchmark-mandated parameters:
 *                step in voxels, no cap, true trilinear, front-to-back
 *                compositing with early ray termination, fp32.
 *   ivro_render_f64: IVRO_CONFIG evaluated in double precision.
 */
#ifndef IVR_ORACLE_H
#define IVR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { IVRO_COMPAT = 0, IVRO_CONFIG = 1 };

/* Per-frame view, in the reference STATE form (connection.h:9-18). */
typedef struct {
    float mvm[16];     /* GL column-major modelview  */
    float pm[16];      /* GL column-major projection */
    float vv[4];       /* view volume l, r, b, t     */
    float newconst[6]; /* cut-plane constants        */
    float alpha;       /* opacity scale (myalpha)    */
    int width, height;
} ivro_view;

typedef struct {
    int mode;          /* IVRO_COMPAT or IVRO_CONFIG                         */
    int xfer;          /* compat: transfer function id, transferfuncs.h:8-13 */
    float step_vox;    /* config: sample spacing in voxels                   */
    float ert;         /* config: stop once accumulated alpha >= ert         */
    float lutscale;    /* LUT index = (int)(s*lutscale) - lutoffset          */
    int lutoffset;
    int maxlutindx;    /* compat SEISMIC2 upper clamp                        */
    int brick;         /* config tile mode: brick edge in voxels (0 = off)   */
    int brick_id[3];   /* config tile mode: brick whose samples are composited */
    int sort_last;     /* config with brick > 0: one pass per ray over all bricks,
                          each brick's partial composited from zero with ERT
                          inside it, partials combined far to near as
                          compositetiles does (brick_id unused) */
} ivro_params;

/* World box of an nx*ny*nz volume (main.c:509-598, utils.c:332-388). */
void ivro_box(int nx, int ny, int nz, float bmin[3], float bmax[3]);

/* Slab test (ray.c:276-318) against min(box planes, newconst). */
int ivro_rayintersect(const float bmin[3], const float bmax[3], const float newconst[6],
                      const float ray[6], float *tin, float *tout);

/* Samplers over a linear x-fastest volume. */
float ivro_sample_ref(const float *vol, int nx, int ny, int nz, const float bmin[3],
                      const float bmax[3], float x, float y, float z);
float ivro_sample_tri(const float *vol, int nx, int ny, int nz, const float bmin[3],
                      const float bmax[3], float x, float y, float z);
/* getnearestsample (ray.c:182-238, non-BGQ): the voxel at the truncated
 * index (clamped to n-2 above, 0 below), its bits read as an unsigned int and
 * converted to float. */
float ivro_sample_nearest(const float *vol, int nx, int ny, int nz, const float bmin[3],
                          const float bmax[3], float x, float y, float z);
double ivro_sample_tri64(const float *vol, int nx, int ny, int nz, const double bmin[3],
                         const double bmax[3], double x, double y, double z);

/* getsamplesonray (ray.c:239-273) with the reference sampler. */
int ivro_samplesonray_compat(const float *vol, int nx, int ny, int nz, const float newconst[6],
                             const float ray[6], int maxs, float *samples);

/* Voxel walk of ray.c:333-483.  ivro_firstvoxel: getfirstvoxel's entry
 * voxel (0) or -1 with first[] untouched.  ivro_voxelsonray: getvoxelsonray's
 * maxdepth offsets maptoffset3d(x, y, z, dims) as floats, or -1 (nothing
 * written).  bmin/bmax: the host's aabb.min/max; dims: volume.localdims. */
int ivro_firstvoxel(const float bmin[3], const float bmax[3], const float ray[6], int first[3]);
int ivro_voxelsonray(const float bmin[3], const float bmax[3], const int dims[3], const float ray[6],
                     int maxdepth, float *samples);

/* Gradient field of a linear x-fastest volume (gradient3d, mathutil.c:242-328,
 * stored at each voxel's own offset): out[4*v] = d/dx, d/dy, d/dz, magnitude;
 * per axis (f[+1] - f[-1]) / 2 inside, one-sided differences / 2 on the faces. */
void ivro_gradient(const float *vol, int nx, int ny, int nz, float *out);

/* The six transfer functions (transferfuncs.c:24-199) over a sample array. */
void ivro_transferfunc(int xfer, const float *lut, int lutoffset, int maxlutindx,
                       const float *samples, int n, float alpha, float out[4]);

/* The five colormaps (colormap.c:21-181) into out[hsize*4]. */
void ivro_colormap(int id, int hsize, float hmin, float hmax, float *out);

/* Ray set-up (mathutil.c:1115-1154, main.c:943-981).  cam[0..2] origin,
 * cam[3..11] the 3x3 view-to-world rotation M (row-major), cam[12] du,
 * cam[13] dv, cam[14..16] compat first-ray dir, cam[17..19] incri,
 * cam[20..22] incrj.  rect = istrt, istop, jstrt, jstop (main.c:959-962). */
void ivro_camera(const float bmin[3], const float bmax[3], const ivro_view *v,
                 float cam[23], int rect[4]);

/* Full frame into rgba[W*H*4] (row 0 = bottom, like colorbuf).  Rows
 * [row0,row1) only; pixels outside stay untouched.  Returns the number of
 * samples classified and composited. */
long long ivro_render(const float *vol, int nx, int ny, int nz, const ivro_view *v,
                      const float *lut, int lutsize, const ivro_params *p,
                      float *rgba, int row0, int row1, int nthreads);
long long ivro_render_f64(const float *vol, int nx, int ny, int nz, const ivro_view *v,
                          const float *lut, int lutsize, const ivro_params *p,
                          double *rgba, int row0, int row1, int nthreads);

/* compositetiles (main.c:1203-1218): dst = tiles[k] over dst, k = num-1..0. */
void ivro_composite_tiles(float *dst, const float *tiles, int num, int npix);

/* RGBA32F -> RGBA8 (main.c:1404), optionally clamped to [0,1]. */
void ivro_quantize(const float *rgba, long long n, unsigned char *out, int clamp);

#ifdef __cplusplus
}
#endif
#endif
