/*
 * ivr_host.h -- host-side frame set-up shared by the C host code and the HIP
 * launch code of libivr.so (internal header).
 */
#ifndef IVR_HOST_H
#define IVR_HOST_H

#include "../../include/ivr_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-frame camera and box quantities, in the float operation order of the
 * reference set-up code. */
typedef struct ivr_cam {
    float bmin[3], bmax[3]; /* world box (main.c:509-598)                      */
    float dconst[6];        /* min(box planes, newconst) (ray.c:44-51)        */
    float o[3];             /* ray origin (mathutil.c:1127-1138)              */
    float M[9];             /* view-to-world rotation mvm^T * vwm, row-major  */
    float m5[3];            /* M[r][2] * 5 (view-plane distance)               */
    float l, b, du, dv;     /* view volume left/bottom and pixel pitch         */
    float d0[3], di[3], dj[3]; /* compat: first ray dir, row and column steps  */
    int rect[4];            /* compat: istrt, istop, jstrt, jstop (main.c:959-962) */
    float vox;              /* voxel edge in world units                       */
    float s[3];             /* voxels per world unit per axis                  */
} ivr_cam;

void ivr_box_of(int nx, int ny, int nz, float bmin[3], float bmax[3]);
void ivr_cam_setup(const ivr_state *st, int nx, int ny, int nz, ivr_cam *cam);

#ifdef __cplusplus
}
#endif
#endif
