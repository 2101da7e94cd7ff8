/*
 * ivr_types.h -- layout-compatible reference types.
 *
 * The C-ABI passes these by pointer exactly as the reference lays them out,
 * so reference host code and the steering wire format keep working:
 *   color4f   types.h:115-117   {r,g,b,a} float, 16 bytes
 *   vector3f  types.h:85-101    {x,y,z} float (u/v/n, w/h/d aliases), 12 bytes
 *   ray3f     types.h:110-113   {orig, dir}, 24 bytes
 *   STATE     connection.h:9-18 steering request, 184 bytes
 */
#ifndef IVR_TYPES_H
#define IVR_TYPES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ivr_color4f {
    float r, g, b, a;
} ivr_color4f;

typedef struct ivr_vector3f {
    union { float x; float u; float w; };
    union { float y; float v; float h; };
    union { float z; float n; float d; };
} ivr_vector3f;

typedef struct ivr_ray3f {
    ivr_vector3f orig;
    ivr_vector3f dir;
} ivr_ray3f;

/* Steering request (connection.h:9-18).  On the wire the three ints are in
 * network byte order and the floats are native little-endian
 * (connection.c:167-200). */
typedef struct ivr_state {
    int finish;
    int winwidth;
    int winheight;
    float alpha;
    float newconst[6];
    float vv[4];
    float mvm[16];
    float pm[16];
} ivr_state;

#ifdef __cplusplus
#define IVR_STATIC_ASSERT static_assert
#else
#define IVR_STATIC_ASSERT _Static_assert
#endif
IVR_STATIC_ASSERT(sizeof(ivr_color4f) == 16, "color4f layout");
IVR_STATIC_ASSERT(sizeof(ivr_vector3f) == 12, "vector3f layout");
IVR_STATIC_ASSERT(sizeof(ivr_ray3f) == 24, "ray3f layout");
IVR_STATIC_ASSERT(sizeof(ivr_state) == 184, "STATE layout");
IVR_STATIC_ASSERT(offsetof(ivr_state, alpha) == 12, "STATE.alpha offset");
IVR_STATIC_ASSERT(offsetof(ivr_state, newconst) == 16, "STATE.newconst offset");
IVR_STATIC_ASSERT(offsetof(ivr_state, vv) == 40, "STATE.vv offset");
IVR_STATIC_ASSERT(offsetof(ivr_state, mvm) == 56, "STATE.mvm offset");
IVR_STATIC_ASSERT(offsetof(ivr_state, pm) == 120, "STATE.pm offset");

#ifdef __cplusplus
}
#endif
#endif
