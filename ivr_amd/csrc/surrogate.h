/*
 * surrogate.h -- learned surrogate objects shared by surrogate.hip and the
 * C-ABI (internal header).
 */
#ifndef IVR_SURROGATE_H
#define IVR_SURROGATE_H

#include "ivr_internal.h"

#define SG_L 16  /* hash-grid levels            */
#define SG_F 2   /* features per level          */
#define SG_IN 32 /* encoding width = SG_L*SG_F  */
#define SG_H 64  /* hidden width                */

/* fp32 parameter block: W1[64][32] b1[64] W2[64][64] b2[64] W3[64] b3 */
#define SG_P_W1 0
#define SG_P_B1 (SG_P_W1 + SG_H * SG_IN)
#define SG_P_W2 (SG_P_B1 + SG_H)
#define SG_P_B2 (SG_P_W2 + SG_H * SG_H)
#define SG_P_W3 (SG_P_B2 + SG_H)
#define SG_P_B3 (SG_P_W3 + SG_H)
#define SG_NPARAM (SG_P_B3 + 1)

/* bf16 A-operand fragment sets (64 lanes x 8 bf16 each) */
#define SG_FRAG_F1 0  /* W1, 4 row tiles                      */
#define SG_FRAG_F2 4  /* W2, 4 row tiles x 2 k-steps, permuted */
#define SG_FRAG_G2 12 /* W2^T, 4 x 2, permuted                 */
#define SG_FRAG_G1 20 /* W1^T, 2 x 2, permuted                 */
#define SG_NFRAG 24

/* macrocells of the surrogate renderer: 4^3 voxels (the plain marches use
 * 16^3), so that fewer samples that cannot contribute reach the network: a
 * network sample costs tens of times a plain one, the finer cells 5 bytes
 * per 64 voxels (half-precision range, per-frame code) */
#define SG_CELL_SHIFT 2

/* bytes allocated past the half-precision tables: the encoding reads every
 * x-neighbour pair of corners with one 8-byte load, which at the last entry
 * of the last table reads 4 bytes beyond it (and discards them) */
#define SG_T16_TAIL 16

struct ivr_surrogate {
    ivr_surrogate_params sp;
    unsigned T = 0;
    /* surrogate bricks: edge B = 2^bshift voxels (the volume's brick edge << sp.log2_group),
     * nb per axis, brick (bx,by,bz) owning floor indices [b*B, b*B+B) like a volume brick */
    int nbricks = 0, B = 0, bshift = 0;
    int n[3] = {0, 0, 0}, nb[3] = {0, 0, 0};
    bool frozen = false; /* ivr_surrogate_freeze: fp32 tables and optimiser state released */
    int res[SG_L] = {};
    unsigned dense_n1[SG_L] = {};
    size_t ntable = 0;
    float *tables = nullptr, *gtables = nullptr, *mt = nullptr, *vt = nullptr;
    void *tables16 = nullptr; /* half-precision copy of tables, read by rendering */
    float *dE = nullptr;      /* training: encoding gradients of two chunks of samples */
    size_t dE_n = 0;          /* samples one chunk of dE holds */
    /* training: the table scatter of a chunk runs on side while the next
     * chunk's network pass runs on the context's stream */
    hipStream_t side = nullptr;
    hipEvent_t ev_net[2] = {nullptr, nullptr}, ev_sc[2] = {nullptr, nullptr};
    float *params = nullptr, *gparams = nullptr, *mp = nullptr, *vp = nullptr;
    void *frag = nullptr;
    float *d_loss = nullptr;
    unsigned char *d_fallback = nullptr;
    int step = 0;
    /* value range of every 4^3 macrocell (SG_CELL_SHIFT) of the trained
     * volume, as half (min, max) rounded outward (4 bytes per 64 voxels):
     * the surrogate skips cells that the frame's transfer function makes
     * transparent, as the plain renderer does with its 16^3 cells */
    int nc[3] = {0, 0, 0};
    size_t ncells = 0;
    __half2 *d_cellrange = nullptr;
    unsigned char *d_cellempty = nullptr;
};

/* is v the volume (dims and brick edge) that s was made for */
bool ivr_surrogate_matches(const ivr_surrogate *s, const ivr_volume *v);

/* a surrogate frame on stream st with the caller's macrocell buffer (ncells
 * bytes, used with skipping) and per-brick fallback flags (device, nbricks
 * bytes; vol must be given when one is set); fr's counters are not cleared */
int ivr_surrogate_launch_on(ivr_ctx *ctx, hipStream_t st, const ivr_surrogate *s, const ivr_volume *vol,
                            const ivr_tf *tf, const ivr_frame_desc *d, ivr_frame *fr, unsigned char *cells,
                            const unsigned char *fallback);

/* several surrogate frames in one launch: ivr_surrogate_prepare runs the
 * frame's macrocell pass on st and fills an item (ivr_surrogate_item_bytes()
 * bytes, trivially copyable) with its workgroup count and LDS bytes; the
 * caller copies the items and the prefix sums of their workgroup counts
 * (nitems + 1 ints) to the device and calls ivr_surrogate_launch_items with
 * the largest LDS size.  Each frame is the one ivr_surrogate_launch_on makes. */
size_t ivr_surrogate_item_bytes(void);
int ivr_surrogate_prepare(ivr_ctx *ctx, hipStream_t st, const ivr_surrogate *s, const ivr_volume *vol,
                          const ivr_tf *tf, const ivr_frame_desc *d, ivr_frame *fr, unsigned char *cells,
                          const unsigned char *fallback, void *item, int *nblocks, size_t *lds);
int ivr_surrogate_launch_items(hipStream_t st, const void *d_items, const int *d_prefix, int nitems, int total,
                               size_t lds);

#endif
