/*
 * batch.hip -- continuous batching of steering requests (ivr_batcher_*).
 *
 * The reference render server handles one STATE at a time: receive, render
 * all pixels, composite, send (main.c:206-250).  Here requests of different
 * sizes, cameras, transfer functions and resident volumes share persistent
 * launches: k_render_batch's workgroups pull pixel blocks from XCD-affine
 * queues over the concatenated block lists of every request in the batch,
 * stage that request's arguments, LUT and brick table in LDS and render the
 * block with the same config-mode march as the single-frame kernel, so each
 * image is bit-identical alone and in a mix.
 *
 * Host side: requests wait in a FIFO, are admitted into a slot while one is
 * free and the HBM budget allows, and the admitted-but-unlaunched slots
 * form the next batch.  Two batch buffers alternate on two streams, so a
 * batch formed while another runs overlaps it on the device; finished
 * batches are retired by polling their events.
 */
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <vector>

#include "render_dev.h"
#include "surrogate.h"

namespace {

using namespace ivr_dev;

#define BATCH_LUT 256

/* The batch's pixel blocks (16 x 16 for fp32 volumes, 64 x 4 for fp16, with
 * the wavefront shapes of the single-frame kernel) are numbered over the
 * concatenated requests and dealt to 8 queues in runs of 8 consecutive
 * blocks, queue q holding runs q, q + 8, ...; a workgroup takes blocks from
 * the queue of its XCD (workgroups are dealt to the XCDs round robin) and,
 * when that is empty, from the others in turn, so each L2 sees runs of
 * neighbouring blocks while every XCD keeps busy.  A block's request
 * arguments, LUT and brick table are staged in LDS when the request
 * changes.  Every workgroup leaves once all eight queues are exhausted. */
__device__ __forceinline__ long long queue_block(int q, unsigned k)
{
    return ((long long)(k >> 3) * 8 + q) * 8 + (k & 7);
}

__global__ __launch_bounds__(IVR_BLOCK) void k_render_batch(const RenderArgs *args, const int *prefix, int nslots,
                                                            int total, unsigned *queues)
{
    __shared__ float4 s_lut[BATCH_LUT];
    __shared__ unsigned long long s_boff[IVR_MAX_BRICKS / 8];
    __shared__ RenderArgs s_a;
    __shared__ int s_tile, s_slot;
    int staged = -1;
    const int xcd = (int)(blockIdx.x & 7);
    for (int qi = 0; qi < 8; qi++) {
        const int q = (xcd + qi) & 7;
        for (;;) {
            if (threadIdx.x == 0) {
                const long long t = queue_block(q, atomicAdd(&queues[q], 1u));
                int sl = 0;
                if (t < total)
                    while (sl + 1 < nslots && prefix[sl + 1] <= (int)t) sl++;
                s_tile = t < total ? (int)t : -1;
                s_slot = sl;
            }
            __syncthreads();
            const int tile = s_tile, slot = s_slot;
            if (tile < 0) {
                __syncthreads(); /* every thread has read s_tile before it is written again */
                break;
            }
            if (slot != staged) {
                const unsigned *src = (const unsigned *)&args[slot];
                unsigned *dst = (unsigned *)&s_a;
                for (int w = threadIdx.x; w < (int)(sizeof(RenderArgs) / 4); w += blockDim.x) dst[w] = src[w];
                __syncthreads();
                for (int t = threadIdx.x; t < s_a.lutsize; t += blockDim.x) s_lut[t] = s_a.lut[t];
                for (int t = threadIdx.x; t < s_a.nbricks; t += blockDim.x) s_boff[t] = s_a.boff[t];
                __syncthreads();
                staged = slot;
            }
            const RenderArgs &a = s_a;
            int i, j;
            block_pixel(a, tile - prefix[slot], i, j); /* a.xcd_chunk = 0: blocks in order */
            Acc acc = {0.0f, 0.0f, 0.0f, 0.0f};
            unsigned cnt = 0, fetched = 0;
            if (i < a.H && j < a.W) {
                float d[3];
                cfg_dir(a, i, j, d);
                const bool skip = a.cellempty != nullptr;
                if (a.dtype_f16) {
                    if (skip) march_config<_Float16, true>(a, s_lut, s_boff, a.o, d, acc, cnt, fetched);
                    else march_config<_Float16, false>(a, s_lut, s_boff, a.o, d, acc, cnt, fetched);
                } else {
                    if (skip) march_config<float, true>(a, s_lut, s_boff, a.o, d, acc, cnt, fetched);
                    else march_config<float, false>(a, s_lut, s_boff, a.o, d, acc, cnt, fetched);
                }
                a.out8[(size_t)i * a.W + j] = make_uchar4(quant8(acc.r), quant8(acc.g), quant8(acc.b), quant8(acc.a));
            }
            count_add(a.counters, 0, cnt);
            count_add(a.counters, 1, fetched);
            __syncthreads();
        }
    }
}

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Slot {
    bool busy = false, launched = false;
    ivr_request req;
    std::vector<float> lut;
    double t_admit = 0.0;
    unsigned long long charge = 0;
    uchar4 *d_img = nullptr;
    float4 *d_lut = nullptr;
    unsigned long long *d_ctr = nullptr;
    unsigned char *d_cells = nullptr;
    size_t cells_cap = 0;
    unsigned char *h_img = nullptr;       /* pinned */
    unsigned long long *h_ctr = nullptr;  /* pinned */
    float4 *h_lut = nullptr;              /* pinned */
};

/* a servable resident: a plain volume, or a surrogate with its fallback
 * flags (snapshot) and the optional plain volume of its fallback bricks */
struct Resident {
    const ivr_volume *vol = nullptr;
    const ivr_surrogate *sg = nullptr;
    unsigned char *d_fallback = nullptr;
    size_t ncells() const { return sg ? sg->ncells : vol->ncells; }
};

struct Batch {
    bool active = false;
    int stream = 0;
    hipEvent_t done = nullptr;
    std::vector<int> slots;
    RenderArgs *d_args = nullptr, *h_args = nullptr;
    int *d_prefix = nullptr, *h_prefix = nullptr;
    unsigned *d_counter = nullptr; /* the 8 block queues of k_render_batch */
    /* surrogate requests of the batch: their frame items, rendered by one launch */
    unsigned char *d_items = nullptr, *h_items = nullptr;
    int *d_sgprefix = nullptr, *h_sgprefix = nullptr;
};

} // namespace

struct ivr_batcher {
    ivr_ctx *ctx = nullptr;
    unsigned long long budget = 0, used = 0;
    unsigned long long base = 0; /* bytes of the resident volumes */
    int max_slots = 0, max_w = 0, max_h = 0;
    hipStream_t streams[2] = {};
    std::vector<Resident> res;
    std::vector<Slot> slots;
    std::deque<Slot> pending; /* requests with their LUT copies */
    Batch batches[2];
    int next_batch = 0;
    std::deque<ivr_completion> done;
    ivr_batcher_stats st = {};
};

double ivr_now(void) { return now_s(); }

namespace {
int batcher_init(ivr_batcher *b);
}

int ivr_batcher_create(ivr_ctx *ctx, unsigned long long budget, int max_slots, int max_w, int max_h,
                       ivr_batcher **out)
{
    if (!ctx || !out || max_slots < 1 || max_slots > 256 || max_w < 1 || max_h < 1)
        return ivr_fail(IVR_ERR_ARG, "bad batcher arguments");
    ivr_batcher *b = new ivr_batcher();
    b->ctx = ctx;
    b->budget = budget;
    b->max_slots = max_slots;
    b->max_w = max_w;
    b->max_h = max_h;
    b->st.budget = budget;
    const int rc = batcher_init(b);
    if (rc) {
        ivr_batcher_destroy(b); /* releases what was allocated */
        return rc;
    }
    *out = b;
    return IVR_OK;
}

namespace {

/* streams, slot and batch buffers of a new batcher */
int batcher_init(ivr_batcher *b)
{
    ivr_ctx *ctx = b->ctx;
    const int max_slots = b->max_slots;
    for (auto &s : b->streams) IVR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    b->slots.resize(max_slots);
    const size_t img = (size_t)b->max_w * b->max_h * 4;
    int rc;
    for (Slot &s : b->slots) {
        if ((rc = ivr_dalloc(ctx, (void **)&s.d_img, img)) || (rc = ivr_dalloc(ctx, (void **)&s.d_lut, BATCH_LUT * 16)) ||
            (rc = ivr_dalloc(ctx, (void **)&s.d_ctr, IVR_CTR_BYTES)))
            return rc;
        IVR_HIP(hipHostMalloc((void **)&s.h_img, img, hipHostMallocDefault));
        IVR_HIP(hipHostMalloc((void **)&s.h_ctr, IVR_CTR_BYTES, hipHostMallocDefault));
        IVR_HIP(hipHostMalloc((void **)&s.h_lut, BATCH_LUT * 16, hipHostMallocDefault));
    }
    for (int k = 0; k < 2; k++) {
        Batch &bt = b->batches[k];
        bt.stream = k;
        IVR_HIP(hipEventCreateWithFlags(&bt.done, hipEventDisableTiming));
        if ((rc = ivr_dalloc(ctx, (void **)&bt.d_args, sizeof(RenderArgs) * max_slots)) ||
            (rc = ivr_dalloc(ctx, (void **)&bt.d_prefix, sizeof(int) * (max_slots + 1))) ||
            (rc = ivr_dalloc(ctx, (void **)&bt.d_counter, 8 * sizeof(unsigned))))
            return rc;
        IVR_HIP(hipHostMalloc((void **)&bt.h_args, sizeof(RenderArgs) * max_slots, hipHostMallocDefault));
        IVR_HIP(hipHostMalloc((void **)&bt.h_prefix, sizeof(int) * (max_slots + 1), hipHostMallocDefault));
        const size_t ib = ivr_surrogate_item_bytes() * max_slots;
        if ((rc = ivr_dalloc(ctx, (void **)&bt.d_items, ib)) ||
            (rc = ivr_dalloc(ctx, (void **)&bt.d_sgprefix, sizeof(int) * (max_slots + 1))))
            return rc;
        IVR_HIP(hipHostMalloc((void **)&bt.h_items, ib, hipHostMallocDefault));
        IVR_HIP(hipHostMalloc((void **)&bt.h_sgprefix, sizeof(int) * (max_slots + 1), hipHostMallocDefault));
    }
    return IVR_OK;
}

} // namespace

int ivr_batcher_destroy(ivr_batcher *b)
{
    if (!b) return IVR_OK;
    for (auto &s : b->streams)
        if (s) (void)hipStreamSynchronize(s);
    const size_t img = (size_t)b->max_w * b->max_h * 4;
    for (Slot &s : b->slots) {
        ivr_dfree(b->ctx, s.d_img, img);
        ivr_dfree(b->ctx, s.d_lut, BATCH_LUT * 16);
        ivr_dfree(b->ctx, s.d_ctr, IVR_CTR_BYTES);
        ivr_dfree(b->ctx, s.d_cells, s.cells_cap);
        if (s.h_img) (void)hipHostFree(s.h_img);
        if (s.h_ctr) (void)hipHostFree(s.h_ctr);
        if (s.h_lut) (void)hipHostFree(s.h_lut);
    }
    for (Batch &bt : b->batches) {
        if (bt.done) (void)hipEventDestroy(bt.done);
        ivr_dfree(b->ctx, bt.d_args, sizeof(RenderArgs) * b->max_slots);
        ivr_dfree(b->ctx, bt.d_prefix, sizeof(int) * (b->max_slots + 1));
        ivr_dfree(b->ctx, bt.d_counter, 8 * sizeof(unsigned));
        if (bt.h_args) (void)hipHostFree(bt.h_args);
        if (bt.h_prefix) (void)hipHostFree(bt.h_prefix);
        ivr_dfree(b->ctx, bt.d_items, ivr_surrogate_item_bytes() * b->max_slots);
        ivr_dfree(b->ctx, bt.d_sgprefix, sizeof(int) * (b->max_slots + 1));
        if (bt.h_items) (void)hipHostFree(bt.h_items);
        if (bt.h_sgprefix) (void)hipHostFree(bt.h_sgprefix);
    }
    for (auto &s : b->streams)
        if (s) (void)hipStreamDestroy(s);
    for (Resident &r : b->res)
        if (r.sg) ivr_dfree(b->ctx, r.d_fallback, r.sg->nbricks);
    delete b;
    return IVR_OK;
}

int ivr_batcher_add_volume(ivr_batcher *b, const ivr_volume *vol)
{
    if (!b || !vol) return ivr_fail(IVR_ERR_ARG, "null argument");
    const unsigned long long bytes = vol->bytes + vol->ncells * 9ull;
    if (b->used + bytes > b->budget)
        return ivr_fail(IVR_ERR_BUDGET, "volume of %llu bytes exceeds the HBM budget", bytes);
    if (vol->nbricks > IVR_MAX_BRICKS / 8) return ivr_fail(IVR_ERR_UNSUPPORTED, "too many bricks to batch");
    b->used += bytes;
    b->base += bytes;
    Resident r;
    r.vol = vol;
    b->res.push_back(r);
    b->st.bytes_in_use = b->used;
    return (int)b->res.size() - 1;
}

int ivr_batcher_add_surrogate(ivr_batcher *b, const ivr_surrogate *sg, const ivr_volume *vol)
{
    if (!b || !sg) return ivr_fail(IVR_ERR_ARG, "null argument");
    if (vol && !ivr_surrogate_matches(sg, vol)) return ivr_fail(IVR_ERR_ARG, "volume does not match surrogate");
    std::vector<unsigned char> h(sg->nbricks);
    IVR_HIP(hipMemcpy(h.data(), sg->d_fallback, sg->nbricks, hipMemcpyDeviceToHost));
    for (unsigned char c : h)
        if (c && !vol) return ivr_fail(IVR_ERR_ARG, "a brick falls back to the plain volume, which was not given");
    /* the network's resident bytes: the half-precision tables it renders from (plus, unless
     * frozen, the fp32 tables, their gradients and Adam moments), parameters, operand
     * fragments, cell ranges */
    const unsigned long long bytes = sg->ntable * (sg->frozen ? 2ull : 18ull) +
                                     SG_NPARAM * (sg->frozen ? 4ull : 16ull) + SG_NFRAG * 1024ull +
                                     sg->ncells * 5ull + sg->nbricks;
    if (b->used + bytes > b->budget)
        return ivr_fail(IVR_ERR_BUDGET, "surrogate of %llu bytes exceeds the HBM budget", bytes);
    Resident r;
    r.vol = vol;
    r.sg = sg;
    int rc = ivr_dalloc(b->ctx, (void **)&r.d_fallback, sg->nbricks);
    if (rc) return rc;
    IVR_HIP(hipMemcpy(r.d_fallback, h.data(), sg->nbricks, hipMemcpyHostToDevice));
    b->used += bytes;
    b->base += bytes;
    b->res.push_back(r);
    b->st.bytes_in_use = b->used;
    return (int)b->res.size() - 1;
}

namespace {

void reject(ivr_batcher *b, const ivr_request &r, int code)
{
    ivr_completion c = {};
    c.id = r.id;
    c.status = code;
    c.t_submit = r.t_submit;
    c.t_done = now_s();
    b->done.push_back(c);
    b->st.rejected++;
}

unsigned long long charge_of(const ivr_batcher *b, const ivr_request &r)
{
    return (unsigned long long)r.desc.state.winwidth * r.desc.state.winheight * 4ull + BATCH_LUT * 16ull +
           sizeof(RenderArgs) + (r.desc.skip ? b->res[r.volume].ncells() : 0ull);
}

int retire(ivr_batcher *b, Batch &bt, bool wait)
{
    if (!bt.active) return 0;
    if (wait) IVR_HIP(hipEventSynchronize(bt.done));
    else {
        const hipError_t e = hipEventQuery(bt.done);
        if (e == hipErrorNotReady) return 0;
        if (e != hipSuccess) return ivr_fail(IVR_ERR_HIP, "batch event: %s", hipGetErrorString(e));
    }
    const double t = now_s();
    for (int si : bt.slots) {
        Slot &s = b->slots[si];
        const size_t bytes = (size_t)s.req.desc.state.winwidth * s.req.desc.state.winheight * 4;
        memcpy(s.req.rgba8, s.h_img, bytes);
        ivr_completion c = {};
        c.id = s.req.id;
        c.status = 0;
        c.t_submit = s.req.t_submit;
        c.t_admit = s.t_admit;
        c.t_done = t;
        c.samples = 0;
        for (int q = 0; q < IVR_CTR_SLOTS; q++) c.samples += s.h_ctr[q * IVR_CTRS];
        b->done.push_back(c);
        b->used -= s.charge;
        s.busy = false;
        s.launched = false;
        b->st.completed++;
    }
    bt.slots.clear();
    bt.active = false;
    b->st.in_flight--;
    b->st.bytes_in_use = b->used;
    return 0;
}

int admit(ivr_batcher *b)
{
    while (!b->pending.empty()) {
        Slot &p = b->pending.front();
        int free_slot = -1;
        for (int k = 0; k < b->max_slots; k++)
            if (!b->slots[k].busy) { free_slot = k; break; }
        if (free_slot < 0) break;
        const unsigned long long ch = charge_of(b, p.req);
        if (b->base + ch > b->budget) { /* could never fit: reject, keep the FIFO moving */
            reject(b, p.req, IVR_ERR_BUDGET);
            b->pending.pop_front();
            continue;
        }
        if (b->used + ch > b->budget) break; /* FIFO: wait for memory to free up */
        Slot &s = b->slots[free_slot];
        const size_t nc = b->res[p.req.volume].ncells();
        if (p.req.desc.skip && s.cells_cap < nc) {
            ivr_dfree(b->ctx, s.d_cells, s.cells_cap);
            s.cells_cap = 0;
            int rc = ivr_dalloc(b->ctx, (void **)&s.d_cells, nc);
            if (rc) return rc;
            s.cells_cap = nc;
        }
        s.req = p.req;
        s.lut.swap(p.lut);
        s.busy = true;
        s.launched = false;
        s.charge = ch;
        s.t_admit = now_s();
        b->used += ch;
        b->pending.pop_front();
        b->st.admitted++;
        b->st.bytes_in_use = b->used;
    }
    b->st.pending = (int)b->pending.size();
    return 0;
}

/* enqueue the batch of slots `ready` on bt's stream (see launch) */
int launch_slots(ivr_batcher *b, Batch *bt, const std::vector<int> &ready);

/* Form the next batch from the admitted, unlaunched slots and enqueue it.
 * When enqueueing fails part way, the batch's stream is drained, each of its
 * requests completes with the error status and its slot is freed, so no
 * client waits for a frame that will never come; the batcher stays usable. */
int launch(ivr_batcher *b)
{
    std::vector<int> ready;
    for (int k = 0; k < b->max_slots; k++)
        if (b->slots[k].busy && !b->slots[k].launched) ready.push_back(k);
    if (ready.empty()) return 0;
    Batch *bt = nullptr;
    for (int q = 0; q < 2; q++) {
        Batch &c = b->batches[(b->next_batch + q) & 1];
        if (!c.active) { bt = &c; break; }
    }
    if (!bt) return 0;
    b->next_batch = (bt->stream + 1) & 1;
    /* FIFO order of admission */
    std::sort(ready.begin(), ready.end(), [&](int x, int y) { return b->slots[x].t_admit < b->slots[y].t_admit; });
    const int rc = launch_slots(b, bt, ready);
    if (!rc) return 0;
    (void)hipStreamSynchronize(b->streams[bt->stream]); /* nothing of it may still write a slot */
    (void)hipGetLastError();
    const double t = now_s();
    for (int si : ready) {
        Slot &s = b->slots[si];
        ivr_completion c = {};
        c.id = s.req.id;
        c.status = rc;
        c.t_submit = s.req.t_submit;
        c.t_admit = s.t_admit;
        c.t_done = t;
        b->done.push_back(c);
        b->used -= s.charge;
        s.busy = false;
        s.launched = false;
        b->st.rejected++;
    }
    bt->slots.clear();
    bt->active = false;
    b->st.bytes_in_use = b->used;
    return 0;
}

int launch_slots(ivr_batcher *b, Batch *bt, const std::vector<int> &ready)
{
    hipStream_t st = b->streams[bt->stream];
    int total = 0, nplain = 0, nsg = 0, sgtotal = 0;
    size_t sglds = 0;
    const size_t ib = ivr_surrogate_item_bytes();
    bt->slots = ready;
    for (size_t q = 0; q < ready.size(); q++) {
        Slot &s = b->slots[ready[q]];
        const Resident &res = b->res[s.req.volume];
        const ivr_volume *v = res.vol;
        const ivr_frame_desc &d = s.req.desc;
        memcpy(s.h_lut, s.lut.data(), s.lut.size() * sizeof(float));
        IVR_HIP(hipMemcpyAsync(s.d_lut, s.h_lut, s.lut.size() * sizeof(float), hipMemcpyHostToDevice, st));
        ivr_tf tf;
        tf.n = (int)(s.lut.size() / 4);
        tf.d_lut = s.d_lut;
        ivr_frame fr;
        fr.W = d.state.winwidth;
        fr.H = d.state.winheight;
        fr.d_rgba8 = s.d_img;
        fr.d_rgba32 = nullptr;
        fr.d_counters = s.d_ctr;
        if (res.sg) { /* the batch's surrogate frames share one launch ahead of the batch kernel */
            IVR_HIP(hipMemsetAsync(s.d_ctr, 0, IVR_CTR_BYTES, st));
            int nblocks = 0;
            size_t lds = 0;
            int rc = ivr_surrogate_prepare(b->ctx, st, res.sg, v, &tf, &d, &fr, d.skip ? s.d_cells : nullptr,
                                           res.d_fallback, bt->h_items + ib * nsg, &nblocks, &lds);
            if (rc) return rc;
            bt->h_sgprefix[nsg++] = sgtotal;
            sgtotal += nblocks;
            sglds = std::max(sglds, lds);
            s.launched = true;
            continue;
        }
        const int q2 = nplain++; /* plain requests fill the batch's argument table in order */
        ivr_cam cam;
        ivr_cam_setup(&d.state, v->n[0], v->n[1], v->n[2], &cam);
        RenderArgs &a = bt->h_args[q2];
        ivr_fill_args(a, cam, v, &tf, &d, &fr);
        ivr_set_grid(a, b->ctx, fr.W, fr.H); /* the single-frame kernel's block and wavefront shapes */
        a.xcd_chunk = 0;                     /* blocks come from the queues in order */
        a.cellempty = d.skip ? s.d_cells : nullptr;
        if (d.skip) {
            int rc = ivr_launch_cells_on(st, v, a, s.d_cells);
            if (rc) return rc;
        }
        IVR_HIP(hipMemsetAsync(s.d_ctr, 0, IVR_CTR_BYTES, st));
        bt->h_prefix[q2] = total;
        total += a.nblk;
        s.launched = true;
    }
    if (nsg > 0) {
        bt->h_sgprefix[nsg] = sgtotal;
        IVR_HIP(hipMemcpyAsync(bt->d_items, bt->h_items, ib * nsg, hipMemcpyHostToDevice, st));
        IVR_HIP(hipMemcpyAsync(bt->d_sgprefix, bt->h_sgprefix, sizeof(int) * (nsg + 1), hipMemcpyHostToDevice, st));
        int rc = ivr_surrogate_launch_items(st, bt->d_items, bt->d_sgprefix, nsg, sgtotal, sglds);
        if (rc) return rc;
    }
    if (nplain > 0) {
        bt->h_prefix[nplain] = total;
        IVR_HIP(hipMemcpyAsync(bt->d_args, bt->h_args, sizeof(RenderArgs) * nplain, hipMemcpyHostToDevice, st));
        IVR_HIP(hipMemcpyAsync(bt->d_prefix, bt->h_prefix, sizeof(int) * (nplain + 1), hipMemcpyHostToDevice, st));
        IVR_HIP(hipMemsetAsync(bt->d_counter, 0, 8 * sizeof(unsigned), st));
        const int grid = std::min(total, 2048);
        hipLaunchKernelGGL(k_render_batch, dim3(grid), dim3(IVR_BLOCK), 0, st, bt->d_args, bt->d_prefix, nplain,
                           total, bt->d_counter);
        IVR_HIP(hipGetLastError());
    }
    for (int si : ready) {
        Slot &s = b->slots[si];
        const size_t bytes = (size_t)s.req.desc.state.winwidth * s.req.desc.state.winheight * 4;
        IVR_HIP(hipMemcpyAsync(s.h_img, s.d_img, bytes, hipMemcpyDeviceToHost, st));
        IVR_HIP(hipMemcpyAsync(s.h_ctr, s.d_ctr, IVR_CTR_BYTES, hipMemcpyDeviceToHost, st));
    }
    IVR_HIP(hipEventRecord(bt->done, st));
    bt->active = true;
    b->st.batches++;
    b->st.in_flight++;
    return 0;
}

} // namespace

int ivr_batcher_submit(ivr_batcher *b, const ivr_request *r)
{
    if (!b || !r) return ivr_fail(IVR_ERR_ARG, "null argument");
    b->st.submitted++;
    const ivr_frame_desc &d = r->desc;
    int code = 0;
    if (r->volume < 0 || r->volume >= (int)b->res.size()) code = IVR_ERR_ARG;
    else if (d.mode != IVR_MODE_CONFIG || d.composite != IVR_FUSED || !(d.step_vox > 0.0f)) code = IVR_ERR_UNSUPPORTED;
    else if (d.state.winwidth < 1 || d.state.winheight < 1 || d.state.winwidth > b->max_w ||
             d.state.winheight > b->max_h)
        code = IVR_ERR_ARG;
    else if (!r->lut || r->lutsize < 1 || r->lutsize > BATCH_LUT || !r->rgba8) code = IVR_ERR_ARG;
    else if (b->base + charge_of(b, *r) > b->budget) code = IVR_ERR_BUDGET; /* never fits */
    if (code) {
        reject(b, *r, code);
        return IVR_OK;
    }
    Slot p;
    p.req = *r;
    p.lut.assign(r->lut, r->lut + 4 * r->lutsize);
    b->pending.push_back(std::move(p));
    b->st.pending = (int)b->pending.size();
    return IVR_OK;
}

int ivr_batcher_progress(ivr_batcher *b, int block)
{
    if (!b) return ivr_fail(IVR_ERR_ARG, "null batcher");
    int rc;
    for (Batch &bt : b->batches)
        if ((rc = retire(b, bt, false))) return rc;
    if ((rc = admit(b)) || (rc = launch(b))) return rc;
    if (block && b->done.empty()) {
        /* wait for the oldest in-flight batch, then refill */
        Batch *oldest = nullptr;
        for (Batch &bt : b->batches)
            if (bt.active && (!oldest || bt.stream == b->next_batch)) oldest = &bt;
        if (oldest) {
            if ((rc = retire(b, *oldest, true))) return rc;
            if ((rc = admit(b)) || (rc = launch(b))) return rc;
        }
    }
    return (int)b->done.size();
}

int ivr_batcher_poll(ivr_batcher *b, ivr_completion *out, int max)
{
    if (!b || !out) return ivr_fail(IVR_ERR_ARG, "null argument");
    int n = 0;
    while (n < max && !b->done.empty()) {
        out[n++] = b->done.front();
        b->done.pop_front();
    }
    return n;
}

int ivr_batcher_stats_get(const ivr_batcher *b, ivr_batcher_stats *st)
{
    if (!b || !st) return ivr_fail(IVR_ERR_ARG, "null argument");
    *st = b->st;
    st->pending = (int)b->pending.size();
    return IVR_OK;
}
