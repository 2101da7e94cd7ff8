/*
 * ivrserver.c -- steering render server on the MI355X render path.
 *
 * Speaks the reference frame protocol (connection.c:61-216, myserver.c):
 *   handshake  the accepting side sends "Hello, world!" (13 bytes), the
 *              connecting side answers "Hello, server!" (14 bytes);
 *   request    a 184-byte STATE, its three ints in network byte order, the
 *              floats native little-endian (connection.c:167-200);
 *   reply      winwidth*winheight*4 bytes of RGBA8, bottom row first
 *              (main.c:1404-1405);
 *   finish     a STATE with finish != 0 ends the session (main.c:211).
 * Two roles:
 *   --relay HOST:PORT  connect to the reference relay (myserver) exactly as
 *                      the reference render server does and serve its one
 *                      client (drop-in for bin/ivrserver);
 *   --listen PORT      accept any number of steering clients directly and
 *                      serve them concurrently: every request is submitted
 *                      to the continuous batcher, so frames of different
 *                      clients share persistent launches.
 * Versioned extension (optional, per message): "IVRX", u32 version = 1,
 * u64 request id, i32 volume index, i32 LUT entries n (0: the volume's
 * default), n*16 bytes RGBA float LUT, then the STATE.  Its reply is "IVRR",
 * u64 id, i32 status, i32 width, i32 height, then the RGBA8 bytes (none when
 * status != 0).  All extension integers are little-endian.
 *
 * Compat serving (--compat COLORMAP XFER, or the reference server's own
 * command line, below): every request is rendered as the reference render
 * server renders it (main.c:206-250): ray.c semantics (IVR_MODE_COMPAT),
 * transfer function XFER, the LUT built as initvolluts builds it
 * (main.c:1075-1200: the volume's histogram on the device, colormap
 * COLORMAP over its [min, max], lutoffset = (int)(min + 0.5), maxlutindx =
 * hsize - 1), one STATE at a time; the reply is the RGBA8 of
 * createlocalimage's colour buffer.  An extension message's LUT, when given,
 * replaces the volume's for that request.
 *
 * The reference server's command line is accepted as is:
 *   ivrserver FILE W H D ENDIAN COLORMAP XFER PX PY PZ TILEW TILEH ANGLE AX AY AZ
 * (parseargs, main.c:1047-1073; ENDIAN 1 little, 0 big as initvolattrs
 * compares it, main.c:386-392): compat serving of the raw volume FILE behind
 * the relay at localhost:1400 (main.c:19-23), or at $IVR_RELAY (HOST:PORT);
 * TILEW x TILEH rounds each window down to whole tiles as setviewport does
 * (mathutil.h:147-163) and that viewport is rendered, as the reference
 * server renders it (see render_compat); the decomposition and view arguments have no
 * effect (one GPU holds the whole volume; the client's STATE sets the
 * view).  --tile TW TH sets the rounding in the named form (default 1 1).
 *
 * --swap-floats byte-swaps the STATE's floats on receipt, as the reference's
 * BIGENDIAN build does (connection.c:191-197), for peers of the other byte
 * order.  --tiff DIR (compat) also writes each frame's colour buffer as
 * DIR/frameNNNNN.tiff in writetotiff's format (utils.c:18-154).
 *
 * Volumes (--volume, repeatable):
 *   gen:C:N[:f16]               synthetic generator of benchmark config C at N^3
 *   raw:PATH:NX:NY:NZ:E[:f16]   raw float32 file, x fastest, E = 1 little /
 *                               0 big endian (readvoldata, main.c:629-706;
 *                               BYTESWAP_FP, sysutil.h:121-132); its LUT index
 *                               spans [min, max] of the data.
 */
#include <arpa/inet.h>
#include <errno.h>
#include <netdb.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <poll.h>
#include <signal.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/socket.h>
#include <unistd.h>

#include "../../include/ivr_gpu.h"

#define MAX_CLIENTS 64
#define MAX_VOLS 8
#define EXT_HDR 24 /* magic, version, id, volume, lut entries */

typedef struct {
    ivr_volume *vol;
    float lut[256 * 4];
    float lutscale;
    int lutoffset;
    /* compat serving: the initvolluts LUT on the device */
    ivr_tf *tf;
    int maxlutindx, hsize;
} served_volume;

typedef struct {
    int used;   /* slot holds a client (or a zombie awaiting its frame) */
    int zombie; /* disconnected while a frame was in flight */
    int fd;
    unsigned char *in;
    size_t have, cap;
    int busy;
    int ext;
    unsigned long long ext_id;
    unsigned char *img;
    size_t img_cap;
    int w, h;
    unsigned long long rid;
} client;

static served_volume g_vol[MAX_VOLS];
static int g_nvol = 0;
static client g_cl[MAX_CLIENTS];
static int g_skip = 1;
static unsigned long long g_next_rid = 1;
/* compat serving */
static int g_compat = 0, g_cmap = 0, g_xfer = 3;
static int g_tile_w = 1, g_tile_h = 1; /* compat: the viewport is a whole number of tiles */
static int g_swap_floats = 0;           /* STATE floats of the other byte order (BIGENDIAN) */
static const char *g_tiff_dir = NULL;   /* compat: each frame also written as a TIFF */
static int g_tiff_count = 0;
static ivr_ctx *g_ctx = NULL;
static ivr_frame *g_frame = NULL; /* compat frames, reused while the size stays */
static int g_fw = 0, g_fh = 0;
static ivr_tf *g_req_tf = NULL;    /* an extension message's own LUT */

static int send_all(int fd, const void *p, size_t n)
{
    const unsigned char *b = (const unsigned char *)p;
    while (n) {
        ssize_t r = send(fd, b, n, MSG_NOSIGNAL);
        if (r <= 0) {
            if (r < 0 && errno == EINTR) continue;
            return -1;
        }
        b += r;
        n -= (size_t)r;
    }
    return 0;
}

static int recv_all(int fd, void *p, size_t n)
{
    unsigned char *b = (unsigned char *)p;
    while (n) {
        ssize_t r = recv(fd, b, n, 0);
        if (r <= 0) {
            if (r < 0 && errno == EINTR) continue;
            return -1;
        }
        b += r;
        n -= (size_t)r;
    }
    return 0;
}

static void nodelay(int fd)
{
    int one = 1;
    setsockopt(fd, IPPROTO_TCP, TCP_NODELAY, &one, sizeof one);
}

static float bswapf(float f)
{
    unsigned char b[4], r[4];
    memcpy(b, &f, 4);
    r[0] = b[3]; r[1] = b[2]; r[2] = b[1]; r[3] = b[0];
    memcpy(&f, r, 4);
    return f;
}

static int add_volume(ivr_ctx *ctx, const char *spec)
{
    if (g_nvol >= MAX_VOLS) return -1;
    served_volume *sv = &g_vol[g_nvol];
    char buf[512];
    snprintf(buf, sizeof buf, "%s", spec);
    int f16 = strstr(spec, ":f16") != NULL;
    int rc;
    if (!strncmp(buf, "gen:", 4)) {
        int c = 0, n = 0;
        if (sscanf(buf + 4, "%d:%d", &c, &n) != 2 || n < 2) return -1;
        int brick = n >= 256 ? 256 : 64;
        if ((rc = ivr_volume_create(ctx, n, n, n, f16 ? IVR_F16 : IVR_F32, brick, &sv->vol))) return rc;
        ivr_gen_params gp;
        memset(&gp, 0, sizeof gp);
        gp.kind = c == 1 ? IVR_GEN_BLOBS : IVR_GEN_NOISE_BLOBS;
        gp.nblobs = c == 1 ? 8 : 32;
        gp.octaves = 4;
        gp.lacunarity = 2.0f;
        gp.gain = 0.5f;
        gp.base_freq = 4.0f;
        gp.empty_frac = c == 3 ? 0.7f : (c == 2 ? 0.4f : 0.0f);
        gp.value_scale = 1.0f;
        gp.seed = 1000 + (uint64_t)c;
        if ((rc = ivr_volume_generate(ctx, sv->vol, &gp))) return rc;
        ivr_tf_build(2000 + (uint64_t)c, 0.05f, 0.0f, 0, sv->lut);
        sv->lutscale = 255.0f;
        sv->lutoffset = 0;
    } else if (!strncmp(buf, "raw:", 4)) {
        char *path = buf + 4, *p = strchr(path, ':');
        int nx, ny, nz, endian;
        if (!p) return -1;
        *p = 0;
        if (sscanf(p + 1, "%d:%d:%d:%d", &nx, &ny, &nz, &endian) != 4) return -1;
        size_t n = (size_t)nx * ny * nz;
        float *data = (float *)malloc(n * sizeof(float));
        FILE *f = fopen(path, "rb");
        if (!data || !f || fread(data, sizeof(float), n, f) != n) {
            fprintf(stderr, "ivrserver: cannot read %s\n", path);
            if (f) fclose(f);
            free(data);
            return -1;
        }
        fclose(f);
        float lo = 0, hi = 0;
        for (size_t i = 0; i < n; i++) {
            if (endian == 0) data[i] = bswapf(data[i]);
            if (i == 0 || data[i] < lo) lo = data[i];
            if (i == 0 || data[i] > hi) hi = data[i];
        }
        int md = nx > ny ? nx : ny;
        md = md > nz ? md : nz;
        int brick = md >= 256 ? 256 : 64;
        if ((rc = ivr_volume_create(ctx, nx, ny, nz, f16 ? IVR_F16 : IVR_F32, brick, &sv->vol)) ||
            (rc = ivr_volume_upload(ctx, sv->vol, data))) {
            free(data);
            return rc;
        }
        free(data);
        ivr_tf_build(2002, 0.05f, 0.0f, 0, sv->lut);
        sv->lutscale = hi > lo ? 255.0f / (hi - lo) : 1.0f;
        sv->lutoffset = (int)(lo * sv->lutscale);
    } else {
        return -1;
    }
    if (g_compat) {
        /* initvolluts: the histogram's range sizes the LUT, colormap over it */
        const float one[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if ((rc = ivr_tf_create(ctx, one, 1, &sv->tf)) ||
            (rc = ivr_volume_luts(ctx, sv->vol, g_cmap, sv->tf, &sv->lutoffset, &sv->maxlutindx, &sv->hsize)))
            return rc;
        fprintf(stderr, "ivrserver: LUT of %d entries (colormap %d), lutoffset %d\n", sv->hsize, g_cmap,
                sv->lutoffset);
    }
    g_nvol++;
    return g_nvol - 1;
}

/* the compat frame's colour buffer as DIR/frameNNNNN.tiff (writetotiff, utils.c:18-154, RGB) */
static int write_tiff(int w, int h)
{
    const size_t np = (size_t)w * h, bytes = 512 + np * 3;
    float *img = (float *)malloc(np * 16);
    unsigned char *out = (unsigned char *)malloc(bytes);
    int rc = (!img || !out) ? IVR_ERR_NOMEM : ivr_frame_download(g_ctx, g_frame, NULL, img);
    if (!rc && ivr_tiff_encode(img, w, h, 3, out, bytes) < 0) rc = IVR_ERR_ARG;
    if (!rc) {
        char path[1024];
        snprintf(path, sizeof path, "%s/frame%05d.tiff", g_tiff_dir, g_tiff_count++);
        FILE *f = fopen(path, "wb");
        if (!f || fwrite(out, 1, bytes, f) != bytes) rc = IVR_ERR_IO;
        if (f) fclose(f);
    }
    free(img);
    free(out);
    return rc;
}

/* One request rendered as the reference render server renders it (compat)
 * into the window's winwidth x winheight RGBA8 bytes.  setviewport
 * (mathutil.h:147-163) rounds the window down to whole tiles and the
 * reference renders and sends that viewport (writetomemory, main.c:1340-1410);
 * for a window of whole tiles these are the window's bytes.  For other
 * windows the reference sends fewer bytes than the relay forwards
 * (myserver.c:147-153 reads winwidth x winheight x 4) and the session stalls;
 * here the viewport's pixels fill the bottom left of the window and the rest
 * is zero, so the protocol keeps going. */
static int render_compat(int vol, const ivr_state *win, const float *lut, int nl, unsigned char *rgba8)
{
    served_volume *sv = &g_vol[vol];
    const int W = win->winwidth, H = win->winheight;
    ivr_state vst = *win;
    vst.winwidth = W / g_tile_w * g_tile_w;
    vst.winheight = H / g_tile_h * g_tile_h;
    const ivr_state *st = &vst;
    const int padded = vst.winwidth != W || vst.winheight != H;
    memset(rgba8, 0, (size_t)W * H * 4);
    if (vst.winwidth == 0 || vst.winheight == 0) return 0;
    int rc;
    if (!g_frame || g_fw != st->winwidth || g_fh != st->winheight) {
        if (g_frame) ivr_frame_destroy(g_ctx, g_frame);
        g_frame = NULL;
        if ((rc = ivr_frame_create(g_ctx, st->winwidth, st->winheight, g_tiff_dir != NULL, &g_frame))) return rc;
        g_fw = st->winwidth;
        g_fh = st->winheight;
    }
    ivr_tf *tf = sv->tf;
    ivr_frame_desc d;
    ivr_desc_config(&d);
    d.state = *st;
    d.mode = IVR_MODE_COMPAT;
    d.xfer = g_xfer;
    d.lutscale = 1.0f;
    d.lutoffset = sv->lutoffset;
    d.maxlutindx = sv->maxlutindx;
    if (lut && nl > 0) {
        if ((rc = g_req_tf ? ivr_tf_update(g_ctx, g_req_tf, lut, nl) : ivr_tf_create(g_ctx, lut, nl, &g_req_tf)))
            return rc;
        tf = g_req_tf;
        d.maxlutindx = nl - 1;
    }
    if ((rc = ivr_render(g_ctx, sv->vol, tf, &d, g_frame))) return rc;
    if (g_tiff_dir && (rc = write_tiff(st->winwidth, st->winheight))) return rc;
    if (!padded) return ivr_frame_download(g_ctx, g_frame, rgba8, NULL);
    const size_t row = (size_t)vst.winwidth * 4;
    unsigned char *view = (unsigned char *)malloc(row * vst.winheight);
    if (!view) return IVR_ERR_NOMEM;
    rc = ivr_frame_download(g_ctx, g_frame, view, NULL);
    for (int y = 0; y < vst.winheight && !rc; y++) memcpy(rgba8 + (size_t)y * W * 4, view + (size_t)y * row, row);
    free(view);
    return rc;
}

static client *new_client(int fd)
{
    for (int i = 0; i < MAX_CLIENTS; i++)
        if (!g_cl[i].used) {
            memset(&g_cl[i], 0, sizeof g_cl[i]);
            g_cl[i].used = 1;
            g_cl[i].fd = fd;
            return &g_cl[i];
        }
    return NULL;
}

/* close a client; its image buffer lives on until an in-flight frame lands */
static void drop_client(client *c)
{
    if (c->fd > 0) close(c->fd);
    c->fd = -1;
    free(c->in);
    c->in = NULL;
    c->have = c->cap = 0;
    if (c->busy) {
        c->zombie = 1;
        return;
    }
    free(c->img);
    memset(c, 0, sizeof *c);
}

/* bytes of the next complete message in c->in, 0 if incomplete */
static size_t message_len(const client *c)
{
    if (c->have >= 4 && !memcmp(c->in, "IVRX", 4)) {
        if (c->have < EXT_HDR) return 0;
        int32_t nl;
        memcpy(&nl, c->in + 20, 4);
        if (nl < 0 || nl > 256) return (size_t)-1;
        size_t need = EXT_HDR + (size_t)nl * 16 + 184;
        return c->have >= need ? need : 0;
    }
    return c->have >= 184 ? 184 : 0;
}

static int reply(client *c, int status)
{
    if (c->ext) {
        unsigned char h[24];
        int32_t st = status, w = c->w, hh = c->h;
        memcpy(h, "IVRR", 4);
        memcpy(h + 4, &c->ext_id, 8);
        memcpy(h + 12, &st, 4);
        memcpy(h + 16, &w, 4);
        memcpy(h + 20, &hh, 4);
        if (send_all(c->fd, h, sizeof h)) return -1;
        if (status) return 0;
    } else if (status) {
        /* the plain protocol has no error reply: send a black frame */
        memset(c->img, 0, (size_t)c->w * c->h * 4);
    }
    return send_all(c->fd, c->img, (size_t)c->w * c->h * 4);
}

/* parse one message; submit it; returns 1 on finish, 0 submitted, -1 error */
static int handle(ivr_batcher *bt, client *c, size_t len)
{
    const unsigned char *m = c->in;
    int vol = 0, nl = 0;
    const float *lut = NULL;
    c->ext = 0;
    if (len != 184) {
        uint32_t ver;
        memcpy(&ver, m + 4, 4);
        if (ver != 1) return -1;
        c->ext = 1;
        memcpy(&c->ext_id, m + 8, 8);
        memcpy(&vol, m + 16, 4);
        memcpy(&nl, m + 20, 4);
        lut = nl ? (const float *)(m + EXT_HDR) : NULL;
        m += EXT_HDR + (size_t)nl * 16;
    }
    ivr_state st;
    ivr_state_unpack(m, &st);
    if (g_swap_floats) ivr_state_swap_floats(&st);
    if (st.finish) return 1;
    c->w = st.winwidth;
    c->h = st.winheight;
    if (c->w < 1 || c->h < 1 || c->w > 8192 || c->h > 8192) return -1;
    size_t need = (size_t)c->w * c->h * 4;
    if (need > c->img_cap) {
        free(c->img);
        c->img = (unsigned char *)malloc(need);
        c->img_cap = c->img ? need : 0;
        if (!c->img) return -1;
    }
    if (g_compat) { /* one STATE at a time, as the reference server loop (main.c:206-250) */
        const int v = (vol >= 0 && vol < g_nvol) ? vol : -1;
        int rc = v < 0 ? IVR_ERR_ARG : render_compat(v, &st, lut, nl, c->img);
        if (rc) fprintf(stderr, "ivrserver: compat frame failed: %s\n", ivr_last_error());
        return reply(c, rc) ? -1 : 0;
    }
    /* the batcher's frames */
    ivr_request r;
    memset(&r, 0, sizeof r);
    r.id = c->rid = g_next_rid++;
    r.volume = (vol >= 0 && vol < g_nvol) ? vol : -1;
    ivr_desc_config(&r.desc);
    r.desc.state = st;
    r.desc.skip = g_skip;
    if (r.volume >= 0) {
        r.desc.lutscale = g_vol[r.volume].lutscale;
        r.desc.lutoffset = g_vol[r.volume].lutoffset;
    }
    if (lut) {
        r.lut = lut;
        r.lutsize = nl;
    } else {
        r.lut = r.volume >= 0 ? g_vol[r.volume].lut : g_vol[0].lut;
        r.lutsize = 256;
    }
    r.rgba8 = c->img;
    r.t_submit = ivr_now();
    c->busy = 1;
    return ivr_batcher_submit(bt, &r) ? -1 : 0;
}

static void consume(client *c, size_t len)
{
    memmove(c->in, c->in + len, c->have - len);
    c->have -= len;
}

/* submit the complete messages buffered in c->in while the client has no
 * frame in flight (a client may send its next STATE, or its finish, before
 * the reply to the last one); -1 on finish or a malformed message */
static int parse_buffered(ivr_batcher *bt, client *c)
{
    while (!c->busy) {
        size_t len = message_len(c);
        if (len == (size_t)-1) return -1;
        if (!len) break;
        int rc = handle(bt, c, len);
        consume(c, len);
        if (rc) return -1; /* finish or malformed */
    }
    return 0;
}

/* retire finished frames, reply to their clients and start their buffered
 * requests; -1 when the batcher reports a device error */
static int pump(ivr_batcher *bt)
{
    ivr_completion done[64];
    if (ivr_batcher_progress(bt, 0) < 0) {
        fprintf(stderr, "ivrserver: batcher: %s\n", ivr_last_error());
        return -1;
    }
    int n = ivr_batcher_poll(bt, done, 64);
    for (int k = 0; k < n; k++)
        for (int i = 0; i < MAX_CLIENTS; i++)
            if (g_cl[i].used && g_cl[i].busy && g_cl[i].rid == done[k].id) {
                g_cl[i].busy = 0;
                if (g_cl[i].zombie || reply(&g_cl[i], done[k].status) || parse_buffered(bt, &g_cl[i]))
                    drop_client(&g_cl[i]);
            }
    return n;
}

static int read_client(ivr_batcher *bt, client *c)
{
    if (c->cap - c->have < 65536) {
        size_t nc = c->cap ? c->cap * 2 : 1 << 17;
        unsigned char *p = (unsigned char *)realloc(c->in, nc);
        if (!p) return -1;
        c->in = p;
        c->cap = nc;
    }
    ssize_t r = recv(c->fd, c->in + c->have, c->cap - c->have, 0);
    if (r <= 0) return -1; /* peer closed mid-STATE or error: drop it */
    c->have += (size_t)r;
    return parse_buffered(bt, c);
}

static void usage(void)
{
    fprintf(stderr, "usage: ivrserver (--listen PORT | --relay HOST:PORT) --volume SPEC [--volume SPEC ...]\n"
                    "       [--budget-gb G] [--slots N] [--max W H] [--skip 0|1] [--device D]\n"
                    "       [--compat COLORMAP XFER] [--tile TW TH] [--tiff DIR] [--swap-floats]\n"
                    "   or: ivrserver FILE W H D ENDIAN COLORMAP XFER PX PY PZ TILEW TILEH ANGLE AX AY AZ\n");
}

int main(int argc, char **argv)
{
    const char *listen_port = NULL, *relay = NULL;
    const char *specs[MAX_VOLS];
    int nspec = 0, slots = 16, maxw = 1920, maxh = 1080, device = 0;
    double budget_gb = 64.0;
    char refspec[600], relay_buf[256];
    if (argc == 17 && argv[1][0] != '-') {
        /* the reference server's command line (parseargs, main.c:1047-1073) */
        const int endian = atoi(argv[5]);
        snprintf(refspec, sizeof refspec, "raw:%s:%s:%s:%s:%d", argv[1], argv[2], argv[3], argv[4], endian ? 1 : 0);
        specs[nspec++] = refspec;
        g_compat = 1;
        g_cmap = atoi(argv[6]);
        g_xfer = atoi(argv[7]);
        g_tile_w = atoi(argv[11]);
        g_tile_h = atoi(argv[12]);
        const char *env = getenv("IVR_RELAY");
        snprintf(relay_buf, sizeof relay_buf, "%s", env && *env ? env : "localhost:1400");
        relay = relay_buf;
        argc = 1; /* nothing else to parse */
    }
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--listen") && i + 1 < argc) listen_port = argv[++i];
        else if (!strcmp(argv[i], "--relay") && i + 1 < argc) relay = argv[++i];
        else if (!strcmp(argv[i], "--volume") && i + 1 < argc && nspec < MAX_VOLS) specs[nspec++] = argv[++i];
        else if (!strcmp(argv[i], "--budget-gb") && i + 1 < argc) budget_gb = atof(argv[++i]);
        else if (!strcmp(argv[i], "--slots") && i + 1 < argc) slots = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--max") && i + 2 < argc) { maxw = atoi(argv[++i]); maxh = atoi(argv[++i]); }
        else if (!strcmp(argv[i], "--skip") && i + 1 < argc) g_skip = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--device") && i + 1 < argc) device = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--swap-floats")) g_swap_floats = 1;
        else if (!strcmp(argv[i], "--tiff") && i + 1 < argc) g_tiff_dir = argv[++i];
        else if (!strcmp(argv[i], "--tile") && i + 2 < argc) {
            g_tile_w = atoi(argv[++i]);
            g_tile_h = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--compat") && i + 2 < argc) {
            g_compat = 1;
            g_cmap = atoi(argv[++i]);
            g_xfer = atoi(argv[++i]);
        }
        else { usage(); return 2; }
    }
    if ((!listen_port && !relay) || nspec == 0) { usage(); return 2; }
    signal(SIGPIPE, SIG_IGN);
    ivr_ctx *ctx;
    ivr_batcher *bt;
    if (ivr_ctx_create(device, &ctx)) { fprintf(stderr, "ivrserver: %s\n", ivr_last_error()); return 1; }
    g_ctx = ctx;
    if (g_tile_w < 1 || g_tile_h < 1) {
        fprintf(stderr, "ivrserver: tile sizes must be positive\n");
        return 2;
    }
    if (g_compat && (g_cmap < 0 || g_cmap > 4 || g_xfer < 0 || g_xfer > 5)) {
        /* the reference builds no LUT for colormap -1 (main.c:340) and renders with none */
        fprintf(stderr, "ivrserver: compat serving needs colormap 0..4 and transfer function 0..5\n");
        return 2;
    }
    if (ivr_batcher_create(ctx, (unsigned long long)(budget_gb * 1073741824.0), slots, maxw, maxh, &bt)) {
        fprintf(stderr, "ivrserver: %s\n", ivr_last_error());
        return 1;
    }
    for (int i = 0; i < nspec; i++) {
        int v = add_volume(ctx, specs[i]);
        if (v < 0 || ivr_batcher_add_volume(bt, g_vol[v].vol) < 0) {
            fprintf(stderr, "ivrserver: volume '%s' failed: %s\n", specs[i], ivr_last_error());
            return 1;
        }
        fprintf(stderr, "ivrserver: volume %d ready (%s)\n", v, specs[i]);
    }
    int lfd = -1;
    if (relay) {
        /* the reference render server's side of the relay (initconnection) */
        char host[256];
        snprintf(host, sizeof host, "%s", relay);
        char *colon = strrchr(host, ':');
        if (!colon) { usage(); return 2; }
        *colon = 0;
        struct addrinfo hints, *res;
        memset(&hints, 0, sizeof hints);
        hints.ai_family = AF_UNSPEC;
        hints.ai_socktype = SOCK_STREAM;
        if (getaddrinfo(host, colon + 1, &hints, &res)) { fprintf(stderr, "ivrserver: bad relay\n"); return 1; }
        int fd = socket(res->ai_family, res->ai_socktype, res->ai_protocol);
        if (fd < 0 || connect(fd, res->ai_addr, res->ai_addrlen)) { perror("ivrserver: connect"); return 1; }
        freeaddrinfo(res);
        nodelay(fd);
        char hello[14];
        if (recv_all(fd, hello, 13) || send_all(fd, "Hello, server!", 14)) { fprintf(stderr, "handshake\n"); return 1; }
        new_client(fd);
    } else {
        struct sockaddr_in a;
        memset(&a, 0, sizeof a);
        a.sin_family = AF_INET;
        a.sin_addr.s_addr = htonl(INADDR_LOOPBACK);
        a.sin_port = htons((uint16_t)atoi(listen_port));
        lfd = socket(AF_INET, SOCK_STREAM, 0);
        int one = 1;
        setsockopt(lfd, SOL_SOCKET, SO_REUSEADDR, &one, sizeof one);
        if (bind(lfd, (struct sockaddr *)&a, sizeof a) || listen(lfd, 64)) { perror("ivrserver: listen"); return 1; }
        fprintf(stderr, "ivrserver: listening on 127.0.0.1:%s\n", listen_port);
    }
    fprintf(stderr, "ivrserver: ready\n");
    for (;;) {
        struct pollfd pf[MAX_CLIENTS + 1];
        int map[MAX_CLIENTS + 1], np = 0, nclients = 0;
        if (lfd >= 0) { pf[np].fd = lfd; pf[np].events = POLLIN; map[np++] = -1; }
        for (int i = 0; i < MAX_CLIENTS; i++)
            if (g_cl[i].used && !g_cl[i].zombie) {
                nclients++;
                pf[np].fd = g_cl[i].fd;
                pf[np].events = g_cl[i].busy ? 0 : POLLIN;
                map[np++] = i;
            }
        int zombies = 0;
        for (int i = 0; i < MAX_CLIENTS; i++) zombies += g_cl[i].zombie;
        if (relay && nclients == 0 && zombies == 0) break; /* relay session over */
        ivr_batcher_stats st;
        ivr_batcher_stats_get(bt, &st);
        const int work = st.pending > 0 || st.in_flight > 0;
        int pr = poll(pf, (nfds_t)np, work ? 0 : 50);
        if (pr < 0 && errno != EINTR) break;
        for (int k = 0; k < np && pr > 0; k++) {
            if (!(pf[k].revents & (POLLIN | POLLHUP | POLLERR))) continue;
            if (map[k] < 0) {
                int fd = accept(lfd, NULL, NULL);
                if (fd < 0) continue;
                nodelay(fd);
                char greet[14];
                /* the relay's greeting (myserver.c:118-123) */
                if (send_all(fd, "Hello, world!", 13) || recv_all(fd, greet, 14) || !new_client(fd)) close(fd);
            } else if (read_client(bt, &g_cl[map[k]])) {
                drop_client(&g_cl[map[k]]);
            }
        }
        if (pump(bt) < 0) break;
        if (work) usleep(100);
    }
    ivr_batcher_destroy(bt);
    if (g_frame) ivr_frame_destroy(ctx, g_frame);
    if (g_req_tf) ivr_tf_destroy(ctx, g_req_tf);
    for (int v = 0; v < g_nvol; v++) {
        if (g_vol[v].tf) ivr_tf_destroy(ctx, g_vol[v].tf);
        ivr_volume_destroy(ctx, g_vol[v].vol);
    }
    ivr_ctx_destroy(ctx);
    return 0;
}
