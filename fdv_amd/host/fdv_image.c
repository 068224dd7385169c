/* fdv_image.c -- key frames with the reference's signatures:
 *
 *   fdvamd_image_encode   = fdv_image_encode (fdv.h:871, body 4845-5038)
 *   fdvamd_iframe_encode  = iframe_encode    (fdv.h:5487-5539), the built-in
 *                           key-frame coder of the sequence and streaming
 *                           encoders
 *
 * The device decides (fdvgpu_image_batch: the rate-distortion quadtree, the
 * leaves' modes and levels, the deblocked reconstruction, every plane of a
 * batch in one pass); this file serialises the decided tree into the codec's
 * twelve streams and writes the container (fdvamd_pack_image_).
 *
 * Stream order: one stream of split flags per tree depth (64, 32, 16, 8), the
 * large-leaf modes, the 4x4 modes under the two contexts (left and above
 * cells agree, or not), then the 4x4 counts and levels, the 8x8 counts and
 * levels, and the region flags (fdv.h:4285-4303).  Every stream holds the
 * surviving nodes' symbols in coding order: a node's split flag (none where
 * the plane's edge implies the split), then its children, or its mode and its
 * coefficients.
 */
#include "fdv_amd.h"
#include "fdvgpu.h"

#include <stdlib.h>
#include <string.h>

extern void fdvamd_seterr_(const char *fmt, ...);
size_t fdvamd_pack_image_(int w, int h, int qp, const uint8_t *const sv[12], const size_t sn[12], uint8_t *out,
                          size_t cap);

enum { ST_SPLIT = 0, ST_MODEB = 4, ST_MODE0, ST_MODE1, ST_CNT, ST_LVL, ST_CNT8, ST_LVL8, ST_FLAG, ST_N };

static const uint8_t ZZ4[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};
static const uint8_t ZZ8[64] = {
     0,  1,  8, 16,  9,  2,  3, 10, 17, 24, 32, 25, 18, 11,  4,  5,
    12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13,  6,  7, 14, 21, 28,
    35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
    58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

typedef struct { uint8_t *b[ST_N]; size_t n[ST_N], cap; } streams;

static void put(streams *s, int k, unsigned v) { if (s->n[k] < s->cap) s->b[k][s->n[k]++] = (uint8_t)v; }
static void put_var(streams *s, int k, uint32_t u) {
    for (; u >= 0x80; u >>= 7) put(s, k, (u & 0x7f) | 0x80);
    put(s, k, u);
}
static uint32_t zig(int v) { return ((uint32_t)v << 1) ^ (uint32_t)(v >> 31); }

/* a transform block's count and levels: n coefficients at (x, y), row stride
 * rs of the level plane, scanned by zz */
static void block_syms(streams *s, const int16_t *lev, size_t rs, int x, int y, int side, const uint8_t *zz,
                       int kcnt, int klvl) {
    const int n = side * side;
    int last = -1;
    for (int k = 0; k < n; ++k) {
        const int i = zz[k];
        if (lev[(size_t)(y + i / side) * rs + x + i % side]) last = k;
    }
    put(s, kcnt, (unsigned)(last + 1));
    for (int k = 0; k <= last; ++k) {
        const int i = zz[k];
        put_var(s, klvl, zig(lev[(size_t)(y + i / side) * rs + x + i % side]));
    }
}

/* The twelve streams of one decided plane: the units in raster order, each
 * tree walked in coding order with an explicit stack of pending nodes. */
static int serialise(const fdvgpu_itree_out *t, int w, int h, streams *s) {
    int tw[4], th[4], toff[4], o = 0;
    for (int d = 0; d < 4; ++d) {
        const int n = 64 >> d;
        tw[d] = (w + n - 1) / n; th[d] = (h + n - 1) / n;
        toff[d] = o; o += 2 * tw[d] * th[d];
    }
    const int cw = w / 4, rw8 = (w + 7) / 8;
    for (int uy = 0; uy < h; uy += 64)
        for (int ux = 0; ux < w; ux += 64) {
            struct { int d, x, y; } stk[64];                  /* pending nodes */
            int sp = 0;
            stk[sp].d = 0; stk[sp].x = ux; stk[sp].y = uy; ++sp;
            while (sp > 0) {
                --sp;
                const int d = stk[sp].d, y = stk[sp].y, x = stk[sp].x, n = 64 >> d;
                if (n == 4) {
                    const int cx = x / 4, cy = y / 4;
                    const int lm = cx ? t->cmode[(size_t)cy * cw + cx - 1] : 0;
                    const int am = cy ? t->cmode[(size_t)(cy - 1) * cw + cx] : 0;
                    put(s, lm == am ? ST_MODE0 : ST_MODE1, t->cmode[(size_t)cy * cw + cx]);
                    block_syms(s, t->lev, (size_t)w, x, y, 4, ZZ4, ST_CNT, ST_LVL);
                    continue;
                }
                const size_t ni = (size_t)(y / n) * tw[d] + x / n;
                const int sf = t->tree[toff[d] + ni];
                if (sf > 2) return -1;
                if (sf != 2) put(s, ST_SPLIT + d, (unsigned)sf);
                if (sf) {
                    for (int k = 3; k >= 0; --k) {                /* pushed in reverse: popped in z-order */
                        const int cx = x + (k & 1) * n / 2, cy = y + (k >> 1) * n / 2;
                        if (cx < w && cy < h) { stk[sp].d = d + 1; stk[sp].x = cx; stk[sp].y = cy; ++sp; }
                    }
                    continue;
                }
                put(s, ST_MODEB, t->tree[toff[d] + (size_t)tw[d] * th[d] + ni]);
                for (int ry = y; ry < y + n; ry += 8)
                    for (int rx = x; rx < x + n; rx += 8) {
                        const int f = t->rflag[(size_t)(ry / 8) * rw8 + rx / 8];
                        put(s, ST_FLAG, (unsigned)f);
                        if (f == 1)
                            for (int q = 0; q < 4; ++q)
                                block_syms(s, t->lev, (size_t)w, rx + (q & 1) * 4, ry + (q >> 1) * 4, 4, ZZ4, ST_CNT,
                                           ST_LVL);
                        else if (f == 2)
                            block_syms(s, t->lev, (size_t)w, rx, ry, 8, ZZ8, ST_CNT8, ST_LVL8);
                    }
            }
        }
    for (int k = 0; k < ST_N; ++k) if (s->n[k] >= s->cap) return -1;
    return 0;
}

static int streams_alloc(streams *s, size_t px) {
    memset(s, 0, sizeof *s);
    s->cap = px * 5 + 1024;
    for (int k = 0; k < ST_N; ++k)
        if (!(s->b[k] = malloc(s->cap))) return -1;
    return 0;
}
static void streams_free(streams *s) { for (int k = 0; k < ST_N; ++k) free(s->b[k]); }

static size_t pack_tree(const fdvgpu_itree_out *t, int w, int h, int qp, uint8_t *out, size_t cap, streams *s) {
    memset(s->n, 0, sizeof s->n);
    if (serialise(t, w, h, s) != 0) { fdvamd_seterr_("stream overflow"); return 0; }
    const uint8_t *sv[ST_N];
    for (int k = 0; k < ST_N; ++k) sv[k] = s->b[k];
    const size_t n = fdvamd_pack_image_(w, h, qp, sv, s->n, out, cap);
    if (!n) fdvamd_seterr_("output capacity too small");
    return n;
}

size_t fdvgpu_image_pack(const fdvgpu_itree_out *t, int w, int h, int qp, uint8_t *out, size_t cap) {
    if (!t || !t->lev || !t->rflag || !t->cmode || !t->tree || !out || w <= 0 || h <= 0 || (w & 3) || (h & 3)) {
        fdvamd_seterr_("fdvgpu_image_pack: bad arguments");
        return 0;
    }
    streams s;
    size_t n = 0;
    if (streams_alloc(&s, (size_t)w * h) == 0) n = pack_tree(t, w, h, qp, out, cap, &s);
    else fdvamd_seterr_("out of memory");
    streams_free(&s);
    return n;
}

typedef struct {
    fdvgpu_itree_out t;
    uint8_t *mem;
} itree_buf;

static int itree_alloc(itree_buf *b, int w, int h) {
    const size_t px = (size_t)w * h, fl = (size_t)((w + 7) / 8) * ((h + 7) / 8), tb = fdvgpu_itree_bytes(w, h);
    b->mem = malloc(px + 2 * px + fl + px / 16 + tb + 64);
    if (!b->mem) return -1;
    uint8_t *q = b->mem;
    b->t.lev = (int16_t *)q; q += 2 * px;
    b->t.rec = q; q += px;
    b->t.rflag = q; q += fl;
    b->t.cmode = q; q += px / 16;
    b->t.tree = q;
    return 0;
}

int fdvamd_image_encode_batch(int n, const uint8_t *const *src, const int *w, const int *h, const int *stride,
                              const int *qp, uint8_t *const *out, const size_t *cap, uint8_t *const *recon,
                              size_t *len) {
    if (n < 1 || !src || !w || !h || !stride || !qp || !out || !cap || !len) {
        fdvamd_seterr_("fdvamd_image_encode_batch: bad arguments");
        return -1;
    }
    int rc = -1;
    itree_buf *tb = calloc((size_t)n, sizeof *tb);
    fdvgpu_iplane_in *in = calloc((size_t)n, sizeof *in);
    fdvgpu_itree_out *ot = calloc((size_t)n, sizeof *ot);
    streams s;
    memset(&s, 0, sizeof s);
    if (!tb || !in || !ot) { fdvamd_seterr_("out of memory"); goto out; }
    size_t maxpx = 0;
    for (int i = 0; i < n; ++i) {
        len[i] = 0;
        if (!src[i] || w[i] <= 0 || h[i] <= 0 || (w[i] & 3) || (h[i] & 3) || stride[i] < w[i] || qp[i] < 0 ||
            qp[i] > 51 || !out[i]) {
            fdvamd_seterr_("fdvamd_image_encode_batch: plane %d: bad arguments", i);
            goto out;
        }
        if (itree_alloc(&tb[i], w[i], h[i]) != 0) { fdvamd_seterr_("out of memory"); goto out; }
        in[i] = (fdvgpu_iplane_in){src[i], w[i], h[i], stride[i], qp[i]};
        ot[i] = tb[i].t;
        if ((size_t)w[i] * h[i] > maxpx) maxpx = (size_t)w[i] * h[i];
    }
    if (fdvgpu_image_batch(n, in, ot) != 0) {
        fdvamd_seterr_("%s", fdvgpu_last_error());
        goto out;
    }
    if (streams_alloc(&s, maxpx) != 0) { fdvamd_seterr_("out of memory"); goto out; }
    for (int i = 0; i < n; ++i) {
        len[i] = pack_tree(&ot[i], w[i], h[i], qp[i], out[i], cap[i], &s);
        if (!len[i]) goto out;
        if (recon && recon[i])
            memcpy(recon[i], ot[i].rec, (size_t)w[i] * h[i]);
    }
    rc = 0;
out:
    streams_free(&s);
    if (tb) for (int i = 0; i < n; ++i) free(tb[i].mem);
    free(tb); free(in); free(ot);
    return rc;
}

size_t fdvamd_image_encode(const uint8_t *src, int w, int h, int stride, int qp, uint8_t *out, size_t cap,
                           uint8_t *recon) {
    size_t len = 0;
    if (w <= 0 || h <= 0 || (w & 3) || (h & 3)) return 0;                /* as fdv.h:4847 */
    const uint8_t *s1[1] = {src};
    uint8_t *o1[1] = {out}, *r1[1] = {recon};
    return fdvamd_image_encode_batch(1, s1, &w, &h, &stride, &qp, o1, &cap, r1, &len) == 0 ? len : 0;
}

/* Key frames of nb bands at once (each band's three planes; the batch holds
 * them all): blob[b] gets band b's I-frame -- per plane a u32 little-endian
 * length and the plane's container -- and rec[b] its tight I420
 * reconstruction (iframe_encode, fdv.h:5487-5539). */
int fdvamd_iframe_bands_(const uint8_t *const *frames, int w, const int *bh, int nb, int qp, uint8_t **blob,
                         size_t bcap, size_t *blen, uint8_t **rec) {
    const int n = 3 * nb;
    int rc = -1;
    const uint8_t **src = calloc((size_t)n, sizeof *src);
    int *pw = calloc((size_t)n, sizeof *pw), *ph = calloc((size_t)n, sizeof *ph), *pq = calloc((size_t)n, sizeof *pq);
    uint8_t **po = calloc((size_t)n, sizeof *po), **pr = calloc((size_t)n, sizeof *pr);
    size_t *pc = calloc((size_t)n, sizeof *pc), *pl = calloc((size_t)n, sizeof *pl);
    if (!src || !pw || !ph || !pq || !po || !pr || !pc || !pl) { fdvamd_seterr_("out of memory"); goto out; }
    for (int b = 0; b < nb; ++b) {
        const size_t ys = (size_t)w * bh[b], cs = (size_t)(w / 2) * (bh[b] / 2);
        for (int p = 0; p < 3; ++p) {
            const int i = 3 * b + p;
            src[i] = frames[b] + (p ? ys + (size_t)(p - 1) * cs : 0);
            pw[i] = p ? w / 2 : w; ph[i] = p ? bh[b] / 2 : bh[b]; pq[i] = qp;
            pr[i] = rec[b] + (p ? ys + (size_t)(p - 1) * cs : 0);
            pc[i] = (size_t)pw[i] * ph[i] * 2 + 4096;                    /* fdv.h:5510-5512 */
            if (!(po[i] = malloc(pc[i]))) { fdvamd_seterr_("out of memory"); goto out; }
        }
    }
    if (fdvamd_image_encode_batch(n, src, pw, ph, pw, pq, po, pc, pr, pl) != 0) goto out;
    for (int b = 0; b < nb; ++b) {
        size_t o = 0;
        for (int p = 0; p < 3; ++p) {
            const int i = 3 * b + p;
            if (o + 4 + pl[i] > bcap) { fdvamd_seterr_("band %d: key frame exceeds its capacity", b); goto out; }
            for (int k = 0; k < 4; ++k) blob[b][o++] = (uint8_t)(pl[i] >> (8 * k));
            memcpy(blob[b] + o, po[i], pl[i]);
            o += pl[i];
        }
        blen[b] = o;
    }
    rc = 0;
out:
    if (po) for (int i = 0; i < n; ++i) free(po[i]);
    free(src); free(pw); free(ph); free(pq); free(po); free(pr); free(pc); free(pl);
    return rc;
}

size_t fdvamd_iframe_encode(const uint8_t *y, const uint8_t *u, const uint8_t *v, int w, int h, int qp, uint8_t *out,
                            size_t cap, uint8_t *ry, uint8_t *ru, uint8_t *rv) {
    if (!y || !u || !v || !out || !ry || !ru || !rv || w <= 0 || h <= 0 || (w & 7) || (h & 7)) {
        fdvamd_seterr_("fdvamd_iframe_encode: bad arguments");
        return 0;
    }
    const uint8_t *src[3] = {y, u, v};
    int pw[3] = {w, w / 2, w / 2}, ph[3] = {h, h / 2, h / 2}, pq[3] = {qp, qp, qp};
    uint8_t *po[3] = {0}, *pr[3] = {ry, ru, rv};
    size_t pc[3], pl[3], o = 0;
    for (int p = 0; p < 3; ++p) {
        pc[p] = (size_t)pw[p] * ph[p] * 2 + 4096;
        if (!(po[p] = malloc(pc[p]))) { fdvamd_seterr_("out of memory"); goto fail; }
    }
    if (fdvamd_image_encode_batch(3, src, pw, ph, pw, pq, po, pc, pr, pl) != 0) goto fail;
    for (int p = 0; p < 3; ++p) {
        if (o + 4 + pl[p] > cap) { fdvamd_seterr_("fdvamd_iframe_encode: output capacity too small"); goto fail; }
        for (int k = 0; k < 4; ++k) out[o++] = (uint8_t)(pl[p] >> (8 * k));
        memcpy(out + o, po[p], pl[p]);
        o += pl[p];
    }
    for (int p = 0; p < 3; ++p) free(po[p]);
    return o;
fail:
    for (int p = 0; p < 3; ++p) free(po[p]);
    return 0;
}
