/* fdv_stream.c -- the reference's file container, streaming decoder, streaming
 * encoder and average-bitrate control on the MI355X path, with the reference's
 * signatures:
 *
 *   fdvamd_wrap / fdvamd_read / fdvamd_decode = fdv_wrap / fdv_read / fdv_decode
 *       (fdv.h:1090-1105, bodies 7448-7520)
 *   fdvamd_dec_*  = fdv_dec_*  (fdv.h:1126-1142, bodies 7526-7795)
 *   fdvamd_enc_*  = fdv_enc_*  (fdv.h:1146-1198, bodies 7798-8459)
 *   fdvamd_rc_*   = fdv_rc_*   (fdv.h:1023-1054, bodies 7880-8040)
 *
 * P-frames of every band of a frame go through the device in one pass
 * (fdvamd_pframe_batch, fdvamd_pframe_decode_batch); key frames through the
 * registered I-frame coder.  The rate control decides each frame's quantizer
 * and lambda scale, and the requantising retry its final attempt, so the
 * streams match the reference's byte for byte only if every step of that
 * arithmetic does: it is restated here in the reference's operation order.
 */
#include "fdv_amd.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

extern void fdvamd_seterr_(const char *fmt, ...);
extern fdvamd_iframe_encode_fn fdvamd_iframe_encoder_(void);
extern fdvamd_iframe_decode_fn fdvamd_iframe_decoder_(void);
extern fdvamd_encoder *fdvamd_sequence_encoder_(int nb, const int *w, const int *h, int me_mode);
extern int fdvamd_intra_bands_(fdvamd_iframe_encode_fn fn, const uint8_t *const *frames, int w, const int *bh, int nb,
                               int qp, int nthreads, uint8_t **blob, size_t bcap, size_t *blen, uint8_t **rec);

/* Little-endian fields (fdv_put16 / fdv_put32 / fdv_get16 / fdv_get32). */
static size_t put16(uint8_t *b, size_t p, unsigned v) { b[p] = (uint8_t)v; b[p + 1] = (uint8_t)(v >> 8); return p + 2; }
static size_t put32(uint8_t *b, size_t p, uint32_t v) {
    for (int i = 0; i < 4; ++i) b[p + i] = (uint8_t)(v >> (8 * i));
    return p + 4;
}
static unsigned get16(const uint8_t *b, size_t *p) { const unsigned v = b[*p] | (unsigned)b[*p + 1] << 8; *p += 2; return v; }
static uint32_t get32(const uint8_t *b, size_t *p) {
    const uint32_t v = b[*p] | (uint32_t)b[*p + 1] << 8 | (uint32_t)b[*p + 2] << 16 | (uint32_t)b[*p + 3] << 24;
    *p += 4;
    return v;
}
static int ceil_div(int a, int b) { return (a + b - 1) / b; }

/* ---- container (fdv.h:7448-7520) ------------------------------------------ */

#define MAGIC "FDV1"
#define VERSION 1

size_t fdvamd_wrap(uint8_t *out, size_t cap, int kind, int fps, const uint8_t *payload, size_t payload_len) {
    if (!out || payload_len > 0xffffffffu || cap < FDVAMD_HEAD + payload_len) return 0;
    if (kind != FDVAMD_VIDEO && kind != FDVAMD_TILED) return 0;
    if (fps <= 0 || fps > 65535) fps = 25;
    memmove(out + FDVAMD_HEAD, payload, payload_len);      /* the payload may already sit there */
    memcpy(out, MAGIC, 4);
    out[4] = (uint8_t)VERSION;
    out[5] = (uint8_t)kind;
    size_t o = put16(out, 6, (unsigned)fps);
    put32(out, o, (uint32_t)payload_len);
    return FDVAMD_HEAD + payload_len;
}

int fdvamd_read(const uint8_t *in, size_t len, fdvamd_info *info) {
    if (!in || !info || len < FDVAMD_HEAD || memcmp(in, MAGIC, 4) != 0) return -1;
    memset(info, 0, sizeof *info);
    info->version = in[4];
    info->kind = in[5];
    if (info->version != VERSION || (info->kind != FDVAMD_VIDEO && info->kind != FDVAMD_TILED)) return -1;
    size_t p = 6;
    info->fps = (int)get16(in, &p);
    const uint32_t plen = get32(in, &p);
    if (plen > len - FDVAMD_HEAD) return -1;
    info->payload = in + FDVAMD_HEAD;
    info->payload_len = plen;
    info->nbands = 1;
    size_t q = 0;
    if (info->kind == FDVAMD_VIDEO) {                  /* the two streams' headers differ */
        if (plen < 8) return -1;
        info->nframes = (int)get16(info->payload, &q);
        info->w = (int)get16(info->payload, &q);
        info->h = (int)get16(info->payload, &q);
        info->qp = info->payload[q++];
        info->keyint = info->payload[q++];
    } else {
        if (plen < 12) return -1;
        info->w = (int)get16(info->payload, &q);
        info->h = (int)get16(info->payload, &q);
        info->nframes = (int)get16(info->payload, &q);
        info->qp = info->payload[q++];
        info->keyint = info->payload[q++];
        info->nbands = (int)get16(info->payload, &q);
    }
    if (info->nframes <= 0 || info->w <= 0 || info->h <= 0 || (info->w & 15) || (info->h & 15)) return -1;
    if (info->w > 8192 || info->h > 8192 || info->qp < 0 || info->qp > 51) return -1;
    if (info->fps <= 0) info->fps = 25;
    return 0;
}

int fdvamd_decode(const fdvamd_info *info, uint8_t *out, int nthreads) {
    if (!info || !out) return -1;
    int nf = 0, w = 0, h = 0;
    const int rc = info->kind == FDVAMD_TILED
                       ? fdvamd_vtile_decode(info->payload, info->payload_len, out, &nf, &w, &h, nthreads)
                       : fdvamd_video_decode(info->payload, info->payload_len, out, &nf, &w, &h);
    if (rc != 0) return -1;
    if (nf != info->nframes || w != info->w || h != info->h) return -1;    /* head and stream disagree */
    return 0;
}

/* ---- streaming decode (fdv.h:7526-7795) ----------------------------------- */

/* One coded video stream walked a frame at a time (a tiled file holds one per
 * band, stepped together).  Its reference pool lives in the device decoder. */
typedef struct {
    const uint8_t *pay;
    size_t paylen, pos;
    int nframes, w, h, qp, keyint, next, nref;
    size_t fsize;
    fdvamd_tabcache tc;
} vstream;

struct fdvamd_dec {
    int kind, w, h, nframes, fps, qp, keyint, next, nbands;
    size_t fsize;
    vstream *s;
    int *band_y0, *band_h;
    uint8_t **bandbuf;            /* tiled: each band's decoded sub-frame */
    uint8_t *scratch;             /* frames a seek passes over */
    fdvamd_decoder *dev;
    fdvamd_iframe_decode_fn idec;
};

static int vs_open(vstream *s, const uint8_t *pay, size_t len) {
    if (len < 8) return -1;
    size_t p = 0;
    s->nframes = (int)get16(pay, &p);
    s->w = (int)get16(pay, &p);
    s->h = (int)get16(pay, &p);
    s->qp = pay[p++];
    s->keyint = pay[p++];
    if (s->nframes <= 0 || s->w <= 0 || s->h <= 0 || (s->w & 15) || (s->h & 15) || s->w > 8192 || s->h > 8192)
        return -1;
    if (s->qp > 51) return -1;
    s->pay = pay; s->paylen = len; s->pos = p; s->next = 0; s->nref = 0;
    s->fsize = (size_t)s->w * s->h + 2 * (size_t)(s->w / 2) * (s->h / 2);
    return 0;
}

/* Back to the first frame: the pool is simply marked empty. */
static void vs_rewind(vstream *s) { s->pos = 8; s->next = 0; s->nref = 0; }

/* Skip the next frame's payload (only before a key frame). */
static int vs_skip(vstream *s) {
    if (s->next >= s->nframes || s->pos + 4 > s->paylen) return -1;
    const uint32_t blob = get32(s->pay, &s->pos);
    if (s->pos + blob > s->paylen) return -1;
    s->pos += blob;
    ++s->next;
    return 0;
}

void fdvamd_dec_close(fdvamd_dec *d) {
    if (!d) return;
    fdvamd_decoder_close(d->dev);
    if (d->bandbuf) for (int b = 0; b < d->nbands; ++b) free(d->bandbuf[b]);
    free(d->bandbuf); free(d->s); free(d->band_y0); free(d->band_h); free(d->scratch);
    free(d);
}

fdvamd_dec *fdvamd_dec_open(const fdvamd_info *info) {
    if (!info || !info->payload) { fdvamd_seterr_("fdvamd_dec_open: bad arguments"); return NULL; }
    fdvamd_dec *d = calloc(1, sizeof *d);
    if (!d) { fdvamd_seterr_("out of memory"); return NULL; }
    d->kind = info->kind;
    d->w = info->w; d->h = info->h; d->nframes = info->nframes;
    d->fps = info->fps; d->qp = info->qp; d->keyint = info->keyint;
    d->fsize = (size_t)d->w * d->h * 3 / 2;
    d->nbands = 1;
    d->idec = fdvamd_iframe_decoder_();
    int *ws = NULL;
    if (!d->idec) { fdvamd_seterr_("no I-frame decoder registered (fdvamd_set_iframe_coder)"); goto fail; }
    if (d->kind == FDVAMD_VIDEO) {
        d->s = calloc(1, sizeof *d->s);
        d->band_y0 = calloc(1, sizeof(int));
        d->band_h = calloc(1, sizeof(int));
        if (!d->s || !d->band_y0 || !d->band_h) goto oom;
        if (vs_open(d->s, info->payload, info->payload_len) != 0) goto bad;
        d->band_h[0] = d->s->h;
    } else {
        /* a header, then one length-prefixed video stream per band */
        size_t p = 0;
        if (info->payload_len < 12) goto bad;
        const int ww = (int)get16(info->payload, &p), hh = (int)get16(info->payload, &p);
        p += 4;                                   /* nframes, qp, keyint: already known */
        const int nbands = (int)get16(info->payload, &p), band_mbrows = (int)get16(info->payload, &p);
        const int bh_full = band_mbrows * 16;
        if (ww != d->w || hh != d->h || nbands < 1 || band_mbrows < 1 || nbands != ceil_div(hh, bh_full)) goto bad;
        d->nbands = nbands;
        d->s = calloc((size_t)nbands, sizeof *d->s);
        d->band_y0 = calloc((size_t)nbands, sizeof(int));
        d->band_h = calloc((size_t)nbands, sizeof(int));
        d->bandbuf = calloc((size_t)nbands, sizeof(uint8_t *));
        if (!d->s || !d->band_y0 || !d->band_h || !d->bandbuf) goto oom;
        for (int b = 0; b < nbands; ++b) {
            const int y0 = b * bh_full, bh = y0 + bh_full <= hh ? bh_full : hh - y0;
            if (p + 4 > info->payload_len) goto bad;
            const uint32_t blob = get32(info->payload, &p);
            if (p + blob > info->payload_len) goto bad;
            if (vs_open(&d->s[b], info->payload + p, blob) != 0) goto bad;
            if (d->s[b].w != d->w || d->s[b].h != bh) goto bad;
            p += blob;
            d->band_y0[b] = y0;
            d->band_h[b] = bh;
            if (!(d->bandbuf[b] = malloc(d->s[b].fsize))) goto oom;
        }
    }
    /* the bands' P-frames are decoded in one device pass: they must share their
     * key interval, as the reference's tiled encoder writes them */
    for (int b = 1; b < d->nbands; ++b)
        if (d->s[b].keyint != d->s[0].keyint) { fdvamd_seterr_("fdvamd_dec_open: bands with different key intervals"); goto fail; }
    ws = malloc((size_t)d->nbands * sizeof(int));
    if (!ws) goto oom;
    for (int b = 0; b < d->nbands; ++b) ws[b] = d->w;
    d->dev = fdvamd_decoder_open(d->nbands, ws, d->band_h);
    free(ws);
    if (!d->dev) goto fail;
    return d;
oom:
    fdvamd_seterr_("out of memory");
    goto fail;
bad:
    fdvamd_seterr_("fdvamd_dec_open: malformed stream");
fail:
    fdvamd_dec_close(d);
    return NULL;
}

int fdvamd_dec_pos(const fdvamd_dec *d) { return d ? d->next : -1; }
size_t fdvamd_dec_frame_size(const fdvamd_dec *d) { return d ? d->fsize : 0; }

/* The next frame of every band, into dst[b]: key frames by the registered
 * decoder, P-frames of all bands in one device pass.  1, 0 at the end, -1. */
static int bands_next(fdvamd_dec *d, uint8_t *const *dst) {
    const int nb = d->nbands;
    const uint8_t *in[256];
    size_t len[256];
    int nav[256];
    fdvamd_tabcache *tcp[256];
    if (nb > 256) { fdvamd_seterr_("too many bands"); return -1; }
    for (int b = 0; b < nb; ++b) {
        vstream *s = &d->s[b];
        if (s->next >= s->nframes) return 0;
        if (s->pos + 4 > s->paylen) return -1;
        const uint32_t blob = get32(s->pay, &s->pos);
        if (s->pos + blob > s->paylen) return -1;
        in[b] = s->pay + s->pos;
        len[b] = blob;
        s->pos += blob;
    }
    const int f = d->s[0].next;
    const int is_intra = f == 0 || (d->s[0].keyint > 0 && f % d->s[0].keyint == 0);
    if (is_intra) {
        for (int b = 0; b < nb; ++b) {
            vstream *s = &d->s[b];
            memset(&s->tc, 0, sizeof s->tc);        /* tables and history restart at a key frame */
            const size_t ys = (size_t)s->w * s->h, cs = (size_t)(s->w / 2) * (s->h / 2);
            if (d->idec(in[b], len[b], dst[b], dst[b] + ys, dst[b] + ys + cs, s->w, s->h) != 0) {
                fdvamd_seterr_("frame %d of band %d rejected", f, b);
                return -1;
            }
            if (fdvamd_decoder_set_ref(d->dev, b, dst[b]) != 0) return -1;
        }
    } else {
        for (int b = 0; b < nb; ++b) { nav[b] = d->s[b].nref >= 2 ? 2 : 1; tcp[b] = &d->s[b].tc; }
        if (fdvamd_pframe_decode_batch(d->dev, in, len, nav, tcp, dst) != 0) return -1;
    }
    for (int b = 0; b < nb; ++b) {
        vstream *s = &d->s[b];
        s->nref = is_intra ? 1 : (s->nref < 2 ? s->nref + 1 : 2);
        ++s->next;
    }
    return 1;
}

int fdvamd_dec_next(fdvamd_dec *d, uint8_t *out) {
    if (!d || !out) return -1;
    if (d->next >= d->nframes) return 0;
    if (d->kind == FDVAMD_VIDEO) {
        uint8_t *dst[1] = {out};
        const int rc = bands_next(d, dst);
        if (rc == 1) ++d->next;
        return rc;
    }
    const int rc = bands_next(d, d->bandbuf);
    if (rc != 1) return rc;
    /* each band's rows into place */
    const int cw = d->w / 2, ch = d->h / 2;
    for (int b = 0; b < d->nbands; ++b) {
        const int y0 = d->band_y0[b], bh = d->band_h[b], bch = bh / 2, cy0 = y0 / 2;
        const uint8_t *sy = d->bandbuf[b], *su = sy + (size_t)d->w * bh, *sv = su + (size_t)cw * bch;
        uint8_t *fy = out, *fu = fy + (size_t)d->w * d->h, *fv = fu + (size_t)cw * ch;
        memcpy(fy + (size_t)y0 * d->w, sy, (size_t)d->w * bh);
        memcpy(fu + (size_t)cy0 * cw, su, (size_t)cw * bch);
        memcpy(fv + (size_t)cy0 * cw, sv, (size_t)cw * bch);
    }
    ++d->next;
    return 1;
}

int fdvamd_dec_seek(fdvamd_dec *d, int frame) {
    if (!d || frame < 0 || frame >= d->nframes) return -1;
    if (frame == d->next) return 0;
    /* the key frame at or before the target; roll forward from here instead
     * when already past it and behind the target */
    const int key = d->keyint > 0 ? (frame / d->keyint) * d->keyint : 0;
    const int from = (d->next > key && d->next <= frame) ? d->next : key;
    if (from == key) {
        for (int b = 0; b < d->nbands; ++b) vs_rewind(&d->s[b]);
        d->next = 0;
        while (d->next < key) {
            for (int b = 0; b < d->nbands; ++b)
                if (vs_skip(&d->s[b]) != 0) return -1;
            ++d->next;
        }
    }
    if (!d->scratch && !(d->scratch = malloc(d->fsize))) return -1;
    while (d->next < frame)
        if (fdvamd_dec_next(d, d->scratch) != 1) return -1;
    return 0;
}

/* ---- average-bitrate control (fdv.h:7830-8040) ---------------------------- */

#define RC_DRAIN 2.0        /* seconds over which the bucket returns to half full */
#define RC_STEP 1           /* largest QP move between consecutive P-frames */
#define RC_IP_MAX 8.0       /* how much finer a key frame may be coded */
#define RC_IBOOST 10.0      /* a key frame's share, in P-frames */
#define RC_GAIN 1.0         /* QP steps per doubling of the bit error */
#define RC_OVER 1.5         /* how far past its allocation a frame may land */
#define RC_LFRAC 1.15       /* the lambda multiplier spanning one QP step */

void fdvamd_rc_init(fdvamd_rc *rc, int bitrate, int fps, int w, int h, int keyint, double bufsecs, int qpmin,
                    int qpmax) {
    memset(rc, 0, sizeof *rc);
    rc->bitrate = bitrate > 0 ? bitrate : 1000000;
    rc->fps = fps > 0 ? fps : 30;
    rc->bufbits = rc->bitrate * (bufsecs > 0.0 ? bufsecs : 1.0);
    rc->qpmin = qpmin > 0 ? qpmin : 8;
    rc->qpmax = qpmax > 0 ? qpmax : 51;
    rc->fill = rc->bufbits * 0.5;
    rc->keyint = keyint > 1 ? keyint : rc->fps * 10;
    /* starting point from bits per pixel per frame */
    const double bpp = (double)rc->bitrate / rc->fps / ((double)w * h);
    double qp = 26.0 - 6.0 * log2(bpp / 0.03);
    if (qp < rc->qpmin) qp = rc->qpmin;
    if (qp > rc->qpmax) qp = rc->qpmax;
    rc->qp_p = qp;
    rc->ip_off = 0.0;
}

/* What the frame may cost: the key-frame interval's bits split between its key
 * frame (RC_IBOOST shares) and P-frames, steered slowly toward a half-full
 * bucket, clamped. */
static double rc_target(fdvamd_rc *rc, int is_intra) {
    const double budget = (double)rc->bitrate / rc->fps;
    const double n = (double)rc->keyint;
    const double denom = RC_IBOOST + (n - 1.0);
    double target = budget * n * (is_intra ? RC_IBOOST : 1.0) / denom;
    target -= (rc->fill - rc->bufbits * 0.5) / (rc->fps * RC_DRAIN);
    const double lo = budget * (is_intra ? 1.0 : 0.4);
    const double hi = budget * (is_intra ? n * RC_IBOOST / denom * 2.0 : 2.5);
    if (target < lo) target = lo;
    if (target > hi) target = hi;
    return target;
}

void fdvamd_rc_split(const fdvamd_rc *rc, double q, int *qp, double *lscale) {
    if (q < rc->qpmin) q = rc->qpmin;
    if (q > rc->qpmax) q = rc->qpmax;
    const int i = (int)floor(q);
    if (i >= rc->qpmax) { *qp = rc->qpmax; *lscale = 1.0; return; }
    *qp = i;
    *lscale = pow(RC_LFRAC, q - (double)i);
}

int fdvamd_rc_pick(fdvamd_rc *rc, int is_intra) {
    rc->last_target = rc_target(rc, is_intra);
    const double q = is_intra ? rc->qp_p - rc->ip_off : rc->qp_p;
    int qp;
    if (is_intra) {
        /* a key frame takes the nearest whole quantizer */
        qp = (int)lround(q);
        if (qp < rc->qpmin) qp = rc->qpmin;
        if (qp > rc->qpmax) qp = rc->qpmax;
        rc->lscale = 1.0;
    } else {
        fdvamd_rc_split(rc, q, &qp, &rc->lscale);
    }
    rc->used_qp = qp;
    return qp;
}

void fdvamd_rc_adopt(fdvamd_rc *rc, int is_intra, double q) {
    if (is_intra) {
        double off = rc->qp_p - q;
        if (off < 0.0) off = 0.0;
        if (off > RC_IP_MAX) off = RC_IP_MAX;
        rc->ip_off = off;
    } else {
        if (q < rc->qpmin) q = rc->qpmin;
        if (q > rc->qpmax) q = rc->qpmax;
        rc->qp_p = q;
    }
}

void fdvamd_rc_update(fdvamd_rc *rc, int is_intra, size_t bytes) {
    const double bits = (double)bytes * 8.0;
    const double err = log2(bits / (rc->last_target > 1.0 ? rc->last_target : 1.0));
    double dq = RC_GAIN * err;
    if (is_intra) {
        /* a key frame moves the offset, not the operating point */
        if (dq > 2.0) dq = 2.0;
        if (dq < -2.0) dq = -2.0;
        rc->ip_off -= dq;
        if (rc->ip_off < 0.0) rc->ip_off = 0.0;
        if (rc->ip_off > RC_IP_MAX) rc->ip_off = RC_IP_MAX;
    } else {
        if (dq > RC_STEP) dq = RC_STEP;
        if (dq < -RC_STEP) dq = -RC_STEP;
        rc->qp_p += dq;
        if (rc->qp_p < rc->qpmin) rc->qp_p = rc->qpmin;
        if (rc->qp_p > rc->qpmax) rc->qp_p = rc->qpmax;
    }
    rc->fill += bits - (double)rc->bitrate / rc->fps;
    if (rc->fill < 0.0) rc->fill = 0.0;
    if (rc->fill > rc->bufbits * 2.0) rc->fill = rc->bufbits * 2.0;
}

void fdvamd_rc_skip(fdvamd_rc *rc, size_t bytes) {
    rc->fill += (double)bytes * 8.0 - (double)rc->bitrate / rc->fps;
    if (rc->fill < 0.0) rc->fill = 0.0;
    if (rc->fill > rc->bufbits * 2.0) rc->fill = rc->bufbits * 2.0;
}

/* ---- streaming encode (fdv.h:7798-8459) ----------------------------------- */

/* One coded sub-stream: the whole video, or one band of a tiled one. */
typedef struct {
    int h;
    size_t fsize;
    fdvamd_tabcache tc;
    uint8_t *out;                 /* FDVAMD_HEAD reserved, then its video stream */
    size_t cap, len;
    uint8_t *src, *blob, *rec;    /* this frame's slice, coded blob, key-frame reconstruction */
    size_t bcap, blen;
} substream;

struct fdvamd_enc {
    int w, h, qp, keyint, n, key_due, nref, tiled, nbands, band_mbrows, nthreads;
    size_t fsize, bmax;           /* bmax: every band's blob buffer (the tallest band's capacity) */
    fdvamd_rc rc;
    int use_rc;
    substream *b;
    fdvamd_encoder *dev;
};

void fdvamd_enc_close(fdvamd_enc *e) {
    if (!e) return;
    fdvamd_encoder_close(e->dev);
    if (e->b)
        for (int k = 0; k < e->nbands; ++k) { free(e->b[k].out); free(e->b[k].src); free(e->b[k].blob); free(e->b[k].rec); }
    free(e->b);
    free(e);
}

static fdvamd_enc *enc_make(int w, int h, int qp, int keyint, int band_mbrows, int nthreads, int tiled) {
    if (w <= 0 || h <= 0 || (w & 15) || (h & 15) || w > 8192 || h > 8192 || qp < 0 || qp > 51) {
        fdvamd_seterr_("fdvamd_enc_open: bad arguments"); return NULL;
    }
    if (keyint < 0) keyint = 0;
    if (keyint > 255) keyint = 255;
    const int bh_full = band_mbrows * 16, nb = tiled ? ceil_div(h, bh_full) : 1;
    if (nb > 65535) { fdvamd_seterr_("fdvamd_enc_open_tiled: too many bands"); return NULL; }
    fdvamd_enc *e = calloc(1, sizeof *e);
    int *ws = malloc((size_t)nb * sizeof(int)), *hs = malloc((size_t)nb * sizeof(int));
    if (!e || !ws || !hs || !(e->b = calloc((size_t)nb, sizeof *e->b))) goto oom;
    e->w = w; e->h = h; e->qp = qp; e->keyint = keyint; e->tiled = tiled;
    e->nbands = nb; e->band_mbrows = band_mbrows; e->nthreads = nthreads < 1 ? 1 : nthreads;
    e->fsize = (size_t)w * h + 2 * (size_t)(w / 2) * (h / 2);
    for (int k = 0; k < nb; ++k) {
        substream *s = &e->b[k];
        const int y0 = k * bh_full;
        s->h = tiled ? (y0 + bh_full <= h ? bh_full : h - y0) : h;
        s->fsize = (size_t)w * s->h + 2 * (size_t)(w / 2) * (s->h / 2);
        s->cap = FDVAMD_HEAD + 8 + s->fsize;
        s->bcap = (size_t)w * s->h * 3 + 8192;          /* the reference's frame buffer (fdv.h:8140) */
        e->bmax = k == 0 ? s->bcap : e->bmax;             /* band 0 is the tallest */
        s->out = malloc(s->cap);
        s->src = malloc(s->fsize);
        s->blob = malloc(e->bmax);
        s->rec = malloc(s->fsize);
        if (!s->out || !s->src || !s->blob || !s->rec) goto oom;
        /* the band's video header, frame count patched at finish */
        size_t o = put16(s->out, FDVAMD_HEAD, 0);
        o = put16(s->out, o, (unsigned)w);
        o = put16(s->out, o, (unsigned)s->h);
        s->out[o++] = (uint8_t)qp;
        s->out[o++] = (uint8_t)keyint;
        s->len = o;
        ws[k] = w;
        hs[k] = s->h;
    }
    if (!(e->dev = fdvamd_sequence_encoder_(nb, ws, hs, -1))) goto fail;
    free(ws); free(hs);
    return e;
oom:
    fdvamd_seterr_("out of memory");
fail:
    free(ws); free(hs);
    fdvamd_enc_close(e);
    return NULL;
}

fdvamd_enc *fdvamd_enc_open(int w, int h, int qp, int keyint) { return enc_make(w, h, qp, keyint, 1, 1, 0); }

fdvamd_enc *fdvamd_enc_open_tiled(int w, int h, int qp, int keyint, int band_mbrows, int nthreads) {
    if (band_mbrows < 1) { fdvamd_seterr_("fdvamd_enc_open_tiled: bad band height"); return NULL; }
    return enc_make(w, h, qp, keyint, band_mbrows, nthreads, 1);
}

int fdvamd_enc_count(const fdvamd_enc *e) { return e ? e->n : 0; }

size_t fdvamd_enc_bytes(const fdvamd_enc *e) {
    if (!e) return 0;
    size_t t = 0;
    for (int k = 0; k < e->nbands; ++k) t += e->b[k].len - FDVAMD_HEAD;
    return t;
}

int fdvamd_enc_set_bitrate(fdvamd_enc *e, int bits_per_second, int fps, double bufsecs) {
    if (!e || bits_per_second <= 0 || e->n != 0) { fdvamd_seterr_("fdvamd_enc_set_bitrate: bad arguments"); return -1; }
    fdvamd_rc_init(&e->rc, bits_per_second, fps, e->w, e->h, e->keyint, bufsecs, 0, 0);
    e->use_rc = 1;
    return 0;
}

/* Every band's frame at quantizer qp, lambda scale ls: key frames by the
 * registered coder (their reconstruction kept for the pool), P-frames in one
 * device pass.  Each band's blob goes to its blob buffer. */
static int code_frame(fdvamd_enc *e, int is_intra, int repeat, int qp, double ls, fdvamd_tabcache *const *tcp) {
    const int nb = e->nbands;
    const uint8_t *src[256];
    uint8_t *blob[256], *rec[256];
    size_t cap[256], blen[256];
    int bh[256], nav[256], qps[256], fs[256];
    double lss[256];
    if (nb > 256) { fdvamd_seterr_("too many bands"); return -1; }
    for (int k = 0; k < nb; ++k) {
        substream *s = &e->b[k];
        src[k] = repeat ? NULL : s->src;
        blob[k] = s->blob; rec[k] = s->rec; cap[k] = s->bcap; bh[k] = s->h;
        nav[k] = e->nref >= 2 ? 2 : 1; qps[k] = qp; lss[k] = ls; fs[k] = repeat;
    }
    if (is_intra) {
        const fdvamd_iframe_encode_fn ienc = fdvamd_iframe_encoder_();
        if (fdvamd_intra_bands_(ienc, src, e->w, bh, nb, qp, e->nthreads, blob, e->bmax, blen, rec) != 0) {
            fdvamd_seterr_("the I-frame coder failed on frame %d", e->n);
            return -1;
        }
    } else if (fdvamd_pframe_batch(e->dev, src, nav, qps, lss, fs, tcp, blob, cap, blen, NULL) != 0) {
        return -1;
    }
    for (int k = 0; k < nb; ++k) e->b[k].blen = blen[k];
    return 0;
}

/* One frame of every band (enc_push, fdv.h:8231-8370, and bands_push,
 * 8065-8125): key-frame decision, quantizer, coding, and -- on a single stream
 * under rate control -- the requantising retry, which keeps the device's motion
 * tables for P-frames. */
static int enc_push(fdvamd_enc *e, const uint8_t *i420, int repeat) {
    if (e->n >= 65535) { fdvamd_seterr_("too many frames"); return -1; }
    const int f = e->n;
    int is_intra = f == 0 || e->key_due || (e->keyint > 0 && f % e->keyint == 0);
    /* a held frame cannot carry a key frame: the next real one does */
    if (repeat) { e->key_due = is_intra; is_intra = 0; }
    else e->key_due = 0;
    const int nb = e->nbands;
    fdvamd_tabcache *tcp[256];
    for (int k = 0; k < nb; ++k) {
        substream *s = &e->b[k];
        tcp[k] = &s->tc;
        if (is_intra) memset(&s->tc, 0, sizeof s->tc);
        if (!repeat) {
            /* the band's rows of the frame, planes contiguous */
            const int y0 = k * e->band_mbrows * 16, cw = e->w / 2, ch = e->h / 2, bch = s->h / 2, cy0 = y0 / 2;
            const uint8_t *fy = i420, *fu = fy + (size_t)e->w * e->h, *fv = fu + (size_t)cw * ch;
            if (!e->tiled) memcpy(s->src, i420, s->fsize);
            else {
                memcpy(s->src, fy + (size_t)y0 * e->w, (size_t)e->w * s->h);
                memcpy(s->src + (size_t)e->w * s->h, fu + (size_t)cy0 * cw, (size_t)cw * bch);
                memcpy(s->src + (size_t)e->w * s->h + (size_t)cw * bch, fv + (size_t)cy0 * cw, (size_t)cw * bch);
            }
        }
    }
    /* the quantizer: the rate control (a tiled encoder's bands all take its
     * one QP at lambda scale 1), or the fixed one */
    int fqp = e->qp;
    double lscale = 1.0;
    if (e->use_rc) {
        fqp = repeat ? e->rc.used_qp : fdvamd_rc_pick(&e->rc, is_intra);
        if (!e->tiled) lscale = e->rc.lscale;
    }
    double qreal = (double)fqp + log(lscale) / log(RC_LFRAC);
    const size_t before = fdvamd_enc_bytes(e);
    const fdvamd_tabcache tc_before = e->b[0].tc;
    if (code_frame(e, is_intra, repeat, fqp, lscale, tcp) != 0) return -1;
    if (e->use_rc && !e->tiled && e->b[0].blen > 0) {
        /* a frame far past its allocation is coded again, coarser, until it
         * fits or the quantizer runs out; the QP it ends at is adopted */
        for (int tries = 0; tries < 3 && qreal < (double)e->rc.qpmax; ++tries) {
            const double alloc = e->rc.last_target > 1.0 ? e->rc.last_target : 1.0;
            const double over = (double)e->b[0].blen * 8.0 / alloc;
            if (over <= RC_OVER) break;
            double step = 2.0 * log2(over);
            if (step < 0.25) step = 0.25;
            if (step > 4.0) step = 4.0;
            double qn = qreal + step;
            if (qn > (double)e->rc.qpmax) qn = (double)e->rc.qpmax;
            int retry;
            double rls;
            fdvamd_rc_split(&e->rc, qn, &retry, &rls);
            e->b[0].tc = tc_before;
            if (!is_intra && fdvamd_rewind_keep_motion(e->dev) != 0) return -1;
            if (code_frame(e, is_intra, repeat, retry, rls, tcp) != 0) return -1;
            fqp = retry; qreal = qn; e->rc.used_qp = retry;
        }
        fdvamd_rc_adopt(&e->rc, is_intra, qreal);
    }
    for (int k = 0; k < nb; ++k) {
        substream *s = &e->b[k];
        if (s->len + 4 + s->blen > s->cap) {
            size_t want = s->cap * 2;
            while (want < s->len + 4 + s->blen) want *= 2;
            uint8_t *bigger = realloc(s->out, want);
            if (!bigger) { fdvamd_seterr_("out of memory"); return -1; }
            s->out = bigger;
            s->cap = want;
        }
        s->len = put32(s->out, s->len, (uint32_t)s->blen);
        memcpy(s->out + s->len, s->blob, s->blen);
        s->len += s->blen;
        /* the key frame's reconstruction becomes the near reference (load_ref
         * and the pool swap); a P-frame's stays on the device */
        if (is_intra && fdvamd_set_ref_planes(e->dev, k, 0, s->rec, s->rec + (size_t)e->w * s->h,
                                              s->rec + (size_t)e->w * s->h + (size_t)(e->w / 2) * (s->h / 2)) != 0)
            return -1;
    }
    if (e->use_rc) {
        const size_t used = e->tiled ? fdvamd_enc_bytes(e) - before : e->b[0].blen;
        if (repeat) fdvamd_rc_skip(&e->rc, used);
        else fdvamd_rc_update(&e->rc, is_intra, used);
    }
    e->nref = is_intra ? 1 : (e->nref < 2 ? e->nref + 1 : 2);
    ++e->n;
    return 0;
}

int fdvamd_enc_frame(fdvamd_enc *e, const uint8_t *i420) {
    if (!e || !i420) { fdvamd_seterr_("fdvamd_enc_frame: bad arguments"); return -1; }
    return enc_push(e, i420, 0);
}

int fdvamd_enc_repeat(fdvamd_enc *e) {
    if (!e || e->n == 0) { fdvamd_seterr_("fdvamd_enc_repeat: nothing to repeat"); return -1; }
    return enc_push(e, NULL, 1);
}

uint8_t *fdvamd_enc_finish(fdvamd_enc *e, int fps, size_t *len) {
    if (!e) return NULL;
    if (e->n == 0) { fdvamd_enc_close(e); return NULL; }
    uint8_t *buf;
    size_t flen;
    if (e->tiled) {
        /* the layout fdv_vtile_encode writes */
        size_t total = 12;
        for (int k = 0; k < e->nbands; ++k) total += 4 + (e->b[k].len - FDVAMD_HEAD);
        const size_t cap = FDVAMD_HEAD + total + 64;
        if (!(buf = malloc(cap))) { fdvamd_enc_close(e); return NULL; }
        size_t o = put16(buf, FDVAMD_HEAD, (unsigned)e->w);
        o = put16(buf, o, (unsigned)e->h);
        o = put16(buf, o, (unsigned)e->n);
        buf[o++] = (uint8_t)e->qp;
        buf[o++] = (uint8_t)e->keyint;
        o = put16(buf, o, (unsigned)e->nbands);
        o = put16(buf, o, (unsigned)e->band_mbrows);
        for (int k = 0; k < e->nbands; ++k) {
            substream *s = &e->b[k];
            put16(s->out, FDVAMD_HEAD, (unsigned)e->n);
            const size_t bl = s->len - FDVAMD_HEAD;
            o = put32(buf, o, (uint32_t)bl);
            memcpy(buf + o, s->out + FDVAMD_HEAD, bl);
            o += bl;
        }
        flen = fdvamd_wrap(buf, cap, FDVAMD_TILED, fps, buf + FDVAMD_HEAD, o - FDVAMD_HEAD);
        fdvamd_enc_close(e);
        if (!flen) { free(buf); return NULL; }
    } else {
        substream *s = &e->b[0];
        put16(s->out, FDVAMD_HEAD, (unsigned)e->n);       /* the one thing a streaming writer cannot know */
        flen = fdvamd_wrap(s->out, s->cap, FDVAMD_VIDEO, fps, s->out + FDVAMD_HEAD, s->len - FDVAMD_HEAD);
        buf = s->out;
        s->out = NULL;                                     /* ownership passes to the caller */
        fdvamd_enc_close(e);
        if (!flen) { free(buf); return NULL; }
    }
    if (len) *len = flen;
    return buf;
}
