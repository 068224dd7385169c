/* fdv_seq.c -- the reference's sequence entry points on the MI355X path, with
 * the reference's own signatures and containers:
 *
 *   fdvamd_video_encode  = fdv_video_encode  (fdv.h:903-904, body 6741-6854)
 *   fdvamd_vtile_encode  = fdv_vtile_encode  (fdv.h:973-975, body 7161-7243)
 *
 * P-frames go through the device encoder (fdvamd_pframe_batch: every band of a
 * frame in one device pass, entropy coded on host threads).  Key frames go
 * through the device key-frame coder (fdvamd_iframe_bands_: every plane of
 * every band in one device pass), or through an I-frame coder the caller
 * registers with fdvamd_set_iframe_coder.  The container bytes -- the sequence
 * header, the per-frame lengths, the tiled header and the per-band lengths in
 * band order -- follow the reference exactly, so a reference decoder reads the
 * output unchanged.
 */
#include "fdv_amd.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

extern void fdvamd_seterr_(const char *fmt, ...);

static fdvamd_iframe_encode_fn g_ienc;
static fdvamd_iframe_decode_fn g_idec;
static int g_mode;                        /* search mode of the sequence entry points */
static float *g_nnw;                      /* proposer weights for the learned mode */
static int g_nnn;
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;

int fdvamd_set_iframe_coder(fdvamd_iframe_encode_fn enc, fdvamd_iframe_decode_fn dec) {
    pthread_mutex_lock(&g_mu);
    g_ienc = enc;
    g_idec = dec;
    pthread_mutex_unlock(&g_mu);
    return 0;
}

/* The registered key-frame decoder, or the built-in one, for the sequence decoders. */
fdvamd_iframe_decode_fn fdvamd_iframe_decoder_(void) {
    pthread_mutex_lock(&g_mu);
    const fdvamd_iframe_decode_fn f = g_idec;
    pthread_mutex_unlock(&g_mu);
    return f ? f : fdvamd_iframe_decode;
}

int fdvamd_set_sequence_mode(int me_mode) {
    if (me_mode < 0 || me_mode > 3) { fdvamd_seterr_("fdvamd_set_sequence_mode: bad mode %d", me_mode); return -1; }
    pthread_mutex_lock(&g_mu);
    g_mode = me_mode;
    pthread_mutex_unlock(&g_mu);
    return 0;
}

int fdvamd_set_sequence_proposer(const float *weights, int n) {
    float *c = NULL;
    if (weights && n > 0) {
        c = malloc((size_t)n * sizeof(float));
        if (!c) { fdvamd_seterr_("out of memory"); return -1; }
        memcpy(c, weights, (size_t)n * sizeof(float));
    }
    pthread_mutex_lock(&g_mu);
    free(g_nnw);
    g_nnw = c;
    g_nnn = c ? n : 0;
    pthread_mutex_unlock(&g_mu);
    return 0;
}

/* Little-endian header fields (fdv_put16 / fdv_put32, fdv.h:1742-1743). */
static size_t put16(uint8_t *b, size_t p, unsigned v) { b[p] = (uint8_t)v; b[p + 1] = (uint8_t)(v >> 8); return p + 2; }
static size_t put32(uint8_t *b, size_t p, uint32_t v) {
    for (int i = 0; i < 4; ++i) b[p + i] = (uint8_t)(v >> (8 * i));
    return p + 4;
}

/* One coded sub-stream (a whole video, or one band of a tiled one): its bytes
 * so far against the capacity the reference gives it. */
typedef struct {
    uint8_t *buf;
    size_t n, cap;
    int ok;
} sstream;

static int ss_open(sstream *s, size_t cap) {
    s->cap = cap;
    s->n = 0;
    s->ok = 1;
    s->buf = malloc(cap ? cap : 1);
    return s->buf ? 0 : -1;
}

/* Append one frame record: u32 length, then the blob (fdv.h:6807-6815). */
static void ss_frame(sstream *s, const uint8_t *blob, size_t len) {
    if (!s->ok || len == 0 || s->n + 4 + len > s->cap) { s->ok = 0; return; }
    s->n = put32(s->buf, s->n, (uint32_t)len);
    memcpy(s->buf + s->n, blob, len);
    s->n += len;
}

/* The registered key-frame coder, for the streaming encoder. */
fdvamd_iframe_encode_fn fdvamd_iframe_encoder_(void) {
    pthread_mutex_lock(&g_mu);
    const fdvamd_iframe_encode_fn f = g_ienc;
    pthread_mutex_unlock(&g_mu);
    return f;
}

/* A device encoder for nb sequences in the given search mode (-1: the
 * process-wide sequence mode), with the registered proposer weights in the
 * learned mode. */
fdvamd_encoder *fdvamd_sequence_encoder_(int nb, const int *w, const int *h, int me_mode) {
    pthread_mutex_lock(&g_mu);
    if (me_mode < 0) me_mode = g_mode;
    pthread_mutex_unlock(&g_mu);
    fdvamd_encoder *e = fdvamd_encoder_open(nb, w, h, me_mode);
    if (!e || me_mode != 2) return e;
    pthread_mutex_lock(&g_mu);
    const int have = g_nnw != NULL;
    const int prc = have ? fdvamd_set_proposer(e, g_nnw, g_nnn) : -1;
    pthread_mutex_unlock(&g_mu);
    if (prc != 0) {
        if (!have) fdvamd_seterr_("learned mode without proposer weights");
        fdvamd_encoder_close(e);
        return NULL;
    }
    return e;
}

/* Key frames of every band in parallel, as the reference's band threads code
 * them (fdv.h:7211-7222): each band's I-frame blob and reconstruction. */
typedef struct {
    fdvamd_iframe_encode_fn fn;
    const uint8_t *const *frames;
    int w, qp, nb, first, step;
    const int *bh;
    uint8_t **blob, **rec;
    size_t *blen, bcap;
} ijob;

static void *ijob_run(void *arg) {
    ijob *j = arg;
    for (int b = j->first; b < j->nb; b += j->step) {
        const size_t ys = (size_t)j->w * j->bh[b], cs = (size_t)(j->w / 2) * (j->bh[b] / 2);
        const uint8_t *y = j->frames[b];
        j->blen[b] = j->fn(y, y + ys, y + ys + cs, j->w, j->bh[b], j->qp, j->blob[b], j->bcap, j->rec[b],
                           j->rec[b] + ys, j->rec[b] + ys + cs);
    }
    return NULL;
}

extern int fdvamd_iframe_bands_(const uint8_t *const *frames, int w, const int *bh, int nb, int qp, uint8_t **blob,
                                size_t bcap, size_t *blen, uint8_t **rec);

/* fn NULL: the device key-frame coder, all bands in one pass. */
int fdvamd_intra_bands_(fdvamd_iframe_encode_fn fn, const uint8_t *const *frames, int w, const int *bh, int nb, int qp,
                        int nthreads, uint8_t **blob, size_t bcap, size_t *blen, uint8_t **rec) {
    if (!fn) return fdvamd_iframe_bands_(frames, w, bh, nb, qp, blob, bcap, blen, rec);
    int nt = nthreads < 1 ? 1 : nthreads > nb ? nb : nthreads;
    pthread_t tid[64];
    ijob jobs[64];
    if (nt > 64) nt = 64;
    int spawned = 0;
    for (int t = 0; t < nt; ++t)
        jobs[t] = (ijob){fn, frames, w, qp, nb, t, nt, bh, blob, rec, blen, bcap};
    for (int t = 1; t < nt; ++t) {
        if (pthread_create(&tid[t], NULL, ijob_run, &jobs[t]) != 0) break;
        spawned = t;
    }
    ijob_run(&jobs[0]);
    for (int t = 1; t <= spawned; ++t) pthread_join(tid[t], NULL);
    for (int t = spawned + 1; t < nt; ++t) ijob_run(&jobs[t]);      /* spawn failed: run here */
    for (int b = 0; b < nb; ++b) if (!blen[b]) return -1;
    return 0;
}

/* The frames of nb bands (or one whole stream, nb = 1) coded as independent
 * video sub-streams in one device batch: frame f of every band in one pass.
 * src[f][b] is band b's tight I420 sub-frame of frame f. */
static int code_bands(const uint8_t *const *const *src, int nframes, int w, const int *bh, int nb, int qp, int keyint,
                      int nthreads, int me_mode, sstream *ss) {
    fdvamd_iframe_encode_fn ienc;
    pthread_mutex_lock(&g_mu);
    ienc = g_ienc;
    pthread_mutex_unlock(&g_mu);
    int *ws = malloc((size_t)nb * sizeof(int)), *nav = malloc((size_t)nb * sizeof(int));
    int *qps = malloc((size_t)nb * sizeof(int)), *nref = calloc((size_t)nb, sizeof(int));
    uint8_t **blob = calloc((size_t)nb, sizeof(uint8_t *)), **rec = calloc((size_t)nb, sizeof(uint8_t *));
    size_t *blen = calloc((size_t)nb, sizeof(size_t)), *caps = malloc((size_t)nb * sizeof(size_t));
    fdvamd_tabcache *tc = calloc((size_t)nb, sizeof(fdvamd_tabcache));
    fdvamd_tabcache **tcp = malloc((size_t)nb * sizeof(fdvamd_tabcache *));
    fdvamd_encoder *e = NULL;
    int rc = -1;
    size_t bcap = 0;
    if (!ws || !nav || !qps || !nref || !blob || !rec || !blen || !caps || !tc || !tcp) {
        fdvamd_seterr_("out of memory");
        goto out;
    }
    for (int b = 0; b < nb; ++b) {
        ws[b] = w;
        qps[b] = qp;
        tcp[b] = &tc[b];
        /* the per-frame blob capacity of fdv_video_encode (fdv.h:6756) */
        caps[b] = (size_t)w * bh[b] * 3 + 8192;
        if (caps[b] > bcap) bcap = caps[b];
    }
    for (int b = 0; b < nb; ++b) {
        blob[b] = malloc(bcap);
        rec[b] = malloc((size_t)w * bh[b] + 2 * (size_t)(w / 2) * (bh[b] / 2));
        if (!blob[b] || !rec[b]) { fdvamd_seterr_("out of memory"); goto out; }
    }
    e = fdvamd_sequence_encoder_(nb, ws, bh, me_mode);
    if (!e) goto out;
    for (int f = 0; f < nframes; ++f) {
        const int is_intra = f == 0 || (keyint > 0 && f % keyint == 0);
        if (is_intra) {
            /* tables and history restart at a key frame (fdv.h:6784) */
            memset(tc, 0, (size_t)nb * sizeof *tc);
            if (fdvamd_intra_bands_(ienc, src[f], w, bh, nb, qp, nthreads, blob, bcap, blen, rec) != 0) {
                fdvamd_seterr_("the I-frame coder failed on frame %d", f);
                goto out;
            }
            for (int b = 0; b < nb; ++b) {
                /* the key frame's reconstruction becomes the near reference
                 * (load_ref + pool swap, fdv.h:6842-6846) */
                if (fdvamd_set_ref(e, b, 0, rec[b]) != 0) goto out;
                nref[b] = 1;
            }
        } else {
            for (int b = 0; b < nb; ++b) nav[b] = nref[b] >= 2 ? 2 : 1;
            if (fdvamd_pframe_batch(e, src[f], nav, qps, NULL, NULL, tcp, blob, caps, blen, NULL) != 0) goto out;
            for (int b = 0; b < nb; ++b) nref[b] = nref[b] < 2 ? nref[b] + 1 : 2;
        }
        for (int b = 0; b < nb; ++b) ss_frame(&ss[b], blob[b], blen[b]);
    }
    rc = 0;
    for (int b = 0; b < nb; ++b)
        if (!ss[b].ok) { fdvamd_seterr_("band %d does not fit its stream capacity", b); rc = -1; }
out:
    fdvamd_encoder_close(e);
    if (blob) for (int b = 0; b < nb; ++b) free(blob[b]);
    if (rec) for (int b = 0; b < nb; ++b) free(rec[b]);
    free(ws); free(nav); free(qps); free(nref); free(blob); free(rec); free(blen); free(caps); free(tc); free(tcp);
    return rc;
}

/* Sequence header (fdv.h:6765-6770). */
static size_t seq_header(uint8_t *b, int nframes, int w, int h, int qp, int keyint) {
    size_t o = 0;
    o = put16(b, o, (unsigned)nframes);
    o = put16(b, o, (unsigned)w);
    o = put16(b, o, (unsigned)h);
    b[o++] = (uint8_t)qp;
    b[o++] = (uint8_t)keyint;
    return o;
}

static int sequence_mode(void) {
    pthread_mutex_lock(&g_mu);
    const int m = g_mode;
    pthread_mutex_unlock(&g_mu);
    return m;
}

size_t fdvamd_video_encode(const uint8_t *const *frames, int nframes, int w, int h, int qp, int keyint,
                           uint8_t *out, size_t cap) {
    if (nframes <= 0 || (w & 15) || (h & 15) || !frames || !out || w <= 0 || h <= 0) {
        fdvamd_seterr_("fdvamd_video_encode: bad arguments");
        return 0;
    }
    if (keyint < 0) keyint = 0;
    if (keyint > 255) keyint = 255;
    if (cap < 8) { fdvamd_seterr_("fdvamd_video_encode: no room for the header"); return 0; }
    const uint8_t *const **src = malloc((size_t)nframes * sizeof *src);
    if (!src) { fdvamd_seterr_("out of memory"); return 0; }
    for (int f = 0; f < nframes; ++f) src[f] = &frames[f];
    sstream ss;
    size_t o = seq_header(out, nframes, w, h, qp, keyint);
    ss.buf = out + o;
    ss.n = 0;
    ss.cap = cap - o;
    ss.ok = 1;
    const int rc = code_bands(src, nframes, w, &h, 1, qp, keyint, 1, sequence_mode(), &ss);
    free(src);
    return rc == 0 ? o + ss.n : 0;
}

size_t fdvamd_vtile_encode(const uint8_t *const *frames, int nframes, int w, int h, int qp, int keyint,
                           int band_mbrows, int nthreads, uint8_t *out, size_t cap) {
    if (nframes <= 0 || w <= 0 || h <= 0 || (w & 15) || (h & 15) || band_mbrows < 1 || !frames || !out) {
        fdvamd_seterr_("fdvamd_vtile_encode: bad arguments");
        return 0;
    }
    if (cap < 12) { fdvamd_seterr_("fdvamd_vtile_encode: no room for the header"); return 0; }
    const int bh_full = band_mbrows * 16, nb = (h + bh_full - 1) / bh_full;
    if (keyint < 0) keyint = 0;
    if (keyint > 255) keyint = 255;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > nb) nthreads = nb;
    const int cw = w / 2, ch = h / 2;
    int *bh = malloc((size_t)nb * sizeof(int));
    sstream *ss = calloc((size_t)nb, sizeof *ss);
    const uint8_t **subp = malloc((size_t)nframes * nb * sizeof *subp);
    const uint8_t *const **src = malloc((size_t)nframes * sizeof *src);
    uint8_t *sub = NULL;
    size_t o = 0, rc = 0;
    size_t per = 0;
    if (!bh || !ss || !subp || !src) { fdvamd_seterr_("out of memory"); goto out; }
    for (int b = 0; b < nb; ++b) {
        const int y0 = b * bh_full;
        bh[b] = y0 + bh_full <= h ? bh_full : h - y0;
        per += (size_t)w * bh[b] + 2 * (size_t)cw * (bh[b] / 2);
    }
    sub = malloc(per * (size_t)nframes);
    if (!sub) { fdvamd_seterr_("out of memory"); goto out; }
    /* band sub-frames (encode_band, fdv.h:7113-7124) */
    for (int f = 0; f < nframes; ++f) {
        uint8_t *d = sub + (size_t)f * per;
        const uint8_t *fy = frames[f], *fu = fy + (size_t)w * h, *fv = fu + (size_t)cw * ch;
        for (int b = 0; b < nb; ++b) {
            const int y0 = b * bh_full, cy0 = y0 / 2, bch = bh[b] / 2;
            memcpy(d, fy + (size_t)y0 * w, (size_t)w * bh[b]);
            memcpy(d + (size_t)w * bh[b], fu + (size_t)cy0 * cw, (size_t)cw * bch);
            memcpy(d + (size_t)w * bh[b] + (size_t)cw * bch, fv + (size_t)cy0 * cw, (size_t)cw * bch);
            subp[(size_t)f * nb + b] = d;
            d += (size_t)w * bh[b] + 2 * (size_t)cw * bch;
        }
        src[f] = &subp[(size_t)f * nb];
    }
    /* each band is its own fdv_video_encode stream, with that call's capacity
     * (fdv.h:7110: two band frames + 64 KiB) */
    for (int b = 0; b < nb; ++b) {
        const size_t bfs = (size_t)w * bh[b] + 2 * (size_t)cw * (bh[b] / 2);
        if (ss_open(&ss[b], bfs * 2 + 65536) != 0) { fdvamd_seterr_("out of memory"); goto out; }
        ss[b].n = seq_header(ss[b].buf, nframes, w, bh[b], qp, keyint);
    }
    if (code_bands(src, nframes, w, bh, nb, qp, keyint, nthreads, sequence_mode(), ss) != 0) goto out;
    /* the tiled header, then the band streams in band order (fdv.h:7177-7184, 7228-7236) */
    o = put16(out, o, (unsigned)w);
    o = put16(out, o, (unsigned)h);
    o = put16(out, o, (unsigned)nframes);
    out[o++] = (uint8_t)qp;
    out[o++] = (uint8_t)keyint;
    o = put16(out, o, (unsigned)nb);
    o = put16(out, o, (unsigned)band_mbrows);
    for (int b = 0; b < nb; ++b) {
        if (o + 4 + ss[b].n > cap) { fdvamd_seterr_("fdvamd_vtile_encode: output does not fit"); goto out; }
        o = put32(out, o, (uint32_t)ss[b].n);
        memcpy(out + o, ss[b].buf, ss[b].n);
        o += ss[b].n;
    }
    rc = o;
out:
    if (ss) for (int b = 0; b < nb; ++b) free(ss[b].buf);
    free(bh); free(ss); free(subp); free(src); free(sub);
    return rc;
}
