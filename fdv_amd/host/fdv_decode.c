/* fdv_decode.c -- the reference's decoders on the MI355X path, with the
 * reference's signatures:
 *
 *   fdvamd_video_decode = fdv_video_decode (fdv.h:907-908, body 6857-6929)
 *   fdvamd_vtile_decode = fdv_vtile_decode (fdv.h:980-981, body 7280-7378)
 *
 * A P-frame is decoded in two halves: on the host, its payload's entropy coding
 * is undone and its symbols parsed into MB codes (fdv_unpack.c); on the device,
 * the frame is reconstructed from them and deblocked (fdvgpu_decode), every
 * sequence or band of a frame in one pass, against a two-frame reference pool
 * held in device memory.  Key frames come from the decoder registered with
 * fdvamd_set_iframe_coder.
 */
#include "fdv_amd.h"
#include "fdvgpu.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

extern void fdvamd_seterr_(const char *fmt, ...);
extern fdvamd_iframe_decode_fn fdvamd_iframe_decoder_(void);

struct fdvamd_decoder {
    fdvgpu *g;
    int nseq;
    int *w, *h, *qp;
    fdvgpu_mbcode **codes;
};

void fdvamd_decoder_close(fdvamd_decoder *d) {
    if (!d) return;
    fdvgpu_close(d->g);
    if (d->codes) for (int s = 0; s < d->nseq; ++s) free(d->codes[s]);
    free(d->codes); free(d->w); free(d->h); free(d->qp);
    free(d);
}

fdvamd_decoder *fdvamd_decoder_open(int nseq, const int *w, const int *h) {
    if (nseq <= 0 || !w || !h) { fdvamd_seterr_("fdvamd_decoder_open: bad arguments"); return NULL; }
    fdvamd_decoder *d = calloc(1, sizeof *d);
    if (!d) { fdvamd_seterr_("out of memory"); return NULL; }
    d->nseq = nseq;
    d->w = malloc((size_t)nseq * sizeof(int));
    d->h = malloc((size_t)nseq * sizeof(int));
    d->qp = malloc((size_t)nseq * sizeof(int));
    d->codes = calloc((size_t)nseq, sizeof *d->codes);
    if (!d->w || !d->h || !d->qp || !d->codes) { fdvamd_seterr_("out of memory"); fdvamd_decoder_close(d); return NULL; }
    for (int s = 0; s < nseq; ++s) {
        d->w[s] = w[s];
        d->h[s] = h[s];
        d->codes[s] = malloc((size_t)(w[s] / 16 > 0 ? w[s] / 16 : 1) * (size_t)(h[s] / 16 > 0 ? h[s] / 16 : 1) *
                             sizeof(fdvgpu_mbcode));
        if (!d->codes[s]) { fdvamd_seterr_("out of memory"); fdvamd_decoder_close(d); return NULL; }
    }
    d->g = fdvgpu_open(nseq, w, h, 1);
    if (!d->g) { fdvamd_seterr_("%s", fdvgpu_last_error()); fdvamd_decoder_close(d); return NULL; }
    return d;
}

int fdvamd_decoder_set_ref(fdvamd_decoder *d, int seq, const uint8_t *i420) {
    if (!d || fdvgpu_put_ref(d->g, seq, 0, i420) != 0) {
        fdvamd_seterr_("%s", d ? fdvgpu_last_error() : "bad arguments");
        return -1;
    }
    return 0;
}

/* The host half of one sequence's P-frame, on a thread of its own. */
typedef struct {
    fdvamd_decoder *d;
    const uint8_t *const *in;
    const size_t *len;
    fdvamd_tabcache *const *tc;
    int first, step, rc;
} unjob;

static void *unjob_run(void *arg) {
    unjob *j = arg;
    for (int s = j->first; s < j->d->nseq; s += j->step) {
        const int w = j->d->w[s], h = j->d->h[s];
        uint8_t *syms = NULL;
        size_t sn[16];
        if (fdvamd_unpack_pframe(j->in[s], j->len[s], (size_t)(w / 16) * (h / 16), j->tc ? j->tc[s] : NULL,
                                 &j->d->qp[s], &syms, sn) != 0 ||
            fdvamd_parse_pframe(syms, sn, w, h, j->d->codes[s]) != 0)
            j->rc = -1;
        free(syms);
    }
    return NULL;
}

int fdvamd_pframe_decode_batch(fdvamd_decoder *d, const uint8_t *const *in, const size_t *len, const int *navail,
                               fdvamd_tabcache *const *tc, uint8_t *const *out) {
    if (!d || !in || !len || !navail) { fdvamd_seterr_("fdvamd_pframe_decode_batch: bad arguments"); return -1; }
    const int nt = d->nseq < 16 ? d->nseq : 16;
    pthread_t tid[16];
    unjob jobs[16];
    int spawned = 0;
    for (int t = 0; t < nt; ++t) jobs[t] = (unjob){d, in, len, tc, t, nt, 0};
    for (int t = 1; t < nt; ++t) {
        if (pthread_create(&tid[t], NULL, unjob_run, &jobs[t]) != 0) break;
        spawned = t;
    }
    unjob_run(&jobs[0]);
    for (int t = 1; t <= spawned; ++t) pthread_join(tid[t], NULL);
    for (int t = spawned + 1; t < nt; ++t) unjob_run(&jobs[t]);
    for (int t = 0; t < nt; ++t)
        if (jobs[t].rc) { fdvamd_seterr_("malformed P-frame payload"); return -1; }
    if (fdvgpu_decode(d->g, (const fdvgpu_mbcode *const *)d->codes, d->qp, navail) != 0) {
        fdvamd_seterr_("%s", fdvgpu_last_error());
        return -1;
    }
    if (out)
        for (int s = 0; s < d->nseq; ++s)
            if (out[s] && fdvgpu_recon(d->g, s, out[s]) != 0) { fdvamd_seterr_("%s", fdvgpu_last_error()); return -1; }
    return 0;
}

static unsigned get16(const uint8_t *b, size_t *p) { unsigned v = b[*p] | (unsigned)b[*p + 1] << 8; *p += 2; return v; }
static uint32_t get32(const uint8_t *b, size_t *p) {
    uint32_t v = 0;
    for (int i = 0; i < 4; ++i) v |= (uint32_t)b[*p + i] << (8 * i);
    *p += 4;
    return v;
}

/* One band (or whole) video stream, located: its frame blobs in order. */
typedef struct {
    const uint8_t *in;
    size_t len;
    int nf, w, h, qp, keyint;
    const uint8_t **blob;
    size_t *blen;
} vstream;

/* The sequence header and frame records of fdv_video_encode's container
 * (fdv.h:6857-6885): -1 where the reference decoder rejects them. */
static int vstream_locate(vstream *v, const uint8_t *in, size_t len) {
    memset(v, 0, sizeof *v);
    if (len < 8) return -1;
    size_t p = 0;
    v->in = in; v->len = len;
    v->nf = (int)get16(in, &p);
    v->w = (int)get16(in, &p);
    v->h = (int)get16(in, &p);
    v->qp = in[p++];
    v->keyint = in[p++];
    if (v->nf <= 0 || (v->w & 15) || (v->h & 15) || v->w > 8192 || v->h > 8192 || v->w <= 0 || v->h <= 0) return -1;
    if (v->qp > 51) return -1;
    v->blob = malloc((size_t)v->nf * sizeof *v->blob);
    v->blen = malloc((size_t)v->nf * sizeof *v->blen);
    if (!v->blob || !v->blen) return -1;
    for (int f = 0; f < v->nf; ++f) {
        if (p + 4 > len) return -1;
        const uint32_t b = get32(in, &p);
        if (p + b > len) return -1;
        v->blob[f] = in + p;
        v->blen[f] = b;
        p += b;
    }
    return 0;
}

static void vstream_free(vstream *v) { free(v->blob); free(v->blen); }

typedef struct {
    fdvamd_iframe_decode_fn fn;
    const vstream *vs;
    int f, nb, first, step, rc;
    uint8_t **dst;
} ikjob;

static void *ikjob_run(void *arg) {
    ikjob *j = arg;
    for (int b = j->first; b < j->nb; b += j->step) {
        const vstream *v = &j->vs[b];
        const size_t ys = (size_t)v->w * v->h, cs = (size_t)(v->w / 2) * (v->h / 2);
        if (j->fn(v->blob[j->f], v->blen[j->f], j->dst[b], j->dst[b] + ys, j->dst[b] + ys + cs, v->w, v->h) != 0)
            j->rc = -1;
    }
    return NULL;
}

/* Decode nb streams of equal frame count and key interval in lockstep, frame f
 * of every stream in one device pass, into frame_out[b] + f * frame_bytes[b]. */
static int decode_lockstep(const vstream *vs, int nb, int nthreads, uint8_t **frame_out, size_t *frame_bytes) {
    fdvamd_iframe_decode_fn idec = fdvamd_iframe_decoder_();
    if (!idec) { fdvamd_seterr_("no I-frame decoder registered (fdvamd_set_iframe_coder)"); return -1; }
    const int nf = vs[0].nf;
    int *w = malloc((size_t)nb * sizeof(int)), *h = malloc((size_t)nb * sizeof(int));
    int *nav = malloc((size_t)nb * sizeof(int)), nref = 0;
    const uint8_t **in = malloc((size_t)nb * sizeof *in);
    size_t *len = malloc((size_t)nb * sizeof *len);
    uint8_t **dst = malloc((size_t)nb * sizeof *dst);
    fdvamd_tabcache *tc = calloc((size_t)nb, sizeof *tc);
    fdvamd_tabcache **tcp = malloc((size_t)nb * sizeof *tcp);
    fdvamd_decoder *d = NULL;
    int rc = -1;
    if (!w || !h || !nav || !in || !len || !dst || !tc || !tcp) { fdvamd_seterr_("out of memory"); goto out; }
    for (int b = 0; b < nb; ++b) { w[b] = vs[b].w; h[b] = vs[b].h; tcp[b] = &tc[b]; }
    d = fdvamd_decoder_open(nb, w, h);
    if (!d) goto out;
    for (int f = 0; f < nf; ++f) {
        const int keyint = vs[0].keyint;
        const int is_intra = f == 0 || (keyint > 0 && f % keyint == 0);
        for (int b = 0; b < nb; ++b) {
            in[b] = vs[b].blob[f];
            len[b] = vs[b].blen[f];
            dst[b] = frame_out[b] + (size_t)f * frame_bytes[b];
        }
        if (is_intra) {
            /* tables and history restart at a key frame (fdv.h:6887) */
            memset(tc, 0, (size_t)nb * sizeof *tc);
            int nt = nthreads < 1 ? 1 : nthreads > nb ? nb : nthreads;
            if (nt > 64) nt = 64;
            pthread_t tid[64];
            ikjob jobs[64];
            int spawned = 0;
            for (int t = 0; t < nt; ++t) jobs[t] = (ikjob){idec, vs, f, nb, t, nt, 0, dst};
            for (int t = 1; t < nt; ++t) {
                if (pthread_create(&tid[t], NULL, ikjob_run, &jobs[t]) != 0) break;
                spawned = t;
            }
            ikjob_run(&jobs[0]);
            for (int t = 1; t <= spawned; ++t) pthread_join(tid[t], NULL);
            for (int t = spawned + 1; t < nt; ++t) ikjob_run(&jobs[t]);
            for (int t = 0; t < nt; ++t)
                if (jobs[t].rc) { fdvamd_seterr_("frame %d: malformed key frame", f); goto out; }
            for (int b = 0; b < nb; ++b)
                if (fdvamd_decoder_set_ref(d, b, dst[b]) != 0) goto out;
            nref = 1;
        } else {
            for (int b = 0; b < nb; ++b) nav[b] = nref >= 2 ? 2 : 1;
            if (fdvamd_pframe_decode_batch(d, in, len, nav, tcp, dst) != 0) goto out;
            nref = nref < 2 ? nref + 1 : 2;
        }
    }
    rc = 0;
out:
    fdvamd_decoder_close(d);
    free(w); free(h); free(nav); free(in); free(len); free(dst); free(tc); free(tcp);
    return rc;
}

int fdvamd_video_decode(const uint8_t *in, size_t len, uint8_t *out, int *nframes, int *w, int *h) {
    if (!in || !out) { fdvamd_seterr_("fdvamd_video_decode: bad arguments"); return -1; }
    vstream v;
    if (vstream_locate(&v, in, len) != 0) {
        vstream_free(&v);
        fdvamd_seterr_("fdvamd_video_decode: malformed stream");
        return -1;
    }
    size_t fb = (size_t)v.w * v.h + 2 * (size_t)(v.w / 2) * (v.h / 2);
    uint8_t *dst = out;
    const int rc = decode_lockstep(&v, 1, 1, &dst, &fb);
    if (rc == 0) {
        if (nframes) *nframes = v.nf;
        if (w) *w = v.w;
        if (h) *h = v.h;
    }
    vstream_free(&v);
    return rc;
}

int fdvamd_vtile_decode(const uint8_t *in, size_t len, uint8_t *out, int *nframes, int *w, int *h, int nthreads) {
    if (!in || !out || len < 12) { fdvamd_seterr_("fdvamd_vtile_decode: bad arguments"); return -1; }
    size_t p = 0;
    const int ww = (int)get16(in, &p), hh = (int)get16(in, &p), nf = (int)get16(in, &p);
    p += 2;                                                    /* qp, keyint: each band carries its own */
    const int nb = (int)get16(in, &p), band_mbrows = (int)get16(in, &p);
    /* the header checks of fdv_vtile_decode (fdv.h:7290-7296) */
    if (nf <= 0 || ww <= 0 || hh <= 0 || (ww & 15) || (hh & 15) || ww > 8192 || hh > 8192 || band_mbrows < 1) {
        fdvamd_seterr_("fdvamd_vtile_decode: malformed header");
        return -1;
    }
    const int bh_full = band_mbrows * 16;
    if (nb != (hh + bh_full - 1) / bh_full || nb < 1 || nb > hh / 16) {
        fdvamd_seterr_("fdvamd_vtile_decode: malformed header");
        return -1;
    }
    vstream *vs = calloc((size_t)nb, sizeof *vs);
    uint8_t **sub = calloc((size_t)nb, sizeof *sub);
    size_t *fb = calloc((size_t)nb, sizeof *fb);
    int rc = -1;
    if (!vs || !sub || !fb) { fdvamd_seterr_("out of memory"); goto out; }
    for (int b = 0; b < nb; ++b) {
        const int y0 = b * bh_full, bh = y0 + bh_full <= hh ? bh_full : hh - y0;
        if (p + 4 > len) { fdvamd_seterr_("fdvamd_vtile_decode: truncated"); goto out; }
        const uint32_t bl = get32(in, &p);
        if (p + bl > len) { fdvamd_seterr_("fdvamd_vtile_decode: truncated"); goto out; }
        /* each band is a video stream of this band's size and the frame count
         * (decode_band, fdv.h:7262-7263) */
        if (vstream_locate(&vs[b], in + p, bl) != 0 || vs[b].nf != nf || vs[b].w != ww || vs[b].h != bh) {
            fdvamd_seterr_("fdvamd_vtile_decode: malformed band %d", b);
            goto out;
        }
        p += bl;
        if (vs[b].keyint != vs[0].keyint) {
            /* bands of one tiled stream share the key interval; the lockstep
             * batch relies on it */
            fdvamd_seterr_("fdvamd_vtile_decode: bands with different key intervals");
            goto out;
        }
        fb[b] = (size_t)ww * bh + 2 * (size_t)(ww / 2) * (bh / 2);
        sub[b] = malloc(fb[b] * (size_t)nf);
        if (!sub[b]) { fdvamd_seterr_("out of memory"); goto out; }
    }
    if (decode_lockstep(vs, nb, nthreads, sub, fb) != 0) goto out;
    /* band sub-frames into the whole frames (decode_band, fdv.h:7265-7276) */
    {
        const int cw = ww / 2, ch = hh / 2;
        const size_t fsize = (size_t)ww * hh + 2 * (size_t)cw * ch;
        for (int b = 0; b < nb; ++b) {
            const int y0 = b * bh_full, bh = vs[b].h, bch = bh / 2, cy0 = y0 / 2;
            for (int f = 0; f < nf; ++f) {
                const uint8_t *s = sub[b] + (size_t)f * fb[b];
                uint8_t *fy = out + (size_t)f * fsize, *fu = fy + (size_t)ww * hh, *fv = fu + (size_t)cw * ch;
                memcpy(fy + (size_t)y0 * ww, s, (size_t)ww * bh);
                memcpy(fu + (size_t)cy0 * cw, s + (size_t)ww * bh, (size_t)cw * bch);
                memcpy(fv + (size_t)cy0 * cw, s + (size_t)ww * bh + (size_t)cw * bch, (size_t)cw * bch);
            }
        }
    }
    if (nframes) *nframes = nf;
    if (w) *w = ww;
    if (h) *h = hh;
    rc = 0;
out:
    if (vs) for (int b = 0; b < nb; ++b) vstream_free(&vs[b]);
    if (sub) for (int b = 0; b < nb; ++b) free(sub[b]);
    free(vs); free(sub); free(fb);
    return rc;
}
