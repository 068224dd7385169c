/* fdv_video.c -- the P-frame encoder on the MI355X path: device analysis through
 * the fdvgpu C ABI, then host entropy packing, one thread per sequence.
 *
 * Reference behaviour mirrored: pframe_encode's contract (fdv.h:5764-6421) and
 * the near/far reference pool of fdv_video_encode (fdv.h:6749-6754, 6842-6846),
 * which here lives in device memory and rotates on every P-frame.
 */
#include "fdv_amd.h"
#include "fdvgpu.h"

#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static _Thread_local char amd_err[256];
static void amd_seterr(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(amd_err, sizeof amd_err, fmt, ap);
    va_end(ap);
}
const char *fdvamd_last_error(void) { return amd_err; }
/* The same message slot, for the library's other host files. */
void fdvamd_seterr_(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(amd_err, sizeof amd_err, fmt, ap);
    va_end(ap);
}

typedef struct {
    fdvamd_encoder *e;
    const int *qp;
    fdvamd_tabcache *const *tc;
    uint8_t *const *out;
    const size_t *cap;
    size_t *len;
    int slot;                /* -1: the last fdvgpu_pframe; 0/1: a collected slot */
    int first, step, rc;
} packjob;

/* Packing threads, created with the encoder: each frame step hands them one
 * job each (generation counter) and waits until all are done. */
typedef struct {
    pthread_mutex_t mu;
    pthread_cond_t go, done;
    pthread_t tid[15];
    int nw;                  /* worker threads running */
    unsigned gen;
    int pending, quit;
    packjob jobs[16];
} packpool;

struct fdvamd_encoder {
    fdvgpu *g;
    int nseq, me_mode;
    int *w, *h;
    uint8_t *tmp;            /* contiguous I420 staging for fdvamd_pframe_encode */
    /* pipelined frames: up to two submitted and not yet collected, the oldest
     * in slot pipe_head; each slot's QPs, for its packing */
    int pipe_n, pipe_head;
    int *pipe_qp[2];
    packpool *pool;
};

static void *pack_worker(void *arg);

typedef struct { packpool *p; int k; } poolarg;

static void *pool_main(void *arg) {
    poolarg a = *(poolarg *)arg;
    free(arg);
    packpool *p = a.p;
    unsigned seen = 0;
    for (;;) {
        pthread_mutex_lock(&p->mu);
        while (!p->quit && p->gen == seen) pthread_cond_wait(&p->go, &p->mu);
        if (p->quit) { pthread_mutex_unlock(&p->mu); return NULL; }
        seen = p->gen;
        pthread_mutex_unlock(&p->mu);
        pack_worker(&p->jobs[a.k]);
        pthread_mutex_lock(&p->mu);
        if (--p->pending == 0) pthread_cond_signal(&p->done);
        pthread_mutex_unlock(&p->mu);
    }
}

static packpool *pool_open(int nseq) {
    packpool *p = calloc(1, sizeof *p);
    if (!p) return NULL;
    pthread_mutex_init(&p->mu, NULL);
    pthread_cond_init(&p->go, NULL);
    pthread_cond_init(&p->done, NULL);
    const int want = (nseq < 16 ? nseq : 16) - 1;
    for (int k = 1; k <= want; ++k) {
        poolarg *a = malloc(sizeof *a);
        if (!a) break;
        *a = (poolarg){p, k};
        if (pthread_create(&p->tid[k - 1], NULL, pool_main, a) != 0) { free(a); break; }
        p->nw = k;
    }
    return p;
}

static void pool_close(packpool *p) {
    if (!p) return;
    pthread_mutex_lock(&p->mu);
    p->quit = 1;
    pthread_cond_broadcast(&p->go);
    pthread_mutex_unlock(&p->mu);
    for (int k = 0; k < p->nw; ++k) pthread_join(p->tid[k], NULL);
    pthread_mutex_destroy(&p->mu);
    pthread_cond_destroy(&p->go);
    pthread_cond_destroy(&p->done);
    free(p);
}

fdvamd_encoder *fdvamd_encoder_open(int nseq, const int *w, const int *h, int me_mode) {
    if (nseq <= 0 || !w || !h || me_mode < 0 || me_mode > 3) { amd_seterr("bad arguments"); return NULL; }
    fdvamd_encoder *e = calloc(1, sizeof *e);
    if (!e) return NULL;
    e->nseq = nseq; e->me_mode = me_mode;
    e->w = malloc(nseq * sizeof(int)); e->h = malloc(nseq * sizeof(int));
    if (!e->w || !e->h) { fdvamd_encoder_close(e); return NULL; }
    memcpy(e->w, w, nseq * sizeof(int)); memcpy(e->h, h, nseq * sizeof(int));
    e->pipe_qp[0] = malloc(nseq * sizeof(int)); e->pipe_qp[1] = malloc(nseq * sizeof(int));
    e->pool = pool_open(nseq);
    if (!e->pipe_qp[0] || !e->pipe_qp[1] || !e->pool) { fdvamd_encoder_close(e); return NULL; }
    e->g = fdvgpu_open(nseq, w, h, 1);
    if (!e->g) { amd_seterr("%s", fdvgpu_last_error()); fdvamd_encoder_close(e); return NULL; }
    return e;
}

void fdvamd_encoder_close(fdvamd_encoder *e) {
    if (!e) return;
    pool_close(e->pool);
    fdvgpu_close(e->g);
    free(e->w); free(e->h); free(e->tmp); free(e->pipe_qp[0]); free(e->pipe_qp[1]);
    free(e);
}

int fdvamd_set_ref(fdvamd_encoder *e, int seq, int which, const uint8_t *i420) {
    if (!e || fdvgpu_put_ref(e->g, seq, which, i420) != 0) {
        amd_seterr("%s", fdvgpu_last_error()); return -1;
    }
    return 0;
}

int fdvamd_set_ref_planes(fdvamd_encoder *e, int seq, int which,
                          const uint8_t *y, const uint8_t *u, const uint8_t *v) {
    if (!e || seq < 0 || seq >= e->nseq || !y || !u || !v) { amd_seterr("bad arguments"); return -1; }
    const size_t ys = (size_t)e->w[seq] * e->h[seq], cs = (size_t)(e->w[seq] / 2) * (e->h[seq] / 2);
    uint8_t *t = malloc(ys + 2 * cs);
    if (!t) { amd_seterr("out of memory"); return -1; }
    memcpy(t, y, ys); memcpy(t + ys, u, cs); memcpy(t + ys + cs, v, cs);
    int rc = fdvamd_set_ref(e, seq, which, t);
    free(t);
    return rc;
}

static void *pack_worker(void *arg) {
    packjob *j = arg;
    for (int s = j->first; s < j->e->nseq; s += j->step) {
        const uint8_t *sv[16];
        size_t sn[16];
        const int got = j->slot < 0 ? fdvgpu_streams(j->e->g, s, sv, sn)
                                    : fdvgpu_collected_streams(j->e->g, j->slot, s, sv, sn);
        if (got != 0) { j->len[s] = 0; j->rc = -1; continue; }
        j->len[s] = fdvamd_pack_pframe(j->qp[s], sv, sn, j->tc ? j->tc[s] : NULL, j->out[s], j->cap[s]);
        if (!j->len[s]) j->rc = -1;
    }
    return NULL;
}

/* Host entropy of every sequence: the sequences are independent, so they pack
 * in parallel, one thread per sequence up to 16. */
static int pack_all(fdvamd_encoder *e, int slot, const int *qp, fdvamd_tabcache *const *tc, uint8_t *const *out,
                    const size_t *cap, size_t *len) {
    packpool *p = e->pool;
    const int nt = p->nw + 1;
    int rc = 0;
    for (int t = 0; t < nt; ++t)
        p->jobs[t] = (packjob){.e = e, .qp = qp, .tc = tc, .out = out, .cap = cap, .len = len,
                               .slot = slot, .first = t, .step = nt, .rc = 0};
    pthread_mutex_lock(&p->mu);
    p->pending = p->nw;
    p->gen++;
    pthread_cond_broadcast(&p->go);
    pthread_mutex_unlock(&p->mu);
    pack_worker(&p->jobs[0]);
    pthread_mutex_lock(&p->mu);
    while (p->pending) pthread_cond_wait(&p->done, &p->mu);
    pthread_mutex_unlock(&p->mu);
    for (int t = 0; t < nt; ++t) if (p->jobs[t].rc) rc = -1;
    return rc;
}

int fdvamd_pframe_submit(fdvamd_encoder *e, const uint8_t *const *frames, const int *navail, const int *qp,
                         const double *lscale, const int *force_skip) {
    if (!e || !frames || !navail || !qp) { amd_seterr("bad arguments"); return -1; }
    if (e->pipe_n >= 2) { amd_seterr("two frames in flight: collect one first"); return -1; }
    double *ls = malloc((size_t)e->nseq * sizeof(double));
    if (!ls) { amd_seterr("out of memory"); return -1; }
    for (int s = 0; s < e->nseq; ++s) ls[s] = lscale ? lscale[s] : 1.0;
    const int slot = (e->pipe_head + e->pipe_n) % 2;
    const int rc = fdvgpu_submit(e->g, slot, frames, qp, ls, navail, force_skip, e->me_mode);
    free(ls);
    if (rc != 0) { amd_seterr("%s", fdvgpu_last_error()); return -1; }
    memcpy(e->pipe_qp[slot], qp, (size_t)e->nseq * sizeof(int));
    e->pipe_n++;
    return 0;
}

int fdvamd_pframe_collect(fdvamd_encoder *e, fdvamd_tabcache *const *tc, uint8_t *const *out, const size_t *cap,
                          size_t *len) {
    if (!e || !out || !cap || !len) { amd_seterr("bad arguments"); return -1; }
    if (!e->pipe_n) { amd_seterr("no frame submitted"); return -1; }
    const int slot = e->pipe_head;
    e->pipe_head ^= 1;
    e->pipe_n--;
    if (fdvgpu_collect(e->g, slot) != 0) { amd_seterr("%s", fdvgpu_last_error()); return -1; }
    const int rc = pack_all(e, slot, e->pipe_qp[slot], tc, out, cap, len);
    if (rc) amd_seterr("a sequence failed to pack");
    return rc;
}

int fdvamd_pframe_batch(fdvamd_encoder *e, const uint8_t *const *frames, const int *navail,
                        const int *qp, const double *lscale, const int *force_skip,
                        fdvamd_tabcache *const *tc, uint8_t *const *out, const size_t *cap,
                        size_t *len, uint8_t *const *recon) {
    if (!e || !frames || !navail || !qp || !out || !cap || !len) { amd_seterr("bad arguments"); return -1; }
    if (e->pipe_n) { amd_seterr("submitted frames not collected"); return -1; }
    int *slots = calloc((size_t)e->nseq, sizeof(int));
    double *ls = malloc((size_t)e->nseq * sizeof(double));
    if (!slots || !ls) { free(slots); free(ls); amd_seterr("out of memory"); return -1; }
    for (int s = 0; s < e->nseq; ++s) ls[s] = lscale ? lscale[s] : 1.0;
    int rc = 0;
    for (int s = 0; s < e->nseq && rc == 0; ++s) {
        /* a held frame (force_skip) reads no source: it may have none */
        if (!frames[s] && !(force_skip && force_skip[s])) { amd_seterr("sequence %d has no frame", s); rc = -2; }
        else if (frames[s] && fdvgpu_put_source(e->g, s, 0, frames[s]) != 0) rc = -1;
    }
    if (rc == 0 && fdvgpu_pframe(e->g, slots, qp, ls, navail, force_skip, e->me_mode) != 0) rc = -1;
    if (rc == 0 && fdvgpu_fetch(e->g) != 0) rc = -1;
    free(slots); free(ls);
    if (rc == -2) return -1;
    if (rc != 0) { amd_seterr("%s", fdvgpu_last_error()); return -1; }

    rc = pack_all(e, -1, qp, tc, out, cap, len);
    if (recon)
        for (int s = 0; s < e->nseq; ++s)
            if (recon[s] && fdvgpu_recon(e->g, s, recon[s]) != 0) { amd_seterr("%s", fdvgpu_last_error()); rc = -1; }
    if (rc) amd_seterr("a sequence failed to pack");
    return rc;
}

size_t fdvamd_pframe_encode(fdvamd_encoder *e, const uint8_t *cy, const uint8_t *cu,
                            const uint8_t *cv, int w, int h, int navail, int qp,
                            uint8_t *out, size_t cap, uint8_t *ry, uint8_t *ru, uint8_t *rv,
                            int force_skip, fdvamd_tabcache *tc, double lscale) {
    if (!e) { amd_seterr("bad arguments"); return 0; }
    if (e->nseq != 1) { amd_seterr("fdvamd_pframe_encode needs a one-sequence encoder (use fdvamd_pframe_batch)"); return 0; }
    if (w != e->w[0] || h != e->h[0]) { amd_seterr("frame size does not match the encoder"); return 0; }
    const size_t ys = (size_t)w * h, cs = (size_t)(w / 2) * (h / 2);
    if (!e->tmp && !(e->tmp = malloc(ys + 2 * cs))) { amd_seterr("out of memory"); return 0; }
    memcpy(e->tmp, cy, ys); memcpy(e->tmp + ys, cu, cs); memcpy(e->tmp + ys + cs, cv, cs);
    const uint8_t *fr[1] = {e->tmp};
    uint8_t *o[1] = {out};
    size_t len = 0;
    fdvamd_tabcache *t[1] = {tc};
    uint8_t *rec = NULL;
    if (ry || ru || rv) {
        rec = malloc(ys + 2 * cs);
        if (!rec) { amd_seterr("out of memory"); return 0; }
    }
    uint8_t *recs[1] = {rec};
    int fs = force_skip;
    if (fdvamd_pframe_batch(e, fr, &navail, &qp, &lscale, &fs, t, o, &cap, &len, recs) != 0) {
        free(rec); return 0;
    }
    if (rec) {
        if (ry) memcpy(ry, rec, ys);
        if (ru) memcpy(ru, rec + ys, cs);
        if (rv) memcpy(rv, rec + ys + cs, cs);
        free(rec);
    }
    return len;
}

int fdvamd_rewind(fdvamd_encoder *e) {
    if (e && e->pipe_n) { amd_seterr("submitted frames not collected"); return -1; }
    if (!e || fdvgpu_rewind(e->g) != 0) { amd_seterr("%s", e ? fdvgpu_last_error() : "bad arguments"); return -1; }
    return 0;
}

int fdvamd_rewind_keep_motion(fdvamd_encoder *e) {
    if (e && e->pipe_n) { amd_seterr("submitted frames not collected"); return -1; }
    if (!e || fdvgpu_rewind_keep_motion(e->g) != 0) {
        amd_seterr("%s", e ? fdvgpu_last_error() : "bad arguments"); return -1;
    }
    return 0;
}

int fdvamd_set_proposer(fdvamd_encoder *e, const float *weights, int n) {
    if (!e || fdvgpu_set_proposer(e->g, weights, n) != 0) {
        amd_seterr("%s", e ? fdvgpu_last_error() : "bad arguments"); return -1;
    }
    return 0;
}

int fdvamd_set_proposer_thresholds(fdvamd_encoder *e, float conf, float skip) {
    if (!e || fdvgpu_set_proposer_thresholds(e->g, conf, skip) != 0) {
        amd_seterr("%s", e ? fdvgpu_last_error() : "bad arguments"); return -1;
    }
    return 0;
}
