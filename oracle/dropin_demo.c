/* dropin_demo.c -- a caller of the reference's sequence, container and
 * streaming API switched to the MI355X path by one include
 * (include/fdv_amd_dropin.h), checked against the reference's own functions in
 * the same program.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference's
 * header where it lies (REF_DIR) into oracle/_ref/; run by
 * tests/test_gpu_dropin.py.  Usage: dropin_demo W H NFRAMES QP KEYINT BAND_MBROWS
 * Prints "dropin ok ..." and exits 0 when the drop-in's output is byte-
 * identical to the reference's -- the plane codec, the whole-sequence coders,
 * and the streaming encoder at a fixed QP and under bitrate control,
 * single-stream and tiled, with held frames -- and both decoders return the
 * same frames, the streaming ones also after seeks.  Key frames go through the
 * drop-in's own coders (nothing is registered).
 */
#define FDV_IMPLEMENTATION
#include "fdv.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* the reference's coders, taken before the include renames the calls below */
static size_t (*const ref_image_encode)(const uint8_t *, int, int, int, int, uint8_t *, size_t, uint8_t *) =
    fdv_image_encode;
static int (*const ref_image_decode)(const uint8_t *, size_t, uint8_t *, int, int *, int *) = fdv_image_decode;
static size_t (*const ref_video_encode)(const uint8_t *const *, int, int, int, int, int, uint8_t *, size_t) =
    fdv_video_encode;
static int (*const ref_video_decode)(const uint8_t *, size_t, uint8_t *, int *, int *, int *) = fdv_video_decode;
static size_t (*const ref_vtile_encode)(const uint8_t *const *, int, int, int, int, int, int, int, uint8_t *,
                                        size_t) = fdv_vtile_encode;
static int (*const ref_vtile_decode)(const uint8_t *, size_t, uint8_t *, int *, int *, int *, int) =
    fdv_vtile_decode;
static fdv_encoder *(*const ref_enc_open)(int, int, int, int) = fdv_enc_open;
static fdv_encoder *(*const ref_enc_open_tiled)(int, int, int, int, int, int) = fdv_enc_open_tiled;
static int (*const ref_enc_frame)(fdv_encoder *, const uint8_t *) = fdv_enc_frame;
static int (*const ref_enc_repeat)(fdv_encoder *) = fdv_enc_repeat;
static int (*const ref_enc_set_bitrate)(fdv_encoder *, int, int, double) = fdv_enc_set_bitrate;
static uint8_t *(*const ref_enc_finish)(fdv_encoder *, int, size_t *) = fdv_enc_finish;
static int (*const ref_read)(const uint8_t *, size_t, fdv_info *) = fdv_read;
static fdv_decoder *(*const ref_dec_open)(const fdv_info *) = fdv_dec_open;
static int (*const ref_dec_next)(fdv_decoder *, uint8_t *) = fdv_dec_next;
static int (*const ref_dec_seek)(fdv_decoder *, int) = fdv_dec_seek;
static void (*const ref_dec_close)(fdv_decoder *) = fdv_dec_close;

#include "fdv_amd_dropin.h"

/* Seeded content: a textured background panning (+2, -1) px a frame and a
 * brighter square moving (+3, +1), with light noise. */
static void make_frames(uint8_t *buf, int w, int h, int n) {
    const size_t fs = (size_t)w * h * 3 / 2;
    uint32_t s = 12345u;
    for (int f = 0; f < n; ++f) {
        uint8_t *y = buf + (size_t)f * fs, *u = y + (size_t)w * h, *v = u + (size_t)(w / 2) * (h / 2);
        for (int r = 0; r < h; ++r)
            for (int c = 0; c < w; ++c) {
                const int X = c + 2 * f, Y = r - f;
                int p = 128 + (int)(48 * ((X * 7 + Y * 13) % 31) / 31) + ((X / 8 + Y / 8) & 1) * 24;
                if (c >= 20 + 3 * f % (w - 40) && c < 52 + 3 * f % (w - 40) && r >= 16 + f % (h - 40) &&
                    r < 48 + f % (h - 40))
                    p = 60 + (c * 5 + r * 3) % 40;
                s = s * 1664525u + 1013904223u;
                p += (int)(s >> 29) - 4;
                y[(size_t)r * w + c] = (uint8_t)(p < 0 ? 0 : p > 255 ? 255 : p);
            }
        for (int r = 0; r < h / 2; ++r)
            for (int c = 0; c < w / 2; ++c) {
                u[(size_t)r * (w / 2) + c] = (uint8_t)(96 + (c + f) % 64);
                v[(size_t)r * (w / 2) + c] = (uint8_t)(160 - (r + f) % 48);
            }
    }
}

/* The streaming encoder through unchanged call sites against the reference's,
 * frames 3 and 5 preceded by a held frame (5 is a key-frame slot when keyint
 * divides it); then both files through both streaming decoders, frame by
 * frame and after seeks.  Returns the number of mismatches. */
static int stream_case(const uint8_t *const *fp, int n, int w, int h, int qp, int keyint, int tiled, int bmr, int bps) {
    fdv_encoder *e = tiled ? fdv_enc_open_tiled(w, h, qp, keyint, bmr, 4) : fdv_enc_open(w, h, qp, keyint);
    fdv_encoder *r = tiled ? ref_enc_open_tiled(w, h, qp, keyint, bmr, 4) : ref_enc_open(w, h, qp, keyint);
    if (!e || !r) { printf("stream open failed: %s\n", fdvamd_last_error()); return 1; }
    if (bps && (fdv_enc_set_bitrate(e, bps, 30, 1.0) != 0 || ref_enc_set_bitrate(r, bps, 30, 1.0) != 0)) return 1;
    int total = 0;
    for (int f = 0; f < n; ++f) {
        if (f == 3 || f == 5) {
            if (fdv_enc_repeat(e) != 0 || ref_enc_repeat(r) != 0) { printf("repeat failed\n"); return 1; }
            ++total;
        }
        if (fdv_enc_frame(e, fp[f]) != 0 || ref_enc_frame(r, fp[f]) != 0) {
            printf("frame %d failed: %s\n", f, fdvamd_last_error());
            return 1;
        }
        ++total;
    }
    size_t la = 0, lb = 0;
    uint8_t *a = fdv_enc_finish(e, 30, &la), *b = ref_enc_finish(r, 30, &lb);
    if (!a || !b) { printf("finish failed\n"); return 1; }
    const int same = la == lb && !memcmp(a, b, la);
    fdv_info ia, ib;
    int dsame = fdv_read(a, la, &ia) == 0 && ref_read(b, lb, &ib) == 0 && ia.nframes == total;
    fdv_decoder *da = dsame ? fdv_dec_open(&ia) : NULL, *db = dsame ? ref_dec_open(&ib) : NULL;
    const size_t fs = (size_t)w * h * 3 / 2;
    uint8_t *x = malloc(fs), *y = malloc(fs);
    if (!da || !db || !x || !y) dsame = 0;
    for (int k = 0; dsame && k < total; ++k)
        dsame = fdv_dec_next(da, x) == 1 && ref_dec_next(db, y) == 1 && !memcmp(x, y, fs);
    dsame = dsame && fdv_dec_next(da, x) == 0 && ref_dec_next(db, y) == 0;
    const int seeks[4] = {total - 2, 1, total / 2, total / 2 + 1};
    for (int k = 0; dsame && k < 4; ++k)
        dsame = fdv_dec_seek(da, seeks[k]) == 0 && ref_dec_seek(db, seeks[k]) == 0 && fdv_dec_next(da, x) == 1 &&
                ref_dec_next(db, y) == 1 && !memcmp(x, y, fs);
    printf("fdv_enc_%s%s: %zu B vs the reference's %zu B, streams %s, streaming decode and seeks %s\n",
           tiled ? "open_tiled" : "open", bps ? " + fdv_enc_set_bitrate" : "", la, lb, same ? "identical" : "DIFFER",
           dsame ? "identical" : "DIFFER");
    if (da) fdv_dec_close(da);
    if (db) ref_dec_close(db);
    free(a); free(b); free(x); free(y);
    return !same + !dsame;
}

int main(int argc, char **argv) {
    if (argc != 7) { fprintf(stderr, "usage: %s W H NFRAMES QP KEYINT BAND_MBROWS\n", argv[0]); return 2; }
    const int w = atoi(argv[1]), h = atoi(argv[2]), n = atoi(argv[3]), qp = atoi(argv[4]);
    const int keyint = atoi(argv[5]), bmr = atoi(argv[6]);
    const size_t fs = (size_t)w * h * 3 / 2, cap = fs * (size_t)n * 2 + 65536;
    uint8_t *frames = malloc(fs * (size_t)n), *a = malloc(cap), *b = malloc(cap);
    uint8_t *da = malloc(fs * (size_t)n), *db = malloc(fs * (size_t)n);
    const uint8_t **fp = malloc((size_t)n * sizeof *fp);
    if (!frames || !a || !b || !da || !db || !fp) return 3;
    make_frames(frames, w, h, n);
    for (int f = 0; f < n; ++f) fp[f] = frames + (size_t)f * fs;
    int fails = 0;
    {
        /* an unchanged call site of the plane codec: frame 0's luma */
        uint8_t *ra = malloc((size_t)w * h), *rb = malloc((size_t)w * h), *xa = malloc((size_t)w * h);
        uint8_t *xb = malloc((size_t)w * h);
        if (!ra || !rb || !xa || !xb) return 3;
        const size_t na = fdv_image_encode(frames, w, h, w, qp, a, cap, ra);
        const size_t nb = ref_image_encode(frames, w, h, w, qp, b, cap, rb);
        int w1, h1, w2, h2;
        const int da_ = fdv_image_decode(a, na, xa, w, &w1, &h1), db_ = ref_image_decode(b, nb, xb, w, &w2, &h2);
        const int same = na && na == nb && !memcmp(a, b, na) && !memcmp(ra, rb, (size_t)w * h);
        const int dsame = da_ == 0 && db_ == 0 && !memcmp(xa, xb, (size_t)w * h) && !memcmp(xa, ra, (size_t)w * h);
        printf("fdv_image_encode: %zu B vs the reference's %zu B, planes %s, decoded %s\n", na, nb,
               same ? "identical" : "DIFFER", dsame ? "identical" : "DIFFER");
        fails += !same + !dsame;
        free(ra); free(rb); free(xa); free(xb);
    }
    for (int tiled = 0; tiled < 2; ++tiled) {
        /* an unchanged call site: fdv_video_encode / fdv_vtile_encode */
        const size_t na = tiled ? fdv_vtile_encode(fp, n, w, h, qp, keyint, bmr, 4, a, cap)
                                : fdv_video_encode(fp, n, w, h, qp, keyint, a, cap);
        const size_t nb = tiled ? ref_vtile_encode(fp, n, w, h, qp, keyint, bmr, 4, b, cap)
                                : ref_video_encode(fp, n, w, h, qp, keyint, b, cap);
        int nfa = 0, wa = 0, ha = 0, nfb = 0, wb = 0, hb = 0;
        const int ra = tiled ? fdv_vtile_decode(a, na, da, &nfa, &wa, &ha, 4) : fdv_video_decode(a, na, da, &nfa, &wa, &ha);
        const int rb = tiled ? ref_vtile_decode(b, nb, db, &nfb, &wb, &hb, 4) : ref_video_decode(b, nb, db, &nfb, &wb, &hb);
        const int same = na && na == nb && !memcmp(a, b, na);
        const int dsame = ra == 0 && rb == 0 && nfa == n && !memcmp(da, db, fs * (size_t)n);
        printf("%s: %zu B vs the reference's %zu B, streams %s, decoded frames %s (%s)\n",
               tiled ? "fdv_vtile_encode" : "fdv_video_encode", na, nb, same ? "identical" : "DIFFER",
               dsame ? "identical" : "DIFFER", na ? "" : fdvamd_last_error());
        fails += !same + !dsame;
    }
    for (int tiled = 0; tiled < 2; ++tiled)
        for (int rc = 0; rc < 2; ++rc)
            fails += stream_case(fp, n, w, h, qp, keyint, tiled, bmr, rc ? w * h * 30 / 200 : 0);
    if (!fails) printf("dropin ok %dx%d %d frames qp %d keyint %d\n", w, h, n, qp, keyint);
    free(frames); free(a); free(b); free(da); free(db); free(fp);
    return fails ? 1 : 0;
}
