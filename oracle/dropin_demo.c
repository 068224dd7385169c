/* dropin_demo.c -- a caller of the reference's sequence API switched to the
 * MI355X path by one include (include/fdv_amd_dropin.h), checked against the
 * reference's own coders in the same program.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference's
 * header where it lies (REF_DIR) into oracle/_ref/; run by
 * tests/test_gpu_dropin.py.  Usage: dropin_demo W H NFRAMES QP KEYINT BAND_MBROWS
 * Prints "dropin ok ..." and exits 0 when the drop-in's streams are byte-
 * identical to the reference's and both decoders return the same frames.
 */
#define FDV_IMPLEMENTATION
#include "fdv.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* the reference's coders, taken before the include renames the calls below */
static size_t (*const ref_video_encode)(const uint8_t *const *, int, int, int, int, int, uint8_t *, size_t) =
    fdv_video_encode;
static int (*const ref_video_decode)(const uint8_t *, size_t, uint8_t *, int *, int *, int *) = fdv_video_decode;
static size_t (*const ref_vtile_encode)(const uint8_t *const *, int, int, int, int, int, int, int, uint8_t *,
                                        size_t) = fdv_vtile_encode;
static int (*const ref_vtile_decode)(const uint8_t *, size_t, uint8_t *, int *, int *, int *, int) =
    fdv_vtile_decode;

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
    if (fdvamd_dropin_register() != 0) return 4;
    int fails = 0;
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
    if (!fails) printf("dropin ok %dx%d %d frames qp %d keyint %d\n", w, h, n, qp, keyint);
    free(frames); free(a); free(b); free(da); free(db); free(fp);
    return fails ? 1 : 0;
}
