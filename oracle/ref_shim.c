/* ref_shim.c -- builds the reference codec (yhbsh/fdv, a single-header C11
 * library) from its own source where it lies, and exports thin C wrappers over
 * the P-frame internals that the header keeps file-static, so the parity tests
 * can drive them with arbitrary inputs.
 *
 * TEST INFRASTRUCTURE ONLY.  This file contains no codec logic of its own: it
 * includes include/fdv.h with FDV_IMPLEMENTATION defined and forwards to
 * functions defined there.  The recipe in oracle/Makefile compiles it with
 * -I pointing at the reference tree; the output goes to oracle/_ref/ only.
 *
 * Reference entry points wrapped (include/fdv.h):
 *   pframe_encode      5764    pframe_decode     6492
 *   code_residual      5417    chroma_block      5631
 *   load_ref           5183    c8_init / C8 / QSTEP8 2585-2612
 *   iframe_encode      5487    iframe_decode     5541
 */
#define FDV_IMPLEMENTATION
#include "fdv.h"

/* Padded reference frame from tight I420 planes (load_ref, fdv.h:5183). */
static int shim_ref_from_i420(fdv_frame *f, const uint8_t *y, const uint8_t *u,
                              const uint8_t *v, int w, int h) {
    if (fdv_frame_alloc(f, w, h) != 0) return -1;
    load_ref(f, y, u, v, w, h);
    return 0;
}

/* pframe_encode over a whole frame (y0=0, bh=h) with references given as tight
 * I420 planes.  tc_blob (may be NULL) is an fdv_tabcache carried between calls:
 * the caller owns sizeof(fdv_tabcache) bytes, zeroed at a key frame. */
size_t fdvref_pframe_encode(const uint8_t *cy, const uint8_t *cu, const uint8_t *cv,
                            int w, int h,
                            const uint8_t *r0y, const uint8_t *r0u, const uint8_t *r0v,
                            const uint8_t *r1y, const uint8_t *r1u, const uint8_t *r1v,
                            int navail, int qp, uint8_t *out, size_t cap,
                            uint8_t *ry, uint8_t *ru, uint8_t *rv, int force_skip,
                            void *tc_blob, double lscale) {
    fdv_frame f0, f1;
    if (shim_ref_from_i420(&f0, r0y, r0u, r0v, w, h) != 0) return 0;
    if (shim_ref_from_i420(&f1, r1y ? r1y : r0y, r1u ? r1u : r0u, r1v ? r1v : r0v, w, h) != 0) {
        fdv_frame_free(&f0); return 0;
    }
    const fdv_frame *refs[2] = {&f0, navail >= 2 ? &f1 : &f0};
    size_t n = pframe_encode(cy, cu, cv, w, h, 0, h, refs, navail, qp, out, cap,
                             ry, ru, rv, force_skip, (fdv_tabcache *)tc_blob, lscale);
    fdv_frame_free(&f0); fdv_frame_free(&f1);
    return n;
}

int fdvref_pframe_decode(const uint8_t *in, size_t len, int w, int h,
                         const uint8_t *r0y, const uint8_t *r0u, const uint8_t *r0v,
                         const uint8_t *r1y, const uint8_t *r1u, const uint8_t *r1v,
                         int navail, uint8_t *dy, uint8_t *du, uint8_t *dv, void *tc_blob) {
    fdv_frame f0, f1;
    if (shim_ref_from_i420(&f0, r0y, r0u, r0v, w, h) != 0) return -1;
    if (shim_ref_from_i420(&f1, r1y ? r1y : r0y, r1u ? r1u : r0u, r1v ? r1v : r0v, w, h) != 0) {
        fdv_frame_free(&f0); return -1;
    }
    const fdv_frame *refs[2] = {&f0, navail >= 2 ? &f1 : &f0};
    int rc = pframe_decode(in, len, refs, dy, du, dv, w, h, 0, h, (fdv_tabcache *)tc_blob);
    fdv_frame_free(&f0); fdv_frame_free(&f1);
    return rc;
}

size_t fdvref_tabcache_size(void) { return sizeof(fdv_tabcache); }

/* code_residual (fdv.h:5417): the residual coder for one n x n block.  The
 * seven coefficient streams are returned through caller buffers of `scap`
 * bytes each, in the fdv_cw order n, l, n8, l8, fl, nc, lc. */
void fdvref_code_residual(const uint8_t *cur, int cstride, int ox, int oy,
                          const uint8_t *pred, int n, int qp, double lambda,
                          int chroma, uint8_t *rec, uint8_t *const bufs[7],
                          size_t lens[7], double *ssd, int *bits) {
    fdv_cw w = { .n = bufs[0], .l = bufs[1], .n8 = bufs[2], .l8 = bufs[3],
                 .fl = bufs[4], .nc = bufs[5], .lc = bufs[6], .chroma = chroma };
    code_residual(cur, cstride, ox, oy, pred, n, qp, lambda, rec, &w, ssd, bits);
    lens[0] = w.np; lens[1] = w.lp; lens[2] = w.np8; lens[3] = w.lp8;
    lens[4] = w.flp; lens[5] = w.ncp; lens[6] = w.lcp;
}

void fdvref_c8_tables(int32_t c8[64], int32_t qstep8[52]) {
    pthread_once(&c8_once, c8_init);
    memcpy(c8, C8, sizeof C8);
    memcpy(qstep8, QSTEP8, sizeof QSTEP8);
}

/* Motion search / compensation on a padded luma plane built from a tight one. */
int fdvref_me_search(const uint8_t *cur, int w, int h, const uint8_t *refy,
                     int bx, int by, int bw, int bh, int range, int pmx, int pmy,
                     int *mvx, int *mvy) {
    fdv_frame f;
    if (fdv_frame_alloc(&f, w, h) != 0) return -1;
    for (int r = 0; r < h; ++r) memcpy(fdv_plane_at(&f.planes[0], 0, r), refy + (size_t)r * w, w);
    fdv_frame_extend_borders(&f);
    int s = fdv_me_search(cur, w, &f.planes[0], bx, by, bw, bh, range, pmx, pmy, mvx, mvy);
    fdv_frame_free(&f);
    return s;
}

void fdvref_mc(const uint8_t *refp, int w, int h, int chroma, int bx, int by, int bw, int bh,
               int mvx, int mvy, uint8_t *dst) {
    fdv_frame f;
    if (fdv_frame_alloc(&f, w, h) != 0) return;
    for (int r = 0; r < h; ++r) memcpy(fdv_plane_at(&f.planes[0], 0, r), refp + (size_t)r * w, w);
    fdv_frame_extend_borders(&f);
    if (chroma) fdv_mc_chroma(&f.planes[0], bx, by, bw, bh, mvx, mvy, dst, bw);
    else        fdv_mc_luma(&f.planes[0], bx, by, bw, bh, mvx, mvy, dst, bw);
    fdv_frame_free(&f);
}

/* The key-frame coder over a whole frame (iframe_encode, fdv.h:5487), with the
 * contract of fdvamd_iframe_encode_fn, and its decoder (iframe_decode, 5541). */
size_t fdvref_iframe_encode(const uint8_t *y, const uint8_t *u, const uint8_t *v, int w, int h, int qp,
                            uint8_t *out, size_t cap, uint8_t *ry, uint8_t *ru, uint8_t *rv) {
    return iframe_encode(y, u, v, w, h, 0, h, qp, out, cap, ry, ru, rv);
}

int fdvref_iframe_decode(const uint8_t *in, size_t len, uint8_t *y, uint8_t *u, uint8_t *v, int w, int h) {
    return iframe_decode(in, len, y, u, v, w, h, 0, h);
}
