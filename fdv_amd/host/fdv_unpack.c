/* fdv_unpack.c -- host half of the P-frame decoder: a payload's entropy coding
 * undone into its 16 symbol streams, and the streams parsed into each MB's
 * decision and quantised levels (fdvgpu_mbcode) for the device reconstruction.
 *
 * Restated from the reference decoder, pframe_decode (yhbsh/fdv include/fdv.h
 * 6492-6737): the entropy part (6504-6610: the four codings, table reuse
 * from the cache, the adaptive streams, the history update), the MB syntax
 * (6626-6729: mode contexts, merge lists, predicted vectors, the vector bounds
 * check), and the residual syntax of decode_residual / decode_block4 /
 * chroma_block_dec (5280, 5426, 6429).  Every malformed input the reference
 * rejects is rejected here (-1).
 */
#include "fdv_amd.h"
#include "fdvgpu.h"

#include <stdlib.h>
#include <string.h>

#define NSTR 16
#define NSTRUCT 9
#define SCALE_BITS 12
#define PROB_SCALE (1u << SCALE_BITS)
#define STATE_LO (1u << 23)
enum { S_REF = 4, S_MVX = 5, S_MVY = 6, S_SUB = 7, S_MERGE = 8, S_CN = 9, S_CL = 10, S_CN8 = 11, S_CL8 = 12,
       S_FL = 13, S_CNC = 14, S_CLC = 15 };

/* ---- bounded readers (fdv.h:1711-1740) ---------------------------------- */

static int rd_byte(const uint8_t *s, size_t *p, size_t end, int *ok) {
    if (*p >= end) { *ok = 0; return 0; }
    return s[(*p)++];
}
static uint32_t rd_count32(const uint8_t *s, size_t *p, size_t end, int *ok) {
    uint32_t v = 0;
    for (int sh = 0;; sh += 7) {
        if (*p >= end || sh > 28) { *ok = 0; return 0; }
        const uint8_t c = s[(*p)++];
        v |= (uint32_t)(c & 0x7f) << sh;
        if (!(c & 0x80)) return v;
    }
}
static int rd_count(const uint8_t *s, size_t *p, size_t end, int max, int *ok) {
    const int c = rd_byte(s, p, end, ok);
    if (c > max) { *ok = 0; return 0; }
    return c;
}
static int rd_level(const uint8_t *s, size_t *p, size_t end, int *ok) {
    uint32_t u = 0;
    for (int k = 0, sh = 0; k < 5; ++k, sh += 7) {          /* at most 5 LEB bytes */
        const int c = rd_byte(s, p, end, ok);
        u |= (uint32_t)(c & 0x7f) << sh;
        if (!(c & 0x80)) break;
    }
    return (int)(u >> 1) ^ -(int)(u & 1);
}

/* ---- static rANS decoding (fdv.h:1851-1935) ------------------------------ */

typedef struct { uint16_t freq[256], cum[256]; uint8_t sym[PROB_SCALE]; } rtable;

static void rtable_build(const uint16_t f[256], rtable *t) {
    uint32_t c = 0;
    for (int s = 0; s < 256; ++s) {
        t->freq[s] = f[s];
        t->cum[s] = (uint16_t)c;
        memset(t->sym + c, s, f[s]);
        c += f[s];
    }
}
/* A table as sent: a u16 count, then (symbol, varint frequency) pairs; the
 * frequencies must sum to 4096 with no symbol twice. */
static size_t read_table(const uint8_t *in, size_t p, size_t end, uint16_t f[256], int *ok) {
    memset(f, 0, 256 * sizeof f[0]);
    if (p + 2 > end) { *ok = 0; return p; }
    const int n = in[p] | (int)in[p + 1] << 8;
    p += 2;
    uint32_t sum = 0;
    for (int k = 0; k < n; ++k) {
        if (p >= end) { *ok = 0; return p; }
        const int s = in[p++];
        uint32_t v = 0;
        uint8_t c;
        int sh = 0;
        do {
            if (p >= end || sh > 28) { *ok = 0; return p; }
            c = in[p++];
            v |= (uint32_t)(c & 0x7f) << sh;
            sh += 7;
        } while (c & 0x80);
        if (f[s]) { *ok = 0; return p; }
        f[s] = (uint16_t)v;
        sum += v;
    }
    if (sum != PROB_SCALE) *ok = 0;
    return p;
}
/* n symbols from `in`; refilling stops at the end of the input, so a corrupt
 * stream yields bounded garbage that the parse then rejects. */
static void rans_unpack(const uint8_t *in, size_t len, size_t n, const rtable *t, uint8_t *out) {
    const uint8_t *p = in, *end = in + len;
    uint32_t x = 0;
    for (int b = 0; b < 4; ++b) x |= (uint32_t)(p < end ? *p++ : 0) << (8 * b);
    for (size_t i = 0; i < n; ++i) {
        const uint8_t s = t->sym[x & (PROB_SCALE - 1)];
        out[i] = s;
        x = t->freq[s] * (x >> SCALE_BITS) + (x & (PROB_SCALE - 1)) - t->cum[s];
        while (x < STATE_LO && p < end) x = (x << 8) | *p++;
    }
}

/* The cached form of a table (fdv.h:5667-5677) and the history (5661): as the
 * encoder keeps them. */
static void table_keep(uint16_t dst[256], const uint16_t src[256]) {
    int missing = 0, big = 0;
    for (int s = 0; s < 256; ++s) { if (!src[s]) ++missing; if (src[s] > src[big]) big = s; }
    if (src[big] <= (uint16_t)missing) { memcpy(dst, src, 256 * sizeof dst[0]); return; }
    for (int s = 0; s < 256; ++s) dst[s] = src[s] ? src[s] : 1;
    dst[big] = (uint16_t)(dst[big] - missing);
}
static void history_add(uint32_t h[256], const uint32_t c[256]) {
    uint64_t tot = 0;
    for (int s = 0; s < 256; ++s) { h[s] += c[s]; tot += h[s]; }
    if (tot > (1u << 21)) for (int s = 0; s < 256; ++s) h[s] >>= 1;
}

/* ---- the adaptive range decoder (fdv.h:2024-2143) ------------------------- */

typedef struct { uint32_t w[256], tree[257], total; } amodel;

static void am_rebuild(amodel *m) {
    memset(m->tree, 0, sizeof m->tree);
    m->total = 0;
    for (int i = 0; i < 256; ++i) { m->tree[i + 1] = m->w[i]; m->total += m->w[i]; }
    for (int i = 1; i <= 256; ++i) { const int j = i + (i & -i); if (j <= 256) m->tree[j] += m->tree[i]; }
}
static void am_seed(amodel *m, const uint32_t hist[256]) {
    uint64_t s = 0;
    for (int i = 0; i < 256; ++i) s += hist[i];
    for (int i = 0; i < 256; ++i)
        m->w[i] = 1 + (s <= 8192 ? hist[i] : (uint32_t)(((uint64_t)hist[i] * 8192 + s / 2) / s));
    am_rebuild(m);
}
static void am_bump(amodel *m, unsigned sym) {
    m->w[sym] += 128;
    m->total += 128;
    if (m->total > 65536) {
        for (int i = 0; i < 256; ++i) m->w[i] = (m->w[i] + 1) >> 1;
        am_rebuild(m);
        return;
    }
    for (int i = (int)sym + 1; i <= 256; i += i & -i) m->tree[i] += 128;
}
/* The symbol whose cumulative range holds v, by descending the tree. */
static unsigned am_find(const amodel *m, uint32_t v, uint32_t *below) {
    unsigned pos = 0;
    uint32_t rem = v;
    for (unsigned step = 128; step; step >>= 1)
        if (pos + step <= 256 && m->tree[pos + step] <= rem) { pos += step; rem -= m->tree[pos]; }
    *below = v - rem;
    return pos;
}

typedef struct { const uint8_t *in; size_t len, p; uint32_t range, code; } rdecoder;

static void rd_init(rdecoder *d, const uint8_t *in, size_t len) {
    d->in = in; d->len = len; d->p = 0; d->range = 0xFFFFFFFFu; d->code = 0;
    for (int i = 0; i < 4; ++i) { d->code = (d->code << 8) | (d->p < d->len ? d->in[d->p] : 0); d->p++; }
}
static unsigned rd_sym(rdecoder *d, amodel *m) {
    const uint32_t r = d->range / m->total;
    uint32_t v = d->code / r, below;
    if (v >= m->total) v = m->total - 1;
    const unsigned s = am_find(m, v, &below);
    d->code -= r * below;
    d->range = r * m->w[s];
    while (d->range < (1u << 24)) {
        d->code = (d->code << 8) | (d->p < d->len ? d->in[d->p] : 0);
        d->p++;
        d->range <<= 8;
    }
    am_bump(m, s);
    return s;
}

/* Streams sent adaptively: shortest first (stable), within 65536 symbols in
 * total (ad_select, fdv.h:482) -- derived from the counts, as the encoder chose. */
static unsigned adaptive_set(const size_t sn[NSTR]) {
    int order[NSTR];
    for (int i = 0; i < NSTR; ++i) order[i] = i;
    for (int i = 1; i < NSTR; ++i) {
        const int k = order[i];
        int j = i;
        while (j > 0 && sn[k] < sn[order[j - 1]]) { order[j] = order[j - 1]; --j; }
        order[j] = k;
    }
    unsigned set = 0;
    size_t used = 0;
    for (int i = 0; i < NSTR; ++i) {
        const int k = order[i];
        if (sn[k] == 0 || used + sn[k] > 65536) continue;
        set |= 1u << k;
        used += sn[k];
    }
    return set;
}

/* Payload -> the 16 streams, one after another in `*syms` (malloc'd; the
 * caller frees it), their lengths in sn[], the frame's qp.  nmb bounds the
 * symbol count as the reference does (fdv.h:6513). */
int fdvamd_unpack_pframe(const uint8_t *in, size_t len, size_t nmb, fdvamd_tabcache *tc, int *qp, uint8_t **syms,
                         size_t sn[16]) {
    *syms = NULL;
    if (!in || len < 2) return -1;
    size_t p = 0;
    const int q = in[p++];
    const int emode = in[p++];
    if (q > 51 || emode > 3) return -1;
    const int reuse = emode == 2;
    if (reuse && (!tc || !tc->have)) return -1;
    const size_t maxsyms = nmb * (size_t)(16 + 384 * 6);
    size_t total = 0, base[NSTR];
    for (int k = 0; k < NSTR; ++k) {
        int ok = 1;
        const uint32_t v = rd_count32(in, &p, len, &ok);
        if (!ok || v > maxsyms) return -1;
        sn[k] = v;
        total += v;
        if (total > maxsyms) return -1;
    }
    for (size_t k = 0, off = 0; k < NSTR; ++k) { base[k] = off; off += sn[k]; }
    uint8_t *out = malloc(total ? total : 1);
    rtable *t = malloc(sizeof *t);
    if (!out || !t) { free(out); free(t); return -1; }
    int ok = 1;
    if (!(emode & 1)) {                                         /* one shared table */
        uint16_t f[256];
        if (reuse) memcpy(f, tc->f1, sizeof f);
        else p = read_table(in, p, len, f, &ok);
        if (!ok || p + 4 > len) goto bad;
        uint32_t l1 = 0;
        for (int i = 0; i < 4; ++i) l1 |= (uint32_t)in[p + i] << (8 * i);
        p += 4;
        if (p + l1 > len) goto bad;
        rtable_build(f, t);
        rans_unpack(in + p, l1, total, t, out);
        if (!reuse && tc) { table_keep(tc->f1, f); tc->have = 1; }
    } else {                                                    /* a table per stream */
        if (p + 2 > len) goto bad;
        const unsigned mask = in[p] | (unsigned)in[p + 1] << 8;
        p += 2;
        unsigned am = 0;
        if (emode == 3) {
            if (!tc || !tc->hn) goto bad;
            am = adaptive_set(sn);
        }
        if ((mask & ~am) && (!tc || !tc->have)) goto bad;
        for (int k = 0; k < NSTR; ++k) {
            if (sn[k] == 0 || ((am >> k) & 1)) continue;
            const int reused = (mask >> k) & 1;
            uint16_t f[256];
            if (reused) memcpy(f, tc->fs[k], sizeof f);
            else p = read_table(in, p, len, f, &ok);
            if (!ok) goto bad;
            const uint32_t n = rd_count32(in, &p, len, &ok);
            if (!ok || p + n > len) goto bad;
            rtable_build(f, t);
            rans_unpack(in + p, n, sn[k], t, out + base[k]);
            p += n;
            if (!reused && tc) table_keep(tc->fs[k], f);
        }
        if (am) {
            const uint32_t n = rd_count32(in, &p, len, &ok);
            if (!ok || p + n > len) goto bad;
            rdecoder d;
            rd_init(&d, in + p, n);
            for (int k = 0; k < NSTR; ++k) {
                if (!((am >> k) & 1)) continue;
                amodel m;
                am_seed(&m, tc->hist[k]);
                for (size_t i = 0; i < sn[k]; ++i) out[base[k] + i] = (uint8_t)rd_sym(&d, &m);
            }
            p += n;
        }
        if (tc) tc->have = 1;
    }
    if (tc) {                                                   /* the next frame's adaptive priors */
        for (int k = 0; k < NSTR; ++k) {
            uint32_t c[256] = {0};
            for (size_t i = 0; i < sn[k]; ++i) c[out[base[k] + i]]++;
            history_add(tc->hist[k], c);
        }
        if (tc->hn < (1 << 20)) ++tc->hn;
    }
    free(t);
    *qp = q;
    *syms = out;
    return 0;
bad:
    free(out);
    free(t);
    return -1;
}

/* ---- the MB syntax ------------------------------------------------------ */

typedef struct { int16_t x, y; uint8_t ref, inter; } mvrec;
typedef struct { const uint8_t *b; size_t cur[NSTR], end[NSTR]; int ok; } cursor;

static int c_byte(cursor *c, int k) { return rd_byte(c->b, &c->cur[k], c->end[k], &c->ok); }
static int c_count(cursor *c, int k, int max) { return rd_count(c->b, &c->cur[k], c->end[k], max, &c->ok); }
static int c_level(cursor *c, int k) { return rd_level(c->b, &c->cur[k], c->end[k], &c->ok); }

static const uint8_t SCAN4[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};
static const uint8_t SCAN8[64] = {
     0,  1,  8, 16,  9,  2,  3, 10, 17, 24, 32, 25, 18, 11,  4,  5,
    12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13,  6,  7, 14, 21, 28,
    35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
    58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

/* One 4x4 block's levels (decode_block4, fdv.h:5280): a count, then that many
 * levels in zigzag order; raster order out. */
static void blk4(cursor *c, int cn, int cl, int16_t *lev) {
    memset(lev, 0, 16 * sizeof *lev);
    const int n = c_count(c, cn, 16);
    for (int k = 0; k < n; ++k) lev[SCAN4[k]] = (int16_t)c_level(c, cl);
}
/* One 8x8 block's levels (fdv_coeff8_decode, fdv.h:2205). */
static void blk8(cursor *c, int16_t *lev) {
    memset(lev, 0, 64 * sizeof *lev);
    const int n = c_count(c, S_CN8, 64);
    for (int k = 0; k < n; ++k) lev[SCAN8[k]] = (int16_t)c_level(c, S_CL8);
}
/* One 8x8 region of decode_residual (fdv.h:5426-5462): its flag, then four
 * 4x4 blocks (flag 1) or one 8x8 (flag 2); returns the flag. */
static int region(cursor *c, int chroma, int16_t *lev) {
    const int fl = c_byte(c, S_FL);
    if (fl == 1) for (int b = 0; b < 4; ++b) blk4(c, chroma ? S_CNC : S_CN, chroma ? S_CLC : S_CL, lev + 16 * b);
    else if (fl == 2) blk8(c, lev);
    else if (fl != 0) c->ok = 0;
    return fl;
}

/* The reference decoder's bounds check of a vector (mv_in_bounds, fdv.h:6480):
 * the 6-tap reads must stay inside the plane's 32-sample border. */
static int mv_ok(int pw, int ph, int x, int y, int bw, int bh, int mvx, int mvy) {
    const int ix = x + (mvx >> 2), iy = y + (mvy >> 2);
    return !(ix - 3 < -32 || iy - 3 < -32 || ix + bw + 3 > pw + 32 || iy + bh + 3 > ph + 32);
}

int fdvamd_parse_pframe(const uint8_t *syms, const size_t sn[16], int w, int h, void *codes_out) {
    fdvgpu_mbcode *codes = codes_out;
    const int mw = w / 16, mh = h / 16, cw = w / 2, ch = h / 2;
    cursor c;
    c.b = syms;
    c.ok = 1;
    for (size_t k = 0, off = 0; k < NSTR; ++k) { c.cur[k] = off; c.end[k] = off + sn[k]; off += sn[k]; }
    uint8_t *mmap = calloc((size_t)mw * mh, 1);
    mvrec *mvm = calloc((size_t)mw * mh, sizeof *mvm);
    if (!mmap || !mvm) { free(mmap); free(mvm); return -1; }
    for (int my = 0; my < mh && c.ok; ++my) {
        int pmx = 0, pmy = 0;
        for (int mx = 0; mx < mw && c.ok; ++mx) {
            const int i = my * mw + mx, bx = mx * 16, by = my * 16, cbx = bx / 2, cby = by / 2;
            fdvgpu_mbcode *R = &codes[i];
            memset(R, 0, sizeof *R);
            /* the mode's model: whether the left and above MBs were coded (fdv.h:4172) */
            const int ctx = (mx > 0 && mmap[i - 1] ? 2 : 0) + (my > 0 && mmap[i - mw] ? 1 : 0);
            const int mode = c_byte(&c, ctx);
            mmap[i] = (uint8_t)mode;
            R->mode = (uint8_t)mode;
            if (mode == 2) {                                  /* INTRA (fdv.h:6536-6554) */
                /* any value: an unknown sub-mode predicts DC (fdv.h:2884-2911) */
                R->sub = (uint8_t)c_byte(&c, S_SUB);
                for (int r = 0; r < 4; ++r) R->lflag[r] = (uint8_t)region(&c, 0, R->lev + 64 * r);
                for (int p = 0; p < 2; ++p) R->cflag[p] = (uint8_t)region(&c, 1, R->clev + 64 * p);
                R->ckind = 2;
                mvm[i] = (mvrec){0, 0, 0, 0};
                continue;                                     /* the predictor is unchanged */
            }
            if (mode == 3) {                                  /* INTER8 (fdv.h:6557-6590) */
                int qx = pmx, qy = pmy;
                for (int q = 0; q < 4 && c.ok; ++q) {
                    const int vx = qx + c_level(&c, S_MVX), vy = qy + c_level(&c, S_MVY);
                    qx = vx; qy = vy;
                    R->mv[q][0] = (int16_t)vx; R->mv[q][1] = (int16_t)vy;
                    if (!c.ok || !mv_ok(w, h, bx + (q & 1) * 8, by + (q >> 1) * 8, 8, 8, vx, vy)) { c.ok = 0; break; }
                    R->lflag[q] = (uint8_t)region(&c, 0, R->lev + 64 * q);
                }
                if (!c.ok) break;
                pmx = R->mv[3][0]; pmy = R->mv[3][1];
                mvm[i] = (mvrec){R->mv[3][0], R->mv[3][1], 0, 1};
                for (int p = 0; p < 2 && c.ok; ++p)
                    for (int q = 0; q < 4 && c.ok; ++q) {
                        if (!mv_ok(cw, ch, cbx + (q & 1) * 4, cby + (q >> 1) * 4, 4, 4, R->mv[q][0] / 2, R->mv[q][1] / 2)) {
                            c.ok = 0; break;
                        }
                        blk4(&c, S_CNC, S_CLC, R->clev + 64 * p + 16 * q);
                    }
                R->ckind = 3;
                continue;
            }
            if (mode != 0 && mode != 1) { c.ok = 0; break; }
            int vx, vy, ref = 0;
            if (mode == 0) {                                  /* SKIP: a merge candidate (fdv.h:6596-6616) */
                mvrec cand[3];
                int n = 0;
                cand[n++] = (mvrec){(int16_t)pmx, (int16_t)pmy, 0, 1};
                for (int k = 0; k < 2 && n < 3; ++k) {
                    const int x = mx + k, y = my - 1;
                    if (y < 0 || x >= mw) continue;
                    const mvrec *m = &mvm[y * mw + x];
                    if (!m->inter) continue;
                    int dup = 0;
                    for (int j = 0; j < n; ++j) dup |= cand[j].x == m->x && cand[j].y == m->y && cand[j].ref == m->ref;
                    if (!dup) cand[n++] = *m;
                }
                const int mi = n > 1 ? c_count(&c, S_MERGE, n - 1) : 0;
                if (!c.ok) break;
                vx = cand[mi].x; vy = cand[mi].y; ref = cand[mi].ref;
                R->merge_idx = (uint8_t)mi; R->ncand = (uint8_t)n;
            } else {                                          /* INTER16 */
                ref = c_byte(&c, S_REF);
                if (ref > 1) { c.ok = 0; break; }
                vx = pmx + c_level(&c, S_MVX);
                vy = pmy + c_level(&c, S_MVY);
            }
            pmx = vx; pmy = vy;
            mvm[i] = (mvrec){(int16_t)vx, (int16_t)vy, (uint8_t)ref, 1};
            R->ref = (uint8_t)ref;
            R->mv[0][0] = (int16_t)vx; R->mv[0][1] = (int16_t)vy;
            if (!c.ok || !mv_ok(w, h, bx, by, 16, 16, vx, vy) || !mv_ok(cw, ch, cbx, cby, 8, 8, vx / 2, vy / 2)) {
                c.ok = 0; break;
            }
            if (mode == 0) { R->ckind = 0; continue; }
            for (int r = 0; r < 4; ++r) R->lflag[r] = (uint8_t)region(&c, 0, R->lev + 64 * r);
            /* chroma: an any-flag per plane, then four 4x4 blocks (fdv.h:6429-6456) */
            for (int p = 0; p < 2 && c.ok; ++p) {
                const int any = c_byte(&c, S_FL);
                if (any > 1) { c.ok = 0; break; }
                R->cflag[p] = (uint8_t)any;
                if (any) for (int b = 0; b < 4; ++b) blk4(&c, S_CNC, S_CLC, R->clev + 64 * p + 16 * b);
            }
            R->ckind = 1;
        }
    }
    free(mmap);
    free(mvm);
    return c.ok ? 0 : -1;
}

/* ---- key frames (fdv_image_decode, fdv.h:5040-5152; iframe_decode 5541) ----
 *
 * The image container undone into its twelve streams, then each 64x64 unit's
 * quadtree walked in coding order: a split flag per node (none where the
 * plane's edge implies the split), a mode per leaf (the 4x4 leaves' under the
 * context of their left and above cells), the leaf predicted from the
 * reconstruction so far and its residual added (decode_residual, 5426).  The
 * plane is deblocked at the end.  Malformed input the reference rejects is
 * rejected (-1). */

#include <pthread.h>

void fdvg_tables8(int32_t c8[64], int32_t qs8[52]);
static int32_t g_c8[64], g_qs8[52];
static pthread_once_t g_c8_once = PTHREAD_ONCE_INIT;
static void c8_load(void) { fdvg_tables8(g_c8, g_qs8); }

enum { I_SPLIT = 0, I_MODEB = 4, I_MODE0, I_MODE1, I_CNT, I_LVL, I_CNT8, I_LVL8, I_FLAG, I_N };

static const int DQ4[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16}, {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};
static int cls4(int i) { const int r = (i >> 2) & 1, c = i & 1; return r == c ? r : 2; }
static int clip255(int v) { return v < 0 ? 0 : v > 255 ? 255 : v; }

/* dequantisation and the inverse 4x4 transform (fdv.h:2246-2292, 2546) */
static void inv4(const int16_t lev[16], int qp, int16_t out[16]) {
    int c[16], t[16];
    for (int i = 0; i < 16; ++i) c[i] = (int32_t)(lev[i] * DQ4[qp % 6][cls4(i)]) * (int32_t)(1 << (qp / 6));
    for (int i = 0; i < 4; ++i) {
        const int a = c[4 * i], b = c[4 * i + 1], d2 = c[4 * i + 2], d = c[4 * i + 3];
        const int e0 = a + d2, e1 = a - d2, e2 = (b >> 1) - d, e3 = b + (d >> 1);
        t[4 * i] = e0 + e3; t[4 * i + 1] = e1 + e2; t[4 * i + 2] = e1 - e2; t[4 * i + 3] = e0 - e3;
    }
    for (int j = 0; j < 4; ++j) {
        const int a = t[j], b = t[4 + j], d2 = t[8 + j], d = t[12 + j];
        const int e0 = a + d2, e1 = a - d2, e2 = (b >> 1) - d, e3 = b + (d >> 1);
        out[j] = (int16_t)((e0 + e3 + 32) >> 6); out[4 + j] = (int16_t)((e1 + e2 + 32) >> 6);
        out[8 + j] = (int16_t)((e1 - e2 + 32) >> 6); out[12 + j] = (int16_t)((e0 - e3 + 32) >> 6);
    }
}
/* dequantisation (saturated) and the inverse 8x8 transform (fdv.h:2639-2685) */
static void inv8(const int16_t lev[64], int qp, int16_t out[64]) {
    pthread_once(&g_c8_once, c8_load);
    int32_t y[64];
    int64_t b[64];
    for (int i = 0; i < 64; ++i) {
        const int64_t v = (int64_t)lev[i] * g_qs8[qp];
        y[i] = v < INT32_MIN ? INT32_MIN : v > INT32_MAX ? INT32_MAX : (int32_t)v;
    }
    for (int r = 0; r < 8; ++r)
        for (int v = 0; v < 8; ++v) {
            int64_t s = 0;
            for (int u = 0; u < 8; ++u) s += (int64_t)g_c8[u * 8 + r] * y[u * 8 + v];
            b[r * 8 + v] = s;
        }
    for (int r = 0; r < 8; ++r)
        for (int x = 0; x < 8; ++x) {
            int64_t s = 0;
            for (int v = 0; v < 8; ++v) s += b[r * 8 + v] * g_c8[v * 8 + x];
            out[r * 8 + x] = (int16_t)((s + ((int64_t)1 << 35)) >> 36);
        }
}

/* Z-order rank of a unit's 4x4 cell and the availability of a sample while
 * the 4x4 block at (bx, by) is reconstructed (fdv.h:3659, 3677). */
static int zrank(int lx, int ly) {
    int z = 0;
    for (int b = 8; b; b >>= 1) z = (z << 2) | ((ly & b) ? 2 : 0) | ((lx & b) ? 1 : 0);
    return z;
}
static int sample_done(int x, int y, int bx, int by) {
    const int ux = bx & ~63, uy = by & ~63;
    if (y < uy || x < ux) return 1;
    if (x >= ux + 64) return 0;
    return zrank((x - ux) >> 2, (y - uy) >> 2) < zrank((bx - ux) >> 2, (by - uy) >> 2);
}
/* The nine 4x4 predictors of H.264 (8.3.1.2) over the edge E (E[3 - i] left
 * sample i, E[4] the corner, E[5 + i] top sample i, top-right included). */
static int pred4_at(const int *E, int m, int x, int y, int dc) {
    int kind = 2, i = 0;
    switch (m) {
    case 1: kind = 0; i = 5 + x; break;
    case 2: kind = 0; i = 3 - y; break;
    case 3: if (x == 3 && y == 3) { kind = 3; i = 11; } else i = 5 + x + y; break;
    case 4: i = 3 + x - y; break;
    case 5: {
        const int z = 2 * x - y, k = x - (y >> 1);
        if (z >= 0) { kind = (z & 1) ? 2 : 1; i = (z & 1) ? 3 + k : 4 + k; }
        else i = z == -1 ? 3 : 4 - y;
        break;
    }
    case 6: {
        const int z = 2 * y - x, k = y - (x >> 1);
        if (z >= 0) { kind = (z & 1) ? 2 : 1; i = 3 - k; }
        else i = z == -1 ? 3 : 2 + x;
        break;
    }
    case 7: kind = (y & 1) ? 2 : 1; i = 5 + x + (y >> 1); break;
    case 8: {
        const int z = x + 2 * y, k = y + (x >> 1);
        if (z < 5) { kind = (z & 1) ? 2 : 1; i = (z & 1) ? 1 - k : 2 - k; }
        else if (z == 5) kind = 4;
        else kind = 0;
        break;
    }
    default: return dc;
    }
    const int a = E[i], b = E[i + 1], c = E[i + 2];
    return kind == 0 ? a : kind == 1 ? (a + b + 1) >> 1 : kind == 2 ? (a + 2 * b + c + 2) >> 2 :
           kind == 3 ? (a + 3 * b + 2) >> 2 : (3 * a + b + 2) >> 2;
}

typedef struct {
    const uint8_t *b;
    size_t cur[I_N], end[I_N];
    int ok, w, h, qp, stride;
    uint8_t *dst, *cmap;
} idec;

static int i_count(idec *d, int k, int max) { return rd_count(d->b, &d->cur[k], d->end[k], max, &d->ok); }
static void i_block4(idec *d, const uint8_t *pred, int ps, uint8_t *o, int os) {
    int16_t lev[16] = {0}, rr[16];
    const int n = i_count(d, I_CNT, 16);
    for (int k = 0; k < n; ++k) lev[SCAN4[k]] = (int16_t)rd_level(d->b, &d->cur[I_LVL], d->end[I_LVL], &d->ok);
    inv4(lev, d->qp, rr);
    for (int i = 0; i < 16; ++i) o[(i >> 2) * os + (i & 3)] = (uint8_t)clip255(pred[(i >> 2) * ps + (i & 3)] + rr[i]);
}

static void i_node(idec *d, int bx, int by, int n, int depth) {
    if (!d->ok) return;
    const int w = d->w, h = d->h, st = d->stride, hn = n / 2;
    uint8_t *D = d->dst;
    if (bx + n > w || by + n > h) {
        for (int k = 0; k < 4; ++k) {
            const int cx = bx + (k & 1) * hn, cy = by + (k >> 1) * hn;
            if (cx < w && cy < h) i_node(d, cx, cy, hn, depth + 1);
        }
        return;
    }
    if (n == 4) {
        const int cx = bx / 4, cy = by / 4, cw = w / 4;
        const int lm = cx ? d->cmap[(size_t)cy * cw + cx - 1] : 0, am = cy ? d->cmap[(size_t)(cy - 1) * cw + cx] : 0;
        const int mode = i_count(d, lm == am ? I_MODE0 : I_MODE1, 8);
        if (!d->ok) return;
        d->cmap[(size_t)cy * cw + cx] = (uint8_t)mode;
        const int ht = by > 0, hl = bx > 0, htl = ht && hl;
        const int htr = ht && bx + 4 < w && sample_done(bx + 4, by - 1, bx, by);
        int E[16] = {0};
        for (int i = 0; i < 4; ++i) E[3 - i] = hl ? D[(size_t)(by + i) * st + bx - 1] : 0;
        E[4] = htl ? D[(size_t)(by - 1) * st + bx - 1] : 128;
        for (int j = 0; j < 8; ++j) E[5 + j] = ht ? D[(size_t)(by - 1) * st + bx + (j < 4 || htr ? j : 3)] : 0;
        int s = 0;
        if (ht) for (int j = 0; j < 4; ++j) s += E[5 + j];
        if (hl) for (int i = 0; i < 4; ++i) s += E[3 - i];
        const int dc = ht && hl ? (s + 4) >> 3 : ht || hl ? (s + 2) >> 2 : 128;
        uint8_t pred[16];
        for (int y = 0; y < 4; ++y)
            for (int x = 0; x < 4; ++x) pred[y * 4 + x] = (uint8_t)pred4_at(E, mode, x, y, dc);
        i_block4(d, pred, 4, D + (size_t)by * st + bx, st);
        return;
    }
    const int split = i_count(d, I_SPLIT + depth, 1);
    if (!d->ok) return;
    if (split) {
        for (int k = 0; k < 4; ++k) i_node(d, bx + (k & 1) * hn, by + (k >> 1) * hn, hn, depth + 1);
        return;
    }
    const int mode = i_count(d, I_MODEB, 3);
    if (!d->ok) return;
    /* the n x n prediction (gather_nb 5201, fdv_intra_nxn_inner 2879): an
     * unavailable mode falls back to DC */
    const int ht = by > 0, hl = bx > 0;
    uint8_t top[64], left[64], pred[64 * 64];
    for (int k = 0; k < n; ++k) {
        top[k] = ht ? D[(size_t)(by - 1) * st + bx + k] : 0;
        left[k] = hl ? D[(size_t)(by + k) * st + bx - 1] : 0;
    }
    const int tl = ht && hl ? D[(size_t)(by - 1) * st + bx - 1] : 128;
    if (mode == 3 && ht && hl && n <= 32) {
        int H = 0, V = 0;
        for (int k = 1; k <= hn; ++k) {
            const int j = hn - 1 - k;
            H += k * (top[hn - 1 + k] - (j < 0 ? tl : top[j]));
            V += k * (left[hn - 1 + k] - (j < 0 ? tl : left[j]));
        }
        const int a = 16 * (left[n - 1] + top[n - 1]);
        const int b = n == 32 ? (11 * H + 512) >> 10 : n == 16 ? (5 * H + 32) >> 6 : (17 * H + 16) >> 5;
        const int c = n == 32 ? (11 * V + 512) >> 10 : n == 16 ? (5 * V + 32) >> 6 : (17 * V + 16) >> 5;
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) pred[y * n + x] = (uint8_t)clip255((a + c * (y - hn + 1) + 16 + b * (x - hn + 1)) >> 5);
    } else if (mode == 1 && ht) {
        for (int y = 0; y < n; ++y) memcpy(pred + y * n, top, (size_t)n);
    } else if (mode == 2 && hl) {
        for (int y = 0; y < n; ++y) memset(pred + y * n, left[y], (size_t)n);
    } else {
        int s = 0, dc = 128;
        if (ht) for (int k = 0; k < n; ++k) s += top[k];
        if (hl) for (int k = 0; k < n; ++k) s += left[k];
        if (ht && hl) dc = (s + n) / (2 * n);
        else if (ht || hl) dc = (s + n / 2) / n;
        memset(pred, dc, (size_t)n * n);
    }
    for (int ry = 0; ry < n && d->ok; ry += 8)
        for (int rx = 0; rx < n && d->ok; rx += 8) {
            const int fl = rd_byte(d->b, &d->cur[I_FLAG], d->end[I_FLAG], &d->ok);
            const uint8_t *pb = pred + ry * n + rx;
            uint8_t *o = D + (size_t)(by + ry) * st + bx + rx;
            if (fl == 0) {
                for (int y = 0; y < 8; ++y) memcpy(o + (size_t)y * st, pb + y * n, 8);
            } else if (fl == 1) {
                for (int q = 0; q < 4; ++q)
                    i_block4(d, pb + (q >> 1) * 4 * n + (q & 1) * 4, n, o + (size_t)((q >> 1) * 4) * st + (q & 1) * 4, st);
            } else if (fl == 2) {
                int16_t lev[64] = {0}, rr[64];
                const int cnt = i_count(d, I_CNT8, 64);
                for (int k = 0; k < cnt; ++k)
                    lev[SCAN8[k]] = (int16_t)rd_level(d->b, &d->cur[I_LVL8], d->end[I_LVL8], &d->ok);
                inv8(lev, d->qp, rr);
                for (int i = 0; i < 64; ++i)
                    o[(size_t)(i >> 3) * st + (i & 7)] = (uint8_t)clip255(pb[(i >> 3) * n + (i & 7)] + rr[i]);
            } else {
                d->ok = 0;
            }
        }
    for (int i = 0; i < n / 4; ++i)
        memset(d->cmap + (size_t)(by / 4 + i) * (w / 4) + bx / 4, mode == 1 ? 1 : mode == 2 ? 2 : 0, (size_t)n / 4);
}

/* the in-loop filter of a plane, vertical edges then horizontal (fdv.h:3479-3627) */
static void i_edge(uint8_t *q, ptrdiff_t s, int al, int be, int tc) {
    const int p1 = q[-2 * s], p0 = q[-s], q0 = q[0], q1 = q[s];
    if (abs(p0 - q0) < al && abs(p1 - p0) < be && abs(q1 - q0) < be) {
        int dd = ((q0 - p0) * 4 + (p1 - q1) + 4) >> 3;
        dd = dd < -tc ? -tc : dd > tc ? tc : dd;
        q[-s] = (uint8_t)clip255(p0 + dd);
        q[0] = (uint8_t)clip255(q0 - dd);
    }
}
static void i_deblock(uint8_t *p, int w, int h, int st, int qp) {
    const int al = 6 + qp, be = 2 + qp / 4, tc = 1 + qp / 10;
    for (int y = 0; y < h; ++y)
        for (int x = 8; x < w; x += 8) i_edge(p + (size_t)y * st + x, 1, al, be, tc);
    for (int y = 8; y < h; y += 8)
        for (int x = 0; x < w; ++x) i_edge(p + (size_t)y * st + x, st, al, be, tc);
}

int fdvamd_image_decode(const uint8_t *in, size_t len, uint8_t *dst, int stride, int *w_out, int *h_out) {
    if (!in || !dst || len < 6) return -1;
    size_t p = 0;
    const int w = in[0] | in[1] << 8, h = in[2] | in[3] << 8, qp = in[4], mode = in[5];
    p = 6;
    if (w <= 0 || h <= 0 || (w & 3) || (h & 3) || w > 8192 || h > 8192 || qp > 51 || mode > 2) return -1;
    if (stride < w) return -1;
    /* what the geometry lets each stream hold (fdv.h:5058-5074) */
    const size_t units = (size_t)((w + 63) / 64) * ((h + 63) / 64), b4 = (size_t)(w / 4) * (h / 4);
    const size_t cnt_cap = units * (256 + 64 + 8), lvl_cap = units * (64 * 64 * 5 + 64);
    size_t cap[I_N] = {units, 4 * units, 16 * units, 64 * units, b4, b4, b4, cnt_cap, lvl_cap, cnt_cap, lvl_cap, cnt_cap};
    size_t sn[I_N], off[I_N], total = 0;
    int ok = 1;
    for (int k = 0; k < I_N; ++k) {
        const uint32_t v = rd_count32(in, &p, len, &ok);
        if (!ok || v > cap[k]) return -1;
        off[k] = total; sn[k] = v; total += v;
    }
    uint8_t *syms = malloc(total ? total : 1);
    rtable *t = malloc(sizeof *t);
    int rc = -1;
    idec d;
    memset(&d, 0, sizeof d);
    if (!syms || !t) goto out;
    if (mode == 2) {
        const uint32_t la = rd_count32(in, &p, len, &ok);
        if (!ok || p + la > len) goto out;
        static const uint32_t flat[256];
        rdecoder rd;
        rd_init(&rd, in + p, la);
        for (int k = 0; k < I_N; ++k) {
            amodel m;
            am_seed(&m, flat);
            for (size_t i = 0; i < sn[k]; ++i) syms[off[k] + i] = (uint8_t)rd_sym(&rd, &m);
        }
    } else {
        for (int k = 0; k < (mode == 0 ? 1 : I_N); ++k) {
            const size_t n = mode == 0 ? total : sn[k];
            if (mode == 1 && n == 0) continue;
            uint16_t f[256];
            p = read_table(in, p, len, f, &ok);
            if (!ok) goto out;
            const uint32_t l = rd_count32(in, &p, len, &ok);
            if (!ok || p + l > len) goto out;
            rtable_build(f, t);
            rans_unpack(in + p, l, n, t, syms + (mode == 0 ? 0 : off[k]));
            p += l;
        }
    }
    d.b = syms; d.ok = 1; d.w = w; d.h = h; d.qp = qp; d.stride = stride; d.dst = dst;
    for (int k = 0; k < I_N; ++k) { d.cur[k] = off[k]; d.end[k] = off[k] + sn[k]; }
    if (!(d.cmap = calloc(b4 ? b4 : 1, 1))) goto out;
    for (int by = 0; by < h && d.ok; by += 64)
        for (int bx = 0; bx < w && d.ok; bx += 64) i_node(&d, bx, by, 64, 0);
    if (!d.ok) goto out;
    i_deblock(dst, w, h, stride, qp);
    if (w_out) *w_out = w;
    if (h_out) *h_out = h;
    rc = 0;
out:
    free(syms); free(t); free(d.cmap);
    return rc;
}

int fdvamd_iframe_decode(const uint8_t *in, size_t len, uint8_t *y, uint8_t *u, uint8_t *v, int w, int h) {
    if (!in || !y || !u || !v) return -1;
    uint8_t *dst[3] = {y, u, v};
    const int pw[3] = {w, w / 2, w / 2}, ph[3] = {h, h / 2, h / 2};
    size_t p = 0;
    for (int i = 0; i < 3; ++i) {
        if (p + 4 > len) return -1;
        const uint32_t blob = (uint32_t)in[p] | (uint32_t)in[p + 1] << 8 | (uint32_t)in[p + 2] << 16 |
                              (uint32_t)in[p + 3] << 24;
        p += 4;
        if (p + blob > len || blob < 4) return -1;
        /* the blob's own dimensions must be the plane's (fdv.h:5556-5561) */
        if ((in[p] | in[p + 1] << 8) != pw[i] || (in[p + 2] | in[p + 3] << 8) != ph[i]) return -1;
        if (fdvamd_image_decode(in + p, blob, dst[i], pw[i], NULL, NULL) != 0) return -1;
        p += blob;
    }
    return 0;
}
