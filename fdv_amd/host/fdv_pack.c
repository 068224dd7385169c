/* fdv_pack.c -- host entropy stage of a P-frame: the 16 symbol streams the
 * device produces become the reference's payload bytes (yhbsh/fdv include/fdv.h
 * 6172-6421), with the table cache and adaptive history carried between frames.
 *
 * Four codings are sized and the smallest kept (ties to the lower mode):
 *   0  one static rANS table over all streams, sent;
 *   1  a table per stream, each either sent or reused from the cache;
 *   2  the cached shared table;
 *   3  as 1, but the shortest streams (up to 65536 symbols) go through one
 *      adaptive range-coded blob primed from the stream history.
 * The coders are the reference's: 12-bit byte rANS with a 2^23 state floor
 * (fdv.h:1760-1801), and a 32-bit carry-propagating range coder over a
 * frequency model that adds 128 per symbol and halves past 65536
 * (fdv.h:2003-2118).
 */
#include "fdv_amd.h"

#include <stdlib.h>
#include <string.h>

#define NSTR 16
#define SCALE_BITS 12
#define PROB_SCALE (1u << SCALE_BITS)
#define STATE_LO (1u << 23)

static int varint_len(uint32_t u) { int n = 1; for (; u >= 0x80; u >>= 7) ++n; return n; }
static size_t varint_put(uint8_t *b, size_t p, uint32_t u) {
    for (; u >= 0x80; u >>= 7) b[p++] = (uint8_t)(u | 0x80);
    b[p++] = (uint8_t)u;
    return p;
}

/* ---- static rANS ---------------------------------------------------------- */

/* Frequencies summing to 4096: proportional, used symbols at least 1, the
 * largest (first on ties) taking the rounding slack (fdv.h:1804-1840). */
static void freqs_from_counts(const uint32_t cnt[256], uint16_t f[256]) {
    uint64_t total = 0;
    for (int s = 0; s < 256; ++s) total += cnt[s];
    memset(f, 0, 256 * sizeof f[0]);
    if (total == 0) { f[0] = PROB_SCALE; return; }
    int big = 0;
    uint32_t sum = 0;
    for (int s = 0; s < 256; ++s) {
        if (!cnt[s]) continue;
        uint32_t v = (uint32_t)(((uint64_t)cnt[s] << SCALE_BITS) / total);
        f[s] = (uint16_t)(v ? v : 1);
        sum += f[s];
        if (f[s] > f[big]) big = s;
    }
    f[big] = (uint16_t)((int)f[big] + (int)PROB_SCALE - (int)sum);
}

/* Encode back to front into the tail of `buf`, then move the result to the
 * front; the final state leads, little-endian (fdv.h:1936-1948).  The state's
 * division by the frequency is a multiplication by M = ceil(2^64 / f) (the
 * high 64 bits of the 128-bit product): with x < 2^31 and f <= 2^12 the error
 * of x M / 2^64 against x / f is below 2^-33, less than the 1/f that separates
 * x / f from the next integer, so the quotient is exact. */
static size_t rans_pack(const uint8_t *sym, size_t n, const uint16_t f[256], uint8_t *buf, size_t cap) {
    uint32_t start[256], acc = 0;
    uint64_t rcp[256];
    for (int s = 0; s < 256; ++s) {
        start[s] = acc;
        acc += f[s];
        rcp[s] = f[s] > 1 ? ~0ull / f[s] + 1 : 0;
    }
    uint8_t *p = buf + cap;
    uint32_t x = STATE_LO;
    for (size_t i = n; i > 0; --i) {
        const unsigned sm = sym[i - 1];
        const uint32_t fs = f[sm];
        const uint64_t bound = ((uint64_t)(STATE_LO >> SCALE_BITS) << 8) * fs;
        while ((uint64_t)x >= bound) { *--p = (uint8_t)(x & 0xff); x >>= 8; }
        const uint32_t q = fs == 1 ? x : (uint32_t)(((unsigned __int128)x * rcp[sm]) >> 64);
        x = (q << SCALE_BITS) + (x - q * fs) + start[sm];
    }
    p -= 4;
    for (int k = 0; k < 4; ++k) p[k] = (uint8_t)(x >> (8 * k));
    size_t len = (size_t)(buf + cap - p);
    memmove(buf, p, len);
    return len;
}

static size_t table_bytes(const uint16_t f[256]) {
    size_t n = 2;
    for (int s = 0; s < 256; ++s) if (f[s]) n += 1 + (size_t)varint_len(f[s]);
    return n;
}
static size_t table_put(uint8_t *b, size_t p, const uint16_t f[256]) {
    unsigned used = 0;
    for (int s = 0; s < 256; ++s) used += f[s] != 0;
    b[p++] = (uint8_t)(used & 0xff);
    b[p++] = (uint8_t)(used >> 8);
    for (int s = 0; s < 256; ++s)
        if (f[s]) { b[p++] = (uint8_t)s; p = varint_put(b, p, f[s]); }
    return p;
}
static int table_codes(const uint16_t f[256], const uint32_t cnt[256]) {
    for (int s = 0; s < 256; ++s) if (cnt[s] && !f[s]) return 0;
    return 1;
}
/* The cached form: every unused symbol one slot, paid by the largest bin,
 * unless that bin cannot spare them (fdv.h:5667-5677). */
static void table_keep(uint16_t dst[256], const uint16_t src[256]) {
    int missing = 0, big = 0;
    for (int s = 0; s < 256; ++s) { if (!src[s]) ++missing; if (src[s] > src[big]) big = s; }
    if (src[big] <= (uint16_t)missing) { memcpy(dst, src, 256 * sizeof dst[0]); return; }
    for (int s = 0; s < 256; ++s) dst[s] = src[s] ? src[s] : 1;
    dst[big] = (uint16_t)(dst[big] - missing);
}
static void history_add(uint32_t h[256], const uint32_t c[256]) {   /* fdv.h:5661 */
    uint64_t tot = 0;
    for (int s = 0; s < 256; ++s) { h[s] += c[s]; tot += h[s]; }
    if (tot > (1u << 21)) for (int s = 0; s < 256; ++s) h[s] >>= 1;
}

/* ---- adaptive range coder --------------------------------------------------- */

typedef struct { uint32_t w[256]; uint32_t tree[257]; uint32_t total; } amodel;

/* Fenwick tree over the weights: cumulative sums in 8 steps. */
static void am_rebuild(amodel *m) {
    memset(m->tree, 0, sizeof m->tree);
    m->total = 0;
    for (int i = 0; i < 256; ++i) { m->tree[i + 1] = m->w[i]; m->total += m->w[i]; }
    for (int i = 1; i <= 256; ++i) { int j = i + (i & -i); if (j <= 256) m->tree[j] += m->tree[i]; }
}
static void am_seed(amodel *m, const uint32_t hist[256]) {        /* fdv.h:2024 */
    uint64_t s = 0;
    for (int i = 0; i < 256; ++i) s += hist[i];
    for (int i = 0; i < 256; ++i)
        m->w[i] = 1 + (s <= 8192 ? hist[i] : (uint32_t)(((uint64_t)hist[i] * 8192 + s / 2) / s));
    am_rebuild(m);
}
static uint32_t am_below(const amodel *m, unsigned sym) {
    uint32_t r = 0;
    for (int i = (int)sym; i > 0; i -= i & -i) r += m->tree[i];
    return r;
}
static void am_bump(amodel *m, unsigned sym) {                     /* fdv.h:2049 */
    m->w[sym] += 128;
    m->total += 128;
    if (m->total > 65536) {
        for (int i = 0; i < 256; ++i) m->w[i] = (m->w[i] + 1) >> 1;
        am_rebuild(m);
        return;
    }
    for (int i = (int)sym + 1; i <= 256; i += i & -i) m->tree[i] += 128;
}

typedef struct {
    uint8_t *out; size_t cap, n;
    uint64_t low; uint32_t range;
    uint8_t held; uint64_t run;      /* the byte not yet written, and 0xFF run length */
    int overflow;
} rcoder;

static void rc_emit(rcoder *c, uint8_t b) {
    if (c->n < c->cap) c->out[c->n++] = b; else c->overflow = 1;
}
static void rc_shift(rcoder *c) {                                  /* fdv.h:2088 */
    const uint32_t lo = (uint32_t)c->low;
    const unsigned carry = (unsigned)(c->low >> 32);
    if (lo < 0xFF000000u || carry) {
        uint8_t b = c->held;
        for (; c->run; --c->run) { rc_emit(c, (uint8_t)(b + carry)); b = 0xFF; }
        c->held = (uint8_t)(lo >> 24);
    }
    c->run++;
    c->low = (uint64_t)(uint32_t)(lo << 8);
}
static void rc_code(rcoder *c, amodel *m, unsigned sym) {
    const uint32_t r = c->range / m->total;
    c->low += (uint64_t)r * am_below(m, sym);
    c->range = r * m->w[sym];
    while (c->range < (1u << 24)) { rc_shift(c); c->range <<= 8; }
    am_bump(m, sym);
}
static size_t rc_finish(rcoder *c) {
    for (int i = 0; i < 5; ++i) rc_shift(c);
    if (c->overflow || c->n < 1) return 0;
    memmove(c->out, c->out + 1, c->n - 1);      /* the leading zero byte is implicit */
    return c->n - 1;
}

/* Streams coded adaptively: shortest first (stable), within 65536 symbols
 * in total (ad_select, fdv.h:482). */
static unsigned adaptive_set(const size_t sn[NSTR], size_t *nsyms) {
    int order[NSTR];
    for (int i = 0; i < NSTR; ++i) order[i] = i;
    for (int i = 1; i < NSTR; ++i) {
        int k = order[i], j = i;
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
    *nsyms = used;
    return set;
}

/* ---- the payload ------------------------------------------------------------ */

typedef struct {
    uint32_t cnt[NSTR][256], call[256];
    uint16_t fk[NSTR][256], f1[256];
    uint8_t *all, *shared, *shared_r, *ada;
    uint8_t *own[NSTR], *reuse[NSTR];
    size_t l_shared, l_shared_r, l_ada, l_own[NSTR], l_reuse[NSTR];
} packwork;

static void packwork_free(packwork *w) {
    free(w->all); free(w->shared); free(w->shared_r); free(w->ada);
    for (int k = 0; k < NSTR; ++k) { free(w->own[k]); free(w->reuse[k]); }
    free(w);
}

size_t fdvamd_pack_pframe(int qp, const uint8_t *const sv[16], const size_t sn[16],
                          fdvamd_tabcache *tc, uint8_t *out, size_t cap) {
    packwork *w = calloc(1, sizeof *w);
    if (!w) return 0;
    size_t total = 0, hdr = 0;
    for (int k = 0; k < NSTR; ++k) { total += sn[k]; hdr += (size_t)varint_len((uint32_t)sn[k]); }
    const size_t bcap = total * 2 + 1024;
    w->all = malloc(total ? total : 1);
    w->shared = malloc(bcap);
    if (!w->all || !w->shared) { packwork_free(w); return 0; }
    size_t o = 0;
    for (int k = 0; k < NSTR; ++k) {
        for (size_t i = 0; i < sn[k]; ++i) w->cnt[k][sv[k][i]]++;
        if (sn[k]) memcpy(w->all + o, sv[k], sn[k]);
        o += sn[k];
    }
    for (int k = 0; k < NSTR; ++k) for (int s = 0; s < 256; ++s) w->call[s] += w->cnt[k][s];

    /* candidate 0: one fresh shared table */
    freqs_from_counts(w->call, w->f1);
    w->l_shared = rans_pack(w->all, total, w->f1, w->shared, bcap);
    const size_t size0 = hdr + table_bytes(w->f1) + 4 + w->l_shared;
    for (int k = 0; k < NSTR; ++k) {
        freqs_from_counts(w->cnt[k], w->fk[k]);
        const size_t c = sn[k] * 2 + 1024;
        if (!(w->own[k] = malloc(c))) { packwork_free(w); return 0; }
        w->l_own[k] = rans_pack(sv[k], sn[k], w->fk[k], w->own[k], c);
    }
    /* cached tables, where they can code this frame */
    size_t size2 = (size_t)-1;
    int per_stream_reuse = 0;
    if (tc && tc->have) {
        if (table_codes(tc->f1, w->call) && (w->shared_r = malloc(bcap)) != NULL) {
            w->l_shared_r = rans_pack(w->all, total, tc->f1, w->shared_r, bcap);
            if (w->l_shared_r) size2 = hdr + 4 + w->l_shared_r;
        }
        int ok = 1;
        for (int k = 0; k < NSTR; ++k) ok &= table_codes(tc->fs[k], w->cnt[k]);
        for (int k = 0; k < NSTR && ok; ++k) {
            const size_t c = sn[k] * 2 + 1024;
            if (!(w->reuse[k] = malloc(c))) { ok = 0; break; }
            w->l_reuse[k] = rans_pack(sv[k], sn[k], tc->fs[k], w->reuse[k], c);
        }
        per_stream_reuse = ok;
    }
    /* candidate 1: per stream, fresh or reused, whichever is smaller */
    size_t size1 = hdr + 2;
    unsigned rmask = 0;
    size_t cost[NSTR];
    for (int k = 0; k < NSTR; ++k) {
        cost[k] = 0;
        if (!sn[k]) continue;
        const size_t fresh = table_bytes(w->fk[k]) + (size_t)varint_len((uint32_t)w->l_own[k]) + w->l_own[k];
        const size_t again = per_stream_reuse ? (size_t)varint_len((uint32_t)w->l_reuse[k]) + w->l_reuse[k]
                                              : (size_t)-1;
        if (again < fresh) { rmask |= 1u << k; cost[k] = again; } else cost[k] = fresh;
        size1 += cost[k];
    }
    /* candidate 3: short streams adaptive against the history */
    size_t size3 = (size_t)-1;
    unsigned amask = 0;
    if (tc && tc->hn) {
        size_t nada = 0;
        amask = adaptive_set(sn, &nada);
        if (amask && (w->ada = malloc(nada * 3 + 64)) != NULL) {
            rcoder rc = {w->ada, nada * 3 + 64, 0, 0, 0xFFFFFFFFu, 0, 1, 0};
            amodel m;
            for (int k = 0; k < NSTR; ++k) {
                if (!((amask >> k) & 1)) continue;
                am_seed(&m, tc->hist[k]);
                for (size_t i = 0; i < sn[k]; ++i) rc_code(&rc, &m, sv[k][i]);
            }
            w->l_ada = rc_finish(&rc);
            if (w->l_ada) {
                size3 = hdr + 2 + (size_t)varint_len((uint32_t)w->l_ada) + w->l_ada;
                for (int k = 0; k < NSTR; ++k)
                    if (sn[k] && !((amask >> k) & 1)) size3 += cost[k];
            }
        }
    }
    int mode = 0;
    size_t best = size0;
    if (size1 < best) { best = size1; mode = 1; }
    if (size2 < best) { best = size2; mode = 2; }
    if (size3 < best) { best = size3; mode = 3; }
    if (best + 2 > cap) { packwork_free(w); return 0; }

    size_t p = 0;
    out[p++] = (uint8_t)qp;
    out[p++] = (uint8_t)mode;
    for (int k = 0; k < NSTR; ++k) p = varint_put(out, p, (uint32_t)sn[k]);
    if (mode == 0 || mode == 2) {
        const uint8_t *bytes = mode == 0 ? w->shared : w->shared_r;
        const size_t n = mode == 0 ? w->l_shared : w->l_shared_r;
        if (mode == 0) p = table_put(out, p, w->f1);
        for (int i = 0; i < 4; ++i) out[p++] = (uint8_t)(n >> (8 * i));
        memcpy(out + p, bytes, n);
        p += n;
    } else {
        const unsigned am = mode == 3 ? amask : 0u;
        const unsigned sent_mask = rmask & ~am;
        out[p++] = (uint8_t)(sent_mask & 0xff);
        out[p++] = (uint8_t)(sent_mask >> 8);
        for (int k = 0; k < NSTR; ++k) {
            if (!sn[k] || ((am >> k) & 1)) continue;
            const int reused = (rmask >> k) & 1;
            if (!reused) p = table_put(out, p, w->fk[k]);
            const size_t n = reused ? w->l_reuse[k] : w->l_own[k];
            p = varint_put(out, p, (uint32_t)n);
            memcpy(out + p, reused ? w->reuse[k] : w->own[k], n);
            p += n;
        }
        if (mode == 3) {
            p = varint_put(out, p, (uint32_t)w->l_ada);
            memcpy(out + p, w->ada, w->l_ada);
            p += w->l_ada;
        }
    }
    if (tc) {
        if (mode == 0) { table_keep(tc->f1, w->f1); tc->have = 1; }
        else if (mode == 1 || mode == 3) {
            const unsigned am = mode == 3 ? amask : 0u;
            for (int k = 0; k < NSTR; ++k)
                if (sn[k] && !((am >> k) & 1) && !((rmask >> k) & 1)) table_keep(tc->fs[k], w->fk[k]);
            tc->have = 1;
        }
        for (int k = 0; k < NSTR; ++k) history_add(tc->hist[k], w->cnt[k]);
        if (tc->hn < (1 << 20)) ++tc->hn;
    }
    packwork_free(w);
    return p;
}

/* ---- the key frame's container (fdv_image_encode, fdv.h:4926-5030) ----------
 *
 * Twelve streams, sized three ways and the smallest kept (ties to the lower
 * mode): 0 one static table over all streams, 1 a static table per non-empty
 * stream, 2 one adaptive range-coded blob with a flat model per stream (no
 * tables; a key frame has no history).  Header: width and height as u16, qp,
 * the mode, a varint symbol count per stream; the table and blob lengths are
 * varints. */
#define NIS 12
size_t fdvamd_pack_image_(int w, int h, int qp, const uint8_t *const sv[NIS], const size_t sn[NIS], uint8_t *out,
                          size_t cap) {
    size_t total = 0, hdr = 0, res = 0;
    for (int k = 0; k < NIS; ++k) { total += sn[k]; hdr += (size_t)varint_len((uint32_t)sn[k]); }
    uint32_t call[256] = {0}, (*cnt)[256] = calloc(NIS, sizeof *cnt);
    uint16_t f1[256], (*fk)[256] = malloc(NIS * sizeof *fk);
    uint8_t *all = malloc(total ? total : 1), *shared = malloc(total * 2 + 1024), *ada = malloc(total * 3 + 4096);
    uint8_t *own[NIS] = {0};
    size_t l_own[NIS] = {0};
    if (!cnt || !fk || !all || !shared || !ada) goto done;
    {
        size_t o = 0;
        for (int k = 0; k < NIS; ++k) {
            for (size_t i = 0; i < sn[k]; ++i) cnt[k][sv[k][i]]++;
            if (sn[k]) memcpy(all + o, sv[k], sn[k]);
            o += sn[k];
        }
    }
    for (int k = 0; k < NIS; ++k) for (int s = 0; s < 256; ++s) call[s] += cnt[k][s];
    freqs_from_counts(call, f1);
    const size_t l_shared = rans_pack(all, total, f1, shared, total * 2 + 1024);
    const size_t size0 = hdr + table_bytes(f1) + (size_t)varint_len((uint32_t)l_shared) + l_shared;
    size_t size1 = hdr;
    for (int k = 0; k < NIS; ++k) {
        freqs_from_counts(cnt[k], fk[k]);
        const size_t c = sn[k] * 2 + 1024;
        if (!(own[k] = malloc(c))) goto done;
        l_own[k] = rans_pack(sv[k], sn[k], fk[k], own[k], c);
        if (sn[k]) size1 += table_bytes(fk[k]) + (size_t)varint_len((uint32_t)l_own[k]) + l_own[k];
    }
    size_t size2 = (size_t)-1, l_ada;
    {
        rcoder rc = {ada, total * 3 + 4096, 0, 0, 0xFFFFFFFFu, 0, 1, 0};
        static const uint32_t flat[256];
        for (int k = 0; k < NIS; ++k) {
            amodel m;
            am_seed(&m, flat);
            for (size_t i = 0; i < sn[k]; ++i) rc_code(&rc, &m, sv[k][i]);
        }
        l_ada = rc_finish(&rc);
        if (l_ada) size2 = hdr + (size_t)varint_len((uint32_t)l_ada) + l_ada;
    }
    int mode = 0;
    size_t best = size0;
    if (size1 < best) { best = size1; mode = 1; }
    if (size2 < best) { best = size2; mode = 2; }
    if (5 + best > cap) goto done;
    size_t o = 0;
    out[o++] = (uint8_t)w; out[o++] = (uint8_t)(w >> 8);
    out[o++] = (uint8_t)h; out[o++] = (uint8_t)(h >> 8);
    out[o++] = (uint8_t)qp;
    out[o++] = (uint8_t)mode;
    for (int k = 0; k < NIS; ++k) o = varint_put(out, o, (uint32_t)sn[k]);
    if (mode == 2) {
        o = varint_put(out, o, (uint32_t)l_ada);
        memcpy(out + o, ada, l_ada); o += l_ada;
    } else if (mode == 0) {
        o = table_put(out, o, f1);
        o = varint_put(out, o, (uint32_t)l_shared);
        memcpy(out + o, shared, l_shared); o += l_shared;
    } else {
        for (int k = 0; k < NIS; ++k) {
            if (!sn[k]) continue;
            o = table_put(out, o, fk[k]);
            o = varint_put(out, o, (uint32_t)l_own[k]);
            memcpy(out + o, own[k], l_own[k]); o += l_own[k];
        }
    }
    res = o;
done:
    for (int k = 0; k < NIS; ++k) free(own[k]);
    free(cnt); free(fk); free(all); free(shared); free(ada);
    return res;
}
