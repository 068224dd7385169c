/* fdv_oracle.h -- CPU restatement of fdv's P-frame encode-analysis and
 * reconstruction path (motion search, motion compensation, 4x4/8x8 integer
 * transforms, quantisation + RDOQ, residual coding, intra-in-P, mode decision,
 * deblocking) and of the host entropy packing of a P-frame payload.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product never links it.
 *
 * Every function cites the reference (yhbsh/fdv, include/fdv.h) line it follows.
 * Planes handed in here are "padded": `data` points at pixel (0,0) of a buffer
 * with a replicated border of ORC_BORDER pixels on every side (fdv.h:1548-1630).
 */
#ifndef FDV_ORACLE_H
#define FDV_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#define ORC_BORDER 32
#define ORC_NS 16            /* coded streams of a P-frame (fdv.h:6198)          */
#define ORC_NSTRUCT 9        /* PS_NSTRUCT: 4 mode contexts + REF MVX MVY SUB MERGE */

/* Stream indices (fdv.h:4170, 6199-6207). */
enum { ORC_S_MODE0 = 0, ORC_S_REF = 4, ORC_S_MVX = 5, ORC_S_MVY = 6, ORC_S_SUB = 7,
       ORC_S_MERGE = 8, ORC_S_CN = 9, ORC_S_CL = 10, ORC_S_CN8 = 11, ORC_S_CL8 = 12,
       ORC_S_FL = 13, ORC_S_CNC = 14, ORC_S_CLC = 15 };

/* Motion-search flavours of the analysis. */
enum { ORC_ME_DIAMOND = 0,   /* the reference's EPZS diamond (fdv.h:3375)      */
       ORC_ME_EXHAUSTIVE = 1 /* full +-range integer search, same sub-pel step */ };

/* One macroblock's decision.  mode: 0 SKIP, 1 INTER16, 2 INTRA, 3 INTER8.
 * A given INTRA decision with sub == ORC_SUB_SCREEN has its sub-mode chosen when
 * the MB is coded, by the SATD screen over its reconstructed neighbours. */
#define ORC_SUB_SCREEN 0xFF
typedef struct {
    uint8_t mode, ref, merge_idx, sub;
    int16_t mv[4][2];        /* [0] for SKIP / INTER16, all four for INTER8 */
} orc_mbinfo;

/* The entropy state carried between P-frames (fdv_tabcache, fdv.h:599-612). */
typedef struct {
    uint16_t f1[256];
    uint16_t fs[16][256];
    int      have;
    uint32_t hist[16][256];
    int      hn;
} orc_tabcache;

/* Padded plane helpers (plane_alloc / plane_extend, fdv.h:1548, 1603). */
int  orc_plane_stride(int w);
void orc_plane_extend(uint8_t *data, int w, int h, int stride);
void orc_plane_from_tight(uint8_t *data, int stride, const uint8_t *src, int w, int h);

/* Block kernels. */
int  orc_bits_val(uint32_t v);
void orc_fdct4(const int16_t r[16], int32_t c[16]);
void orc_idct4(const int32_t c[16], int16_t r[16]);
void orc_quant4(const int32_t c[16], int16_t l[16], int qp);
void orc_dequant4(const int16_t l[16], int32_t c[16], int qp);
void orc_rdoq4(int16_t l[16], const int16_t res[16], int qp, double lambda);
void orc_tables8(int32_t c8[64], int32_t qstep8[52]);
void orc_fdct8(const int16_t x[64], int32_t y[64]);
void orc_idct8(const int32_t y[64], int16_t x[64]);
void orc_quant8(const int32_t y[64], int16_t l[64], int qp);
void orc_dequant8(const int16_t l[64], int32_t y[64], int qp);
int  orc_satd4(const uint8_t *src, int sstride, const uint8_t *pred);
void orc_intra_nxn(int mode, const uint8_t *top, const uint8_t *left, uint8_t tl, int n,
                   int ht, int hl, uint8_t *pred);
void orc_mc_luma(const uint8_t *ref, int stride, int bx, int by, int bw, int bh,
                 int mvx, int mvy, uint8_t *dst, int dstride);
void orc_mc_chroma(const uint8_t *ref, int stride, int bx, int by, int bw, int bh,
                   int mvx, int mvy, uint8_t *dst, int dstride);
int  orc_me_search(const uint8_t *cur, int cstride, const uint8_t *ref, int rstride,
                   int bx, int by, int bw, int bh, int range, int pmx, int pmy,
                   int me_mode, int *mvx, int *mvy);
void orc_deblock_plane(uint8_t *p, int w, int h, int stride, int qp);

/* code_residual (fdv.h:5300): returns the 7 coefficient streams' bytes through
 * bufs[0..6] (n l n8 l8 fl nc lc) and their lengths. */
void orc_code_residual(const uint8_t *cur, int cstride, int ox, int oy,
                       const uint8_t *pred, int n, int qp, double lambda, int chroma,
                       uint8_t *rec, uint8_t *const bufs[7], size_t lens[7],
                       double *ssd, int *bits);

/* The lambda of a P-frame (fdv.h:5804). */
double orc_lambda(int qp, double lscale);

/* P-frame analysis + reconstruction over a whole w x h frame (pframe_encode,
 * fdv.h:5764-6170).  refs are padded planes (ystride for luma, cstride for
 * chroma).  Outputs: deblocked tight recon ry/ru/rv; per-MB decisions (may be
 * NULL); the 16 streams in bufs[k] (each of capacity scap) with lengths sn[k].
 * `given` (may be NULL) replaces search and mode decision by fixed decisions.
 * Returns 0, or -1 if a stream buffer would overflow. */
int orc_pframe_analyse(const uint8_t *cy, const uint8_t *cu, const uint8_t *cv,
                       int w, int h,
                       const uint8_t *const ry_ref[2], const uint8_t *const ru_ref[2],
                       const uint8_t *const rv_ref[2], int ystride, int cstride,
                       int navail, int qp, double lambda, int force_skip, int me_mode,
                       const orc_mbinfo *given,
                       uint8_t *ry, uint8_t *ru, uint8_t *rv, orc_mbinfo *mbinfo,
                       uint8_t *const bufs[ORC_NS], size_t scap, size_t sn[ORC_NS]);

/* Entropy packing of a P-frame payload from its 16 streams (fdv.h:6172-6421).
 * tc may be NULL.  Returns the payload length, or 0 on overflow. */
size_t orc_pframe_pack(int qp, const uint8_t *const sv[ORC_NS], const size_t sn[ORC_NS],
                       orc_tabcache *tc, uint8_t *out, size_t cap);

/* Convenience: analyse + pack with tight I420 references (built into padded
 * planes internally).  Mirrors pframe_encode's argument meaning. */
size_t orc_pframe_encode(const uint8_t *cy, const uint8_t *cu, const uint8_t *cv,
                         int w, int h,
                         const uint8_t *r0y, const uint8_t *r0u, const uint8_t *r0v,
                         const uint8_t *r1y, const uint8_t *r1u, const uint8_t *r1v,
                         int navail, int qp, uint8_t *out, size_t cap,
                         uint8_t *ry, uint8_t *ru, uint8_t *rv, int force_skip,
                         orc_tabcache *tc, double lscale, int me_mode,
                         orc_mbinfo *mbinfo);

/* ---- motion tables (the library's exhaustive, parallel and learned modes; not
 * in the reference).  A per-MB motion table, as the GPU's table kernels write it. */
typedef struct {
    int16_t mv16[2][2];
    int32_t sad16[2];
    int16_t mv8[4][2];
    int32_t sad8[4];
    int32_t flags;           /* bit0: SKIP expected -- INTRA and INTER8 are not evaluated */
} orc_metab;

/* The learned proposer's output for one MB (fdvg_nnout). */
typedef struct {
    int8_t cx[8], cy[8];     /* whole-pel vectors on the nearest reference */
    int8_t kx[4], ky[4];     /* the correlation map's 4 lowest-cost positions */
    int8_t qx[4][2], qy[4][2];   /* per quadrant: its map's 2 lowest-cost positions */
    float conf, pskip;
} orc_nnout;

/* Exhaustive-search table of a frame (16x16 per reference, 8x8 quadrants on 0). */
void orc_metab_exhaustive(const uint8_t *cy, int w, int h, const uint8_t *const RY[2], int ys,
                          int navail, orc_metab *mt);
/* The table of the learned mode: where the proposer's confidence is at least
 * thr, the whole-pel search covers the 3x3 neighbourhoods of its vectors (near
 * reference: the network's 8, the correlation map's 4, the quadrant maps' 8;
 * far reference: the first 12 scaled by 2) and the zero vector, with the exhaustive
 * search's tie order, then the sub-pel cross; elsewhere the exhaustive search.
 * flags bit 0 = SKIP probability >= sthr: the decision then weighs SKIP against
 * INTER16 only. */
void orc_metab_candidates(const uint8_t *cy, int w, int h, const uint8_t *const RY[2], int ys, int navail,
                          const orc_nnout *nn, float thr, float sthr, orc_metab *mt);
void orc_set_no_deblock(int on);

/* The MB loop of pframe_encode (orc_pframe_analyse) with the motion vectors
 * and SADs taken from a per-MB table instead of a search: 16x16 per reference,
 * the four 8x8 quadrants on reference 0; an MB whose flags bit 0 is set weighs
 * SKIP against INTER16 only.  With orc_metab_exhaustive's table this is the
 * exhaustive mode's result; the block-parallel modes reproduce it. */
int orc_pframe_analyse_table(const uint8_t *cy, const uint8_t *cu, const uint8_t *cv,
                             int w, int h,
                             const uint8_t *const RY[2], const uint8_t *const RU[2],
                             const uint8_t *const RV[2], int ys, int cs,
                             int navail, int qp, double lambda, const orc_metab *mt,
                             uint8_t *ry, uint8_t *ru, uint8_t *rv, orc_mbinfo *info,
                             uint8_t *const bufs[ORC_NS], size_t scap, size_t sn[ORC_NS]);

/* ---- the key frame (fdv_image_encode, fdv.h:4845; iframe_encode 5487) */
/* The decided quadtree of a plane, in position-indexed planes.  Depth d has
 * nodes of side 64 >> d on a ceil(w/n) x ceil(h/n) grid; split[d] is 0 for a
 * leaf, 1 for a split, 2 for a split implied by the plane's edge; nmode[d] is
 * a leaf's DC/V/H/PLANE mode.  cmode holds per 4x4 cell the mode context map:
 * a 4x4 leaf's mode (0..8), or a large leaf's mode mapped to 4x4 (V 1, H 2,
 * else DC 0).  rflag holds per 8x8 region of a large leaf 0 / 1 (four 4x4) / 2
 * (one 8x8).  lev holds each transform block's levels in raster order at the
 * block's own position; rec the reconstruction before deblocking. */
typedef struct {
    uint8_t *rec;
    int16_t *lev;
    uint8_t *rflag, *cmode;
    uint8_t *split[4], *nmode[4];
} orc_itree;
int orc_itree_alloc(orc_itree *t, int w, int h);
void orc_itree_free(orc_itree *t);
void orc_intra4(int mode, const uint8_t top[8], const uint8_t left[4], uint8_t tl, int ht, int hl,
                uint8_t pred[16]);
void orc_image_analyse(const uint8_t *src, int w, int h, int stride, int qp, orc_itree *t);
size_t orc_image_pack(const orc_itree *t, int w, int h, int qp, uint8_t *out, size_t cap);
size_t orc_image_encode(const uint8_t *src, int w, int h, int stride, int qp, uint8_t *out, size_t cap,
                        uint8_t *recon);
size_t orc_iframe_encode(const uint8_t *y, const uint8_t *u, const uint8_t *v, int w, int h, int qp,
                         uint8_t *out, size_t cap, uint8_t *ry, uint8_t *ru, uint8_t *rv);

#endif
