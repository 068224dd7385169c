// fdvgpu_common.h -- device-side layouts shared by the encode-analysis kernels
// and their host driver.  Plain C structs only: the C-ABI never exposes them.
//
// Terms follow the reference codec (yhbsh/fdv include/fdv.h): a *sequence* is one
// independently coded picture stream -- a whole video, or one band of a
// tile-parallel video (fdv.h:7092-7138) -- and its frames are split into 16x16
// macroblocks (MB) coded in raster order.
#pragma once
#include <stdint.h>

#define FDVG_BORDER 32          // replicated border of every padded plane (fdv.h:217)
#define FDVG_RANGE 16           // integer search radius in pels (fdv.h:5169)
#define FDVG_LIM (FDVG_RANGE * 4)
#define FDVG_NS 16              // coded streams per P-frame (fdv.h:6198)

// Stream indices (fdv.h:4170 structure models, then fdv_cw order 618-627).
enum { SG_REF = 4, SG_MVX = 5, SG_MVY = 6, SG_SUB = 7, SG_MERGE = 8, SG_CN = 9,
       SG_CL = 10, SG_CN8 = 11, SG_CL8 = 12, SG_FL = 13, SG_CNC = 14, SG_CLC = 15 };

// Motion-search flavours.
enum { FDVG_ME_DIAMOND = 0,    // the reference's EPZS diamond, run inside the wavefront
       FDVG_ME_TABLE = 1 };    // per-MB vectors precomputed block-parallel (exhaustive
                               // search, or a learned proposal refined on the GPU)

// One MB decision as supplied by a caller (given-decisions mode) or reported back.
typedef struct {
    uint8_t mode, ref, merge_idx, sub;   // mode: 0 SKIP 1 INTER16 2 INTRA 3 INTER8
    int16_t mv[4][2];
} fdvg_mbinfo;

// Precomputed motion for one MB: 16x16 vector and SAD per reference, the four
// 8x8 quadrant vectors on reference 0, and a skip hint (learned proposer only).
typedef struct {
    int16_t mv16[2][2];
    int32_t sad16[2];
    int16_t mv8[4][2];
    int32_t sad8[4];
    int32_t flags;      // bit0: the proposer expects SKIP -- INTRA and INTER8 are not evaluated
} fdvg_metab;

// The learned proposer's output for one MB: 8 whole-pel vectors on the nearest
// reference (pels, best first), the 4 lowest-cost positions of its half-res
// correlation map (pels; ties to the smaller |dx| + |dy|, then the lower
// position), the 2 lowest of each 8x8 quadrant's own map (same ties), the 8
// vectors' total probability, the SKIP probability.
typedef struct {
    int8_t cx[8], cy[8];
    int8_t kx[4], ky[4];
    int8_t qx[4][2], qy[4][2];   // per 8x8 quadrant: the 2 lowest-cost positions of its map
    float conf, pskip;
} fdvg_nnout;

// Block-parallel decisions (parallel and learned modes).  Per MB: the residual
// SSD and bits of the INTER16 and INTER8 candidates, without the
// vector-difference bits (which depend on the predictor), and the SKIP SSDs of
// a set of likely merge candidates (the MB's own and its neighbours' table
// vectors).  INTRA is costed in the replay, from the true reconstruction.
#define FDVG_NSC 12
// A given INTRA decision whose sub-mode is chosen when the MB is coded, by the
// SATD screen over its reconstructed neighbours (fdv.h:5950-5970).
#define FDVG_SUB_SCREEN 0xFF
typedef struct {
    int32_t d16[2], b16[2];           // INTER16 on each reference (bits incl. the 6 mode/ref bits)
    int32_t d8, b8;                   // INTER8 (bits incl. the 3 mode bits)
    int16_t scx[FDVG_NSC], scy[FDVG_NSC];
    int32_t sce[FDVG_NSC];
    uint8_t scr[FDVG_NSC];
    uint8_t nsc, pad[3];
} fdvg_pcost;

// The true vector predictor and merge list of an MB once its neighbours are
// decided (fdv.h:5680-5732).
typedef struct {
    int16_t pmx, pmy;
    int16_t cx[3], cy[3];
    uint8_t cr[3], ncand;
} fdvg_pctx;

// What the wavefront hands to the serialiser for one MB: the decision and the
// winning candidate's quantised levels in symbolic form (region flags + levels).
// Luma region r (raster 8x8 of the MB) keeps its levels at lev[r*64..]: four
// raster 4x4 blocks of 16 (flag 1) or one raster 8x8 (flag 2).  Chroma plane p
// keeps clev[p*64..] in the same region layout.
typedef struct {
    uint8_t mode, ref, merge_idx, ncand;
    uint8_t sub, ckind, pad0, pad1;   // ckind: 0 none(SKIP) 1 any-flag 4x4 (INTER16)
                                      //        2 region flag (INTRA) 3 flagless 4x4 (INTER8)
    uint8_t lflag[4];
    uint8_t cflag[2];
    uint8_t pad2[2];
    int16_t mvd[4][2];                // coded vector differences (INTER16: [0]; INTER8: 4)
    int16_t mv[4][2];                 // the vectors themselves
    int16_t lev[256];
    int16_t clev[128];
} fdvg_mbrec;

// One sequence's frame job.
typedef struct {
    const uint8_t *cy, *cu, *cv;       // tight source planes (w, w/2 strides)
    const uint8_t *ry[2], *ru[2], *rv[2];  // padded reference planes, at pixel (0,0)
    uint8_t *oy, *ou, *ov;             // padded output planes (the next reference)
    const fdvg_mbinfo *given;          // NULL, or one decision per MB
    const fdvg_metab *metab;           // used when me_mode == FDVG_ME_TABLE
    fdvg_mbinfo *info;                 // per-MB decisions out (may be NULL)
    double lambda;
    int w, h, ys, cs;                  // luma dims, padded luma / chroma strides
    int navail, qp, me_mode, force_skip;
    int mb0;                           // first MB of this sequence in the global MB arrays
    int tag;                           // 1..255, differs between consecutive frames
    unsigned step;                     // the context's frame-step id: nonzero, new for every step
    int pad;
} fdvg_seq;

// Per-MB handoff to the MB row below: the bottom recon rows (16 Y, 8 U, 8 V),
// the motion record and the mode, packed 7 bytes to a 64-bit word whose top byte
// is the frame tag.  Written and read with relaxed device-scope atomics, so a
// reader on any XCD sees either the old frame's word or this frame's whole word.
#define FDVG_HWORDS 6

// A plane for the block-parallel passes (deblock, borders, padding).
typedef struct { uint8_t *p; int w, h, stride, qp; } PlaneRef;

// Where one sequence's MBs live, for the serialiser.
typedef struct { int mb0, mw, mh, pad; } SymSeq;

// One plane of a key frame for the intra kernels (fdv_image_encode, fdv.h:4845).
// Coding tree units of 64x64, cols x rows of them; depth d of the quadtree has
// nodes of side 64 >> d on a tw[d] x th[d] grid, whose split flags (0 leaf,
// 1 split, 2 split implied by the plane's edge) start at tree + toff[d] and
// leaf modes at tree + toff[d] + tw[d] * th[d].  cmode: per 4x4 cell the mode
// context value (a 4x4 leaf's mode, or a large leaf's mode mapped to 4x4).
// lev: the levels of each surviving transform block in raster order at the
// block's own position; rflag: per 8x8 region of a large leaf 0 / 1 / 2.
// edge: per unit 32 tagged words, its bottom row then its right column, four
// samples a word.
typedef struct {
    const uint8_t *src;
    uint8_t *rec;
    int16_t *lev;
    uint8_t *rflag, *cmode, *tree;
    uint64_t *edge;
    double lambda;
    int w, h, qp, cols, rows, pad;
    int toff[4], tw[4], th[4];
} fdvg_iplane;
