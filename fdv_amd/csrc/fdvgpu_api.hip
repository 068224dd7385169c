// fdvgpu_api.hip -- host side of the C ABI declared in include/fdvgpu.h.
//
// Device memory per sequence (sized for the frame; 288 GB of HBM holds thousands):
//   3 padded frames (fdv.h plane layout: stride = align32(w + 64), h + 64 rows,
//     data at (32, 32)) rotating as near / far / output reference,
//   nsrc tight I420 source slots,
//   per-MB records (decision + levels), handoff words, stream counts.
// One HIP stream (default priority) per context; every frame step is the same
// fixed launch sequence, so it can be captured into a graph by a caller.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "fdvgpu_common.h"
#include "../../include/fdvgpu.h"

extern "C" int fdvg_set_tables(const int32_t *c8, const int32_t *qs8);
extern "C" int fdvg_set_tables_decode(const int32_t *c8, const int32_t *qs8);
extern "C" int fdvg_launch_decode(const fdvg_seq *, int, int, const fdvg_mbinfo *, int, const fdvg_mbrec *, unsigned *,
                                  unsigned *, uint64_t *, hipStream_t);
extern "C" int fdvg_launch_analyse(const fdvg_seq *, const int2 *, int, unsigned *, fdvg_mbrec *,
                                   uint64_t *, hipStream_t);
extern "C" int fdvg_launch_deblock(const void *, int, int, int, hipStream_t);
extern "C" int fdvg_launch_border(const void *, int, int, hipStream_t);
extern "C" int fdvg_launch_pad(const void *, const uint8_t *const *, int, int, hipStream_t);
extern "C" int fdvg_launch_zero(uint32_t *, int, hipStream_t);
extern "C" int fdvg_launch_stage(const uint32_t *, const unsigned *, int, const uint8_t *, size_t, uint32_t *, uint8_t *,
                                 hipStream_t);
extern "C" int fdvg_launch_symbols(const void *, int, int, const fdvg_mbrec *, uint32_t *, uint32_t *,
                                   uint32_t *, uint8_t *, size_t, int, hipStream_t);

// Launch shapes the autotuner chooses among (fdvgpu_autotune).
struct Tune {
    int sym_waves;      // serialiser MBs (wavefronts) per workgroup: 1, 2, 4, 8
    int deblock_bs;     // deblock threads per workgroup: 64, 128, 256
    int nn_wpc;         // proposer workgroups per CU: 1, 2
    int pd_wpc;         // coding workers of the decision stage per CU: 1, 2
};
extern "C" int fdvg_launch_metab(const fdvg_seq *, int, int, void *, const void *, float, float, hipStream_t);
extern "C" int fdvg_launch_nn_infer(const fdvg_seq *, int, int, const float *, fdvg_nnout *, int, hipStream_t);
extern "C" int fdvg_nn_nparam(void);
extern "C" int fdvg_nn_setup(void);

// The proposer's confidence threshold: below it an MB gets the exhaustive search.
#define FDVG_NN_THRESHOLD 0.7f     // default confidence threshold of the proposer
#define FDVG_NN_SKIP_OFF 2.0f      // default SKIP threshold: no pruning
extern "C" int fdvg_launch_parallel(const fdvg_seq *, int, int, const int2 *, int, const unsigned *, const int *,
                                    int, unsigned *, uint64_t *, uint64_t *, fdvg_mbinfo *, fdvg_pctx *, uint64_t *,
                                    uint64_t *, uint64_t *, unsigned *, uint64_t *, int, fdvg_mbrec *, hipStream_t);
extern "C" int fdvg_launch_pcosts(const fdvg_seq *, int, const unsigned *, const int *, int, unsigned *, uint64_t *,
                                  uint64_t *, hipStream_t);
extern "C" int fdvg_launch_pxrep(const fdvg_seq *, const int2 *, int, unsigned *, const uint64_t *, const uint64_t *,
                                 fdvg_mbinfo *, fdvg_pctx *, uint64_t *, uint64_t *, uint64_t *, hipStream_t);
extern "C" int fdvg_pdecide_plan(const int *, const int *, const int *, int, int2 *, unsigned *);
extern "C" int fdvg_xcd_order(const int *, const int *, const int *, int, const unsigned *, int, unsigned *, int *);
extern "C" int fdvg_pdecide_grid(int, int, int);
extern "C" int fdvg_launch_pdecide(const fdvg_seq *, int, const unsigned *, const int *, int, unsigned *,
                                   const uint64_t *, const uint64_t *, uint64_t *, fdvg_mbrec *, hipStream_t);
// One captured graph per (stage 0-3, search mode 0-3, block-parallel or not,
// decision structure 0-4 for the block-parallel decision stage); distinct
// for every combination.
#define NGRAPH (4 * 4 * 2 * 5)
static inline int graph_key(int stage, int me_mode, bool par, int overlap) {
    return ((stage * 4 + me_mode) * 2 + (par ? 1 : 0)) * 5 + (stage == 1 && par ? overlap : 0);
}
// the control words of a frame step: [0] row ticket, [1] error flag, [3..7]
// kernel tickets, [8..15] and [16..23] the per-XCD tickets of the cost and
// coding kernels
#define CTL_WORDS 24
#define PC_WORDS 34                     // a MB's candidate costs as tagged words
#define EW_WORDS 24                     // a MB's candidates' reconstruction edges as tagged words
#define BW_WORDS 8                      // a decided MB's reconstruction edges, tagged: bottom row, right column
#define PD_WORDS 10                     // a MB's decision and context as tagged words

static thread_local char g_err[256];
static void seterr(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
extern "C" const char *fdvgpu_last_error(void) { return g_err; }
// The same message slot, for the library's other source files.
extern "C" void fdvg_seterr(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    seterr("%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); return -1; } } while (0)
#define CKN(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    seterr("%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); goto fail; } } while (0)

struct SeqMem {
    int w, h, cw, ch, ys, cs, mw, mh, nmb, mb0;
    size_t ysz, csz, fsz;          // padded Y plane bytes, chroma plane bytes, frame bytes
    uint8_t *frames;               // 3 padded frames
    int near_, far_, out_;
    uint8_t *src;                  // nsrc tight frames
    size_t srcsz;
    int has_given, tag;
};

struct fdvgpu {
    int nseq, nsrc, nrows, total_mb, maxmb, maxpx, maxmh, maxmw;
    int xgrid;                      // grid width of the one-workgroup-per-MB kernels (XCD strips)
    SeqMem *S;
    hipStream_t st;
    fdvg_seq *seq_d, *seq_h;        // seq_h: RING staging copies, one per frame in flight
    int2 *rowtab_d;
    unsigned *ticket_d;             // CTL_WORDS control words (tickets, error flag)
    fdvg_mbrec *recs_d;
    uint64_t *hand_d;
    fdvgpu_mbinfo *info_d, *given_d;
    void *metab_d;
    PlaneRef *planes_d, *planes_h;  // planes_h: RING staging copies
    hipEvent_t ring_ev[8];
    int ring_pos;
    SymSeq *sym_d;
    uint32_t *cnt_d, *totals_d, *bases_d, *bases_h;
    uint8_t *compact_d, *compact_h;
    // block-parallel decisions
    uint64_t *pcw_d;                // per MB: the candidate costs as tagged words
    uint64_t *edgew_d;              // per MB: the coded candidates' bottom rows and right columns, tagged
    uint64_t *botw_d;               // per MB: the decided reconstruction's bottom row, tagged (for the row below)
    unsigned *xorder_d;             // per XCD: the MBs of its rows, in the order the replays decide them
    int *xoff_d;                    // list k of xorder_d: xoff_d[k] .. xoff_d[k + 1]
    int xmax;                       // the longest list
    fdvg_pctx *pctx_d;
    fdvg_mbinfo *pgiven_d;
    uint64_t *mrec_d;               // motion-record handoff of the decision wavefront
    unsigned *ilist_d;              // intra MBs in dependency order (count first)
    uint64_t *ih_d;                 // intra handoff words, 16 per MB (bottom row, right column)
    unsigned step_ctr;              // frame-step ids (never rewound)
    // The decision stage: the costs (k_pcosts), the replays (k_pxrep, one MB
    // row a workgroup), the coding of the decisions.  Structure 0: separate
    // launches, the coding by k_pencode / k_pintra; 1: the coding by the
    // workers of k_pdecide after the replays; 2: the replays on a second stream
    // beside the costs, the workers beside the replays.
    int2 *units_d;                  // replay units (sequence, row)
    uint64_t *decw_d;               // per MB: the decision and its context as tagged words
    int nunit;
    int overlap, overlap_avail;     // the decision structure; whether 2 fits beside the replays
    int overlap_fixed;              // set by the caller: the autotuner keeps it
    hipStream_t st2;
    hipEvent_t ev_fork, ev_join;
    // launch shapes, and one instantiated graph per (search mode, decision structure)
    Tune tune;
    int graphs;
    hipGraphExec_t gexec[NGRAPH];   // per (stage, search mode, block-parallel, decision structure): graph_key
    int warm[NGRAPH];               // the key has run once eagerly (launch caches filled)
    // learned proposer
    float *nnw_d;
    fdvg_nnout *nn_d;
    int have_nn;
    float nn_thr, nn_sthr;          // proposer thresholds: candidate search, SKIP pruning
    size_t compact_cap;
    uint32_t *flags_h;
    hipEvent_t ev[5];
    // frame steps queued from resident sources (fdvgpu_submit_slots): the
    // serialisation and staging run on st3 beside the next step's tables
    hipStream_t st3;
    hipEvent_t ev_db, ev_ser;       // deblock done (st) / serialisation and staging done (st3)
    hipEvent_t pev[2][5];           // per staging slot: the step's timing events
    hipEvent_t *cev;                // the timing events of the step being queued
    unsigned *flags_snap_d;         // per staging slot: the step's ticket words 0, 1
    int ser_async;                  // the step being queued serialises on st3
    int ser_slot;                   // ... into this staging slot
    int ser_pending;                // a serialisation may still run on st3
    // pipelined frame steps (fdvgpu_submit / fdvgpu_collect), two slots: the
    // pinned upload copies of the sources, the host buffers the staging kernel
    // writes (mapped for the device), and the event after that kernel
    uint8_t *pipe_up[2];
    uint8_t *pipe_sc[2], *pipe_sc_d[2];
    uint32_t *pipe_sb[2], *pipe_sb_d[2];
    hipEvent_t pipe_ev[2];
    int pipe_nev;                   // events created
    int pipe_ok;                    // all of the above allocated
    int have_frame, fetched;
    int can_rewind;                 // the last call was fdvgpu_pframe: its rotation can be undone
    int last_mode;                  // the search mode of the last fdvgpu_pframe
    int keep_tables;                // the next fdvgpu_pframe reuses the motion tables of the undone one
};

static uint8_t *fy(const SeqMem &m, int f) { return m.frames + (size_t)f * m.fsz + (size_t)FDVG_BORDER * m.ys + FDVG_BORDER; }
static uint8_t *fu(const SeqMem &m, int f) { return m.frames + (size_t)f * m.fsz + m.ysz + (size_t)FDVG_BORDER * m.cs + FDVG_BORDER; }
static uint8_t *fv(const SeqMem &m, int f) { return m.frames + (size_t)f * m.fsz + m.ysz + m.csz + (size_t)FDVG_BORDER * m.cs + FDVG_BORDER; }

// 8x8 DCT basis and quantiser steps, from libm exactly as the reference builds
// them (c8_init, fdv.h:2597-2612).
extern "C" void fdvg_tables8(int32_t c8[64], int32_t qs8[52]) {
    const double pi = 3.14159265358979323846;
    for (int u = 0; u < 8; ++u) {
        double a = u ? sqrt(2.0 / 8.0) : sqrt(1.0 / 8.0);
        for (int x = 0; x < 8; ++x) c8[u * 8 + x] = (int32_t)lround(a * cos((2 * x + 1) * u * pi / 16.0) * 512.0);
    }
    for (int u = 1; u < 8; ++u) {
        int32_t s = 0;
        for (int x = 0; x < 8; ++x) s += c8[u * 8 + x];
        c8[u * 8] -= s;
    }
    for (int q = 0; q < 52; ++q) qs8[q] = (int32_t)lround(262144.0 * (0.625 * pow(2.0, q / 6.0)));
}

extern "C" void fdvgpu_close(fdvgpu *g) {
    if (!g) return;
    // every queued step, on each of the context's streams, ends before its memory goes
    if (g->st) (void)hipStreamSynchronize(g->st);
    if (g->st2) (void)hipStreamSynchronize(g->st2);
    if (g->st3) (void)hipStreamSynchronize(g->st3);
    if (g->S) for (int s = 0; s < g->nseq; ++s) { (void)hipFree(g->S[s].frames); (void)hipFree(g->S[s].src); }
    (void)hipFree(g->seq_d); (void)hipHostFree(g->seq_h); (void)hipFree(g->rowtab_d); (void)hipFree(g->ticket_d);
    (void)hipFree(g->recs_d); (void)hipFree(g->hand_d); (void)hipFree(g->info_d); (void)hipFree(g->given_d);
    (void)hipFree(g->metab_d); (void)hipFree(g->planes_d); (void)hipHostFree(g->planes_h); (void)hipFree(g->sym_d);
    (void)hipFree(g->cnt_d); (void)hipFree(g->totals_d); (void)hipFree(g->bases_d); (void)hipHostFree(g->bases_h);
    (void)hipFree(g->compact_d); (void)hipHostFree(g->compact_h); (void)hipHostFree(g->flags_h);
    (void)hipFree(g->pcw_d); (void)hipFree(g->pctx_d); (void)hipFree(g->pgiven_d); (void)hipFree(g->mrec_d);
    (void)hipFree(g->ilist_d); (void)hipFree(g->ih_d);
    (void)hipFree(g->nnw_d); (void)hipFree(g->nn_d);
    for (int k = 0; k < NGRAPH; ++k) if (g->gexec[k]) (void)hipGraphExecDestroy(g->gexec[k]);
    for (int i = 0; i < 5; ++i) if (g->ev[i]) (void)hipEventDestroy(g->ev[i]);
    for (int k = 0; k < 2; ++k)
        for (int i = 0; i < 5; ++i) if (g->pev[k][i]) (void)hipEventDestroy(g->pev[k][i]);
    if (g->ev_db) (void)hipEventDestroy(g->ev_db);
    if (g->ev_ser) (void)hipEventDestroy(g->ev_ser);
    if (g->st3) (void)hipStreamDestroy(g->st3);
    (void)hipFree(g->flags_snap_d);
    for (int i = 0; i < 8; ++i) if (g->ring_ev[i]) (void)hipEventDestroy(g->ring_ev[i]);
    if (g->st) (void)hipStreamDestroy(g->st);
    if (g->st2) (void)hipStreamDestroy(g->st2);
    if (g->ev_fork) (void)hipEventDestroy(g->ev_fork);
    if (g->ev_join) (void)hipEventDestroy(g->ev_join);
    (void)hipFree(g->units_d); (void)hipFree(g->xorder_d); (void)hipFree(g->xoff_d); (void)hipFree(g->decw_d);
    (void)hipFree(g->edgew_d); (void)hipFree(g->botw_d);
    for (int k = 0; k < 2; ++k) {
        if (g->pipe_up[k]) (void)hipHostFree(g->pipe_up[k]);
        if (g->pipe_sc[k]) (void)hipHostFree(g->pipe_sc[k]);
        if (g->pipe_sb[k]) (void)hipHostFree(g->pipe_sb[k]);
        if (k < g->pipe_nev) (void)hipEventDestroy(g->pipe_ev[k]);
    }
    free(g->S);
    free(g);
}

extern "C" fdvgpu *fdvgpu_open(int nseq, const int *w, const int *h, int nsrc) {
    if (nseq <= 0 || !w || !h || nsrc <= 0) { seterr("fdvgpu_open: bad arguments"); return NULL; }
    for (int s = 0; s < nseq; ++s)
        if (w[s] <= 0 || h[s] <= 0 || (w[s] & 15) || (h[s] & 15) || w[s] > 8192 || h[s] > 8192) {
            seterr("fdvgpu_open: sequence %d has size %dx%d (multiples of 16, at most 8192)", s, w[s], h[s]);
            return NULL;
        }
    fdvgpu *g = (fdvgpu *)calloc(1, sizeof *g);
    if (!g) { seterr("out of host memory"); return NULL; }
    g->nseq = nseq; g->nsrc = nsrc;
    g->S = (SeqMem *)calloc((size_t)nseq, sizeof(SeqMem));
    int maxmh = 0;
    int2 *rowtab_h = NULL;
    SymSeq *sym_h = NULL;
    int32_t c8[64], qs8[52];
    if (!g->S) { seterr("out of host memory"); goto fail; }
    for (int s = 0; s < nseq; ++s) {
        SeqMem &m = g->S[s];
        m.w = w[s]; m.h = h[s]; m.cw = m.w / 2; m.ch = m.h / 2;
        m.ys = (m.w + 2 * FDVG_BORDER + 31) & ~31;
        m.cs = (m.cw + 2 * FDVG_BORDER + 31) & ~31;
        m.mw = m.w / 16; m.mh = m.h / 16; m.nmb = m.mw * m.mh;
        m.ysz = (size_t)m.ys * (m.h + 2 * FDVG_BORDER);
        m.csz = (size_t)m.cs * (m.ch + 2 * FDVG_BORDER);
        m.fsz = (m.ysz + 2 * m.csz + 255) & ~(size_t)255;
        m.srcsz = ((size_t)m.w * m.h + 2 * (size_t)m.cw * m.ch + 255) & ~(size_t)255;
        m.mb0 = g->total_mb;
        g->total_mb += m.nmb;
        g->nrows += m.mh;
        if (m.mh > maxmh) maxmh = m.mh;
        if (m.mh > g->maxmh) g->maxmh = m.mh;
        if (m.mw > g->maxmw) g->maxmw = m.mw;
        if (m.nmb > g->maxmb) g->maxmb = m.nmb;
        // one-workgroup-per-MB kernels: 8 column strips of ceil(mw / 8), one per XCD
        if (m.mh * ((m.mw + 7) >> 3) * 8 > g->xgrid) g->xgrid = m.mh * ((m.mw + 7) >> 3) * 8;
        if ((int)((size_t)m.ys * (m.h + 2 * FDVG_BORDER)) > g->maxpx) g->maxpx = (int)((size_t)m.ys * (m.h + 2 * FDVG_BORDER));
        m.near_ = 0; m.far_ = 1; m.out_ = 2;
        CKN(hipMalloc(&m.frames, 3 * m.fsz));
        CKN(hipMemset(m.frames, 0, 3 * m.fsz));
        CKN(hipMalloc(&m.src, (size_t)nsrc * m.srcsz));
        g->compact_cap += (size_t)m.nmb * 2048 + 4096;
    }
    CKN(hipStreamCreate(&g->st));
    CKN(hipStreamCreate(&g->st2));
    CKN(hipEventCreateWithFlags(&g->ev_fork, hipEventDisableTiming));
    CKN(hipEventCreateWithFlags(&g->ev_join, hipEventDisableTiming));
    for (int i = 0; i < 5; ++i) CKN(hipEventCreate(&g->ev[i]));
    CKN(hipStreamCreate(&g->st3));
    CKN(hipEventCreateWithFlags(&g->ev_db, hipEventDisableTiming));
    CKN(hipEventCreateWithFlags(&g->ev_ser, hipEventDisableTiming));
    for (int k = 0; k < 2; ++k)
        for (int i = 0; i < 5; ++i) CKN(hipEventCreate(&g->pev[k][i]));
    CKN(hipMalloc(&g->flags_snap_d, 4 * sizeof(unsigned)));
    g->cev = g->ev;
    for (int i = 0; i < 8; ++i) CKN(hipEventCreateWithFlags(&g->ring_ev[i], hipEventDisableTiming));
    CKN(hipMalloc(&g->seq_d, nseq * sizeof(fdvg_seq)));
    CKN(hipHostMalloc(&g->seq_h, 8 * nseq * sizeof(fdvg_seq), hipHostMallocDefault));
    CKN(hipMalloc(&g->rowtab_d, g->nrows * sizeof(int2)));
    CKN(hipMalloc(&g->ticket_d, CTL_WORDS * sizeof(unsigned)));
    CKN(hipMalloc(&g->recs_d, (size_t)g->total_mb * sizeof(fdvg_mbrec)));
    CKN(hipMalloc(&g->hand_d, (size_t)g->total_mb * FDVG_HWORDS * sizeof(uint64_t)));
    CKN(hipMemset(g->hand_d, 0, (size_t)g->total_mb * FDVG_HWORDS * sizeof(uint64_t)));
    CKN(hipMalloc(&g->info_d, (size_t)g->total_mb * sizeof(fdvgpu_mbinfo)));
    CKN(hipMalloc(&g->given_d, (size_t)g->total_mb * sizeof(fdvgpu_mbinfo)));
    CKN(hipMalloc(&g->metab_d, (size_t)g->total_mb * sizeof(fdvg_metab)));
    CKN(hipMemset(g->metab_d, 0, (size_t)g->total_mb * sizeof(fdvg_metab)));
    CKN(hipMalloc(&g->planes_d, 3 * nseq * sizeof(PlaneRef)));
    CKN(hipHostMalloc(&g->planes_h, 8 * 3 * nseq * sizeof(PlaneRef), hipHostMallocDefault));
    CKN(hipMalloc(&g->sym_d, nseq * sizeof(SymSeq)));
    CKN(hipMalloc(&g->cnt_d, (size_t)g->total_mb * FDVG_NS * sizeof(uint32_t)));
    CKN(hipMalloc(&g->totals_d, (size_t)nseq * FDVG_NS * sizeof(uint32_t)));
    CKN(hipMalloc(&g->bases_d, ((size_t)nseq * FDVG_NS + 1) * sizeof(uint32_t)));
    CKN(hipHostMalloc(&g->bases_h, ((size_t)nseq * FDVG_NS + 1) * sizeof(uint32_t), hipHostMallocDefault));
    CKN(hipMalloc(&g->compact_d, g->compact_cap));
    CKN(hipHostMalloc(&g->compact_h, g->compact_cap, hipHostMallocDefault));
    CKN(hipHostMalloc(&g->flags_h, 64, hipHostMallocDefault));
    // tagged words: id 0 is never a frame step
    CKN(hipMalloc(&g->pcw_d, (size_t)g->total_mb * PC_WORDS * sizeof(uint64_t)));
    CKN(hipMemset(g->pcw_d, 0, (size_t)g->total_mb * PC_WORDS * sizeof(uint64_t)));
    CKN(hipMalloc(&g->edgew_d, (size_t)g->total_mb * EW_WORDS * sizeof(uint64_t)));
    CKN(hipMemset(g->edgew_d, 0, (size_t)g->total_mb * EW_WORDS * sizeof(uint64_t)));
    CKN(hipMalloc(&g->botw_d, (size_t)g->total_mb * BW_WORDS * sizeof(uint64_t)));
    CKN(hipMemset(g->botw_d, 0, (size_t)g->total_mb * BW_WORDS * sizeof(uint64_t)));
    CKN(hipMalloc(&g->pctx_d, (size_t)g->total_mb * sizeof(fdvg_pctx)));
    CKN(hipMalloc(&g->pgiven_d, (size_t)g->total_mb * sizeof(fdvg_mbinfo)));
    CKN(hipMalloc(&g->mrec_d, (size_t)g->total_mb * sizeof(uint64_t)));
    CKN(hipMemset(g->mrec_d, 0, (size_t)g->total_mb * sizeof(uint64_t)));
    CKN(hipMalloc(&g->ilist_d, ((size_t)g->total_mb + 1) * sizeof(unsigned)));
    CKN(hipMalloc(&g->ih_d, (size_t)g->total_mb * 16 * sizeof(uint64_t)));
    CKN(hipMemset(g->ih_d, 0, (size_t)g->total_mb * 16 * sizeof(uint64_t)));   // id 0: never a step
    {
        // the block-parallel order and the replay units (one a MB row)
        int *wv = (int *)malloc(3 * nseq * sizeof(int));
        int2 *units_h = (int2 *)malloc(((size_t)g->nrows + nseq) * sizeof(int2));
        unsigned *corder_h = (unsigned *)malloc(2 * (size_t)g->total_mb * sizeof(unsigned));
        int xoff_h[9];
        if (!wv || !units_h || !corder_h) { free(wv); free(units_h); free(corder_h); seterr("out of host memory"); goto fail; }
        for (int k = 0; k < nseq; ++k) { wv[k] = g->S[k].w; wv[nseq + k] = g->S[k].h; wv[2 * nseq + k] = g->S[k].mb0; }
        g->nunit = fdvg_pdecide_plan(wv, wv + nseq, wv + 2 * nseq, nseq, units_h, corder_h);
        unsigned *xorder_h = corder_h + g->total_mb;
        g->xmax = g->nunit > 0 ? fdvg_xcd_order(wv, wv + nseq, wv + 2 * nseq, nseq, corder_h, g->total_mb, xorder_h, xoff_h) : 0;
        // side by side only where the workers leave room for every replay
        g->overlap_avail = g->nunit > 0 && fdvg_pdecide_grid(g->nunit, g->total_mb, -1) > 0;
        g->overlap = g->overlap_avail ? 4 : 3;
        hipError_t e = g->nunit <= 0 ? hipErrorInvalidValue : hipSuccess;
        if (e == hipSuccess) e = hipMalloc(&g->xorder_d, (size_t)g->total_mb * sizeof(unsigned));
        if (e == hipSuccess)
            e = hipMemcpy(g->xorder_d, xorder_h, (size_t)g->total_mb * sizeof(unsigned), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&g->xoff_d, 9 * sizeof(int));
        if (e == hipSuccess) e = hipMemcpy(g->xoff_d, xoff_h, 9 * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&g->units_d, (size_t)g->nunit * sizeof(int2));
        if (e == hipSuccess) e = hipMalloc(&g->decw_d, (size_t)g->total_mb * PD_WORDS * sizeof(uint64_t));
        if (e == hipSuccess) e = hipMemset(g->decw_d, 0, (size_t)g->total_mb * PD_WORDS * sizeof(uint64_t));
        if (e == hipSuccess) e = hipMemcpy(g->units_d, units_h, (size_t)g->nunit * sizeof(int2), hipMemcpyHostToDevice);
        free(wv); free(units_h); free(corder_h);
        if (e != hipSuccess) { seterr("block-parallel plan: %s", hipGetErrorString(e)); goto fail; }
    }
    CKN(hipMalloc(&g->nnw_d, (size_t)fdvg_nn_nparam() * sizeof(float)));
    if (fdvg_nn_setup() != 0) { seterr("proposer kernel setup failed"); goto fail; }
    g->tune = Tune{4, 256, 1, 2};
    g->graphs = 1;
    g->nn_thr = FDVG_NN_THRESHOLD;
    g->nn_sthr = FDVG_NN_SKIP_OFF;
    CKN(hipMalloc(&g->nn_d, (size_t)g->total_mb * sizeof(fdvg_nnout)));

    // rows in (row, sequence) order: every sequence's row r before any row r+1
    rowtab_h = (int2 *)malloc(g->nrows * sizeof(int2));
    sym_h = (SymSeq *)malloc(nseq * sizeof(SymSeq));
    if (!rowtab_h || !sym_h) { seterr("out of host memory"); goto fail; }
    {
        int k = 0;
        for (int r = 0; r < maxmh; ++r)
            for (int s = 0; s < nseq; ++s)
                if (r < g->S[s].mh) rowtab_h[k++] = make_int2(s, r);
        for (int s = 0; s < nseq; ++s) sym_h[s] = SymSeq{g->S[s].mb0, g->S[s].mw, g->S[s].mh, 0};
    }
    CKN(hipMemcpy(g->rowtab_d, rowtab_h, g->nrows * sizeof(int2), hipMemcpyHostToDevice));
    CKN(hipMemcpy(g->sym_d, sym_h, nseq * sizeof(SymSeq), hipMemcpyHostToDevice));
    free(rowtab_h); rowtab_h = NULL;
    free(sym_h); sym_h = NULL;
    fdvg_tables8(c8, qs8);
    if (fdvg_set_tables(c8, qs8) != 0 || fdvg_set_tables_decode(c8, qs8) != 0) {
        seterr("could not load the 8x8 tables"); goto fail;
    }
    return g;
fail:
    free(rowtab_h); free(sym_h);
    fdvgpu_close(g);
    return NULL;
}

static int seq_ok(fdvgpu *g, int seq) {
    if (!g || seq < 0 || seq >= g->nseq) { seterr("bad context or sequence index"); return 0; }
    return 1;
}

extern "C" int fdvgpu_put_source(fdvgpu *g, int seq, int slot, const uint8_t *i420) {
    if (!seq_ok(g, seq) || slot < 0 || slot >= g->nsrc || !i420) { seterr("fdvgpu_put_source: bad arguments"); return -1; }
    SeqMem &m = g->S[seq];
    CK(hipMemcpyAsync(m.src + (size_t)slot * m.srcsz, i420, (size_t)m.w * m.h + 2 * (size_t)m.cw * m.ch,
                      hipMemcpyHostToDevice, g->st));
    CK(hipStreamSynchronize(g->st));
    return 0;
}

extern "C" int fdvgpu_put_ref(fdvgpu *g, int seq, int which, const uint8_t *i420) {
    if (!seq_ok(g, seq) || (which != 0 && which != 1) || !i420) { seterr("fdvgpu_put_ref: bad arguments"); return -1; }
    SeqMem &m = g->S[seq];
    g->can_rewind = 0;
    g->keep_tables = 0;
    int f = which == 0 ? m.near_ : m.far_;
    size_t tsz = (size_t)m.w * m.h + 2 * (size_t)m.cw * m.ch;
    uint8_t *tmp = NULL;
    uint8_t **srcs_d = NULL;
    PlaneRef *pl_d = NULL;
    CK(hipMalloc(&tmp, tsz));
    PlaneRef pl[3] = {{fy(m, f), m.w, m.h, m.ys, 0}, {fu(m, f), m.cw, m.ch, m.cs, 0}, {fv(m, f), m.cw, m.ch, m.cs, 0}};
    const uint8_t *srcs[3] = {tmp, tmp + (size_t)m.w * m.h, tmp + (size_t)m.w * m.h + (size_t)m.cw * m.ch};
    int rc = -1;
    if (hipMalloc(&srcs_d, sizeof srcs) != hipSuccess || hipMalloc(&pl_d, sizeof pl) != hipSuccess) {
        seterr("fdvgpu_put_ref: device allocation failed");
    } else if (hipMemcpy(tmp, i420, tsz, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(srcs_d, srcs, sizeof srcs, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(pl_d, pl, sizeof pl, hipMemcpyHostToDevice) != hipSuccess) {
        seterr("fdvgpu_put_ref: copy failed");
    } else if (fdvg_launch_pad(pl_d, srcs_d, 3, (int)(m.ysz), g->st) != 0 ||
               hipStreamSynchronize(g->st) != hipSuccess) {
        seterr("fdvgpu_put_ref: padding kernel failed");
    } else rc = 0;
    (void)hipFree(tmp); (void)hipFree(srcs_d); (void)hipFree(pl_d);
    return rc;
}

extern "C" int fdvgpu_set_given(fdvgpu *g, int seq, const fdvgpu_mbinfo *dec) {
    if (!seq_ok(g, seq)) return -1;
    SeqMem &m = g->S[seq];
    if (!dec) { m.has_given = 0; return 0; }
    for (int i = 0; i < m.nmb; ++i) {
        const fdvgpu_mbinfo &d = dec[i];
        int nq = d.mode == 3 ? 4 : 1;
        if (d.mode > 3 || d.ref > 1 || d.merge_idx > 2 || d.sub > 3) { seterr("given decision %d out of range", i); return -1; }
        for (int k = 0; k < nq; ++k)
            if (d.mv[k][0] < -FDVG_LIM - 3 || d.mv[k][0] > FDVG_LIM + 3 || d.mv[k][1] < -FDVG_LIM - 3 || d.mv[k][1] > FDVG_LIM + 3) {
                seterr("given vector of MB %d outside the search range", i); return -1;
            }
    }
    CK(hipMemcpy(g->given_d + m.mb0, dec, (size_t)m.nmb * sizeof(fdvgpu_mbinfo), hipMemcpyHostToDevice));
    m.has_given = 1;
    return 0;
}

// The device work of one frame step after its descriptor copies, in four
// stages (motion tables, decisions, deblock + borders, serialisation): each a
// fixed sequence of launches on the context's stream for a given search mode
// and decision structure (par: block-parallel decisions).
static int launch_stage(fdvgpu *g, int stage, int me_mode, bool par) {
    if (stage == 0) {
        if (me_mode == FDVGPU_ME_LEARNED &&
            fdvg_launch_nn_infer(g->seq_d, g->nseq, g->total_mb, g->nnw_d, g->nn_d, g->tune.nn_wpc, g->st) != 0) {
            seterr("proposer launch failed"); return -1;
        }
        if (me_mode != FDVGPU_ME_DIAMOND &&
            fdvg_launch_metab(g->seq_d, g->nseq, g->xgrid, g->metab_d, me_mode == FDVGPU_ME_LEARNED ? g->nn_d : NULL,
                              g->nn_thr, g->nn_sthr, g->st) != 0) {
            seterr("motion table launch failed"); return -1;
        }
        if (fdvg_launch_zero(g->ticket_d, CTL_WORDS, g->st) != 0) { seterr("ticket reset launch failed"); return -1; }
    } else if (stage == 1) {
        if (par && (g->overlap == 2 || g->overlap == 4)) {
            // the costs on this stream, the replays beside them on the second
            // one (they wait for each chunk's cost words), then, after the
            // costs, the coding workers on this stream beside the replays.  The
            // costs go first: where the two streams share a hardware queue the
            // launches simply run one after the other.  The costs never wait;
            // the workers run once the costs are done, and are few enough that
            // every replay still finds a CU of its own.
            const int wgrid = fdvg_pdecide_grid(g->nunit, g->total_mb, -1);
            if (wgrid <= 0) { seterr("parallel analysis: no launch shape"); return -1; }
            CK(hipEventRecord(g->ev_fork, g->st));
            CK(hipStreamWaitEvent(g->st2, g->ev_fork, 0));
            if (fdvg_launch_pcosts(g->seq_d, g->nseq, g->xorder_d, g->xoff_d, g->xmax, g->ticket_d, g->pcw_d,
                                   g->edgew_d, g->st) != 0 ||
                fdvg_launch_pxrep(g->seq_d, g->units_d, g->nunit, g->ticket_d, g->pcw_d, g->edgew_d, g->pgiven_d,
                                  g->pctx_d, g->mrec_d, g->botw_d, g->decw_d, g->st2) != 0 ||
                fdvg_launch_pdecide(g->seq_d, g->nseq, g->xorder_d, g->xoff_d, wgrid, g->ticket_d, g->decw_d,
                                    g->botw_d, g->ih_d, g->recs_d, g->st) != 0) {
                seterr("parallel analysis launch failed: %s", hipGetErrorString(hipGetLastError())); return -1;
            }
            if (g->overlap == 4) {
                // 4: behind the replays on their stream, as many more workers as
                // in 3 join those beside them on the same tickets; they start once
                // every replay is done, so none can keep a replay from a CU
                const int dgrid = fdvg_pdecide_grid(g->nunit, g->total_mb, g->tune.pd_wpc);
                if (dgrid <= 0 ||
                    fdvg_launch_pdecide(g->seq_d, g->nseq, g->xorder_d, g->xoff_d, dgrid, g->ticket_d, g->decw_d,
                                        g->botw_d, g->ih_d, g->recs_d, g->st2) != 0) {
                    seterr("parallel analysis launch failed: %s", hipGetErrorString(hipGetLastError())); return -1;
                }
            }
            CK(hipEventRecord(g->ev_join, g->st2));
            CK(hipStreamWaitEvent(g->st, g->ev_join, 0));
        } else if (par && g->overlap == 3) {
            // the replays beside the costs on the second stream (as in 2), then
            // every coding worker after both (as many as in 1): for batches whose
            // MB count makes the workers' throughput, not the replays' chain, the
            // bound
            const int dgrid = fdvg_pdecide_grid(g->nunit, g->total_mb, g->tune.pd_wpc);
            if (dgrid <= 0) { seterr("parallel analysis: no launch shape"); return -1; }
            CK(hipEventRecord(g->ev_fork, g->st));
            CK(hipStreamWaitEvent(g->st2, g->ev_fork, 0));
            if (fdvg_launch_pcosts(g->seq_d, g->nseq, g->xorder_d, g->xoff_d, g->xmax, g->ticket_d, g->pcw_d,
                                   g->edgew_d, g->st) != 0 ||
                fdvg_launch_pxrep(g->seq_d, g->units_d, g->nunit, g->ticket_d, g->pcw_d, g->edgew_d, g->pgiven_d,
                                  g->pctx_d, g->mrec_d, g->botw_d, g->decw_d, g->st2) != 0) {
                seterr("parallel analysis launch failed: %s", hipGetErrorString(hipGetLastError())); return -1;
            }
            CK(hipEventRecord(g->ev_join, g->st2));
            CK(hipStreamWaitEvent(g->st, g->ev_join, 0));
            if (fdvg_launch_pdecide(g->seq_d, g->nseq, g->xorder_d, g->xoff_d, dgrid, g->ticket_d, g->decw_d,
                                    g->botw_d, g->ih_d, g->recs_d, g->st) != 0) {
                seterr("parallel analysis launch failed: %s", hipGetErrorString(hipGetLastError())); return -1;
            }
        } else if (par && g->overlap == 1) {
            // the costs, the replays, then the coding workers, one stream
            const int dgrid = fdvg_pdecide_grid(g->nunit, g->total_mb, g->tune.pd_wpc);
            if (dgrid <= 0 ||
                fdvg_launch_pcosts(g->seq_d, g->nseq, g->xorder_d, g->xoff_d, g->xmax, g->ticket_d, g->pcw_d,
                                   g->edgew_d, g->st) != 0 ||
                fdvg_launch_pxrep(g->seq_d, g->units_d, g->nunit, g->ticket_d, g->pcw_d, g->edgew_d, g->pgiven_d,
                                  g->pctx_d, g->mrec_d, g->botw_d, g->decw_d, g->st) != 0 ||
                fdvg_launch_pdecide(g->seq_d, g->nseq, g->xorder_d, g->xoff_d, dgrid, g->ticket_d, g->decw_d,
                                    g->botw_d, g->ih_d, g->recs_d, g->st) != 0) {
                seterr("parallel analysis launch failed: %s", hipGetErrorString(hipGetLastError())); return -1;
            }
        } else if (par) {
            if (fdvg_launch_zero(g->ilist_d, 1, g->st) != 0 ||
                fdvg_launch_parallel(g->seq_d, g->nseq, g->xgrid, g->units_d, g->nunit, g->xorder_d, g->xoff_d,
                                     g->xmax, g->ticket_d, g->pcw_d, g->edgew_d, g->pgiven_d, g->pctx_d, g->mrec_d, g->botw_d, g->decw_d,
                                     g->ilist_d, g->ih_d, g->total_mb, g->recs_d, g->st) != 0) {
                seterr("parallel analysis launch failed: %s", hipGetErrorString(hipGetLastError())); return -1;
            }
        } else if (fdvg_launch_analyse(g->seq_d, g->rowtab_d, g->nrows, g->ticket_d, g->recs_d, g->hand_d, g->st) != 0) {
            seterr("analysis launch failed: %s", hipGetErrorString(hipGetLastError())); return -1;
        }
    } else if (stage == 2) {
        if (fdvg_launch_deblock(g->planes_d, 3 * g->nseq, g->maxpx, g->tune.deblock_bs, g->st) != 0 ||
            fdvg_launch_border(g->planes_d, 3 * g->nseq, g->maxpx, g->st) != 0) {
            seterr("deblock launch failed"); return -1;
        }
    } else if (fdvg_launch_symbols(g->sym_d, g->nseq, g->maxmb, g->recs_d, g->cnt_d, g->totals_d, g->bases_d,
                                   g->compact_d, g->compact_cap, g->tune.sym_waves,
                                   g->ser_async ? g->st3 : g->st) != 0) {
        seterr("serialiser launch failed"); return -1;
    }
    return 0;
}

// One stage: eagerly the first time its key is seen (which also fills the
// launchers' one-time caches), from then on as a captured hipGraph.
static int run_stage(fdvgpu *g, int stage, int me_mode, bool par) {
    const int key = graph_key(stage, me_mode, par, g->overlap);
    // the side-by-side stage runs eagerly: a captured graph need not run its
    // two branches at the same time, and run first alone the fused launch
    // would wait for costs that do not come
    if (stage == 1 && par && (g->overlap == 2 || g->overlap == 4)) return launch_stage(g, stage, me_mode, par);
    if (!g->graphs || !g->warm[key]) {
        if (launch_stage(g, stage, me_mode, par) != 0) return -1;
        g->warm[key] = 1;
        return 0;
    }
    // the serialisation of a step queued by fdvgpu_submit_slots runs on st3
    hipStream_t sst = stage == 3 && g->ser_async ? g->st3 : g->st;
    if (!g->gexec[key]) {
        hipGraph_t gr = NULL;
        CK(hipStreamBeginCapture(sst, hipStreamCaptureModeThreadLocal));
        const int rc = launch_stage(g, stage, me_mode, par);
        const hipError_t e = hipStreamEndCapture(sst, &gr);
        if (rc != 0 || e != hipSuccess) {
            if (gr) (void)hipGraphDestroy(gr);
            seterr("frame graph capture failed: %s", rc != 0 ? fdvgpu_last_error() : hipGetErrorString(e));
            return -1;
        }
        const hipError_t ei = hipGraphInstantiate(&g->gexec[key], gr, NULL, NULL, 0);
        (void)hipGraphDestroy(gr);
        if (ei != hipSuccess) { g->gexec[key] = NULL; seterr("frame graph: %s", hipGetErrorString(ei)); return -1; }
    }
    CK(hipGraphLaunch(g->gexec[key], sst));
    return 0;
}

extern "C" int fdvgpu_pframe(fdvgpu *g, const int *src_slot, const int *qp, const double *lscale,
                             const int *navail, const int *force_skip, int me_mode) {
    if (!g || !src_slot || !qp || !lscale || !navail) { seterr("fdvgpu_pframe: bad arguments"); return -1; }
    if (me_mode < FDVGPU_ME_DIAMOND || me_mode > FDVGPU_ME_PARALLEL) { seterr("fdvgpu_pframe: bad me_mode"); return -1; }
    if (me_mode == FDVGPU_ME_LEARNED && !g->have_nn) { seterr("fdvgpu_pframe: learned mode without proposer weights"); return -1; }
    // block-parallel decisions (learned / parallel modes) need neither held
    // frames nor given decisions; those frames take the wavefront, whose
    // table mode they share
    bool par = me_mode == FDVGPU_ME_LEARNED || me_mode == FDVGPU_ME_PARALLEL;
    for (int s = 0; s < g->nseq; ++s)
        if ((force_skip && force_skip[s]) || g->S[s].has_given) par = false;
    for (int s = 0; s < g->nseq; ++s) {
        if (src_slot[s] < 0 || src_slot[s] >= g->nsrc || qp[s] < 0 || qp[s] > 51 ||
            navail[s] < 1 || navail[s] > 2 || !(lscale[s] > 0.0)) {
            seterr("fdvgpu_pframe: bad parameters for sequence %d", s); return -1;
        }
    }
    g->step_ctr = g->step_ctr + 1 ? g->step_ctr + 1 : 1;       // never 0 (the flags' initial value)
    const unsigned step = g->step_ctr;
    hipEvent_t *ev = g->cev;
    CK(hipEventRecord(ev[0], g->st));
    // staging slot for this frame's descriptors: the copy of the frame 8 steps
    // back must have run before the slot is rewritten
    const int slot = g->ring_pos;
    g->ring_pos = (g->ring_pos + 1) % 8;
    CK(hipEventSynchronize(g->ring_ev[slot]));
    fdvg_seq *seq_h = g->seq_h + (size_t)slot * g->nseq;
    PlaneRef *planes_h = g->planes_h + (size_t)slot * 3 * g->nseq;
    for (int s = 0; s < g->nseq; ++s) {
        SeqMem &m = g->S[s];
        fdvg_seq &q = seq_h[s];
        memset(&q, 0, sizeof q);
        const uint8_t *src = m.src + (size_t)src_slot[s] * m.srcsz;
        q.cy = src; q.cu = src + (size_t)m.w * m.h; q.cv = q.cu + (size_t)m.cw * m.ch;
        const int r1 = navail[s] >= 2 ? m.far_ : m.near_;
        q.ry[0] = fy(m, m.near_); q.ru[0] = fu(m, m.near_); q.rv[0] = fv(m, m.near_);
        q.ry[1] = fy(m, r1);      q.ru[1] = fu(m, r1);      q.rv[1] = fv(m, r1);
        q.oy = fy(m, m.out_); q.ou = fu(m, m.out_); q.ov = fv(m, m.out_);
        q.given = m.has_given ? (const fdvg_mbinfo *)(g->given_d + m.mb0) : NULL;
        q.metab = me_mode != FDVGPU_ME_DIAMOND ? (const fdvg_metab *)g->metab_d + m.mb0 : NULL;
        q.info = (fdvg_mbinfo *)(g->info_d + m.mb0);
        // lambda = lscale * 0.425 * 2^((qp-12)/3), in the reference's order (fdv.h:5804)
        q.lambda = lscale[s] * 0.425 * pow(2.0, (qp[s] - 12) / 3.0);
        q.w = m.w; q.h = m.h; q.ys = m.ys; q.cs = m.cs;
        q.navail = navail[s]; q.qp = qp[s];
        q.me_mode = me_mode == FDVGPU_ME_DIAMOND ? FDVG_ME_DIAMOND : FDVG_ME_TABLE;
        q.force_skip = force_skip ? force_skip[s] : 0;
        q.mb0 = m.mb0;
        // The wavefront's handoff words carry the tag of the frame that wrote
        // them, and only that kernel writes them: advancing the tag on its frames
        // alone keeps consecutive writers' tags distinct however many
        // block-parallel frames come between them.
        if (!par) m.tag = m.tag % 255 + 1;
        q.tag = m.tag;
        q.step = step;
        PlaneRef *pl = &planes_h[3 * s];
        pl[0] = PlaneRef{q.oy, m.w, m.h, m.ys, qp[s]};
        pl[1] = PlaneRef{q.ou, m.cw, m.ch, m.cs, qp[s]};
        pl[2] = PlaneRef{q.ov, m.cw, m.ch, m.cs, qp[s]};
    }
    CK(hipMemcpyAsync(g->seq_d, seq_h, g->nseq * sizeof(fdvg_seq), hipMemcpyHostToDevice, g->st));
    CK(hipMemcpyAsync(g->planes_d, planes_h, 3 * g->nseq * sizeof(PlaneRef), hipMemcpyHostToDevice, g->st));
    CK(hipEventRecord(g->ring_ev[slot], g->st));
    // the stages, with the timing events between them (outside any graph); a
    // retry after fdvgpu_rewind_keep_motion keeps the undone frame's motion
    // tables and only clears the counters of stage 0
    const bool reuse = g->keep_tables && me_mode == g->last_mode && me_mode != FDVGPU_ME_DIAMOND;
    g->keep_tables = 0;
    for (int stage = 0; stage < 4; ++stage) {
        if (stage == 3 && g->ser_async) {
            // the serialisation on st3 once the deblock is done, so that the
            // next step's tables run beside it; its MB records stay untouched
            // until then (the next step's decisions wait for ev_ser)
            CK(hipEventRecord(g->ev_db, g->st));
            CK(hipStreamWaitEvent(g->st3, g->ev_db, 0));
            if (run_stage(g, stage, me_mode, par) != 0) return -1;
            CK(hipEventRecord(ev[4], g->st3));
            continue;
        }
        if (stage == 0 && reuse) {
            if (fdvg_launch_zero(g->ticket_d, CTL_WORDS, g->st) != 0) { seterr("ticket reset launch failed"); return -1; }
        } else if (run_stage(g, stage, me_mode, par) != 0) return -1;
        // the decisions rewrite the MB records: an earlier step's serialisation
        // on st3 must be done (counted in the tables' time)
        if (stage == 0 && g->ser_pending) CK(hipStreamWaitEvent(g->st, g->ev_ser, 0));
        if (stage == 1 && g->ser_async)
            CK(hipMemcpyAsync(g->flags_snap_d + 2 * g->ser_slot, g->ticket_d, 2 * sizeof(unsigned),
                              hipMemcpyDeviceToDevice, g->st));
        CK(hipEventRecord(ev[stage + 1], g->st));
    }
    g->last_mode = me_mode;
    for (int s = 0; s < g->nseq; ++s) {
        SeqMem &m = g->S[s];
        int old_far = m.far_;
        m.far_ = m.near_; m.near_ = m.out_; m.out_ = old_far;
        m.has_given = 0;
    }
    g->have_frame = 1;
    g->fetched = 0;
    g->can_rewind = 1;
    return 0;
}

// ---- decoding

extern "C" int fdvgpu_decode(fdvgpu *g, const fdvgpu_mbcode *const *codes, const int *qp, const int *navail) {
    static_assert(sizeof(fdvgpu_mbcode) == sizeof(fdvg_mbrec), "MB code layout");
    if (!g || !codes || !qp || !navail) { seterr("fdvgpu_decode: bad arguments"); return -1; }
    for (int s = 0; s < g->nseq; ++s)
        if (!codes[s] || qp[s] < 0 || qp[s] > 51 || navail[s] < 1 || navail[s] > 2) {
            seterr("fdvgpu_decode: bad parameters for sequence %d", s); return -1;
        }
    fdvgpu_mbinfo *modes = (fdvgpu_mbinfo *)calloc((size_t)g->total_mb, sizeof(fdvgpu_mbinfo));
    if (!modes) { seterr("out of host memory"); return -1; }
    g->step_ctr = g->step_ctr + 1 ? g->step_ctr + 1 : 1;
    const int slot = g->ring_pos;
    g->ring_pos = (g->ring_pos + 1) % 8;
    int rc = -1;
    if (hipEventSynchronize(g->ring_ev[slot]) != hipSuccess) { seterr("fdvgpu_decode: staging wait failed"); goto out; }
    {
        fdvg_seq *seq_h = g->seq_h + (size_t)slot * g->nseq;
        PlaneRef *planes_h = g->planes_h + (size_t)slot * 3 * g->nseq;
        for (int s = 0; s < g->nseq; ++s) {
            SeqMem &m = g->S[s];
            fdvg_seq &q = seq_h[s];
            memset(&q, 0, sizeof q);
            const int r1 = navail[s] >= 2 ? m.far_ : m.near_;
            q.ry[0] = fy(m, m.near_); q.ru[0] = fu(m, m.near_); q.rv[0] = fv(m, m.near_);
            q.ry[1] = fy(m, r1);      q.ru[1] = fu(m, r1);      q.rv[1] = fv(m, r1);
            q.oy = fy(m, m.out_); q.ou = fu(m, m.out_); q.ov = fv(m, m.out_);
            q.w = m.w; q.h = m.h; q.ys = m.ys; q.cs = m.cs;
            q.navail = navail[s]; q.qp = qp[s];
            q.mb0 = m.mb0;
            q.step = g->step_ctr;
            planes_h[3 * s + 0] = PlaneRef{q.oy, m.w, m.h, m.ys, qp[s]};
            planes_h[3 * s + 1] = PlaneRef{q.ou, m.cw, m.ch, m.cs, qp[s]};
            planes_h[3 * s + 2] = PlaneRef{q.ov, m.cw, m.ch, m.cs, qp[s]};
            for (int i = 0; i < m.nmb; ++i) modes[m.mb0 + i].mode = codes[s][i].mode;
            if (hipMemcpyAsync(g->recs_d + m.mb0, codes[s], (size_t)m.nmb * sizeof(fdvg_mbrec), hipMemcpyHostToDevice,
                               g->st) != hipSuccess) { seterr("fdvgpu_decode: upload failed"); goto out; }
        }
        if (hipMemcpyAsync(g->seq_d, seq_h, g->nseq * sizeof(fdvg_seq), hipMemcpyHostToDevice, g->st) != hipSuccess ||
            hipMemcpyAsync(g->planes_d, planes_h, 3 * g->nseq * sizeof(PlaneRef), hipMemcpyHostToDevice, g->st) !=
                hipSuccess ||
            hipMemcpyAsync(g->pgiven_d, modes, (size_t)g->total_mb * sizeof(fdvgpu_mbinfo), hipMemcpyHostToDevice,
                           g->st) != hipSuccess ||
            hipEventRecord(g->ring_ev[slot], g->st) != hipSuccess) {
            seterr("fdvgpu_decode: upload failed"); goto out;
        }
    }
    if (fdvg_launch_zero(g->ticket_d, CTL_WORDS, g->st) != 0 ||
        fdvg_launch_decode(g->seq_d, g->nseq, g->xgrid, (const fdvg_mbinfo *)g->pgiven_d, g->total_mb, g->recs_d,
                           g->ticket_d, g->ilist_d, g->ih_d, g->st) != 0) {
        seterr("fdvgpu_decode: launch failed: %s", hipGetErrorString(hipGetLastError())); goto out;
    }
    if (launch_stage(g, 2, FDVGPU_ME_DIAMOND, false) != 0) goto out;
    // the decoded frames are read right after: wait here, and check the
    // intra pass's bounded waits
    if (hipMemcpyAsync(g->flags_h, g->ticket_d, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, g->st) != hipSuccess ||
        hipStreamSynchronize(g->st) != hipSuccess) {
        seterr("fdvgpu_decode: %s", hipGetErrorString(hipGetLastError())); goto out;
    }
    if (g->flags_h[1]) { seterr("fdvgpu_decode: an intra MB waited too long for its neighbours"); goto out; }
    for (int s = 0; s < g->nseq; ++s) {
        SeqMem &m = g->S[s];
        const int old_far = m.far_;
        m.far_ = m.near_; m.near_ = m.out_; m.out_ = old_far;
    }
    g->have_frame = 0;
    g->can_rewind = 0;
    rc = 0;
out:
    free(modes);
    return rc;
}

// ---- pipelined frame steps

static int pipe_alloc(fdvgpu *g) {
    if (g->pipe_ok) return 0;
    if (g->pipe_nev || g->pipe_up[0]) { seterr("pipelined buffers: an earlier allocation failed"); return -1; }
    size_t up = 0;
    for (int s = 0; s < g->nseq; ++s) up += (size_t)g->S[s].w * g->S[s].h + 2 * (size_t)g->S[s].cw * g->S[s].ch;
    const size_t nb = ((size_t)g->nseq * FDVG_NS + 3) * sizeof(uint32_t);
    for (int k = 0; k < 2; ++k) {
        CK(hipHostMalloc(&g->pipe_up[k], up, hipHostMallocDefault));
        CK(hipHostMalloc(&g->pipe_sc[k], g->compact_cap, hipHostMallocCoherent | hipHostMallocMapped));
        CK(hipHostMalloc(&g->pipe_sb[k], nb, hipHostMallocCoherent | hipHostMallocMapped));
        CK(hipHostGetDevicePointer((void **)&g->pipe_sc_d[k], g->pipe_sc[k], 0));
        CK(hipHostGetDevicePointer((void **)&g->pipe_sb_d[k], g->pipe_sb[k], 0));
        CK(hipEventCreateWithFlags(&g->pipe_ev[k], hipEventDisableTiming));
        g->pipe_nev = k + 1;
    }
    g->pipe_ok = 1;
    return 0;
}

extern "C" int fdvgpu_submit(fdvgpu *g, int slot, const uint8_t *const *frames, const int *qp, const double *lscale,
                             const int *navail, const int *force_skip, int me_mode) {
    if (!g || slot < 0 || slot > 1 || !frames) { seterr("fdvgpu_submit: bad arguments"); return -1; }
    for (int s = 0; s < g->nseq; ++s) if (!frames[s]) { seterr("fdvgpu_submit: no frame for sequence %d", s); return -1; }
    if (pipe_alloc(g) != 0) return -1;
    // the slot's previous use (its upload and its staging) is complete
    CK(hipEventSynchronize(g->pipe_ev[slot]));
    size_t off = 0;
    for (int s = 0; s < g->nseq; ++s) {
        SeqMem &m = g->S[s];
        const size_t fsz = (size_t)m.w * m.h + 2 * (size_t)m.cw * m.ch;
        memcpy(g->pipe_up[slot] + off, frames[s], fsz);
        CK(hipMemcpyAsync(m.src, g->pipe_up[slot] + off, fsz, hipMemcpyHostToDevice, g->st));
        off += fsz;
    }
    int *slots = (int *)calloc((size_t)g->nseq, sizeof(int));
    if (!slots) { seterr("fdvgpu_submit: out of memory"); return -1; }
    const int rc = fdvgpu_pframe(g, slots, qp, lscale, navail, force_skip, me_mode);
    free(slots);
    if (rc != 0) return -1;
    g->can_rewind = 0;
    if (fdvg_launch_stage(g->bases_d, g->ticket_d, g->nseq * FDVG_NS, g->compact_d, g->compact_cap,
                          g->pipe_sb_d[slot], g->pipe_sc_d[slot], g->st) != 0) {
        seterr("fdvgpu_submit: staging launch failed: %s", hipGetErrorString(hipGetLastError())); return -1;
    }
    CK(hipEventRecord(g->pipe_ev[slot], g->st));
    return 0;
}

extern "C" int fdvgpu_collect(fdvgpu *g, int slot) {
    if (!g || slot < 0 || slot > 1 || !g->pipe_ok) { seterr("fdvgpu_collect: nothing submitted"); return -1; }
    CK(hipEventSynchronize(g->pipe_ev[slot]));
    const int n = g->nseq * FDVG_NS;
    if (g->pipe_sb[slot][n + 1]) { seterr("MB wavefront: a row waited too long for the row above"); return -1; }
    if (g->pipe_sb[slot][n + 2]) { seterr("symbol streams overflowed (%u bytes)", g->pipe_sb[slot][n]); return -1; }
    return 0;
}

// A frame step from sources already in the context's slots, queued as
// fdvgpu_submit queues an uploaded one; its serialisation and the staging of
// its streams run on a third stream beside the next step's motion tables.
extern "C" int fdvgpu_submit_slots(fdvgpu *g, int slot, const int *src_slot, const int *qp, const double *lscale,
                                   const int *navail, const int *force_skip, int me_mode) {
    if (!g || slot < 0 || slot > 1 || !src_slot) { seterr("fdvgpu_submit_slots: bad arguments"); return -1; }
    if (pipe_alloc(g) != 0) return -1;
    // the slot's previous staging (and with it its timing events) is complete
    CK(hipEventSynchronize(g->pipe_ev[slot]));
    g->ser_async = 1; g->ser_slot = slot; g->cev = g->pev[slot];
    const int rc = fdvgpu_pframe(g, src_slot, qp, lscale, navail, force_skip, me_mode);
    g->ser_async = 0; g->cev = g->ev;
    if (rc != 0) return -1;
    g->can_rewind = 0;
    if (fdvg_launch_stage(g->bases_d, g->flags_snap_d + 2 * slot, g->nseq * FDVG_NS, g->compact_d, g->compact_cap,
                          g->pipe_sb_d[slot], g->pipe_sc_d[slot], g->st3) != 0) {
        seterr("fdvgpu_submit_slots: staging launch failed: %s", hipGetErrorString(hipGetLastError())); return -1;
    }
    CK(hipEventRecord(g->pipe_ev[slot], g->st3));
    CK(hipEventRecord(g->ev_ser, g->st3));
    g->ser_pending = 1;
    return 0;
}

extern "C" int fdvgpu_slot_timings(fdvgpu *g, int slot, float ms[4]) {
    if (!g || slot < 0 || slot > 1 || !ms) { seterr("fdvgpu_slot_timings: bad arguments"); return -1; }
    CK(hipEventSynchronize(g->pev[slot][4]));
    for (int i = 0; i < 4; ++i) CK(hipEventElapsedTime(&ms[i], g->pev[slot][i], g->pev[slot][i + 1]));
    return 0;
}

extern "C" int fdvgpu_collected_streams(fdvgpu *g, int slot, int seq, const uint8_t **sv, size_t *sn) {
    if (!seq_ok(g, seq) || slot < 0 || slot > 1 || !g->pipe_ok || !sv || !sn) return -1;
    for (int k = 0; k < FDVG_NS; ++k) {
        const uint32_t a = g->pipe_sb[slot][seq * FDVG_NS + k], b = g->pipe_sb[slot][seq * FDVG_NS + k + 1];
        sv[k] = g->pipe_sc[slot] + a;
        sn[k] = b - a;
    }
    return 0;
}

extern "C" int fdvgpu_rewind_keep_motion(fdvgpu *g) {
    if (fdvgpu_rewind(g) != 0) return -1;
    g->keep_tables = 1;
    return 0;
}

extern "C" int fdvgpu_rewind(fdvgpu *g) {
    if (!g || !g->can_rewind) { seterr("fdvgpu_rewind: no P-frame to undo"); return -1; }
    // inverse of the rotation in fdvgpu_pframe: the pre-frame far reference still
    // sits untouched in the output slot until the next frame writes it
    for (int s = 0; s < g->nseq; ++s) {
        SeqMem &m = g->S[s];
        int out = m.near_;
        m.near_ = m.far_; m.far_ = m.out_; m.out_ = out;
    }
    g->can_rewind = 0;
    return 0;
}

extern "C" int fdvgpu_fetch(fdvgpu *g) {
    if (!g || !g->have_frame) { seterr("fdvgpu_fetch: no frame analysed"); return -1; }
    if (g->fetched) return 0;
    const int n = g->nseq * FDVG_NS;
    // a step queued by fdvgpu_submit_slots serialises on st3
    if (g->ser_pending) CK(hipStreamWaitEvent(g->st, g->ev_ser, 0));
    CK(hipMemcpyAsync(g->bases_h, g->bases_d, (n + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, g->st));
    CK(hipMemcpyAsync(g->flags_h, g->ticket_d, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, g->st));
    CK(hipStreamSynchronize(g->st));
    if (g->flags_h[1]) { seterr("MB wavefront: a row waited too long for the row above"); return -1; }
    size_t total = g->bases_h[n];
    if (total > g->compact_cap) { seterr("symbol streams overflowed (%zu bytes)", total); return -1; }
    if (total) {
        CK(hipMemcpyAsync(g->compact_h, g->compact_d, total, hipMemcpyDeviceToHost, g->st));
        CK(hipStreamSynchronize(g->st));
    }
    g->fetched = 1;
    return 0;
}

extern "C" int fdvgpu_streams(fdvgpu *g, int seq, const uint8_t **sv, size_t *sn) {
    if (!seq_ok(g, seq) || !sv || !sn) return -1;
    if (fdvgpu_fetch(g) != 0) return -1;
    for (int k = 0; k < FDVG_NS; ++k) {
        uint32_t a = g->bases_h[seq * FDVG_NS + k], b = g->bases_h[seq * FDVG_NS + k + 1];
        sv[k] = g->compact_h + a;
        sn[k] = b - a;
    }
    return 0;
}

extern "C" int fdvgpu_recon(fdvgpu *g, int seq, uint8_t *i420) {
    if (!seq_ok(g, seq) || !i420) return -1;
    SeqMem &m = g->S[seq];
    CK(hipMemcpy2DAsync(i420, m.w, fy(m, m.near_), m.ys, m.w, m.h, hipMemcpyDeviceToHost, g->st));
    uint8_t *u = i420 + (size_t)m.w * m.h, *v = u + (size_t)m.cw * m.ch;
    CK(hipMemcpy2DAsync(u, m.cw, fu(m, m.near_), m.cs, m.cw, m.ch, hipMemcpyDeviceToHost, g->st));
    CK(hipMemcpy2DAsync(v, m.cw, fv(m, m.near_), m.cs, m.cw, m.ch, hipMemcpyDeviceToHost, g->st));
    CK(hipStreamSynchronize(g->st));
    return 0;
}

extern "C" int fdvgpu_decisions(fdvgpu *g, int seq, fdvgpu_mbinfo *out) {
    if (!seq_ok(g, seq) || !out) return -1;
    SeqMem &m = g->S[seq];
    CK(hipMemcpyAsync(out, g->info_d + m.mb0, (size_t)m.nmb * sizeof(fdvgpu_mbinfo), hipMemcpyDeviceToHost, g->st));
    CK(hipStreamSynchronize(g->st));
    return 0;
}

extern "C" int fdvgpu_timings(fdvgpu *g, float ms[4]) {
    if (!g || !ms || !g->have_frame) return -1;
    CK(hipEventSynchronize(g->ev[4]));
    for (int i = 0; i < 4; ++i) CK(hipEventElapsedTime(&ms[i], g->ev[i], g->ev[i + 1]));
    return 0;
}


extern "C" int fdvgpu_set_proposer(fdvgpu *g, const float *w, int n) {
    if (!g || !w || n != fdvg_nn_nparam()) { seterr("fdvgpu_set_proposer: expected %d weights", fdvg_nn_nparam()); return -1; }
    CK(hipMemcpyAsync(g->nnw_d, w, (size_t)n * sizeof(float), hipMemcpyHostToDevice, g->st));
    CK(hipStreamSynchronize(g->st));
    g->have_nn = 1;
    return 0;
}

static void drop_graphs(fdvgpu *g);

extern "C" int fdvgpu_set_proposer_thresholds(fdvgpu *g, float conf, float skip) {
    if (!g || !(conf >= 0.f) || !(skip >= 0.f)) { seterr("fdvgpu_set_proposer_thresholds: bad arguments"); return -1; }
    g->nn_thr = conf;
    g->nn_sthr = skip;
    drop_graphs(g);                 // the thresholds are launch arguments of captured graphs
    return 0;
}

extern "C" int fdvgpu_proposals(fdvgpu *g, int seq, fdvgpu_proposal *out) {
    if (!seq_ok(g, seq) || !out) { seterr("fdvgpu_proposals: bad arguments"); return -1; }
    static_assert(sizeof(fdvgpu_proposal) == sizeof(fdvg_nnout), "proposal layout");
    CK(hipMemcpyAsync(out, g->nn_d + g->S[seq].mb0, (size_t)g->S[seq].nmb * sizeof(fdvg_nnout),
                      hipMemcpyDeviceToHost, g->st));
    CK(hipStreamSynchronize(g->st));
    return 0;
}

extern "C" int fdvgpu_motion(fdvgpu *g, int seq, fdvgpu_motion_entry *out) {
    if (!seq_ok(g, seq) || !out) { seterr("fdvgpu_motion: bad arguments"); return -1; }
    static_assert(sizeof(fdvgpu_motion_entry) == sizeof(fdvg_metab), "motion table layout");
    CK(hipMemcpyAsync(out, (const fdvg_metab *)g->metab_d + g->S[seq].mb0, (size_t)g->S[seq].nmb * sizeof(fdvg_metab),
                      hipMemcpyDeviceToHost, g->st));
    CK(hipStreamSynchronize(g->st));
    return 0;
}

extern "C" int fdvgpu_planes(fdvgpu *g, int seq, int src_slot, const uint8_t **cur_y, int *cur_stride,
                             const uint8_t **ref_y, int *ref_stride) {
    if (!seq_ok(g, seq) || src_slot < 0 || src_slot >= g->nsrc || !cur_y || !cur_stride || !ref_y || !ref_stride) {
        seterr("fdvgpu_planes: bad arguments"); return -1;
    }
    SeqMem &m = g->S[seq];
    *cur_y = m.src + (size_t)src_slot * m.srcsz; *cur_stride = m.w;
    *ref_y = fy(m, m.near_); *ref_stride = m.ys;
    return 0;
}

extern "C" int fdvgpu_set_overlap(fdvgpu *g, int on) {
    if (!g || on < 0 || on > 4) { seterr("fdvgpu_set_overlap: bad arguments"); return -1; }
    if ((on == 2 || on == 4) && !g->overlap_avail) {
        seterr("fdvgpu_set_overlap: the replays and the coding workers do not fit side by side"); return -1;
    }
    g->overlap = on;
    g->overlap_fixed = 1;
    return 0;
}

extern "C" int fdvgpu_get_overlap(fdvgpu *g) {
    if (!g) { seterr("fdvgpu_get_overlap: bad context"); return -1; }
    return g->overlap;
}

extern "C" int fdvgpu_set_graphs(fdvgpu *g, int on) {
    if (!g) { seterr("fdvgpu_set_graphs: bad context"); return -1; }
    g->graphs = on ? 1 : 0;
    return 0;
}

static void drop_graphs(fdvgpu *g) {
    for (int k = 0; k < NGRAPH; ++k)
        if (g->gexec[k]) { (void)hipGraphExecDestroy(g->gexec[k]); g->gexec[k] = NULL; }
}

extern "C" int fdvgpu_get_tune(fdvgpu *g, int out[4]) {
    if (!g || !out) { seterr("fdvgpu_get_tune: bad arguments"); return -1; }
    out[0] = g->tune.sym_waves; out[1] = g->tune.deblock_bs; out[2] = g->tune.nn_wpc; out[3] = g->tune.pd_wpc;
    return 0;
}

extern "C" int fdvgpu_set_tune(fdvgpu *g, const int in[4]) {
    if (!g || !in) { seterr("fdvgpu_set_tune: bad arguments"); return -1; }
    if (in[0] <= 0 || in[0] > 8 || (in[1] != 64 && in[1] != 128 && in[1] != 256) || in[2] < 1 || in[2] > 2 ||
        in[3] < 1 || in[3] > 2) {
        seterr("fdvgpu_set_tune: shapes out of range"); return -1;
    }
    g->tune = Tune{in[0], in[1], in[2], in[3]};
    drop_graphs(g);
    return 0;
}

// Each candidate launch shape times `reps` frame steps of the given frame (the
// reference pool is rewound after each), one parameter at a time, keeping the
// fastest; the chosen shapes, and the decision structure where two are
// available and none was set explicitly (fdvgpu_set_overlap), stay in the
// context.
extern "C" int fdvgpu_autotune(fdvgpu *g, const int *src_slot, const int *qp, const double *lscale,
                               const int *navail, int me_mode, int reps) {
    if (!g || reps <= 0) { seterr("fdvgpu_autotune: bad arguments"); return -1; }
    // parameter 4, tried first: the decision structure (2: the coding workers
    // beside the replays, as many as leave every replay a CU, where that is
    // available; 3: every worker after the replays, which run beside the costs;
    // 4: 2 joined by 3's workers once the replays are done)
    static const int opts[5][3] = {{2, 4, 8}, {64, 128, 256}, {1, 2, 0}, {1, 2, 0}, {2, 3, 4}};
    static const int order[5] = {4, 0, 1, 2, 3};
    const bool par = me_mode == FDVGPU_ME_LEARNED || me_mode == FDVGPU_ME_PARALLEL;
    for (int pi = 0; pi < 5; ++pi) {
        const int prm = order[pi];
        if (prm == 2 && me_mode != FDVGPU_ME_LEARNED) continue;
        if (prm == 3 && !(par && (g->overlap == 1 || g->overlap == 3 || g->overlap == 4))) continue;
        if (prm == 4 && !(par && !g->overlap_fixed)) continue;
        float best = 1e30f;
        int bestv = 0;
        for (int k = 0; k < 3 && opts[prm][k]; ++k) {
            Tune t = g->tune;
            if (prm == 0) t.sym_waves = opts[0][k];
            else if (prm == 1) t.deblock_bs = opts[1][k];
            else if (prm == 2) t.nn_wpc = opts[2][k];
            else if (prm == 3) t.pd_wpc = opts[3][k];
            else if ((opts[4][k] == 2 || opts[4][k] == 4) && !g->overlap_avail) continue;
            else g->overlap = opts[4][k];
            g->tune = t;
            drop_graphs(g);
            float tot = 0.f;
            for (int r = 0; r <= reps; ++r) {           // the first run warms up (and captures)
                if (fdvgpu_pframe(g, src_slot, qp, lscale, navail, NULL, me_mode) != 0) return -1;
                CK(hipEventSynchronize(g->ev[4]));
                float ms = 0.f;
                CK(hipEventElapsedTime(&ms, g->ev[0], g->ev[4]));
                if (r) tot += ms;
                if (fdvgpu_rewind(g) != 0) return -1;
            }
            if (tot < best) { best = tot; bestv = opts[prm][k]; }
        }
        if (prm == 0) g->tune.sym_waves = bestv;
        else if (prm == 1) g->tune.deblock_bs = bestv;
        else if (prm == 2) g->tune.nn_wpc = bestv;
        else if (prm == 3) g->tune.pd_wpc = bestv;
        else g->overlap = bestv;
        drop_graphs(g);
    }
    g->have_frame = 0;
    return 0;
}
