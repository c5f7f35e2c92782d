/*
 * dev_types.h — device-side data layout shared by the host planner and the
 * HIP kernels.
 *
 * HBM layout
 *   Frame buffer: one allocation per frame, planes back to back:
 *     Y  [sy  x H ]   sy  = W rounded up to 64 bytes
 *     U  [suv x ch]   suv = ceil(W/2) rounded up to 64 bytes, ch = ceil(H/2)
 *     V  [suv x ch]
 *   Every row starts 64-byte aligned, so 4/16-byte vector accesses at
 *   4-aligned x never straddle alignment.
 *
 *   Prepared frame ("prep"): the parsed frame staged once into HBM:
 *     DBlock[nblk]   24-byte block records (clamped, normalised modes)
 *     mc_tab[]       motion-compensation units: block << 8 | unit in block
 *     intra[]        local indices of intra blocks, by dependency level;
 *                    inside a level small blocks (at most 16x16) first,
 *                    each group by the length of its chain of later
 *                    dependents, longest first
 *     tx_tab[c][]    transforms of size class c (inter and intra blocks, list
 *                    order), 8 bytes each: coefficient offset; x | y << 14,
 *                    bit 31 set for an intra block
 *     coeffs[]       the parser's int16 coefficients, verbatim
 *
 *   Job: one frame of one stream in a batch: output frame, the 8 reference
 *   slots, the prep, and per-stream scratch (owner map, intra edge and
 *   neighbour records, intra residuals).
 *   A batch is an array of jobs plus per-stage prefix sums of work items, so
 *   every stage is one launch over all jobs of the batch.
 */
#ifndef CVP9_DEV_TYPES_H
#define CVP9_DEV_TYPES_H

#include <stdint.h>

#define CVP9_CELL 4                 /* owner-map granularity in luma px */
#define CVP9_OWNER_NONE 0xffffffffu
#define CVP9_MAX_BATCH 4096
#define CVP9_MAX_HEAD_LEVELS 64     /* leading intra levels launched one by one */

/* Edge record of an intra block: its final bottom row and right column of
 * Y, U and V, 4 samples per 64-bit word with the frame epoch in the upper 32
 * bits, so each word carries its own validity. Word ranges:
 *   Y bottom [0,16)  Y right [16,32)  U bottom [32,40)  U right [40,48)
 *   V bottom [48,56) V right [56,64)   (intra blocks are at most 64x64) */
#define CVP9_EDGE_WORDS 64
#define CVP9_MAX_INTRA_DIM 64

/* Neighbour record of an intra block (one per entry of the intra list),
 * written by the neighbour stage before the ordered intra stage. Cells are
 * the 4x4 luma cells next to the block: above row (ceil(bw/4) cells), left
 * column (ceil(bh/4)), then the above-left cell; one 64-bit word each, then
 * a mask with bit c set when cell c belongs to an earlier intra block.
 *   resolved cell: bytes 0-3 luma, 4-5 U, 6-7 V samples of the cell on the
 *     block's side (above: left to right; left: top to bottom; above-left:
 *     the single samples in bytes 3, 5, 7)
 *   intra cell: owner block index (bits 0-31), edge side (bit 32: 0 bottom
 *     row, 1 right column), luma edge word (bits 33-36), index of the
 *     cell's first chroma sample on the owner's chroma edge (bits 40-44)
 * followed by the block's local index and a copy of its DBlock, so the
 * ordered stage reads everything it needs about a block in one round. */
#define CVP9_NB_CELLS 33
#define CVP9_NB_MASK  33   /* cells of earlier intra blocks (edge-record pointers) */
#define CVP9_NB_IMASK 34   /* cells of inter blocks (samples read from the frame) */
#define CVP9_NB_INDEX 35
#define CVP9_NB_BLOCK 36   /* 3 words */
#define CVP9_NB_WORDS 39

/* Motion compensation splits an inter block into units of
 * cvp9_mc_unit_rows(bw) luma rows, four units per wave. A unit is 16 lanes;
 * each lane takes one column of 4-sample quads over a run of up to
 * CVP9_MC_RUN rows (luma, then chroma), so a reference row is read once
 * for two output rows. Blocks are at most 64x64: an 8x8 or a 16x16 block is
 * one unit, 32x32 four, 64x64 sixteen. */
#define CVP9_MC_UNIT 16
#define CVP9_MC_RUN  4
/* Units per workgroup of the motion-compensation kernel. Each job's unit
 * count is padded to a multiple of it in the batch prefix, so a workgroup
 * works for one job. */
#define CVP9_MC_WG_UNITS 16

/* Grid barrier of the fused kernel: words from DBatch::bar, zeroed by the
 * descriptor copy before every launch. The workgroups on one XCD form a
 * group (the XCD's id read from the hardware, at most CVP9_BAR_GROUPS):
 * each registers in its group's count at start; at a barrier each arrives
 * on its group's counter, the last of a group arrives on the global
 * counter and, once every group is in, releases its group through the
 * group's own word. Words that are polled sit 128 bytes apart. */
#define CVP9_BAR_GROUPS  16
#define CVP9_BAR_GLOBAL  0      /* groups arrived, all phases */
#define CVP9_BAR_STARTED 32     /* workgroups registered */
#define CVP9_BAR_COUNT   64     /* CVP9_BAR_GROUPS words: workgroups per group */
#define CVP9_BAR_ARRIVE  96     /* + 32 * group: arrivals, all phases */
#define CVP9_BAR_RELEASE (CVP9_BAR_ARRIVE + 32 * CVP9_BAR_GROUPS) /* + 32 * group: last phase released */
#define CVP9_BAR_BYTES   4608   /* >= 4 * (CVP9_BAR_RELEASE + 32 * CVP9_BAR_GROUPS), 256-aligned */
#if defined(__HIPCC__)
#define CVP9_HD __host__ __device__
#else
#define CVP9_HD
#endif
CVP9_HD static inline uint32_t cvp9_mc_unit_rows(uint32_t bw)
{
    const uint32_t nq = (bw + 3) / 4;
    uint32_t p2 = 1;
    while (p2 < nq) p2 <<= 1;
    return p2 >= CVP9_MC_UNIT ? CVP9_MC_RUN : CVP9_MC_RUN * CVP9_MC_UNIT / p2;
}

/* The neighbours an intra block's prediction reads (vp9_reconstruct.c:
 * 63-91): DC the above row and the left column, V the above row, H the left
 * column, TM both and the above-left sample; a block reads the union over
 * its luma and chroma modes (mode classes 0 DC, 1 V, 2 H, 3 TM). A cell the
 * prediction does not read is no dependency: the planner orders blocks by
 * the cells they read, and the neighbour stage makes only those wait. */
#define CVP9_READS_ABOVE  1u
#define CVP9_READS_LEFT   2u
#define CVP9_READS_CORNER 4u
CVP9_HD static inline uint32_t cvp9_mode_reads(uint32_t ymode, uint32_t uvmode)
{
    const uint32_t r[4] = {CVP9_READS_ABOVE | CVP9_READS_LEFT, CVP9_READS_ABOVE, CVP9_READS_LEFT,
                           CVP9_READS_ABOVE | CVP9_READS_LEFT | CVP9_READS_CORNER};
    return r[ymode & 3] | r[uvmode & 3];
}

/* Compact coefficients: chunks of this many pool entries, one 64-bit
 * nonzero mask each. */
#define CVP9_COEF_CHUNK 64

/* block flags */
#define DB_INTRA   1u
#define DB_SKIP    2u   /* no residual (skip, or transform grid empty) */
#define DB_NOREF   4u   /* inter block whose reference slot is empty */

struct DBlock {          /* 24 bytes */
    uint16_t x, y;       /* luma position */
    uint8_t  bw, bh;     /* clamped to the frame, > 0 */
    uint8_t  tx;         /* transform size class 0..3 (4 << tx px) */
    uint8_t  flags;
    uint8_t  ymode, uvmode; /* 0 DC, 1 V, 2 H, 3 TM */
    uint8_t  ref;        /* reference slot 0..7 */
    uint8_t  pad;
    int16_t  mvx, mvy;   /* 1/8 pel */
    uint32_t coeff;      /* element offset of the first transform */
    uint32_t nx;         /* transforms per row (clamped width / T) */
};

#ifdef __cplusplus
static_assert(sizeof(DBlock) == 24, "DBlock is copied as three 64-bit words");
#endif

struct DFrame {          /* 40 bytes */
    uint8_t *y, *u, *v;
    uint32_t w, h;       /* luma size */
    uint32_t sy, suv;    /* plane strides */
};

struct DJob {
    DFrame out;
    DFrame ref[8];                 /* y == nullptr: empty slot */
    const DBlock   *blocks;
    const int16_t  *coeffs;
    const uint4    *mc_tab;        /* per MC unit, a copy of the block's MC fields:
                                      x | y << 16,
                                      bw | bh << 8 | (ref | unit << 4) << 16 |
                                        (tx | skip << 2 | nx << 3) << 24,
                                      (uint16)mvx | (uint16)mvy << 16,
                                      coefficient offset of the block's first transform */
    const uint32_t *intra;
    const uint32_t *tx_tab[4];
    uint32_t       *owner;         /* ow x oh cells, local block index */
    uint64_t       *edge;          /* per local block: CVP9_EDGE_WORDS edge words */
    uint64_t       *nbr;           /* per intra-list entry: CVP9_NB_WORDS words */
    int16_t        *resid;         /* intra residuals, indexed like coeffs */
    uint8_t        *nv12;          /* optional NV12 target, pitch W */
    uint32_t       *err;           /* error word (device memory, polled by waits): byte 0
                                      a dependency wait gave up, byte 1 a neighbour record
                                      was inconsistent, byte 2 a grid barrier timed out */
    uint32_t       *err_host;      /* the same flags in host-visible memory, set with
                                      them, so the host reads them without a copy */
    uint32_t       *ectr;          /* the stream's epoch counter (device) */
    /* compact coefficients (n_chunks > 0): per chunk of 64 pool entries its
     * nonzero mask and the position of its first value among cvals; the
     * expansion stage writes the dense pool (coeffs, n_chunks * 64 entries) */
    const uint64_t *cmask;
    const uint32_t *cstart;
    const int16_t  *cvals;
    uint32_t        n_chunks;
    /* chains of intra blocks for chain workers (n_chains > 0): n_chains + 1
     * offsets, the chained intra-list entries chain by chain, then every
     * other entry in list order */
    const uint32_t *chain;
    uint32_t        n_chains;
    /* intra-list entries for the deferred sample pass: the n_px_small blocks
     * of at most 8x8 first, then the others */
    const uint32_t *px;
    uint32_t        n_px_small;
    uint32_t n_blocks, n_inter, n_intra, n_mc;
    uint32_t n_mc_a;               /* motion-compensation units [0, n_mc_a) run in the
                                      early launch, the rest in the late one */
    uint32_t cls_ntx[4];
    uint32_t ow, oh;
    uint32_t epoch;                /* written by the batch's first kernel: *ectr + 1 */
    int32_t  lf_thr;               /* 0: no deblocking */
    uint32_t lf_wx, lf_wy;         /* deblocking: waves of 64 units (8x8 at 8k-2, 8m-4) per
                                      band of 8 rows, and bands */
    uint32_t fill;                 /* 1: frame not fully covered by blocks */
};

/* Work-item counts per job, as exclusive prefixes (length njobs + 1). */
struct DBatch {
    const DJob     *jobs;
    uint32_t        njobs;
    const uint32_t *pre_blocks;    /* owner-map scatter: one wave per block */
    const uint32_t *pre_inter;     /* motion compensation, early units (per job padded to
                                      whole workgroups) */
    const uint32_t *pre_inter2;    /* motion compensation, late units */
    const uint32_t *pre_grp[4];    /* residual: groups of 64/N transforms */
    const uint32_t *pre_intra;     /* intra blocks per job */
    const uint32_t *pre_lf;        /* deblocking: waves of 64 units */
    const uint32_t *pre_nv12;      /* NV12 pack: 64x64 luma tiles */
    const uint32_t *pre_fill;      /* 128-fill: 64 KiB chunks */
    const uint32_t *pre_exp;       /* compact coefficients: waves of 4 chunks */
    const uint32_t *pre_px;        /* deferred intra samples: waves (four blocks of at most
                                      8x8 to a wave, one larger block) */
    uint32_t       *jticket;       /* per-job intra ticket counters (zeroed by the descriptor
                                      copy), for jobs no persistent worker starts on */
    const uint32_t *jstart;        /* per job, the intra-list entry where the persistent
                                      workers start (after the level launches) */
    uint32_t       *bar;           /* grid-barrier counter of the fused kernel (zeroed likewise) */
    /* leading levels of the intra stage, per dependency level l (rows of
     * njobs + 1): exclusive prefixes over jobs of the small (at most 16x16)
     * and of the large blocks on level l, and each job's intra-list
     * position where level l starts (its small blocks come first) */
    const uint32_t *lvl_pre;
    const uint32_t *lvl_pre2;
    const uint32_t *lvl_start;
    uint32_t        spin_limit;    /* bound on every dependency wait */
    uint32_t        resid_all;     /* 1: every transform's residual goes to the int16
                                      residual buffer and motion compensation adds the
                                      inter ones (the transform stage then does not depend
                                      on the previous frame); 0: inter transforms add to
                                      the prediction in place */
    uint32_t        fused_lanes;   /* fused kernel: threads per intra worker (64 or 256) */
    uint32_t        fused_workers; /* fused kernel: persistent intra workers */
};

#endif
