/* kernels.h — host-side interface of the reconstruction kernels. */
#ifndef CVP9_KERNELS_H
#define CVP9_KERNELS_H

#include <hip/hip_runtime.h>
#include "dev_types.h"

enum Stage {
    STAGE_FILL = 0,
    STAGE_OWNER,
    STAGE_MC,
    STAGE_RES4,
    STAGE_RES8,
    STAGE_RES16,
    STAGE_RES32,
    STAGE_NBR,   /* neighbour records of intra blocks */
    STAGE_INTRA,
    STAGE_LF,
    STAGE_NV12,
    STAGE_FUSED, /* the fused kernel (all of the above in one launch) */
    STAGE_EXPAND, /* compact coefficients to the dense pool (runs first) */
    STAGE_COUNT
};

/* Launch shape knobs, chosen per workload class by the autotuner. */
struct Tune {
    uint8_t res_split[4];    /* waves sharing one transform group, N = 4,8,16,32: 1, 2 or 4 */
    uint16_t intra_lanes;    /* threads per intra worker: 64 (one wave) or 256 */
    uint32_t intra_workers;  /* persistent intra workers */
    uint32_t intra_level_min; /* leading dependency levels holding at least this many
                                 blocks over the batch run as one launch per level (up
                                 to CVP9_MAX_HEAD_LEVELS), the rest on persistent
                                 workers; 0: persistent workers only */
    uint32_t intra_defer;    /* 1: the ordered intra stage publishes edges only and a
                                parallel pass (k_intra_px) writes the blocks' samples */
    uint32_t intra_chains;   /* 1: a one-job batch whose frame lists chains of intra blocks
                                runs them on chain workers (k_intra_chain) */
};

/* Total work items of each stage over the whole batch (the last element of
 * each prefix array in DBatch), plus launch parameters. */
struct StageCounts {
    uint32_t exp;             /* compact-coefficient waves (4 chunks each) */
    uint32_t fill, blocks, inter, inter2, grp[4], intra, lf, nv12; /* inter: early motion-
                                      compensation units, inter2: late ones (padded) */
    uint32_t levels;          /* leading levels launched one by one */
    const uint32_t *level_n;  /* host arrays, per leading level over the batch: large and */
    const uint32_t *level_ns; /* small (at most 16x16) intra blocks */
    uint32_t intra_tail;      /* intra blocks left to the persistent workers */
    uint32_t chains;          /* chains listed by a one-job batch's frame (0 otherwise) */
    uint32_t chained;         /* and the intra blocks on them */
    uint32_t px;              /* waves of the deferred sample pass (k_intra_px) */
    Tune tune;
};

/* Default launch shape for a device with `cus` compute units. */
Tune cvp9_default_tune(int cus);

hipError_t cvp9_set_idct_tables(const double *m4, const double *m8, const double *m16,
                                const double *m32);
hipError_t cvp9_launch_copy_probe(const void *src, void *dst, size_t bytes, hipStream_t s);
/* done (optional): when *done is set and the stage is one kernel (motion
 * compensation, deblocking, NV12 pack), that kernel's own completion records
 * the event, with no marker behind it on the queue, and *done is cleared;
 * otherwise the caller records it. Give it only to a phase's last launch. */
hipError_t cvp9_launch_stage(int stage, const DBatch &b, const StageCounts &n, hipStream_t s,
                             hipEvent_t *done = nullptr);

/* Consecutive stages that can share one launch, starting at `stage`: the
 * owner map with motion compensation (2), and all four transform sizes (4)
 * while they need at most CVP9_RES_SMALL_MAX_WGS workgroups; 1 otherwise.
 * cvp9_launch_stages runs stages [stage, stage + span) with that launch. */
#define CVP9_RES_SMALL_MAX_WGS 2048
int cvp9_stage_span(int stage, const DBatch &b, const StageCounts &n);
hipError_t cvp9_launch_stages(int stage, int span, const DBatch &b, const StageCounts &n,
                              hipStream_t s, hipEvent_t *done = nullptr);
/* Motion compensation of each job's late units [n_mc_a, n_mc) (STAGE_MC
 * runs the early ones). */
hipError_t cvp9_launch_mc_late(const DBatch &b, const StageCounts &n, hipStream_t s,
                               hipEvent_t *done = nullptr);

/* Engines alive in this process (each may have a waiting launch on each of
 * its two queues): the persistent launches of all of them share the
 * device's resident capacity. */
int cvp9_live_engines();

/* Fused persistent kernel: every stage of the batch in one launch of `wgs`
 * co-resident workgroups (at most cvp9_fused_max_wgs()). */
hipError_t cvp9_launch_fused(const DBatch &b, uint32_t wgs, hipStream_t s);
uint32_t cvp9_fused_max_wgs(int device);

/* *dst = *src on queue s (a stream's error word into a per-frame host word). */
hipError_t cvp9_launch_copy_word(const uint32_t *src, uint32_t *dst, hipStream_t s);

#endif
