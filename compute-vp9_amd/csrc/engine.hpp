/*
 * engine.hpp — host side of the reconstruction engine (internal C++).
 *
 *   plan_*      validate a parsed frame and lay out its device image
 *   StreamState one decode stream: device reference pool + scratch
 *   BatchBuilder one launch per stage over the frames of many streams
 */
#ifndef CVP9_ENGINE_HPP
#define CVP9_ENGINE_HPP

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "cvp9_hip/abi_types.h"
#include "cvp9_hip/engine.h"
#include "dev_types.h"
#include "kernels.h"

namespace cvp9 {

/* A persistent pool of host threads for the planner: run(n, fn) calls
 * fn(0) .. fn(n - 1) on the pool's threads and the calling thread and
 * returns when all have returned. One job at a time: a caller that finds
 * the pool busy (another thread's job) runs its tasks itself. */
class HostPool {
public:
    static HostPool &get();
    int threads() const { return (int)th_.size() + 1; }
    void run(int n, const std::function<void(int)> &fn);
    ~HostPool();

private:
    struct Job;
    HostPool();
    void worker();
    std::vector<std::thread> th_;
    std::mutex mu_, run_mu_;
    std::condition_variable cv_, idle_cv_;
    Job *job_ = nullptr;
    std::atomic<uint64_t> gen_{0};
    int users_ = 0;
    bool stop_ = false;
};

/* An intra block as the ordering pass sees it: its local block index and
 * its clamped rectangle in 4x4 cells, inclusive. */
struct IntraGeom {
    uint32_t b;
    uint16_t cx0, cy0, cx1, cy1;
    uint32_t reads; /* CVP9_READS_* of its modes */
};

/* Counts of a contiguous range [b0, b1) of the parsed block list, from
 * which the planner's parallel pass finds where each range writes, and the
 * range's intra blocks (local indices counted from the range's first
 * block). */
struct PlanChunk {
    uint32_t b0 = 0, b1 = 0;
    uint32_t n_blocks = 0, n_intra = 0, n_mc_not0 = 0, n_mc_0 = 0;
    uint32_t ntx[4] = {0, 0, 0, 0};
    std::vector<IntraGeom> intra;
};

/* Chains of intra blocks a prepared frame lists at most (see PrepLayout). */
constexpr uint32_t kMaxChains = 256;

/* Byte layout of a prepared frame (host staging and device copy alike). */
struct PrepLayout {
    uint32_t W = 0, H = 0;
    int key = 0;
    int lf_thr = 0;
    uint32_t fill = 0;
    uint32_t n_blocks = 0, n_inter = 0, n_intra = 0;
    uint32_t cls_ntx[4] = {0, 0, 0, 0};
    uint32_t num_coeffs = 0;
    uint32_t n_mc = 0;              /* motion-compensation units (16 lanes each) */
    uint32_t n_mc_not0 = 0;         /* units of blocks not predicted from slot 0: first in
                                       the unit table, the ones the previous frame (unless
                                       KEY) leaves alone */
    uint64_t off_blocks = 0, off_intra = 0, off_mc_tab = 0;
    uint64_t off_px = 0;            /* intra-list entries, blocks of at most 8x8 first */
    uint32_t n_px_small = 0;
    uint64_t off_tx_tab[4] = {0, 0, 0, 0};
    uint64_t off_coeffs = 0, bytes = 0;
    /* Compact coefficients (set compact before plan_measure): the parser's
     * pool in chunks of 64, a 64-bit mask of each chunk's nonzero entries
     * (off_cmask), the position of each chunk's first nonzero value among
     * the packed values (off_cstart) and the nonzero values in pool order
     * (off_cvals, room for the whole pool); the dense pool at off_coeffs,
     * n_chunks * 64 entries, is written on the device from them. Only the
     * image's first `upload` bytes (set by plan_write) need to reach the
     * device, and only its first `host_bytes` are written on the host. */
    bool compact = false;
    uint32_t n_chunks = 0, nnz = 0;
    uint64_t off_cmask = 0, off_cstart = 0, off_cvals = 0;
    uint64_t upload = 0, host_bytes = 0;
    /* written by plan_write: intra-list position where each dependency
     * level starts, then the list's end (empty without intra blocks), and
     * the number of small blocks (at most 16x16, first in their level) */
    std::vector<uint32_t> levels;
    std::vector<uint32_t> level_small;
    /* Chains of intra blocks (set chains before plan_measure): up to
     * kMaxChains chains, each block continuing the chain of its
     * highest-level parent, the most critical first (dependency level plus
     * height of the chain's first block). At off_chain: n_chains + 1
     * offsets into the chained entries, the chained entries (intra-list
     * positions in chain order), then every other entry in list order. */
    bool chains = false;
    uint32_t n_chains = 0, n_chained = 0;
    uint64_t off_chain = 0;
    /* the block list in ranges, one per planner thread (one range when the
     * frame is planned on the calling thread) */
    std::vector<PlanChunk> chunks;
};

/* Validates pf and fills L (keeping L.compact, the caller's choice of
 * coefficient layout). Returns CVP9_OK, CVP9_ERR_UNSUPPORTED (layout
 * this backend does not take: the caller falls back to the CPU path) or
 * CVP9_ERR_INVALID_DATA (coefficient references outside the pool): the code
 * of the first offending block in list order, as a serial pass finds it.
 * Frames of at least kParallelBlocks blocks are validated by up to
 * `threads` host threads (0: all of the pool's). */
constexpr uint32_t kParallelBlocks = 2048;
cvp9_err_t plan_measure(const vp9_parsed_frame_t *pf, PrepLayout &L, int threads = 0);
/* Writes the device image of pf into dst (L.bytes bytes) and L.levels, on
 * as many threads as plan_measure used. The image does not depend on the
 * number of threads. */
void plan_write(const vp9_parsed_frame_t *pf, PrepLayout &L, uint8_t *dst);

/* Device frame buffer: planes Y, U, V back to back, 64-byte row pitch. */
struct FrameBuf {
    uint8_t *dev = nullptr;
    size_t cap = 0;
    uint32_t w = 0, h = 0, sy = 0, suv = 0;
    int refcnt = 0;
    DFrame dframe() const;
};

size_t frame_bytes(uint32_t w, uint32_t h, uint32_t *sy, uint32_t *suv);

class StreamState {
public:
    explicit StreamState(int device) : device_(device) {}
    ~StreamState();
    StreamState(const StreamState &) = delete;
    StreamState &operator=(const StreamState &) = delete;

    /* frame for the next output; hipSuccess or an allocation error. With
     * avoid_prev, never a frame the previous frame reads as a reference
     * (that frame's later stages may still run while this one's first
     * phase writes its output) */
    hipError_t acquire(uint32_t w, uint32_t h, FrameBuf **out, bool may_free = true,
                       bool avoid_prev = false);
    /* motion-compensation units of a frame with layout L that do not read a
     * slot the previous frame writes: all without a previous frame, none
     * after a KEY frame (it writes every slot), else those not predicted
     * from slot 0 */
    uint32_t mc_independent(const PrepLayout &L) const
    {
        return last_written_ < 0 ? L.n_mc : (last_written_ == 1 ? 0 : L.n_mc_not0);
    }
    /* scratch already large enough for L (no allocation needed) */
    bool scratch_fits(const PrepLayout &L) const;
    /* Frame pipelining: the owner map, the neighbour records and the
     * residual buffer come in two sets, alternated frame by frame, so a
     * frame's owner map, neighbour records and transforms can run while the
     * previous frame still reads its set. With two_sets off every frame uses
     * set 0. */
    void set_two_sets(bool on) { two_sets_ = on; if (!on) par_ = 0; }
    int parity() const { return par_; }
    /* serial of the batch whose later stages last read set p (0: none) */
    uint64_t reader(int p) const { return reader_[p]; }
    void release(FrameBuf *f);
    /* reference-pool rule of the CPU path: KEY -> all 8 slots, else slot 0
     * (src/decoder/vp9_decoder.c:224-231) */
    void update_refs(bool key, FrameBuf *out);
    /* back to the state of a new stream (empty slots, no error), keeping
     * device memory; the caller has synchronised with the stream's work */
    hipError_t reset();
    /* owner map, edge and neighbour records and residual scratch large
     * enough for L;
     * may synchronize s */
    hipError_t ensure_scratch(const PrepLayout &L, hipStream_t s);
    /* job descriptor for one frame of this stream, on the current set; the
     * batch with serial `serial` reads that set, and the next frame takes
     * the other one (with two sets) */
    DJob make_job(const PrepLayout &L, const uint8_t *dev_prep, FrameBuf *out, uint8_t *nv12,
                  uint64_t serial = 0);
    /* bytes of HBM this stream holds */
    size_t resident_bytes() const;
    /* the stream's error word in device memory (the kernels poll it), and
     * its host-visible copy's value read on the host: nonzero once a
     * dependency wait of the stream gave up (bit 0), a neighbour record was
     * inconsistent (bit 8) or a grid barrier timed out (bit 16) */
    uint32_t *err_word() const { return errd_; }
    uint32_t error() const { return err_ ? __atomic_load_n(err_, __ATOMIC_ACQUIRE) : 0u; }
    FrameBuf *ref(int i) const { return refs_[i]; }

private:
    int device_;
    std::vector<FrameBuf *> all_;
    FrameBuf *refs_[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    uint32_t *owner_[2] = {nullptr, nullptr};
    size_t owner_cap_[2] = {0, 0};
    uint64_t *edge_ = nullptr;
    size_t edge_cap_ = 0;
    uint64_t *nbr_[2] = {nullptr, nullptr};
    size_t nbr_cap_[2] = {0, 0}; /* intra blocks */
    int16_t *resid_[2] = {nullptr, nullptr};
    size_t resid_cap_[2] = {0, 0};
    /* error word: errd_ in device memory (polled by waits), err_ its copy
     * in pinned, coherent host memory that the kernels set with
     * system-scope byte stores (err_dev_ its device view), so the host reads
     * it without a copy */
    uint32_t *errd_ = nullptr;
    uint32_t *err_ = nullptr, *err_dev_ = nullptr;
    /* device word: the stream's frame epoch, advanced on the device by the
     * first kernel of every batch (so a replayed graph gets new epochs);
     * never reset, so edge words of earlier frames never match a new one */
    uint32_t *ectr_ = nullptr;
    bool two_sets_ = false;
    int par_ = 0;
    uint64_t reader_[2] = {0, 0};
    FrameBuf *prev_read_[8] = {nullptr, nullptr, nullptr, nullptr,
                               nullptr, nullptr, nullptr, nullptr}; /* refs of the last frame */
    int last_written_ = -1; /* slots the last frame wrote: -1 none yet, 0 slot 0, 1 all */
    bool read_by_prev(const FrameBuf *f) const
    {
        for (const FrameBuf *p : prev_read_)
            if (p == f) return true;
        return false;
    }
};

/* Packs job descriptors + per-stage prefix sums for one batch into a host
 * buffer mirrored at a device address, and launches the stage sequence. */
class BatchBuilder {
public:
    /* descriptor bytes for njobs jobs */
    static size_t bytes_for(uint32_t njobs);
    /* host/dev point to bytes_for(njobs) bytes each; layouts[i] is the plan
     * of jobs[i]; fused: the batch runs in the fused kernel (whose barrier
     * words the descriptor then carries) */
    void build(const DJob *jobs, const PrepLayout *const *layouts, uint32_t njobs, uint8_t *host,
               uint8_t *dev, const Tune &tune, uint32_t spin_limit, bool fused = false);
    size_t used() const { return used_; }
    const DBatch &batch() const { return b_; }
    void set_resid_all(bool on) { b_.resid_all = on ? 1u : 0u; }
    const StageCounts &counts() const { return n_; }
    static hipError_t launch(const DBatch &b, const StageCounts &n, hipStream_t s);

private:
    DBatch b_{};
    StageCounts n_{};
    size_t used_ = 0;
    std::vector<uint32_t> level_n_, level_ns_;
};

/* One frame of engine stream sid from a prepared image (layout L) staged in
 * host memory at host, copied to dev on the batch's first queue: the drop-in
 * backend's path through the engine. On success res names the output frame
 * (the stream's new slot 0), the HIP queue the batch finishes on and the
 * stream's error word. */
struct StagedResult {
    DFrame out;
    hipStream_t queue;
    uint32_t *err;
};
/* Reference slot `slot` of engine stream sid as the host pool tracks it
 * (y == nullptr: empty). */
DFrame engine_stream_ref(cvp9_engine_t *e, int sid, int slot);
cvp9_err_t engine_run_staged(cvp9_engine_t *e, int sid, const PrepLayout &L, const uint8_t *host,
                             uint8_t *dev, uint8_t *nv12, StagedResult *res);

/* Once per device: upload the IDCT matrices (computed with the C library's
 * cos exactly as the reference builds them). */
hipError_t init_device(int device);
/* Host copy of the N x N matrix the kernels use. */
void idct_matrix(int n, double *m);

} // namespace cvp9

/* A parsed frame staged in HBM (cvp9_engine_prepare). */
struct cvp9_prep {
    cvp9::PrepLayout L;
    uint8_t *dev = nullptr;
    uint64_t algo = 0;
};

#endif
