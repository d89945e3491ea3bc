// batch.hpp -- resident rank requests and the tick runner shared by the
// batched and served modes.
#pragma once
#include "common.hpp"
#include "pmbasis.hpp"
#include "spmm.hpp"
#include <vector>

namespace rr {

// Rows [row_beg, row_end) of a block vector whose t = 0 projection x^T x is
// wanted; self-described so a single launch can mix requests.
struct GramTile {
    const uint32_t *in;  // the block vector (stride B)
    uint32_t *part;      // this tile's b(b+1)/2 projection partial
    uint32_t row_beg, row_end;
    uint32_t b, pad;
};

struct RedJob {  // S_t of one request = sum of its tiles' partials
    const uint32_t *part;  // ntiles consecutive partials of P words
    uint32_t ntiles, P;
    uint32_t *out;
};

// STAGE_EXTRACT: the basis is complete and the rank inputs are on their way to the host
enum : int { STAGE_KRYLOV = 0, STAGE_PM = 1, STAGE_DONE = 2, STAGE_FAILED = 3, STAGE_EXTRACT = 4 };

// resident device footprint of one request (admission control), bytes
uint64_t request_bytes(uint64_t m, uint64_t n, uint64_t nnz, uint32_t b, uint64_t L, uint64_t d, bool packed);

// Steps (nonzeros per row group) a wavefront of a request is given when its
// operands are sliced: small requests get few wavefronts, large ones up to a
// full device.  A batch is throughput-bound (512 matrices: 2.43 s at 48 or
// 64, 2.46 s at 32, 2.50 s at 128, 2.82 s at 16); the server also answers
// for latency, and under partial load a request with twice the wavefronts
// finishes sooner (256 served requests at 0.7x their saturated rate: p50 108
// ms and p99 1.24 s at 32, 152 ms and 1.58 s at 64, the same throughput).
constexpr uint32_t BATCH_STEPS_PER_WAVE = 64;
constexpr uint32_t SERVER_STEPS_PER_WAVE = 32;
// Ticks per chunk in rr_batch_rank: their passes are launched back to back
// and their partials reduced by one launch.
constexpr uint32_t BATCH_TICKS = 16;
// PM-basis rounds rr_batch_rank keeps queued ahead of the device.
constexpr uint32_t BATCH_PM_AHEAD = 8;
// Ticks per chunk of the server: few while requests wait to join (so they
// join soon), more while none does (fewer launches per entry).
constexpr uint32_t SERVER_TICKS = 4;
constexpr uint32_t SERVER_TICKS_IDLE = 16;
// Chunks and PM-basis rounds the server keeps queued ahead of the device.
constexpr uint32_t SERVER_K_AHEAD = 2;
constexpr uint32_t SERVER_PM_AHEAD = 8;
// The server's PM-basis scheduler keeps twiddle tables for transforms of up to
// 2^SERVER_MAX_LOGN points (orders below 2^20); longer requests are refused
// at submission (rr_rank and rr_batch_rank take orders up to 2^26 - 1).
constexpr uint32_t SERVER_MAX_LOGN = 20;
// The batched and served modes start the PM-basis of a request of at least
// this order while its sequence is still being produced: the leftmost chain
// of the recursion reads only leading coefficients of G.
constexpr uint64_t PM_STREAM_D = 1024;
// rr_rank streams only matrices of at least this many nonzeros: on smaller
// ones the chunked sequence costs more than the overlap saves.
constexpr uint64_t RANK_STREAM_NNZ = 1ull << 21;

struct Request {
    uint32_t p = 0, b = 0, B = 0, P = 0, su = 4, nwg = 0;
    uint64_t m = 0, n = 0, nnz = 0, L = 0, d = 0, t = 0;
    bool packed = false;
    int stage = STAGE_KRYLOV;
    rr_status fail = RR_OK;  // why the request failed (stage STAGE_FAILED)
    int64_t rank = -1;
    uint64_t id = 0, budget_bytes = 0;
    double t_arrive = 0, t_admit = 0, t_done = 0;
    SliceMat sa, sat;    // A and A^T in sliced form, nwg workgroups each
    SpmmArgs args[2];    // pass by A, by A^T (gram_off set per tick)
    DevBuf dr_d, dc_d, dc2_d, x0, z, y, seq, G, M, delta, rank_in;
    std::vector<GramTile> tiles_gram;
    PmJob pm;

    // su: column block size 3 or 4 (0: the one that pads least)
    rr_status init(const rr_csr *A, uint32_t p, uint32_t b, const uint32_t *dr, const uint32_t *dc,
                   const uint32_t *V, uint64_t L, uint64_t d, uint32_t tile_rows, hipStream_t st, uint32_t su = 0,
                   uint32_t steps_per_wave = BATCH_STEPS_PER_WAVE);
    // PM-basis buffers; the job then reads the coefficients of G built so far
    rr_status alloc_pm();
    // G's coefficients for the sequence entries produced so far (t), on st
    rr_status grow_G(hipStream_t st);
    // sequence complete: build G for the PM scheduler; with release, wait for
    // the stream and free the sliced operands
    rr_status start_pm(hipStream_t st, bool release = true);
    rr_status finish_rank(hipStream_t st);  // PM done: rank formula from the extracted coefficients
    uint64_t device_bytes() const;          // resident footprint estimate (admission control)
};

rr_status rank_from_device_basis(const uint32_t *M, const int64_t *delta, uint32_t b, uint64_t len, uint32_t p,
                                 uint32_t *scratch, hipStream_t st, int64_t *rank);
// The same in two steps: the 2b^2 + 1 words the formula needs, extracted on
// the device and queued for copying into host (pinned) memory on st; then the
// rank from those words once the copy is done.
rr_status rank_extract_async(const uint32_t *M, const int64_t *delta, uint32_t b, uint64_t len, uint32_t *scratch,
                             uint32_t *host, hipStream_t st);
int64_t rank_from_extracted(const uint32_t *host, uint32_t b, uint32_t p);
rr_status rank_from_device_seq(const uint32_t *seq_dev, uint32_t b, uint64_t d, uint32_t p, hipStream_t st,
                               int64_t *rank);

const uint32_t *krylov_device_seq(rr_krylov *c);  // the last rr_krylov_run's entries, from index 0
hipStream_t krylov_stream(rr_krylov *c);

// The L-entry sequence of context k (copied to seq_out) and the rank from its
// PM-basis of order d: the sequence is produced in chunks, and the PM-basis
// runs on a second stream on the coefficients of G produced so far.
rr_status rank_streamed(rr_krylov *k, uint32_t b, uint64_t L, uint64_t d, uint32_t p, uint32_t *seq_out,
                        int64_t *rank);

// Events marking the last `cap` items issued to one stream, oldest first.
struct EventRing {
    std::vector<hipEvent_t> ev;
    size_t head = 0, used = 0;
    rr_status init(size_t cap) {
        ev.assign(cap, nullptr);
        for (auto &e : ev) RR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        return RR_OK;
    }
    ~EventRing() {
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    bool room() {
        retire();
        return used < ev.size();
    }
    rr_status push(hipStream_t st) {
        RR_HIP(hipEventRecord(ev[(head + used) % ev.size()], st));
        used++;
        return RR_OK;
    }
    bool retire() {  // drops the items the device has finished; true if any
        bool any = false;
        while (used && hipEventQuery(ev[head]) == hipSuccess) {
            head = (head + 1) % ev.size();
            used--;
            any = true;
        }
        return any;
    }
};

struct TickRunner {
    static constexpr size_t STAGE_BYTES = 4u << 20;
    static constexpr int NSTAGE = 8;
    hipStream_t st = nullptr;
    int cus = 0;
    char *stage_host = nullptr;
    hipEvent_t stage_ev[NSTAGE] = {};
    int stage_next = 0;
    DevBuf desc_dev, part;
    size_t part_cap = 0;
    uint64_t ticks = 0;
    // after tick(): whether any of its work reached the stream (a tick that
    // fails with this false changed nothing, and may be retried)
    bool launched = false;
    // tests of failure handling: a tick holding this request id fails with
    // RR_ENOMEM before any of its work is queued (0: never)
    uint64_t fail_id = 0;

    rr_status init(hipStream_t st, int cus);
    // nticks entries of every request; descriptors beyond one staging slot
    // are split over several staged launches
    rr_status tick(std::vector<Request *> &reqs, uint32_t p, uint32_t nticks = 1);
    ~TickRunner();

  private:
    rr_status tick_staged(std::vector<Request *> &reqs, uint32_t p, uint32_t nticks);
};

}  // namespace rr
