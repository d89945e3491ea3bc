// server.cpp -- continuous-batching rank server on one GPU.
//
// No counterpart in the reference (one matrix per process,
// src/main.rs:254-639).  Requests arrive at any time through
// rr_server_submit; a worker thread admits them in arrival order while their
// resident footprint fits the HBM budget, keeps each admitted request's
// matrix, vectors and partial sequence on the device, and loops:
//   - one Krylov entry for every request still producing its sequence
//     (stream 1; requests at different entries and of different sizes share
//     the launches),
//   - one batched PM-basis round for every request whose sequence is
//     complete (stream 2, so it overlaps the Krylov launches),
//   - completion: the rank inputs are extracted on the device, the result is
//     published and the request's memory returns to the budget.
// A request that fails (invalid input, allocation) is reported with its own
// status and does not affect the others.
#include "common.hpp"
#include "batch.hpp"
#include <thread>
#include <mutex>
#include <condition_variable>
#include <deque>
#include <memory>
#include <chrono>
#include <cstring>

namespace rr {

struct Pending {
    uint64_t id;
    double t_arrive;
    uint32_t b;
    uint64_t m, n, L, d;
    std::vector<uint64_t> rp;
    std::vector<uint32_t> col, val, dr, dc, V;
};

static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace rr

struct rr_server {
    uint32_t p = 0;
    uint64_t budget = 0, used = 0;
    int device = 0, cus = 0;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::unique_ptr<rr::Pending>> queue;
    std::vector<rr_result> results;
    bool stop = false;
    uint64_t next_id = 1;
    // owned by the worker thread
    hipStream_t st_k = nullptr, st_pm = nullptr;
    rr::TickRunner ticks;
    rr::PmScheduler pm;
    std::vector<std::unique_ptr<rr::Request>> active;
    uint64_t completed = 0, failed = 0, pm_rounds = 0, max_active = 0;

    void run();
    void publish(rr::Request *R, int32_t status, uint64_t bytes);
};

void rr_server::publish(rr::Request *R, int32_t status, uint64_t bytes) {
    rr_result res;
    res.id = R->id;
    res.rank = status == RR_OK ? R->rank : -1;
    res.status = status;
    res.latency_ms = rr::now_ms() - R->t_arrive;
    res.queue_ms = R->t_admit - R->t_arrive;
    std::lock_guard<std::mutex> g(mu);
    results.push_back(res);
    used -= std::min(used, bytes);
    if (status == RR_OK) completed++;
    else failed++;
    cv.notify_all();
}

void rr_server::run() {
    using namespace rr;
    (void)hipSetDevice(device);
    std::vector<Request *> kry;
    std::vector<PmJob *> jobs;
    for (;;) {
        // admission, arrival order, while the footprint fits the budget
        std::vector<std::unique_ptr<Pending>> admit;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (stop) return;
            if (active.empty() && queue.empty()) {
                cv.wait_for(lk, std::chrono::milliseconds(20));
                continue;
            }
            while (!queue.empty()) {
                Pending &q = *queue.front();
                uint64_t need = request_bytes(q.m, q.n, q.rp[q.m], q.b, q.L, q.d, false);
                // a request larger than the whole budget runs alone
                if (used + need > budget && !(active.empty() && admit.empty())) break;
                used += need;
                admit.push_back(std::move(queue.front()));
                queue.pop_front();
            }
        }
        for (auto &q : admit) {
            std::unique_ptr<Request> R(new Request());
            rr_csr A{q->m, q->n, q->rp.data(), q->col.data(), q->val.data()};
            R->id = q->id;
            R->t_arrive = q->t_arrive;
            R->t_admit = now_ms();
            R->budget_bytes = request_bytes(q->m, q->n, q->rp[q->m], q->b, q->L, q->d, false);
            rr_status s = R->init(&A, p, q->b, q->dr.data(), q->dc.data(), q->V.data(), q->L, q->d, 2048, st_k);
            // a long sequence's PM-basis starts on its leading coefficients
            if (s == RR_OK && R->d >= PM_STREAM_D && (s = R->alloc_pm()) == RR_OK) s = pm.add(&R->pm);
            if (s != RR_OK) {
                publish(R.get(), s, R->budget_bytes);
                continue;
            }
            active.push_back(std::move(R));
        }
        max_active = std::max<uint64_t>(max_active, active.size());
        // the next Krylov entries (up to SERVER_TICKS, none past a sequence's
        // end; one while a request is at its first) of every request still
        // producing its sequence
        kry.clear();
        bool first = false;
        uint64_t left = ~0ull;
        for (auto &R : active)
            if (R->stage == STAGE_KRYLOV && R->t < R->L) {
                kry.push_back(R.get());
                first |= R->t == 0;
                left = std::min<uint64_t>(left, R->L - R->t);
            }
        const uint32_t nt = first ? 1u : (uint32_t)std::min<uint64_t>(SERVER_TICKS, left);
        if (!kry.empty() && ticks.tick(kry, p, nt) != RR_OK)
            for (auto *R : kry) R->stage = STAGE_FAILED;
        // sequences completed: build G on the Krylov stream, hand over to the
        // PM stream; PM-basis jobs already running get the coefficients their
        // next operation waits for as soon as the sequence has them
        bool handed = false;
        for (auto &R : active) {
            if (R->stage != STAGE_KRYLOV) continue;
            const bool running = R->G.ptr != nullptr;
            if (R->t >= R->L) {
                if (R->start_pm(st_k) != RR_OK || (!running && pm.add(&R->pm) != RR_OK)) {
                    R->stage = STAGE_FAILED;
                    continue;
                }
                handed = true;
            } else if (running && R->pm.has_op && R->pm.op_need > R->pm.avail && R->t >= R->pm.op_need) {
                if (R->grow_G(st_k) != RR_OK) {
                    R->stage = STAGE_FAILED;
                    continue;
                }
                handed = true;
            }
        }
        if (handed) {
            hipEvent_t ev;
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess) {
                (void)hipEventRecord(ev, st_k);
                (void)hipStreamWaitEvent(st_pm, ev, 0);
                (void)hipEventDestroy(ev);
            }
        }
        // one PM-basis round for every request whose sequence is complete or
        // whose PM-basis runs on its leading coefficients
        auto in_pm = [](const Request *R) {
            return !R->pm.done && (R->stage == STAGE_PM || (R->stage == STAGE_KRYLOV && R->G.ptr));
        };
        jobs.clear();
        for (auto &R : active)
            if (in_pm(R.get())) jobs.push_back(&R->pm);
        if (!jobs.empty()) {
            bool launched = false;
            if (pm.round(jobs, &launched) != RR_OK)
                for (auto &R : active)
                    if (in_pm(R.get())) R->stage = STAGE_FAILED;
            pm_rounds += launched;
        }
        // completion
        for (size_t i = 0; i < active.size();) {
            Request *R = active[i].get();
            bool finished = false;
            int32_t status = RR_OK;
            if (R->stage == STAGE_FAILED) {
                finished = true;
                status = RR_EDEVICE;
            } else if (R->stage == STAGE_PM && R->pm.done) {
                status = R->finish_rank(st_pm);
                finished = true;
            }
            if (finished) {
                R->t_done = now_ms();
                publish(R, status, R->budget_bytes);
                active.erase(active.begin() + i);
            } else {
                i++;
            }
        }
    }
}

using namespace rr;

extern "C" {

rr_status rr_server_create(uint32_t p, uint64_t hbm_budget_bytes, rr_server **out) {
    clear_error();
    RR_CHECK(out, RR_EINVAL, "null argument");
    RR_CHECK(p > 2 && p < (1u << 31) && h_is_prime(p), RR_EINVAL, "p=%u must be a prime in (2, 2^31)", p);
    std::unique_ptr<rr_server> S(new rr_server());
    S->p = p;
    S->budget = hbm_budget_bytes;
    RR_HIP(hipGetDevice(&S->device));
    RR_HIP(hipDeviceGetAttribute(&S->cus, hipDeviceAttributeMultiprocessorCount, S->device));
    RR_HIP(hipStreamCreateWithFlags(&S->st_k, hipStreamNonBlocking));
    RR_HIP(hipStreamCreateWithFlags(&S->st_pm, hipStreamNonBlocking));
    rr_status s;
    if ((s = S->ticks.init(S->st_k, S->cus)) != RR_OK) return s;
    if ((s = S->pm.init(p, S->st_pm, NTT_MAX_LOGN, 128, 64)) != RR_OK) return s;
    rr_server *raw = S.release();
    raw->worker = std::thread([raw] { raw->run(); });
    *out = raw;
    return RR_OK;
}

rr_status rr_server_submit(rr_server *S, const rr_csr *A, uint32_t b, const uint32_t *row_precond,
                           const uint32_t *col_precond, const uint32_t *V, uint64_t seq_len, uint64_t order,
                           uint64_t *id_out) {
    clear_error();
    RR_CHECK(S && A && row_precond && col_precond && V && id_out, RR_EINVAL, "null argument");
    RR_CHECK(b >= 1 && b <= 64, RR_EINVAL, "b=%u must be in 1..64", b);
    RR_CHECK(A->m >= 1 && A->n >= 1 && A->row_ptr, RR_EINVAL, "empty matrix");
    const uint64_t nnz = A->row_ptr[A->m];
    RR_CHECK(nnz == 0 || (A->col && A->val), RR_EINVAL, "null col/val");
    std::unique_ptr<Pending> q(new Pending());
    q->b = b;
    q->m = A->m;
    q->n = A->n;
    q->L = seq_len ? seq_len : 2 * std::min(A->m, A->n) / b + 8;
    q->d = order ? order : q->L - 1;
    q->rp.assign(A->row_ptr, A->row_ptr + A->m + 1);
    q->col.assign(A->col, A->col + nnz);
    q->val.assign(A->val, A->val + nnz);
    q->dr.assign(row_precond, row_precond + A->m);
    q->dc.assign(col_precond, col_precond + A->n);
    q->V.assign(V, V + A->n * b);
    q->t_arrive = now_ms();
    std::lock_guard<std::mutex> g(S->mu);
    q->id = S->next_id++;
    *id_out = q->id;
    S->queue.push_back(std::move(q));
    S->cv.notify_all();
    return RR_OK;
}

rr_status rr_server_poll(rr_server *S, rr_result *out, uint32_t cap, uint32_t *n_out, int wait_ms) {
    RR_CHECK(S && n_out && (out || cap == 0), RR_EINVAL, "null argument");
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->results.empty() && wait_ms > 0)
        S->cv.wait_for(lk, std::chrono::milliseconds(wait_ms), [S] { return !S->results.empty(); });
    uint32_t n = (uint32_t)std::min<size_t>(cap, S->results.size());
    for (uint32_t i = 0; i < n; i++) out[i] = S->results[i];
    S->results.erase(S->results.begin(), S->results.begin() + n);
    *n_out = n;
    return RR_OK;
}

rr_status rr_server_stats_get(rr_server *S, rr_server_stats *out) {
    RR_CHECK(S && out, RR_EINVAL, "null argument");
    std::lock_guard<std::mutex> g(S->mu);
    out->queued = S->queue.size();
    out->used_bytes = S->used;
    out->completed = S->completed;
    out->failed = S->failed;
    out->ticks = S->ticks.ticks;
    out->pm_rounds = S->pm_rounds;
    out->max_active = S->max_active;
    return RR_OK;
}

void rr_server_destroy(rr_server *S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
        S->queue.clear();
        S->cv.notify_all();
    }
    if (S->worker.joinable()) S->worker.join();
    if (S->st_k) (void)hipStreamSynchronize(S->st_k);
    if (S->st_pm) (void)hipStreamSynchronize(S->st_pm);
    S->active.clear();
    delete S;
}

}  // extern "C"
