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
// Host set-up of an admitted request (validation, transpose, slicing,
// uploads) runs on set-up threads with streams of their own, so admissions do
// not stall the ticks of the requests in flight.  A request that fails
// (invalid input, allocation, a failing operation of its own) is reported
// with its own status; a tick or PM-basis round that fails before any of its
// work reached the device is retried request by request, so only the
// requests that fail alone are dropped.
#include "common.hpp"
#include "batch.hpp"
#include <thread>
#include <mutex>
#include <condition_variable>
#include <deque>
#include <memory>
#include <chrono>
#include <cstring>
#include <string>
#include <map>

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

// A request between admission and the tick loop: set up on a set-up thread.
struct Staged {
    std::unique_ptr<rr::Request> R;
    std::unique_ptr<rr::Pending> q;
    rr_status status = RR_OK;
    std::string msg;
    hipEvent_t ev = nullptr;  // its uploads are done
};

constexpr int SERVER_SETUP_THREADS = 4;
// Pinned host slots the rank inputs of completed bases are copied into (2b^2 + 1 words each, b <= 64).
constexpr int SERVER_EXTRACT_SLOTS = 64;
constexpr size_t EXTRACT_WORDS = 2 * 64 * 64 + 4;

struct rr_server {
    uint32_t p = 0;
    uint64_t budget = 0, used = 0;
    int device = 0, cus = 0;
    std::thread worker;
    std::vector<std::thread> setup;
    std::vector<hipStream_t> setup_st;
    std::mutex mu;
    std::condition_variable cv, cv_setup;
    std::deque<std::unique_ptr<rr::Pending>> queue;     // submitted, not admitted
    std::deque<std::unique_ptr<Staged>> to_setup;       // admitted, set-up pending
    std::deque<std::unique_ptr<Staged>> set_up;         // set up, joining the tick loop
    uint64_t in_setup = 0;                              // admitted and not yet joined (any of the three steps)
    std::vector<rr_result> results;
    bool stop = false, broken = false;
    uint64_t next_id = 1;
    // owned by the worker thread
    hipStream_t st_k = nullptr, st_pm = nullptr;
    rr::TickRunner ticks;
    rr::PmScheduler pm;
    std::vector<std::unique_ptr<rr::Request>> active;
    uint64_t completed = 0, failed = 0, pm_rounds = 0, max_active = 0;
    uint32_t *pinned = nullptr;     // SERVER_EXTRACT_SLOTS x EXTRACT_WORDS
    std::vector<int> free_slots;
    struct Extract {                // a request whose rank inputs are being copied back
        int slot;
        hipEvent_t ev;
    };
    std::map<const rr::Request *, Extract> extracting;

    void run();
    void setup_loop(int w);
    void publish(rr::Request *R, int32_t status, uint64_t bytes);
};

void rr_server::setup_loop(int w) {
    using namespace rr;
    const bool dev_ok = hipSetDevice(device) == hipSuccess;
    for (;;) {
        std::unique_ptr<Staged> S;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_setup.wait(lk, [this] { return stop || !to_setup.empty(); });
            if (stop) return;
            S = std::move(to_setup.front());
            to_setup.pop_front();
        }
        Pending &q = *S->q;
        Request &R = *S->R;
        const rr_status e = guarded([&]() -> rr_status {
            if (!dev_ok) {
                set_error("set-up thread could not select the device");
                return RR_EDEVICE;
            }
            rr_csr A{q.m, q.n, q.rp.data(), q.col.data(), q.val.data()};
            rr_status s = R.init(&A, p, q.b, q.dr.data(), q.dc.data(), q.V.data(), q.L, q.d, 2048, setup_st[w], 0,
                                 SERVER_STEPS_PER_WAVE);
            if (s == RR_OK && (hipEventCreateWithFlags(&S->ev, hipEventDisableTiming) != hipSuccess ||
                               hipEventRecord(S->ev, setup_st[w]) != hipSuccess)) {
                set_error("set-up event failed");
                s = RR_EDEVICE;
            }
            return s;
        });
        S->status = e;
        if (e != RR_OK) S->msg = rr_last_error();
        S->q.reset();  // the host copy of the inputs is no longer needed
        std::lock_guard<std::mutex> g(mu);
        set_up.push_back(std::move(S));
        cv.notify_all();
    }
}

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
    std::vector<Request *> kry, fresh;
    std::vector<PmJob *> jobs;
    EventRing kq, pq;
    hipEvent_t handoff = nullptr;
    if (kq.init(SERVER_K_AHEAD) != RR_OK || pq.init(SERVER_PM_AHEAD) != RR_OK ||
        hipEventCreateWithFlags(&handoff, hipEventDisableTiming) != hipSuccess) {
        std::lock_guard<std::mutex> g(mu);  // cannot run: every request fails
        broken = true;
        cv.notify_all();
        return;
    }
    struct EvGuard {
        hipEvent_t e;
        ~EvGuard() { (void)hipEventDestroy(e); }
    } handoff_guard{handoff};
    bool pm_idle = true;
    auto fail = [](Request *R, rr_status e) {
        if (!R || R->stage == STAGE_FAILED) return;
        R->stage = STAGE_FAILED;
        R->fail = e;
    };
    for (;;) {
        // admission, arrival order, while the footprint fits the budget; the
        // admitted requests go to the set-up threads
        std::vector<std::unique_ptr<Staged>> joining;
        {
            std::unique_lock<std::mutex> lk(mu);
            if (stop) return;
            if (active.empty() && queue.empty() && in_setup == 0) {
                cv.wait_for(lk, std::chrono::milliseconds(20));
                continue;
            }
            while (!queue.empty()) {
                Pending &q = *queue.front();
                uint64_t need = request_bytes(q.m, q.n, q.rp[q.m], q.b, q.L, q.d, false);
                // a request larger than the whole budget runs alone
                const bool alone = active.empty() && in_setup == 0;
                if (used + need > budget && !alone) break;
                used += need;
                in_setup++;
                std::unique_ptr<Staged> S(new Staged());
                S->R.reset(new Request());
                S->R->id = q.id;
                S->R->t_arrive = q.t_arrive;
                S->R->t_admit = now_ms();
                S->R->budget_bytes = need;
                S->q = std::move(queue.front());
                queue.pop_front();
                to_setup.push_back(std::move(S));
                cv_setup.notify_one();
            }
            // set-up done: join once the uploads are complete
            for (size_t i = 0; i < set_up.size();) {
                Staged &S = *set_up[i];
                if (S.status == RR_OK && hipEventQuery(S.ev) == hipErrorNotReady) {
                    i++;
                    continue;
                }
                joining.push_back(std::move(set_up[i]));
                set_up.erase(set_up.begin() + i);
                in_setup--;
            }
            if (active.empty() && joining.empty()) {  // only set-ups in flight: wait for one to end
                cv.wait_for(lk, std::chrono::milliseconds(1));
                continue;
            }
        }
        for (auto &S : joining) {
            std::unique_ptr<Request> R = std::move(S->R);
            if (S->ev) (void)hipEventDestroy(S->ev);
            rr_status s = S->status;
            if (s != RR_OK) {
                set_error("%s", S->msg.c_str());
                publish(R.get(), s, R->budget_bytes);
                continue;
            }
            R->pm.tag = R->id;
            // a long sequence's PM-basis starts on its leading coefficients
            if (R->d >= PM_STREAM_D && (s = R->alloc_pm()) == RR_OK) s = pm.add(&R->pm);
            if (s != RR_OK) {
                publish(R.get(), s, R->budget_bytes);
                continue;
            }
            if (R->d >= PM_STREAM_D) pm_idle = false;
            active.push_back(std::move(R));
        }
        max_active = std::max<uint64_t>(max_active, active.size());
        bool issued = false;
        // The next Krylov chunk, while fewer than SERVER_K_AHEAD are queued:
        // requests at their first entry take a tick of their own; the others
        // up to SERVER_TICKS_IDLE entries (SERVER_TICKS while requests wait
        // to join), none past a sequence's end, partials within 64 M words.
        if (kq.room()) {
            kry.clear();
            fresh.clear();
            uint64_t left = ~0ull;
            size_t words = 0;
            for (auto &R : active)
                if (R->stage == STAGE_KRYLOV && R->t < R->L) {
                    if (R->t == 0) {
                        fresh.push_back(R.get());
                        continue;
                    }
                    kry.push_back(R.get());
                    left = std::min<uint64_t>(left, R->L - R->t);
                    words += (size_t)R->nwg * R->P;
                }
            bool waiting_join;
            {
                std::lock_guard<std::mutex> g(mu);
                waiting_join = !queue.empty() || in_setup > 0;
            }
            auto run_tick = [&](std::vector<Request *> &rs, uint32_t nt) {
                const rr_status ts = ticks.tick(rs, p, nt);
                if (ts != RR_OK && !ticks.launched) {
                    // nothing reached the device: each request alone, so only those failing alone drop out
                    for (auto *R : rs) {
                        std::vector<Request *> one{R};
                        const rr_status e = ticks.tick(one, p, nt);
                        if (e != RR_OK) fail(R, e);
                    }
                } else if (ts != RR_OK) {
                    for (auto *R : rs) fail(R, ts);  // the stream holds part of this tick's work
                }
            };
            if (!fresh.empty()) run_tick(fresh, 1);
            if (!kry.empty()) {
                uint64_t nt = std::min<uint64_t>(waiting_join ? SERVER_TICKS : SERVER_TICKS_IDLE, left);
                nt = std::max<uint64_t>(1, std::min<uint64_t>(nt, (64ull << 20) / std::max<size_t>(words, 1)));
                run_tick(kry, (uint32_t)nt);
            }
            if (!fresh.empty() || !kry.empty()) {
                (void)kq.push(st_k);
                issued = true;
            }
        }
        // sequences completed: build G on the Krylov stream, hand over to the
        // PM stream; PM-basis jobs already running get the coefficients their
        // next operation waits for as soon as the sequence has them
        bool handed = false;
        for (auto &R : active) {
            if (R->stage != STAGE_KRYLOV) continue;
            const bool running = R->G.ptr != nullptr;
            rr_status e = RR_OK;
            if (R->t >= R->L) {
                if ((e = R->start_pm(st_k)) != RR_OK || (!running && (e = pm.add(&R->pm)) != RR_OK)) {
                    fail(R.get(), e);
                    continue;
                }
                handed = true;
            } else if (running && R->pm.has_op && R->pm.op_need > R->pm.avail && R->t >= R->pm.op_need) {
                if ((e = R->grow_G(st_k)) != RR_OK) {
                    fail(R.get(), e);
                    continue;
                }
                handed = true;
            }
        }
        if (handed) {
            (void)hipEventRecord(handoff, st_k);
            (void)hipStreamWaitEvent(st_pm, handoff, 0);
            pm_idle = false;
            issued = true;
        }
        // PM-basis rounds for every request whose sequence is complete or whose
        // PM-basis runs on its leading coefficients, while fewer than
        // SERVER_PM_AHEAD are queued
        auto in_pm = [](const Request *R) {
            return !R->pm.done && (R->stage == STAGE_PM || (R->stage == STAGE_KRYLOV && R->G.ptr));
        };
        auto owner = [&](const PmJob *J) -> Request * {
            for (auto &R : active)
                if (&R->pm == J) return R.get();
            return nullptr;
        };
        while (!pm_idle && pq.room()) {
            jobs.clear();
            for (auto &R : active)
                if (in_pm(R.get())) jobs.push_back(&R->pm);
            if (jobs.empty()) {
                pm_idle = true;
                break;
            }
            bool launched = false;
            const rr_status ps = pm.round(jobs, &launched);
            if (ps != RR_OK && pm.last.empty()) {
                // no job's operation was taken: each job alone
                for (auto *J : jobs) {
                    std::vector<PmJob *> one{J};
                    bool l1 = false;
                    const rr_status e = pm.round(one, &l1);
                    launched |= l1;
                    if (e != RR_OK) fail(owner(J), e);
                }
            } else if (ps != RR_OK) {
                for (auto *J : pm.last) fail(owner(J), ps);  // their operations were taken and not completed
            }
            for (auto *J : jobs)  // a job's own failure (its recursion could not go on)
                if (J->err != RR_OK) fail(owner(J), J->err);
            if (!launched) {
                pm_idle = true;  // every job is done, failed or waits for its sequence
                break;
            }
            pm_rounds++;
            (void)pq.push(st_pm);
            issued = true;
        }
        // bases complete: their rank inputs go to pinned slots on the PM stream
        for (auto &R : active) {
            if (R->stage != STAGE_PM || !R->pm.done || free_slots.empty()) continue;
            const int slot = free_slots.back();
            uint32_t *h = pinned + (size_t)slot * EXTRACT_WORDS;
            hipEvent_t ev = nullptr;
            rr_status e = rank_extract_async(R->M.as<uint32_t>(), R->delta.as<int64_t>(), R->b, R->d + 1,
                                             R->rank_in.as<uint32_t>(), h, st_pm);
            if (e == RR_OK && (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
                               hipEventRecord(ev, st_pm) != hipSuccess)) {
                set_error("completion event failed");
                e = RR_EDEVICE;
            }
            if (e != RR_OK) {
                if (ev) (void)hipEventDestroy(ev);
                fail(R.get(), e);
                continue;
            }
            free_slots.pop_back();
            extracting[R.get()] = Extract{slot, ev};
            R->stage = STAGE_EXTRACT;
            issued = true;
        }
        // completion
        for (size_t i = 0; i < active.size();) {
            Request *R = active[i].get();
            bool finished = false;
            int32_t status = RR_OK;
            if (R->stage == STAGE_FAILED) {
                finished = true;
                status = R->fail != RR_OK ? R->fail : RR_EDEVICE;
            } else if (R->stage == STAGE_EXTRACT) {
                auto it = extracting.find(R);
                const hipError_t q = hipEventQuery(it->second.ev);
                if (q != hipErrorNotReady) {
                    if (q == hipSuccess) {
                        R->rank = rank_from_extracted(pinned + (size_t)it->second.slot * EXTRACT_WORDS, R->b, p);
                        R->stage = STAGE_DONE;
                    } else {
                        set_error("completion event: %s", hipGetErrorString(q));
                        status = RR_EDEVICE;
                    }
                    (void)hipEventDestroy(it->second.ev);
                    free_slots.push_back(it->second.slot);
                    extracting.erase(it);
                    finished = true;
                }
            }
            if (finished) {
                R->t_done = now_ms();
                publish(R, status, R->budget_bytes);
                active.erase(active.begin() + i);
            } else {
                i++;
            }
        }
        // the device holds enough queued work (or is idle): wait a little
        if (!issued && !kq.retire() && !pq.retire()) std::this_thread::yield();
    }
}

using namespace rr;

extern "C" {

rr_status rr_server_create(uint32_t p, uint64_t hbm_budget_bytes, rr_server **out) {
    return rr::guarded([&]() -> rr_status {
        clear_error();
        RR_CHECK(out, RR_EINVAL, "null argument");
        RR_CHECK(p > 2 && p < (1u << 31) && h_is_prime(p), RR_EINVAL, "p=%u must be a prime in (2, 2^31)", p);
        std::unique_ptr<rr_server> S(new rr_server());
        S->p = p;
        S->budget = hbm_budget_bytes;
        RR_HIP(hipGetDevice(&S->device));
        RR_HIP(hipDeviceGetAttribute(&S->cus, hipDeviceAttributeMultiprocessorCount, S->device));
        RR_HIP(rr::stream_create(&S->st_k));
        RR_HIP(rr::stream_create(&S->st_pm));
        rr_status s;
        if ((s = S->ticks.init(S->st_k, S->cus)) != RR_OK) return s;
        if ((s = S->pm.init(p, S->st_pm, SERVER_MAX_LOGN, 128, 64)) != RR_OK) return s;
        // RR_SERVER_FAIL_AT=tick:<id> or pm:<id> (tests of failure handling): the
        // ticks or PM-basis rounds holding that request fail before queueing work
        if (const char *f = getenv("RR_SERVER_FAIL_AT")) {
            const uint64_t id = strtoull(strchr(f, ':') ? strchr(f, ':') + 1 : f, nullptr, 10);
            if (!strncmp(f, "tick:", 5)) S->ticks.fail_id = id;
            if (!strncmp(f, "pm:", 3)) S->pm.fail_tag = id;
        }
        RR_HIP(hipHostMalloc((void **)&S->pinned, SERVER_EXTRACT_SLOTS * EXTRACT_WORDS * 4, hipHostMallocDefault));
        for (int i = 0; i < SERVER_EXTRACT_SLOTS; i++) S->free_slots.push_back(i);
        for (int w = 0; w < SERVER_SETUP_THREADS; w++) {
            hipStream_t ws;
            RR_HIP(rr::stream_create(&ws));
            S->setup_st.push_back(ws);
        }
        rr_server *raw = S.release();
        for (int w = 0; w < SERVER_SETUP_THREADS; w++) raw->setup.emplace_back([raw, w] { raw->setup_loop(w); });
        raw->worker = std::thread([raw] { raw->run(); });
        *out = raw;
        return RR_OK;
    });
}

rr_status rr_server_submit(rr_server *S, const rr_csr *A, uint32_t b, const uint32_t *row_precond,
                           const uint32_t *col_precond, const uint32_t *V, uint64_t seq_len, uint64_t order,
                           uint64_t *id_out) {
    return rr::guarded([&]() -> rr_status {
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
        RR_CHECK(q->d + 1 <= (1ull << SERVER_MAX_LOGN), RR_EINVAL,
                 "order %llu above the server's 2^%u-point transforms (rr_rank takes it)", (unsigned long long)q->d,
                 SERVER_MAX_LOGN);
        q->rp.assign(A->row_ptr, A->row_ptr + A->m + 1);
        q->col.assign(A->col, A->col + nnz);
        q->val.assign(A->val, A->val + nnz);
        q->dr.assign(row_precond, row_precond + A->m);
        q->dc.assign(col_precond, col_precond + A->n);
        q->V.assign(V, V + A->n * b);
        q->t_arrive = now_ms();
        std::lock_guard<std::mutex> g(S->mu);
        RR_CHECK(!S->broken, RR_EDEVICE, "the server's worker could not start");
        q->id = S->next_id++;
        *id_out = q->id;
        S->queue.push_back(std::move(q));
        S->cv.notify_all();
        return RR_OK;
    });
}

rr_status rr_server_poll(rr_server *S, rr_result *out, uint32_t cap, uint32_t *n_out, int wait_ms) {
    return rr::guarded([&]() -> rr_status {
        RR_CHECK(S && n_out && (out || cap == 0), RR_EINVAL, "null argument");
        std::unique_lock<std::mutex> lk(S->mu);
        if (S->results.empty() && wait_ms > 0)
            S->cv.wait_for(lk, std::chrono::milliseconds(wait_ms), [S] { return !S->results.empty(); });
        uint32_t n = (uint32_t)std::min<size_t>(cap, S->results.size());
        for (uint32_t i = 0; i < n; i++) out[i] = S->results[i];
        S->results.erase(S->results.begin(), S->results.begin() + n);
        *n_out = n;
        return RR_OK;
    });
}

rr_status rr_server_stats_get(rr_server *S, rr_server_stats *out) {
    return rr::guarded([&]() -> rr_status {
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
    });
}

void rr_server_destroy(rr_server *S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
        S->queue.clear();
        S->cv.notify_all();
        S->cv_setup.notify_all();
    }
    if (S->worker.joinable()) S->worker.join();
    for (auto &t : S->setup)
        if (t.joinable()) t.join();
    for (auto ws : S->setup_st) {
        (void)hipStreamSynchronize(ws);
        rr::stream_destroy(ws);
    }
    for (auto *dq : {&S->to_setup, &S->set_up})
        for (auto &st : *dq)
            if (st->ev) (void)hipEventDestroy(st->ev);
    S->to_setup.clear();
    S->set_up.clear();
    if (S->st_k) (void)hipStreamSynchronize(S->st_k);
    if (S->st_pm) (void)hipStreamSynchronize(S->st_pm);
    for (auto &kv : S->extracting) (void)hipEventDestroy(kv.second.ev);
    S->extracting.clear();
    S->active.clear();
    if (S->pinned) (void)hipHostFree(S->pinned);
    hipStream_t sk = S->st_k, sp = S->st_pm;
    delete S;  // its tick runner and scheduler still use the streams while they go
    rr::stream_destroy(sk);
    rr::stream_destroy(sp);
}

}  // extern "C"
