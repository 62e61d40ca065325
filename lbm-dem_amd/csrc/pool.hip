// Run pool: many independent impact runs on one GPU (batched and served workloads).
//
// Memory: one device allocation of the HBM budget, carved first-fit into one block per run
// (the run's context places all its buffers in its block, so admission is one host-side
// bookkeeping step and release never synchronises the device). A run is admitted, in arrival
// order, when its block fits and fewer than maxConcurrent runs are active.
//
// Stepping: a run first advances alone on its own stream (its start-up steps: the impactor wait,
// the tile tuner, the first LB step). Runs past their start-up then advance together in a batch
// per storage precision (lbdem_batch: every kernel launched once for all members, step pairs
// replayed from a captured graph), in chunks of stepsPerChunk steps with at most two chunks in
// flight. A batch is re-formed only when it has no chunk in flight: when members finish or are
// cancelled, or when runs are waiting to join (a quarter of the batch's size, or one that has
// waited 50 ms). A run with a single step left, or with the other step parity, takes that step
// alone (a batch may hold a single run). The batches' argument tables live in a pool-owned
// region, so re-forming a batch frees nothing on the device.
//
// Threads: the stepper thread admits runs, enqueues chunks and polls their events; HELPERS helper
// threads build contexts (host-side node typing, uploads) and finish runs (result download,
// release), so neither stalls the other's work. Every context is used by one thread at a time.
// Runs share nothing on the device: a run's result is the one it has alone.
#include "lbdem_internal.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t BLOCK_ALIGN = 4096;
constexpr int IN_FLIGHT = 2;
constexpr size_t BATCH_REGION = size_t(64) << 20;  // argument tables of the batches (all precisions)
constexpr double JOIN_WAIT_S = 0.05;                // longest a ready run waits for a re-formed batch
constexpr int HELPERS = 4;                          // threads building and finishing runs

size_t round_block(size_t b) { return (std::max<size_t>(b, 1) + BLOCK_ALIGN - 1) & ~(BLOCK_ALIGN - 1); }

// first-fit blocks over [0, size) with coalescing of neighbouring free blocks
class Arena {
  public:
    void reset(size_t size) {
        free_.clear();
        free_[0] = size;
        used_ = 0;
    }
    bool alloc(size_t bytes, size_t& off) {
        bytes = round_block(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const size_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            used_ += bytes;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = round_block(bytes);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }

  private:
    std::map<size_t, size_t> free_;
    size_t used_ = 0;
};

struct Chunk {
    hipEvent_t ev;
    int steps;
};

struct Run {
    int id = 0;
    lbdem_params params{};
    std::vector<int> size;
    std::vector<double> r, x0, objR, objX0, objX1;
    int nsteps = 0;
    size_t bytes = 0, off = 0;
    lbdem_ctx* ctx = nullptr;
    int enqueued = 0;
    std::atomic<int> done{0};  // written by the stepper, read by lbdem_pool_poll
    bool cancel = false, failed = false;
    int error = 0;
    std::string message;
    std::deque<Chunk> inflight;
    lbdem_run_result res{};
    // stepper-owned
    bool batched = false;  // stepped by a batch
    bool solo = false;     // a batch could not take it: stepped alone to the end
    double readySince = -1.0;
};

// a batch of runs of one storage precision (stepper-owned)
struct PoolBatch {
    lbdem_batch* b = nullptr;
    std::vector<Run*> members;
    std::deque<Chunk> inflight;
    int tuningSize = 0;  // members when tuning started (0: not tuning)
};

}  // namespace

struct lbdem_pool {
    int device = 0;
    int maxConcurrent = 1;
    int chunk = 32;
    bool batching = true;
    int batchTile = 0;  // walk of the batches (0: tuned on the batches themselves)
    // stepper-owned, per storage precision: the tuned walk and the size of the batch it was tuned
    // on (a batch at least twice that size is tuned again, until one of 8 runs or more was: the
    // best walk changes with the batch's size while it is small)
    int tunedTile[3] = {0, 0, 0}, tunedSize[3] = {0, 0, 0};
    size_t budget = 0;
    char* base = nullptr;
    char* batchRegion = nullptr;
    Arena arena;
    Clock::time_point t0;

    std::mutex mu;
    std::condition_variable cvStep, cvHelp, cvDone;
    std::vector<std::unique_ptr<Run>> runs;  // by id
    std::deque<int> queue;                   // waiting for admission, in arrival order
    std::deque<int> toInit, toFinish;        // handed to the helper thread
    std::vector<int> stepping;               // runs the stepper advances
    int active = 0;                          // admitted and not finished
    bool stop = false;
    lbdem_pool_stats stats{};
    std::string err;

    std::thread stepper;
    std::vector<std::thread> helpers;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void set_message(Run& r, const std::string& m) {
    std::strncpy(r.res.message, m.c_str(), sizeof(r.res.message) - 1);
    r.res.message[sizeof(r.res.message) - 1] = 0;
}

std::string ctx_error(lbdem_ctx* c) {
    const char* m = lbdem_last_error(c);
    return m ? std::string(m) : std::string();
}

// helper thread: context of an admitted run
void init_run(lbdem_pool* P, Run& r) {
    int rc = lbdem_create(&r.params, &r.ctx);
    if (!rc) rc = lbdem::ctx_use_region(r.ctx, P->base + r.off, r.bytes);
    if (!rc) rc = lbdem_set_particles(r.ctx, int(r.r.size()), r.size.data(), r.r.data(), r.x0.data());
    if (!rc) rc = lbdem_set_objects(r.ctx, int(r.objR.size()), r.objR.data(), r.objX0.data(), r.objX1.data());
    if (!rc) rc = lbdem_init(r.ctx);
    if (!rc && P->batching) rc = lbdem_use_graphs(r.ctx, 0);  // alone only for start-up steps
    // A pinned walk (lbdem_pool_set_batch_tile) is every step's, batched or not. Otherwise a batch
    // runs every member with the batch's z-walk (tuned on the batch itself), so a run's own tile
    // tuner would time start-up steps for nothing: runs that will be batched take their start-up
    // steps with the 8-node walk and can join a batch after them.
    if (!rc && (P->batchTile || P->batching)) rc = lbdem_set_tile(r.ctx, P->batchTile ? P->batchTile : 8);
    if (rc) {
        r.failed = true;
        r.error = rc;
        r.message = "initialisation: " + ctx_error(r.ctx);
    }
}

// helper thread: results of a run whose chunks have all completed (or that failed/was cancelled)
void finish_run(Run& r) {
    if (r.ctx) {
        const int rc = lbdem_synchronize(r.ctx);  // waits for the stream, reports device error flags
        if (rc && !r.failed) {
            r.failed = true;
            r.error = rc;
            r.message = ctx_error(r.ctx);
        }
        if (!r.failed && !r.cancel) {
            lbdem_info info{};
            int e = lbdem_get_objects(r.ctx, r.res.objects);
            if (!e) e = lbdem_total_mass(r.ctx, &r.res.totalMass);
            if (!e) e = lbdem_get_info(r.ctx, &info);
            if (e) {
                r.failed = true;
                r.error = e;
                r.message = "result download: " + ctx_error(r.ctx);
            } else {
                r.res.nActive = info.nActive;
                r.res.nodeSteps = info.totNodes * int64_t(r.done.load());
            }
        }
        (void)lbdem_device_bytes(r.ctx, &r.res.deviceBytes);
        lbdem_tuning tu{};
        if (lbdem_get_tuning(r.ctx, &tu) == 0) r.res.tileZ = tu.tile[2];
        lbdem_destroy(r.ctx);  // synchronises its stream before anything is released
        r.ctx = nullptr;
    }
    for (const Chunk& ch : r.inflight) (void)hipEventDestroy(ch.ev);  // stream drained above
    r.inflight.clear();
}

void helper_loop(lbdem_pool* P) {
    (void)hipSetDevice(P->device);
    std::unique_lock<std::mutex> lk(P->mu);
    for (;;) {
        P->cvHelp.wait(lk, [&] { return !P->toFinish.empty() || !P->toInit.empty() || (P->stop && P->active == 0); });
        if (!P->toFinish.empty()) {  // first: finishing frees memory for admissions
            Run& r = *P->runs[P->toFinish.front()];
            P->toFinish.pop_front();
            lk.unlock();
            finish_run(r);
            lk.lock();
            P->arena.release(r.off, r.bytes);
            --P->active;
            r.res.stepsDone = r.done.load();
            r.res.tDone = P->now();
            if (r.failed) {
                r.res.state = LBDEM_RUN_FAILED;
                r.res.error = r.error ? r.error : LBDEM_E_DEVICE;
                set_message(r, r.message);
                ++P->stats.failed;
            } else if (r.cancel) {
                r.res.state = LBDEM_RUN_CANCELLED;
                set_message(r, "cancelled");
                ++P->stats.cancelled;
            } else {
                r.res.state = LBDEM_RUN_DONE;
                ++P->stats.done;
            }
            P->cvDone.notify_all();
            P->cvStep.notify_all();
            continue;
        }
        if (!P->toInit.empty()) {
            Run& r = *P->runs[P->toInit.front()];
            P->toInit.pop_front();
            lk.unlock();
            init_run(P, r);
            lk.lock();
            if (r.failed || r.cancel) {
                P->toFinish.push_back(r.id);
            } else {
                P->stepping.push_back(r.id);
                P->cvStep.notify_all();
            }
            continue;
        }
        if (P->stop && P->active == 0) break;
    }
}

// events of chunks: reused, created on demand (stepper thread only)
struct EventPool {
    std::vector<hipEvent_t> spare;
    hipEvent_t get() {
        if (!spare.empty()) {
            hipEvent_t e = spare.back();
            spare.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        return hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess ? e : nullptr;
    }
    void put(hipEvent_t e) { spare.push_back(e); }
    ~EventPool() {
        for (hipEvent_t e : spare) (void)hipEventDestroy(e);
    }
};

void fail_run(Run& r, int code, const std::string& m) {
    if (r.failed) return;
    r.failed = true;
    r.error = code;
    r.message = m;
}

// completed chunks at the front of a queue; returns the steps they held (-1 on a device error)
int poll_chunks(std::deque<Chunk>& q, EventPool& ev, std::string& err) {
    int steps = 0;
    while (!q.empty()) {
        const hipError_t e = hipEventQuery(q.front().ev);
        if (e == hipErrorNotReady) break;
        if (e != hipSuccess) {
            err = std::string("HIP error: ") + hipGetErrorString(e);
            return -1;
        }
        steps += q.front().steps;
        ev.put(q.front().ev);
        q.pop_front();
    }
    return steps;
}

// stepper thread: admission, chunk enqueueing and completion polling
void stepper_loop(lbdem_pool* P) {
    (void)hipSetDevice(P->device);
    EventPool ev;
    std::vector<Run*> mine;
    PoolBatch batches[3];  // by storage precision
    int batchBuilds = 0, batchedPeak = 0;
    std::unique_lock<std::mutex> lk(P->mu);
    for (;;) {
        // admission in arrival order: the head waits until its block fits
        while (!P->queue.empty() && P->active < P->maxConcurrent) {
            Run& r = *P->runs[P->queue.front()];
            size_t off = 0;
            if (!P->arena.alloc(r.bytes, off)) break;
            P->queue.pop_front();
            r.off = off;
            r.res.state = LBDEM_RUN_ACTIVE;
            r.res.tStart = P->now();
            ++P->active;
            P->stats.activePeak = std::max(P->stats.activePeak, P->active);
            P->stats.bytesPeak = std::max(P->stats.bytesPeak, P->arena.used());
            P->toInit.push_back(r.id);
            P->cvHelp.notify_all();
        }
        if (P->stop && P->active == 0 && P->queue.empty()) break;
        mine.clear();
        for (int id : P->stepping) mine.push_back(P->runs[id].get());
        std::vector<char> cancelled(mine.size());
        for (size_t k = 0; k < mine.size(); ++k) cancelled[k] = mine[k]->cancel;
        P->stats.batchBuilds = batchBuilds;
        P->stats.batchedPeak = batchedPeak;
        lk.unlock();
        const double now = P->now();

        bool progress = false;
        std::vector<int> finished;
        // completed chunks of runs stepping alone, then of the batches
        for (Run* r : mine) {
            std::string err;
            const int k = poll_chunks(r->inflight, ev, err);
            if (k < 0) fail_run(*r, LBDEM_E_DEVICE, err);
            else if (k) { r->done.fetch_add(k); progress = true; }
        }
        for (PoolBatch& B : batches) {
            std::string err;
            const int k = poll_chunks(B.inflight, ev, err);
            if (k < 0)
                for (Run* m : B.members) fail_run(*m, LBDEM_E_DEVICE, err);
            else if (k) {
                for (Run* m : B.members) m->done.fetch_add(k);
                progress = true;
            }
        }
        auto is_cancelled = [&](Run* r) {
            for (size_t k = 0; k < mine.size(); ++k)
                if (mine[k] == r) return bool(cancelled[k]);
            return false;
        };
        // runs ready to join a batch: start-up over, idle, two or more steps to go
        auto ready = [&](Run* r) {
            return !r->batched && !r->solo && !r->failed && !is_cancelled(r) && r->inflight.empty() &&
                   r->nsteps - r->enqueued >= 2 && lbdem::ctx_steady(r->ctx);
        };
        bool reform[3] = {false, false, false};  // a batch waits for its chunks, then is re-formed
        for (int pr = 0; P->batching && pr < 3; ++pr) {
            PoolBatch& B = batches[pr];
            std::vector<Run*> waiting;
            for (Run* r : mine)
                if (lbdem::ctx_precision(r->ctx) == pr && ready(r)) {
                    if (r->readySince < 0) r->readySince = now;
                    waiting.push_back(r);
                }
            bool leave = false;
            for (Run* m : B.members)
                leave = leave || m->failed || is_cancelled(m) || m->nsteps - m->enqueued < 2;
            double oldest = 1e300;
            for (Run* r : waiting) oldest = std::min(oldest, r->readySince);
            const bool join = !waiting.empty() &&
                              (waiting.size() * 4 >= std::max<size_t>(B.members.size(), 1) || now - oldest > JOIN_WAIT_S ||
                               B.members.empty());
            if (!(leave || join)) continue;
            reform[pr] = true;
            if (!B.inflight.empty()) continue;  // re-formed once its chunks are done (no new ones go out)
            // members and joiners at one step parity (the batch's, or the first joiner's)
            std::vector<Run*> keep;
            for (Run* m : B.members)
                if (!m->failed && !is_cancelled(m) && m->nsteps - m->enqueued >= 2) keep.push_back(m);
            const int phase = !keep.empty() ? lbdem::ctx_phase(keep[0]->ctx)
                                            : (!waiting.empty() ? lbdem::ctx_phase(waiting[0]->ctx) : -1);
            for (Run* r : waiting)
                if (lbdem::ctx_phase(r->ctx) == phase) keep.push_back(r);
            if (B.b) lbdem_batch_destroy(B.b);
            B.b = nullptr;
            for (Run* m : B.members) m->batched = false;
            B.members.clear();
            if (!keep.empty()) {
                std::vector<lbdem_ctx*> ctxs;
                for (Run* r : keep) ctxs.push_back(r->ctx);
                if (lbdem_batch_create(ctxs.data(), int(ctxs.size()), &B.b) == 0) {
                    const size_t slice = (BATCH_REGION / 3) & ~size_t(255);  // tables are read as aligned records
                    lbdem::batch_use_region(B.b, P->batchRegion + pr * slice, slice);
                    // the batch's walk: fixed, tuned on a batch of comparable size, or tuned now
                    const int nm = int(keep.size());
                    int tz = P->batchTile;
                    if (!tz && P->tunedTile[pr] && (nm < 2 * P->tunedSize[pr] || P->tunedSize[pr] >= 8))
                        tz = P->tunedTile[pr];
                    B.tuningSize = tz ? 0 : nm;
                    (void)lbdem_batch_set_tile(B.b, tz);
                    B.members = keep;
                    for (Run* r : keep) { r->batched = true; r->readySince = -1.0; }
                    ++batchBuilds;
                    batchedPeak = std::max(batchedPeak, int(keep.size()));
                } else {
                    B.b = nullptr;
                    for (Run* r : keep) r->solo = true;  // step alone
                }
            }
            progress = true;
        }
        // batch chunks: none while the batch is due to be re-formed
        for (int pr = 0; pr < 3; ++pr) {
            PoolBatch& B = batches[pr];
            if (!B.b || reform[pr]) continue;
            while (int(B.inflight.size()) < IN_FLIGHT) {
                int n = P->chunk;
                for (Run* m : B.members) n = std::min(n, m->nsteps - m->enqueued);
                n -= n % 2;
                if (n < 2) break;
                hipEvent_t e = ev.get();
                int rc = e ? lbdem_batch_cycle(B.b, n) : LBDEM_E_DEVICE;
                if (!rc && hipEventRecord(e, lbdem::batch_stream(B.b)) != hipSuccess) rc = LBDEM_E_DEVICE;
                if (rc) {
                    if (e) ev.put(e);
                    const std::string m = lbdem_batch_last_error(B.b);
                    for (Run* r : B.members) fail_run(*r, rc, "batch: " + m);
                    break;
                }
                B.inflight.push_back({e, n});
                for (Run* m : B.members) m->enqueued += n;
                progress = true;
                int tz = 0, done = 0;
                if (B.tuningSize && lbdem_batch_get_tile(B.b, &tz, &done) == 0 && done) {
                    P->tunedTile[pr] = tz;
                    P->tunedSize[pr] = B.tuningSize;
                    B.tuningSize = 0;
                }
            }
        }
        // runs stepping alone: start-up steps, a parity step, the last single step, or everything
        // when batching is off
        for (Run* rp : mine) {
            Run& r = *rp;
            const bool cancel = is_cancelled(rp);
            if (r.batched) continue;
            while (!r.failed && !cancel && int(r.inflight.size()) < IN_FLIGHT && r.enqueued < r.nsteps) {
                int n = std::min(P->chunk, r.nsteps - r.enqueued);
                if (P->batching && !r.solo) {
                    if (lbdem::ctx_steady(r.ctx) && r.inflight.empty()) {
                        if (r.nsteps - r.enqueued >= 2) {
                            // wants a batch; a run at the other parity than the batch takes one step
                            const PoolBatch& B = batches[lbdem::ctx_precision(r.ctx)];
                            const bool parityOff = !B.members.empty() &&
                                                   lbdem::ctx_phase(B.members[0]->ctx) != lbdem::ctx_phase(r.ctx);
                            if (!parityOff) break;
                            n = 1;
                        }
                    } else if (lbdem::ctx_steady(r.ctx)) {
                        break;  // wait for its chunks, then join
                    } else {
                        n = std::min(n, 12);  // start-up steps in small chunks
                    }
                }
                hipEvent_t e = ev.get();
                int rc = e ? lbdem_cycle(r.ctx, n) : LBDEM_E_DEVICE;
                if (!rc && hipEventRecord(e, lbdem::ctx_stream(r.ctx)) != hipSuccess) rc = LBDEM_E_DEVICE;
                if (rc) {
                    if (e) ev.put(e);
                    fail_run(r, rc, ctx_error(r.ctx));
                    break;
                }
                r.inflight.push_back({e, n});
                r.enqueued += n;
                progress = true;
            }
            if (r.failed || (r.inflight.empty() && (cancel || r.done >= r.nsteps))) finished.push_back(r.id);
        }

        lk.lock();
        for (int id : finished) {
            P->stepping.erase(std::find(P->stepping.begin(), P->stepping.end(), id));
            P->toFinish.push_back(id);
        }
        if (!finished.empty()) P->cvHelp.notify_all();
        if (!progress && finished.empty()) {
            // nothing moved: sleep until a submission, an admission or a short poll interval
            P->cvStep.wait_for(lk, std::chrono::microseconds(P->stepping.empty() ? 20000 : 100));
        }
    }
    lk.unlock();
    for (PoolBatch& B : batches) {
        if (B.b) lbdem_batch_destroy(B.b);
        for (const Chunk& c : B.inflight) (void)hipEventDestroy(c.ev);
    }
}

}  // namespace

extern "C" int lbdem_pool_create(size_t budget, int maxConcurrent, int stepsPerChunk, int device, lbdem_pool** out) {
    if (!out || maxConcurrent < 1 || stepsPerChunk < 0) return LBDEM_E_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return LBDEM_E_DEVICE;
    if (budget == 0) {
        size_t freeB = 0, totalB = 0;
        if (hipMemGetInfo(&freeB, &totalB) != hipSuccess) return LBDEM_E_DEVICE;
        budget = size_t(double(freeB) * 0.85);
    }
    budget &= ~(BLOCK_ALIGN - 1);
    void* base = nullptr;
    if (budget == 0 || hipMalloc(&base, budget) != hipSuccess) {
        (void)hipGetLastError();
        return LBDEM_E_NOMEM;
    }
    lbdem_pool* P = new lbdem_pool();
    P->device = device;
    P->maxConcurrent = maxConcurrent;
    P->chunk = stepsPerChunk ? stepsPerChunk : 32;
    P->budget = budget;
    P->base = static_cast<char*>(base);
    P->arena.reset(budget);
    P->stats.budget = budget;
    P->t0 = Clock::now();
    void* region = nullptr;
    if (hipMalloc(&region, BATCH_REGION) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(base);
        delete P;
        return LBDEM_E_NOMEM;
    }
    P->batchRegion = static_cast<char*>(region);
    if (const char* e = std::getenv("LBDEM_POOL_BATCH")) P->batching = std::atoi(e) != 0;
    P->stepper = std::thread(stepper_loop, P);
    for (int h = 0; h < HELPERS; ++h) P->helpers.emplace_back(helper_loop, P);
    *out = P;
    return 0;
}

extern "C" int lbdem_pool_submit(lbdem_pool* P, const lbdem_run_desc* d, int* runId) {
    if (!P || !d || !runId) return LBDEM_E_ARG;
    const int np = d->nParticles, no = d->nObjects;
    if (np < 0 || no < 0 || d->nsteps < 0 || (np > 0 && (!d->r || !d->x0)) || (no > 0 && (!d->objR || !d->objX0 || !d->objX1))) {
        std::lock_guard<std::mutex> g(P->mu);
        P->err = "invalid run description";
        return LBDEM_E_ARG;
    }
    if (no > 2) {
        std::lock_guard<std::mutex> g(P->mu);
        P->err = "at most 2 objects per run";
        return LBDEM_E_ARG;
    }
    auto r = std::make_unique<Run>();
    r->params = d->params;
    r->params.device = P->device;
    r->size.assign(size_t(np), 1);
    if (d->size) r->size.assign(d->size, d->size + np);
    r->r.assign(d->r, d->r + np);
    r->x0.assign(d->x0, d->x0 + 3 * size_t(np));
    r->objR.assign(d->objR, d->objR + no);
    r->objX0.assign(d->objX0, d->objX0 + 3 * size_t(no));
    r->objX1.assign(d->objX1, d->objX1 + 3 * size_t(no));
    r->nsteps = d->nsteps;
    size_t bytes = 0;
    const int rc = lbdem_estimate_bytes(&r->params, np, r->r.data(), no, r->objR.data(), &bytes);
    std::lock_guard<std::mutex> g(P->mu);
    if (rc) {
        P->err = "run parameters rejected (code " + std::to_string(rc) + ")";
        return rc;
    }
    if (round_block(bytes) > P->budget) {
        P->err = "run needs " + std::to_string(bytes) + " device bytes, above the pool budget";
        return LBDEM_E_CAPACITY;
    }
    if (P->stop) {
        P->err = "pool is shutting down";
        return LBDEM_E_STATE;
    }
    r->bytes = bytes;
    r->id = int(P->runs.size());
    r->res.state = LBDEM_RUN_QUEUED;
    r->res.nObjects = no;
    r->res.tSubmit = P->now();
    *runId = r->id;
    P->queue.push_back(r->id);
    P->runs.push_back(std::move(r));
    ++P->stats.submitted;
    P->cvStep.notify_all();
    return 0;
}

static bool finished_state(int s) { return s == LBDEM_RUN_DONE || s == LBDEM_RUN_FAILED || s == LBDEM_RUN_CANCELLED; }

extern "C" int lbdem_pool_poll(lbdem_pool* P, int id, lbdem_run_result* out) {
    if (!P || !out) return LBDEM_E_ARG;
    std::lock_guard<std::mutex> g(P->mu);
    if (id < 0 || id >= int(P->runs.size())) return LBDEM_E_ARG;
    const Run& r = *P->runs[id];
    *out = r.res;
    if (!finished_state(r.res.state)) out->stepsDone = r.done.load();  // advisory while running
    return 0;
}

extern "C" int lbdem_pool_wait(lbdem_pool* P, int id, double timeout, lbdem_run_result* out) {
    if (!P) return LBDEM_E_ARG;
    {
        std::unique_lock<std::mutex> lk(P->mu);
        if (id < 0 || id >= int(P->runs.size())) return LBDEM_E_ARG;
        const Run& r = *P->runs[id];
        const auto pred = [&] { return finished_state(r.res.state); };
        if (!P->cvDone.wait_for(lk, std::chrono::duration<double>(timeout), pred)) {
            P->err = "timed out waiting for run " + std::to_string(id);
            return LBDEM_E_STATE;
        }
    }
    return out ? lbdem_pool_poll(P, id, out) : 0;
}

extern "C" int lbdem_pool_drain(lbdem_pool* P, double timeout) {
    if (!P) return LBDEM_E_ARG;
    std::unique_lock<std::mutex> lk(P->mu);
    const auto pred = [&] {
        for (const auto& r : P->runs)
            if (!finished_state(r->res.state)) return false;
        return true;
    };
    if (!P->cvDone.wait_for(lk, std::chrono::duration<double>(timeout), pred)) {
        P->err = "timed out draining the pool";
        return LBDEM_E_STATE;
    }
    return 0;
}

extern "C" int lbdem_pool_set_batch_tile(lbdem_pool* P, int tileZ) {
    if (!P) return LBDEM_E_ARG;
    if (tileZ != 0 && tileZ != 1 && tileZ != 2 && tileZ != 4 && tileZ != 8) {
        std::lock_guard<std::mutex> lk(P->mu);
        P->err = "tile z-walk length must be 0 (tune), 1, 2, 4 or 8";
        return LBDEM_E_ARG;
    }
    std::lock_guard<std::mutex> lk(P->mu);
    if (P->stats.submitted > 0) {
        P->err = "set the batch tile before the first submission";
        return LBDEM_E_STATE;
    }
    P->batchTile = tileZ;
    return 0;
}

extern "C" int lbdem_pool_cancel(lbdem_pool* P, int id) {
    if (!P) return LBDEM_E_ARG;
    std::lock_guard<std::mutex> g(P->mu);
    if (id < 0 || id >= int(P->runs.size())) return LBDEM_E_ARG;
    Run& r = *P->runs[id];
    if (finished_state(r.res.state)) return 0;
    r.cancel = true;
    auto it = std::find(P->queue.begin(), P->queue.end(), id);
    if (it != P->queue.end()) {  // never admitted: nothing on the device
        P->queue.erase(it);
        r.res.state = LBDEM_RUN_CANCELLED;
        r.res.tDone = P->now();
        set_message(r, "cancelled");
        ++P->stats.cancelled;
        P->cvDone.notify_all();
    }
    P->cvStep.notify_all();
    return 0;
}

extern "C" int lbdem_pool_get_stats(lbdem_pool* P, lbdem_pool_stats* s) {
    if (!P || !s) return LBDEM_E_ARG;
    std::lock_guard<std::mutex> g(P->mu);
    *s = P->stats;
    s->queued = int(P->queue.size());
    s->active = P->active;
    s->bytesInUse = P->arena.used();
    return 0;
}

extern "C" double lbdem_pool_clock(const lbdem_pool* P) { return P ? P->now() : 0.0; }

// the message is copied under the pool's lock into storage of the calling thread (valid until the
// thread's next call)
extern "C" const char* lbdem_pool_last_error(const lbdem_pool* P) {
    if (!P) return "null pool";
    thread_local std::string copy;
    std::lock_guard<std::mutex> g(const_cast<lbdem_pool*>(P)->mu);
    copy = P->err;
    return copy.c_str();
}

extern "C" int lbdem_pool_destroy(lbdem_pool* P) {
    if (!P) return 0;
    {
        std::lock_guard<std::mutex> g(P->mu);
        P->stop = true;
        for (int id : P->queue) {
            Run& r = *P->runs[id];
            r.cancel = true;
            r.res.state = LBDEM_RUN_CANCELLED;
            ++P->stats.cancelled;
        }
        P->queue.clear();
        for (auto& r : P->runs) r->cancel = true;
        P->cvStep.notify_all();
        P->cvHelp.notify_all();
        P->cvDone.notify_all();
    }
    P->stepper.join();
    {
        std::lock_guard<std::mutex> g(P->mu);
        P->cvHelp.notify_all();
    }
    for (std::thread& h : P->helpers) h.join();
    (void)hipSetDevice(P->device);
    (void)hipFree(P->batchRegion);
    (void)hipFree(P->base);
    delete P;
    return 0;
}

extern "C" int lbdem_pool_set_batching(lbdem_pool* P, int enable) {
    if (!P) return LBDEM_E_ARG;
    std::lock_guard<std::mutex> g(P->mu);
    if (!P->runs.empty()) {
        P->err = "batching is set before the first submission";
        return LBDEM_E_STATE;
    }
    P->batching = enable != 0;
    return 0;
}
