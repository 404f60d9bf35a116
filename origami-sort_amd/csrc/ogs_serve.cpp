// ogs_serve.cpp -- continuous-batching sort server (ogs_server_* in include/ogs.h).
//
// The reference is a library call with no request loop; this is the serving
// form of its sort entry points (Origami/sorter.h:3851, 4194).  Requests are
// ORDER-BY-style sorts of device or host arrays of any supported key type.
//
//   submit ──► pending FIFO ──► scheduler thread ──► batch on a free stream ──► the stream's waiter
//                               (admission by HBM      (ogs_sort_batch over       (event wait, host
//                                footprint, pooled      the batch's requests)      copy-back, release)
//                                device arena)
//
// Admission: a batch takes pending requests of one key type in FIFO order
// while their footprint fits the arena (device request: one scratch copy of
// the payload; host request: payload staging + scratch = 2x) and the batch
// caps allow.  The key type is the oldest request's, unless a batch of that
// type is already on the GPU and a request of a type with none waits among
// the next 64: then that type goes first, so the types overlap on the GPU
// while requests of a running type gather into its next batch.  Requests
// longer than coalesce_max run alone as plain sorts.  Batches run on
// `streams` HIP streams, one batch per stream at a time, and each stream has
// its own waiter thread, so a stream is free again as soon as its own batch
// ends (not when an older batch on another stream does); requests that arrive
// while others are on the GPU are admitted as soon as a stream and arena space
// free up, and their kernels overlap with the running batches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "ogs_internal.hpp"

namespace ogs {

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over one device block (offsets in bytes, 256-aligned).
class Arena {
  public:
    void init(size_t cap) {
        cap_ = cap;
        free_.clear();
        free_[0] = cap;
    }
    // returns offset or SIZE_MAX
    size_t alloc(size_t size) {
        size = (size + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= size) {
                const size_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > size) free_[off + size] = len - size;
                used_[off] = size;
                return off;
            }
        }
        return SIZE_MAX;
    }
    void release(size_t off) {
        auto u = used_.find(off);
        if (u == used_.end()) return;
        size_t start = off, len = u->second;
        used_.erase(u);
        auto next = free_.lower_bound(start);
        if (next != free_.end() && start + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = len;
    }
    size_t capacity() const { return cap_; }

  private:
    size_t cap_ = 0;
    std::map<size_t, size_t> free_, used_;
};

struct Request {
    uint64_t ticket;
    ogs_key_t kt;
    const void *src;
    void *dst;
    uint64_t n;
    int flags;
    int status = OGS_OK;
    bool done = false;
    double t_submit = 0, t_start = 0, t_done = 0;
    size_t dev_off = 0;  // staging offset within the batch block (host requests)
};

struct Batch {
    std::vector<Request *> reqs;
    size_t arena_off = SIZE_MAX;
    int stream = -1;
    hipEvent_t ev = nullptr;
    int status = OGS_OK;
    bool issued = false;  // launches done and the event recorded (its waiter may wait on it)
};

}  // namespace

struct Server {
    Ctx *ctx = nullptr;
    ogs_server_config cfg{};
    void *arena_base = nullptr;
    Arena arena;
    std::vector<hipStream_t> streams;
    std::vector<Batch *> slot;  // per stream: its batch on the GPU, or nullptr
    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_slot;
    std::deque<Request *> pending;
    std::map<uint64_t, Request *> reqs;
    std::map<int, int> running;  // key type -> batches of it on the GPU
    std::thread sched;
    std::vector<std::thread> waiters;
    bool stopping = false;     // no new submissions; drain what is queued
    bool sched_done = false;   // the scheduler has issued its last batch
    uint64_t releases = 0;     // arena releases so far (admission retries on change)
    uint64_t next_ticket = 1;
    uint64_t coalesce_max = 1u << 22;  // longest request that is batched with others
    double t0 = 0;
    ogs_server_stats stats{};

    static size_t footprint(const Request *r) {
        size_t w = ogs_item_size(r->kt);
        size_t one = ((size_t)r->n * w + 255) & ~(size_t)255;
        return (r->flags & OGS_SUBMIT_HOST) ? 2 * one : one;
    }

    int free_stream() const {
        for (size_t i = 0; i < slot.size(); i++)
            if (!slot[i]) return (int)i;
        return -1;
    }
    void schedule_loop();
    void wait_loop(int si);
    void finish(Batch *b);
};

void Server::schedule_loop() {
    (void)hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_sched.wait(lk, [&] {
            if (stopping && pending.empty()) return true;
            if (pending.empty()) return false;
            return free_stream() >= 0;
        });
        if (pending.empty()) {
            if (stopping) break;
            continue;
        }
        const int si = free_stream();
        // the batch's key type: the oldest request's, unless that type is on
        // the GPU and one that is not waits close behind
        Request *head = pending.front();
        if (running[head->kt] > 0) {
            size_t k = 0;
            for (auto it = pending.begin(); it != pending.end() && k < 64; ++it, ++k)
                if (running[(*it)->kt] == 0) {
                    head = *it;
                    break;
                }
        }
        if (footprint(head) > arena.capacity()) {
            pending.erase(std::find(pending.begin(), pending.end(), head));
            head->status = fail(OGS_ERR_NOMEM, "request of %llu items exceeds the server's HBM budget",
                                (unsigned long long)head->n);
            head->done = true;
            head->t_start = head->t_done = now_s() - t0;
            stats.rejected++;
            cv_done.notify_all();
            continue;
        }
        std::vector<Request *> take;
        size_t bytes = 0;
        uint64_t items = 0;
        size_t scanned = 0;
        // long requests run alone (a batch finishes when its last request
        // does, so coalescing them would only delay the short ones)
        const bool alone = head->n > coalesce_max;
        if (alone) {
            head->dev_off = 0;
            take.push_back(head);
            bytes = footprint(head);
            items = head->n;
        }
        for (auto it = pending.begin(); !alone && it != pending.end() && scanned < 256; ++it, ++scanned) {
            Request *r = *it;
            if (r->kt != head->kt || r->n > coalesce_max) continue;
            const size_t f = footprint(r);
            if (!take.empty() && (take.size() >= cfg.max_batch_requests || items + r->n > cfg.max_batch_items ||
                                  bytes + f > arena.capacity()))
                break;
            r->dev_off = bytes;
            take.push_back(r);
            bytes += f;
            items += r->n;
        }
        size_t off = arena.alloc(bytes);
        while (off == SIZE_MAX && take.size() > 1) {  // shrink the batch to what fits now
            bytes -= footprint(take.back());
            items -= take.back()->n;
            take.pop_back();
            off = arena.alloc(bytes);
        }
        if (off == SIZE_MAX) {
            // wait until a running batch releases arena space
            const uint64_t seen = releases;
            cv_sched.wait(lk, [&] { return releases != seen; });
            continue;
        }
        for (Request *r : take) pending.erase(std::find(pending.begin(), pending.end(), r));
        Batch *b = new Batch();
        b->reqs = take;
        b->arena_off = off;
        b->stream = si;
        slot[si] = b;
        running[head->kt]++;
        const double ts = now_s() - t0;
        for (Request *r : take) r->t_start = ts;
        lk.unlock();

        // issue: host staging uploads, the batched sort, the completion event
        hipStream_t s = streams[si];
        char *base = (char *)arena_base + off;
        std::vector<ogs_batch_item> items_v;
        int rc = OGS_OK;
        for (Request *r : take) {
            const size_t w = ogs_item_size(r->kt);
            const size_t one = ((size_t)r->n * w + 255) & ~(size_t)255;
            char *blk = base + r->dev_off;
            if (r->flags & OGS_SUBMIT_HOST) {
                hipError_t e = hipMemcpyAsync(blk, r->src, (size_t)r->n * w, hipMemcpyHostToDevice, s);
                if (e != hipSuccess && rc == OGS_OK) rc = hip_fail(e, "request upload");
                items_v.push_back({blk, blk, blk + one, r->n});
            } else {
                items_v.push_back({r->src, r->dst, blk, r->n});
            }
        }
        if (rc == OGS_OK && alone) {
            // a long request runs as one plain sort: its tile and pass shapes
            // are the plain ones, and no per-tile descriptors are built
            const ogs_batch_item &it = items_v[0];
            void *res = nullptr;
            rc = ogs_sort_keep(reinterpret_cast<ogs_ctx *>(ctx), head->kt, it.in, it.out, it.scratch, it.n, &res,
                               (void *)s);
            if (rc == OGS_OK && res != it.out) {
                hipError_t ec = hipMemcpyAsync(it.out, res, (size_t)it.n * ogs_item_size(head->kt),
                                               hipMemcpyDeviceToDevice, s);
                if (ec != hipSuccess) rc = hip_fail(ec, "result copy");
            }
        } else if (rc == OGS_OK) {
            rc = ogs_sort_batch(reinterpret_cast<ogs_ctx *>(ctx), head->kt, items_v.data(), items_v.size(), (void *)s);
        }
        hipError_t e = hipEventCreateWithFlags(&b->ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(b->ev, s);
        if (e != hipSuccess && rc == OGS_OK) rc = hip_fail(e, "batch event");
        b->status = rc;

        lk.lock();
        stats.batches++;
        stats.batched_requests += take.size();
        b->issued = true;
        cv_slot.notify_all();
    }
}

void Server::finish(Batch *b) {
    // host requests: copy results back (outside the lock)
    if (b->status == OGS_OK) {
        for (Request *r : b->reqs) {
            if (!(r->flags & OGS_SUBMIT_HOST)) continue;
            const size_t w = ogs_item_size(r->kt);
            hipError_t e = hipMemcpy(r->dst, (char *)arena_base + b->arena_off + r->dev_off, (size_t)r->n * w,
                                     hipMemcpyDeviceToHost);
            if (e != hipSuccess) r->status = hip_fail(e, "request download");
        }
    }
    const double td = now_s() - t0;
    std::lock_guard<std::mutex> g(mu);
    for (Request *r : b->reqs) {
        if (b->status != OGS_OK) r->status = b->status;
        r->t_done = td;
        r->done = true;
        stats.completed++;
        stats.items += r->n;
    }
    arena.release(b->arena_off);
    releases++;
    running[b->reqs.front()->kt]--;
    slot[b->stream] = nullptr;
    if (b->ev) (void)hipEventDestroy(b->ev);
    delete b;
    cv_done.notify_all();
    cv_sched.notify_all();
}

// Stream si's waiter: waits for each batch issued on the stream, then
// finishes it (results copied back, requests done, stream and arena freed).
void Server::wait_loop(int si) {
    (void)hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_slot.wait(lk, [&] { return (slot[si] && slot[si]->issued) || sched_done; });
        if (!slot[si] || !slot[si]->issued) break;  // sched_done and nothing left on this stream
        Batch *b = slot[si];
        lk.unlock();
        if (b->ev) (void)hipEventSynchronize(b->ev);
        finish(b);
        lk.lock();
    }
}

}  // namespace ogs

using namespace ogs;

extern "C" {

int ogs_server_create(ogs_server **srv, ogs_ctx *ctx, const ogs_server_config *cfg) {
    if (!srv || !ctx || !cfg) return fail(OGS_ERR_ARG, "NULL argument");
    if (cfg->hbm_budget_bytes < 4096 || cfg->streams < 1 || cfg->streams > 16)
        return fail(OGS_ERR_ARG, "bad server config (budget >= 4096 B, 1..16 streams)");
    Server *s = new Server();
    s->ctx = reinterpret_cast<Ctx *>(ctx);
    s->cfg = *cfg;
    if (!s->cfg.max_batch_requests) s->cfg.max_batch_requests = 1024;
    if (!s->cfg.max_batch_items) s->cfg.max_batch_items = UINT64_MAX;
    if (cfg->coalesce_max_items) s->coalesce_max = cfg->coalesce_max_items;
    hipError_t e = hipSetDevice(s->ctx->device);
    if (e == hipSuccess) e = hipMalloc(&s->arena_base, cfg->hbm_budget_bytes);
    if (e != hipSuccess) {
        delete s;
        return fail(OGS_ERR_NOMEM, "server arena of %llu B: %s", (unsigned long long)cfg->hbm_budget_bytes,
                    hipGetErrorString(e));
    }
    s->arena.init(cfg->hbm_budget_bytes);
    s->streams.resize(cfg->streams);
    s->slot.assign(cfg->streams, nullptr);
    // Each stream's working memory (sample arrays and split tables of its
    // sorts, batch descriptor tables, co-rank tables) is reserved now, at
    // budget / 8 per stream within [64 MiB, 1 GiB]: growing it while batches
    // run would free the old block, and hipFree waits for the whole device.
    const size_t pre = std::min<size_t>(std::max<size_t>(cfg->hbm_budget_bytes / 8, 64u << 20), 1u << 30);
    int rc = OGS_OK;
    size_t made = 0;
    for (; made < s->streams.size() && rc == OGS_OK; made++) {
        hipStream_t &st = s->streams[made];
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) {
            rc = hip_fail(e, "hipStreamCreate");
            break;
        }
        Workspace *w = s->ctx->workspace(st);
        rc = w->reserve(pre);
        if (rc == OGS_OK) rc = w->reserve_part(16u << 20);
    }
    if (rc != OGS_OK) {  // undo: the streams made so far, their workspaces, the arena
        for (size_t i = 0; i < made; i++) {
            s->ctx->drop_workspace(s->streams[i]);
            (void)hipStreamDestroy(s->streams[i]);
        }
        (void)hipFree(s->arena_base);
        delete s;
        return rc;
    }
    s->t0 = now_s();
    s->sched = std::thread([s] { s->schedule_loop(); });
    for (int i = 0; i < cfg->streams; i++) s->waiters.emplace_back([s, i] { s->wait_loop(i); });
    *srv = reinterpret_cast<ogs_server *>(s);
    return OGS_OK;
}

int ogs_server_destroy(ogs_server *srv) {
    if (!srv) return OGS_OK;
    Server *s = reinterpret_cast<Server *>(srv);
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stopping = true;
    }
    s->cv_sched.notify_all();
    s->sched.join();
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->sched_done = true;
    }
    s->cv_slot.notify_all();
    for (auto &w : s->waiters) w.join();
    for (auto &st : s->streams) {
        (void)hipStreamSynchronize(st);
        s->ctx->drop_workspace(st);
        (void)hipStreamDestroy(st);
    }
    (void)hipFree(s->arena_base);
    for (auto &kv : s->reqs) delete kv.second;
    delete s;
    return OGS_OK;
}

int ogs_server_submit(ogs_server *srv, ogs_key_t kt, const void *src, void *dst, uint64_t n, int flags,
                      uint64_t *ticket) {
    if (!srv || !ticket) return fail(OGS_ERR_ARG, "NULL argument");
    if (!ogs_item_size(kt)) return fail(OGS_ERR_ARG, "unknown key type %d", (int)kt);
    if (n && (!src || !dst)) return fail(OGS_ERR_ARG, "NULL src or dst");
    Server *s = reinterpret_cast<Server *>(srv);
    Request *r = new Request();
    r->kt = kt;
    r->src = src;
    r->dst = dst;
    r->n = n;
    r->flags = flags;
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stopping) {
        delete r;
        return fail(OGS_ERR_STATE, "server is shutting down");
    }
    r->ticket = s->next_ticket++;
    r->t_submit = now_s() - s->t0;
    s->reqs[r->ticket] = r;
    s->stats.submitted++;
    if (n == 0) {
        r->done = true;
        r->t_start = r->t_done = r->t_submit;
        s->stats.completed++;
    } else {
        s->pending.push_back(r);
        s->cv_sched.notify_all();
    }
    *ticket = r->ticket;
    return OGS_OK;
}

int ogs_server_wait(ogs_server *srv, uint64_t ticket, double timeout_s) {
    if (!srv) return fail(OGS_ERR_ARG, "NULL server");
    Server *s = reinterpret_cast<Server *>(srv);
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(OGS_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    Request *r = it->second;
    auto pred = [&] { return r->done; };
    if (timeout_s < 0) s->cv_done.wait(lk, pred);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
        return fail(OGS_ERR_TIMEOUT, "ticket %llu not done after %.3f s", (unsigned long long)ticket, timeout_s);
    return r->status;
}

int ogs_server_poll(ogs_server *srv, uint64_t ticket, int *done) {
    if (!srv || !done) return fail(OGS_ERR_ARG, "NULL argument");
    Server *s = reinterpret_cast<Server *>(srv);
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(OGS_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    *done = it->second->done ? 1 : 0;
    return it->second->done ? it->second->status : OGS_OK;
}

int ogs_server_times(ogs_server *srv, uint64_t ticket, double *t_submit, double *t_start, double *t_done) {
    if (!srv) return fail(OGS_ERR_ARG, "NULL server");
    Server *s = reinterpret_cast<Server *>(srv);
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(OGS_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    if (t_submit) *t_submit = it->second->t_submit;
    if (t_start) *t_start = it->second->t_start;
    if (t_done) *t_done = it->second->t_done;
    return OGS_OK;
}

int ogs_server_release(ogs_server *srv, uint64_t ticket) {
    if (!srv) return fail(OGS_ERR_ARG, "NULL server");
    Server *s = reinterpret_cast<Server *>(srv);
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return fail(OGS_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    if (!it->second->done) return fail(OGS_ERR_STATE, "ticket %llu still running", (unsigned long long)ticket);
    delete it->second;
    s->reqs.erase(it);
    return OGS_OK;
}

int ogs_server_get_stats(ogs_server *srv, ogs_server_stats *out) {
    if (!srv || !out) return fail(OGS_ERR_ARG, "NULL argument");
    Server *s = reinterpret_cast<Server *>(srv);
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->stats;
    return OGS_OK;
}

}  // extern "C"
