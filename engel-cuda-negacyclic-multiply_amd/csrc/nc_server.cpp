// Serving loop (include/negacyclic_batch.h): continuous batching of single
// product requests of mixed ring degree over resident per-degree plans.
//
// One scheduler thread owns the device work.  Per degree it keeps a queue of
// waiting requests and `lanes_per_degree` lanes; a lane is one in-flight
// batch with its own stream and event.  Every batch's device memory
// (operands, results, residue workspace) is one block of a single device
// arena of at most hbm_budget bytes, allocated at creation: a request is
// admitted only when its batch's block fits in the arena, so admitted
// requests never hold more device memory than the budget.  Each turn the
// scheduler (1) moves newly submitted requests to their degree's queue,
// (2) retires lanes whose event has fired (requests complete, their block
// returns to the arena), and (3) fills every idle lane with waiting requests
// of its degree, as many as max_batch and the arena allow: the operands in,
// the batched pipeline, each result out followed by its own event, all on
// the lane's stream, so transfers of one batch overlap kernels of others.
// Requests below kMapMaxBytes per buffer whose buffers are mapped pinned
// memory move by the GPU's own loads and stores (one gather launch for all
// operands of a batch, one scatter launch per result); the others by
// hipMemcpyAsync.  Requests that arrive while lanes are busy are admitted as soon as
// a lane and arena space free up (the next turn), not after the queue
// drains.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <iterator>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/negacyclic_batch.h"
#include "nc_internal.hpp"

namespace {

using nc::set_error;
using Clock = std::chrono::steady_clock;

double now_s() { return std::chrono::duration<double>(Clock::now().time_since_epoch()).count(); }

uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

// Buffers below this size move by the transfer kernels, larger ones by
// hipMemcpyAsync.  DMA copies pay a fixed ~13 us each (240 KB copies reach
// ~13 GB/s) but take no compute units and reach 45-53 GB/s from 2 MB up,
// where transfer kernels would compete with the pipeline's kernels for
// compute-unit slots.  On MI355X, serving degrees 2^11..2^16 at 6000
// requests/s: this split gave p50 0.37 ms and p99 2.3-3.1 ms against 0.53-0.64
// and 7.5-9.8 ms with copies only and 0.64-0.67 and 3.8-4.7 ms with kernels
// only, and held p99 near 4 ms at 8000/s, where either alone reached 18-41 ms.
constexpr uint64_t kMapMaxBytes = 1u << 20;

// Device-visible address of [p, p + bytes) if it lies in one mapped pinned
// host allocation and is 8-byte aligned (what the transfer kernels need),
// else null.  HIP maps pinned memory at its host address; anything else
// (pageable memory, device memory, another mapping) takes the copy path.
// One allocation: the first and the last word are both mapped pinned memory
// with the same range start (the range size attribute holds only 32 bits:
// it reads 0 for a 24 GiB allocation).
const void* mapped(const void* p, uint64_t bytes) {
    if (reinterpret_cast<uintptr_t>(p) % 8 || bytes % 8 || bytes == 0) return nullptr;
    void* start[2] = {nullptr, nullptr};
    for (int k = 0; k < 2; ++k) {
        const char* x = static_cast<const char*>(p) + (k ? bytes - 8 : 0);
        hipPointerAttribute_t at{};
        if (hipPointerGetAttributes(&at, x) != hipSuccess || at.type != hipMemoryTypeHost || at.devicePointer != x ||
            hipPointerGetAttribute(&start[k], HIP_POINTER_ATTRIBUTE_RANGE_START_ADDR, const_cast<char*>(x)) != hipSuccess)
            return nullptr;
    }
    return start[0] && start[0] == start[1] ? p : nullptr;
}

struct Req {
    uint64_t id = 0;
    uint32_t N = 0;
    const uint32_t* a = nullptr;
    const uint32_t* b = nullptr;
    uint32_t* c = nullptr;
    // device-visible addresses of a, b, c when all three are mapped pinned
    // memory (the transfer kernels move them), else null (copies do)
    const void* da = nullptr;
    const void* db = nullptr;
    void* dc = nullptr;
    uint64_t bytes = 0;
    double t_submit = 0, t_done = 0;
    bool done = false;
    int status = NC_OK;
};

// The device memory of admitted batches: one allocation, handed out as
// first-fit blocks; a freed block merges with free neighbours.  Used by the
// scheduler thread only.
struct Arena {
    char* base = nullptr;
    uint64_t size = 0, used = 0;
    std::map<uint64_t, uint64_t> free_blocks;  // offset -> bytes

    void reset(char* b, uint64_t n) {
        base = b;
        size = n;
        used = 0;
        free_blocks.clear();
        if (n) free_blocks[0] = n;
    }
    bool alloc(uint64_t bytes, uint64_t* off) {
        for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const uint64_t rest = it->second - bytes, at = it->first + bytes;
            free_blocks.erase(it);
            if (rest) free_blocks[at] = rest;
            used += bytes;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        if (!bytes) return;
        used -= bytes;
        auto next = free_blocks.lower_bound(off);
        if (next != free_blocks.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_blocks.erase(next);
        }
        if (next != free_blocks.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_blocks[off] = bytes;
    }
};

struct Lane {
    hipStream_t s = nullptr;
    // one event per request, recorded after its own result copy: a request
    // completes when its result is home, not when its batch's last one is
    std::vector<hipEvent_t> req_ev;
    size_t done_upto = 0;  // requests of the batch already completed (they complete in order)
    std::vector<Req*> batch;
    std::vector<const void*> xsrc;  // gather lists, kept to spare allocations
    std::vector<void*> xdst;
    uint64_t off = 0, bytes = 0;  // the batch's arena block
    double t_enqueue = 0;         // when the batch's work was enqueued
    bool busy = false;
};

struct Degree {
    uint32_t N = 0;
    nc_plan* plan = nullptr;
    uint64_t words = 0;      // u32 limbs of one operand
    uint64_t ws = 0;         // residue bytes of one operand
    uint64_t footprint = 0;  // block bytes of one request
    uint32_t cap = 0;
    std::deque<Req*> queue;
    std::vector<Lane> lanes;

    // Block of n products: a | b | c (n N L u32 each, product-major), then
    // the residue rows of a and of b (n ws bytes each), every part 256-byte
    // aligned; block_bytes(n) <= n block_bytes(1).
    uint64_t block_bytes(uint32_t n) const { return 3 * align256(n * words * 4) + 2 * align256(n * ws); }
    struct Parts {
        uint32_t *a, *b, *c;
        void *ws_a, *ws_b;
    };
    Parts parts(char* block, uint32_t n) const {
        const uint64_t io = align256(n * words * 4), w = align256(n * ws);
        return Parts{reinterpret_cast<uint32_t*>(block), reinterpret_cast<uint32_t*>(block + io),
                     reinterpret_cast<uint32_t*>(block + 2 * io), block + 3 * io, block + 3 * io + w};
    }
};

}  // namespace

struct nc_server {
    nc_server_config cfg{};
    std::vector<uint32_t> q;
    int device = 0;
    std::map<uint32_t, Degree> deg;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> inbox;                  // under mu
    std::unordered_map<uint64_t, Req*> reqs;  // under mu
    uint64_t next_id = 1;                     // under mu
    bool stop = false;                        // under mu
    bool ready = false;                       // under mu: lanes warmed (or failed)
    int warm_rc = NC_OK;                      // under mu
    nc_server_stats st{};                     // under mu
    nc_tableset* tables = nullptr;            // one table set for every served degree
    Arena arena;                              // scheduler thread only (after creation)
    void* arena_mem = nullptr;
    std::thread th;
};

namespace {

int hip_err(hipError_t e, const char* what) { return set_error(NC_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e)); }

// Completes the batch's requests [done_upto, upto) with `status`; once all
// are complete the lane's arena block returns and the lane goes idle.
void complete(nc_server* sv, Lane& ln, size_t upto, int status) {
    if (upto <= ln.done_upto) return;
    const double t = now_s();
    {
        std::lock_guard<std::mutex> g(sv->mu);
        for (size_t i = ln.done_upto; i < upto; ++i) {
            Req* r = ln.batch[i];
            r->t_done = t;
            r->status = status;
            r->done = true;
        }
        sv->st.completed += upto - ln.done_upto;
        if (upto == ln.batch.size())
            sv->st.max_batch_us = std::max<uint64_t>(sv->st.max_batch_us, (uint64_t)((t - ln.t_enqueue) * 1e6));
    }
    sv->cv_done.notify_all();
    ln.done_upto = upto;
    if (upto < ln.batch.size()) return;
    sv->arena.release(ln.off, ln.bytes);
    ln.bytes = 0;
    ln.batch.clear();
    ln.done_upto = 0;
    ln.busy = false;
}

void finish(nc_server* sv, Lane& ln, int status) { complete(sv, ln, ln.batch.size(), status); }

// Completes the lane's requests whose result copies have landed; returns
// true if any did (or an error ended the batch).
bool poll_lane(nc_server* sv, Lane& ln) {
    size_t upto = ln.done_upto;
    while (upto < ln.batch.size()) {
        const hipError_t q = hipEventQuery(ln.req_ev[upto]);
        if (q == hipErrorNotReady) break;
        if (q != hipSuccess) {
            complete(sv, ln, upto, NC_OK);
            finish(sv, ln, hip_err(q, "lane"));
            return true;
        }
        ++upto;
    }
    if (upto == ln.done_upto) return false;
    complete(sv, ln, upto, NC_OK);
    return true;
}

// Enqueues one batch on an idle lane: the largest batch (up to max_batch)
// whose block fits in the arena now.  Returns false if none fits.
bool dispatch(nc_server* sv, Degree& d, Lane& ln) {
    uint32_t n = (uint32_t)std::min<size_t>(d.queue.size(), d.cap);
    for (; n > 0; --n)
        if (sv->arena.alloc(d.block_bytes(n), &ln.off)) break;
    if (n == 0) return false;
    ln.bytes = d.block_bytes(n);
    for (uint32_t i = 0; i < n; ++i) {
        ln.batch.push_back(d.queue.front());
        d.queue.pop_front();
    }
    const size_t words = d.words;
    const Degree::Parts pt = d.parts(sv->arena.base + ln.off, n);
    ln.t_enqueue = now_s();
    hipError_t e = hipSuccess;
    ln.xsrc.clear();
    ln.xdst.clear();
    uint64_t n_mapped = 0;
    for (uint32_t i = 0; i < n && e == hipSuccess; ++i) {
        const Req* r = ln.batch[i];
        if (r->dc) {
            ln.xsrc.insert(ln.xsrc.end(), {r->da, r->db});
            ln.xdst.insert(ln.xdst.end(), {pt.a + i * words, pt.b + i * words});
            ++n_mapped;
            continue;
        }
        e = hipMemcpyAsync(pt.a + i * words, r->a, words * 4, hipMemcpyHostToDevice, ln.s);
        if (e == hipSuccess) e = hipMemcpyAsync(pt.b + i * words, r->b, words * 4, hipMemcpyHostToDevice, ln.s);
    }
    if (e == hipSuccess && !ln.xsrc.empty())
        e = nc::launch_host_xfer(ln.xsrc.data(), ln.xdst.data(), (uint32_t)ln.xsrc.size(), words * 4, false, ln.s);
    int rc = e == hipSuccess ? NC_OK : hip_err(e, "request operands in");
    if (rc == NC_OK)
        rc = nc::run_pipeline_ws(d.plan, pt.ws_a, pt.ws_b, pt.a, pt.b, pt.c, n, ln.s, nc::choose_for(d.plan, n));
    for (uint32_t i = 0; i < n && rc == NC_OK; ++i) {
        Req* r = ln.batch[i];
        const void* src = pt.c + i * words;
        e = r->dc ? nc::launch_host_xfer(&src, &r->dc, 1, words * 4, true, ln.s)
                  : hipMemcpyAsync(r->c, src, words * 4, hipMemcpyDeviceToHost, ln.s);
        if (e == hipSuccess) e = hipEventRecord(ln.req_ev[i], ln.s);
        if (e != hipSuccess) rc = hip_err(e, "request result out");
    }
    {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->st.max_enqueue_us = std::max<uint64_t>(sv->st.max_enqueue_us, (uint64_t)((now_s() - ln.t_enqueue) * 1e6));
        sv->st.launches += 1;
        sv->st.max_batch_seen = std::max<uint64_t>(sv->st.max_batch_seen, n);
        sv->st.mapped_requests += n_mapped;
        sv->st.peak_bytes = std::max<uint64_t>(sv->st.peak_bytes, sv->arena.used);
    }
    if (rc != NC_OK) {
        (void)hipStreamSynchronize(ln.s);
        finish(sv, ln, rc);
        return true;
    }
    ln.busy = true;
    return true;
}

// Warms every lane before the first request, on the scheduler thread (the
// thread that issues all lane work later): copies and transfer kernels in
// from pinned memory, the pipeline at every batch-size bucket (1, 2, 4, .., max_batch, or the
// largest whose block fits the arena: each bucket's tuning may use kernels
// nothing has launched yet, and a kernel's first launch in a process is
// slow), copy and transfer kernel out, all on the lane's own stream, in arena blocks that
// are returned afterwards.  Measured at 6000 requests/s: without it the
// first ~0.6 s of a process ran behind (p99 ~40 ms, against ~4 ms later);
// with it the first ~0.1 s still does.
int warm_lanes(nc_server* sv) {
    const uint32_t L = sv->cfg.L;
    size_t max_words = 0;
    for (auto& kv : sv->deg) max_words = std::max(max_words, (size_t)kv.first * L);
    uint32_t* h = nullptr;
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&h), 3 * max_words * 4, hipHostMallocDefault);
    if (e != hipSuccess) return hip_err(e, "warm-up buffer");
    std::fill(h, h + 3 * max_words, 0u);
    // the same buffer as the transfer kernels see it (null: they are not warmed)
    uint32_t* hd = static_cast<uint32_t*>(const_cast<void*>(mapped(h, 3 * max_words * 4)));
    int rc = NC_OK;
    for (auto& kv : sv->deg) {
        Degree& d = kv.second;
        const size_t words = d.words;
        uint32_t top = d.cap;  // the arena is all free here
        while (top > 1 && d.block_bytes(top) > sv->arena.size) --top;
        for (Lane& ln : d.lanes) {
            uint64_t off = 0;
            if (!sv->arena.alloc(d.block_bytes(top), &off)) {
                rc = set_error(NC_ERR_CAPACITY, "request arena below one request");
                break;
            }
            char* block = sv->arena.base + off;
            e = hipMemsetAsync(block, 0, d.block_bytes(top), ln.s);  // zero operands are valid (0 < q)
            for (uint32_t n = 1; rc == NC_OK; n = std::min(2 * n, top)) {
                const Degree::Parts pt = d.parts(block, n);
                if (e == hipSuccess) e = hipMemcpyAsync(pt.a, h, words * 4, hipMemcpyHostToDevice, ln.s);
                if (e == hipSuccess) e = hipMemcpyAsync(pt.b, h + max_words, words * 4, hipMemcpyHostToDevice, ln.s);
                if (e == hipSuccess && hd) {
                    const void* src[2] = {hd, hd + max_words};
                    void* dst[2] = {pt.a, pt.b};
                    e = nc::launch_host_xfer(src, dst, 2, words * 4, false, ln.s);
                }
                if (e != hipSuccess) rc = hip_err(e, "warm-up copy in");
                if (rc == NC_OK)
                    rc = nc::run_pipeline_ws(d.plan, pt.ws_a, pt.ws_b, pt.a, pt.b, pt.c, n, ln.s,
                                             nc::choose_for(d.plan, n));
                if (rc == NC_OK) {
                    e = hipMemcpyAsync(h + 2 * max_words, pt.c, words * 4, hipMemcpyDeviceToHost, ln.s);
                    if (e == hipSuccess && hd) {
                        const void* src = pt.c;
                        void* dst = hd + 2 * max_words;
                        e = nc::launch_host_xfer(&src, &dst, 1, words * 4, true, ln.s);
                    }
                    if (e != hipSuccess) rc = hip_err(e, "warm-up copy out");
                }
                if (n == top) break;
            }
            const hipError_t es = hipStreamSynchronize(ln.s);  // before the block returns to the arena
            if (rc == NC_OK && es != hipSuccess) rc = hip_err(es, "warm-up");
            sv->arena.release(off, d.block_bytes(top));
            if (rc != NC_OK) break;
        }
        if (rc != NC_OK) break;
    }
    // no copy may still use h when it is freed (an error above can leave
    // work queued on a lane)
    for (auto& kv : sv->deg)
        for (Lane& ln : kv.second.lanes) (void)hipStreamSynchronize(ln.s);
    (void)hipHostFree(h);
    return rc;
}

void scheduler(nc_server* sv) {
    (void)hipSetDevice(sv->device);
    {
        const int rc = warm_lanes(sv);
        std::lock_guard<std::mutex> g(sv->mu);
        sv->warm_rc = rc;
        sv->ready = true;
    }
    sv->cv_done.notify_all();
    if (sv->warm_rc != NC_OK) return;
    // a gap between turns counts only when work was pending at the end of the
    // earlier turn (an idle scheduler sleeps up to 2 ms on purpose)
    double last_turn = now_s();
    bool was_active = false;
    for (;;) {
        bool any_busy = false, waiting = false;
        const double turn = now_s();
        {
            std::lock_guard<std::mutex> g(sv->mu);
            if (was_active)
                sv->st.max_turn_gap_us =
                    std::max<uint64_t>(sv->st.max_turn_gap_us, (uint64_t)((turn - last_turn) * 1e6));
            while (!sv->inbox.empty()) {
                Req* r = sv->inbox.front();
                sv->inbox.pop_front();
                sv->deg[r->N].queue.push_back(r);
            }
        }
        bool progress = false;
        for (auto& kv : sv->deg)
            for (Lane& ln : kv.second.lanes)
                if (ln.busy && poll_lane(sv, ln)) progress = true;
        for (auto& kv : sv->deg) {
            Degree& d = kv.second;
            for (Lane& ln : d.lanes) {
                if (d.queue.empty()) break;
                if (ln.busy) continue;
                if (dispatch(sv, d, ln)) {
                    progress = true;
                } else {
                    std::lock_guard<std::mutex> g(sv->mu);
                    sv->st.admission_waits += 1;
                    break;
                }
            }
            for (Lane& ln : d.lanes) any_busy = any_busy || ln.busy;
            waiting = waiting || !d.queue.empty();
        }
        last_turn = now_s();
        was_active = any_busy || waiting;
        std::unique_lock<std::mutex> lk(sv->mu);
        if (sv->stop && !any_busy && !waiting && sv->inbox.empty()) break;
        if (!progress && sv->inbox.empty()) {
            // poll in-flight lanes often; sleep longer when idle (a submit wakes us)
            sv->cv_work.wait_for(lk, any_busy ? std::chrono::microseconds(20) : std::chrono::milliseconds(2));
        }
    }
}

void free_degree(Degree& d) {
    for (Lane& ln : d.lanes) {
        if (ln.s) (void)hipStreamSynchronize(ln.s);
        for (hipEvent_t ev : ln.req_ev)
            if (ev) (void)hipEventDestroy(ev);
        if (ln.s) (void)hipStreamDestroy(ln.s);
    }
    if (d.plan) nc_plan_destroy(d.plan);
}

}  // namespace

extern "C" {

void* nc_host_alloc(size_t bytes) {
    void* p = nullptr;
    return hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess ? p : nullptr;
}

void nc_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

void nc_server_destroy(nc_server* sv) {
    if (!sv) return;
    if (sv->th.joinable()) {
        {
            std::lock_guard<std::mutex> g(sv->mu);
            sv->stop = true;
        }
        sv->cv_work.notify_all();
        sv->th.join();
    }
    for (auto& kv : sv->deg) free_degree(kv.second);  // every lane stream is drained first
    nc_tableset_destroy(sv->tables);
    if (sv->arena_mem) (void)hipFree(sv->arena_mem);
    for (auto& kv : sv->reqs) delete kv.second;
    delete sv;
}

int nc_server_create(const nc_server_config* cfg, nc_server** out) {
    if (!out) return set_error(NC_ERR_INVALID, "out is null");
    *out = nullptr;
    if (!cfg || !cfg->q || !cfg->degrees || cfg->n_degrees == 0) return set_error(NC_ERR_INVALID, "bad config");
    if (cfg->L == 0 || cfg->L > 32) return set_error(NC_ERR_INVALID, "L must be in [1, 32]");
    if (cfg->max_batch == 0 || cfg->lanes_per_degree == 0) return set_error(NC_ERR_INVALID, "max_batch and lanes must be >= 1");
    if (cfg->lanes_per_degree > 64) return set_error(NC_ERR_INVALID, "lanes_per_degree above 64");
    nc_server* sv = new (std::nothrow) nc_server();
    if (!sv) return set_error(NC_ERR_NOMEM, "server");
    sv->cfg = *cfg;
    sv->q.assign(cfg->q, cfg->q + cfg->L);
    sv->cfg.q = sv->q.data();
    sv->cfg.degrees = nullptr;
    auto fail_out = [&](int rc) {
        nc_server_destroy(sv);
        return rc;
    };
    hipError_t e = hipGetDevice(&sv->device);
    if (e != hipSuccess) return fail_out(hip_err(e, "hipGetDevice"));
    if (int rc = nc_tableset_create(cfg->L, sv->q.data(), &sv->tables)) return fail_out(rc);
    uint64_t most = 0;  // every lane of every degree holding a full batch at once
    for (uint32_t i = 0; i < cfg->n_degrees; ++i) {
        const uint32_t N = cfg->degrees[i];
        if (sv->deg.count(N)) continue;
        if ((uint64_t)cfg->max_batch * N > (1ull << 31)) return fail_out(set_error(NC_ERR_CAPACITY, "max_batch * N above 2^31"));
        Degree& d = sv->deg[N];
        d.N = N;
        if (int rc = nc_plan_create_shared(sv->tables, N, &d.plan)) return fail_out(rc);
        d.words = (uint64_t)N * cfg->L;
        d.ws = nc::workspace_bytes_per_product(d.plan);
        d.footprint = d.block_bytes(1);
        if (d.footprint > cfg->hbm_budget)
            return fail_out(set_error(NC_ERR_CAPACITY, "hbm_budget below one request of degree " + std::to_string(N)));
        d.cap = cfg->max_batch;
        most += (uint64_t)cfg->lanes_per_degree * d.block_bytes(d.cap);
        if (cfg->autotune) {
            for (uint32_t bsz : {1u, d.cap})
                if (int rc = nc_plan_autotune(d.plan, bsz, nullptr, nullptr, nullptr)) return fail_out(rc);
        }
        d.lanes.resize(cfg->lanes_per_degree);
        for (Lane& ln : d.lanes) {
            e = hipStreamCreateWithFlags(&ln.s, hipStreamNonBlocking);
            ln.req_ev.assign(d.cap, nullptr);
            for (uint32_t i = 0; i < d.cap && e == hipSuccess; ++i)
                e = hipEventCreateWithFlags(&ln.req_ev[i], hipEventDisableTiming);
            if (e != hipSuccess) return fail_out(hip_err(e, "lane streams"));
        }
    }
    // the arena: hbm_budget bytes, or what every lane holding a full batch
    // at once needs, if that is less
    const uint64_t arena = std::min<uint64_t>(cfg->hbm_budget & ~255ull, most);
    e = hipMalloc(&sv->arena_mem, arena);
    if (e != hipSuccess)
        return fail_out(set_error(NC_ERR_NOMEM, "request arena of " + std::to_string(arena) + " bytes: " + hipGetErrorString(e)));
    sv->arena.reset(static_cast<char*>(sv->arena_mem), arena);
    sv->st.device_bytes = arena;
    sv->th = std::thread(scheduler, sv);
    {  // the scheduler thread warms its lanes before taking requests; creation reports how that went
        std::unique_lock<std::mutex> lk(sv->mu);
        sv->cv_done.wait(lk, [&] { return sv->ready; });
    }
    if (sv->warm_rc != NC_OK) return fail_out(sv->warm_rc);
    *out = sv;
    return NC_OK;
}

int nc_server_footprint(const nc_server* sv, uint32_t N, uint64_t* bytes) {
    if (!sv || !bytes) return set_error(NC_ERR_INVALID, "null argument");
    auto it = sv->deg.find(N);
    if (it == sv->deg.end()) return set_error(NC_ERR_INVALID, "degree not served");
    *bytes = it->second.footprint;
    return NC_OK;
}

int nc_server_submit(nc_server* sv, uint32_t N, const uint32_t* a, const uint32_t* b, uint32_t* c, uint64_t* ticket) {
    if (!sv || !a || !b || !c || !ticket) return set_error(NC_ERR_INVALID, "null argument");
    auto it = sv->deg.find(N);  // the degree map is fixed after creation
    if (it == sv->deg.end()) return set_error(NC_ERR_INVALID, "degree " + std::to_string(N) + " not served");
    Req* r = new (std::nothrow) Req();
    if (!r) return set_error(NC_ERR_NOMEM, "request");
    r->N = N;
    r->a = a;
    r->b = b;
    r->c = c;
    const uint64_t io = it->second.words * 4;
    r->da = io < kMapMaxBytes ? mapped(a, io) : nullptr;
    r->db = r->da ? mapped(b, io) : nullptr;
    r->dc = r->db ? const_cast<void*>(mapped(c, io)) : nullptr;
    r->bytes = it->second.footprint;
    r->t_submit = now_s();
    {
        std::lock_guard<std::mutex> g(sv->mu);
        if (sv->stop) {
            delete r;
            return set_error(NC_ERR_STATE, "server is stopping");
        }
        r->id = sv->next_id++;
        sv->reqs[r->id] = r;
        sv->inbox.push_back(r);
        sv->st.submitted += 1;
    }
    sv->cv_work.notify_one();
    *ticket = r->id;
    return NC_OK;
}

int nc_server_wait(nc_server* sv, uint64_t ticket, int timeout_ms, double* latency_s) {
    if (!sv) return set_error(NC_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return set_error(NC_ERR_STATE, "unknown ticket");
    Req* r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) sv->cv_done.wait(lk, ready);
    else if (!sv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return set_error(NC_ERR_STATE, "timed out");
    if (latency_s) *latency_s = r->t_done - r->t_submit;
    const int status = r->status;
    sv->reqs.erase(it);
    delete r;
    return status;
}

int nc_server_get_stats(const nc_server* sv, nc_server_stats* out) {
    if (!sv || !out) return set_error(NC_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(const_cast<nc_server*>(sv)->mu);
    *out = sv->st;
    return NC_OK;
}

int nc_server_reset_stats(nc_server* sv) {
    if (!sv) return set_error(NC_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> g(sv->mu);
    const uint64_t dev = sv->st.device_bytes;
    sv->st = nc_server_stats{};
    sv->st.device_bytes = dev;
    return NC_OK;
}

int nc_server_replay(nc_server* sv, uint32_t n, const uint32_t* degrees, const double* arrival_s,
                     const uint32_t* const* a, const uint32_t* const* b, uint32_t* const* c, double* latency_s,
                     double* makespan_s) {
    if (!sv || !degrees || !arrival_s || !a || !b || !c || !latency_s) return set_error(NC_ERR_INVALID, "null argument");
    std::vector<uint64_t> tickets(n);
    std::vector<double> sched(n);
    const double t0 = now_s();
    int rc = NC_OK;
    uint32_t sent = 0;
    for (; sent < n; ++sent) {
        if (sent && arrival_s[sent] < arrival_s[sent - 1]) {
            rc = set_error(NC_ERR_INVALID, "arrival times must ascend");
            break;
        }
        sched[sent] = t0 + arrival_s[sent];
        while (now_s() < sched[sent]) std::this_thread::yield();
        rc = nc_server_submit(sv, degrees[sent], a[sent], b[sent], c[sent], &tickets[sent]);
        if (rc) break;
    }
    double last = t0;
    for (uint32_t i = 0; i < sent; ++i) {  // every submitted request is waited for, even after an error
        std::unique_lock<std::mutex> lk(sv->mu);
        Req* r = sv->reqs[tickets[i]];
        sv->cv_done.wait(lk, [&] { return r->done; });
        latency_s[i] = r->t_done - sched[i];
        last = std::max(last, r->t_done);
        if (rc == NC_OK && r->status != NC_OK) rc = r->status;
        sv->reqs.erase(tickets[i]);
        delete r;
    }
    if (makespan_s) *makespan_s = last - t0;
    return rc;
}

}  // extern "C"
