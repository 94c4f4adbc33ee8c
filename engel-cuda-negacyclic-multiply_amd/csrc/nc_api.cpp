// C ABI of the negacyclic product library (include/negacyclic.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/negacyclic.h"
#include "nc_internal.hpp"
#include "nc_kernels.hpp"
#include "nc_plan.hpp"

using nc::Choice;
using nc::kMfmaAll;
using nc::kMfmaCrt;
using nc::kMfmaDecompose;

static_assert(nc::kScratchBytes == NC_NOALLOC_SCRATCH_BYTES, "scratch size in the header and the library agree");

namespace {

thread_local std::string g_last_error;

int fail(int status, const std::string& msg) {
    g_last_error = msg;
    return status;
}

int hip_fail(hipError_t e, const char* what) {
    return fail(NC_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define NC_HIP(call, what)                               \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

uint32_t bucket_of(uint32_t batch) {
    uint32_t b = 0;
    while ((1u << b) < batch && b < 31) ++b;
    return b;
}

uint32_t k_of(const nc_plan* p, uint32_t width) {
    return width == 62 ? (uint32_t)p->host.s62.primes.size() : (uint32_t)p->host.s30.primes.size();
}

// Untuned default: 30-bit primes, and the smallest outer split that still
// gives about 1024 workgroup-sized blocks (k * batch * 2^s1): fewer, larger
// blocks cost fewer HBM round trips but must still fill 256 compute units.
uint32_t mfma_default(const nc_plan* p) { return p->mfma_mode < 0 ? kMfmaAll : (uint32_t)p->mfma_mode; }

// Matrix-core paths this plan can run (30-bit set).
uint32_t mfma_available(const nc_plan* p) {
    uint32_t m = 0;
    if (p->dev.w30.mfma_tab && p->host.N >= 32) m |= kMfmaDecompose;
    if (nc::crt_mfma_available(p->dev)) m |= kMfmaCrt;
    return m;
}

// Paths a call with choice c actually runs.
// The last two inverse NTT stages run inside the CRT (split 2, matrix-core CRT).
bool inv_in_crt(const nc_plan* p, const Choice& c) {
    return nc::crt_fuses_outer_inv(p->dev, (int)c.width, (c.mfma & kMfmaCrt) != 0, c.log_blocks);
}

uint32_t mfma_in_use(const nc_plan* p, const Choice& c) { return c.width == 30 ? (c.mfma & mfma_available(p)) : 0u; }

Choice heuristic(const nc_plan* p, uint32_t batch, uint32_t width = 30) {
    const uint32_t log_n = p->host.log_n;
    Choice c;
    c.width = width;
    c.mfma = mfma_default(p);
    if (log_n < 4) return c;
    const uint32_t lo = nc::min_log_blocks(log_n), hi = std::max(lo, nc::max_log_blocks(log_n));
    uint32_t s = lo;
    while (s < hi && (uint64_t)k_of(p, width) * batch * (1ull << s) < 1024) ++s;
    c.log_blocks = s;
    return c;
}

// unchunked: the configuration for callers that run one pipeline on their
// own workspace (graph capture, the server), which never chunk.
Choice choose(const nc_plan* p, uint32_t batch, bool unchunked = false) {
    Choice c;
    const nc_plan::Tuned& t = p->tune[bucket_of(batch)];
    if (p->forced) c = p->forced_choice;
    else if (t.tuned) c = unchunked ? t.unchunked : t.choice;
    else c = heuristic(p, batch);
    if (p->mfma_mode >= 0) c.mfma = (uint32_t)p->mfma_mode;
    if (p->chunks_forced) c.chunks = p->chunks_forced;
    if (unchunked) c.chunks = 1;
    if (p->persist_mode >= 0) c.persist = (uint32_t)p->persist_mode;
    return c;
}

int grow_workspace(nc_plan* p, uint32_t batch) {
    if (batch <= p->ws_cap) return NC_OK;
    NC_HIP(hipDeviceSynchronize(), "workspace sync");
    if (p->d_res_a) (void)hipFree(p->d_res_a);
    if (p->d_res_b) (void)hipFree(p->d_res_b);
    p->d_res_a = p->d_res_b = nullptr;
    p->ws_cap = 0;
    const size_t bytes = (size_t)batch * p->res_bytes;
    NC_HIP(hipMalloc(&p->d_res_a, bytes), "workspace a");
    NC_HIP(hipMalloc(&p->d_res_b, bytes), "workspace b");
    p->ws_cap = batch;
    return NC_OK;
}

int check_batch(const nc_plan* p, uint32_t batch) {
    if (batch == 0) return fail(NC_ERR_INVALID, "batch must be >= 1");
    if ((uint64_t)batch * p->host.N > (1ull << 31)) return fail(NC_ERR_CAPACITY, "batch * N must not exceed 2^31");
    return NC_OK;
}

int check_width(uint32_t width) {
    if (width != 62 && width != 30) return fail(NC_ERR_INVALID, "prime width must be 62 or 30");
    return NC_OK;
}

// Chunked batches: chunk i runs on internal stream i % 2; the second stream
// starts once chunk 0's decomposition is done, so one chunk's VALU-bound NTT
// overlaps another chunk's memory-bound decomposition, outer stages and CRT.
// Chunks use disjoint slices of the workspace and of a, b, c.
int run_chunked(nc_plan* p, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, uint32_t batch, hipStream_t s,
                const Choice& c) {
    const uint32_t n = std::min(c.chunks, batch);
    const uint32_t per = (batch + n - 1) / n;
    const size_t words = (size_t)p->host.N * p->host.L;
    NC_HIP(hipEventRecord(p->ev_fork, s), "fork");
    NC_HIP(hipStreamWaitEvent(p->sub[0], p->ev_fork, 0), "fork");
    NC_HIP(hipStreamWaitEvent(p->sub[1], p->ev_fork, 0), "fork");
    const bool dm = (c.mfma & kMfmaDecompose) != 0;
    const bool fuse = nc::decompose_fuses_outer(p->dev, (int)c.width, dm, c.log_blocks);
    for (uint32_t i = 0; i * per < batch; ++i) {
        const uint32_t off = i * per, cnt = std::min(per, batch - off);
        hipStream_t st = p->sub[i & 1];
        void* wa = static_cast<char*>(p->d_res_a) + (size_t)off * p->res_bytes;
        void* wb = static_cast<char*>(p->d_res_b) + (size_t)off * p->res_bytes;
        if (i == 1) NC_HIP(hipStreamWaitEvent(st, p->ev_first, 0), "stagger");
        NC_HIP(nc::launch_decompose(p->dev, (int)c.width, d_a + off * words, wa, cnt, d_b + off * words, wb, cnt, st,
                                    dm, fuse ? c.log_blocks : 0u),
               "decompose launch");
        if (i == 0) NC_HIP(hipEventRecord(p->ev_first, st), "stagger");
        const bool inv = inv_in_crt(p, c);
        NC_HIP(nc::launch_ntt_product(p->dev, (int)c.width, wa, wb, cnt, c.log_blocks, st, fuse, inv), "ntt launch");
        if (inv)
            NC_HIP(nc::launch_crt_inv2(p->dev, static_cast<const uint32_t*>(wa), d_c + off * words, cnt, st),
                   "crt launch");
        else
            NC_HIP(nc::launch_crt(p->dev, (int)c.width, wa, d_c + off * words, cnt, st, (c.mfma & kMfmaCrt) != 0),
                   "crt launch");
    }
    for (int k = 0; k < 2; ++k) {
        NC_HIP(hipEventRecord(p->ev_join[k], p->sub[k]), "join");
        NC_HIP(hipStreamWaitEvent(s, p->ev_join[k], 0), "join");
    }
    return NC_OK;
}

int run_pipeline(nc_plan* p, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, uint32_t batch, hipStream_t s,
                 const Choice& c) {
    const bool persist = c.persist && nc::persist_available(p->dev, batch);
    if (c.chunks > 1 && batch > 1 && !persist) return run_chunked(p, d_a, d_b, d_c, batch, s, c);
    return nc::run_pipeline_ws(p, p->d_res_a, p->d_res_b, d_a, d_b, d_c, batch, s, c);
}

// Orders every use of the plan's workspace across streams: the stream
// waits for the previous user's work before it enqueues its own (a no-op on
// the same stream), and the caller records ev_ws after enqueuing.  Under
// stream capture neither happens: a captured graph orders its own nodes, and
// its replays are ordered by the stream they are launched on.
bool capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

int ws_acquire(nc_plan* p, hipStream_t s) {
    if (!capturing(s)) NC_HIP(hipStreamWaitEvent(s, p->ev_ws, 0), "workspace order");
    return NC_OK;
}

int ws_release(nc_plan* p, hipStream_t s) {
    if (!capturing(s)) NC_HIP(hipEventRecord(p->ev_ws, s), "workspace order");
    return NC_OK;
}

// After a persistent launch of `batch` products on the plan's workspace
// (stream s): did every XCD run all of its items, with no readiness wait
// reaching its bound?  Reads the control block and the plan's error word
// (clearing it).
bool persist_drained(nc_plan* p, uint32_t batch, hipStream_t s) {
    if (hipStreamSynchronize(s) != hipSuccess) return false;
    const uint32_t nxcd = p->dev.xcds, words = nc::persist_xcd_words(batch, nxcd);
    const size_t rows = (size_t)batch * p->dev.w30.k * p->dev.N * sizeof(uint32_t);
    std::vector<uint32_t> h((size_t)nxcd * words);
    uint32_t err = 0;
    if (hipMemcpy(h.data(), static_cast<char*>(p->d_res_b) + rows, h.size() * sizeof(uint32_t),
                  hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&err, p->d_err, sizeof err, hipMemcpyDeviceToHost) != hipSuccess)
        return false;
    if (err) (void)hipMemset(p->d_err, 0, sizeof err);
    const nc::PersistShape sh = nc::persist_shape_for(p->dev.log_n, p->dev.w30.k);
    bool ok = err == 0;
    for (uint32_t x = 0; x < nxcd; ++x) {
        const uint32_t* xb = h.data() + (size_t)x * words;
        ok = ok && xb[nc::kCtlErr] == 0 && xb[nc::kCtlDone] == nc::persist_products_on(batch, nxcd, x) * sh.R;
    }
    return ok;
}

// ---- persisted autotuner choices ------------------------------------------
// One text line per tuned (device, shape, batch bucket, pinned settings):
//   nc-tune-v3 <arch> <cus> <N> <L> <q_bits> <k62> <k30> <bucket> <mfma_mode>
//   <chunks_forced> <persist_mode> : <width> <log_blocks> <mfma> <chunks>
//   <persist> <width1> <log_blocks1> <mfma1> <persist1> <median_ms>
// (the second group is the unchunked choice).  The file is appended to; the
// last line for a key wins.  v3 lines come from timing on random operands;
// v2 lines (timed on zeros) no longer match any key.  Path: nc_set_tune_cache, else the environment
// variable NC_TUNE_CACHE, else no cache.
std::mutex g_cache_mu;
std::string g_cache_path;
bool g_cache_set = false;

std::string cache_path() {
    std::lock_guard<std::mutex> g(g_cache_mu);
    if (g_cache_set) return g_cache_path;
    const char* env = std::getenv("NC_TUNE_CACHE");
    return env ? std::string(env) : std::string();
}

std::string cache_key(const nc_plan* p, uint32_t batch) {
    if (cache_path().empty()) return std::string();
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, p->device) != hipSuccess) return std::string();
    std::string arch(prop.gcnArchName);
    for (char& ch : arch)
        if (ch == ' ' || ch == ':') ch = '_';
    char buf[256];
    std::snprintf(buf, sizeof buf, "nc-tune-v3 %s %d %u %u %u %u %u %u %d %u %d", arch.c_str(),
                  prop.multiProcessorCount, p->host.N, p->host.L, p->host.q_bits, k_of(p, 62), k_of(p, 30),
                  bucket_of(batch), p->mfma_mode, p->chunks_forced, p->persist_mode);
    return buf;
}

bool valid_choice(const nc_plan* p, const Choice& c) {
    if (c.width != 30 && c.width != 62) return false;
    if (c.chunks < 1 || c.chunks > 64 || c.mfma > kMfmaAll || c.persist > 1) return false;
    if (p->host.log_n >= 4 &&
        (c.log_blocks < nc::min_log_blocks(p->host.log_n) || c.log_blocks > nc::max_log_blocks(p->host.log_n)))
        return false;
    return true;
}

std::mutex g_file_mu;  // this process's reads and appends of the cache file

bool cache_lookup(const nc_plan* p, const std::string& key, nc_plan::Tuned* out) {
    const std::string path = cache_path();
    std::lock_guard<std::mutex> g(g_file_mu);
    FILE* f = path.empty() ? nullptr : std::fopen(path.c_str(), "r");
    if (!f) return false;
    char line[512];
    bool found = false;
    while (std::fgets(line, sizeof line, f)) {
        const char* colon = std::strstr(line, " : ");
        if (!colon || (size_t)(colon - line) != key.size() || std::strncmp(line, key.c_str(), key.size()) != 0) continue;
        nc_plan::Tuned t;
        float med = 0.f;
        if (std::sscanf(colon + 3, "%u %u %u %u %u %u %u %u %u %f", &t.choice.width, &t.choice.log_blocks,
                        &t.choice.mfma, &t.choice.chunks, &t.choice.persist, &t.unchunked.width,
                        &t.unchunked.log_blocks, &t.unchunked.mfma, &t.unchunked.persist, &med) != 10)
            continue;
        t.unchunked.chunks = 1;
        t.median_ms = med;
        if (!valid_choice(p, t.choice) || !valid_choice(p, t.unchunked)) continue;
        *out = t;
        found = true;  // later lines override earlier ones
    }
    std::fclose(f);
    return found;
}

void cache_store(const std::string& key, const nc_plan::Tuned& t) {
    const std::string path = cache_path();
    if (path.empty()) return;
    std::lock_guard<std::mutex> g(g_file_mu);
    char buf[512];
    const int len = std::snprintf(buf, sizeof buf, "%s : %u %u %u %u %u %u %u %u %u %.5f\n", key.c_str(),
                                  t.choice.width, t.choice.log_blocks, t.choice.mfma, t.choice.chunks,
                                  t.choice.persist, t.unchunked.width, t.unchunked.log_blocks, t.unchunked.mfma,
                                  t.unchunked.persist, t.median_ms);
    // one write of the whole line in append mode: lines from concurrent
    // processes do not interleave
    FILE* f = std::fopen(path.c_str(), "a");
    if (!f) return;  // an unwritable cache only costs the next process a retune
    std::fwrite(buf, 1, (size_t)len, f);
    std::fclose(f);
}

// Upload one host table vector at `off` of the table allocation.
template <class V>
hipError_t upload(char* base, size_t off, const V& v) {
    return hipMemcpy(base + off, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice);
}

}  // namespace

namespace nc {

int set_error(int status, const std::string& msg) { return fail(status, msg); }

Choice choose_for(const nc_plan* p, uint32_t batch) { return choose(p, batch, /*unchunked=*/true); }

size_t workspace_bytes_per_product(const nc_plan* p) { return p->res_bytes; }

int run_pipeline_ws(nc_plan* p, void* ws_a, void* ws_b, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c,
                    uint32_t batch, hipStream_t s, const Choice& c) {
    if (c.persist && persist_available(p->dev, batch)) {
        // one launch for the batch; its control block sits after b's
        // residue rows, in the trailer every product's workspace carries
        const size_t rows = (size_t)batch * p->dev.w30.k * p->dev.N * sizeof(uint32_t);
        if (rows + persist_ctl_bytes(p->dev, batch) > (size_t)batch * p->res_bytes)
            return fail(NC_ERR_CAPACITY, "workspace trailer too small for the persistent control block");
        uint32_t* ctl = reinterpret_cast<uint32_t*>(static_cast<char*>(ws_b) + rows);
        NC_HIP(launch_persist30(p->dev, d_a, d_b, d_c, static_cast<uint32_t*>(ws_a), static_cast<uint32_t*>(ws_b),
                                batch, ctl, p->d_err, p->persist_tmo, s),
               "persistent launch");
        return NC_OK;
    }
    const bool dm = (c.mfma & kMfmaDecompose) != 0;
    const bool fuse = decompose_fuses_outer(p->dev, (int)c.width, dm, c.log_blocks);
    NC_HIP(launch_decompose(p->dev, (int)c.width, d_a, ws_a, batch, d_b, ws_b, batch, s, dm, fuse ? c.log_blocks : 0u),
           "decompose launch");
    const bool inv = inv_in_crt(p, c);
    NC_HIP(launch_ntt_product(p->dev, (int)c.width, ws_a, ws_b, batch, c.log_blocks, s, fuse, inv), "ntt launch");
    if (inv)
        NC_HIP(launch_crt_inv2(p->dev, static_cast<const uint32_t*>(ws_a), d_c, batch, s), "crt launch");
    else
        NC_HIP(launch_crt(p->dev, (int)c.width, ws_a, d_c, batch, s, (c.mfma & kMfmaCrt) != 0), "crt launch");
    return NC_OK;
}

int validate_batch(const nc_plan* p, uint32_t batch) { return check_batch(p, batch); }


}  // namespace nc

// One table set for every degree of one modulus (negacyclic_batch.h): the
// master tables of its largest degree, resident on one device, shared by the
// plans created from it (each holds a reference, as does the creator).
struct nc_tableset {
    nc::HostTables host;
    nc::DevPlan dev{};
    int device = 0;
    void* d_tables = nullptr;
    size_t table_bytes = 0;
    float create_ms = 0.f;
    std::atomic<int> refs{1};
};

namespace {

void tableset_release(nc_tableset* ts) {
    if (ts && ts->refs.fetch_sub(1) == 1) {
        if (ts->d_tables) (void)hipFree(ts->d_tables);
        delete ts;
    }
}

}  // namespace

extern "C" {

const char* nc_status_string(int status) {
    switch (status) {
        case NC_OK: return "ok";
        case NC_ERR_INVALID: return "invalid argument";
        case NC_ERR_SIZE: return "a and b must contain N*L limbs";
        case NC_ERR_DEVICE: return "HIP runtime error";
        case NC_ERR_NOMEM: return "out of memory";
        case NC_ERR_PRIMES: return "not enough NTT primes for this modulus";
        case NC_ERR_SELFTEST: return "init self test failed";
        case NC_ERR_CAPACITY: return "capacity exceeded";
        case NC_ERR_STATE: return "invalid state";
        case NC_ERR_UNSUPPORTED: return "switched off in this build";
        default: return "unknown status";
    }
}

const char* nc_last_error(void) { return g_last_error.c_str(); }

void nc_plan_destroy(nc_plan* p) {
    if (!p) return;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    if (p->d_tables) (void)hipFree(p->d_tables);
    if (p->d_res_a) (void)hipFree(p->d_res_a);
    if (p->d_res_b) (void)hipFree(p->d_res_b);
    if (p->d_io) (void)hipFree(p->d_io);
    if (p->d_err) (void)hipFree(p->d_err);
    if (p->h_pin) (void)hipHostFree(p->h_pin);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    for (int k = 0; k < 2; ++k) {
        if (p->sub[k]) (void)hipStreamSynchronize(p->sub[k]);
        if (p->sub[k]) (void)hipStreamDestroy(p->sub[k]);
        if (p->ev_join[k]) (void)hipEventDestroy(p->ev_join[k]);
    }
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    for (hipEvent_t ev : p->ev_piece)
        if (ev) (void)hipEventDestroy(ev);
    if (p->ev_first) (void)hipEventDestroy(p->ev_first);
    if (p->ev_ws) {
        (void)hipEventSynchronize(p->ev_ws);
        (void)hipEventDestroy(p->ev_ws);
    }
    tableset_release(p->shared);
    delete p;
}

namespace {

// Sampled entries of the device-built twiddle tables (both directions,
// both sets: first, middle and last prime; x = 1, 2, N / 2 + 1, N - 1 and
// pseudo-random x < N) against host arithmetic; rows are 2^tw_log entries
// apart (a shared set's rows are longer than N: degree N reads their prefix).
int check_device_twiddles(const nc::HostTables& t, uint32_t tw_log, const char* f62, const char* i62, const char* f30,
                          const char* i30) {
    const uint32_t k62 = (uint32_t)t.s62.primes.size(), k30 = (uint32_t)t.s30.primes.size(), N = t.N;
    uint64_t lcg = 0x9e3779b97f4a7c15ull;
    for (int set = 0; set < 2; ++set) {
        const uint32_t k = set == 0 ? k62 : k30;
        for (uint32_t j : {0u, k / 2, k - 1}) {
            uint32_t xs[9] = {1u % N, 2u % N, (N / 2 + 1) % N, N - 1, 0, 0, 0, 0, 0};
            for (int r = 4; r < 9; ++r) {
                lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
                xs[r] = (uint32_t)(lcg >> 33) & (N - 1);
            }
            for (uint32_t x : xs) {
                for (int inv = 0; inv < 2; ++inv) {
                    const size_t at = ((size_t)j << tw_log) + x;
                    bool same;
                    hipError_t e;
                    if (set == 0) {
                        nc::Tw got{};
                        e = hipMemcpy(&got, (inv ? i62 : f62) + at * sizeof(nc::Tw), sizeof(got), hipMemcpyDeviceToHost);
                        const nc::Tw want = nc::twiddle_entry62(t, j, x, inv != 0);
                        same = got.w == want.w && got.ws == want.ws;
                    } else {
                        nc::Tw32 got{};
                        e = hipMemcpy(&got, (inv ? i30 : f30) + at * sizeof(nc::Tw32), sizeof(got),
                                      hipMemcpyDeviceToHost);
                        const nc::Tw32 want = nc::twiddle_entry30(t, j, x, inv != 0);
                        same = got.w == want.w && got.ws == want.ws;
                    }
                    if (e != hipSuccess) return hip_fail(e, "twiddle table check");
                    if (!same) return fail(NC_ERR_SELFTEST, "self test: device twiddle table");
                }
            }
        }
    }
    return NC_OK;
}

// Host tables -> one resident device allocation: every table of both prime
// sets, the [k][N] twiddle tables built in place on the device and checked
// against host arithmetic; fills the device view d (strides tw_log = log N).
int upload_tables(const nc::HostTables& t, int device, void** d_tables, size_t* bytes, nc::DevPlan& d) {
    const nc::HostSet62& a = t.s62;
    const nc::HostSet30& b = t.s30;
    size_t off = 0;
    auto slot = [&](size_t n) {
        const size_t at = off;
        off = align_up(off + n, 256);
        return at;
    };
    const size_t k62 = a.primes.size(), k30 = b.primes.size();
    const size_t o62_twf = slot(k62 * t.N * sizeof(nc::Tw)), o62_twi = slot(k62 * t.N * sizeof(nc::Tw));
    const size_t o62_pow = slot(a.tw_pow.size() * sizeof(nc::Tw));
    const size_t o62_pc = slot(a.pc.size() * sizeof(nc::PrimeConst)), o62_pl = slot(a.place.size() * 8);
    const size_t o62_cc = slot(a.cc.size() * sizeof(nc::CrtConst)), o62_mq = slot(a.miq.size() * 4);
    const size_t o62_qc = slot(sizeof(nc::QConst));
    const size_t o30_twf = slot(k30 * t.N * sizeof(nc::Tw32)), o30_twi = slot(k30 * t.N * sizeof(nc::Tw32));
    const size_t o30_pow = slot(b.tw_pow.size() * sizeof(nc::Tw32));
    const size_t o30_pc = slot(b.pc.size() * sizeof(nc::PrimeConst32)), o30_pl = slot(b.place.size() * 4);
    const size_t o30_cc = slot(b.cc.size() * sizeof(nc::CrtConst32)), o30_mq = slot(b.miq.size() * 4);
    const size_t o30_qc = slot(sizeof(nc::QConst));
    const size_t o30_mt = slot(b.mfma_tab.size()), o30_mg = slot(b.mfma_bias.size() * sizeof(int32_t));
    const size_t o30_ct = slot(b.crt_tab.size()), o30_cc2 = slot(b.crt_c.size() * sizeof(int64_t));
    const size_t o30_cm = slot(sizeof(nc::CrtMfmaConst));
    *bytes = off;
    NC_HIP(hipMalloc(d_tables, off), "table allocation");
    char* base = static_cast<char*>(*d_tables);
    const hipError_t ups[] = {
        upload(base, o62_pow, a.tw_pow), upload(base, o62_pc, a.pc),
        upload(base, o62_pl, a.place),   upload(base, o62_cc, a.cc),      upload(base, o62_mq, a.miq),
        hipMemcpy(base + o62_qc, &a.qc, sizeof(nc::QConst), hipMemcpyHostToDevice),
        upload(base, o30_pow, b.tw_pow), upload(base, o30_pc, b.pc),
        upload(base, o30_pl, b.place),   upload(base, o30_cc, b.cc),      upload(base, o30_mq, b.miq),
        hipMemcpy(base + o30_qc, &b.qc, sizeof(nc::QConst), hipMemcpyHostToDevice),
        upload(base, o30_mt, b.mfma_tab), upload(base, o30_mg, b.mfma_bias),
        upload(base, o30_ct, b.crt_tab),  upload(base, o30_cc2, b.crt_c),
        hipMemcpy(base + o30_cm, &b.crt_mc, sizeof(nc::CrtMfmaConst), hipMemcpyHostToDevice),
    };
    for (hipError_t u : ups)
        if (u != hipSuccess) return hip_fail(u, "table upload");
    // the [k][N] twiddle tables, built in place on the device
    hipError_t e = nc::launch_twiddles62(reinterpret_cast<nc::Tw*>(base + o62_twf),
                                         reinterpret_cast<nc::Tw*>(base + o62_twi),
                                         reinterpret_cast<const nc::Tw*>(base + o62_pow),
                                         reinterpret_cast<const nc::PrimeConst*>(base + o62_pc), t.log_n,
                                         (uint32_t)k62, nullptr);
    if (e == hipSuccess)
        e = nc::launch_twiddles30(
            reinterpret_cast<nc::Tw32*>(base + o30_twf), reinterpret_cast<nc::Tw32*>(base + o30_twi),
            reinterpret_cast<const nc::Tw32*>(base + o30_pow), reinterpret_cast<const nc::PrimeConst32*>(base + o30_pc),
            t.log_n, (uint32_t)k30, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return hip_fail(e, "twiddle tables");
    if (int rc = check_device_twiddles(t, t.log_n, base + o62_twf, base + o62_twi, base + o30_twf, base + o30_twi))
        return rc;
    d.N = t.N;
    d.L = t.L;
    d.log_n = t.log_n;
    d.w62.k = (uint32_t)a.primes.size();
    d.w62.tw_log = d.w30.tw_log = t.log_n;  // these [k][N] tables
    d.w62.tw_fwd = reinterpret_cast<const nc::Tw*>(base + o62_twf);
    d.w62.tw_inv = reinterpret_cast<const nc::Tw*>(base + o62_twi);
    d.w62.pc = reinterpret_cast<const nc::PrimeConst*>(base + o62_pc);
    d.w62.place = reinterpret_cast<const uint64_t*>(base + o62_pl);
    d.w62.cc = reinterpret_cast<const nc::CrtConst*>(base + o62_cc);
    d.w62.miq = reinterpret_cast<const uint32_t*>(base + o62_mq);
    d.w62.qc = reinterpret_cast<const nc::QConst*>(base + o62_qc);
    d.w30.k = (uint32_t)b.primes.size();
    d.w30.tw_fwd = reinterpret_cast<const nc::Tw32*>(base + o30_twf);
    d.w30.tw_inv = reinterpret_cast<const nc::Tw32*>(base + o30_twi);
    d.w30.pc = reinterpret_cast<const nc::PrimeConst32*>(base + o30_pc);
    d.w30.place = reinterpret_cast<const uint32_t*>(base + o30_pl);
    d.w30.cc = reinterpret_cast<const nc::CrtConst32*>(base + o30_cc);
    d.w30.miq = reinterpret_cast<const uint32_t*>(base + o30_mq);
    d.w30.qc = reinterpret_cast<const nc::QConst*>(base + o30_qc);
    if (!b.mfma_tab.empty()) {
        d.w30.mfma_tab = reinterpret_cast<const int8_t*>(base + o30_mt);
        d.w30.mfma_bias = reinterpret_cast<const int32_t*>(base + o30_mg);
        d.w30.mfma_tiles = (uint32_t)(b.mfma_tab.size() / (32 * nc::kMfmaK));
    }
    if (!b.crt_tab.empty()) {
        d.w30.crt_tab = reinterpret_cast<const int8_t*>(base + o30_ct);
        d.w30.crt_c = reinterpret_cast<const int64_t*>(base + o30_cc2);
        d.w30.crt_mc = reinterpret_cast<const nc::CrtMfmaConst*>(base + o30_cm);
        d.w30.crt_rt = b.crt_rt;
        d.w30.crt_kc = b.crt_kc;
    }
    int cus = 0;
    NC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "device attribute");
    d.cus = cus > 0 ? (uint32_t)cus : 1u;
    int xccs = 0;  // XCDs; unknown (0) leaves the persistent form unavailable
    if (hipDeviceGetAttribute(&xccs, hipDeviceAttributeNumberOfXccs, device) != hipSuccess) {
        (void)hipGetLastError();
        xccs = 0;
    }
    d.xcds = xccs > 0 ? (uint32_t)xccs : 0u;
    return NC_OK;
}

// The rest of a plan once its tables are resident: the scratch address of
// the allocation-free entry point, streams and events, host-path staging and
// a one-product workspace.
int finish_plan(nc_plan* p) {
    const nc::HostTables& t = p->host;
    hipError_t e = nc::scratch_address(&p->scratch);  // this device's scratch array, looked up once per device
    if (e != hipSuccess) return hip_fail(e, "scratch address");
    // residue rows of the wider set, and a trailer for the persistent form's
    // control block
    p->res_bytes = std::max((size_t)p->dev.w62.k * 8, (size_t)p->dev.w30.k * 4) * t.N + nc::kPersistTrailer;
    e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "stream");
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = hipStreamCreateWithFlags(&p->sub[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_join[k], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_first, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_ws, hipEventDisableTiming);
    for (int i = 0; i < nc_plan::kHostPieces && e == hipSuccess; ++i)
        e = hipEventCreateWithFlags(&p->ev_piece[i], hipEventDisableTiming);
    if (e != hipSuccess) return hip_fail(e, "chunk streams");
    const size_t io = (size_t)t.N * t.L * sizeof(uint32_t);
    NC_HIP(hipMalloc(&p->d_err, 256), "error word");
    NC_HIP(hipMemset(p->d_err, 0, 256), "error word");
    NC_HIP(hipMalloc(&p->d_io, 3 * io), "staging");
    NC_HIP(hipHostMalloc(reinterpret_cast<void**>(&p->h_pin), 3 * io, hipHostMallocDefault), "pinned staging");
    void* hd = nullptr;  // the transfer kernels' view of it; without one, copies move it
    if (hipHostGetDevicePointer(&hd, p->h_pin, 0) == hipSuccess) p->h_pin_dev = static_cast<uint32_t*>(hd);
    else (void)hipGetLastError();
    return grow_workspace(p, 1);
}

// The device half of a table set's initialisation (the host does it in
// fherma_init, solve.cu:1174-1204, 1342-1369): for each prime width, a
// Miller-Rabin search over the candidates p = 1 mod 2N below the word bound
// with each prime's psi, the prime rule applied to the results on the host,
// then the big-integer CRT products mod q of the chosen primes and the
// 30-bit set's int8 MFMA CRT table.  build_tables checks all of it.
int device_init(uint32_t L, const uint32_t* q, uint32_t log_n, nc::DeviceInit& init) {
    uint32_t q_bits = 0;
    for (int l = (int)L - 1; l >= 0 && !q_bits; --l)
        if (q[l]) q_bits = 32 * (uint32_t)l + 32 - (uint32_t)__builtin_clz(q[l]);
    const uint64_t step = 2ull << log_n;
    constexpr uint32_t kCands = 8192;
    std::vector<void*> bufs;
    auto dev_alloc = [&](size_t bytes, void** out) {
        const hipError_t e = hipMalloc(out, bytes);
        if (e == hipSuccess) bufs.push_back(*out);
        return e;
    };
    auto run = [&]() -> int {
        uint64_t* d_cand = nullptr;
        NC_HIP(dev_alloc(kCands * sizeof(uint64_t), reinterpret_cast<void**>(&d_cand)), "prime search buffer");
        for (int bits : {62, 30}) {
            uint64_t top = (1ull << bits) - 1;
            top -= (top - 1) % step;
            const uint32_t count = (uint32_t)std::min<uint64_t>(kCands, (top - 1) / step);
            NC_HIP(nc::launch_prime_search(top, step, count, d_cand, nullptr), "prime search");
            std::vector<uint64_t> psi(count);
            NC_HIP(hipMemcpy(psi.data(), d_cand, count * sizeof(uint64_t), hipMemcpyDeviceToHost), "prime search");
            std::vector<std::pair<uint64_t, uint64_t>> cands(count);
            for (uint32_t i = 0; i < count; ++i) cands[i] = {top - (uint64_t)i * step, psi[i]};
            const bool ok = nc::select_primes(cands, log_n, q_bits, bits == 62 ? (size_t)nc::kMaxPrimes
                                                                              : (size_t)nc::kMaxPrimes30,
                                              bits == 62 ? init.primes62 : init.primes30,
                                              bits == 62 ? init.psi62 : init.psi30);
            if (!ok) return fail(NC_ERR_PRIMES, "not enough NTT primes for this modulus");
        }
        uint32_t* d_q = nullptr;
        NC_HIP(dev_alloc(L * sizeof(uint32_t), reinterpret_cast<void**>(&d_q)), "modulus");
        NC_HIP(hipMemcpy(d_q, q, L * sizeof(uint32_t), hipMemcpyHostToDevice), "modulus");
        for (int bits : {62, 30}) {
            const std::vector<uint64_t>& pr = bits == 62 ? init.primes62 : init.primes30;
            std::vector<uint32_t>& prod = bits == 62 ? init.prod62 : init.prod30;
            const uint32_t k = (uint32_t)pr.size();
            uint64_t* d_p = nullptr;
            uint32_t* d_prod = nullptr;
            NC_HIP(dev_alloc(k * sizeof(uint64_t), reinterpret_cast<void**>(&d_p)), "primes");
            NC_HIP(dev_alloc((k + 1) * nc::kMaxLimbs * sizeof(uint32_t), reinterpret_cast<void**>(&d_prod)),
                   "CRT products");
            NC_HIP(hipMemcpy(d_p, pr.data(), k * sizeof(uint64_t), hipMemcpyHostToDevice), "primes");
            NC_HIP(nc::launch_crt_products(d_p, k, d_q, L, d_prod, nullptr), "CRT products");
            prod.resize((size_t)(k + 1) * nc::kMaxLimbs);
            NC_HIP(hipMemcpy(prod.data(), d_prod, prod.size() * sizeof(uint32_t), hipMemcpyDeviceToHost),
                   "CRT products");
            if (bits != 30 || q_bits + 54 < 32 * L) continue;  // no MFMA CRT for this shape
            // (-M) mod q from M mod q, then the int8 table over (M / p_j) mod q
            std::vector<uint32_t> neg_m(nc::kMaxLimbs, 0);
            bool zero = true;
            for (uint32_t l = 0; l < L; ++l) zero = zero && prod[(size_t)k * nc::kMaxLimbs + l] == 0;
            uint32_t br = 0;
            for (uint32_t l = 0; l < L && !zero; ++l) {
                const uint64_t d = (uint64_t)q[l] - prod[(size_t)k * nc::kMaxLimbs + l] - br;
                neg_m[l] = (uint32_t)d;
                br = (uint32_t)(d >> 63);
            }
            const uint32_t rt = (L + 4 + 7) / 8, kc = (k + 7) / 8;
            int8_t* d_tab = nullptr;
            int64_t* d_corr = nullptr;
            uint32_t *d_neg = nullptr, *d_negr = nullptr;
            const size_t tab_bytes = (size_t)rt * 32 * 32 * kc;
            NC_HIP(dev_alloc(tab_bytes, reinterpret_cast<void**>(&d_tab)), "CRT table");
            NC_HIP(dev_alloc(8 * rt * sizeof(int64_t), reinterpret_cast<void**>(&d_corr)), "CRT table");
            NC_HIP(dev_alloc(nc::kMaxLimbs * sizeof(uint32_t), reinterpret_cast<void**>(&d_neg)), "CRT table");
            NC_HIP(dev_alloc(nc::kMaxLimbs * sizeof(uint32_t), reinterpret_cast<void**>(&d_negr)), "CRT table");
            NC_HIP(hipMemset(d_tab, 0, tab_bytes), "CRT table");
            NC_HIP(hipMemcpy(d_neg, neg_m.data(), nc::kMaxLimbs * sizeof(uint32_t), hipMemcpyHostToDevice), "CRT table");
            NC_HIP(nc::launch_crt_mfma_tables(d_prod, d_p, k, d_q, L, rt, kc, d_neg, d_tab, d_corr, d_negr, nullptr),
                   "CRT table");
            init.crt_tab.resize(tab_bytes);
            init.crt_c.resize(8 * rt);
            init.neg_m_r.resize(nc::kMaxLimbs);
            NC_HIP(hipMemcpy(init.crt_tab.data(), d_tab, tab_bytes, hipMemcpyDeviceToHost), "CRT table");
            NC_HIP(hipMemcpy(init.crt_c.data(), d_corr, 8 * rt * sizeof(int64_t), hipMemcpyDeviceToHost), "CRT table");
            NC_HIP(hipMemcpy(init.neg_m_r.data(), d_negr, nc::kMaxLimbs * sizeof(uint32_t), hipMemcpyDeviceToHost),
                   "CRT table");
        }
        return NC_OK;
    };
    const int rc = run();
    for (void* b : bufs) (void)hipFree(b);
    return rc;
}

// Runs a host-table step, mapping its exceptions to status codes.
extern "C++" template <class F>
int guarded_tables(F&& f) {
    try {
        f();
    } catch (const std::invalid_argument& e) {
        return fail(NC_ERR_INVALID, e.what());
    } catch (const std::runtime_error& e) {
        const std::string m = e.what();
        return fail(m.rfind("self test", 0) == 0 ? NC_ERR_SELFTEST : NC_ERR_PRIMES, m);
    } catch (const std::bad_alloc&) {
        return fail(NC_ERR_NOMEM, "host tables");
    }
    return NC_OK;
}

}  // namespace

int nc_check_tables(uint32_t N, uint32_t L, const uint32_t* q) {
    if (!q) return fail(NC_ERR_INVALID, "q is null");
    return guarded_tables([&] { nc::self_test(nc::build_tables(N, L, q, /*host_twiddles=*/false)); });
}

int nc_plan_create(uint32_t N, uint32_t L, const uint32_t* q, nc_plan** out) {
    if (!out) return fail(NC_ERR_INVALID, "out is null");
    *out = nullptr;
    if (!q) return fail(NC_ERR_INVALID, "q is null");
    nc_plan* p = new (std::nothrow) nc_plan();
    if (!p) return fail(NC_ERR_NOMEM, "plan allocation");
    int rc = guarded_tables([&] {
        p->host = nc::build_tables(N, L, q, /*host_twiddles=*/false);
        nc::self_test(p->host);
    });
    if (rc) {
        delete p;
        return rc;
    }
    const hipError_t e = hipGetDevice(&p->device);
    if (e != hipSuccess) rc = hip_fail(e, "hipGetDevice");
    if (rc == NC_OK) rc = upload_tables(p->host, p->device, &p->d_tables, &p->table_bytes, p->dev);
    if (rc == NC_OK) rc = finish_plan(p);
    if (rc) {
        nc_plan_destroy(p);
        return rc;
    }
    *out = p;
    return NC_OK;
}

int nc_tableset_create(uint32_t L, const uint32_t* q, nc_tableset** out) {
    if (!out) return fail(NC_ERR_INVALID, "out is null");
    *out = nullptr;
    if (!q) return fail(NC_ERR_INVALID, "q is null");
    const auto t0 = std::chrono::steady_clock::now();
    nc_tableset* ts = new (std::nothrow) nc_tableset();
    if (!ts) return fail(NC_ERR_NOMEM, "table set allocation");
    int rc = guarded_tables([&] { nc::validate_modulus(L, q); });  // before any device work
    hipError_t e = hipSuccess;
    if (rc == NC_OK) e = hipGetDevice(&ts->device);
    if (rc == NC_OK && e != hipSuccess) rc = hip_fail(e, "hipGetDevice");
    // prime search, psi and the big-integer CRT tables on the device; the host
    // checks them, builds the word-level constants and runs the self tests
    nc::DeviceInit init;
    if (rc == NC_OK) rc = device_init(L, q, nc::kMaxLogN, init);
    if (rc == NC_OK)
        rc = guarded_tables([&] {
            ts->host = nc::build_tables(1u << nc::kMaxLogN, L, q, /*host_twiddles=*/false, &init);
            nc::self_test(ts->host);
        });
    if (rc == NC_OK) rc = upload_tables(ts->host, ts->device, &ts->d_tables, &ts->table_bytes, ts->dev);
    if (rc) {
        tableset_release(ts);
        return rc;
    }
    ts->create_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = ts;
    return NC_OK;
}

void nc_tableset_destroy(nc_tableset* ts) { tableset_release(ts); }

int nc_tableset_get_info(const nc_tableset* ts, uint32_t* k62, uint32_t* k30, uint64_t* table_bytes,
                         float* create_ms) {
    if (!ts) return fail(NC_ERR_INVALID, "null table set");
    if (k62) *k62 = ts->dev.w62.k;
    if (k30) *k30 = ts->dev.w30.k;
    if (table_bytes) *table_bytes = ts->table_bytes;
    if (create_ms) *create_ms = ts->create_ms;
    return NC_OK;
}

int nc_plan_create_shared(nc_tableset* ts, uint32_t N, nc_plan** out) {
    if (!out) return fail(NC_ERR_INVALID, "out is null");
    *out = nullptr;
    if (!ts) return fail(NC_ERR_INVALID, "null table set");
    int dev = 0;
    NC_HIP(hipGetDevice(&dev), "hipGetDevice");
    if (dev != ts->device) return fail(NC_ERR_INVALID, "the table set lives on another device");
    nc_plan* p = new (std::nothrow) nc_plan();
    if (!p) return fail(NC_ERR_NOMEM, "plan allocation");
    int rc = guarded_tables([&] {
        p->host = nc::derive_tables(ts->host, N);
        nc::self_test(p->host);
    });
    if (rc) {
        delete p;
        return rc;
    }
    ts->refs.fetch_add(1);
    p->shared = ts;
    p->device = dev;
    // the plan's own device tables: the N-dependent per-prime constants of both sets
    const nc::HostTables& t = p->host;
    const size_t b62 = align_up(t.s62.pc.size() * sizeof(nc::PrimeConst), 256);
    const size_t b30 = t.s30.pc.size() * sizeof(nc::PrimeConst32);
    p->table_bytes = b62 + b30;
    hipError_t e = hipMalloc(&p->d_tables, p->table_bytes);
    char* base = static_cast<char*>(p->d_tables);
    if (e == hipSuccess) e = upload(base, 0, t.s62.pc);
    if (e == hipSuccess) e = upload(base, b62, t.s30.pc);
    if (e != hipSuccess) rc = hip_fail(e, "plan constants");
    if (rc == NC_OK) {
        p->dev = ts->dev;  // the set's tables, rows 2^16 entries apart
        p->dev.N = t.N;
        p->dev.log_n = t.log_n;
        p->dev.w62.pc = reinterpret_cast<const nc::PrimeConst*>(base);
        p->dev.w30.pc = reinterpret_cast<const nc::PrimeConst32*>(base + b62);
        // degree N reads the prefix of every row: sampled entries against host arithmetic at N
        rc = check_device_twiddles(t, ts->dev.w30.tw_log, reinterpret_cast<const char*>(ts->dev.w62.tw_fwd),
                                   reinterpret_cast<const char*>(ts->dev.w62.tw_inv),
                                   reinterpret_cast<const char*>(ts->dev.w30.tw_fwd),
                                   reinterpret_cast<const char*>(ts->dev.w30.tw_inv));
    }
    if (rc == NC_OK) rc = finish_plan(p);
    if (rc) {
        nc_plan_destroy(p);
        return rc;
    }
    *out = p;
    return NC_OK;
}

int nc_plan_get_info(const nc_plan* p, nc_plan_info* info) {
    if (!p || !info) return fail(NC_ERR_INVALID, "null argument");
    const Choice c = choose(p, 1);
    info->N = p->host.N;
    info->L = p->host.L;
    info->q_bits = p->host.q_bits;
    info->k62 = k_of(p, 62);
    info->k30 = k_of(p, 30);
    info->width = c.width;
    info->log_blocks = c.log_blocks;
    info->mfma_paths = mfma_in_use(p, c);
    info->table_bytes = p->table_bytes;
    info->bytes_per_product = 2ull * p->res_bytes;
    return NC_OK;
}

int nc_plan_primes(const nc_plan* p, uint32_t width, uint64_t* out) {
    if (!p || !out) return fail(NC_ERR_INVALID, "null argument");
    if (int rc = check_width(width)) return rc;
    const std::vector<uint64_t>& v = width == 62 ? p->host.s62.primes : p->host.s30.primes;
    std::memcpy(out, v.data(), v.size() * sizeof(uint64_t));
    return NC_OK;
}

int nc_plan_copy_twiddles(const nc_plan* p, uint32_t width, int inverse, void* out, size_t bytes) {
    if (!p || !out) return fail(NC_ERR_INVALID, "null argument");
    if (int rc = check_width(width)) return rc;
    const size_t n = (size_t)k_of(p, width) * p->host.N;
    const void* src;
    size_t need;
    if (width == 62) {
        src = inverse ? (const void*)p->dev.w62.tw_inv : (const void*)p->dev.w62.tw_fwd;
        need = n * sizeof(nc::Tw);
    } else {
        src = inverse ? (const void*)p->dev.w30.tw_inv : (const void*)p->dev.w30.tw_fwd;
        need = n * sizeof(nc::Tw32);
    }
    if (bytes != need) return fail(NC_ERR_SIZE, "bytes must equal the twiddle table size");
    // rows of N entries, 2^tw_log entries apart (a shared set's rows are longer)
    const size_t entry = need / n, row = entry * p->host.N;
    const size_t pitch = entry << (width == 62 ? p->dev.w62.tw_log : p->dev.w30.tw_log);
    NC_HIP(hipMemcpy2D(out, row, src, pitch, row, k_of(p, width), hipMemcpyDeviceToHost), "twiddle copy");
    return NC_OK;
}

int nc_host_twiddles(uint32_t N, uint32_t L, const uint32_t* q, uint32_t width, int inverse, void* out,
                     size_t* bytes) {
    if (!q || !bytes) return fail(NC_ERR_INVALID, "null argument");
    if (int rc = check_width(width)) return rc;
    try {
        const nc::HostTables t = nc::build_tables(N, L, q, /*host_twiddles=*/true);
        const void* src;
        size_t need;
        if (width == 62) {
            const auto& v = inverse ? t.s62.tw_inv : t.s62.tw_fwd;
            src = v.data();
            need = v.size() * sizeof(nc::Tw);
        } else {
            const auto& v = inverse ? t.s30.tw_inv : t.s30.tw_fwd;
            src = v.data();
            need = v.size() * sizeof(nc::Tw32);
        }
        const size_t cap = *bytes;
        *bytes = need;
        if (!out) return NC_OK;
        if (cap < need) return fail(NC_ERR_SIZE, "buffer smaller than the twiddle table");
        std::memcpy(out, src, need);
    } catch (const std::invalid_argument& e) {
        return fail(NC_ERR_INVALID, e.what());
    } catch (const std::runtime_error& e) {
        return fail(NC_ERR_PRIMES, e.what());
    } catch (const std::bad_alloc&) {
        return fail(NC_ERR_NOMEM, "host tables");
    }
    return NC_OK;
}

int nc_plan_set_tuning(nc_plan* p, uint32_t width, uint32_t log_blocks) {
    if (!p) return fail(NC_ERR_INVALID, "null plan");
    if (width == 0) {  // back to autotuned / heuristic choices
        p->forced = false;
        return NC_OK;
    }
    if (int rc = check_width(width)) return rc;
    Choice c;
    c.width = width;
    c.mfma = mfma_default(p);
    if (p->host.log_n >= 4) {
        if (log_blocks == 0xffffffffu) {
            c.log_blocks = heuristic(p, 1, width).log_blocks;
        } else {
            if (log_blocks < nc::min_log_blocks(p->host.log_n) || log_blocks > nc::max_log_blocks(p->host.log_n))
                return fail(NC_ERR_INVALID, "log_blocks out of range for this N");
            c.log_blocks = log_blocks;
        }
    }
    p->forced_choice = c;
    p->forced = true;
    return NC_OK;
}

int nc_plan_set_chunks(nc_plan* p, uint32_t chunks) {
    if (!p) return fail(NC_ERR_INVALID, "null plan");
    if (chunks > 64) return fail(NC_ERR_INVALID, "chunks must be in [0, 64]");
    p->chunks_forced = chunks;
    return NC_OK;
}

int nc_plan_get_choice(const nc_plan* p, uint32_t batch, uint32_t* width, uint32_t* log_blocks,
                       uint32_t* mfma_paths, uint32_t* chunks) {
    if (!p) return fail(NC_ERR_INVALID, "null plan");
    if (batch == 0) return fail(NC_ERR_INVALID, "batch must be >= 1");
    const Choice c = choose(p, batch);
    if (width) *width = c.width;
    if (log_blocks) *log_blocks = c.log_blocks;
    if (mfma_paths) *mfma_paths = mfma_in_use(p, c);
    if (chunks) *chunks = std::min(c.chunks, batch);
    return NC_OK;
}

int nc_plan_set_mfma(nc_plan* p, int mode) {
    if (!p) return fail(NC_ERR_INVALID, "null plan");
    if (mode < -1 || mode > (int)kMfmaAll) return fail(NC_ERR_INVALID, "mode must be -1 or a mask of 1 and 2");
    p->mfma_mode = mode;
    return NC_OK;
}

int nc_plan_set_persist(nc_plan* p, int mode, uint32_t wait_bound_us) {
    if (!p) return fail(NC_ERR_INVALID, "null plan");
    if (mode < -1 || mode > 1) return fail(NC_ERR_INVALID, "mode must be -1, 0 or 1");
    if (mode == 1 && !nc::persist_launch_enabled())
        return fail(NC_ERR_UNSUPPORTED, "the persistent batched form is switched off in this build");
    p->persist_mode = mode;
    p->persist_tmo = wait_bound_us ? (uint64_t)wait_bound_us * 100u : 100000000ull;  // 100 MHz ticks
    return NC_OK;
}

int nc_plan_get_persist(const nc_plan* p, uint32_t batch, uint32_t* persistent) {
    if (!p) return fail(NC_ERR_INVALID, "null plan");
    if (batch == 0) return fail(NC_ERR_INVALID, "batch must be >= 1");
    const Choice c = choose(p, batch);
    if (persistent) *persistent = (c.persist && nc::persist_available(p->dev, batch)) ? 1u : 0u;
    return NC_OK;
}

int nc_plan_persist_status(nc_plan* p, void* stream, uint32_t* flags) {
    if (!p) return fail(NC_ERR_INVALID, "null plan");
    NC_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "status sync");
    NC_HIP(hipEventSynchronize(p->ev_ws), "status sync");
    uint32_t v = 0;
    NC_HIP(hipMemcpy(&v, p->d_err, sizeof v, hipMemcpyDeviceToHost), "error word");
    if (v) NC_HIP(hipMemset(p->d_err, 0, sizeof v), "error word");
    if (flags) *flags = v;
    return NC_OK;
}

int nc_persist_emulate(uint32_t N, uint32_t k, uint32_t batch, uint32_t xcds, uint32_t workgroups, uint64_t seed,
                       double* makespan, double* wait_fraction) {
    if (N < 1024 || N > 16384 || (N & (N - 1)))
        return fail(NC_ERR_INVALID, "the persistent form covers N = 2^10 .. 2^14");
    nc::PersistEmuStats st;
    const int r = nc::persist_emulate((uint32_t)__builtin_ctz(N), k, batch, xcds, workgroups, seed, &st);
    if (r) return fail(NC_ERR_STATE, "queue emulation failed check " + std::to_string(r));
    if (makespan) *makespan = st.makespan;
    if (wait_fraction) *wait_fraction = st.wait_fraction;
    return NC_OK;
}

int nc_set_tune_cache(const char* path) {
    std::lock_guard<std::mutex> g(g_cache_mu);
    g_cache_path = path ? std::string(path) : std::string();
    g_cache_set = true;
    return NC_OK;
}

int nc_plan_autotune(nc_plan* p, uint32_t batch, void* stream, uint32_t* chosen_width, uint32_t* chosen_log_blocks) {
    if (!p) return fail(NC_ERR_INVALID, "null plan");
    int rc = check_batch(p, batch);
    if (rc) return rc;
    const auto wall0 = std::chrono::steady_clock::now();
    auto wall_ms = [&] {
        return std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    };
    nc_plan::Tuned& t = p->tune[bucket_of(batch)];
    const std::string key = cache_key(p, batch);
    if (!key.empty()) {
        nc_plan::Tuned hit;
        if (cache_lookup(p, key, &hit)) {
            t = hit;
            t.tuned = true;
            t.from_cache = true;
            t.tune_ms = wall_ms();
            p->forced = false;
            if (chosen_width) *chosen_width = t.choice.width;
            if (chosen_log_blocks) *chosen_log_blocks = t.choice.log_blocks;
            return NC_OK;
        }
    }
    rc = grow_workspace(p, batch);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (capturing(s)) return fail(NC_ERR_STATE, "autotune cannot run under stream capture");
    const size_t n = (size_t)batch * p->host.N * p->host.L;
    uint32_t* buf = nullptr;
    NC_HIP(hipMalloc(&buf, 3 * n * sizeof(uint32_t)), "autotune buffers");
    if (int rc2 = ws_acquire(p, s)) {
        (void)hipFree(buf);
        return rc2;
    }
    // random operands below q: the kernels' instructions do not depend on
    // the data, but their speed does (zero operands draw less power and
    // time about 10 % faster), so candidates are timed as callers run them
    const uint32_t* ql = p->host.s30.qc.q;
    uint32_t q_hi = p->host.L - 1;
    while (q_hi > 0 && ql[q_hi] == 0) --q_hi;
    hipError_t e = nc::launch_fill_operands(buf, 2 * n, p->host.L, q_hi, ql[q_hi], 0x6e63u, s);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    const uint32_t log_n = p->host.log_n;
    const uint32_t lo = log_n >= 4 ? nc::min_log_blocks(log_n) : 0u;
    const uint32_t hi = log_n >= 4 ? std::max(lo, nc::max_log_blocks(log_n)) : 0u;
    const uint32_t avail = mfma_available(p);
    struct Cand {
        uint32_t width;
        uint32_t mfma;
    };
    std::vector<Cand> cands;
    if (p->mfma_mode < 0) {  // every distinct subset of the available paths
        for (uint32_t m = 0; m <= kMfmaAll; ++m)
            if ((m & avail) == m) cands.push_back({30u, m});
    } else {
        cands.push_back({30u, mfma_default(p)});
    }
    cands.push_back({62u, 0u});
    // one warm-up run of choice c, then `runs` timed runs appended to `out`
    auto time_runs = [&](const Choice& c, int runs, std::vector<float>& out) {
        for (int rep = 0; rep <= runs && e == hipSuccess; ++rep) {  // rep 0 warms up
            e = hipEventRecord(t0, s);
            if (e == hipSuccess && run_pipeline(p, buf, buf + n, buf + 2 * n, batch, s, c) != NC_OK)
                e = hipErrorLaunchFailure;
            if (e == hipSuccess) e = hipEventRecord(t1, s);
            if (e == hipSuccess) e = hipEventSynchronize(t1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
            if (rep > 0 && e == hipSuccess) out.push_back(ms);
        }
    };
    auto median = [](std::vector<float> v) {
        std::sort(v.begin(), v.end());
        return v.empty() ? 0.f : v[v.size() / 2];
    };
    struct Timed {
        float ms;
        Choice c;
    };
    std::vector<Timed> timed;
    std::vector<uint32_t> chunk_opts = {1u};
    if (batch >= 8) chunk_opts.push_back(2u);
    if (batch >= 32) chunk_opts.push_back(4u);
    if (p->chunks_forced) chunk_opts = {p->chunks_forced};
    for (const Cand& cd : cands) {
        for (uint32_t lb = lo; lb <= hi && e == hipSuccess; ++lb)
            for (uint32_t ch : chunk_opts) {
                if (ch > 1 && cd.width != 30 && !p->chunks_forced) continue;
                Choice c;
                c.width = cd.width;
                c.mfma = cd.mfma;
                c.log_blocks = lb;
                c.chunks = ch;
                std::vector<float> runs;
                time_runs(c, 2, runs);
                if (e == hipSuccess) timed.push_back({median(runs), c});
            }
    }
    // the persistent batched form (its own configuration: 30-bit set, both
    // matrix-core paths, split 2), kept only if its timed runs drained every
    // XCD's queue with no readiness wait reaching its bound
    if (e == hipSuccess && p->persist_mode != 0 && !p->chunks_forced && nc::persist_available(p->dev, batch) &&
        (p->mfma_mode < 0 || (uint32_t)p->mfma_mode == kMfmaAll)) {
        Choice c;
        c.width = 30;
        c.mfma = kMfmaAll;
        c.log_blocks = 2;
        c.persist = 1;
        std::vector<float> runs;
        time_runs(c, 2, runs);
        if (e == hipSuccess && persist_drained(p, batch, s)) timed.push_back({median(runs), c});
    }
    // The quick pass above times each candidate twice; runs on one device
    // vary by several percent, as much as the leaders differ.  The three
    // fastest (and the fastest unchunked one, which graphs and the server
    // use) are timed again, in turn, over three rounds of three runs, and the
    // lowest median of those nine runs decides: a choice that repeats from
    // process to process, not one run's luck.
    std::sort(timed.begin(), timed.end(), [](const Timed& x, const Timed& y) { return x.ms < y.ms; });
    std::vector<Timed> top(timed.begin(), timed.begin() + std::min<size_t>(3, timed.size()));
    for (const Timed& x : timed)
        if (x.c.chunks == 1) {
            bool in = false;
            for (const Timed& y : top) in = in || (y.c.width == x.c.width && y.c.log_blocks == x.c.log_blocks &&
                                                   y.c.mfma == x.c.mfma && y.c.chunks == 1 &&
                                                   y.c.persist == x.c.persist);
            if (!in) top.push_back(x);
            break;
        }
    std::vector<std::vector<float>> samples(top.size());
    std::vector<bool> drained(top.size(), true);
    for (int round = 0; round < 3 && e == hipSuccess; ++round)
        for (size_t i = 0; i < top.size() && e == hipSuccess; ++i) {
            time_runs(top[i].c, 3, samples[i]);
            if (e == hipSuccess && top[i].c.persist && !persist_drained(p, batch, s)) drained[i] = false;
        }
    for (size_t i = 0; i < top.size(); ++i)
        if (!drained[i]) samples[i].clear();
    int best = -1, best1 = -1;
    std::vector<float> med(top.size());
    for (size_t i = 0; i < top.size(); ++i) {
        med[i] = median(samples[i]);
        if (samples[i].empty()) continue;
        if (best < 0 || med[i] < med[best]) best = (int)i;
        if (top[i].c.chunks == 1 && (best1 < 0 || med[i] < med[best1])) best1 = (int)i;
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (e == hipSuccess) e = hipEventRecord(p->ev_ws, s);
    (void)hipStreamSynchronize(s);
    (void)hipFree(buf);
    if (e != hipSuccess) return hip_fail(e, "autotune");
    if (best < 0) return fail(NC_ERR_STATE, "autotune timed no configuration");
    t.choice = top[best].c;
    t.unchunked = best1 >= 0 ? top[best1].c : heuristic(p, batch);
    t.median_ms = med[best];
    t.tuned = true;
    t.from_cache = false;
    t.tune_ms = wall_ms();
    p->forced = false;
    if (!key.empty()) cache_store(key, t);
    if (chosen_width) *chosen_width = t.choice.width;
    if (chosen_log_blocks) *chosen_log_blocks = t.choice.log_blocks;
    return NC_OK;
}

int nc_plan_get_tune_info(const nc_plan* p, uint32_t batch, uint32_t* from_cache, float* tune_ms, float* median_ms) {
    if (!p) return fail(NC_ERR_INVALID, "null plan");
    if (batch == 0) return fail(NC_ERR_INVALID, "batch must be >= 1");
    const nc_plan::Tuned& t = p->tune[bucket_of(batch)];
    if (!t.tuned) return fail(NC_ERR_STATE, "no tuned configuration for this batch size");
    if (from_cache) *from_cache = t.from_cache ? 1u : 0u;
    if (tune_ms) *tune_ms = t.tune_ms;
    if (median_ms) *median_ms = t.median_ms;
    return NC_OK;
}

int nc_multiply(nc_plan* p, const uint32_t* a, const uint32_t* b, uint32_t* c, size_t count) {
    if (!p || !a || !b || !c) return fail(NC_ERR_INVALID, "null argument");
    const size_t n = (size_t)p->host.N * p->host.L;
    if (count != n) return fail(NC_ERR_SIZE, "a and b must contain N*L limbs");
    uint32_t* ha = p->h_pin;
    uint32_t* hb = p->h_pin + n;
    uint32_t* hc = p->h_pin + 2 * n;
    uint32_t* da = p->d_io;
    uint32_t* db = p->d_io + n;
    uint32_t* dc = p->d_io + 2 * n;
    // Operands of 1 MiB and more are staged in pieces: the transfer of one
    // piece overlaps the host copy of the next into pinned memory, on the way
    // in and on the way out.  Pieces move by the GPU's own loads and stores on
    // the mapped staging buffer, which skips a DMA set-up of ~13 us per copy;
    // a piece the transfer kernel cannot take (not 8-byte aligned) moves by
    // hipMemcpyAsync.  Measured p50 on MI355X: N = 2^10 0.083 -> 0.064 ms
    // (kernel transfers), N = 2^14 0.26 -> 0.22 ms (kernel transfers, two
    // pieces; four pieces raised p99), N = 2^16 0.93 -> 0.70 ms (four pieces,
    // kernel transfers; two pieces 0.75 ms).
    const size_t bytes = n * sizeof(uint32_t);
    const int pieces = bytes >= (2u << 20) ? nc_plan::kHostPieces : bytes >= (1u << 20) ? 2 : 1;
    const size_t pw = (n + pieces - 1) / pieces;
    auto move = [&](uint32_t* dst, const uint32_t* dst_dev, const uint32_t* src, const uint32_t* src_dev, size_t len,
                    hipMemcpyKind kind) -> hipError_t {
        if (p->h_pin_dev && len % 8 == 0 && reinterpret_cast<uintptr_t>(dst_dev) % 8 == 0 &&
            reinterpret_cast<uintptr_t>(src_dev) % 8 == 0) {
            const void* s1 = src_dev;
            void* d1 = const_cast<uint32_t*>(dst_dev);
            return nc::launch_host_xfer(&s1, &d1, 1, len, kind == hipMemcpyDeviceToHost, p->stream);
        }
        return hipMemcpyAsync(dst, src, len, kind, p->stream);
    };
    // the GPU's address of a pinned staging word
    auto dev_of = [&](const uint32_t* h) { return p->h_pin_dev ? p->h_pin_dev + (h - p->h_pin) : nullptr; };
    for (int op = 0; op < 2; ++op) {  // a and b may be the same host buffer
        const uint32_t* src = op == 0 ? a : b;
        uint32_t* h = op == 0 ? ha : hb;
        uint32_t* d = op == 0 ? da : db;
        for (int i = 0; i < pieces; ++i) {
            const size_t off = i * pw, len = std::min(pw, n - off) * sizeof(uint32_t);
            std::memcpy(h + off, src + off, len);
            NC_HIP(move(d + off, d + off, h + off, dev_of(h + off), len, hipMemcpyHostToDevice), "a/b copy");
        }
    }
    int rc = ws_acquire(p, p->stream);
    if (rc == NC_OK) rc = run_pipeline(p, da, db, dc, 1, p->stream, choose(p, 1));
    if (rc == NC_OK) rc = ws_release(p, p->stream);
    if (rc) return rc;
    for (int i = 0; i < pieces; ++i) {
        const size_t off = i * pw, len = std::min(pw, n - off) * sizeof(uint32_t);
        NC_HIP(move(hc + off, dev_of(hc + off), dc + off, dc + off, len, hipMemcpyDeviceToHost), "c copy");
        NC_HIP(hipEventRecord(p->ev_piece[i], p->stream), "c copy event");
    }
    for (int i = 0; i < pieces; ++i) {
        const size_t off = i * pw, len = std::min(pw, n - off) * sizeof(uint32_t);
        NC_HIP(hipEventSynchronize(p->ev_piece[i]), "sync");
        std::memcpy(c + off, hc + off, len);
    }
    return NC_OK;
}

int nc_multiply_device(nc_plan* p, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, uint32_t batch,
                       void* stream) {
    if (!p || !d_a || !d_b || !d_c) return fail(NC_ERR_INVALID, "null argument");
    int rc = check_batch(p, batch);
    if (rc) return rc;
    rc = grow_workspace(p, batch);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    rc = ws_acquire(p, s);
    if (rc == NC_OK) rc = run_pipeline(p, d_a, d_b, d_c, batch, s, choose(p, batch));
    if (rc == NC_OK) rc = ws_release(p, s);
    return rc;
}

int nc_multiply_device_noalloc(nc_plan* p, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, uint32_t batch,
                               void* stream) {
    if (!p || !d_a || !d_b || !d_c) return fail(NC_ERR_INVALID, "null argument");
    if (int rc = check_batch(p, batch)) return rc;
    if (!p->scratch) return fail(NC_ERR_STATE, "no scratch address for this plan's device");
    if ((uint64_t)batch * 2 * p->res_bytes > nc::kScratchBytes)
        return fail(NC_ERR_CAPACITY, "batch * bytes_per_product (" + std::to_string((uint64_t)batch * 2 * p->res_bytes) +
                                         ") exceeds the scratch array (" + std::to_string(nc::kScratchBytes) + " bytes)");
    // residue rows of a in the first half of what the batch uses, b after
    // them: nothing is allocated, and nothing needs to start at zero (every
    // residue is written before it is read; the persistent form's control
    // block, in b's trailer, is cleared by that form's reset kernel)
    char* ws = static_cast<char*>(p->scratch);
    Choice c = choose(p, batch);
    c.chunks = 1;
    return nc::run_pipeline_ws(p, ws, ws + (size_t)batch * p->res_bytes, d_a, d_b, d_c, batch,
                               static_cast<hipStream_t>(stream), c);
}

int nc_stage_residues(nc_plan* p, uint32_t width, const uint32_t* d_poly, void* d_res, uint32_t batch, void* stream) {
    if (!p || !d_poly || !d_res) return fail(NC_ERR_INVALID, "null argument");
    if (int rc = check_width(width)) return rc;
    if (int rc = check_batch(p, batch)) return rc;
    NC_HIP(nc::launch_decompose(p->dev, (int)width, d_poly, d_res, batch, nullptr, nullptr, 0,
                                static_cast<hipStream_t>(stream), (choose(p, batch).mfma & kMfmaDecompose) != 0),
           "decompose launch");
    return NC_OK;
}

int nc_stage_product_residues(nc_plan* p, uint32_t width, const uint32_t* d_a, const uint32_t* d_b, void* d_res,
                              uint32_t batch, void* stream) {
    if (!p || !d_a || !d_b || !d_res) return fail(NC_ERR_INVALID, "null argument");
    if (int rc = check_width(width)) return rc;
    int rc = check_batch(p, batch);
    if (rc) return rc;
    rc = grow_workspace(p, batch);
    if (rc) return rc;
    Choice c = choose(p, batch);
    if (c.width != width) c = heuristic(p, batch, width);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool dm = (c.mfma & kMfmaDecompose) != 0;
    const bool fuse = nc::decompose_fuses_outer(p->dev, (int)width, dm, c.log_blocks);
    if (int rc2 = ws_acquire(p, s)) return rc2;
    NC_HIP(nc::launch_decompose(p->dev, (int)width, d_a, d_res, batch, d_b, p->d_res_b, batch, s, dm,
                                fuse ? c.log_blocks : 0u),
           "decompose launch");
    NC_HIP(nc::launch_ntt_product(p->dev, (int)width, d_res, p->d_res_b, batch, c.log_blocks, s, fuse), "ntt launch");
    return ws_release(p, s);
}

int nc_stage_crt(nc_plan* p, uint32_t width, const void* d_res, uint32_t* d_c, uint32_t batch, void* stream) {
    if (!p || !d_res || !d_c) return fail(NC_ERR_INVALID, "null argument");
    if (int rc = check_width(width)) return rc;
    if (int rc = check_batch(p, batch)) return rc;
    NC_HIP(nc::launch_crt(p->dev, (int)width, d_res, d_c, batch, static_cast<hipStream_t>(stream),
                          (choose(p, batch).mfma & kMfmaCrt) != 0),
           "crt launch");
    return NC_OK;
}

int nc_profile_stages(nc_plan* p, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, uint32_t batch,
                      uint32_t reps, void* stream, float* ms) {
    if (!p || !d_a || !d_b || !d_c || !ms || reps == 0) return fail(NC_ERR_INVALID, "bad argument");
    if (reps > 4096) return fail(NC_ERR_INVALID, "reps above 4096");
    int rc = check_batch(p, batch);
    if (rc) return rc;
    rc = grow_workspace(p, batch);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const Choice c = choose(p, batch);
    const int w = (int)c.width;
    const bool dm = (c.mfma & kMfmaDecompose) != 0;
    const bool fuse = nc::decompose_fuses_outer(p->dev, w, dm, c.log_blocks);
    // reps run back to back, as a batched step's launches do; events are read
    // after the last one
    std::vector<hipEvent_t> ev(6 * (size_t)reps, nullptr);
    int status = NC_OK;
    for (size_t i = 0; i < ev.size() && status == NC_OK; ++i) {
        const hipError_t e = hipEventCreate(&ev[i]);
        if (e != hipSuccess) status = hip_fail(e, "event");
    }
    double acc[5] = {0, 0, 0, 0, 0};
    const bool inv = inv_in_crt(p, c);  // then the CRT time includes the last inverse stages
    if (status == NC_OK) status = ws_acquire(p, s);
    hipError_t e = hipSuccess;
    for (uint32_t r = 0; r < reps && status == NC_OK && e == hipSuccess; ++r) {
        hipEvent_t* er = &ev[6 * (size_t)r];
        e = hipEventRecord(er[0], s);
        if (e == hipSuccess)
            e = nc::launch_decompose(p->dev, w, d_a, p->d_res_a, batch, d_b, p->d_res_b, batch, s, dm,
                                     fuse ? c.log_blocks : 0u);
        if (e == hipSuccess) e = hipEventRecord(er[1], s);
        if (e == hipSuccess && !fuse)
            e = nc::launch_ntt_outer_fwd(p->dev, w, p->d_res_a, p->d_res_b, batch, c.log_blocks, s);
        if (e == hipSuccess) e = hipEventRecord(er[2], s);
        if (e == hipSuccess) e = nc::launch_ntt_inner(p->dev, w, p->d_res_a, p->d_res_b, batch, c.log_blocks, s);
        if (e == hipSuccess) e = hipEventRecord(er[3], s);
        if (e == hipSuccess && !inv) e = nc::launch_ntt_outer_inv(p->dev, w, p->d_res_a, batch, c.log_blocks, s);
        if (e == hipSuccess) e = hipEventRecord(er[4], s);
        if (e == hipSuccess)
            e = inv ? nc::launch_crt_inv2(p->dev, static_cast<const uint32_t*>(p->d_res_a), d_c, batch, s)
                    : nc::launch_crt(p->dev, w, p->d_res_a, d_c, batch, s, (c.mfma & kMfmaCrt) != 0);
        if (e == hipSuccess) e = hipEventRecord(er[5], s);
    }
    if (status == NC_OK && e == hipSuccess) status = ws_release(p, s);
    if (status == NC_OK && e == hipSuccess) e = hipStreamSynchronize(s);
    for (uint32_t r = 0; r < reps && status == NC_OK && e == hipSuccess; ++r)
        for (int i = 0; i < 5 && e == hipSuccess; ++i) {
            float t = 0;
            e = hipEventElapsedTime(&t, ev[6 * (size_t)r + i], ev[6 * (size_t)r + i + 1]);
            acc[i] += t;
        }
    if (status == NC_OK && e != hipSuccess) status = hip_fail(e, "profile");
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (status) return status;
    const bool outer = p->host.log_n >= 4 && c.log_blocks > 0;
    for (int i = 0; i < 5; ++i) ms[i] = (float)(acc[i] / reps);
    if (!outer) ms[1] = ms[3] = 0.f;
    if (fuse) ms[1] = 0.f;  // the outer forward stages ran inside the decomposition
    if (inv_in_crt(p, c)) ms[3] = 0.f;  // the last inverse stages ran inside the CRT
    return NC_OK;
}

int nc_measure_mul32_peak(double* out) {
    if (!out) return fail(NC_ERR_INVALID, "null argument");
    hipDeviceProp_t prop;
    int dev = 0;
    NC_HIP(hipGetDevice(&dev), "device");
    NC_HIP(hipGetDeviceProperties(&prop, dev), "device properties");
    const uint32_t blocks = (uint32_t)prop.multiProcessorCount * 16;  // 16 waves of 64 lanes per CU
    const uint32_t iters = 4096;
    uint32_t* d_in = nullptr;
    uint64_t* d_out = nullptr;
    NC_HIP(hipMalloc(&d_in, 64 * sizeof(uint32_t)), "probe input");
    hipError_t e = hipMalloc(&d_out, (size_t)blocks * 256 * sizeof(uint64_t));
    uint32_t h_in[64];
    for (int i = 0; i < 64; ++i) h_in[i] = 0x9E3779B9u * (uint32_t)(i + 1);
    if (e == hipSuccess) e = hipMemcpy(d_in, h_in, sizeof h_in, hipMemcpyHostToDevice);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    float best = 0;
    for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {  // first launch warms up
        e = hipEventRecord(t0, nullptr);
        if (e == hipSuccess) e = nc::launch_mul32_peak(d_in, d_out, blocks, iters, nullptr);
        if (e == hipSuccess) e = hipEventRecord(t1, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
        if (rep > 0 && (best == 0 || ms < best)) best = ms;
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    (void)hipFree(d_in);
    if (d_out) (void)hipFree(d_out);
    if (e != hipSuccess) return hip_fail(e, "mul32 probe");
    const double muls = (double)blocks * 256.0 * iters * 8.0;
    *out = muls / (best * 1e-3);
    return NC_OK;
}

int nc_gen_uniform(uint64_t seed, uint32_t N, uint32_t L, const uint32_t* q, uint32_t* out) {
    if (!q || !out || L == 0 || L > (uint32_t)nc::kMaxLimbs || N == 0) return fail(NC_ERR_INVALID, "bad argument");
    int qbits = 0;
    for (int l = (int)L - 1; l >= 0; --l)
        if (q[l]) {
            qbits = l * 32 + 32 - __builtin_clz(q[l]);
            break;
        }
    if (qbits == 0) return fail(NC_ERR_INVALID, "q is zero");
    const int top = qbits - 32 * ((int)L - 1);
    const uint32_t mask = top >= 32 ? 0xffffffffu : (top <= 0 ? 0u : ((1u << top) - 1u));
    const uint32_t words = (L + 1) / 2;
    uint64_t s = seed;
    uint32_t cand[nc::kMaxLimbs + 1];
    for (uint32_t i = 0; i < N; ++i) {
        for (;;) {
            for (uint32_t w = 0; w < words; ++w) {
                uint64_t z = (s += 0x9E3779B97F4A7C15ull);
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                z ^= z >> 31;
                cand[2 * w] = (uint32_t)z;
                cand[2 * w + 1] = (uint32_t)(z >> 32);
            }
            cand[L - 1] &= mask;
            int cmp = 0;
            for (int l = (int)L - 1; l >= 0 && !cmp; --l)
                if (cand[l] != q[l]) cmp = cand[l] > q[l] ? 1 : -1;
            if (cmp < 0) break;
        }
        std::memcpy(out + (size_t)i * L, cand, L * sizeof(uint32_t));
    }
    return NC_OK;
}

}  // extern "C"
