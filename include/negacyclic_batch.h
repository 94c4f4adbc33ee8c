/* Multi-degree batches and serving on top of the negacyclic product library
 * (negacyclic.h).  Neither has a counterpart in the reference, whose only
 * entry points run one product per call (fherma.h:45-47, solve.cu:1489-1528);
 * they are the batched and served forms of that same product, bit-identical
 * to it per product.
 *
 * Thread safety: an nc_batch is used from one host thread at a time; an
 * nc_server may be called from any number of threads. */
#ifndef NEGACYCLIC_BATCH_H
#define NEGACYCLIC_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "negacyclic.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- batches of mixed ring degree --------------------------------------
 * A group is `count` products of one degree N whose operands lie contiguous
 * in device memory (count x N x L u32 each, product-major). */
typedef struct nc_group {
    uint32_t N;
    uint32_t count;
    const uint32_t* d_a;
    const uint32_t* d_b;
    uint32_t* d_c;
} nc_group;

typedef struct nc_batch nc_batch;
typedef struct nc_graph nc_graph;

/* Per-degree plans for one modulus q (L limbs), created on first use, all on
 * one table set (nc_tableset_create) made with the first plan. */
int nc_batch_create(uint32_t L, const uint32_t* q, nc_batch** out);
/* Destroys the plans; every graph captured from this batch must be gone. */
void nc_batch_destroy(nc_batch* batch);
/* The plan of degree N (created if needed); owned by the batch. */
int nc_batch_plan(nc_batch* batch, uint32_t N, nc_plan** out);
/* Autotunes each group's (N, count) on `stream` (nc_plan_autotune). */
int nc_batch_autotune(nc_batch* batch, const nc_group* groups, uint32_t n, void* stream);
/* Enqueues every group on `stream`, one after another, without a graph. */
int nc_batch_run(nc_batch* batch, const nc_group* groups, uint32_t n, void* stream);
/* Captures every group's pipeline into one executable HIP graph: one branch
 * per distinct degree (branches run concurrently; groups of one degree run in
 * order within its branch), each with the configuration its plan has for that
 * count, and workspace owned by the graph.  The graph is bound to the
 * groups' buffers; replays recompute from whatever they then hold. */
int nc_batch_capture(nc_batch* batch, const nc_group* groups, uint32_t n, nc_graph** out);
/* One launch of the whole graph on `stream` (returns without waiting). */
int nc_graph_launch(nc_graph* graph, void* stream);
void nc_graph_destroy(nc_graph* graph);

/* ---- serving -------------------------------------------------------------
 * A request is one product of any served degree, with operands and result in
 * host memory.  A scheduler thread batches waiting requests per degree and
 * keeps up to `lanes_per_degree` batches of each degree in flight on their
 * own streams (continuous batching: requests that arrive while others run
 * join the next batch as lanes free up); each batch's copies in and out
 * overlap other batches' kernels.  A request completes as soon as its own
 * result copy has landed, not when its batch's last one has.  The device memory of requests is one
 * arena the server allocates at creation: hbm_budget bytes, or less when
 * every lane of every degree holding max_batch requests at once needs less
 * (nc_server_stats.device_bytes reports it).  Each batch's operands, results
 * and residue workspace are one block of that arena (at most
 * count x nc_server_footprint bytes), and a waiting request is admitted only
 * when its batch's block fits, so the device bytes of admitted, unfinished
 * requests never exceed hbm_budget; plan tables (nc_plan_info.table_bytes
 * per degree) are outside it.  When a request's a, b and c are each below
 * 1 MiB (N L 4 bytes; N <= 8192 at L = 28) and all lie in mapped pinned
 * memory (nc_host_alloc, or any hipHostMalloc allocation), the GPU moves
 * them with its own loads and stores over PCIe: one launch for all operands
 * of a batch, one per result, with no per-copy DMA set-up.  Larger requests
 * in pinned memory move by hipMemcpyAsync on the DMA engines, which overlap
 * kernels; pageable memory works too but copies synchronously. */
typedef struct nc_server_config {
    uint32_t L;
    const uint32_t* q;
    const uint32_t* degrees;   /* ring degrees served (one resident table
                                  set for all of them, nc_tableset_create) */
    uint32_t n_degrees;
    uint32_t max_batch;        /* products per launch (per lane)            */
    uint32_t lanes_per_degree; /* batches of one degree in flight at once   */
    uint64_t hbm_budget;       /* bound on the request arena (device bytes)  */
    uint32_t autotune;         /* 1: autotune each degree at creation         */
} nc_server_config;

typedef struct nc_server_stats {
    uint64_t submitted;
    uint64_t completed;
    uint64_t launches;           /* batches run                              */
    uint64_t max_batch_seen;     /* largest batch launched                   */
    uint64_t peak_bytes;         /* most arena bytes in use at once          */
    uint64_t admission_waits;    /* times a waiting request did not fit      */
    uint64_t device_bytes;       /* the request arena's size                 */
    /* where latency goes (microseconds, maxima since creation or the last
     * nc_server_reset_stats): the longest gap between two scheduler turns,
     * the longest host time inside one batch's enqueue (copies, kernels,
     * events), and the longest time from a batch's enqueue to the turn that
     * saw its last request complete */
    uint64_t max_turn_gap_us;
    uint64_t max_enqueue_us;
    uint64_t max_batch_us;
    uint64_t mapped_requests;    /* requests moved by the transfer kernels     */
} nc_server_stats;

typedef struct nc_server nc_server;

/* Creates the per-degree plans (autotuned at batch 1 and max_batch when
 * config->autotune), lanes and scheduler thread, and returns once that
 * thread has run every lane at every batch-size bucket (1, 2, 4, ..,
 * max_batch) on zeros, so that no request pays for a kernel's or a
 * stream's first use. */
int nc_server_create(const nc_server_config* config, nc_server** out);
/* Waits for every submitted request, then frees everything (null safe). */
void nc_server_destroy(nc_server* server);
/* Device bytes one request of degree N counts against hbm_budget. */
int nc_server_footprint(const nc_server* server, uint32_t N, uint64_t* bytes);
/* Queues one product (returns at once).  a, b (N x L u32) are read and c is
 * written by the device at some later time: keep them alive and untouched
 * until nc_server_wait reports completion. */
int nc_server_submit(nc_server* server, uint32_t N, const uint32_t* a, const uint32_t* b, uint32_t* c,
                     uint64_t* ticket);
/* Waits for one ticket (timeout_ms < 0: no limit; NC_ERR_STATE on timeout or
 * an unknown ticket); latency_s (nullable) receives completion - submission.
 * A ticket can be waited for once. */
int nc_server_wait(nc_server* server, uint64_t ticket, int timeout_ms, double* latency_s);
int nc_server_get_stats(const nc_server* server, nc_server_stats* out);
/* Zeroes every counter and maximum of nc_server_stats except device_bytes
 * (to measure one stretch of traffic, e.g. after a warm-up). */
int nc_server_reset_stats(nc_server* server);
/* Open-loop replay: request i is submitted at arrival_s[i] seconds after the
 * call starts (arrival_s ascending); latency_s[i] receives its completion
 * time minus its scheduled arrival, makespan_s (nullable) the time from start
 * to the last completion. */
int nc_server_replay(nc_server* server, uint32_t n, const uint32_t* degrees, const double* arrival_s,
                     const uint32_t* const* a, const uint32_t* const* b, uint32_t* const* c, double* latency_s,
                     double* makespan_s);

/* Pinned host memory (hipHostMalloc) for request buffers; NULL on failure. */
void* nc_host_alloc(size_t bytes);
void nc_host_free(void* p);

#ifdef __cplusplus
}
#endif

#endif /* NEGACYCLIC_BATCH_H */
