/*
 * negacyclic.h -- C ABI of the MI355X (gfx950) negacyclic product library.
 *
 *   c = a * b  mod (X^N + 1, q)
 *
 * a, b, c are N x L arrays of little-endian u32 limbs, coefficient major:
 * limb l of coefficient i sits at [i*L + l] (the reference layout,
 * negacyclic.fhk:3-10, main.cpp:25-37, solve.cu:520/835).  Inputs must be
 * canonical in [0, q); the output is canonical in [0, q).  q is any odd
 * modulus of L <= 32 limbs with bits(q) + 70 <= 1088 (the reference's caps,
 * solve.cu:1170 and 1339-1341); N is a power of two in [2, 65536] (the
 * reference stops at 32768, solve.cu:23).
 *
 * The arithmetic is exact: residue-number-system products over k NTT-friendly
 * primes, negacyclic NTTs, a pointwise product and a centred CRT back to L
 * limbs mod q.  Two prime sets are built per plan, both by the reference's
 * prime rule (solve.cu:1179-1204: p = 1 mod 2N, descending, until
 * bits(prod p) >= log2 N + 2 bits(q) + 4): 62-bit primes below 2^62 (the
 * reference's own set) and 30-bit primes below 2^30 (about twice as many,
 * cheaper 32-bit arithmetic).  Which one a call uses is a tuning choice
 * (nc_plan_autotune / nc_plan_set_tuning); the output never depends on it.
 * Every entry point returns NC_OK (0) or a negative status; nc_last_error()
 * gives the detail of the last failure on the calling thread.
 *
 * Threading: a plan's host state is not re-entrant (like the reference's
 * State, solve.cu:28-65): calls on one plan from several host threads must
 * be serialised.  Its device work may go to any streams: every call that
 * uses the plan's residue workspace (nc_multiply, nc_multiply_device,
 * nc_stage_product_residues, nc_profile_stages, nc_plan_autotune) first
 * makes its stream wait for the previous such call's work, so calls on
 * different streams reach the workspace one after another, as the
 * reference's calls do on its one stream (solve.cu:1489-1521).  Under
 * stream capture the captured graph carries no such wait.
 */
#ifndef NEGACYCLIC_H
#define NEGACYCLIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
#define NC_OK 0
#define NC_ERR_INVALID -1    /* bad N, L, q, batch, or a null pointer        */
#define NC_ERR_SIZE -2       /* a/b/c element count != N*L                   */
#define NC_ERR_DEVICE -3     /* a HIP runtime error                          */
#define NC_ERR_NOMEM -4      /* device or host allocation failed             */
#define NC_ERR_PRIMES -5     /* no prime set covers the coefficient bound    */
#define NC_ERR_SELFTEST -6   /* an init-time known-answer test failed        */
#define NC_ERR_CAPACITY -7   /* request exceeds a configured capacity        */
#define NC_ERR_STATE -8      /* call not valid in the object's current state */
#define NC_ERR_UNSUPPORTED -9 /* the requested form is switched off in this build */

typedef struct nc_plan nc_plan;

/* Plan information (nc_plan_get_info). */
typedef struct nc_plan_info {
    uint32_t N;            /* ring degree                                   */
    uint32_t L;            /* limbs per coefficient                         */
    uint32_t q_bits;       /* bits(q)                                       */
    uint32_t k62;          /* primes in the 62-bit set                      */
    uint32_t k30;          /* primes in the 30-bit set                      */
    uint32_t width;        /* prime width used for one product per call     */
    uint32_t log_blocks;   /* its NTT outer split (log2 blocks per row)     */
    uint32_t mfma_paths;   /* matrix-core paths it runs (NC_MFMA_* mask)    */
    uint64_t table_bytes;  /* resident device bytes of the plan's tables    */
    uint64_t bytes_per_product; /* device workspace bytes per product       */
} nc_plan_info;

/* Text for a status code. */
const char* nc_status_string(int status);
/* Detail of the last failure on this host thread ("" if none). */
const char* nc_last_error(void);

/* ---- plan: replaces fherma_init / fherma_free ------------------------
 * nc_plan_create replaces the set-up half of fherma_init
 * (solve.cu:1150-1487): validation (1167-1172), prime rule (1174-1204),
 * twiddle / Montgomery / CRT tables (1224-1369, 1396-1398), the
 * known-answer self tests (1151-1166, 1285-1314, 1370-1415) and the
 * upload.  The [k][N] twiddle tables (the bulk of the tables, 1224-1257)
 * are built on the device from log2(N) powers per prime, and sampled
 * entries are checked against host arithmetic; the rest is built on the
 * host.  nc_plan_destroy replaces fherma_free (solve.cu:1530-1567) and is
 * null safe. */
int nc_plan_create(uint32_t N, uint32_t L, const uint32_t* q, nc_plan** out);
/* Builds the host tables of (N, L, q) and runs the known-answer self tests
 * without touching a device (the host half of nc_plan_create). */
int nc_check_tables(uint32_t N, uint32_t L, const uint32_t* q);
void nc_plan_destroy(nc_plan* plan);
int nc_plan_get_info(const nc_plan* plan, nc_plan_info* info);
/* The primes of one set (width 62 or 30), host copy; out holds k values. */
int nc_plan_primes(const nc_plan* plan, uint32_t width, uint64_t* out);
/* ---- shared table sets (no counterpart in the reference) ---------------
 * One set of resident tables for every degree N <= 65536 of one modulus q
 * (L limbs): both prime sets chosen once by the prime rule at N = 65536
 * (p = 1 mod 2^17, so every smaller degree's NTTs exist in them), the
 * 65536-entry twiddle rows of every prime built on the device, the CRT and
 * matrix-core tables built once.  A plan of degree N on the set
 * (nc_plan_create_shared) uses psi_N = psi_65536^(65536 / N), whose twiddle
 * table is the first N entries of every row, and owns only its N-dependent
 * per-prime constants and its staging: the table work the reference does in
 * fherma_init (solve.cu:1150-1487) is paid once per modulus instead of once
 * per degree, and one set serves every degree.  Products are bit-identical
 * to those of nc_plan_create's plans; the primes differ from the per-N rule
 * (solve.cu:1179-1204) below N = 65536, so stage-level residues of a shared
 * plan are over the set's primes (nc_plan_primes). */
typedef struct nc_tableset nc_tableset;
int nc_tableset_create(uint32_t L, const uint32_t* q, nc_tableset** out);
/* Drops the creator's reference (null safe); the set is freed once every
 * plan created from it has been destroyed as well. */
void nc_tableset_destroy(nc_tableset* set);
/* Primes per set, resident device bytes, and the wall time nc_tableset_create
 * took (any out pointer may be NULL). */
int nc_tableset_get_info(const nc_tableset* set, uint32_t* k62, uint32_t* k30, uint64_t* table_bytes,
                         float* create_ms);
/* A plan of degree N (power of two in [2, 65536]) on the set's tables, on
 * the set's device (the current device must be that one).  Its
 * nc_plan_info.table_bytes counts only its own per-prime constants. */
int nc_plan_create_shared(nc_tableset* set, uint32_t N, nc_plan** out);

/* Copies one device twiddle table of a plan to host memory: width 62 gives
 * k x N {w, ws} pairs of uint64 (16 bytes each), width 30 k x N pairs of
 * uint32 (8 bytes, forward w stored as 2^32 - w); inverse selects psi^-1.
 * bytes must equal the table size.  Inspection for tests and tools; the
 * reference keeps the same tables (tw_fwd / tw_fwd_shoup, solve.cu:1224-1257)
 * on its device with no such accessor. */
int nc_plan_copy_twiddles(const nc_plan* plan, uint32_t width, int inverse, void* out, size_t bytes);
/* The same table built entirely on the host (no device), in the same
 * layout: the reference's host construction (solve.cu:1224-1257) in this
 * library's merged-psi form.  *bytes: capacity of out on entry, the table
 * size on return; out == NULL only reports the size. */
int nc_host_twiddles(uint32_t N, uint32_t L, const uint32_t* q, uint32_t width, int inverse, void* out,
                     size_t* bytes);
/* Pin every call to prime width `width` (62 or 30) and NTT outer split
 * `log_blocks` (0xffffffff: a default for that width).  width 0 returns to
 * the autotuned / default choices.  Any accepted setting yields bit-identical
 * output. */
int nc_plan_set_tuning(nc_plan* plan, uint32_t width, uint32_t log_blocks);

/* The configuration calls with `batch` products use now (tuned, forced or
 * default): prime width, NTT outer split, matrix-core paths (NC_MFMA_*),
 * chunks (any out pointer may be NULL). */
int nc_plan_get_choice(const nc_plan* plan, uint32_t batch, uint32_t* width, uint32_t* log_blocks,
                       uint32_t* mfma_paths, uint32_t* chunks);

/* Batches of nc_multiply_device run as `chunks` slices on two internal
 * streams, the second started one decomposition later, so that one slice's
 * NTT (VALU-bound) overlaps another's decomposition and CRT (memory-bound).
 * chunks 0 lets nc_plan_autotune decide (default 1: no split); 1..64 pins it.
 * Output bits never depend on it. */
int nc_plan_set_chunks(nc_plan* plan, uint32_t chunks);

/* Matrix-core (int8 MFMA) paths of the 30-bit prime set:
 *   NC_MFMA_DECOMPOSE  limbs -> residues (L % 4 == 0, N >= 32)
 *   NC_MFMA_CRT        weighted residues -> limbs (N >= 32,
 *                      bits(q) + 54 >= 32 L)
 * mode: a mask of these (0 = VALU kernels only), or -1 to let
 * nc_plan_autotune decide (default: all on until tuned).  A path that is not
 * available for the plan's shape runs on the VALU.  Output bits never
 * depend on it. */
#define NC_MFMA_DECOMPOSE 1
#define NC_MFMA_CRT 2
int nc_plan_set_mfma(nc_plan* plan, int mode);

/* Persistent batched form: the whole batch in one launch, replacing the
 * reference's fixed launch sequence per product (solve.cu:1489-1528).  Every
 * XCD drains its own queue of work items -- decompositions (with the first
 * two forward NTT stages), inner NTT blocks with the pointwise product, CRT
 * groups (with the last two inverse stages) -- of the products assigned to
 * it, so the stages of different products share every compute unit; an NTT
 * item waits for its product's decompositions, a CRT item for its product's
 * NTT items, through counters polled inside the launch.  Available for the
 * 30-bit set with both matrix-core paths (the fused shape: L = 28, 57..64
 * primes), N = 2^10 .. 2^14, when the device reports its XCD count and the
 * batch has at least one product per XCD; elsewhere calls run the launch
 * sequence.  mode: -1 lets nc_plan_autotune decide (default off until
 * tuned), 0 never, 1 wherever available.  wait_bound_us bounds one readiness
 * wait inside the launch (0: 1 s); a wait that reaches it stops its XCD's
 * queue, so the launch still ends, and sets the plan's error word
 * (nc_plan_persist_status).  Output bits never depend on the mode.  A
 * library built with the form switched off returns NC_ERR_UNSUPPORTED for
 * mode 1 (-1 and 0 then run the launch sequence).
 * (No counterpart in the reference.) */
int nc_plan_set_persist(nc_plan* plan, int mode, uint32_t wait_bound_us);
/* *persistent = 1 if calls with `batch` products run the persistent form. */
int nc_plan_get_persist(const nc_plan* plan, uint32_t batch, uint32_t* persistent);
/* Synchronises `stream` and the plan's last workspace user, then returns in
 * *flags and clears the error bits the plan's persistent launches set since
 * the last call (1: a readiness wait reached its bound; the products of that
 * launch are not valid).  0 when every launch drained its queues. */
int nc_plan_persist_status(nc_plan* plan, void* stream, uint32_t* flags);
/* CPU emulation of the persistent form's queues (no device): `batch`
 * products of degree N with k primes over `xcds` XCDs, `workgroups`
 * workgroups per XCD taking queue positions in order (each fetching its next
 * position before running the current one) with random item durations from
 * `seed`.  NC_OK when every item ran exactly once, only after its inputs,
 * and every queue drained (no deadlock); NC_ERR_STATE naming the failed
 * check otherwise.  makespan (emulated time units) and wait_fraction (share
 * of workgroup time spent waiting for inputs) may be NULL. */
int nc_persist_emulate(uint32_t N, uint32_t k, uint32_t batch, uint32_t xcds, uint32_t workgroups, uint64_t seed,
                       double* makespan, double* wait_fraction);

/* Autotuner: times every (prime width, NTT split, matrix-core paths, chunks)
 * for `batch` products on `stream` (zero inputs; the kernels' work does not
 * depend on the data), twice each, then the three fastest and the fastest
 * unchunked one again in turn over three rounds of three runs, and keeps the
 * lowest median of those nine runs for batches of that size bucket
 * (ceil(log2 batch)); the best unchunked configuration is kept as well, for
 * graphs and the server, which run one pipeline per batch.  With a tuning
 * cache (nc_set_tune_cache) a choice found there for this device, shape,
 * bucket and pinned settings is taken without any device work, and a new
 * choice is appended to it.  Clears a setting forced by nc_plan_set_tuning.
 * Output bits never depend on it.  (The reference has a single fixed
 * schedule, solve.cu:1006-1038.) */
int nc_plan_autotune(nc_plan* plan, uint32_t batch, void* stream, uint32_t* chosen_width,
                     uint32_t* chosen_log_blocks);
/* Process-wide tuning cache: a text file of one line per tuned (device
 * architecture and compute-unit count, N, L, bits(q), primes per set, batch
 * bucket, pinned matrix-core mode and chunks), appended to by every
 * nc_plan_autotune that times, read by every one that can skip timing.  path
 * NULL (or "") disables it; without a call the environment variable
 * NC_TUNE_CACHE names it, if set.  (No counterpart in the reference.) */
int nc_set_tune_cache(const char* path);
/* How the tuned configuration of batch-size bucket `batch` came about:
 * from_cache 1 if nc_plan_autotune took it from the cache, tune_ms the wall
 * time of that call, median_ms the chosen configuration's median run (any
 * out pointer may be NULL; NC_ERR_STATE if that bucket is not tuned). */
int nc_plan_get_tune_info(const nc_plan* plan, uint32_t batch, uint32_t* from_cache, float* tune_ms,
                          float* median_ms);

/* ---- plain product, host buffers: replaces fherma_run ----------------
 * solve.cu:1489-1528: a, b in host memory (N*L u32 each), c written to host
 * memory; count must equal N*L (NC_ERR_SIZE otherwise, like the
 * reference's runtime_error at 1492-1494).  Returns after c is complete. */
int nc_multiply(nc_plan* plan, const uint32_t* a, const uint32_t* b, uint32_t* c, size_t count);

/* ---- device-resident batch --------------------------------------------
 * `batch` independent products; d_a, d_b, d_c are device pointers to
 * batch x N x L u32 (product-major).  Work is enqueued on `stream` (a
 * hipStream_t, or NULL for the null stream) and the call returns without
 * waiting.  Output is bit-identical to nc_multiply on each product. */
int nc_multiply_device(nc_plan* plan, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, uint32_t batch,
                       void* stream);

/* ---- allocation-free device batch ---------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, allocation-free serving loops,
 * harnesses that forbid hipMalloc.  The same products as nc_multiply_device
 * (bit-identical, same arguments, one chunk), with the residue workspace in
 * a fixed device array of NC_NOALLOC_SCRATCH_BYTES per device that the
 * library holds for the life of the process, instead of the plan's
 * allocated workspace.  Calls no hipMalloc / hipFree and no allocating
 * library routine.  batch * nc_plan_info.bytes_per_product must not exceed
 * NC_NOALLOC_SCRATCH_BYTES (NC_ERR_CAPACITY otherwise): one product at
 * N = 65536, four at N = 16384.  Every call on one device shares that
 * device's array and nothing in it needs to start at zero (the persistent
 * form's control block, in the array's per-product trailer, is cleared by
 * that form's own first kernel, not by a memset); calls that
 * overlap on one device (other streams, plans or threads) race on it, so
 * the caller orders them. */
#define NC_NOALLOC_SCRATCH_BYTES (32ull << 20)
int nc_multiply_device_noalloc(nc_plan* plan, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c,
                               uint32_t batch, void* stream);

/* ---- stage entry points (device pointers, enqueue on stream) ---------
 * For stage-level parity checks.  `width` picks the prime set; residues are
 * u64 for width 62 and u32 for width 30, layout [batch][k][N].
 *  nc_stage_residues:          res[p][j][i] = poly_p[i] mod p_j (canonical)
 *                              (the value solve.cu:533-570 computes before
 *                              its Montgomery lift and twist)
 *  nc_stage_product_residues:  res[p][j][i] = (c_int[i] mod p_j) y_j mod p_j,
 *                              c_int the exact integer negacyclic product and
 *                              y_j = (M / p_j)^-1 mod p_j, M the product of
 *                              the set's primes (canonical; the inverse
 *                              transform carries the CRT weight y_j that
 *                              solve.cu:768-776 applies in the CRT)
 *                              (solve.cu:1500-1514 up to untwist)
 *  nc_stage_crt:               d_c from such weighted residues
 *                              (solve.cu:752-836) */
int nc_stage_residues(nc_plan* plan, uint32_t width, const uint32_t* d_poly, void* d_res, uint32_t batch,
                      void* stream);
int nc_stage_product_residues(nc_plan* plan, uint32_t width, const uint32_t* d_a, const uint32_t* d_b, void* d_res,
                              uint32_t batch, void* stream);
int nc_stage_crt(nc_plan* plan, uint32_t width, const void* d_res, uint32_t* d_c, uint32_t batch, void* stream);

/* ---- diagnostics -------------------------------------------------------
 * nc_profile_stages runs the batched pipeline `reps` times (at most 4096)
 * back to back on `stream` and returns the mean duration in milliseconds of
 * each kernel launch, measured with HIP events recorded on that stream
 * between the launches:
 *   ms[0] decompose, ms[1] NTT outer forward, ms[2] NTT inner (forward,
 *   pointwise, inverse), ms[3] NTT outer inverse, ms[4] CRT.
 * A stage the plan's split does not launch reports 0, and so does a stage
 * that runs inside a neighbour (the outer forward stages inside the
 * decomposition, the outer inverse stages inside the CRT): its time is in
 * that neighbour's figure.  Blocks until done.
 * (No counterpart in the reference, which has no timing or tracing.) */
int nc_profile_stages(nc_plan* plan, const uint32_t* d_a, const uint32_t* d_b, uint32_t* d_c, uint32_t batch,
                      uint32_t reps, void* stream, float* ms);
/* Measured 32x32->64-bit integer multiply throughput of the current device
 * (v_mad_u64_u32 chains, all compute units busy), in multiplies per second.
 * The integer multiplier is what bounds this library's kernels; this is the
 * peak their roofline fraction is quoted against.  Blocks until done. */
int nc_measure_mul32_peak(double* mul32_per_second);

/* ---- synthetic inputs ---------------------------------------------------
 * Fills out (N x L) with values uniform in [0, q): SplitMix64(seed) stream,
 * ceil(L/2) u64 words per candidate (word w -> limbs 2w, 2w+1), top limb
 * masked to bits(q) - 32(L-1) bits, candidates >= q redrawn.  Host only. */
int nc_gen_uniform(uint64_t seed, uint32_t N, uint32_t L, const uint32_t* q, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif /* NEGACYCLIC_H */
