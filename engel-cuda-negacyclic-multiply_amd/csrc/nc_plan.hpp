// Host-side plan: the resident per-(N, q) state behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "nc_common.hpp"
#include "nc_kernels.hpp"

namespace nc {

// Host copy of one prime set's tables.
struct HostSet62 {
    std::vector<uint64_t> primes;
    std::vector<Tw> tw_fwd, tw_inv;  // [k][N] (host-built tables only)
    std::vector<Tw> tw_pow;          // [k][2][log_n]: psi^(2^i), psi^-(2^i) with Shoup quotients
    std::vector<PrimeConst> pc;      // [k]
    std::vector<uint64_t> place;     // [k][kMaxLimbs]
    std::vector<CrtConst> cc;        // [k]
    std::vector<uint32_t> miq;       // [k][kMaxLimbs]
    QConst qc{};
};

struct HostSet30 {
    std::vector<int8_t> mfma_tab;  // [tiles][32][kMfmaK], empty when L % 4 != 0
    std::vector<int32_t> mfma_bias;  // [tiles][32]: 128 sum_m t_um of each row
    // int8 MFMA CRT: [crt_rt * 32][crt_kc * 32] table (stored offset by -128)
    // and the per-pseudo-limb corrections [8 crt_rt]; empty when the
    // modulus is too narrow for its limb count (bits(q) + 54 < 32 L)
    std::vector<int8_t> crt_tab;
    std::vector<int64_t> crt_c;
    CrtMfmaConst crt_mc{};
    uint32_t crt_rt = 0, crt_kc = 0;
    std::vector<uint64_t> primes;
    std::vector<Tw32> tw_fwd, tw_inv;  // [k][N] (host-built tables only; forward w stored negated)
    std::vector<Tw32> tw_pow;          // [k][2][log_n]: psi^(2^i), psi^-(2^i) (not negated)
    std::vector<PrimeConst32> pc;
    std::vector<uint32_t> place;
    std::vector<CrtConst32> cc;
    std::vector<uint32_t> miq;
    QConst qc{};
};

struct HostTables {
    uint32_t N = 0, L = 0, log_n = 0, q_bits = 0;
    // false: the [k][N] twiddle tables are left to the device builder
    // (launch_twiddles62/30) and only tw_pow is filled
    bool host_twiddles = true;
    HostSet62 s62;  // 62-bit primes: the reference's prime rule (solve.cu:1179-1204)
    HostSet30 s30;  // 30-bit primes: the same rule below 2^30
};

// Pieces of a table build made on the device (nc_tableset_create): the
// chosen primes of each set (descending) with their psi_N, the big-integer
// CRT products mod q, and the 30-bit set's int8 MFMA CRT table.  The builder
// checks every prime (Miller-Rabin, p = 1 mod 2N, the word bound, the
// coefficient bound) and recomputes sampled products with host arithmetic;
// the init self tests check psi and emulate the MFMA CRT over the table.
struct DeviceInit {
    std::vector<uint64_t> primes62, psi62, primes30, psi30;
    std::vector<uint32_t> prod62, prod30;  // [k + 1][kMaxLimbs]: (M / p_j) mod q for j < k, then M mod q
    std::vector<int8_t> crt_tab;           // empty: build it on the host
    std::vector<int64_t> crt_c;
    std::vector<uint32_t> neg_m_r;         // [kMaxLimbs]
};

// Builds both prime sets and every table (the [k][N] twiddles only when
// host_twiddles), from `init`'s device-made pieces where given; throws
// std::invalid_argument / std::runtime_error with a message on failure.
HostTables build_tables(uint32_t N, uint32_t L, const uint32_t* q, bool host_twiddles = true,
                        const DeviceInit* init = nullptr);

// Throws std::invalid_argument unless q is an odd modulus of L <= 32 limbs
// with 2 <= bits(q) and bits(q) + 70 <= 1088.
void validate_modulus(uint32_t L, const uint32_t* q);

// The prime rule over device-searched candidates: cands[i] = (p, psi or 0
// when p is not prime), descending; takes primes while bits(prod p) < log2 N
// + 2 bits(q) + 4 (as pick_primes does).  Returns false if the candidates
// run out first.
bool select_primes(const std::vector<std::pair<uint64_t, uint64_t>>& cands, uint32_t log_n, uint32_t q_bits,
                   size_t max_k, std::vector<uint64_t>& primes, std::vector<uint64_t>& psi);

// Degree-N tables on a shared set whose master tables `m` were built (without
// host twiddles) for its largest degree: same primes and N-independent
// tables, the powers table and N-dependent per-prime constants for N.
HostTables derive_tables(const HostTables& m, uint32_t N);

// The stored twiddle entry x of prime j (direction inverse), by host
// arithmetic: {psi^(+-bitrev(x)), Shoup quotient}, {0, 0} at x = 0, the
// 30-bit forward value negated.  Checks samples of device-built tables.
Tw twiddle_entry62(const HostTables& t, uint32_t j, uint32_t x, bool inverse);
Tw32 twiddle_entry30(const HostTables& t, uint32_t j, uint32_t x, bool inverse);

// Init-time known-answer checks of both sets (the analogue of
// solve.cu:1151-1166, 1285-1314, 1370-1415 for this formulation).
void self_test(const HostTables& t);

// Kernel configuration of one call.
// Matrix-core paths of the 30-bit set (bitmask).
constexpr uint32_t kMfmaDecompose = 1u, kMfmaCrt = 2u, kMfmaAll = 3u;

struct Choice {
    uint32_t width = 62;        // prime width: 62 or 30
    uint32_t log_blocks = 0;    // NTT outer split
    uint32_t mfma = kMfmaAll;   // int8 matrix-core paths where available
    uint32_t chunks = 1;        // batch split into this many chunks on two staggered streams
    uint32_t persist = 0;       // 1: the persistent batched form (one launch per batch) where available
};

}  // namespace nc

struct nc_tableset;

struct nc_plan {
    nc::HostTables host;
    nc_tableset* shared = nullptr;  // the table set this plan's tables live in (nullptr: its own)
    nc::DevPlan dev{};
    int device = 0;
    // resident device tables (one allocation)
    void* d_tables = nullptr;
    size_t table_bytes = 0;
    // workspace: residue rows of a and b for `ws_cap` products, sized for
    // the wider of the two prime sets
    void* d_res_a = nullptr;
    void* d_res_b = nullptr;
    uint32_t ws_cap = 0;
    size_t res_bytes = 0;  // bytes of one product's residues (max over widths)
    // device address of this device's scratch array (nc_multiply_device_noalloc)
    void* scratch = nullptr;
    // staging for the host-pointer path (one product)
    uint32_t* d_io = nullptr;   // a | b | c, N*L u32 each
    uint32_t* h_pin = nullptr;  // pinned a | b | c
    uint32_t* h_pin_dev = nullptr;  // h_pin as the GPU addresses it (null: copies move it)
    hipStream_t stream = nullptr;
    // chunked batches: two internal streams (default priority) and the events
    // that fork them from, stagger them and join them back to the caller's stream
    hipStream_t sub[2] = {nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_first = nullptr, ev_join[2] = {nullptr, nullptr};
    // the last enqueued work that uses d_res_a / d_res_b: every call that
    // uses them first makes its stream wait for it, then records it, so
    // calls on different streams reach the workspace one after another
    hipEvent_t ev_ws = nullptr;
    // host-path staging pipeline: one event per output piece
    static constexpr int kHostPieces = 4;
    hipEvent_t ev_piece[kHostPieces] = {};
    // Autotuned choice per batch-size bucket (ceil(log2 batch), 0..31);
    // `forced` pins every bucket (nc_plan_set_tuning).
    struct Tuned {
        nc::Choice choice;     // nc_multiply_device (may be chunked)
        nc::Choice unchunked;  // callers that run one pipeline on their own workspace (graphs, server)
        bool tuned = false;
        bool from_cache = false;
        float median_ms = 0.f;  // the chosen configuration's median run
        float tune_ms = 0.f;    // wall time of the nc_plan_autotune call
    };
    Tuned tune[32];
    bool forced = false;
    nc::Choice forced_choice;
    int mfma_mode = -1;  // -1: tuned (default all on), else a kMfma* mask (nc_plan_set_mfma)
    // persistent batched form: the error word its launches set when a
    // readiness wait reaches its bound (read and cleared by
    // nc_plan_persist_status), the pinned mode and the wait bound
    uint32_t* d_err = nullptr;
    int persist_mode = -1;                // -1: tuned (default off), 0: never, 1: wherever available
    uint64_t persist_tmo = 100000000ull;  // s_memrealtime ticks (100 MHz): 1 s
    uint32_t chunks_forced = 0;  // 0: tuned (default 1), else every batch uses this many chunks
};
