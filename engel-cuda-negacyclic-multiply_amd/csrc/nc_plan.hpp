// Host-side plan: the resident per-(N, q) state behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "nc_common.hpp"
#include "nc_kernels.hpp"

namespace nc {

// Host copy of the tables (also what the tests can inspect through the ABI).
struct HostTables {
    uint32_t N = 0, L = 0, k = 0, log_n = 0, q_bits = 0;
    std::vector<uint64_t> primes;
    std::vector<Tw> tw_fwd, tw_inv;  // [k][N]
    std::vector<PrimeConst> pc;      // [k]
    std::vector<uint64_t> place;     // [k][kMaxLimbs]
    std::vector<CrtConst> cc;        // [k]
    std::vector<uint32_t> miq;       // [k][kMaxLimbs]
    QConst qc{};
};

// Builds the prime set (solve.cu:1174-1204 rule) and every table; throws
// std::invalid_argument / std::runtime_error with a message on failure.
HostTables build_tables(uint32_t N, uint32_t L, const uint32_t* q);

// Runs the init-time known-answer checks on the host copy (the analogue of
// solve.cu:1151-1166, 1285-1314, 1370-1415 for this formulation).
void self_test(const HostTables& t);

}  // namespace nc

struct nc_plan {
    nc::HostTables host;
    nc::DevPlan dev{};
    int device = 0;
    // resident device tables (one allocation)
    void* d_tables = nullptr;
    size_t table_bytes = 0;
    // workspace: residue rows of a and b for `ws_cap` products
    uint64_t* d_res_a = nullptr;
    uint64_t* d_res_b = nullptr;
    uint32_t ws_cap = 0;
    // staging for the host-pointer path (one product)
    uint32_t* d_io = nullptr;     // a | b | c, N*L u32 each
    uint32_t* h_pin = nullptr;    // pinned a | b | c
    hipStream_t stream = nullptr;
    // Tuning per batch-size bucket (bucket = ceil(log2 batch), 0..31): NTT
    // outer split and CRT threads per coefficient.  `forced` pins every
    // bucket to log_blocks / crt_split (nc_plan_set_tuning).
    struct Tuning {
        uint8_t log_blocks = 0;
        uint8_t crt_split = 1;
        uint8_t tuned = 0;
    };
    Tuning tune[32];
    bool forced = false;
    uint32_t log_blocks = 0;      // forced / reported NTT outer split
    uint32_t crt_split = 1;
};
