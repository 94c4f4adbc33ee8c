// Work queue of the persistent batched form (one launch per batch).
//
// The reference runs a product as a fixed sequence of launches
// (solve.cu:1489-1528).  The persistent form runs a whole batch in one
// launch: every XCD drains its own queue of work items of three kinds,
//   D  decomposition (with the first two forward NTT stages) of 512
//      coefficients of one operand, all primes;
//   T  the inner NTT of one product's block(s): both forward transforms,
//      the pointwise product and the inverse transform;
//   C  the CRT (with the last two inverse NTT stages) of 256 positions;
// for the products assigned to it (product P goes to XCD P mod nxcd, so that
// a product's hand-offs stay inside one L2).  The queue is cut into rounds
// of R slots; round r holds the D items of the XCD's local product r, the T
// items of product r - lag and the C items of product r - 2 lag, interleaved
// evenly, so the matrix-core / memory-bound stages and the VALU-bound NTT of
// different products run side by side on every compute unit.  The lag puts
// an item's inputs at least (lag - 1) R >= 3 W queue positions earlier (W:
// the workgroups of one XCD, which hold about W consecutive positions at
// once), so they are normally complete when it is taken.  A T item
// waits until all D items of its product are done, a C item until all T
// items are; both precede it in the queue, and workgroups take items in
// queue order, so the smallest unfinished item always has its inputs
// complete (no deadlock with any number of workgroups).
//
// This header is shared by the kernel and the host (schedule construction
// and a CPU emulation of the queue that the tests run).
#pragma once

#include <cstdint>

#ifndef __HIP__
#define NC_HD
#else
#define NC_HD __host__ __device__
#endif

namespace nc {

constexpr uint32_t kPersistMaxR = 512;     // slots per round (the schedule travels as a kernel argument)
constexpr uint32_t kPersistMaxXcd = 16;
constexpr uint32_t kPersistTrailer = 512;  // workspace bytes per product reserved for the control block
constexpr uint32_t kItemD = 0, kItemT = 1, kItemC = 2;
constexpr uint32_t kPersistMaxLag = 64;

// Per-XCD control block, xcd_words words (a multiple of 32: one 128-byte
// line group per XCD):
//   [0] queue head   [1] abort flag   [2] items done   [3] error bits
//   [4, 4 + m_max)              D items done, per local product
//   [4 + m_max, 4 + 2 m_max)    T items done, per local product
constexpr uint32_t kCtlHead = 0, kCtlAbort = 1, kCtlDone = 2, kCtlErr = 3, kCtlCounts = 4;
constexpr uint32_t kPersistErrTimeout = 1u;  // a readiness wait reached its bound

// Items per product of each kind.
struct PersistShape {
    uint32_t nD = 0, nT = 0, nC = 0, R = 0;
    uint32_t bpw = 1;  // NTT blocks per T item (one per sub-block of a workgroup)
};

NC_HD inline uint32_t persist_products_on(uint32_t batch, uint32_t nxcd, uint32_t x) {
    return batch > x ? (batch - x + nxcd - 1) / nxcd : 0u;
}

NC_HD inline uint32_t persist_xcd_words(uint32_t batch, uint32_t nxcd) {
    const uint32_t m_max = (batch + nxcd - 1) / nxcd;
    return (kCtlCounts + 2 * m_max + 31) / 32 * 32;
}

// Slot entry: kind << 12 | index.
NC_HD inline uint32_t persist_entry(uint32_t kind, uint32_t idx) { return kind << 12 | idx; }

struct PersistItem {
    uint32_t kind, idx;  // kind and index within its product's items of that kind
    uint32_t local;      // the product's index among the XCD's products
    bool valid;          // false: an empty slot (rounds before the first / after the last product)
};

// Rounds between a product's D, T and C items for W workgroups per XCD.
NC_HD inline uint32_t persist_lag(uint32_t R, uint32_t W) {
    const uint32_t lag = 1 + (3 * W + R - 1) / R;
    return lag < 2 ? 2 : lag > kPersistMaxLag ? kPersistMaxLag : lag;
}

// Queue position `it` of an XCD with m products.
NC_HD inline PersistItem persist_decode(uint32_t it, uint32_t R, const uint32_t* order, uint32_t m, uint32_t lag) {
    const uint32_t r = it / R, e = order[it - r * R];
    PersistItem x;
    x.kind = e >> 12;
    x.idx = e & 0xfffu;
    const uint32_t back = x.kind * lag;  // D: round r, T: r - lag, C: r - 2 lag
    x.valid = r >= back && r - back < m;
    x.local = r - back;
    return x;
}

// Total queue positions of an XCD with m products.
NC_HD inline uint32_t persist_slots(uint32_t R, uint32_t m, uint32_t lag) { return R * (m + 2 * lag); }

// Host side (nc_persist.cpp).
// Item counts of N = 2^log_n (blocks of N / 4 values) with k primes.
PersistShape persist_shape_for(uint32_t log_n, uint32_t k);
// One round's R slots.
void persist_order(const PersistShape& sh, uint32_t* order);

struct PersistEmuStats {
    uint64_t items = 0;          // queue positions holding an item, all XCDs
    double makespan = 0.0;       // emulated time units (relative item durations)
    double wait_fraction = 0.0;  // share of workgroup time spent waiting for inputs
    uint32_t R = 0, nD = 0, nT = 0, nC = 0;
};
// CPU emulation of the queues of `batch` products over nxcd XCDs with
// `workers` workgroups each; 0 when every item ran once, in dependency
// order, and the queues drained; else the first failed check (1 bad
// arguments, 2 round too long, 3 order not a permutation, 4 deadlock,
// 5 item ran twice, 6 item missing).
int persist_emulate(uint32_t log_n, uint32_t k, uint32_t batch, uint32_t nxcd, uint32_t workers, uint64_t seed,
                    PersistEmuStats* st);

}  // namespace nc
