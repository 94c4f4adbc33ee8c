// Host side of the persistent batched form's work queue (nc_persist.hpp):
// the item counts of a shape, one round's slot order, and a CPU emulation of
// the queue that the tests run (the kernel decodes queue positions with the
// same persist_decode and the same order).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <queue>
#include <random>
#include <vector>

#include "nc_persist.hpp"

namespace nc {

PersistShape persist_shape_for(uint32_t log_n, uint32_t k) {
    PersistShape sh;
    const uint32_t N = 1u << log_n;
    const uint32_t logb = log_n - 2;                   // NTT block of N / 4 values
    sh.bpw = logb >= 12 ? 1u : 1u << (12 - logb);      // blocks of one 256-thread workgroup
    const uint32_t chunks = N / 128;                   // 32-column chunks of the four spread tiles
    sh.nD = 2 * ((chunks + 3) / 4);                    // four chunks (one per wave) per item, both operands
    sh.nT = (4 * k + sh.bpw - 1) / sh.bpw;             // k primes x 4 blocks
    sh.nC = (N + 255) / 256;                           // 256 positions per item
    sh.R = sh.nD + sh.nT + sh.nC;
    return sh;
}

// Each kind's items spread evenly over the round: item i of a kind with n
// items sits at (4 i + 1 + kind) / (4 n), and the slots follow that order
// (ties: D, then T, then C).
void persist_order(const PersistShape& sh, uint32_t* order) {
    const uint32_t n[3] = {sh.nD, sh.nT, sh.nC};
    uint32_t next[3] = {0, 0, 0};
    for (uint32_t s = 0; s < sh.R; ++s) {
        int best = -1;
        for (int kind = 0; kind < 3; ++kind) {
            if (next[kind] >= n[kind]) continue;
            if (best < 0) {
                best = kind;
                continue;
            }
            // (4 i + 1 + kind) / (4 n_kind) < (4 i' + 1 + best) / (4 n_best) ?
            const uint64_t lhs = (uint64_t)(4 * next[kind] + 1 + kind) * n[best];
            const uint64_t rhs = (uint64_t)(4 * next[best] + 1 + best) * n[kind];
            if (lhs < rhs) best = kind;
        }
        order[s] = persist_entry((uint32_t)best, next[best]++);
    }
}

// Discrete-event emulation of one XCD's queue: `workers` workgroups take
// queue positions in order (an empty slot costs no time), wait for their
// item's inputs, and run it for a random duration of its kind.  Checks
// every item runs exactly once and only after its inputs are complete, and
// that the queue drains (a state where every workgroup waits is a deadlock).
int persist_emulate(uint32_t log_n, uint32_t k, uint32_t batch, uint32_t nxcd, uint32_t workers, uint64_t seed,
                    PersistEmuStats* st) {
    if (log_n < 10 || log_n > 14 || k == 0 || batch == 0 || nxcd == 0 || nxcd > kPersistMaxXcd || workers == 0)
        return 1;
    const PersistShape sh = persist_shape_for(log_n, k);
    if (sh.R > kPersistMaxR) return 2;
    std::vector<uint32_t> order(sh.R);
    persist_order(sh, order.data());
    // the order holds every item of every kind exactly once
    {
        std::vector<uint32_t> seen[3] = {std::vector<uint32_t>(sh.nD), std::vector<uint32_t>(sh.nT),
                                         std::vector<uint32_t>(sh.nC)};
        for (uint32_t s = 0; s < sh.R; ++s) {
            const uint32_t kind = order[s] >> 12, idx = order[s] & 0xfffu;
            if (kind > 2 || idx >= seen[kind].size() || seen[kind][idx]++) return 3;
        }
    }
    const uint32_t lag = persist_lag(sh.R, workers);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double dur[3] = {25.0, 12.0, 17.0};  // typical relative durations (D, T, C)
    PersistEmuStats out{};
    for (uint32_t x = 0; x < nxcd; ++x) {
        const uint32_t m = persist_products_on(batch, nxcd, x);
        const uint32_t slots = persist_slots(sh.R, m, lag);
        std::vector<uint32_t> dcnt(m, 0), tcnt(m, 0), ccnt(m, 0);
        std::vector<std::vector<uint8_t>> ran(3 * (size_t)m);
        for (uint32_t p = 0; p < m; ++p) {
            ran[3 * p + 0].assign(sh.nD, 0);
            ran[3 * p + 1].assign(sh.nT, 0);
            ran[3 * p + 2].assign(sh.nC, 0);
        }
        // a workgroup holds its current item and, as the kernel does, the
        // queue position it fetched for afterwards (taken before the current
        // item's wait and run)
        uint32_t head = 0;
        struct W {
            bool busy = false, waiting = false;
            PersistItem item{};
            uint32_t reserved = 0;
            double t_end = 0.0;
        };
        std::vector<W> w(workers);
        for (W& ww : w) ww.reserved = head++;  // the fetch before the first iteration
        double now = 0.0;
        // completion events (time, worker)
        using Ev = std::pair<double, uint32_t>;
        std::priority_queue<Ev, std::vector<Ev>, std::greater<Ev>> ev;
        auto ready = [&](const PersistItem& it) {
            if (it.kind == kItemT) return dcnt[it.local] >= sh.nD;
            if (it.kind == kItemC) return tcnt[it.local] >= sh.nT;
            return true;
        };
        auto start = [&](uint32_t i) {
            const PersistItem& it = w[i].item;
            // inputs complete at start (the check the kernel's wait makes)
            if (!ready(it)) return false;
            w[i].waiting = false;
            w[i].busy = true;
            w[i].t_end = now + dur[it.kind] * (0.5 + U(rng));
            ev.push({w[i].t_end, i});
            return true;
        };
        auto take = [&](uint32_t i) {  // the kernel's loop: current = reserved, fetch the next
            for (;;) {
                const uint32_t pos = w[i].reserved;
                if (pos >= slots) {
                    w[i].busy = w[i].waiting = false;
                    return;
                }
                w[i].reserved = head++;
                const PersistItem it = persist_decode(pos, sh.R, order.data(), m, lag);
                if (!it.valid) continue;
                w[i].item = it;
                w[i].waiting = true;
                start(i);
                return;
            }
        };
        for (uint32_t i = 0; i < workers; ++i) take(i);
        double wait_sum = 0.0;
        for (;;) {
            if (ev.empty()) {
                bool any_wait = false;
                for (const W& ww : w) any_wait = any_wait || ww.waiting;
                if (any_wait) return 4;  // every workgroup waits: deadlock
                break;
            }
            const Ev e = ev.top();
            ev.pop();
            const double prev = now;
            now = e.first;
            for (const W& ww : w) wait_sum += ww.waiting ? now - prev : 0.0;
            W& fin = w[e.second];
            const PersistItem it = fin.item;
            uint8_t& r = ran[3 * (size_t)it.local + it.kind][it.idx];
            if (r++) return 5;  // ran twice
            if (it.kind == kItemD) ++dcnt[it.local];
            else if (it.kind == kItemT) ++tcnt[it.local];
            else ++ccnt[it.local];
            fin.busy = false;
            // waiting workgroups whose inputs are now complete start
            for (uint32_t i = 0; i < workers; ++i)
                if (w[i].waiting) start(i);
            take(e.second);
        }
        for (uint32_t p = 0; p < m; ++p) {
            if (dcnt[p] != sh.nD || tcnt[p] != sh.nT || ccnt[p] != sh.nC) return 6;  // items missing
            for (int kind = 0; kind < 3; ++kind)
                for (uint8_t v : ran[3 * p + kind])
                    if (v != 1) return 6;
        }
        out.items += (uint64_t)m * sh.R;
        out.makespan = std::max(out.makespan, now);
        out.wait_fraction += wait_sum / std::max(1e-9, now * workers) / nxcd;
    }
    out.R = sh.R;
    out.nD = sh.nD;
    out.nT = sh.nT;
    out.nC = sh.nC;
    if (st) *st = out;
    return 0;
}

}  // namespace nc
