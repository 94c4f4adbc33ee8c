// Kernel launchers of the gfx950 negacyclic product (host-callable).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "nc_common.hpp"
#include "nc_persist.hpp"

namespace nc {

// Resident tables of the 62-bit prime set (the reference's prime rule).
struct DevSet62 {
    uint32_t k = 0;
    uint32_t tw_log = 0;            // twiddle rows are 2^tw_log entries apart (log2 N, or the shared set's 16)
    const Tw* tw_fwd = nullptr;     // [k][2^tw_log] psi^bitrev(x) for x in [1, N)
    const Tw* tw_inv = nullptr;     // [k][2^tw_log] psi^-bitrev(x)
    const PrimeConst* pc = nullptr; // [k]
    const uint64_t* place = nullptr;// [k][kMaxLimbs] 2^(32 l) 2^64 mod p_j
    const CrtConst* cc = nullptr;   // [k]
    const uint32_t* miq = nullptr;  // [k][kMaxLimbs] (M / p_j) mod q
    const QConst* qc = nullptr;     // q, (-M) mod q, Barrett constant
};

// Resident tables of the 30-bit prime set.
struct DevSet30 {
    uint32_t k = 0;
    uint32_t tw_log = 0;  // twiddle row stride, as in DevSet62
    const Tw32* tw_fwd = nullptr;
    const Tw32* tw_inv = nullptr;
    const PrimeConst32* pc = nullptr;
    const uint32_t* place = nullptr;  // [k][kMaxLimbs] 2^(32 l) 2^32 mod p_j
    const CrtConst32* cc = nullptr;
    const uint32_t* miq = nullptr;
    const QConst* qc = nullptr;
    // MFMA decomposition: [tiles][32][kMfmaK] int8 and the per-row bias
    // [tiles][32] (int32), or null
    const int8_t* mfma_tab = nullptr;
    const int32_t* mfma_bias = nullptr;
    uint32_t mfma_tiles = 0;
    // MFMA CRT: [crt_rt * 32][crt_kc * 32] int8 table and [8 crt_rt]
    // corrections, or null
    const int8_t* crt_tab = nullptr;
    const int64_t* crt_c = nullptr;
    const CrtMfmaConst* crt_mc = nullptr;
    uint32_t crt_rt = 0, crt_kc = 0;
};

// Device-side view of a plan: dimensions and both prime sets.
struct DevPlan {
    uint32_t N = 0, L = 0, log_n = 0;
    uint32_t cus = 256;  // compute units of the device
    uint32_t xcds = 0;   // its XCDs (0: unknown)
    DevSet62 w62;
    DevSet30 w30;
};

// Residue buffers are u64 (width 62) or u32 (width 30), layout [p][k][N].

// limbs -> canonical residues; polys a (na of them) then b (nb) in one launch.
// mfma: use the int8 matrix-core contraction where available (30-bit primes,
// L % 4 == 0, N >= 32); otherwise, and when false, the VALU kernels.
// fuse_s1 > 0 (only where decompose_fuses_outer) also runs the first fuse_s1
// forward NTT stages, leaving what launch_ntt_outer_fwd would.
bool decompose_fuses_outer(const DevPlan& d, int width, bool mfma, uint32_t log_blocks);
hipError_t launch_decompose(const DevPlan& d, int width, const uint32_t* a, void* res_a, uint32_t na,
                            const uint32_t* b, void* res_b, uint32_t nb, hipStream_t s, bool mfma = true,
                            uint32_t fuse_s1 = 0);

// Residue rows of a and b -> CRT-weighted residues of the exact negacyclic
// product, v_j = (c mod p_j) (M / p_j)^-1 mod p_j (the inverse transform's
// merged N^-1 constant carries the CRT weight), in place in ra (canonical
// [0, p)); rb is clobbered.  log_blocks is the outer
// split (0 = the whole row in one workgroup).  outer_fwd_done: the
// decomposition already ran the outer forward stages.  outer_inv_in_crt:
// stop after the inner kernel (lazy [0, 2p] values) and let
// launch_crt_inv2 run the outer inverse stages.
hipError_t launch_ntt_product(const DevPlan& d, int width, void* ra, void* rb, uint32_t nprod, uint32_t log_blocks,
                              hipStream_t s, bool outer_fwd_done = false, bool outer_inv_in_crt = false);

// The three parts of launch_ntt_product, for per-stage timing.
hipError_t launch_ntt_outer_fwd(const DevPlan& d, int width, void* ra, void* rb, uint32_t nprod, uint32_t log_blocks,
                                hipStream_t s);
hipError_t launch_ntt_inner(const DevPlan& d, int width, void* ra, const void* rb, uint32_t nprod,
                            uint32_t log_blocks, hipStream_t s);
hipError_t launch_ntt_outer_inv(const DevPlan& d, int width, void* ra, uint32_t nprod, uint32_t log_blocks,
                                hipStream_t s);

// CRT-weighted residues v_j = (c mod p_j) (M / p_j)^-1 mod p_j, canonical,
// -> L limbs mod q.  mfma: the int8 matrix-core kernel where available
// (30-bit primes, N >= 32, table present); otherwise the VALU kernels.
bool crt_mfma_available(const DevPlan& d);
hipError_t launch_crt(const DevPlan& d, int width, const void* res, uint32_t* out, uint32_t nprod, hipStream_t s,
                      bool mfma = true);

// Twiddle tables [k][N] built on the device from pow [k][2][log_n]
// (psi^(2^i), psi^-(2^i) with Shoup quotients) and the primes in pc.
hipError_t launch_twiddles62(Tw* fwd, Tw* inv, const Tw* pow, const PrimeConst* pc, uint32_t log_n, uint32_t k,
                             hipStream_t s);
hipError_t launch_twiddles30(Tw32* fwd, Tw32* inv, const Tw32* pow, const PrimeConst32* pc, uint32_t log_n,
                             uint32_t k, hipStream_t s);

// The 30-bit matrix-core CRT that also runs the last two inverse NTT stages
// (outer split 2): res holds the inner kernel's output (lazy [0, 2p]), and
// launch_ntt_product's outer inverse is skipped.  Available for the fixed
// shape L = 28 with 57..64 primes.
bool crt_fuses_outer_inv(const DevPlan& d, int width, bool mfma, uint32_t log_blocks);
hipError_t launch_crt_inv2(const DevPlan& d, const uint32_t* res, uint32_t* out, uint32_t nprod, hipStream_t s);

// The persistent batched form (nc_persist.hpp): one launch per batch of
// `batch` products at this plan's N.  Available for the 30-bit set with both
// matrix-core paths at outer split 2, L = 28, 57..64 primes, N = 2^10..2^14,
// a known XCD count (at most kPersistMaxXcd) and batch >= that count.
bool persist_available(const DevPlan& d, uint32_t batch);
bool persist_launch_enabled();  // false: every call runs the launch sequence
// Control block bytes of one launch (placed by the caller in device memory
// that no other launch uses meanwhile).
size_t persist_ctl_bytes(const DevPlan& d, uint32_t batch);
// Clears the control block and launches the kernel: ra / rb hold `batch`
// products' residue rows ([batch][k][N] u32 each), sticky is the plan's
// error word, tmo the bound of one readiness wait in s_memrealtime ticks
// (100 MHz).
hipError_t launch_persist30(const DevPlan& d, const uint32_t* a, const uint32_t* b, uint32_t* c, uint32_t* ra,
                            uint32_t* rb, uint32_t batch, uint32_t* ctl, uint32_t* sticky, uint64_t tmo,
                            hipStream_t s);

// Table-set initialisation on the device.
// Candidates top - i step, i < count (top = 1 mod step): out[i] = 0 if not
// prime, else the primitive step-th root psi = g^((p - 1) / step) of the
// smallest quadratic non-residue g.
hipError_t launch_prime_search(uint64_t top, uint64_t step, uint32_t count, uint64_t* out, hipStream_t s);
// out [k + 1][kMaxLimbs]: row j < k (M / p_j) mod q, row k M mod q (primes[k], q[L]).
hipError_t launch_crt_products(const uint64_t* primes, uint32_t k, const uint32_t* q, uint32_t L, uint32_t* out,
                               hipStream_t s);
// The int8 MFMA CRT table [rt * 32][kc * 32], its corrections [8 rt] and
// ((-M) mod q) 2^96 mod q [kMaxLimbs] from miq [k][kMaxLimbs] and neg_m.
hipError_t launch_crt_mfma_tables(const uint32_t* miq, const uint64_t* primes, uint32_t k, const uint32_t* q,
                                  uint32_t L, uint32_t rt, uint32_t kc, const uint32_t* neg_m, int8_t* tab,
                                  int64_t* corr, uint32_t* neg_m_r, hipStream_t s);

// Moves count buffers of `bytes` bytes each (a multiple of 8; every pointer
// 8-byte aligned) from src[i] to dst[i] with the GPU's own loads and stores:
// to_host, device memory -> mapped pinned host memory; else the reverse.
// Host pointers must be device-visible addresses of mapped pinned memory
// whose allocation holds the whole buffer.
constexpr uint32_t kHostXferMax = 64;  // buffers per launch
hipError_t launch_host_xfer(const void* const* src, void* const* dst, uint32_t count, uint64_t bytes, bool to_host,
                            hipStream_t s);

// n limbs of pseudo-random operands for timing runs: coefficients of L limbs
// below q, given q's highest non-zero limb index hi < L and its value top.
hipError_t launch_fill_operands(uint32_t* out, uint64_t n, uint32_t L, uint32_t hi, uint32_t top, uint64_t seed,
                                hipStream_t s);

// 32x32->64 multiply throughput probe (blocks x 256 threads x iters x 8 multiplies).
hipError_t launch_mul32_peak(const uint32_t* in, uint64_t* out, uint32_t blocks, uint32_t iters, hipStream_t s);

// Device address of the scratch array (kScratchBytes) on the current
// device, looked up once per device and cached; the array's host-side name
// is not a device address and is never passed to a kernel.
hipError_t scratch_address(void** out);

// Largest / smallest outer split the NTT kernels accept for this N.
uint32_t max_log_blocks(uint32_t log_n);
uint32_t min_log_blocks(uint32_t log_n);

}  // namespace nc
