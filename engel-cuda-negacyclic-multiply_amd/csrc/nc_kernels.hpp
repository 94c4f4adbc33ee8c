// Kernel launchers of the gfx950 negacyclic product (host-callable).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "nc_common.hpp"

namespace nc {

// Device-side view of a plan: dimensions and resident tables.
struct DevPlan {
    uint32_t N, L, k, log_n;
    const Tw* tw_fwd;          // [k][N] psi^bitrev(x) for x in [1, N)
    const Tw* tw_inv;          // [k][N] psi^-bitrev(x)
    const PrimeConst* pc;      // [k]
    const uint64_t* place;     // [k][kMaxLimbs] 2^(32 l) * 2^64 mod p_j
    const CrtConst* cc;        // [k]
    const uint32_t* miq;       // [k][kMaxLimbs] (M / p_j) mod q
    const QConst* qc;          // q, (-M) mod q, Barrett constant
};

// limbs -> canonical residues. polys a (na of them) then b (nb) in one launch:
// res_a[p][j][i] = a_p[i] mod p_j, res_b likewise.
hipError_t launch_decompose(const DevPlan& d, const uint32_t* a, uint64_t* res_a, uint32_t na, const uint32_t* b,
                            uint64_t* res_b, uint32_t nb, hipStream_t s);

// Residue rows of a and b -> residues of the exact negacyclic product, in
// place in ra (canonical [0, p)).  rb is clobbered.  log_blocks is the outer
// split (0 = the whole row in one workgroup).
hipError_t launch_ntt_product(const DevPlan& d, uint64_t* ra, uint64_t* rb, uint32_t nprod, uint32_t log_blocks,
                              hipStream_t s);

// The three parts of launch_ntt_product, for per-stage timing.
hipError_t launch_ntt_outer_fwd(const DevPlan& d, uint64_t* ra, uint64_t* rb, uint32_t nprod, uint32_t log_blocks,
                                hipStream_t s);
hipError_t launch_ntt_inner(const DevPlan& d, uint64_t* ra, const uint64_t* rb, uint32_t nprod, uint32_t log_blocks,
                            hipStream_t s);
hipError_t launch_ntt_outer_inv(const DevPlan& d, uint64_t* ra, uint32_t nprod, uint32_t log_blocks, hipStream_t s);

// 32x32->64 multiply throughput probe (blocks x 256 threads x iters x 8 multiplies).
hipError_t launch_mul32_peak(const uint32_t* in, uint64_t* out, uint32_t blocks, uint32_t iters, hipStream_t s);

// canonical residues -> L limbs mod q. split = threads per coefficient (1 or 4).
hipError_t launch_crt(const DevPlan& d, const uint64_t* res, uint32_t* out, uint32_t nprod, uint32_t split,
                      hipStream_t s);

// Largest outer split the NTT kernels accept for this N (and smallest).
uint32_t max_log_blocks(uint32_t log_n);
uint32_t min_log_blocks(uint32_t log_n);

}  // namespace nc
